"""Training data, training and loading of the surrogate network.

The network (portal_mlp_*, BF16-MFMA kernels in csrc/mlp.hip) maps a node's local
state after the first pass -- log n_H2, log T_kin, per line log(J^0_0/B_nu(2.73 K))
and 100 J^2_0/J^0_0, and the first SE solution rho_1 (per level log10(sqrt(2J+1)
rho^0_0), per level with J > 0 1e3 rho^2_0/rho^0_0) -- to the converged lambda
iteration, as corrections to rho_1: per level log10(rho^0_0 / rho_1^0_0) and per level
with J > 0 1e3 (rho^2_0/rho^0_0 - rho_1^2_0/rho_1^0_0).  Corrections span a far
smaller range than the populations themselves, which is what a BF16 network with
about three significant digits per activation resolves best.  The method produces its
own training data: seeded clouds are iterated to convergence with the mixed-precision
path, and the first-pass features are paired with the converged solution.

    python -m portal_amd.surrogate --levels 10 --clouds 10 --nodes 200000

writes portal_amd/surrogate_weights/<species>.npz, which ``load_into`` finds again.  The
networks of the batch's other rotors (5 to 9 and 11 to 20 levels) are trained on clouds of
the batch recipe (density slopes in [-2, -1], 48 directions, 65 channels):

    python -m portal_amd.surrogate --levels 5 6 7 8 9 11 12 13 14 15 16 17 18 19 20 \
        --clouds 6 --nodes 40000 --nth 6 --nph 11 --nvel 32 --n-exp -2 -1
"""
from __future__ import annotations

import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "surrogate_weights")
WIDTH, DEPTH = 256, 4
LOG_FLOOR = -12.0


def species_key(sp) -> str:
    h = hashlib.sha1()
    for a in (sp.j_lev, sp.r_trans, sp.c_trans):
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
    for a in (sp.e_levs_hz, sp.t_Aji, sp.nu_hz, sp.t_Cji.ravel(order="F"), sp.col_temps):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    h.update(str(sp.kmax).encode())
    return f"L{sp.e_num}_R{sp.r_num}_{h.hexdigest()[:12]}"


FORMAT = 2  # corrections to the first SE solution


def dims(sp):
    nj = int((np.asarray(sp.j_lev) > 0).sum())
    return 2 + 2 * sp.r_num + sp.e_num + nj, sp.e_num + nj


def _state(pol_pops, nodes, sp):
    """log10(sqrt(2J+1) rho00) per level and 1e3 rho20/rho00 per level with J > 0."""
    rows = np.asarray(nodes) - 1
    g = np.sqrt(2.0 * np.asarray(sp.j_lev) + 1.0)
    r0 = pol_pops[rows][:, 0::2]
    r2 = pol_pops[rows][:, 1::2]
    lp = np.log10(np.clip(r0 * g[None, :], 10.0 ** LOG_FLOOR, None))
    cols = [i for i in range(sp.e_num) if sp.j_lev[i] > 0]
    with np.errstate(divide="ignore", invalid="ignore"):
        rat = np.where(r0[:, cols] != 0, r2[:, cols] / r0[:, cols], 0.0) * 1e3
    return lp, rat


def targets(pol_conv, pol_first, nodes, sp):
    """Network targets of the given 1-based nodes: the converged solution (pol_pops
    table, N x 2L) as corrections to the first SE solution."""
    lp, rat = _state(pol_conv, nodes, sp)
    lp1, rat1 = _state(pol_first, nodes, sp)
    return np.concatenate([lp - lp1, rat - rat1], axis=1).astype(np.float32)


def weights_path(sp) -> str:
    return os.path.join(WEIGHTS, species_key(sp) + ".npz")


def load_into(pt, sp) -> bool:
    """Create the network on the context and load trained weights; False when none exist."""
    path = weights_path(sp)
    if not os.path.exists(path):
        return False
    z = np.load(path, allow_pickle=False)
    fmt = int(z["format"]) if "format" in z.files else 1
    if fmt != FORMAT:  # weights of an earlier output definition
        return False
    n_in, n_out = dims(sp)
    pt.mlp_create(n_in, int(z["width"]), int(z["depth"]), n_out, seed=0)
    pt.mlp_set_params(z["params"], z["norm"])
    return True


def make_dataset(pt, sp, g, xth, nvel, dx, max_iter=80, tol=1e-6, tol_align=1e-8):
    """(features, targets, iterations) of one cloud: first-pass features, converged solution."""
    from .lib import PORTAL_MIXED
    pt.upload_grid(g)
    pt.set_levpop(g.lev_p)
    pp1, _ = pt.se_pass(xth, nvel, dx, precision=PORTAL_MIXED)
    feats, nodes = pt.surrogate_features()
    pt.set_levpop(g.lev_p)
    # Ng-accelerated: the same fixed point as the plain iteration, in fewer passes
    st = pt.lambda_iterate(xth, nvel, dx, max_iter, tol_pop=tol, tol_align=tol_align, precision=PORTAL_MIXED,
                           outputs=False, ng_accel=1)
    pp = pt.get_pol_pops()
    return feats, targets(pp, pp1, nodes, sp), st["iterations"]


def train(levels=10, clouds=10, nodes=200000, nth=9, nph=32, nvel=64, epochs=100, batch=4096, lr=1e-3,
          seed0=101, out=None, log=print, n_exp=(-1.7, -1.3)):
    from . import cloud
    from .api import Portal
    sp = cloud.co_like_rotor(levels)
    xth = cloud.legendre_all(nth, nph)
    pt = Portal(0)
    pt.upload_species(sp)
    X, Y = [], []
    for k in range(clouds + 1):  # the last cloud is held out for validation
        t = time.time()
        rng = np.random.default_rng(seed0 + k)
        # the production recipe with the density slope and field direction varied per cloud
        g = cloud.make_cloud(sp, nodes, seed=seed0 + k, n_exp=float(rng.uniform(*n_exp)))
        f, y, it = make_dataset(pt, sp, g, xth, nvel, cloud.DX_INWARD)
        X.append(f)
        Y.append(y)
        log(f"cloud {k}: {len(f)} samples, converged in {it} iterations, {time.time() - t:.1f}s")
    Xv, Yv = X.pop(), Y.pop()
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    n_in, n_out = dims(sp)
    norm = np.concatenate([X.mean(0), np.maximum(X.std(0), 1e-3), Y.mean(0), np.maximum(Y.std(0), 1e-3)])
    pt.mlp_create(n_in, WIDTH, DEPTH, n_out, seed=seed0)
    pt.mlp_set_params(None, norm.astype(np.float32))
    perm = np.random.default_rng(seed0).permutation(len(X))
    X, Y = X[perm], Y[perm]
    # step schedule: `epochs` at lr, then a quarter as many at lr/3, lr/10 and lr/30
    hist, ms = [], 0.0
    for k, (ep, f) in enumerate(((epochs, 1.0), (epochs // 4, 1 / 3), (epochs // 4, 0.1), (epochs // 4, 1 / 30))):
        if ep <= 0:
            continue
        h, t = pt.mlp_train(X, Y, epochs=ep, batch=batch, lr=lr * f, seed=seed0 + k)
        hist.append(h)
        ms += t
    hist = np.concatenate(hist)
    log(f"trained on {len(X)} samples: loss {hist[0]:.3e} -> {hist[-1]:.3e} in {ms / 1e3:.1f}s")
    # held-out error: log10 of sqrt(2J+1) rho00 per level, and 1e3 rho20/rho00
    P = pt.mlp_forward(Xv)
    L = sp.e_num
    e_pop = np.abs(P[:, :L] - Yv[:, :L])
    e_al = np.abs(P[:, L:] - Yv[:, L:])
    val = np.array([np.median(e_pop), np.quantile(e_pop, 0.99), e_pop.max(), np.median(e_al), e_al.max()])
    log(f"held-out cloud: |d log10 pop| median {val[0]:.2e}, 99% {val[1]:.2e}, max {val[2]:.2e}; "
        f"|d 1e3 rho20/rho00| median {val[3]:.2e}, max {val[4]:.2e}")
    params, nm = pt.mlp_get_params()
    path = out or weights_path(sp)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, params=params, norm=nm, width=WIDTH, depth=DEPTH, samples=len(X),
                        loss=hist, key=species_key(sp), heldout=val, format=FORMAT)
    pt.close()
    return path


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--levels", type=int, nargs="+", default=[10])
    ap.add_argument("--clouds", type=int, default=10)
    ap.add_argument("--nodes", type=int, default=200000)
    ap.add_argument("--nth", type=int, default=9)
    ap.add_argument("--nph", type=int, default=32)
    ap.add_argument("--nvel", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--n-exp", type=float, nargs=2, default=[-1.7, -1.3], help="range of the density slope")
    ap.add_argument("--out", default=None, help="weights file (default: the species' file in surrogate_weights/)")
    a = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    for lv in a.levels:
        t0 = time.time()
        print(train(lv, a.clouds, a.nodes, a.nth, a.nph, a.nvel, a.epochs, out=a.out if len(a.levels) == 1 else None,
                    log=lambda *x: print(*x, flush=True), n_exp=tuple(a.n_exp)), f"{time.time() - t0:.0f}s",
              flush=True)
