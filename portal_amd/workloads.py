"""The other reference configurations, run on one GPU and reported as JSON lines.

    python -m portal_amd.workloads small        # 4,000 nodes, 4 levels: FP64, then mixed with its errors
    python -m portal_amd.workloads batch        # 64 independent clouds in one batched iteration (Ng)
    python -m portal_amd.workloads batch --surrogate   # the same with each rotor's network
    python -m portal_amd.workloads million      # the 4,000-node recipe at 1e6 nodes with the surrogate

`small` runs the 4,000-node cloud (48 directions, 65 channels, 30 iterations, a 129 x 129
x 65 image) in FP64 and again in mixed precision, and reports the mixed run's max and
median relative errors of the level populations, of rho^2_0/rho^0_0 and of the Stokes
images against the FP64 run.  `batch` builds 64 seeded clouds (node counts log-uniform
in [5e3, 1e5], rotors of 4..20 levels, density slopes in [-2, -1], random fields) and
iterates all of them to convergence in mixed precision with one persistent launch per
kernel per iteration; with --surrogate every model whose rotor has a trained network
starts from its prediction.  `million` iterates the 1e6-node cloud to convergence with and
without the surrogate and reports iterations and the Stokes error against FP64.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import cloud, driver, surrogate
from .api import Portal, batch_lambda_iterate
from .lib import PORTAL_FP64, PORTAL_MIXED

TOL_POP, TOL_ALIGN = 1e-6, 1e-8


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _image(pt, sp, g, nvel_img, dvel, im_pix, precision=PORTAL_FP64):
    pp = pt.get_pol_pops()
    rp = driver.rescale_pol_pops(pp, pt.get_levpop(), sp.lev_g)
    lines = driver.image_lines(sp, [1], [2])
    vel = driver.channel_velocities(nvel_img, dvel)
    (I, Q, U, T), st = pt.image_pass(rp, lines, np.pi / 4, 0.0, nvel_img, vel, im_pix, 0.7, cloud.DX_IMAGE,
                                     n_depth_rt=4096, precision=precision)
    return (I, Q, U), st


def stokes_error(ref, got):
    """Max and median |difference| of I, Q and U on the scale of max |I_ref|."""
    s = np.abs(ref[0]).max()
    out = {}
    for name, a, b in zip("IQU", ref, got):
        d = np.abs(b - a) / s
        out[f"{name}_max"] = float(d.max())
        out[f"{name}_median"] = float(np.median(d))
    return out


def _rel_errors(pp_ref, pp, idx):
    r0, q0 = pp_ref[idx][:, 0::2], pp[idx][:, 0::2]
    big = r0 > 1e-8 * r0.max(axis=0)
    e_pop = np.abs(q0 - r0)[big] / r0[big]
    a_ref = pp_ref[idx][:, 3::2] / pp_ref[idx][:, 2::2]
    a = pp[idx][:, 3::2] / pp[idx][:, 2::2]
    e_al = np.abs(a - a_ref)
    return {"pop_rel_max": float(e_pop.max()), "pop_rel_median": float(np.median(e_pop)),
            "align_abs_max": float(e_al.max()), "align_abs_median": float(np.median(e_al)),
            "align_scale": float(np.abs(a_ref).max())}


def small(seed=1):
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 4000, seed=seed)
    xth = cloud.legendre_all(6, 11)
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    res = {}
    for name, prec in (("fp64", PORTAL_FP64), ("mixed", PORTAL_MIXED)):
        pt.set_levpop(g.lev_p)
        t = time.perf_counter()
        pp, lev, st = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 30, tol_pop=0.0, tol_align=0.0, precision=prec,
                                        rebuild_paths=1)
        t = time.perf_counter() - t
        img, sti = _image(pt, sp, g, 32, 93.75, 64, precision=prec)
        res[name] = {"pp": pp, "img": img, "seconds": t, "iterations": st["iterations"],
                     "rt_ms": st["t_rt_ms"], "se_ms": st["t_se_ms"], "walk_ms": st["t_walk_ms"],
                     "image_ms": sti["t_img_ms"], "last_change": st["max_rel_change"]}
        _log(f"small {name}: 30 iterations {t:.3f}s, image {sti['t_img_ms']:.1f} ms")
    out = {"workload": "4,000-node cloud, 4-level rotor, 48 directions, 65 channels, 30 iterations, "
                       "129x129x65 image", "nodes": int(len(idx))}
    for name in ("fp64", "mixed"):
        r = res[name]
        out[name] = {k: r[k] for k in ("seconds", "iterations", "rt_ms", "se_ms", "walk_ms", "image_ms",
                                       "last_change")}
        out[name]["nodes_per_s_30_iterations"] = len(idx) / r["seconds"]
    out["mixed_vs_fp64"] = _rel_errors(res["fp64"]["pp"], res["mixed"]["pp"], idx)
    out["mixed_vs_fp64"]["stokes"] = stokes_error(res["fp64"]["img"], res["mixed"]["img"])
    pt.close()
    return out


def build_batch(n_models=64, seed=4000):
    """The batch configuration's models: n_models seeded clouds with node counts
    log-uniform in [5e3, 1e5], 4..20-level rotors, density slopes uniform in [-2, -1] and
    random fields, each uploaded to its own context at LTE.  Returns [(species, grid,
    Portal)]."""
    rng = np.random.default_rng(seed)
    out = []
    t0 = time.perf_counter()
    for m in range(n_models):
        n = int(round(np.exp(rng.uniform(np.log(5e3), np.log(1e5)))))
        L = int(rng.integers(4, 21))
        n_exp = float(rng.uniform(-2.0, -1.0))
        sp = cloud.co_like_rotor(L)
        g = cloud.make_cloud(sp, n, seed=seed + 1 + m, n_exp=n_exp)
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        out.append((sp, g, pt))
        if (m + 1) % 8 == 0:
            _log(f"batch: {m + 1} clouds built, {time.perf_counter() - t0:.1f}s")
    _log(f"batch: {n_models} clouds, {sum(g.node_num for _, g, _ in out)} nodes, built in "
         f"{time.perf_counter() - t0:.1f}s")
    return out


def batch(n_models=64, seed=4000, nth=6, nph=11, nvel=32, ng=True, surr=False):
    xth = cloud.legendre_all(nth, nph)
    models = build_batch(n_models, seed)
    pts = [pt for _, _, pt in models]
    have = [surr and surrogate.load_into(pt, sp) for sp, _, pt in models]
    sizes = [(g.node_num, sp.e_num, int((g.ab_all[: g.bulk_node_num] > 0).sum())) for sp, g, _ in models]
    t = time.perf_counter()
    # calls of at most 10 iterations, each continuing the last (the same iteration as one
    # call), with a line per call
    done, iters_at, dev_ms, rt_ms = 0, [0] * n_models, 0.0, 0.0
    while True:
        per, tot = batch_lambda_iterate(pts, [xth] * n_models, [nvel] * n_models, cloud.DX_INWARD, 10,
                                        tol_pop=TOL_POP, tol_align=TOL_ALIGN, precision=PORTAL_MIXED,
                                        rebuild_paths=1 if done == 0 else 0, ng_accel=1 if ng else 0,
                                        surrogate=1 if surr and done == 0 else 0)
        if done == 0:
            fb = [p["fallbacks"] for p in per]
        for m, p in enumerate(per):
            if not iters_at[m] and p["converged"]:
                iters_at[m] = done + p["iterations"]
        done += tot["iterations"]
        dev_ms += tot["t_total_ms"]
        rt_ms += tot["t_rt_ms"]
        _log(f"batch: {done} iterations, {sum(1 for v in iters_at if v)} of {n_models} models converged, "
             f"{time.perf_counter() - t:.1f}s")
        if tot["converged"] or done >= 400:
            break
    t = time.perf_counter() - t
    n_solved = sum(s[2] for s in sizes)
    out = {"workload": f"{n_models} clouds, nodes log-uniform in [5e3, 1e5], 4..20 levels, "
                       f"{xth.shape[1]} directions, {2 * nvel + 1} channels, mixed{' + surrogate' if surr else ''}"
                       f"{' + Ng' if ng else ''}, to {TOL_POP:g}/{TOL_ALIGN:g}",
           "models_with_surrogate": int(sum(have)),
           "surrogate_fallbacks": int(sum(fb)) if surr else None,
           "nodes": n_solved, "seconds": t, "converged_nodes_per_s": n_solved / t,
           "iterations_max": done, "all_converged": bool(all(iters_at)),
           "iterations_per_model": [v if v else done for v in iters_at],
           "levels_per_model": [s[1] for s in sizes], "nodes_per_model": [s[2] for s in sizes],
           "device_ms": dev_ms, "iteration_ms": rt_ms,
           "inward_form": {1: "staged persistent", 2: "cooperative persistent",
                           3: "cooperative, one workgroup per item"}.get(tot["rt_variant"], "?"),
           "se_form": {0: "one persistent launch", 1: "a launch per system size"}.get(tot["se_form"], "?"),
           "se_waves_used": [w for w in (1, 2, 4) if tot["se_waves"] & w]}
    out["unconverged"] = [{"model": m, "levels": sizes[m][1], "nodes": sizes[m][2],
                           "pop_change": per[m]["max_rel_change"], "align_change": per[m]["max_align_change"]}
                          for m in range(n_models) if not iters_at[m]]
    for pt in pts:
        pt.close()
    return out


def million(n_int=1000000, seed=5):
    sp = cloud.co_like_rotor(4)
    t0 = time.perf_counter()
    g = cloud.make_cloud(sp, n_int, seed=seed, log=_log)
    xth = cloud.legendre_all(6, 11)
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    _log(f"million: cloud of {g.node_num} nodes built in {time.perf_counter() - t0:.1f}s")
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    have = surrogate.load_into(pt, sp)
    out = {"workload": f"{n_int}-node cloud of the 4,000-node recipe, 4-level rotor, 48 directions, 65 channels, "
                       f"to {TOL_POP:g}/{TOL_ALIGN:g}", "nodes": int(len(idx)), "surrogate": bool(have)}
    runs = [("fp64", PORTAL_FP64, 0, 0), ("mixed", PORTAL_MIXED, 0, 0), ("mixed+ng", PORTAL_MIXED, 0, 1)]
    if have:
        runs += [("mixed+surrogate", PORTAL_MIXED, 1, 0), ("mixed+surrogate+ng", PORTAL_MIXED, 1, 1)]
    ref = None
    for name, prec, surr, ng in runs:
        pt.set_levpop(g.lev_p)
        t = time.perf_counter()
        # the plain FP64 iteration in calls of at most 10 iterations, each continuing the
        # last (the same iteration as one call), with a line per call; the others in one call
        done, walk = 0, 0.0
        chunk = 10 if prec == PORTAL_FP64 and not ng else 300
        while True:
            st = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, chunk, tol_pop=TOL_POP, tol_align=TOL_ALIGN,
                                   precision=prec, surrogate=surr if done == 0 else 0, ng_accel=ng,
                                   rebuild_paths=1 if done == 0 else 0, outputs=False)
            done += st["iterations"]
            walk += st["t_walk_ms"]
            _log(f"million {name}: {done} iterations, {time.perf_counter() - t:.1f}s")
            if st["converged"] or done >= 300:
                break
        st["iterations"], st["t_walk_ms"] = done, walk
        t = time.perf_counter() - t
        img, _ = _image(pt, sp, g, 32, 93.75, 64, precision=prec)
        if ref is None:
            ref = img
        out[name] = {"iterations": st["iterations"], "converged": bool(st["converged"]), "seconds": t,
                     "converged_nodes_per_s": len(idx) / t, "ng_steps": st["ng_steps"],
                     "surrogate_fallbacks": st["fallbacks"], "walk_ms": st["t_walk_ms"],
                     "stokes_err_vs_fp64": stokes_error(ref, img)}
    pt.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["small", "batch", "million"])
    ap.add_argument("--no-ng", action="store_true", help="batch: plain iteration")
    ap.add_argument("--surrogate", action="store_true", help="batch: start each model with a network from its "
                                                             "rotor's prediction")
    a = ap.parse_args()
    res = batch(ng=not a.no_ng, surr=a.surrogate) if a.which == "batch" else {"small": small, "million": million}[a.which]()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
