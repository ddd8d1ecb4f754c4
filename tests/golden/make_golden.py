"""Write the golden fixtures of tests/golden/ (run in the build container only).

Two sources:

* the reference's own routines, compiled from its Fortran sources by
  ``make -C oracle ref`` into ``oracle/_ref/libportal_ref.so``: quadrature,
  find_dir, black_body, trace_out, inverse_rt, make_AC, rotate_all_save,
  trace_out_raytrace, polarize_prop, rot_stokes(_back), ray_in_raytrace, and the
  input stage eat_lime_dim / eat_lime_no_delaunay on the model files in lime_model/;
* exact arithmetic (sympy) for what the reference cannot be built for here
  (pop_eqs.f90 / em_coefs.f90 need the WIGXJPF library, which this image lacks):
  Wigner 3j/6j symbols and the L&L04 7.20 coupling constants re-evaluated as
  exact sums.

The fixtures hold inputs (or the seed that regenerates them) and outputs only.
Usage:  python tests/golden/make_golden.py [star | star_image | node_pixels | eat_lime]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.bind import Reference  # noqa: E402
from portal_amd import cloud  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cloud_digest(g) -> str:
    h = hashlib.sha256()
    for a in (g.XYZ, g.vel_all, g.b_all, g.lev_p, g.nbr_ptr, g.nbr_idx):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    R = Reference()
    rng = np.random.default_rng(20240601)

    # -- quadrature, find_dir, black_body -------------------------------------------
    quad = {f"xth_{a}_{b}": R.legendre_all(a, b) for a, b in [(6, 11), (9, 32), (4, 8), (64, 64)]}
    bs, gs, ths, phs, dirs = [], [], [], [], []
    for _ in range(32):
        b = rng.normal(size=3)
        b /= np.linalg.norm(b)
        g = rng.random(3)
        g /= np.linalg.norm(g)
        th, ph = rng.random() * np.pi, rng.random() * 2 * np.pi
        bs.append(b); gs.append(g); ths.append(th); phs.append(ph)
        dirs.append(R.find_dir(b, g, th, ph))
    Ts = np.array([0.0, 2.73, 5.0, 20.0, 100.0, 1e3])
    nus = np.array([115.2712e9, 230.538e9, 345.796e9, 1.0e12])
    bb = np.array([[R.black_body(T, nu) for nu in nus] for T in Ts])
    np.savez_compressed(os.path.join(OUT, "geometry.npz"), **quad, b=np.array(bs), g=np.array(gs), th=np.array(ths),
                        ph=np.array(phs), dir=np.array(dirs), T=Ts, nu=nus, bb=bb)

    # -- config-1 cloud: walk and inward transfer ---------------------------------
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 4000, seed=1)
    xth = cloud.legendre_all(6, 11)
    dne = g.del_ne()
    nodes = np.unique(np.concatenate([[1, 2, 3, 4000, 4001 - 1], rng.integers(1, g.bulk_node_num, size=59)]))
    nodes = nodes[g.ab_all[nodes - 1] > 0]
    J0 = np.zeros((sp.r_num, len(nodes)))
    J2 = np.zeros((sp.r_num, len(nodes)))
    for i, n in enumerate(nodes):
        J0[:, i], J2[:, i] = R.inverse_rt(int(n), g, sp, dne, xth, 32, cloud.DX_INWARD, g.bvec_all[:, n - 1], g.gvec,
                                          4096)
    walks, wofs, wnode, wdir = [], [0], [], []
    for n in nodes[:24]:
        for k in range(0, xth.shape[1], 7):
            d = R.find_dir(g.bvec_all[:, n - 1], g.gvec, np.arccos(xth[0, k]), xth[1, k])
            dl = R.trace_out(int(n), g, dne, d, cloud.DX_INWARD, 4096)
            walks.append(dl)
            wofs.append(wofs[-1] + dl.shape[1])
            wnode.append(n)
            wdir.append(d)
    np.savez_compressed(os.path.join(OUT, "inward_4000_nodes.npz"), digest=cloud_digest(g), nodes=nodes, J0=J0, J2=J2,
                        walk=np.concatenate(walks, axis=1), walk_ofs=np.array(wofs), walk_node=np.array(wnode),
                        walk_dir=np.array(wdir))

    # -- make_AC --------------------------------------------------------------------
    acs = {}
    for nl in (4, 10):
        s = cloud.co_like_rotor(nl)
        J0v = rng.random(s.r_num) * 1e-16
        J2v = (rng.random(s.r_num) - 0.5) * 1e-18
        for T in (0.5, 1.0, 7.3, 55.0, 999.0):
            acs[f"ac_{nl}_{T}"] = R.make_ac(s, 3.2e11, T, J0v, J2v)
        acs[f"J0_{nl}"] = J0v
        acs[f"J2_{nl}"] = J2v
    # two collision partners (ortho/para mix 0.25/0.75), temperature-dependent rates
    s2 = cloud.co_like_rotor(5)
    ct = s2.col_temps
    C = np.stack([np.outer(np.linspace(1.0, 2.0, len(ct)), np.ones(s2.c_num)) * 1e-11,
                  np.outer(np.linspace(3.0, 1.0, len(ct)), np.ones(s2.c_num)) * 1e-11], axis=2)
    s2.t_Cji = np.asfortranarray(C)
    J0v = rng.random(s2.r_num) * 1e-16
    J2v = (rng.random(s2.r_num) - 0.5) * 1e-18
    for T in (0.5, 3.0, 42.0, 5e3):
        acs[f"ac2p_{T}"] = R.make_ac(s2, 1.7e10, T, J0v, J2v)
    acs["J0_2p"] = J0v
    acs["J2_2p"] = J2v
    acs["C_2p"] = C
    np.savez_compressed(os.path.join(OUT, "make_ac.npz"), **acs)

    # -- image pass pieces ------------------------------------------------------------
    th_inc, ph_az = np.pi / 4, 0.3
    vn, bn, xn = R.rotate_all_save(g, th_inc, ph_az, g.bvec_all)
    sel = np.arange(0, g.node_num, 97)
    im_pix, im_size = 64, 0.7
    pix = []
    for i in range(-im_pix, im_pix + 1, 9):
        for j in range(-im_pix, im_pix + 1, 13):
            pix.append((i, j))
    traces, tofs = [], [0]
    nu0 = np.array([sp.e_levs_hz[1] - sp.e_levs_hz[0]])
    for (i, j) in pix:
        dl, I_in = R.trace_out_raytrace(i * im_size / im_pix, j * im_size / im_pix, g, xn, dne, cloud.DX_IMAGE, 4096,
                                        nu0)
        traces.append(dl)
        tofs.append(tofs[-1] + dl.shape[1])
    pp = []
    for _ in range(40):
        eI = 10 ** rng.uniform(-16, -12)
        eQ = eI * rng.uniform(-0.3, 0.3)
        pI = 10 ** rng.uniform(-30, -28)
        pQ = pI * rng.uniform(-0.3, 0.3)
        ds = 10 ** rng.uniform(10, 14)
        I, Q, U = rng.random(3) * 1e-18
        pp.append([eI, eQ, pI, pQ, ds, I, Q, U, *R.polarize_prop(eI, eQ, pI, pQ, ds, I, Q, U)])
    rs = []
    for _ in range(20):
        b = rng.normal(size=3)
        b /= np.linalg.norm(b)
        Q, U = rng.normal(size=2)
        rs.append([*b, Q, U, *R.rot_stokes(Q, U, b), *R.rot_stokes(Q, U, b, back=True)])
    # ray_in_raytrace on synthetic paths, with and without dust
    rir = []
    rir_in = []
    for t in range(12):
        nc = int(rng.integers(1, 30))
        pol = np.zeros((4, nc))
        pol[0] = 10 ** rng.uniform(-13, -11, nc)
        pol[1] = pol[0] * rng.uniform(-0.2, 0.2, nc)
        pol[2] = 10 ** rng.uniform(-30, -29, nc)
        pol[3] = pol[2] * rng.uniform(-0.2, 0.2, nc)
        bl = rng.uniform(150, 600, nc)
        bv = rng.normal(size=(3, nc))
        bv /= np.linalg.norm(bv, axis=0)
        dust = (rng.random(nc) * 1e-16) if t % 3 == 0 else np.zeros(nc)
        bbl = rng.uniform(5, 40, nc)
        vl = rng.uniform(-2000, 2000, nc)
        ll = np.cumsum(rng.integers(0, 800, nc)).astype(np.int32)
        vo = rng.uniform(-3000, 3000)
        out = R.ray_in_raytrace(vo, nc, pol, bl, bv, dust, 0.15, 2.0, bbl, vl, ll, 1.5e11 * cloud.DX_IMAGE, 115.27e9,
                                1e-19)
        rir_in.append(dict(vo=vo, nc=nc, pol=pol.tolist(), bl=bl.tolist(), bv=bv.tolist(), dust=dust.tolist(),
                           bbl=bbl.tolist(), vl=vl.tolist(), ll=ll.tolist()))
        rir.append(out)
    np.savez_compressed(os.path.join(OUT, "image_pieces.npz"), th_inc=th_inc, ph_az=ph_az, sel=sel, vel_rot=vn[:, sel],
                        b_rot=bn[:, sel], xyz_rot=xn[:, sel], pix=np.array(pix), trace=np.concatenate(traces, axis=1),
                        trace_ofs=np.array(tofs), polprop=np.array(pp), rot=np.array(rs), rir_out=np.array(rir))
    with open(os.path.join(OUT, "ray_in_raytrace_inputs.json"), "w") as f:
        json.dump(rir_in, f)

    # -- exact Wigner symbols and L&L04 7.20 constants (sympy) -----------------------
    from sympy import Rational, sqrt
    from sympy.physics.wigner import wigner_3j, wigner_6j

    w3 = []
    for _ in range(150):
        j1, j2 = int(rng.integers(0, 21)), int(rng.integers(0, 21))
        j3 = int(rng.integers(abs(j1 - j2), j1 + j2 + 1))
        m1 = int(rng.integers(-j1, j1 + 1))
        m2 = int(rng.integers(-j2, j2 + 1))
        m3 = -m1 - m2
        if abs(m3) > j3:
            m3 = 0
            m1 = -m2 if abs(m2) <= j1 else 0
            m2 = -m1
        w3.append([j1, j2, j3, m1, m2, m3, float(wigner_3j(j1, j2, j3, m1, m2, m3).evalf(30))])
    for j in range(0, 21):
        for k in (0, 2, 4, 6, 8):
            for m in sorted({0, min(j, 1), j}):
                w3.append([j, j, k, m, -m, 0, float(wigner_3j(j, j, k, m, -m, 0).evalf(30))])
    w6 = []
    for j in range(0, 21):
        for jp in (j - 1, j, j + 1):
            if jp < 0:
                continue
            for k in (0, 2, 4, 6, 8):
                w6.append([jp, jp, k, j, j, 1, float(wigner_6j(jp, jp, k, j, j, 1).evalf(30))])
            w6.append([1, 1, 2, j, j, jp, float(wigner_6j(1, 1, 2, j, j, jp).evalf(30))])

    def coupling_exact(kind, j, k, j2, k2):
        """The m,q sums of pop_eqs.f90 evaluated exactly."""
        dj, dk = j2 - j, k2 - k
        if abs(dj) > 1 or not (dk == 0 or abs(dk) == 2):
            return 0.0, 0.0
        f1 = sqrt(3 * (2 * k + 1) * (2 * k2 + 1))
        s0 = Rational(0)
        s2 = Rational(0)
        for m in range(-j, j + 1):
            for q in (-1, 0, 1):
                sg = 1
                if kind == 0:
                    ml = q + m
                    a, b, c = wigner_3j(j, j, k, m, -m, 0), wigner_3j(j2, j2, k2, ml, -ml, 0), wigner_3j(j, j2, 1, -m, ml, -q)
                elif kind == 1:
                    mu = m - q
                    a, b, c = wigner_3j(j, j, k, m, -m, 0), wigner_3j(j2, j2, k2, mu, -mu, 0), wigner_3j(j2, j, 1, -mu, m, -q)
                elif kind == 2:
                    mu = m - q
                    a, b, c = wigner_3j(j, j, k, m, -m, 0), wigner_3j(j, j, k2, m, -m, 0), wigner_3j(j2, j, 1, -mu, m, -q)
                    sg = (-1) ** (1 + q)
                else:
                    ml = m + q
                    a, b, c = wigner_3j(j, j, k, m, -m, 0), wigner_3j(j, j, k2, m, -m, 0), wigner_3j(j, j2, 1, -m, ml, -q)
                    sg = (-1) ** (1 - q)
                ww = a * b * c * c
                if dk == 0:
                    s0 += sg * ww * wigner_3j(1, 1, 0, q, -q, 0)
                s2 += sg * ww * wigner_3j(1, 1, 2, q, -q, 0)
        if kind <= 1:
            fac = (2 * j2 + 1) * f1 * (-1) ** (1 - dj)
            return float((s0 * fac).evalf(30)), float((s2 * sqrt(5) * fac).evalf(30))
        fac = (2 * j + 1) * f1
        return float((s0 * fac).evalf(30)), float((s2 * fac * sqrt(5)).evalf(30))

    coup = []
    for kind in range(4):
        for j in range(0, 5):
            for j2 in (j - 1, j, j + 1):
                if j2 < 0:
                    continue
                for k in range(0, min(2 * j, 6) + 1, 2):
                    for k2 in range(0, min(2 * j2, 6) + 1, 2):
                        if kind >= 2 and k2 > min(2 * j, 6):
                            continue
                        c0, c2 = coupling_exact(kind, j, k, j2, k2)
                        coup.append([kind, j, k, j2, k2, c0, c2])
    tsp = []
    for j in range(0, 6):
        for ju in (j - 1, j, j + 1):
            if ju < 0:
                continue
            for k in range(0, min(2 * j, 2 * ju, 6) + 1, 2):
                v = (2 * ju + 1) * (-1) ** (1 + j + ju + k) * wigner_6j(ju, ju, k, j, j, 1)
                tsp.append([j, k, ju, k, float(v.evalf(30))])
    with open(os.path.join(OUT, "wigner_exact.json"), "w") as f:
        json.dump({"w3j": w3, "w6j": w6, "coupling": coup, "t_spont": tsp}, f)
    print("fixtures written to", OUT)
    make_star(R)
    make_star_image(R)
    make_node_pixels(R)


def make_star(R):
    """Star mode: star_all/circle_points/find_angles vectors, and inverse_rt in star mode
    (inverse_rt.f90:1-137 with star_mode = 1) on nodes of a 2,500-node cloud around a
    central star of two sizes, so that circle_points' one-point, ring and Vogel branches
    all occur."""
    rng = np.random.default_rng(77)

    def unit(v):
        return v / np.linalg.norm(v)

    sa_in, sa_ns, sa_x, sa_ofs = [], [], [], [0]
    for t in range(24):
        k, gv, b = unit(rng.normal(size=3)), unit(rng.normal(size=3)), unit(rng.normal(size=3))
        dOm = float(10 ** rng.uniform(-4, 0.3))
        nth, nph = (6, 11) if t % 2 else (9, 32)
        ns, x = R.star_all(nth, nph, k, gv, b, dOm)
        sa_in.append(np.concatenate([[nth, nph, dOm], k, gv, b]))
        sa_ns.append(ns)
        sa_x.append(x)
        sa_ofs.append(sa_ofs[-1] + x.shape[1])
    cp = []
    for n in (1, 4, 9, 10, 23):
        th_s, dth_s = float(rng.uniform(0.01, 0.5)), float(rng.uniform(0.001, 0.1))
        cp.append(np.concatenate([[n, th_s, dth_s], R.circle_points(n, th_s, dth_s).ravel(order="F")]))
    fa = []
    for _ in range(16):
        b, gv, d = unit(rng.normal(size=3)), unit(rng.normal(size=3)), unit(rng.normal(size=3))
        fa.append(np.concatenate([b, gv, d, R.find_angles(b, gv, d)]))
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    del_ne = g.del_ne()
    r = np.linalg.norm(g.XYZ[:, : g.bulk_node_num - 1], axis=0)
    Xs, Ts = np.zeros(3), 6000.0
    rows = []
    for Rs in (0.03, 0.15):
        cand = np.where((r > Rs + 0.02) & (g.ab_all[: g.bulk_node_num - 1] > 0))[0]
        near = cand[np.argsort(r[cand])[:4]]
        far = rng.choice(cand, 6, replace=False)
        for n in np.concatenate([near, far]) + 1:
            X = g.XYZ[:, n - 1]
            dv = Xs - X
            dn = np.linalg.norm(dv)
            dOm = np.pi * (Rs / dn) ** 2
            b = g.bvec_all[:, n - 1]
            ns, xs = R.star_all(6, 11, dv / dn, g.gvec, b, dOm)
            J0, J2 = R.inverse_rt_star(n, g, sp, del_ne, xs, ns, 32, cloud.DX_INWARD, b, g.gvec, 4096, Xs, Rs, Ts)
            rows.append(np.concatenate([[n, Rs, ns, xs.shape[1]], J0, J2]))
    np.savez_compressed(os.path.join(OUT, "star.npz"), sa_in=np.array(sa_in), sa_ns=np.array(sa_ns),
                        sa_x=np.concatenate(sa_x, axis=1), sa_ofs=np.array(sa_ofs),
                        circle=np.array([np.pad(c, (0, 3 + 3 * 23 - len(c))) for c in cp]),
                        find_angles=np.array(fa), inverse=np.array(rows), digest=cloud_digest(g), T_star=Ts)
    print("star fixtures written to", OUT)


def make_star_image(R):
    """Star mode of the image pass: rotate_all_save's rotated star position
    (miscellaneous.f90:900-905) and trace_out_raytrace with star_mod = 1
    (raytrace.f90:103-114: a pixel within R_star of the star's centre starts at
    z = Xs_z + R_star with I_in = B(T_star)) for pixels inside, on the edge of and
    outside the star's disc, on the 2,500-node cloud of the star fixtures."""
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    dne = g.del_ne()
    X_star, Rs, Ts = np.array([0.12, -0.07, 0.2]), 0.15, 6000.0
    th_inc, ph_az = 0.7, 0.4
    _, _, xn, xs = R.rotate_all_save(g, th_inc, ph_az, g.bvec_all, X_star=X_star)
    im_pix, im_size = 64, 0.7
    nu0 = np.array([sp.e_levs_hz[1] - sp.e_levs_hz[0], sp.e_levs_hz[3] - sp.e_levs_hz[2]])
    pix, traces, tofs, iin = [], [], [0], []
    for i in range(-im_pix, im_pix + 1, 3):
        for j in range(-im_pix, im_pix + 1, 3):
            x, y = i * im_size / im_pix, j * im_size / im_pix
            r = np.hypot(x - xs[0], y - xs[1])
            if abs(r - Rs) > 0.05 and (i * 7 + j * 3) % 11:
                continue  # keep the disc's edge and a sparse sample elsewhere
            dl, I_in = R.trace_out_raytrace(x, y, g, xn, dne, cloud.DX_IMAGE, 4096, nu0,
                                            star=dict(R_star=Rs, T_star=Ts, Xs=xs))
            pix.append((i, j))
            traces.append(dl)
            tofs.append(tofs[-1] + dl.shape[1])
            iin.append(I_in)
    bb = np.array([[R.black_body(T, nu) for nu in nu0] for T in (2.73, Ts)])
    np.savez_compressed(os.path.join(OUT, "star_image.npz"), X_star=X_star, R_star=Rs, T_star=Ts, th_inc=th_inc,
                        ph_az=ph_az, xs_rot=xs, im_pix=im_pix, im_size=im_size, nu0=nu0, pix=np.array(pix),
                        trace=np.concatenate(traces, axis=1), trace_ofs=np.array(tofs), I_in=np.array(iin), bb=bb,
                        digest=cloud_digest(g))
    print("star image fixtures written to", OUT, len(pix), "pixels")


def make_node_pixels(R):
    """The reference's pix = 0 image mode (para_finish.f90:245-330): trace_out_raytrace
    started at the rotated (x, y) of bulk nodes, on the 2,500-node cloud of the star
    fixtures, with and without a star."""
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    dne = g.del_ne()
    X_star, Rs, Ts = np.array([0.12, -0.07, 0.2]), 0.15, 6000.0
    th_inc, ph_az = 1.1, 2.3
    _, _, xn, xs = R.rotate_all_save(g, th_inc, ph_az, g.bvec_all, X_star=X_star)
    nu0 = np.array([sp.e_levs_hz[1] - sp.e_levs_hz[0]])
    rng = np.random.default_rng(5)
    nodes = np.sort(rng.choice(g.bulk_node_num - 1, 40, replace=False)) + 1
    traces, tofs, iin = [], [0], []
    for k, n in enumerate(nodes):
        star = dict(R_star=Rs, T_star=Ts, Xs=xs) if k % 2 else None
        dl, I_in = R.trace_out_raytrace(xn[0, n - 1], xn[1, n - 1], g, xn, dne, cloud.DX_IMAGE, 4096, nu0, star=star)
        traces.append(dl)
        tofs.append(tofs[-1] + dl.shape[1])
        iin.append(I_in[0])
    np.savez_compressed(os.path.join(OUT, "node_pixels.npz"), X_star=X_star, R_star=Rs, T_star=Ts, th_inc=th_inc,
                        ph_az=ph_az, nodes=nodes, xy=xn[:2, nodes - 1], trace=np.concatenate(traces, axis=1),
                        trace_ofs=np.array(tofs), I_in=np.array(iin), digest=cloud_digest(g))
    print("node-pixel fixtures written to", OUT)


LIME_DIR = os.path.join(OUT, "lime_model")
EAT_LIME_KEYS = ("XYZ", "vel_all", "temp_all", "lev_p", "e_levs", "t_Aji", "nu", "t_Cji", "col_temps", "h2_dens",
                 "b_all", "ab_all", "j_lev", "r_trans", "c_trans", "del_num_ne", "del_ne", "kd_v", "dust_p")


def lime_model_files():
    """A small model in LIME/LAMDA/dust-table formats: 200 interior nodes and their sinks,
    a 5-level rotor with two collision partners, a 67-row dust table, and a gas-to-dust
    ratio on about half of the nodes.  Returns the four paths (written once)."""
    from scipy.spatial import Delaunay
    from portal_amd import lime_io
    sp = cloud.co_like_rotor(5)
    sp2 = cloud.Species(e_num=sp.e_num, j_lev=sp.j_lev, e_levs_cm=sp.e_levs_cm, mass_mol=sp.mass_mol,
                        r_trans=sp.r_trans, t_Aji=sp.t_Aji, nu_ghz=sp.nu_ghz, c_trans=sp.c_trans,
                        col_temps=sp.col_temps,
                        t_Cji=np.asfortranarray(np.concatenate([sp.t_Cji, 0.7 * sp.t_Cji], axis=2)), kmax=6)
    g = cloud.make_cloud(sp2, 200, seed=7)
    pos = (g.XYZ * g.tx).T.copy()
    tets = Delaunay(pos).simplices
    lam = np.logspace(0, 4, 67)
    rng = np.random.default_rng(3)
    gd = np.where(rng.random(g.bulk_node_num) < 0.5, 100.0 * rng.uniform(0.5, 2, g.bulk_node_num), 0.0)
    return lime_io.write_model(LIME_DIR, g, sp2, tets, pos, dust_table=np.array([lam, 1e3 * lam ** -1.5]),
                               gas_to_dust=gd)


def make_eat_lime(R):
    """The reference's eat_lime_dim + eat_lime_no_delaunay on the committed model files."""
    paths = lime_model_files()
    rel = {k: os.path.relpath(v, OUT) for k, v in paths.items()}
    ref = R.eat_lime(paths["pop"], paths["grid"], paths["mol"], paths["dust"])
    scal = {k: np.array(ref[k]) for k in ("node_num", "del_num", "e_num", "r_num", "c_num", "c_temp", "c_par",
                                          "maxne", "rmax", "bulk_node_num", "mass_mol")}
    np.savez_compressed(os.path.join(OUT, "eat_lime.npz"), files=json.dumps(rel),
                        **{k: ref[k] for k in EAT_LIME_KEYS}, **scal)
    print("eat_lime fixtures written to", OUT)


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["eat_lime"]:
        make_eat_lime(Reference())
    elif sys.argv[1:] == ["star"]:
        make_star(Reference())
    elif sys.argv[1:] == ["star_image"]:
        make_star_image(Reference())
    elif sys.argv[1:] == ["node_pixels"]:
        make_node_pixels(Reference())
    else:
        main()
