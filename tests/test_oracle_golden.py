"""The oracle against the reference's own routines (golden fixtures) and closed forms.

The fixtures were written by tests/golden/make_golden.py from the reference's
Fortran compiled from its sources (oracle/_ref) and from exact sympy arithmetic.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from oracle.bind import _d, _f64
from portal_amd import cloud
import ctypes as C


def _digest(g):
    h = hashlib.sha256()
    for a in (g.XYZ, g.vel_all, g.b_all, g.lev_p, g.nbr_ptr, g.nbr_idx):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def cloud4k():
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 4000, seed=1)
    return sp, g


def test_quadrature_matches_reference(golden_dir):
    z = np.load(os.path.join(golden_dir, "geometry.npz"))
    for key in [k for k in z.files if k.startswith("xth_")]:
        _, a, b = key.split("_")
        mine = cloud.legendre_all(int(a), int(b))
        ref = z[key]
        assert mine.shape == ref.shape
        np.testing.assert_allclose(mine, ref, rtol=0, atol=1e-14)
        assert abs(ref[2].sum() - 1.0) < 1e-13
    # the direction counts the configs rely on
    assert cloud.legendre_all(6, 11).shape[1] == 48
    assert cloud.legendre_all(9, 32).shape[1] == 200


def test_oracle_quadrature_nodes(oracle):
    for n in (4, 6, 9, 64):
        x = np.zeros(n)
        w = np.zeros(n)
        oracle.lib.orc_legendre_nodes(n, _d(x), _d(w))
        xr, wr = np.polynomial.legendre.leggauss(n)
        np.testing.assert_allclose(x, xr, atol=1e-15)
        np.testing.assert_allclose(w, wr, atol=1e-14)


def test_find_dir_and_black_body(oracle, golden_dir):
    z = np.load(os.path.join(golden_dir, "geometry.npz"))
    for b, g, th, ph, d in zip(z["b"], z["g"], z["th"], z["ph"], z["dir"]):
        out = np.zeros(3)
        oracle.lib.orc_find_dir(_d(_f64(b)), _d(_f64(g)), C.c_double(th), C.c_double(ph), _d(out))
        np.testing.assert_allclose(out, d, rtol=0, atol=1e-15)
        # closed form: unit vector at angle th from B
        assert abs(np.linalg.norm(out) - 1) < 1e-14
        assert abs(math.acos(max(-1.0, min(1.0, out @ b))) - th) < 1e-7
    for i, T in enumerate(z["T"]):
        for j, nu in enumerate(z["nu"]):
            v = oracle.lib.orc_black_body(T, nu)
            assert v == pytest.approx(z["bb"][i, j], rel=1e-15, abs=0)
            if T > 0:  # Planck's law with the reference's constants 2h/c^2, h/k
                assert v == pytest.approx(2 * 6.62607015e-34 * nu**3 / 299792458.0**2 / math.expm1(6.62607015e-34 * nu / (1.380649e-23 * T)), rel=2e-6)


def test_walk_matches_reference(oracle, cloud4k, golden_dir):
    sp, g = cloud4k
    z = np.load(os.path.join(golden_dir, "inward_4000_nodes.npz"))
    if str(z["digest"]) != _digest(g):
        pytest.skip("cloud generator output differs from the fixture's")
    G = oracle.grid(g)
    walk, ofs = z["walk"], z["walk_ofs"]
    for k, (n, d) in enumerate(zip(z["walk_node"], z["walk_dir"])):
        cells = np.zeros(4097, dtype=np.int32)
        lens = np.zeros(4097, dtype=np.int32)
        nc = oracle.lib.orc_trace_out(int(n), C.byref(G), _d(_f64(d)), C.c_double(cloud.DX_INWARD), 4096,
                                      cells.ctypes.data_as(C.POINTER(C.c_int)), lens.ctypes.data_as(C.POINTER(C.c_int)))
        ref = walk[:, ofs[k]:ofs[k + 1]]
        assert nc == ref.shape[1]
        np.testing.assert_array_equal(cells[:nc], ref[0])
        np.testing.assert_array_equal(lens[:nc], ref[1])


def test_inverse_rt_matches_reference(oracle, cloud4k, golden_dir):
    sp, g = cloud4k
    z = np.load(os.path.join(golden_dir, "inward_4000_nodes.npz"))
    if str(z["digest"]) != _digest(g):
        pytest.skip("cloud generator output differs from the fixture's")
    xth = cloud.legendre_all(6, 11)
    nodes = z["nodes"]
    lo, hi = int(nodes.min()), int(nodes.max())
    for n in nodes[:20]:
        _, J0, J2 = oracle.se_pass(g, sp, xth, 32, cloud.DX_INWARD, node_lo=int(n), node_hi=int(n))
        i = list(nodes).index(n)
        np.testing.assert_allclose(J0[:, n - 1], z["J0"][:, i], rtol=1e-13, atol=0)
        np.testing.assert_allclose(J2[:, n - 1], z["J2"][:, i], rtol=0, atol=1e-13 * np.abs(z["J0"][:, i]).max())
    assert lo >= 1 and hi <= g.bulk_node_num


def test_make_ac_matches_reference(oracle, golden_dir):
    z = np.load(os.path.join(golden_dir, "make_ac.npz"))
    for nl in (4, 10):
        s = cloud.co_like_rotor(nl)
        for T in (0.5, 1.0, 7.3, 55.0, 999.0):
            ac = oracle.make_ac(s, 3.2e11, T, z[f"J0_{nl}"], z[f"J2_{nl}"])
            np.testing.assert_allclose(ac, z[f"ac_{nl}_{T}"], rtol=1e-15, atol=0)
    s2 = cloud.co_like_rotor(5)
    s2.t_Cji = np.asfortranarray(z["C_2p"])
    for T in (0.5, 3.0, 42.0, 5e3):
        ac = oracle.make_ac(s2, 1.7e10, T, z["J0_2p"], z["J2_2p"])
        np.testing.assert_allclose(ac, z[f"ac2p_{T}"], rtol=1e-15, atol=0)


def test_wigner_exact(oracle, golden_dir):
    d = json.load(open(os.path.join(golden_dir, "wigner_exact.json")))
    for j1, j2, j3, m1, m2, m3, v in d["w3j"]:
        got = oracle.wig3jj(2 * j1, 2 * j2, 2 * j3, 2 * m1, 2 * m2, 2 * m3)
        assert got == pytest.approx(v, rel=1e-14, abs=1e-15)
    for a, b, c, e, f, g, v in d["w6j"]:
        got = oracle.wig6jj(2 * a, 2 * b, 2 * c, 2 * e, 2 * f, 2 * g)
        assert got == pytest.approx(v, rel=1e-14, abs=1e-15)


def test_coupling_constants_exact(oracle, golden_dir):
    d = json.load(open(os.path.join(golden_dir, "wigner_exact.json")))
    names = ["t_abs_const", "t_stim_const", "r_abs_const", "r_stim_const"]
    for kind, j, k, j2, k2, c0, c2 in d["coupling"]:
        g0, g2 = oracle.coef(names[kind], j, k, j2, k2)
        assert g0 == pytest.approx(c0, rel=1e-13, abs=1e-14)
        assert g2 == pytest.approx(c2, rel=1e-13, abs=1e-14)
    for j, k, ju, ku, v in d["t_spont"]:
        assert oracle.coef("t_spont_const", j, k, ju, ku) == pytest.approx(v, rel=1e-13, abs=1e-14)


def test_coupling_selection_rules(oracle):
    # zero unless |dJ| <= 1 and dK in {0, +-2}
    for j in range(0, 5):
        for j2 in range(0, 7):
            for k in range(0, 7, 2):
                for k2 in range(0, 7, 2):
                    t0, t2 = oracle.coef("t_abs_const", j, k, j2, k2)
                    if abs(j2 - j) > 1 or abs(k2 - k) > 2:
                        assert t0 == 0 and t2 == 0
                    if k2 != k:
                        assert t0 == 0


def test_image_pieces_match_reference(oracle, cloud4k, golden_dir):
    sp, g = cloud4k
    z = np.load(os.path.join(golden_dir, "image_pieces.npz"))
    N = g.node_num
    vn, bn, xn = (np.zeros((3, N), order="F") for _ in range(3))
    oracle.lib.orc_rotate_all_save(N, C.c_double(float(z["th_inc"])), C.c_double(float(z["ph_az"])), _d(_f64(g.vel_all)),
                                   _d(_f64(g.bvec_all)), _d(_f64(g.XYZ)), _d(vn), _d(bn), _d(xn), None, None)
    sel = z["sel"]
    np.testing.assert_allclose(vn[:, sel], z["vel_rot"], rtol=1e-15, atol=1e-12)
    np.testing.assert_allclose(bn[:, sel], z["b_rot"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(xn[:, sel], z["xyz_rot"], rtol=0, atol=1e-15)
    ptr, idx = _f64(g.nbr_ptr).astype(np.int32), np.asarray(g.nbr_idx, dtype=np.int32)
    im_pix, im_size = 64, 0.7
    for k, (i, j) in enumerate(z["pix"]):
        cells = np.zeros(4096, dtype=np.int32)
        steps = np.zeros(4096, dtype=np.int32)
        nc = C.c_int()
        rc = oracle.lib.orc_trace_out_raytrace(C.c_double(i * im_size / im_pix), C.c_double(j * im_size / im_pix), N,
                                               _d(xn), ptr.ctypes.data_as(C.POINTER(C.c_int)),
                                               idx.ctypes.data_as(C.POINTER(C.c_int)), C.c_double(cloud.DX_IMAGE), 4096,
                                               0, C.c_double(0.0), None, cells.ctypes.data_as(C.POINTER(C.c_int)),
                                               steps.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nc), None)
        assert rc == 0
        ref = z["trace"][:, z["trace_ofs"][k]:z["trace_ofs"][k + 1]]
        assert nc.value == ref.shape[1]
        np.testing.assert_array_equal(cells[:nc.value], ref[0])
        np.testing.assert_array_equal(steps[:nc.value], ref[1])
    for row in z["polprop"]:
        eI, eQ, pI, pQ, ds, I, Q, U = row[:8]
        a = [C.c_double(I), C.c_double(Q), C.c_double(U)]
        oracle.lib.orc_polarize_prop(*[C.c_double(v) for v in (eI, eQ, pI, pQ, ds)], *[C.byref(v) for v in a])
        np.testing.assert_allclose([v.value for v in a], row[8:], rtol=1e-13, atol=1e-300)
    for row in z["rot"]:
        b = _f64(row[:3])
        qn, un = C.c_double(), C.c_double()
        oracle.lib.orc_rot_stokes(C.c_double(row[3]), C.c_double(row[4]), C.byref(qn), C.byref(un), _d(b))
        np.testing.assert_allclose([qn.value, un.value], row[5:7], rtol=1e-12, atol=1e-13)
        oracle.lib.orc_rot_stokes_back(C.c_double(row[3]), C.c_double(row[4]), C.byref(qn), C.byref(un), _d(b))
        np.testing.assert_allclose([qn.value, un.value], row[7:9], rtol=1e-12, atol=1e-13)


def test_ray_in_raytrace_matches_reference(oracle, golden_dir):
    z = np.load(os.path.join(golden_dir, "image_pieces.npz"))
    inputs = json.load(open(os.path.join(golden_dir, "ray_in_raytrace_inputs.json")))
    for inp, ref in zip(inputs, z["rir_out"]):
        out = np.zeros(8)
        nc = inp["nc"]
        pol = _f64(np.array(inp["pol"]))  # (4, nc) column-major as pol_rayt_list
        bv = _f64(np.array(inp["bv"]))    # (3, nc)
        oracle.lib.orc_ray_in_raytrace(C.c_double(inp["vo"]), nc, _d(pol),
                                       _d(_f64(inp["bl"])), _d(bv), _d(_f64(inp["dust"])),
                                       C.c_double(0.15), C.c_double(2.0), _d(_f64(inp["bbl"])), _d(_f64(inp["vl"])),
                                       np.asarray(inp["ll"], dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int)),
                                       C.c_double(1.5e11 * cloud.DX_IMAGE), C.c_double(115.27e9), C.c_double(1e-19), _d(out))
        scale = max(abs(ref[0]), abs(ref[4]), 1e-300)
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-13 * scale)


def test_star_mode_matches_reference(oracle, golden_dir):
    # star_all, circle_points and find_angles (miscellaneous.f90:86-134, 1139-1265) and
    # inverse_rt in star mode (inverse_rt.f90:1-137, check_neigh_fast's star test
    # miscellaneous.f90:1056-1069) against the compiled reference
    from tests.golden.make_golden import cloud_digest
    z = np.load(os.path.join(golden_dir, "star.npz"))
    for k, row in enumerate(z["sa_in"]):
        nth, nph, dOm = int(row[0]), int(row[1]), row[2]
        kv, gv, bv = row[3:6], row[6:9], row[9:12]
        ns, x = oracle.star_all(nth, nph, kv, gv, bv, dOm)
        ref = z["sa_x"][:, z["sa_ofs"][k]:z["sa_ofs"][k + 1]]
        assert ns == z["sa_ns"][k] and x.shape == ref.shape
        np.testing.assert_allclose(x, ref, rtol=0, atol=1e-14)
    for row in z["circle"]:
        n, th_s, dth_s = int(row[0]), row[1], row[2]
        x = np.zeros((3, n), order="F")
        oracle.lib.orc_circle_points(n, C.c_double(th_s), C.c_double(dth_s), _d(x))
        np.testing.assert_allclose(x.ravel(order="F"), row[3:3 + 3 * n], rtol=0, atol=1e-15)
    for row in z["find_angles"]:
        th, ph = C.c_double(), C.c_double()
        oracle.lib.orc_find_angles(_d(_f64(row[0:3])), _d(_f64(row[3:6])), _d(_f64(row[6:9])), C.byref(th),
                                   C.byref(ph))
        np.testing.assert_allclose([th.value, ph.value], row[9:11], rtol=0, atol=1e-14)
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    assert cloud_digest(g) == str(z["digest"])
    inv = z["inverse"]
    assert len(set(inv[:, 2])) > 1  # more than one circle_points branch among the nodes
    xth = cloud.legendre_all(6, 11)
    for Rs in sorted(set(inv[:, 1])):
        rows = inv[inv[:, 1] == Rs]
        nodes = rows[:, 0].astype(np.int32)
        star = dict(nth=6, nph=11, R_star=float(Rs), T_star=float(z["T_star"]), X_star=np.zeros(3))
        J0, J2 = oracle.inverse_nodes(g, sp, xth, 32, cloud.DX_INWARD, nodes, star=star)
        np.testing.assert_allclose(J0.T, rows[:, 4:7], rtol=1e-13, atol=0)
        np.testing.assert_allclose(J2.T, rows[:, 7:10], rtol=0, atol=1e-13 * np.abs(rows[:, 4:7]).max())


def test_star_image_matches_reference(oracle, golden_dir):
    # the image pass in star mode: rotate_all_save's rotated star (miscellaneous.f90:900-905)
    # and trace_out_raytrace's star start (raytrace.f90:103-114) against the compiled
    # reference, for pixels on, near and off the star's disc
    from portal_amd import driver
    from tests.golden.make_golden import cloud_digest
    z = np.load(os.path.join(golden_dir, "star_image.npz"))
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    assert cloud_digest(g) == str(z["digest"])
    N = g.node_num
    vn, bn, xn = (np.zeros((3, N), order="F") for _ in range(3))
    xs = np.zeros(3)
    oracle.lib.orc_rotate_all_save(N, C.c_double(float(z["th_inc"])), C.c_double(float(z["ph_az"])),
                                   _d(_f64(g.vel_all)), _d(_f64(g.bvec_all)), _d(_f64(g.XYZ)), _d(vn), _d(bn), _d(xn),
                                   _d(_f64(z["X_star"])), _d(xs))
    np.testing.assert_allclose(xs, z["xs_rot"], rtol=0, atol=1e-16)
    ptr, idx = _f64(g.nbr_ptr).astype(np.int32), np.asarray(g.nbr_idx, dtype=np.int32)
    im_pix, im_size, Rs = int(z["im_pix"]), float(z["im_size"]), float(z["R_star"])
    hits = 0
    for k, (i, j) in enumerate(z["pix"]):
        cells = np.zeros(4096, dtype=np.int32)
        steps = np.zeros(4096, dtype=np.int32)
        nc, hit = C.c_int(), C.c_int()
        rc = oracle.lib.orc_trace_out_raytrace(C.c_double(i * im_size / im_pix), C.c_double(j * im_size / im_pix), N,
                                               _d(xn), ptr.ctypes.data_as(C.POINTER(C.c_int)),
                                               idx.ctypes.data_as(C.POINTER(C.c_int)), C.c_double(cloud.DX_IMAGE), 4096,
                                               1, C.c_double(Rs), _d(xs), cells.ctypes.data_as(C.POINTER(C.c_int)),
                                               steps.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nc), C.byref(hit))
        assert rc == 0
        ref = z["trace"][:, z["trace_ofs"][k]:z["trace_ofs"][k + 1]]
        assert nc.value == ref.shape[1]
        np.testing.assert_array_equal(cells[:nc.value], ref[0])
        np.testing.assert_array_equal(steps[:nc.value], ref[1])
        np.testing.assert_array_equal(z["I_in"][k], z["bb"][hit.value])  # B(T_star) exactly where the star starts
        hits += hit.value
    assert 0 < hits < len(z["pix"])
    # the whole pass: with no line emission a star pixel reads B(T_star) - B(2.73) and every
    # other pixel 0; with populations, pixels off the disc are those of the pass without star
    star = dict(R_star=Rs, T_star=float(z["T_star"]), X_star=z["X_star"])
    lines = driver.image_lines(sp, [1, 2], [2, 3])
    vel_out = driver.channel_velocities(3, 300.0)
    th, ph, ip = float(z["th_inc"]), float(z["ph_az"]), 12
    zero = np.zeros((N, 2 * sp.e_num), order="F")
    I0, Q0, U0, T0 = oracle.image_pass(g, sp, zero, lines, th, ph, 3, vel_out, ip, im_size, cloud.DX_IMAGE, star=star)
    n_side = 2 * ip + 1
    ii, jj = np.divmod(np.arange(n_side * n_side), n_side)
    x, y = (ii - ip) * im_size / ip, (jj - ip) * im_size / ip
    on = np.hypot(x - xs[0], y - xs[1]) < Rs
    assert 10 < on.sum() < on.size // 4
    nu = np.array([ln[2] for ln in lines])
    bb = np.array([[orc_black_body(oracle, T, v) for v in nu] for T in (2.73, star["T_star"])])
    np.testing.assert_array_equal(I0[~on], 0.0)
    np.testing.assert_allclose(I0[on], np.broadcast_to((bb[1] - bb[0])[None, :, None], I0[on].shape), rtol=1e-15)
    assert not Q0.any() and not U0.any() and not T0.any()
    pp = np.zeros((N, 2 * sp.e_num), order="F")
    rng = np.random.default_rng(8)
    nb = g.bulk_node_num - 1
    pp[:nb, 0::2] = g.lev_p[:, :nb].T
    pp[:nb, 3::2] = rng.uniform(-1e-3, 1e-3, (nb, sp.e_num - 1))
    pp = driver.rescale_pol_pops(pp, g.lev_p, sp.lev_g)
    with_star = oracle.image_pass(g, sp, pp, lines, th, ph, 3, vel_out, ip, im_size, cloud.DX_IMAGE, star=star)
    without = oracle.image_pass(g, sp, pp, lines, th, ph, 3, vel_out, ip, im_size, cloud.DX_IMAGE)
    for a, b in zip(with_star, without):
        np.testing.assert_array_equal(a[~on], b[~on])
    # the star shines through except at optically thick line centres
    assert (with_star[0][on] > without[0][on]).mean() > 0.5


def test_node_pixels_match_reference(oracle, golden_dir):
    # the reference's pix = 0 image mode (para_finish.f90:245-330): the march starts at each
    # bulk node's rotated (x, y); traces with and without the star against the reference
    from portal_amd import driver
    from tests.golden.make_golden import cloud_digest
    z = np.load(os.path.join(golden_dir, "node_pixels.npz"))
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    assert cloud_digest(g) == str(z["digest"])
    N = g.node_num
    vn, bn, xn = (np.zeros((3, N), order="F") for _ in range(3))
    xs = np.zeros(3)
    oracle.lib.orc_rotate_all_save(N, C.c_double(float(z["th_inc"])), C.c_double(float(z["ph_az"])),
                                   _d(_f64(g.vel_all)), _d(_f64(g.bvec_all)), _d(_f64(g.XYZ)), _d(vn), _d(bn), _d(xn),
                                   _d(_f64(z["X_star"])), _d(xs))
    nodes = z["nodes"]
    np.testing.assert_array_equal(xn[:2, nodes - 1], z["xy"])
    ptr, idx = _f64(g.nbr_ptr).astype(np.int32), np.asarray(g.nbr_idx, dtype=np.int32)
    Rs = float(z["R_star"])
    nu = sp.e_levs_hz[1] - sp.e_levs_hz[0]
    for k, n in enumerate(nodes):
        cells = np.zeros(4096, dtype=np.int32)
        steps = np.zeros(4096, dtype=np.int32)
        nc, hit = C.c_int(), C.c_int()
        rc = oracle.lib.orc_trace_out_raytrace(C.c_double(xn[0, n - 1]), C.c_double(xn[1, n - 1]), N, _d(xn),
                                               ptr.ctypes.data_as(C.POINTER(C.c_int)),
                                               idx.ctypes.data_as(C.POINTER(C.c_int)), C.c_double(cloud.DX_IMAGE), 4096,
                                               k % 2, C.c_double(Rs), _d(xs), cells.ctypes.data_as(C.POINTER(C.c_int)),
                                               steps.ctypes.data_as(C.POINTER(C.c_int)), C.byref(nc), C.byref(hit))
        assert rc == 0
        ref = z["trace"][:, z["trace_ofs"][k]:z["trace_ofs"][k + 1]]
        assert nc.value == ref.shape[1]
        np.testing.assert_array_equal(cells[:nc.value], ref[0])
        np.testing.assert_array_equal(steps[:nc.value], ref[1])
        assert z["I_in"][k] == pytest.approx(orc_black_body(oracle, float(z["T_star"]) if hit.value else 2.73, nu),
                                             rel=1e-14)
    # the pass in this mode: one output row per bulk node, the first one's trace started at
    # node 1's rotated position; with no line emission every row is 0
    lines = driver.image_lines(sp, [1], [2])
    vel_out = driver.channel_velocities(2, 300.0)
    zero = np.zeros((N, 2 * sp.e_num), order="F")
    out = oracle.image_pass(g, sp, zero, lines, float(z["th_inc"]), float(z["ph_az"]), 2, vel_out, 0, 0.0,
                            cloud.DX_IMAGE, node_pixels=True)
    assert all(o.shape == (g.bulk_node_num, 1, 5) for o in out)
    assert not any(o.any() for o in out)


def orc_black_body(oracle, T, nu):
    oracle.lib.orc_black_body.restype = C.c_double
    return oracle.lib.orc_black_body(C.c_double(T), C.c_double(nu))
