"""The configurations other than the bench line, at their full sizes, against the oracle
on seeded node subsets (needs an MI355X):

* the 4,000-node cloud's full image, 129 x 129 pixels x 65 channels, in FP64;
* a batch of clouds spanning 5,000 to 100,000 nodes and 4 to 20 levels: FP64 batch
  iterations against the oracle's SE pass on 16 consecutive nodes of every model, a
  continued call against the oracle from the batch's own populations, and the mixed
  batch against the FP64 one;
* the 10^6-node cloud through the surrogate path: first-pass FP64 J of 32 seeded nodes
  against the oracle, then the mixed surrogate run against the plain mixed run (fewer
  iterations, the same fixed point, every node normalised, the fallback count) and the
  Stokes image of each;
* the 200,000-node, 10-level cloud's 513 x 513 x 129 image: FP64 against the oracle on
  seeded pixel bands, mixed against FP64 on every pixel;
* a model read from LIME/LAMDA/dust-table files through the FP64 SE pass;
* the 64-cloud batch at its full size, converged, with property checks and an oracle
  fixed-point check on two of its models.
"""
import numpy as np
import pytest

from portal_amd import cloud, driver, surrogate
from portal_amd.api import Portal, batch_lambda_iterate
from portal_amd.lib import PORTAL_FP64, PORTAL_MIXED

pytestmark = pytest.mark.gpu

NTH = 16  # oracle threads on the GPU host


def solved(g):
    return np.where(g.ab_all[: g.bulk_node_num] > 0)[0]


def assert_pol_close(pp, ref, idx, L, rtol):
    """rho^0_0 per node and level to rtol, rho^2_0 to rtol of that level's rho^0_0 (of 1e-6
    of the node's largest rho^0_0 for levels holding less)."""
    node_max = np.max(np.abs(ref[idx][:, 0::2]), axis=1)
    for i in range(L):
        o0 = ref[idx, 2 * i]
        np.testing.assert_allclose(pp[idx, 2 * i], o0, rtol=rtol, atol=0)
        scale = np.maximum(np.abs(o0), 1e-6 * node_max)
        d2 = np.abs(pp[idx, 2 * i + 1] - ref[idx, 2 * i + 1])
        assert np.all(d2 <= rtol * scale), f"level {i}: max |d rho20| / rho00 = {np.max(d2 / scale):.2e}"


def test_small_cloud_full_image_fp64(oracle):
    """The 4,000-node cloud: SE pass, then the whole 129^2 x 65 cube of the J=1-0 line."""
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 4000, seed=1)
    xth = cloud.legendre_all(6, 11)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pp, _ = pt.se_pass(xth, 32, cloud.DX_INWARD)
    rp = driver.rescale_pol_pops(pp, g.lev_p, sp.lev_g)
    lines = driver.image_lines(sp, [1], [2])
    vel_out = driver.channel_velocities(32, 93.75)
    (I, Q, U, T), st = pt.image_pass(rp, lines, np.pi / 4, 0.0, 32, vel_out, 64, 0.7, cloud.DX_IMAGE,
                                     n_depth_rt=4096)
    pt.close()
    assert I.shape == (129 * 129, 1, 65)
    Io, Qo, Uo, To = oracle.image_pass(g, sp, rp, lines, np.pi / 4, 0.0, 32, vel_out, 64, 0.7, cloud.DX_IMAGE,
                                       n_depth_rt=4096, nthreads=NTH)
    Imax, Qmax, Umax = np.abs(Io).max(), np.abs(Qo).max(), np.abs(Uo).max()
    print(f"129^2 x 65 image: max |Q|/max |I| {Qmax / Imax:.2e}, max |U|/max |I| {Umax / Imax:.2e}; "
          f"errors I {np.abs(I - Io).max() / Imax:.1e} Q {np.abs(Q - Qo).max() / Qmax:.1e} "
          f"U {np.abs(U - Uo).max() / Umax:.1e} (own scales)")
    np.testing.assert_allclose(I, Io, rtol=1e-10, atol=1e-10 * Imax)
    np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-10 * Qmax)
    np.testing.assert_allclose(U, Uo, rtol=0, atol=1e-10 * Umax)
    np.testing.assert_allclose(T, To, rtol=1e-10, atol=1e-12)
    cube = driver.cube(I, 0, 64, nu0=lines[0][2])
    assert cube.shape == (129, 129, 65)


BATCH = [(5000, 4, -1.1), (20000, 9, -1.9), (60000, 14, -1.4), (100000, 20, -1.6)]


@pytest.fixture(scope="module")
def batch_models():
    out = []
    for k, (n, L, n_exp) in enumerate(BATCH):
        sp = cloud.co_like_rotor(L)
        out.append((sp, cloud.make_cloud(sp, n, seed=5100 + k, n_exp=n_exp)))
    return out


def _contexts(models):
    pts = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    return pts


def test_batch_of_clouds_against_the_oracle(oracle, batch_models):
    xth = cloud.legendre_all(6, 11)
    n = len(batch_models)
    pts = _contexts(batch_models)
    kw = dict(tol_pop=0.0, tol_align=0.0, precision=PORTAL_FP64)
    per, tot = batch_lambda_iterate(pts, [xth] * n, [32] * n, cloud.DX_INWARD, 1, rebuild_paths=1, **kw)
    first = [(pt.get_pol_pops(), pt.get_levpop()) for pt in pts]
    per, tot2 = batch_lambda_iterate(pts, [xth] * n, [32] * n, cloud.DX_INWARD, 1, **kw)
    second = [pt.get_pol_pops() for pt in pts]
    print(f"batch: {sum(len(solved(g)) for _, g in batch_models)} solved nodes, rt_variant {tot['rt_variant']}, "
          f"se_form {tot['se_form']}")
    rng = np.random.default_rng(12)
    for m, (sp, g) in enumerate(batch_models):
        lo = int(rng.integers(1, g.bulk_node_num - 20))
        hi = lo + 15
        idx = np.array([i for i in range(lo - 1, hi) if g.ab_all[i] > 0])
        # iteration 1 from the starting populations, iteration 2 from the batch's own first iterate
        ref1 = oracle.se_pass(g, sp, xth, 32, cloud.DX_INWARD, node_lo=lo, node_hi=hi, nthreads=NTH)[0]
        ref2 = oracle.se_pass(g, sp, xth, 32, cloud.DX_INWARD, node_lo=lo, node_hi=hi, nthreads=NTH,
                              lev_p=first[m][1])[0]
        assert_pol_close(first[m][0], ref1, idx, sp.e_num, 1e-10)
        assert_pol_close(second[m], ref2, idx, sp.e_num, 1e-10)
    # the mixed batch (the configuration's precision) against the FP64 one, first iteration
    for pt, (sp, g) in zip(pts, batch_models):
        pt.set_levpop(g.lev_p)
    batch_lambda_iterate(pts, [xth] * n, [32] * n, cloud.DX_INWARD, 1, tol_pop=0.0, tol_align=0.0,
                         precision=PORTAL_MIXED)
    for m, (pt, (sp, g)) in enumerate(zip(pts, batch_models)):
        pm = pt.get_pol_pops()
        idx = solved(g)
        big = first[m][0][idx][:, 0::2] > 1e-6 * first[m][0][idx][:, 0::2].max(axis=1, keepdims=True)
        e = np.abs(pm[idx][:, 0::2] / first[m][0][idx][:, 0::2] - 1)[big].max()
        print(f"model {m} ({g.node_num} nodes, {sp.e_num} levels): mixed vs FP64 max rel rho00 {e:.2e}")
        assert e < 1e-5
        pt.close()


@pytest.fixture(scope="module")
def million():
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 1_000_000, seed=5)
    return sp, g, cloud.legendre_all(6, 11)


def test_million_node_first_pass_against_the_oracle(oracle, million):
    sp, g, xth = million
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    J0, J2, st = pt.inverse_rt(xth, 32, cloud.DX_INWARD)
    pt.close()
    rng = np.random.default_rng(31)
    sub = np.sort(rng.choice(solved(g), 32, replace=False)) + 1
    J0o, J2o = oracle.inverse_nodes(g, sp, xth, 32, cloud.DX_INWARD, sub, nthreads=NTH)
    print(f"10^6 nodes: {st['n_cross']} crossings per pass, longest path {st['max_path']} cells")
    np.testing.assert_allclose(J0[:, sub - 1], J0o, rtol=1e-10, atol=0)
    for k in range(32):
        assert np.abs(J2[:, sub[k] - 1] - J2o[:, k]).max() <= 1e-10 * np.abs(J0o[:, k]).max()


def test_million_node_surrogate_run(million):
    sp, g, xth = million
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    if not surrogate.load_into(pt, sp):
        pt.close()
        pytest.skip("no trained weights for this species")
    kw = dict(tol_pop=1e-6, tol_align=1e-8, precision=PORTAL_MIXED, ng_accel=1, rebuild_paths=1)
    pt.set_levpop(g.lev_p)
    pp0, lev0, st0 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=0, **kw)
    lines = driver.image_lines(sp, [1], [2])
    vel_out = driver.channel_velocities(32, 93.75)
    img = lambda pp, lev: pt.image_pass(driver.rescale_pol_pops(pp, lev, sp.lev_g), lines, np.pi / 4, 0.0, 32,  # noqa
                                        vel_out, 64, 0.7, cloud.DX_IMAGE, n_depth_rt=4096,
                                        precision=PORTAL_MIXED)[0]
    I0, Q0, U0, _ = img(pp0, lev0)
    pt.set_levpop(g.lev_p)
    pp1, lev1, st1 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=1, fallback_tol=1e-8, **kw)
    I1, Q1, U1, _ = img(pp1, lev1)
    pt.close()
    idx = solved(g)
    print(f"10^6 nodes, mixed + Ng: plain {st0['iterations']} iterations; surrogate {st1['iterations']} "
          f"iterations, {st1['fallbacks']} fallbacks of {len(idx)} nodes")
    assert st0["converged"] and st1["converged"]
    assert st1["iterations"] < st0["iterations"]
    assert st1["fallbacks"] < 0.01 * len(idx)
    # every solved node normalised: sum_i sqrt(2J_i+1) rho^0_0(i) = 1 (em_coefs.f90:158-164)
    sg = np.sqrt(2.0 * sp.j_lev + 1.0)
    norm = pp1[idx][:, 0::2] @ sg
    assert np.abs(norm - 1).max() < 1e-12
    # the same fixed point
    big = lev0[:, idx] > 1e-6 * lev0[:, idx].max(axis=0)
    rel = np.abs(lev1[:, idx] - lev0[:, idx]) / lev0[:, idx]
    jl = [i for i in range(sp.e_num) if sp.j_lev[i] > 0]
    a0 = pp0[idx][:, [2 * i + 1 for i in jl]] / pp0[idx][:, [2 * i for i in jl]]
    a1 = pp1[idx][:, [2 * i + 1 for i in jl]] / pp1[idx][:, [2 * i for i in jl]]
    da = np.abs(a1 - a0)[big[jl].T]
    s = np.abs(I0).max()
    e = [np.abs(I1 - I0).max() / s, np.abs(Q1 - Q0).max() / s, np.abs(U1 - U0).max() / s]
    print(f"same fixed point: population max rel {rel[big].max():.2e}, alignment max abs {da.max():.2e}; "
          f"Stokes I/Q/U max error {e[0]:.1e} {e[1]:.1e} {e[2]:.1e} of max |I|")
    assert rel[big].max() < 5e-5
    assert da.max() < 5e-6
    assert max(e) < 1e-4


def test_production_cloud_image_against_the_oracle(oracle):
    """The 200,000-node, 10-level cloud's 513 x 513 x 129 image of J=1-0 after one FP64 SE
    pass: the FP64 cube against the oracle on three seeded bands of 171 pixels (every
    channel, Q and U on their own scales), and the mixed cube against the FP64 one."""
    sp = cloud.co_like_rotor(10)
    g = cloud.make_cloud(sp, 200_000, seed=2024)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pp, _ = pt.se_pass(cloud.legendre_all(9, 32), 64, cloud.DX_INWARD)
    rp = driver.rescale_pol_pops(pp, g.lev_p, sp.lev_g)
    lines = driver.image_lines(sp, [1], [2])
    vel_out = driver.channel_velocities(64, 46.875)
    im_pix, n_side = 256, 513
    args = (rp, lines, np.pi / 4, 0.0, 64, vel_out, im_pix, 0.7, cloud.DX_IMAGE)
    (I, Q, U, T), _ = pt.image_pass(*args, n_depth_rt=2048, precision=PORTAL_FP64)
    (Im, Qm, Um, Tm), _ = pt.image_pass(*args, n_depth_rt=2048, precision=PORTAL_MIXED)
    pt.close()
    assert I.shape == (n_side * n_side, 1, 129)
    rng = np.random.default_rng(43)
    for ii in (int(x) for x in rng.integers(-200, 200, 3)):
        lo = (ii + im_pix) * n_side + int(rng.integers(0, n_side - 171)) + 1
        hi = lo + 170
        Io, Qo, Uo, To = oracle.image_pass(g, sp, rp, lines, np.pi / 4, 0.0, 64, vel_out, im_pix, 0.7,
                                           cloud.DX_IMAGE, n_depth_rt=2048, pix_lo=lo, pix_hi=hi, nthreads=NTH)
        s = slice(lo - 1, hi)
        Imax, Qmax, Umax = np.abs(Io[s]).max(), np.abs(Qo[s]).max(), np.abs(Uo[s]).max()
        print(f"row {ii:+d}, pixels {lo}..{hi}: max |Q|/max |I| {Qmax / Imax:.1e}; errors I "
              f"{np.abs(I[s] - Io[s]).max() / Imax:.1e} Q {np.abs(Q[s] - Qo[s]).max() / Qmax:.1e} "
              f"U {np.abs(U[s] - Uo[s]).max() / Umax:.1e} (own scales)")
        assert Imax > 0 and Qmax > 0 and Umax > 0
        np.testing.assert_allclose(I[s], Io[s], rtol=1e-10, atol=1e-10 * Imax)
        np.testing.assert_allclose(Q[s], Qo[s], rtol=0, atol=1e-10 * Qmax)
        np.testing.assert_allclose(U[s], Uo[s], rtol=0, atol=1e-10 * Umax)
        np.testing.assert_allclose(T[s], To[s], rtol=1e-10, atol=1e-12)
    e = [np.abs(Im - I).max() / np.abs(I).max(), np.abs(Qm - Q).max() / np.abs(Q).max(),
         np.abs(Um - U).max() / np.abs(U).max()]
    print(f"mixed 513^2 x 129 cube: max |dI|/max |I| {e[0]:.1e}, |dQ|/max |Q| {e[1]:.1e}, |dU|/max |U| {e[2]:.1e}")
    assert max(e) < 1e-4


def test_lime_model_files_through_the_c_abi(oracle):
    """A model read from LIME/LAMDA/dust-table files (tests/golden/lime_model/, read as
    the reference's eat_lime reads them: dust on half of the nodes, two collision
    partners) through the FP64 SE pass, against the oracle on the same arrays."""
    import json
    import os
    from portal_amd import lime_io
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    files = {k: os.path.join(gold, v) for k, v in json.loads(str(np.load(os.path.join(gold, "eat_lime.npz"))["files"])).items()}
    g, sp, _ = lime_io.read_model(files["pop"], files["grid"], files["mol"], files["dust"],
                                  bvec=np.array([0.36, -0.48, 0.8]), gvec=np.array([0.6, 0.8, 0.0]))
    xth = cloud.legendre_all(4, 8)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pp, st = pt.se_pass(xth, 16, cloud.DX_INWARD)
    pt.close()
    ref = oracle.se_pass(g, sp, xth, 16, cloud.DX_INWARD, nthreads=NTH)[0]
    idx = solved(g)
    assert (g.dust_p[idx] > 0).sum() > 50 and sp.c_par == 2
    assert_pol_close(pp, ref, idx, sp.e_num, 1e-10)


def test_the_64_cloud_batch_at_full_size(oracle):
    """The batch configuration at its full size: 64 seeded clouds (about 2e6 nodes,
    log-uniform in [5e3, 1e5], 4..20-level rotors), mixed precision with Ng, to 1e-6 /
    1e-8.  Every model converges; every solved node is normalised with finite, positive
    populations; and on 8 nodes of the smallest and of a 20-level model the converged
    state is a fixed point of the oracle's FP64 SE pass to the mixed path's accuracy."""
    from portal_amd import workloads
    models = workloads.build_batch()
    xth = cloud.legendre_all(6, 11)
    pts = [pt for _, _, pt in models]
    n = len(pts)
    done, conv_at = 0, [0] * n
    while done < 400:
        per, tot = batch_lambda_iterate(pts, [xth] * n, [32] * n, cloud.DX_INWARD, 10, tol_pop=1e-6, tol_align=1e-8,
                                        precision=PORTAL_MIXED, rebuild_paths=1 if done == 0 else 0, ng_accel=1)
        for m, p in enumerate(per):
            if not conv_at[m] and p["converged"]:
                conv_at[m] = done + p["iterations"]
        done += tot["iterations"]
        print(f"64-cloud batch: {done} iterations, {sum(1 for v in conv_at if v)} of {n} converged", flush=True)
        if tot["converged"]:
            break
    print(f"iterations per model: min {min(conv_at)}, max {max(conv_at)}")
    assert all(conv_at)
    for sp, g, pt in models:
        idx = solved(g)
        pp, lev = pt.get_pol_pops(), pt.get_levpop()
        w = np.sqrt(2 * sp.j_lev + 1.0)
        np.testing.assert_allclose(pp[idx][:, 0::2] @ w, 1.0, atol=1e-12)
        assert np.all(np.isfinite(lev[:, idx])) and np.all(lev[:, idx] > 0)
        # alignment within what any sublevel populations allow (|rho20/rho00| <= sqrt 2 for
        # J = 1), on levels holding more than 1e-10 of their node
        assert np.all(np.isfinite(pp[idx]))
        r0, r2 = pp[idx][:, 2::2], pp[idx][:, 3::2]
        held = r0 * w[None, 1:] > 1e-10
        assert np.all(np.abs(r2[held]) < 2.0 * r0[held])
    smallest = min(range(n), key=lambda m: models[m][1].node_num)
    big = [m for m in range(n) if models[m][0].e_num == 20]
    for m in [smallest] + big[:1]:
        sp, g, pt = models[m]
        lo = int(solved(g)[len(solved(g)) // 2]) + 1
        ref = oracle.se_pass(g, sp, xth, 32, cloud.DX_INWARD, node_lo=lo, node_hi=lo + 7, nthreads=NTH,
                             lev_p=pt.get_levpop())[0]
        idx = np.array([i for i in range(lo - 1, lo + 7) if g.ab_all[i] > 0])
        got = pt.get_pol_pops()
        big_lv = ref[idx][:, 0::2] > 1e-6 * ref[idx][:, 0::2].max(axis=1, keepdims=True)
        e = np.abs(got[idx][:, 0::2] / ref[idx][:, 0::2] - 1)[big_lv].max()
        print(f"model {m} ({g.node_num} nodes, {sp.e_num} levels): fixed point vs oracle, max rel rho00 {e:.1e}")
        assert e < 1e-5
    for pt in pts:
        pt.close()
