"""Batched persistent lambda iteration over independent models (needs an MI355X)."""
import numpy as np
import pytest

from portal_amd import cloud
from portal_amd.api import Portal, batch_lambda_iterate
from portal_amd.lib import PORTAL_FP64, PORTAL_MIXED

pytestmark = pytest.mark.gpu


def _models():
    specs = [(600, 4, 31), (1500, 6, 32), (900, 9, 33)]
    out = []
    for n, L, seed in specs:
        sp = cloud.co_like_rotor(L)
        g = cloud.make_cloud(sp, n, seed=seed, n_exp=-1.2 - 0.1 * L / 4)
        out.append((sp, g))
    return out


@pytest.mark.parametrize("precision,tol", [(PORTAL_FP64, 1e-10), (PORTAL_MIXED, 1e-5)])
def test_batch_equals_single_model_iterations(precision, tol):
    models = _models()
    xth = cloud.legendre_all(4, 8)
    iters = 4
    singles = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pp, lev, st = pt.lambda_iterate(xth, 8, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0,
                                        precision=precision)
        singles.append((pp, lev))
        pt.close()
    pts = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    per, tot = batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0,
                                    precision=precision)
    assert tot["iterations"] == iters
    for (sp, g), pt, (pp, lev) in zip(models, pts, singles):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        lb = pt.get_levpop()
        np.testing.assert_allclose(lb[:, idx], lev[:, idx], rtol=tol)
        pb = pt.get_pol_pops()
        np.testing.assert_allclose(pb[idx][:, 0::2], pp[idx][:, 0::2], rtol=tol)
        pt.close()


def test_batch_with_depolarisation_equals_single_models():
    """Models with collisional depolarisation switched on (two of three; the third keeps
    the reference's D = 0) iterate in the batch as they do alone (FP64, 1e-10)."""
    models = _models()
    xth = cloud.legendre_all(4, 8)

    def ctx(sp, g, m):
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        if m != 1:
            pt.set_depol(30.0 * pt.depol_table())
        return pt

    singles = []
    for m, (sp, g) in enumerate(models):
        pt = ctx(sp, g, m)
        singles.append(pt.lambda_iterate(xth, 8, cloud.DX_INWARD, 3, tol_pop=0.0, tol_align=0.0,
                                         precision=PORTAL_FP64)[:2])
        pt.close()
    pts = [ctx(sp, g, m) for m, (sp, g) in enumerate(models)]
    batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, 3, tol_pop=0.0, tol_align=0.0,
                         precision=PORTAL_FP64)
    for (sp, g), pt, (pp, lev) in zip(models, pts, singles):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=1e-10)
        pb = pt.get_pol_pops()
        np.testing.assert_allclose(pb[idx][:, 0::2], pp[idx][:, 0::2], rtol=1e-10)
        scale = np.repeat(np.abs(pp[idx][:, 0::2]), 1, axis=1)
        assert np.all(np.abs(pb[idx][:, 1::2] - pp[idx][:, 1::2]) <= 1e-10 * scale + 1e-300)
        pt.close()


def test_batch_converges_each_model():
    models = _models()
    xth = cloud.legendre_all(4, 8)
    pts = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    per, tot = batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, 80, tol_pop=1e-5, tol_align=1e-7)
    print(per, tot)
    assert all(p["iterations"] <= tot["iterations"] for p in per)
    assert all(p["max_rel_change"] < 1e-3 for p in per)
    for pt in pts:
        pt.close()


@pytest.mark.parametrize("ng,se_form", [(0, -1), (1, -1), (1, 0)])
def test_batch_models_stop_where_their_single_runs_stop(ng, se_form):
    # a model leaves the batch's iteration once it meets the tolerance: its iteration
    # count, level densities and SE solution equal those of its single-model run to the
    # same tolerance, with Ng extrapolation on the same schedule or without, and with the
    # SE solves in the timed form or in the one persistent launch over all sizes
    models = _models()
    xth = cloud.legendre_all(4, 8)
    kw = dict(tol_pop=1e-6, tol_align=1e-8, precision=PORTAL_FP64, ng_accel=ng)
    singles = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pp, lev, st = pt.lambda_iterate(xth, 8, cloud.DX_INWARD, 200, **kw)
        assert st["converged"]
        singles.append((pp, lev, st["iterations"]))
        pt.close()
    pts = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    if se_form >= 0:
        pts[0].force_se(se_form, 0)
    per, tot = batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, 200, **kw)
    if se_form >= 0:
        assert tot["se_form"] == se_form
    its = [s[2] for s in singles]
    print(f"ng={ng}: single-model iterations {its}, batch {[p['iterations'] for p in per]}")
    assert tot["converged"] and tot["iterations"] == max(its)
    assert len(set(its)) > 1  # the models leave at different iterations
    for (sp, g), pt, p, (pp, lev, n_it) in zip(models, pts, per, singles):
        assert p["converged"] and p["iterations"] == n_it
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=1e-10)
        pb = pt.get_pol_pops()
        np.testing.assert_allclose(pb[idx][:, 0::2], pp[idx][:, 0::2], rtol=1e-10)
        scale = np.abs(pp[idx][:, 0::2]).max(axis=1, keepdims=True)
        assert np.all(np.abs(pb[idx][:, 1::2] - pp[idx][:, 1::2]) <= 1e-10 * scale + 1e-300)
        pt.close()


def test_batch_continued_calls_are_one_iteration():
    # 4 + 3 batch iterations in two calls equal 7 in one (the second call continues each
    # model's iteration), and the iteration at which each model meets the tolerance does
    # not depend on where the calls split
    models = _models()
    xth = cloud.legendre_all(4, 8)

    def fresh():
        pts = []
        for sp, g in models:
            pt = Portal(0)
            pt.upload_grid(g)
            pt.upload_species(sp)
            pt.set_levpop(g.lev_p)
            pts.append(pt)
        return pts

    a = fresh()
    # one inward form for both (a timed choice may differ between the two sets of contexts
    # in mixed precision's last bits)
    a[0].force_batch_rt(3)
    batch_lambda_iterate(a, [xth] * 3, [8] * 3, cloud.DX_INWARD, 7, tol_pop=0.0, tol_align=0.0)
    b = fresh()
    b[0].force_batch_rt(3)
    batch_lambda_iterate(b, [xth] * 3, [8] * 3, cloud.DX_INWARD, 4, tol_pop=0.0, tol_align=0.0)
    batch_lambda_iterate(b, [xth] * 3, [8] * 3, cloud.DX_INWARD, 3, tol_pop=0.0, tol_align=0.0)
    for pa, pb in zip(a, b):
        np.testing.assert_array_equal(pa.get_levpop(), pb.get_levpop())
    kw = dict(tol_pop=1e-5, tol_align=1e-7)
    c = fresh()
    per, tot = batch_lambda_iterate(c, [xth] * 3, [8] * 3, cloud.DX_INWARD, 200, **kw)
    assert tot["converged"]
    d = fresh()
    at, done = [0, 0, 0], 0
    while True:
        per2, tot2 = batch_lambda_iterate(d, [xth] * 3, [8] * 3, cloud.DX_INWARD, 5, **kw)
        for m, p in enumerate(per2):
            if not at[m] and p["converged"]:
                at[m] = done + p["iterations"]
        done += tot2["iterations"]
        if tot2["converged"] or done > 200:
            break
    assert at == [p["iterations"] for p in per]
    for pt in a + b + c + d:
        pt.close()


@pytest.mark.parametrize("precision,tol,bound", [(PORTAL_FP64, 1e-7, 2e-5), (PORTAL_MIXED, 1e-6, 5e-5)])
def test_batch_ng_reaches_the_plain_fixed_point(precision, tol, bound):
    # Ng acceleration per model in the batch: fewer iterations, same fixed point.  Mixed
    # precision at the production tolerances: an alignment change of 1e-9 (tol_pop 1e-7)
    # is about two FP32 roundings of rho20/rho00 ~ 1e-2, below what the FP32 transfer
    # resolves from one iteration to the next
    models = _models()
    xth = cloud.legendre_all(4, 8)
    kw = dict(tol_pop=tol, tol_align=tol * 1e-2, precision=precision)
    out = []
    for ng in (0, 1):
        pts = []
        for sp, g in models:
            pt = Portal(0)
            pt.upload_grid(g)
            pt.upload_species(sp)
            pt.set_levpop(g.lev_p)
            pts.append(pt)
        per, tot = batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, 300, ng_accel=ng, **kw)
        assert tot["converged"]
        out.append((tot, [pt.get_levpop() for pt in pts]))
        for pt in pts:
            pt.close()
    (t0, lev0), (t1, lev1) = out
    print(f"batch: plain {t0['iterations']} iterations, Ng {t1['iterations']} ({t1['ng_steps']} extrapolations)")
    assert t1["ng_steps"] > 0 and t1["iterations"] < t0["iterations"]
    for (sp, g), a, b in zip(models, lev0, lev1):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        big = a[:, idx] > 1e-8 * a[:, idx].max(axis=0)
        assert (np.abs(b[:, idx] - a[:, idx])[big] / a[:, idx][big]).max() < bound


def _open(models):
    pts = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    return pts


def test_se_waves_per_system_give_the_same_bits():
    # one, two and four waves per SE system in the single-model solve; the 20-level
    # rotor's 74-dimensional system gives some lanes two rows and the waves a column
    # split that is not even
    sp = cloud.co_like_rotor(20)
    g = cloud.make_cloud(sp, 500, seed=44, n_exp=-1.5)
    xth = cloud.legendre_all(4, 8)
    out = []
    for waves in (1, 2, 4):
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pt.force_se(-1, waves)
        pp, lev, st = pt.lambda_iterate(xth, 8, cloud.DX_INWARD, 3, tol_pop=0.0, tol_align=0.0,
                                        precision=PORTAL_FP64)
        assert st["se_waves"] == waves and st["se_form"] == 0
        out.append((pp, lev))
        pt.close()
    assert np.isfinite(out[0][0]).all() and np.abs(out[0][0][:, 0]).max() > 0
    for pp, lev in out[1:]:
        assert np.array_equal(pp, out[0][0])
        assert np.array_equal(lev, out[0][1])


@pytest.mark.parametrize("precision", [PORTAL_FP64, PORTAL_MIXED])
def test_batch_se_forms_give_the_same_bits(precision):
    # the one persistent SE launch and the launch per class at 1, 2 and 4 waves per
    # system (and whatever the autotuner picks) solve each model's systems identically
    specs = [(500, 4, 41), (700, 12, 42), (600, 20, 43)]
    models = []
    for n, L, seed in specs:
        sp = cloud.co_like_rotor(L)
        models.append((sp, cloud.make_cloud(sp, n, seed=seed, n_exp=-1.4)))
    xth = cloud.legendre_all(4, 8)
    runs = []
    for form, waves in [(0, 0), (1, 1), (1, 2), (1, 4), (-1, 0)]:
        pts = _open(models)
        pts[0].force_se(form, waves)
        pts[0].force_batch_rt(3)  # one inward form for every run: the SE forms are compared
        per, tot = batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, 3, tol_pop=0.0, tol_align=0.0,
                                        precision=precision)
        print(form, waves, tot["se_form"], tot["se_waves"])
        assert tot["iterations"] == 3
        if form >= 0:
            assert tot["se_form"] == form
        if form == 1 and waves:
            assert tot["se_waves"] == waves
        runs.append([(pt.get_pol_pops(), pt.get_levpop()) for pt in pts])
        for pt in pts:
            pt.close()
    for r in runs[1:]:
        for (pa, la), (pb, lb) in zip(runs[0], r):
            assert np.array_equal(pa, pb)
            assert np.array_equal(la, lb)


@pytest.mark.parametrize("precision,tol", [(PORTAL_FP64, 1e-10), (PORTAL_MIXED, 1e-5)])
def test_batch_with_dust_and_star_equals_single_models(precision, tol):
    # a model with dust, one in star mode and a plain one in one batch: every model's
    # iterations equal its single-model iterations (dust and the star's rays take the
    # cooperative kernel's dust variant and per-node direction sets)
    sp1 = cloud.co_like_rotor(4)
    g1 = cloud.make_cloud(sp1, 1200, seed=51)
    nb = g1.bulk_node_num - 1
    g1.dust_p[:nb] = 1e-17 * g1.h2_dens[:nb] / 1e12
    sp1.kd_v = np.array([3000.0, 6000.0, 9000.0])
    sp2 = cloud.co_like_rotor(5)
    g2 = cloud.make_cloud(sp2, 1500, seed=52)
    star = dict(nth=6, nph=11, R_star=0.15, T_star=6000.0, X_star=np.zeros(3))
    r = np.linalg.norm(g2.XYZ, axis=0)
    g2.ab_all[np.where(r[: g2.bulk_node_num] < 1.3 * star["R_star"])[0]] = 0.0
    sp3 = cloud.co_like_rotor(6)
    g3 = cloud.make_cloud(sp3, 900, seed=53)
    models = [(sp1, g1, None), (sp2, g2, star), (sp3, g3, None)]
    xth = cloud.legendre_all(6, 11)
    iters = 3

    def open_model(sp, g, st):
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pt.star = st
        return pt

    singles = []
    for sp, g, st in models:
        pt = open_model(sp, g, st)
        pp, lev, _ = pt.lambda_iterate(xth, 8, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0,
                                       precision=precision)
        singles.append((pp, lev))
        pt.close()
    pts = [open_model(*m) for m in models]
    per, tot = batch_lambda_iterate(pts, [xth] * 3, [8] * 3, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0,
                                    precision=precision)
    assert tot["iterations"] == iters and tot["rt_variant"] in (2, 3)
    for (sp, g, _), pt, (pp, lev) in zip(models, pts, singles):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=tol)
        np.testing.assert_allclose(pt.get_pol_pops()[idx][:, 0::2], pp[idx][:, 0::2], rtol=tol)
        pt.close()
    # the dust and the star matter: the dust model without dust iterates differently
    sp1.kd_v = np.zeros(3)
    pt = open_model(sp1, g1, None)
    _, lev0, _ = pt.lambda_iterate(xth, 8, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0, precision=precision)
    idx = np.where(g1.ab_all[: g1.bulk_node_num] > 0)[0]
    assert np.abs(lev0[:, idx] / singles[0][1][:, idx] - 1).max() > 1e-4
    pt.close()


@pytest.mark.parametrize("to_tol", [False, True])
def test_batch_staged_form_when_a_model_does_not_fit_the_cooperative_kernel(to_tol):
    """A 13-level rotor with 255 channels has 12 x 255 pairs, more than the cooperative
    kernel's 256 lanes x 12 slots hold, so the whole batch takes the staged persistent
    form (rt_variant 1), whose counters the batch clears in four ints; in FP64 each model
    equals its single-model iterations to 1e-10, over three iterations or to a tolerance
    (where the models leave the staged form's work list at different iterations)."""
    specs = [(500, 13, 41, 127), (700, 4, 42, 127)]
    models = []
    for n, L, seed, nvel in specs:
        sp = cloud.co_like_rotor(L)
        models.append((sp, cloud.make_cloud(sp, n, seed=seed), nvel))
    xth = cloud.legendre_all(3, 6)
    iters, kw = (200, dict(tol_pop=1e-5, tol_align=1e-7)) if to_tol else (3, dict(tol_pop=0.0, tol_align=0.0))
    singles, its = [], []
    for sp, g, nvel in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pp, lev, st = pt.lambda_iterate(xth, nvel, cloud.DX_INWARD, iters, precision=PORTAL_FP64, **kw)
        singles.append((pp, lev))
        its.append(st["iterations"])
        pt.close()
    pts = []
    for sp, g, nvel in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    per, tot = batch_lambda_iterate(pts, [xth] * 2, [m[2] for m in models], cloud.DX_INWARD, iters,
                                    precision=PORTAL_FP64, **kw)
    print(tot, "single-model iterations", its, "batch", [p["iterations"] for p in per])
    assert tot["rt_variant"] == 1 and tot["iterations"] == max(its)
    assert [p["iterations"] for p in per] == its
    if to_tol:
        assert tot["converged"] and len(set(its)) > 1
    for (sp, g, nvel), pt, (pp, lev) in zip(models, pts, singles):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=1e-10)
        pb = pt.get_pol_pops()
        np.testing.assert_allclose(pb[idx][:, 0::2], pp[idx][:, 0::2], rtol=1e-10)
        pt.close()


def test_batch_surrogate_equals_single_models():
    """The surrogate iterate in a batch: models whose context holds a network (the 4- and
    10-level rotors) get the prediction after the first replay and the residual test
    after the second, as they do alone; a model without a network (a 6-level rotor with
    another dipole moment, for which none is trained) iterates plainly.  FP64, four
    iterations, against single-model runs to 1e-10, with the same fallbacks."""
    from portal_amd import surrogate
    specs = [(3000, 4, 41, 0.11), (2000, 10, 42, 0.11), (1500, 6, 43, 0.12)]
    models = []
    for n, L, seed, mu in specs:
        sp = cloud.co_like_rotor(L, dipole_debye=mu)
        models.append((sp, cloud.make_cloud(sp, n, seed=seed)))
    xth = cloud.legendre_all(4, 8)

    def ctx(sp, g):
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        return pt, surrogate.load_into(pt, sp)

    singles, has = [], []
    for sp, g in models:
        pt, h = ctx(sp, g)
        has.append(h)
        pp, lev, st = pt.lambda_iterate(xth, 16, cloud.DX_INWARD, 4, tol_pop=0.0, tol_align=0.0,
                                        precision=PORTAL_FP64, surrogate=1 if h else 0)
        singles.append((pp, lev, st["fallbacks"], pt.fallback_map() if h else None))
        pt.close()
    if not (has[0] and has[1]):
        pytest.skip("no trained weights for the 4- and 10-level rotors")
    assert not has[2]
    pts = [ctx(sp, g)[0] for sp, g in models]
    per, tot = batch_lambda_iterate(pts, [xth] * 3, [16] * 3, cloud.DX_INWARD, 4, tol_pop=0.0, tol_align=0.0,
                                    precision=PORTAL_FP64, surrogate=1)
    print("fallbacks per model:", [p["fallbacks"] for p in per], "single:", [s[2] for s in singles])
    for m, ((sp, g), pt, (pp, lev, nfb, fmap)) in enumerate(zip(models, pts, singles)):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=1e-10)
        pb = pt.get_pol_pops()
        np.testing.assert_allclose(pb[idx][:, 0::2], pp[idx][:, 0::2], rtol=1e-10)
        assert per[m]["fallbacks"] == nfb
        if fmap is not None:
            np.testing.assert_array_equal(pt.fallback_map(), fmap)
        pt.close()
    assert tot["fallbacks"] == sum(s[2] for s in singles)


def test_batch_with_large_systems_equals_single_models():
    """A batch holding a model with 158-unknown SE systems (41 levels, kmax = 6: the class
    solved with its factors in the slab) beside a 4-level model iterates as they do alone
    (FP64, 1e-10)."""
    specs = [(sp, cloud.make_cloud(sp, n, seed=seed)) for sp, n, seed in
             ((cloud.co_like_rotor(41, kmax=6), 800, 51), (cloud.co_like_rotor(4), 600, 52))]
    xth = cloud.legendre_all(3, 6)

    def ctx(sp, g):
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        return pt

    singles = []
    for sp, g in specs:
        pt = ctx(sp, g)
        singles.append(pt.lambda_iterate(xth, 8, cloud.DX_INWARD, 2, tol_pop=0.0, tol_align=0.0,
                                         precision=PORTAL_FP64)[:2])
        pt.close()
    pts = [ctx(sp, g) for sp, g in specs]
    per, tot = batch_lambda_iterate(pts, [xth] * 2, [8] * 2, cloud.DX_INWARD, 2, tol_pop=0.0, tol_align=0.0,
                                    precision=PORTAL_FP64)
    assert tot["se_form"] == 1
    for (sp, g), pt, (pp, lev) in zip(specs, pts, singles):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=1e-10)
        np.testing.assert_allclose(pt.get_pol_pops()[idx][:, 0::2], pp[idx][:, 0::2], rtol=1e-10)
        pt.close()


@pytest.mark.parametrize("form", [0, 1, 2, 3], ids=["persistent", "flat", "persistent-btf", "flat-btf"])
@pytest.mark.parametrize("precision,tol", [(PORTAL_FP64, 1e-10), (PORTAL_MIXED, 1e-5)])
def test_batch_forms_equal_single_models(precision, tol, form):
    """Each cooperative form of the batch's inward integration (persistent workgroups or
    one per item; rays front to back or back to front) at 65 channels, where the 6- and
    10-level models take 2 and 3 channels per lane and the mixed path's profile
    recurrence: three iterations equal the single-model ones."""
    specs = [(700, 4, 41), (1200, 10, 42), (900, 7, 43)]
    models = []
    for n, L, seed in specs:
        sp = cloud.co_like_rotor(L)
        models.append((sp, cloud.make_cloud(sp, n, seed=seed)))
    xth = cloud.legendre_all(4, 8)
    iters = 3
    singles = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pp, lev, _ = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0,
                                       precision=precision)
        singles.append((pp, lev))
        pt.close()
    pts = []
    for sp, g in models:
        pt = Portal(0)
        pt.upload_grid(g)
        pt.upload_species(sp)
        pt.set_levpop(g.lev_p)
        pts.append(pt)
    pts[0].force_batch_rt(form)
    per, tot = batch_lambda_iterate(pts, [xth] * 3, [32] * 3, cloud.DX_INWARD, iters, tol_pop=0.0, tol_align=0.0,
                                    precision=precision)
    assert tot["iterations"] == iters
    assert tot["rt_variant"] == (3 if form & 1 else 2) and tot["rt_order"] == (form >> 1)
    for (sp, g), pt, (pp, lev) in zip(models, pts, singles):
        idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
        np.testing.assert_allclose(pt.get_levpop()[:, idx], lev[:, idx], rtol=tol)
        np.testing.assert_allclose(pt.get_pol_pops()[idx][:, 0::2], pp[idx][:, 0::2], rtol=tol)
        pt.close()
