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
    batch_lambda_iterate(a, [xth] * 3, [8] * 3, cloud.DX_INWARD, 7, tol_pop=0.0, tol_align=0.0)
    b = fresh()
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


def test_batch_ng_reaches_the_plain_fixed_point():
    # Ng acceleration per model in the batch: fewer iterations, same fixed point
    models = _models()
    xth = cloud.legendre_all(4, 8)
    kw = dict(tol_pop=1e-7, tol_align=1e-9)
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
        assert (np.abs(b[:, idx] - a[:, idx])[big] / a[:, idx][big]).max() < 2e-5
