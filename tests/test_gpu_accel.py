"""Ng and Anderson acceleration and the accelerated lambda iteration (own-cell operator)
of the lambda iteration (needs an MI355X).

The accelerated iteration must reach the fixed point of the plain one, in fewer
iterations, in both precisions.
"""
import numpy as np
import pytest

from portal_amd import cloud
from portal_amd.api import Portal
from portal_amd.lib import PORTAL_FP64, PORTAL_MIXED

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 4000, seed=31)
    xth = cloud.legendre_all(6, 11)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    yield sp, g, xth, pt
    pt.close()


@pytest.mark.parametrize("accel,ali", [(1, 0), (2, 0), (0, 1), (1, 1)], ids=["ng", "anderson", "ali", "ali+ng"])
@pytest.mark.parametrize("precision,tol", [(PORTAL_FP64, 1e-9), (PORTAL_MIXED, 1e-6)])
def test_acceleration_reaches_the_plain_fixed_point(small, precision, tol, accel, ali):
    sp, g, xth, pt = small
    kw = dict(tol_pop=tol, tol_align=tol * 1e-2, precision=precision)
    pt.set_levpop(g.lev_p)
    pp0, lev0, st0 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 400, **kw)
    pt.set_levpop(g.lev_p)
    pp1, lev1, st1 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 400, ng_accel=accel, ali=ali, **kw)
    print(f"precision {precision}: plain {st0['iterations']} iterations, accel {accel} ali {ali}: "
          f"{st1['iterations']} ({st1['ng_steps']} extrapolations)")
    assert st0["iterations"] < 400 and st1["iterations"] < 400
    assert st1["ng_steps"] > 0 or accel == 0
    assert st1["iterations"] < st0["iterations"]
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    big = lev0[:, idx] > 1e-8 * lev0[:, idx].max(axis=0)
    rel = np.abs(lev1[:, idx] - lev0[:, idx]) / np.abs(lev0[:, idx])
    # both stop within their tolerance of the fixed point; the remaining distance of a
    # geometric iteration is tol / (1 - ratio), ratio ~0.9 on this cloud
    assert rel[big].max() < 50 * tol
    al0 = pp0[idx][:, 3::2] / pp0[idx][:, 2::2]
    al1 = pp1[idx][:, 3::2] / pp1[idx][:, 2::2]
    assert np.abs(al1 - al0).max() < 50 * tol


def test_a_continued_call_is_the_same_iteration(small):
    # 4 + 3 iterations in two calls equal 7 in one; the tolerance test of the second call
    # compares with the first call's last SE solution
    sp, g, xth, pt = small
    pt.set_levpop(g.lev_p)
    pp1, lev1, st1 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 7, tol_pop=0.0, tol_align=0.0)
    pt.set_levpop(g.lev_p)
    pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 4, tol_pop=0.0, tol_align=0.0)
    pp2, lev2, st2 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 3, tol_pop=0.0, tol_align=0.0)
    np.testing.assert_array_equal(lev2, lev1)
    np.testing.assert_array_equal(pp2, pp1)
    assert st2["max_rel_change"] == st1["max_rel_change"]
    # the iteration at which a tolerance is met does not depend on where calls split
    kw = dict(tol_pop=1e-4, tol_align=1e-6)
    pt.set_levpop(g.lev_p)
    n_one = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, outputs=False, **kw)["iterations"]
    pt.set_levpop(g.lev_p)
    done = 0
    while True:
        st = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 5, outputs=False, **kw)
        done += st["iterations"]
        if st["converged"] or done > 200:
            break
    assert done == n_one


def test_ali_fixed_point_is_exact(small):
    # the own-cell operator changes the path, not the solution: iterated far below the
    # tolerances above, the plain and accelerated FP64 iterations agree to rounding level
    sp, g, xth, pt = small
    kw = dict(tol_pop=1e-13, tol_align=1e-15, precision=PORTAL_FP64)
    pt.set_levpop(g.lev_p)
    pp0, lev0, st0 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 600, ng_accel=1, **kw)
    pt.set_levpop(g.lev_p)
    pp1, lev1, st1 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 600, ng_accel=1, ali=1, **kw)
    print(f"to 1e-13: Ng {st0['iterations']} iterations, Ng + ALI {st1['iterations']}")
    assert st0["converged"] and st1["converged"]
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    big = lev0[:, idx] > 1e-8 * lev0[:, idx].max(axis=0)
    rel = np.abs(lev1[:, idx] - lev0[:, idx]) / np.abs(lev0[:, idx])
    print(f"max rel difference of the level densities {rel[big].max():.2e}")
    assert rel[big].max() < 1e-10
    al0 = pp0[idx][:, 3::2] / pp0[idx][:, 2::2]
    al1 = pp1[idx][:, 3::2] / pp1[idx][:, 2::2]
    assert np.abs(al1 - al0).max() < 1e-12


@pytest.mark.parametrize("precision,tol", [(PORTAL_FP64, 1e-12), (PORTAL_MIXED, 2e-5)])
def test_ali_operator_is_the_same_in_both_orders(small, precision, tol):
    # the own-cell operator of the accelerated iteration comes from the cooperative
    # kernel in either order of a ray's cells (variant 2 front to back, 3 back to front):
    # five accelerated iterations give the same level densities
    sp, g, xth, pt = small
    out = []
    try:
        for variant in (2, 3):
            pt.force_rt_kernel(precision, variant, 4)
            pt.set_levpop(g.lev_p)
            _, lev, st = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 5, tol_pop=0.0, tol_align=0.0,
                                           precision=precision, ali=1)
            assert st["iterations"] == 5 and st["rt_variant"] == variant
            out.append(lev)
    finally:
        pt.force_rt_kernel(precision, -1)
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    big = out[0][:, idx] > 1e-8 * out[0][:, idx].max(axis=0)
    rel = np.abs(out[1][:, idx] - out[0][:, idx]) / np.abs(out[0][:, idx])
    print(f"precision {precision}: max rel difference between the orders {rel[big].max():.2e}")
    assert rel[big].max() < tol
