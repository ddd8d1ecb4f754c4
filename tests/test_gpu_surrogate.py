"""What the trained surrogate gains, and its fallback (needs an MI355X and trained weights)."""
import numpy as np
import pytest

from portal_amd import cloud, surrogate
from portal_amd.api import Portal
from portal_amd.lib import PORTAL_MIXED

pytestmark = pytest.mark.gpu


def _setup(levels, n, seed):
    sp = cloud.co_like_rotor(levels)
    g = cloud.make_cloud(sp, n, seed=seed)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    return sp, g, pt


def test_surrogate_reduces_iterations_to_the_same_answer():
    sp, g, pt = _setup(4, 4000, seed=777)  # a cloud the network was not trained on
    if not surrogate.load_into(pt, sp):
        pt.close()
        pytest.skip("no trained weights for this species")
    xth = cloud.legendre_all(6, 11)
    kw = dict(tol_pop=1e-5, tol_align=1e-7, precision=PORTAL_MIXED)
    pt.set_levpop(g.lev_p)
    pp0, lev0, st0 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=0, **kw)
    pt.set_levpop(g.lev_p)
    pp1, lev1, st1 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=1, **kw)
    print(f"plain: {st0['iterations']} iterations; surrogate: {st1['iterations']} iterations, "
          f"{st1['fallbacks']} fallbacks")
    assert st1["iterations"] < st0["iterations"]
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    rel = np.abs(lev1[:, idx] - lev0[:, idx]) / np.maximum(np.abs(lev0[:, idx]), 1e-30)
    big = lev0[:, idx] > 1e-6 * lev0[:, idx].max(axis=0)
    # the alignment, which sets Stokes Q and U: rho^2_0/rho^0_0 per level with J > 0
    jl = [i for i in range(sp.e_num) if sp.j_lev[i] > 0]
    a0 = pp0[idx][:, [2 * i + 1 for i in jl]] / pp0[idx][:, [2 * i for i in jl]]
    a1 = pp1[idx][:, [2 * i + 1 for i in jl]] / pp1[idx][:, [2 * i for i in jl]]
    da = np.abs(a1 - a0)[big[jl].T]
    print(f"same answer: population max rel {rel[big].max():.2e}, alignment max abs {da.max():.2e} "
          f"(tolerances {kw['tol_pop']:g} / {kw['tol_align']:g} per iteration)")
    assert rel[big].max() < 2e-4
    assert da.max() < 2e-5
    pt.close()


def test_invalid_network_output_falls_back_to_the_plain_iterate():
    sp, g, pt = _setup(4, 1500, seed=778)
    n_in, n_out = surrogate.dims(sp)
    pt.mlp_create(n_in, 64, 2, n_out, seed=1)
    p, nm = pt.mlp_get_params()
    nm[2 * n_in: 2 * n_in + n_out] = 1e6  # every decoded population overflows
    pt.mlp_set_params(p, nm)
    xth = cloud.legendre_all(4, 8)
    pt.set_levpop(g.lev_p)
    pp0, lev0, st0 = pt.lambda_iterate(xth, 16, cloud.DX_INWARD, 3, tol_pop=0, tol_align=0, surrogate=0)
    pt.set_levpop(g.lev_p)
    pp1, lev1, st1 = pt.lambda_iterate(xth, 16, cloud.DX_INWARD, 3, tol_pop=0, tol_align=0, surrogate=1)
    n_solved = int((g.ab_all[: g.bulk_node_num] > 0).sum())
    assert st1["fallbacks"] == n_solved
    np.testing.assert_array_equal(lev1, lev0)
    pt.close()


def test_prediction_that_inverts_a_line_falls_back():
    # network outputs fixed through the output normalisation (zero spread): all-zero
    # corrections predict rho_1 itself and every node accepts it; a correction of -1 on
    # the J=0 level's log10 population inverts the 1-0 line wherever n1/n0 > 0.3, and
    # every node falls back, including those whose renormalisation alone would pass
    sp, g, pt = _setup(4, 1500, seed=780)
    n_in, n_out = surrogate.dims(sp)
    xth = cloud.legendre_all(4, 8)
    pt.set_levpop(g.lev_p)
    pp1, _ = pt.se_pass(xth, 16, cloud.DX_INWARD)  # rho_1: the state the network corrects
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    f0 = pp1[idx, 0] * 1.0 / (pp1[idx][:, 0::2] * np.sqrt(2.0 * sp.j_lev + 1.0)).sum(axis=1)
    n1_n0 = pp1[idx, 2] * np.sqrt(3.0) / pp1[idx, 0]
    assert n1_n0.min() > 0.3
    sum_passes = (1.0 - 0.9 * f0 > 0.5).sum()
    assert sum_passes > len(idx) // 2

    def run(y_mean):
        pt.mlp_create(n_in, 64, 2, n_out, seed=1)
        p, nm = pt.mlp_get_params()
        nm[2 * n_in: 2 * n_in + n_out] = y_mean
        nm[2 * n_in + n_out:] = 0.0
        pt.mlp_set_params(p, nm)
        pt.set_levpop(g.lev_p)
        return pt.lambda_iterate(xth, 16, cloud.DX_INWARD, 1, tol_pop=0, tol_align=0, surrogate=1)

    y = np.zeros(n_out, dtype=np.float32)
    pp_id, _, st_id = run(y)
    assert st_id["fallbacks"] == 0
    np.testing.assert_allclose(pp_id[idx][:, 0::2], pp1[idx][:, 0::2], rtol=1e-12)
    y[0] = -1.0
    _, _, st_inv = run(y)
    print(f"{len(idx)} nodes, {sum_passes} pass the normalisation test; fallbacks {st_inv['fallbacks']}")
    assert st_inv["fallbacks"] == len(idx)
    pt.close()


def test_features_are_the_first_pass_state():
    # the device features against the same quantities formed on the host
    sp, g, pt = _setup(4, 1200, seed=779)
    xth = cloud.legendre_all(4, 8)
    pt.set_levpop(g.lev_p)
    J0, J2, _ = pt.inverse_rt(xth, 16, cloud.DX_INWARD)
    pt.set_levpop(g.lev_p)
    pp, _ = pt.se_pass(xth, 16, cloud.DX_INWARD)
    f, nodes = pt.surrogate_features()
    n_in, _ = surrogate.dims(sp)
    assert f.shape == (len(nodes), n_in)
    r = nodes - 1
    nu = np.asarray(sp.nu_hz, dtype=np.float64)
    iin = 1.4744995e-50 * nu ** 3 / (np.exp(4.79924e-11 / 2.73 * nu) - 1.0)  # B_nu(2.73 K), miscellaneous.f90:263-285
    exp = [np.log10(g.h2_dens[r] * 1e-6), np.log10(g.temp_all[0, r])]
    for k in range(sp.r_num):
        exp += [np.log10(J0[k, r] / iin[k]), 100.0 * J2[k, r] / J0[k, r]]
    lp, rat = surrogate._state(pp, nodes, sp)
    want = np.column_stack(exp + list(lp.T) + list(rat.T))
    np.testing.assert_allclose(f, want, rtol=2e-6, atol=2e-6)
    pt.close()


def test_residual_test_reverts_exactly_the_wrong_predictions():
    """A network made wrong at every 7th node (its first level's population x1e3,
    renormalised: it passes the validity tests) is caught by the residual test one
    iteration later: every such node reverts to its plain iterate, the count is
    stats.fallbacks, the other nodes keep the prediction about as often as without the
    perturbation, and the iteration reaches the plain fixed point in no more iterations
    than without the surrogate and in fewer than with the residual test off."""
    sp, g, pt = _setup(4, 4000, seed=777)
    if not surrogate.load_into(pt, sp):
        pt.close()
        pytest.skip("no trained weights for this species")
    xth = cloud.legendre_all(6, 11)
    kw = dict(tol_pop=1e-5, tol_align=1e-7, precision=PORTAL_MIXED)
    pt.set_levpop(g.lev_p)
    _, lev0, st0 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=0, **kw)
    pt.set_levpop(g.lev_p)
    _, lev1, st1 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=1, fallback_tol=1e-8, **kw)
    clean = pt.fallback_map()
    pt.surrogate_perturb(7, 1e3)
    pt.set_levpop(g.lev_p)
    _, lev2, st2 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=1, fallback_tol=1e-8, **kw)
    fm = pt.fallback_map()
    # the residual test switched off: the same wrong predictions stay in
    pt.set_levpop(g.lev_p)
    _, lev3, st3 = pt.lambda_iterate(xth, 32, cloud.DX_INWARD, 200, surrogate=1, fallback_tol=float("inf"), **kw)
    off = pt.fallback_map()
    pt.surrogate_perturb(0)
    pt.close()
    w = np.arange(len(fm))
    pert = w % 7 == 0
    print(f"plain {st0['iterations']} it; surrogate {st1['iterations']} it, {st1['fallbacks']} fallbacks; "
          f"perturbed {st2['iterations']} it, {st2['fallbacks']} fallbacks ({pert.sum()} perturbed); "
          f"residual test off {st3['iterations']} it")
    assert len(fm) == int((g.ab_all[: g.bulk_node_num] > 0).sum())
    assert np.all(fm[pert] == 2)
    assert st2["fallbacks"] == int((fm != 0).sum())
    assert st1["fallbacks"] == int((clean != 0).sum())
    # unperturbed nodes: the neighbours' wrong first iterate moves a few across the line
    extra = int(((fm != 0) & ~pert).sum()) - int(((clean != 0) & ~pert).sum())
    assert extra <= 0.02 * (~pert).sum()
    assert np.all(off[pert] == 0)
    # the same fixed point, and still fewer iterations than the plain iteration
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    big = lev0[:, idx] > 1e-6 * lev0[:, idx].max(axis=0)
    rel = np.abs(lev2[:, idx] - lev0[:, idx]) / np.maximum(np.abs(lev0[:, idx]), 1e-30)
    assert rel[big].max() < 1e-3
    # a seventh of the nodes restarting from their plain iterate costs the surrogate's
    # gain, but no more; keeping the wrong predictions costs more than that
    assert st2["iterations"] <= st0["iterations"]
    assert st2["iterations"] < st3["iterations"]
