"""HIP path against the oracle, through the C ABI (needs an MI355X).

FP64 mode must agree with the oracle to 1e-10 relative: J^0_0 per line and node,
rho^0_0 per level, and rho^2_0 to 1e-10 of the level's rho^0_0 (rho^2_0 is a small
difference of large rates, so it is compared on the scale of its level).  Mixed
mode reports its error and must stay within the stated tolerances.
"""
import numpy as np
import pytest

from portal_amd import cloud
from portal_amd.api import Portal, PortalError
from portal_amd.lib import PORTAL_FP64, PORTAL_MIXED

pytestmark = pytest.mark.gpu

NTH = 16  # oracle threads on the GPU host


@pytest.fixture(scope="module")
def cloud4k():
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 4000, seed=1)
    xth = cloud.legendre_all(6, 11)
    return sp, g, xth


@pytest.fixture(scope="module")
def portal1(cloud4k):
    sp, g, xth = cloud4k
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    yield pt
    pt.close()


@pytest.fixture(scope="module")
def oracle_cloud4k(oracle, cloud4k):
    sp, g, xth = cloud4k
    pp, J0, J2, rates, jj = oracle.se_pass(g, sp, xth, 32, cloud.DX_INWARD, extra_info=1, nthreads=NTH)
    return pp, J0, J2, rates, jj


def solved(g):
    return np.where(g.ab_all[: g.bulk_node_num] > 0)[0]


def assert_pops_close(pp, ref, sp, idx, rtol):
    """rho^0_0 per node and level to rtol; rho^2_0 per node and level to rtol of that
    level's own rho^0_0 at that node (of 1e-6 of the node's largest rho^0_0 for levels
    holding less: their rho^2_0 is a rounding-level remainder of the solve)."""
    node_max = np.max(np.abs(ref[idx][:, 0::2]), axis=1)
    for i in range(sp.e_num):
        r0, o0 = pp[idx, 2 * i], ref[idx, 2 * i]
        np.testing.assert_allclose(r0, o0, rtol=rtol, atol=0)
        scale = np.maximum(np.abs(o0), 1e-6 * node_max)
        d2 = np.abs(pp[idx, 2 * i + 1] - ref[idx, 2 * i + 1])
        assert np.all(d2 <= rtol * scale), f"level {i}: max |d rho20| / rho00 = {np.max(d2 / scale):.2e}"


def test_inverse_rt_fp64_parity(portal1, cloud4k, oracle_cloud4k):
    sp, g, xth = cloud4k
    _, J0o, J2o, _, _ = oracle_cloud4k
    J0, J2, st = portal1.inverse_rt(xth, 32, cloud.DX_INWARD)
    idx = solved(g)
    np.testing.assert_allclose(J0[:, idx], J0o[:, idx], rtol=1e-10, atol=0)
    np.testing.assert_allclose(J2[:, idx], J2o[:, idx], rtol=0, atol=1e-10 * np.abs(J0o[:, idx]).max(axis=0).min())
    # nodes that are not solved get zero (main_para.f90:206-208)
    mask = np.ones(g.node_num, bool)
    mask[idx] = False
    assert np.all(J0[:, mask] == 0)
    assert st["n_cross"] > 0 and st["n_paths"] == len(idx) * xth.shape[1]


def test_walk_forms_agree():
    # the single-pass walk (chunked entries + compaction, store sized from a sampled count),
    # its overflow fallback (a second walk at exact offsets) and the two-pass walk build
    # the same paths: the inward pass gives identical J
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 20000, seed=7)
    xth = cloud.legendre_all(4, 8)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    out = []
    for scale in (1.1, 0.05, 0.0):
        pt.walk_chunks(scale)
        J0, J2, st = pt.inverse_rt(xth, 16, cloud.DX_INWARD)
        out.append((J0, J2, st["n_cross"], st["max_path"]))
    pt.close()
    for J0, J2, nc, mp in out[1:]:
        np.testing.assert_array_equal(J0, out[0][0])
        np.testing.assert_array_equal(J2, out[0][1])
        assert nc == out[0][2] and mp == out[0][3]


@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("precision", [PORTAL_FP64, PORTAL_MIXED])
def test_cooperative_units_agree(precision, variant):
    # the cooperative kernel stages rays in units of at most 64 cells; the unit size
    # changes where the staging breaks a ray, not the arithmetic: J is identical (back
    # to front, variant 3, the units of a ray are staged last to first)
    sp = cloud.co_like_rotor(6)
    g = cloud.make_cloud(sp, 20000, seed=9)
    xth = cloud.legendre_all(4, 8)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pt.force_rt_kernel(precision, variant, 4)
    out = []
    for unit in (0, 8, 4096):
        pt.coop_unit(unit)
        J0, J2, st = pt.inverse_rt(xth, 16, cloud.DX_INWARD, precision=precision)
        out.append((J0, J2))
    assert st["max_path"] > 8  # the 8-cell unit splits most rays
    pt.close()
    for J0, J2 in out[1:]:
        np.testing.assert_array_equal(J0, out[0][0])
        np.testing.assert_array_equal(J2, out[0][1])


def test_se_pass_fp64_parity(portal1, cloud4k, oracle_cloud4k):
    sp, g, xth = cloud4k
    ppo, _, _, rates_o, jj_o = oracle_cloud4k
    pp, rates, jj, st = portal1.se_pass(xth, 32, cloud.DX_INWARD, extra_info=1)
    idx = solved(g)
    assert_pops_close(pp, ppo, sp, idx, 1e-10)
    # normalisation sum_i sqrt(2J+1) rho00 = 1 on every solved node
    w = np.sqrt(2 * sp.j_lev + 1.0)
    np.testing.assert_allclose(pp[idx][:, 0::2] @ w, 1.0, atol=1e-13)
    np.testing.assert_allclose(jj[idx], jj_o[idx], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(rates[idx], rates_o[idx], rtol=1e-9, atol=1e-9 * np.abs(rates_o[idx]).max())


def test_se_pass_with_collisional_depolarisation(oracle):
    """The depolarisation option (depol.f90:242-359; the reference switches it off at
    main_sub.f90:96): the library's D^(K) table equals the oracle's, the FP64 SE pass with
    it equals the oracle's to 1e-10 (populations, alignment of every level, the collisional
    rates that carry -D^(K)), and switching it off again gives the plain pass."""
    sp = cloud.co_like_rotor(6)
    g = cloud.make_cloud(sp, 1500, seed=21)
    xth = cloud.legendre_all(4, 8)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    D = pt.depol_table()
    np.testing.assert_allclose(D, oracle.depol_table(sp), rtol=1e-13, atol=0)
    assert D[:, 1:, 0].min() > 0
    D = 30.0 * D  # stronger than these rotors' rates give, so that the alignment changes visibly
    pt.set_depol(D)
    pp, rates, _, _ = pt.se_pass(xth, 16, cloud.DX_INWARD, extra_info=1)
    ppo, _, _, rates_o, _ = oracle.se_pass(g, sp, xth, 16, cloud.DX_INWARD, extra_info=1, nthreads=NTH, depol=D)
    idx = solved(g)
    assert_pops_close(pp, ppo, sp, idx, 1e-10)
    np.testing.assert_allclose(rates[idx], rates_o[idx], rtol=1e-9, atol=1e-9 * np.abs(rates_o[idx]).max())
    pt.set_depol(None)
    pp0, _ = pt.se_pass(xth, 16, cloud.DX_INWARD)
    ppo0 = oracle.se_pass(g, sp, xth, 16, cloud.DX_INWARD, nthreads=NTH)[0]
    pt.close()
    assert_pops_close(pp0, ppo0, sp, idx, 1e-10)
    al = lambda p: np.abs(p[idx][:, 3::2]) / p[idx][:, 2::2]  # noqa: E731  (levels with J > 0)
    print(f"mean |rho20/rho00|: {al(pp0).mean():.3e} without, {al(pp).mean():.3e} with depolarisation")
    assert al(pp).mean() < 0.8 * al(pp0).mean()


def test_se_solve_from_oracle_J(portal1, cloud4k, oracle, oracle_cloud4k):
    sp, g, xth = cloud4k
    _, J0o, J2o, _, _ = oracle_cloud4k
    pp = portal1.se_solve(J0o, J2o)
    ref = oracle.solve_from_J(g, sp, J0o, J2o)
    assert_pops_close(pp, ref, sp, solved(g), 1e-11)


@pytest.mark.parametrize("waves", [1, 2, 4])
def test_se_solve_74_dimensional_systems(oracle, waves):
    # a 20-level rotor: 74-dimensional systems, two rows per lane in the factorisation and
    # the substitutions, split over 1, 2 or 4 waves; same J into both solvers
    sp = cloud.co_like_rotor(20)
    g = cloud.make_cloud(sp, 400, seed=61, n_exp=-1.5)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    assert pt.species_info()[1] == 74
    J0, J2, _ = pt.inverse_rt(cloud.legendre_all(4, 8), 8, cloud.DX_INWARD)
    pt.force_se(-1, waves)
    pp = pt.se_solve(J0, J2)
    pt.close()
    ref = oracle.solve_from_J(g, sp, J0, J2)
    idx = solved(g)
    w = np.sqrt(2 * sp.j_lev + 1.0)
    for i in range(sp.e_num):
        o0 = ref[idx, 2 * i]
        big = w[i] * o0 > 1e-10  # levels that hold a share of the population
        assert big.any() or i > 10
        np.testing.assert_allclose(pp[idx, 2 * i][big], o0[big], rtol=1e-10, atol=0)
        np.testing.assert_allclose(pp[idx, 2 * i][~big], o0[~big], rtol=0, atol=1e-20)
        np.testing.assert_allclose(pp[idx, 2 * i + 1], ref[idx, 2 * i + 1], rtol=0, atol=1e-10 * np.abs(o0).max())


def test_mixed_precision_error_bounds(portal1, cloud4k, oracle_cloud4k):
    sp, g, xth = cloud4k
    ppo, J0o, J2o, _, _ = oracle_cloud4k
    idx = solved(g)
    J0, J2, _ = portal1.inverse_rt(xth, 32, cloud.DX_INWARD, precision=PORTAL_MIXED)
    e0 = np.abs(J0[:, idx] / J0o[:, idx] - 1).max()
    e2 = np.abs(J2[:, idx] - J2o[:, idx]).max() / np.abs(J0o[:, idx]).max()
    print(f"mixed: max rel J0 err {e0:.3e}, max J2 err / max J0 {e2:.3e}")
    assert e0 < 2e-5 and e2 < 1e-6
    pp, _ = portal1.se_pass(xth, 32, cloud.DX_INWARD, precision=PORTAL_MIXED)
    r0 = np.abs(pp[idx][:, 0::2] / ppo[idx][:, 0::2] - 1).max()
    print(f"mixed: max rel rho00 err {r0:.3e}")
    assert r0 < 1e-5


def test_lambda_iteration_matches_oracle(cloud4k, oracle):
    # two closure steps lev_p <- ab n sqrt(2J+1) rho00 on a small cloud
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 800, seed=9)
    xth = cloud.legendre_all(4, 8)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pp, lev, st = pt.lambda_iterate(xth, 16, cloud.DX_INWARD, max_iter=3, tol_pop=0.0, tol_align=0.0)
    assert st["iterations"] == 3
    lev_o = g.lev_p.copy()
    idx = solved(g)
    for _ in range(3):
        ppo, _, _ = oracle.se_pass(g, sp, xth, 16, cloud.DX_INWARD, lev_p=lev_o, nthreads=NTH)
        for i in range(sp.e_num):
            lev_o[i, idx] = g.ab_all[idx] * g.h2_dens[idx] * np.sqrt(2 * sp.j_lev[i] + 1.0) * ppo[idx, 2 * i]
    np.testing.assert_allclose(lev[:, idx], lev_o[:, idx], rtol=1e-10)
    assert_pops_close(pp, ppo, sp, idx, 1e-10)
    pt.close()


def test_full_size_subset_parity(oracle):
    """200,000-node cloud with the 10-level rotor, 200 directions, 129 channels:
    all nodes on the GPU, a seeded subset through the oracle."""
    sp = cloud.co_like_rotor(10)
    g = cloud.make_cloud(sp, 200000, seed=2)
    xth = cloud.legendre_all(9, 32)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    J0, J2, st = pt.inverse_rt(xth, 64, cloud.DX_INWARD)
    pp = pt.se_solve(J0, J2)
    idx = solved(g)
    # size-independent properties on every node: J0 > 0, normalisation
    assert np.all(J0[:, idx] > 0)
    w = np.sqrt(2 * sp.j_lev + 1.0)
    np.testing.assert_allclose(pp[idx][:, 0::2] @ w, 1.0, atol=1e-12)
    rng = np.random.default_rng(77)
    sub = np.sort(rng.choice(idx, 32, replace=False)) + 1
    J0o, J2o = oracle.inverse_nodes(g, sp, xth, 64, cloud.DX_INWARD, sub, nthreads=NTH)
    np.testing.assert_allclose(J0[:, sub - 1], J0o, rtol=1e-10)
    np.testing.assert_allclose(J2[:, sub - 1], J2o, rtol=0, atol=1e-10 * J0o.max(axis=0).min())
    pt.close()


# variant 3: the cooperative kernel taking the cells of a ray back to front
KERNELS = [(0, 4), (1, 4), (1, 2), (2, 2), (2, 4), (2, 5), (2, 8), (3, 4), (3, 5), (3, 8)]


@pytest.mark.parametrize("variant,waves", KERNELS)
def test_each_inward_kernel_4000_nodes(portal1, cloud4k, oracle_cloud4k, variant, waves):
    # every kernel the autotuner may pick, FP64 against the oracle and mixed within bounds
    sp, g, xth = cloud4k
    _, J0o, J2o, _, _ = oracle_cloud4k
    idx = solved(g)
    try:
        portal1.force_rt_kernel(PORTAL_FP64, variant, waves)
        portal1.force_rt_kernel(PORTAL_MIXED, variant, waves)
        J0, J2, st = portal1.inverse_rt(xth, 32, cloud.DX_INWARD)
        assert (st["rt_variant"], st["rt_waves"]) == (variant, waves)
        np.testing.assert_allclose(J0[:, idx], J0o[:, idx], rtol=1e-10, atol=0)
        np.testing.assert_allclose(J2[:, idx], J2o[:, idx], rtol=0,
                                   atol=1e-10 * np.abs(J0o[:, idx]).max(axis=0).min())
        J0m, J2m, _ = portal1.inverse_rt(xth, 32, cloud.DX_INWARD, precision=PORTAL_MIXED)
        assert np.abs(J0m[:, idx] / J0o[:, idx] - 1).max() < 2e-5
        assert np.abs(J2m[:, idx] - J2o[:, idx]).max() / np.abs(J0o[:, idx]).max() < 1e-6
    finally:
        portal1.force_rt_kernel(PORTAL_FP64, -1)
        portal1.force_rt_kernel(PORTAL_MIXED, -1)


@pytest.fixture(scope="module")
def cloud10(oracle):
    # the 10-level rotor with the production angular and velocity grids on a smaller cloud
    sp = cloud.co_like_rotor(10)
    g = cloud.make_cloud(sp, 20000, seed=21)
    xth = cloud.legendre_all(9, 32)
    rng = np.random.default_rng(5)
    sub = np.sort(rng.choice(solved(g), 16, replace=False)) + 1
    J0o, J2o = oracle.inverse_nodes(g, sp, xth, 64, cloud.DX_INWARD, sub, nthreads=NTH)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    yield sp, g, xth, sub, J0o, J2o, pt
    pt.close()


@pytest.mark.parametrize("variant,waves", KERNELS)
def test_each_inward_kernel_10_levels(cloud10, variant, waves):
    sp, g, xth, sub, J0o, J2o, pt = cloud10
    try:
        pt.force_rt_kernel(PORTAL_FP64, variant, waves)
        pt.force_rt_kernel(PORTAL_MIXED, variant, waves)
        J0, J2, st = pt.inverse_rt(xth, 64, cloud.DX_INWARD)
        np.testing.assert_allclose(J0[:, sub - 1], J0o, rtol=1e-10)
        np.testing.assert_allclose(J2[:, sub - 1], J2o, rtol=0, atol=1e-10 * J0o.max(axis=0).min())
        J0m, J2m, _ = pt.inverse_rt(xth, 64, cloud.DX_INWARD, precision=PORTAL_MIXED)
        idx = solved(g)
        e0 = np.abs(J0m[:, idx] / J0[:, idx] - 1).max()
        e2 = np.abs(J2m[:, idx] - J2[:, idx]).max() / np.abs(J0[:, idx]).max()
        print(f"kernel {variant}/{waves}: mixed max rel J0 err {e0:.3e}, J2 err {e2:.3e}")
        assert e0 < 2e-5 and e2 < 1e-6
    finally:
        pt.force_rt_kernel(PORTAL_FP64, -1)
        pt.force_rt_kernel(PORTAL_MIXED, -1)


def test_error_paths(cloud4k):
    sp, g, xth = cloud4k
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    with pytest.raises(PortalError) as e:
        pt.inverse_rt(xth, 32, cloud.DX_INWARD, n_depth=3)
    assert e.value.code == -3  # PORTAL_ERR_DEPTH
    # a kinetic temperature above the collision table (the reference leaves C undefined)
    g2 = cloud.make_cloud(sp, 300, seed=4)
    g2.temp_all[0, 5] = 2e4
    pt.upload_grid(g2)
    pt.set_levpop(g2.lev_p)
    with pytest.raises(PortalError) as e:
        pt.se_pass(cloud.legendre_all(4, 8), 8, cloud.DX_INWARD)
    assert e.value.code == -6  # PORTAL_ERR_TEMP
    # no node to solve: refused before any launch
    g3 = cloud.make_cloud(sp, 300, seed=5)
    g3.ab_all[:] = 0.0
    pt.upload_grid(g3)
    pt.set_levpop(g3.lev_p)
    with pytest.raises(PortalError) as e:
        pt.se_pass(cloud.legendre_all(4, 8), 8, cloud.DX_INWARD)
    assert e.value.code == -1  # PORTAL_ERR_ARG
    pt.close()


def _disconnected_species():
    """A 3-level rotor whose J = 2 level has no radiative or collisional transition: its
    rows of the SE matrix are empty, so the matrix is singular."""
    ct = np.array([1.0, 10.0, 100.0, 1e3, 1e4])
    return cloud.Species(e_num=3, j_lev=np.array([0, 1, 2], dtype=np.int32), e_levs_cm=np.array([0.0, 3.845, 11.535]),
                         mass_mol=28.0, r_trans=np.array([[1], [2]], dtype=np.int32), t_Aji=np.array([7.2e-8]),
                         nu_ghz=np.array([115.27]), c_trans=np.array([[1], [2]], dtype=np.int32), col_temps=ct,
                         t_Cji=np.asfortranarray(np.full((5, 1, 1), 1e-11)), kmax=2)


def test_singular_se_system_is_reported(oracle):
    """A singular SE matrix stops the reference (em_coefs.f90:170-178); the oracle and the
    library report it (PORTAL_ERR_LU) in both precisions, single model and batch, instead
    of returning a solution."""
    sp = _disconnected_species()
    g = cloud.make_cloud(sp, 300, seed=6)
    xth = cloud.legendre_all(4, 8)
    with pytest.raises(RuntimeError):
        oracle.se_pass(g, sp, xth, 8, cloud.DX_INWARD, nthreads=NTH)
    from portal_amd.api import batch_lambda_iterate
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    for prec in (PORTAL_FP64, PORTAL_MIXED):
        pt.set_levpop(g.lev_p)
        with pytest.raises(PortalError) as e:
            pt.se_pass(xth, 8, cloud.DX_INWARD, precision=prec)
        assert e.value.code == -5  # PORTAL_ERR_LU
        pt.set_levpop(g.lev_p)
        with pytest.raises(PortalError) as e:
            batch_lambda_iterate([pt], [xth], [8], cloud.DX_INWARD, 2, precision=prec)
        assert e.value.code == -5
    pt.close()


def test_species_beyond_the_solver_size_is_refused():
    """A 110-level rotor at kmax = 8 has 540 unknowns, beyond the 512 the solver takes: the
    upload is refused with PORTAL_ERR_ARG, and the context stays usable."""
    pt = Portal(0)
    with pytest.raises(PortalError) as e:
        pt.upload_species(cloud.co_like_rotor(110, kmax=8))
    assert e.value.code == -1
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 300, seed=7)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pp, _ = pt.se_pass(cloud.legendre_all(4, 8), 8, cloud.DX_INWARD)
    assert np.all(np.isfinite(pp))
    pt.close()


@pytest.mark.parametrize("kmax,unknowns", [(4, 120), (6, 158), (8, 195)])
def test_large_se_systems_take_the_slab_form(oracle, kmax, unknowns):
    """A 41-level rotor (LAMDA's CO has 41 levels, 820 transitions) at kmax = 4, 6 and 8.
    At 120 unknowns the matrix, its copy and the scalars do not fit one workgroup's LDS,
    so the copy and the scalars go to a global slab; above 128 the factors too, with the
    substitutions on all waves.  FP64 against the oracle to 1e-10."""
    sp = cloud.co_like_rotor(41, kmax=kmax)
    g = cloud.make_cloud(sp, 300, seed=63)
    xth = cloud.legendre_all(3, 6)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    assert pt.species_info()[1] == unknowns
    pp, _ = pt.se_pass(xth, 8, cloud.DX_INWARD)
    pt.close()
    ref = oracle.se_pass(g, sp, xth, 8, cloud.DX_INWARD, nthreads=NTH)[0]
    assert_pops_close(pp, ref, sp, solved(g), 1e-10)


def test_se_forms_give_the_same_bits():
    """On 74-unknown systems the LDS form, the slab form and the all-global form (factors
    in the slab, substitutions on all waves), at 1 and 4 waves per system, give the same
    bits: every row sees the same operations in the same order."""
    sp = cloud.co_like_rotor(20)
    g = cloud.make_cloud(sp, 400, seed=64)
    xth = cloud.legendre_all(3, 6)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    out = []
    for form, waves in ((-1, 1), (2, 4), (2, 1), (3, 4), (3, 1)):
        pt.force_se(form, waves)
        pt.set_levpop(g.lev_p)
        out.append(pt.se_pass(xth, 8, cloud.DX_INWARD)[0])
    pt.close()
    for o in out[1:]:
        np.testing.assert_array_equal(o, out[0])


def test_image_pass_fp64_parity(portal1, cloud4k, oracle, oracle_cloud4k):
    from portal_amd import driver
    sp, g, xth = cloud4k
    ppo = oracle_cloud4k[0]
    pp = driver.rescale_pol_pops(ppo, g.lev_p, sp.lev_g)
    lines = driver.image_lines(sp, [1], [2])
    vel_out = driver.channel_velocities(32, 93.75)
    args = (pp, lines, np.pi / 4, 0.0, 32, vel_out, 24, 0.7, cloud.DX_IMAGE)
    (I, Q, U, T), st = portal1.image_pass(*args, n_depth_rt=2048)
    Io, Qo, Uo, To = oracle.image_pass(g, sp, pp, lines, np.pi / 4, 0.0, 32, vel_out, 24, 0.7, cloud.DX_IMAGE,
                                       n_depth_rt=2048, nthreads=NTH)
    Imax = np.abs(Io).max()
    assert Imax > 0
    np.testing.assert_allclose(I, Io, rtol=1e-10, atol=1e-10 * Imax)
    # Q and U on their own scales (max |Q| is a few 1e-3 of max |I| here)
    np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-10 * np.abs(Qo).max())
    np.testing.assert_allclose(U, Uo, rtol=0, atol=1e-10 * np.abs(Uo).max())
    np.testing.assert_allclose(T, To, rtol=1e-10, atol=1e-12)
    assert np.abs(Qo).max() > 1e-6 * Imax and np.abs(Uo).max() > 0  # polarisation is present


def test_image_pass_mixed_error(portal1, cloud4k, oracle_cloud4k):
    # FP32 Stokes transfer on the same populations: error on the scale of max |I|
    from portal_amd import driver
    sp, g, xth = cloud4k
    pp = driver.rescale_pol_pops(oracle_cloud4k[0], g.lev_p, sp.lev_g)
    lines = driver.image_lines(sp, [1, 2], [2, 3])
    vel_out = driver.channel_velocities(32, 93.75)
    args = (pp, lines, np.pi / 4, 0.0, 32, vel_out, 24, 0.7, cloud.DX_IMAGE)
    (I, Q, U, T), _ = portal1.image_pass(*args, n_depth_rt=2048)
    (Im, Qm, Um, Tm), _ = portal1.image_pass(*args, n_depth_rt=2048, precision=PORTAL_MIXED)
    s = np.abs(I).max()
    e = [np.abs(Im - I).max() / s, np.abs(Qm - Q).max() / s, np.abs(Um - U).max() / s]
    # Q and U on their own scales: the FP32 update keeps them free of I's rounding
    eo = [np.abs(Qm - Q).max() / np.abs(Q).max(), np.abs(Um - U).max() / np.abs(U).max()]
    print(f"mixed image: max |dI|, |dQ|, |dU| / max |I| = {e[0]:.2e} {e[1]:.2e} {e[2]:.2e}; "
          f"|dQ|, |dU| on their own scales {eo[0]:.2e} {eo[1]:.2e} (max |Q|/max |I| {np.abs(Q).max() / s:.1e}); "
          f"max rel dTAU {np.abs(Tm - T).max() / np.abs(T).max():.2e}")
    assert max(e) < 1e-5
    assert max(eo) < 2e-5
    np.testing.assert_allclose(Tm, T, rtol=1e-5, atol=1e-6 * np.abs(T).max())
    assert np.abs(Qm).max() > 0


@pytest.mark.parametrize("variant,waves", [(2, 2), (2, 4), (3, 4)])
def test_inward_pass_with_dust(oracle, variant, waves):
    # dust absorption and emission in the inward integration (inverse_rt.f90:273-297)
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=13)
    nb = g.bulk_node_num - 1
    g.dust_p[:nb] = 1e-17 * g.h2_dens[:nb] / 1e12
    sp.kd_v = np.array([3000.0, 6000.0, 9000.0])  # enough dust to change J by ~1e-2
    xth = cloud.legendre_all(5, 9)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pt.force_rt_kernel(PORTAL_FP64, variant, waves)
    pt.force_rt_kernel(PORTAL_MIXED, variant, waves)
    idx = solved(g)
    _, J0o, J2o, _, _ = oracle.se_pass(g, sp, xth, 16, cloud.DX_INWARD, extra_info=1, nthreads=NTH)
    J0, J2, _ = pt.inverse_rt(xth, 16, cloud.DX_INWARD)
    np.testing.assert_allclose(J0[:, idx], J0o[:, idx], rtol=1e-10, atol=0)
    np.testing.assert_allclose(J2[:, idx], J2o[:, idx], rtol=0, atol=1e-10 * np.abs(J0o[:, idx]).max(axis=0).min())
    J0m, J2m, _ = pt.inverse_rt(xth, 16, cloud.DX_INWARD, precision=PORTAL_MIXED)
    e0 = np.abs(J0m[:, idx] / J0o[:, idx] - 1).max()
    print(f"dust: mixed max rel J0 err {e0:.3e}")
    assert e0 < 2e-5
    # the dust changes J: without it the result differs
    kd = sp.kd_v.copy()
    sp.kd_v = np.zeros(3)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    J0n, _, _ = pt.inverse_rt(xth, 16, cloud.DX_INWARD)
    sp.kd_v = kd
    assert np.abs(J0n[:, idx] / J0[:, idx] - 1).max() > 1e-3
    pt.close()


@pytest.mark.parametrize("variant", [-1, 3])
def test_star_mode(oracle, variant):
    # star mode (main_sub.f90:56-79): per-node star_all directions, star rays ending on the
    # star with B_nu(T_star) behind them; nodes inside the star are not solved (variant
    # -1: the autotuner's choice, which must be a cooperative form)
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=42)
    star = dict(nth=6, nph=11, R_star=0.15, T_star=6000.0, X_star=np.zeros(3))
    r = np.linalg.norm(g.XYZ, axis=0)
    inside = np.where(r[: g.bulk_node_num] < 1.3 * star["R_star"])[0]
    g.ab_all[inside] = 0.0
    xth = cloud.legendre_all(6, 11)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pt.star = star
    if variant >= 0:
        pt.force_rt_kernel(PORTAL_FP64, variant, 4)
        pt.force_rt_kernel(PORTAL_MIXED, variant, 4)
    idx = solved(g)
    rng = np.random.default_rng(8)
    sub = np.sort(np.concatenate([idx[np.argsort(r[idx])[:8]], rng.choice(idx, 24, replace=False)])) + 1
    sub = np.unique(sub)
    J0o, J2o = oracle.inverse_nodes(g, sp, xth, 32, cloud.DX_INWARD, sub, star=star, nthreads=NTH)
    J0, J2, st = pt.inverse_rt(xth, 32, cloud.DX_INWARD)
    assert st["rt_variant"] in ((2, 3) if variant < 0 else (variant,))
    np.testing.assert_allclose(J0[:, sub - 1], J0o, rtol=1e-10)
    np.testing.assert_allclose(J2[:, sub - 1], J2o, rtol=0, atol=1e-10 * np.abs(J0o).max(axis=1).min())
    J0m, _, _ = pt.inverse_rt(xth, 32, cloud.DX_INWARD, precision=PORTAL_MIXED)
    e0 = np.abs(J0m[:, idx] / J0[:, idx] - 1).max()
    print(f"star: mixed max rel J0 err {e0:.3e}")
    assert e0 < 2e-5
    # the star is seen: without it J differs
    pt.star = None
    J0n, _, _ = pt.inverse_rt(xth, 32, cloud.DX_INWARD)
    assert np.abs(J0n[:, idx] / J0[:, idx] - 1).max() > 1e-3
    pt.star = star
    pp, st = pt.se_pass(xth, 32, cloud.DX_INWARD)
    w = np.sqrt(2 * sp.j_lev + 1.0)
    np.testing.assert_allclose(pp[idx][:, 0::2] @ w, 1.0, atol=1e-12)
    pt.close()


def test_image_pass_star(oracle):
    # star mode of the image pass (raytrace.f90:103-114): pixels on the rotated star's disc
    # start on its near side with B_nu(T_star); both precisions against the oracle
    from portal_amd import driver
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    N, nb = g.node_num, g.bulk_node_num - 1
    pp = np.zeros((N, 2 * sp.e_num), order="F")
    rng = np.random.default_rng(8)
    pp[:nb, 0::2] = g.lev_p[:, :nb].T
    pp[:nb, 3::2] = rng.uniform(-1e-3, 1e-3, (nb, sp.e_num - 1))
    pp = driver.rescale_pol_pops(pp, g.lev_p, sp.lev_g)
    star = dict(R_star=0.15, T_star=6000.0, X_star=np.array([0.12, -0.07, 0.2]))
    lines = driver.image_lines(sp, [1, 2], [2, 3])
    vel_out = driver.channel_velocities(8, 200.0)
    args = (pp, lines, 0.7, 0.4, 8, vel_out, 20, 0.7, cloud.DX_IMAGE)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    (I, Q, U, T), _ = pt.image_pass(*args, n_depth_rt=2048, star=star)
    (Im, Qm, Um, Tm), _ = pt.image_pass(*args, n_depth_rt=2048, star=star, precision=PORTAL_MIXED)
    (In, _, _, _), _ = pt.image_pass(*args, n_depth_rt=2048)
    pt.close()
    Io, Qo, Uo, To = oracle.image_pass(g, sp, pp, lines, 0.7, 0.4, 8, vel_out, 20, 0.7, cloud.DX_IMAGE,
                                       n_depth_rt=2048, nthreads=NTH, star=star)
    on = (I != In).any(axis=(1, 2))
    assert 20 < on.sum() < len(on) // 4  # pixels on the star's disc
    Imax = np.abs(Io).max()
    np.testing.assert_allclose(I, Io, rtol=1e-10, atol=1e-10 * Imax)
    np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-10 * np.abs(Qo).max())
    np.testing.assert_allclose(U, Uo, rtol=0, atol=1e-10 * np.abs(Uo).max())
    np.testing.assert_allclose(T, To, rtol=1e-10, atol=1e-12)
    # FP32 transfer: the star's light in units of its own B_nu, the line on the CMB's
    s_on, s_off = np.abs(I[on]).max(), np.abs(I[~on]).max()
    e_on = max(np.abs(Im - I)[on].max(), np.abs(Qm - Q)[on].max(), np.abs(Um - U)[on].max()) / s_on
    e_off = max(np.abs(Im - I)[~on].max(), np.abs(Qm - Q)[~on].max(), np.abs(Um - U)[~on].max()) / s_off
    print(f"star image: mixed max err {e_on:.2e} on the disc, {e_off:.2e} off it")
    assert e_on < 1e-5 and e_off < 1e-5


def test_image_pass_node_pixels(oracle):
    # the reference's pix = 0 mode (para_finish.f90:245-330): one pixel per bulk node at its
    # rotated (x, y), with and without the star, both precisions against the oracle
    from portal_amd import driver
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 2500, seed=41)
    N, nb = g.node_num, g.bulk_node_num - 1
    pp = np.zeros((N, 2 * sp.e_num), order="F")
    rng = np.random.default_rng(9)
    pp[:nb, 0::2] = g.lev_p[:, :nb].T
    pp[:nb, 3::2] = rng.uniform(-1e-3, 1e-3, (nb, sp.e_num - 1))
    pp = driver.rescale_pol_pops(pp, g.lev_p, sp.lev_g)
    lines = driver.image_lines(sp, [1, 2], [2, 3])
    vel_out = driver.channel_velocities(6, 250.0)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    for star in (None, dict(R_star=0.15, T_star=6000.0, X_star=np.array([0.12, -0.07, 0.2]))):
        args = (pp, lines, 1.1, 2.3, 6, vel_out, 0, 0.0, cloud.DX_IMAGE)
        (I, Q, U, T), _ = pt.image_pass(*args, n_depth_rt=2048, star=star, node_pixels=True)
        (Im, Qm, Um, _), _ = pt.image_pass(*args, n_depth_rt=2048, star=star, node_pixels=True,
                                           precision=PORTAL_MIXED)
        Io, Qo, Uo, To = oracle.image_pass(g, sp, pp, lines, 1.1, 2.3, 6, vel_out, 0, 0.0, cloud.DX_IMAGE,
                                           n_depth_rt=2048, nthreads=NTH, star=star, node_pixels=True)
        assert I.shape == (g.bulk_node_num, 2, 13)
        Imax = np.abs(Io).max()
        np.testing.assert_allclose(I, Io, rtol=1e-10, atol=1e-10 * Imax)
        np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-10 * np.abs(Qo).max())
        np.testing.assert_allclose(U, Uo, rtol=0, atol=1e-10 * np.abs(Uo).max())
        np.testing.assert_allclose(T, To, rtol=1e-10, atol=1e-12)
        e = max(np.abs(Im - I).max(), np.abs(Qm - Q).max(), np.abs(Um - U).max()) / Imax
        print(f"node pixels (star {star is not None}): mixed max err {e:.2e} of max |I|")
        assert e < 1e-5
    pt.close()


def test_image_pass_dust_branch(cloud4k, oracle):
    # dust present in the image pass: both branches agree
    from portal_amd import driver
    sp, g0, xth = cloud4k
    g = cloud.make_cloud(sp, 1500, seed=12)
    g.dust_p[: g.bulk_node_num - 1] = 1e-17 * g.h2_dens[: g.bulk_node_num - 1] / 1e12
    sp.kd_v = np.full(sp.r_num, 0.3)
    pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
    rng = np.random.default_rng(3)
    pp[: g.bulk_node_num - 1, 0::2] = 0.2
    pp[: g.bulk_node_num - 1, 3::2] = rng.uniform(-1e-3, 1e-3, (g.bulk_node_num - 1, sp.e_num - 1))
    pp = driver.rescale_pol_pops(pp, g.lev_p, sp.lev_g)
    lines = driver.image_lines(sp, [1, 2], [2, 3])
    vel_out = driver.channel_velocities(8, 300.0)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    (I, Q, U, T), dust, st = pt.image_pass(pp, lines, 1.1, 0.4, 8, vel_out, 10, 0.68, cloud.DX_IMAGE,
                                           n_depth_rt=2048, with_dust=True)
    (Io, Qo, Uo, To), dusto = oracle.image_pass(g, sp, pp, lines, 1.1, 0.4, 8, vel_out, 10, 0.68, cloud.DX_IMAGE,
                                                n_depth_rt=2048, nthreads=NTH, with_dust=True)
    sp.kd_v = np.zeros(sp.r_num)
    Imax = np.abs(Io).max()
    np.testing.assert_allclose(I, Io, rtol=1e-10, atol=1e-10 * Imax)
    np.testing.assert_allclose(Q, Qo, rtol=0, atol=1e-10 * np.abs(Qo).max())
    dmax = np.abs(dusto[0]).max()
    assert dmax > 0
    np.testing.assert_allclose(dust, dusto, rtol=1e-9, atol=1e-10 * dmax)
    pt.close()


def test_image_depth_overflow(portal1, cloud4k, oracle_cloud4k):
    from portal_amd import driver
    sp, g, xth = cloud4k
    pp = driver.rescale_pol_pops(oracle_cloud4k[0], g.lev_p, sp.lev_g)
    with pytest.raises(PortalError) as e:
        portal1.image_pass(pp, driver.image_lines(sp, [1], [2]), 0.3, 0.0, 2, driver.channel_velocities(2, 100.0),
                           4, 0.7, cloud.DX_IMAGE, n_depth_rt=3)
    assert e.value.code == -4  # PORTAL_ERR_DEPTH_RT


@pytest.mark.parametrize("precision,tol", [(PORTAL_FP64, 1e-10), (PORTAL_MIXED, 2e-5)])
def test_forty_line_species_at_129_channels(oracle, precision, tol):
    """LAMDA's CO size: 41 levels, 40 lines, 129 channels.  The cooperative kernel needs 8
    waves (12 slots) for 5,160 (line, channel) pairs; the autotuner chooses among the
    variants that take them.  J of 16 seeded nodes against the oracle, and the SE pass
    (81 unknowns at kmax = 2).  The lines above ~3 THz have backgrounds below FP32's
    range (the mixed path's per-line intensity unit keeps them), and lines in the Wien
    tail of the background take thin cells' emission, which the reference's FP64
    1 - e^-tau drops; the mixed path is checked against the oracle's exact form there."""
    sp = cloud.co_like_rotor(41, kmax=2)
    g = cloud.make_cloud(sp, 1500, seed=71)
    xth = cloud.legendre_all(4, 8)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    J0, J2, st = pt.inverse_rt(xth, 64, cloud.DX_INWARD, precision=precision)
    rng = np.random.default_rng(72)
    sub = np.sort(rng.choice(solved(g), 16, replace=False)) + 1
    J0o, J2o = oracle.inverse_nodes(g, sp, xth, 64, cloud.DX_INWARD, sub, nthreads=NTH,
                                    thin_exact=precision == PORTAL_MIXED)
    print(f"40 lines x 129 channels: variant {st['rt_variant']}/{st['rt_waves']}, "
          f"max rel J0 err {np.abs(J0[:, sub - 1] / J0o - 1).max():.2e}")
    np.testing.assert_allclose(J0[:, sub - 1], J0o, rtol=tol, atol=0)
    for k in range(16):
        assert np.abs(J2[:, sub[k] - 1] - J2o[:, k]).max() <= tol * np.abs(J0o[:, k]).max()
    if precision == PORTAL_FP64:
        pt.set_levpop(g.lev_p)
        pp, _ = pt.se_pass(xth, 64, cloud.DX_INWARD)
        lo = int(sub[0])
        ref = oracle.se_pass(g, sp, xth, 64, cloud.DX_INWARD, node_lo=lo, node_hi=lo + 5, nthreads=NTH)[0]
        idx = np.array([i for i in range(lo - 1, lo + 5) if g.ab_all[i] > 0])
        assert_pops_close(pp, ref, sp, idx, 1e-10)
    pt.close()


def test_forty_line_species_image(oracle):
    """Images of lines from 115 GHz to 3.5 THz of the 40-line species.  FP64 follows the
    reference's propagator; on lines 25 and 30 (2.9 and 3.5 THz at 20 K) that form drops
    the thin cells' emission, which exceeds the 2.73 K background there, and the mixed
    path's cancellation-free form keeps it: the mixed image is checked against the oracle's
    exact form on every line, FP64 against the reference's form where the two agree."""
    from portal_amd import driver
    sp = cloud.co_like_rotor(41, kmax=2)
    g = cloud.make_cloud(sp, 1500, seed=71)
    L = sp.e_num
    pp = np.zeros((g.node_num, 2 * L), order="F")
    for i in range(L):
        pp[:, 2 * i] = g.lev_p[i] / np.sqrt(sp.lev_g[i])
        if i:
            pp[:, 2 * i + 1] = 0.01 * pp[:, 2 * i]
    lines = driver.image_lines(sp, [1, 17, 21, 23, 25, 30], [2, 18, 22, 24, 26, 31], nu0_overwrite=False)
    vel = driver.channel_velocities(8, 150.0)
    args = (pp, lines, 0.6, 0.2, 8, vel, 10, 0.7, cloud.DX_IMAGE)
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    (I, Q, U, T), _ = pt.image_pass(*args, n_depth_rt=2048)
    (Im, Qm, Um, Tm), _ = pt.image_pass(*args, n_depth_rt=2048, precision=PORTAL_MIXED)
    pt.close()
    Io, Qo, Uo, To = oracle.image_pass(g, sp, *args, n_depth_rt=2048, nthreads=NTH)
    It, Qt, Ut, _ = oracle.image_pass(g, sp, *args, n_depth_rt=2048, nthreads=NTH, thin_exact=True)
    assert np.isfinite(Im).all() and np.isfinite(Qm).all() and np.isfinite(Um).all()
    for j in range(len(lines)):
        sI, sQ, sU = (np.abs(x[:, j]).max() for x in (It, Qt, Ut))
        e = [np.abs(Im[:, j] - It[:, j]).max() / sI, np.abs(Qm[:, j] - Qt[:, j]).max() / sQ,
             np.abs(Um[:, j] - Ut[:, j]).max() / sU]
        print(f"line {j} ({lines[j][2] / 1e9:.0f} GHz): max |I| {sI:.2e}; mixed against the exact form "
              f"{e[0]:.1e} {e[1]:.1e} {e[2]:.1e} (each on its own scale); "
              f"reference form off by {np.abs(Io[:, j] - It[:, j]).max() / sI:.1e}")
        assert sI > 0 and sQ > 0
        assert max(e) < 2e-5
        if j < 3:
            np.testing.assert_allclose(I[:, j], Io[:, j], rtol=1e-10, atol=1e-10 * sI)
            np.testing.assert_allclose(Q[:, j], Qo[:, j], rtol=0, atol=1e-10 * np.abs(Qo[:, j]).max())
    np.testing.assert_allclose(Tm, T, rtol=1e-5, atol=1e-6 * np.abs(T).max())
    np.testing.assert_allclose(T, To, rtol=1e-10, atol=1e-12)
