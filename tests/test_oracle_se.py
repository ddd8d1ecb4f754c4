"""Statistical-equilibrium oracle pinned by invariants and closed forms.

pop_eqs.f90 / em_coefs.f90 cannot be compiled here (they need the WIGXJPF
library), so the oracle's SE is pinned by: the normalisation row, isotropic
radiation => no alignment, detailed balance => Boltzmann populations, the
collision-dominated limit => LTE, and an independent dense solve of the
oracle's own assembled matrix.  The coupling constants themselves are pinned
against exact sympy sums in test_oracle_golden.py.
"""
import math

import numpy as np
import pytest

from portal_amd import cloud

H_OVER_K = 4.79924e-11  # the reference's h/k_B [K/Hz] (pop_eqs.f90:535)


def planck(T, nu):
    return 1.4744995e-50 * nu**3 / math.expm1(H_OVER_K * nu / T)


def levels_from_rho(sp, rho):
    # sqrt(2J+1) rho^0_0 = fractional population
    return np.array([math.sqrt(2 * sp.j_lev[i] + 1) * rho[sp.count_array[i]] for i in range(sp.e_num)])


def ac_for(oracle, sp, nH2, T, J0, J2):
    return oracle.make_ac(sp, nH2, T, J0, J2)


@pytest.mark.parametrize("nlev", [4, 10])
def test_normalisation_and_isotropy(oracle, nlev):
    sp = cloud.co_like_rotor(nlev)
    rng = np.random.default_rng(3)
    J0 = 10 ** rng.uniform(-18, -16, sp.r_num)
    rho = oracle.get_rhovec(sp, ac_for(oracle, sp, 3e11, 25.0, J0, np.zeros(sp.r_num)), 25.0)
    assert levels_from_rho(sp, rho).sum() == pytest.approx(1.0, abs=1e-14)
    # isotropic radiation: every K>0 component vanishes
    k0 = set(sp.count_array.tolist())
    others = [rho[i] for i in range(sp.c_dim) if i not in k0]
    assert max(abs(x) for x in others) < 1e-15


@pytest.mark.parametrize("nlev,T,nH2", [(4, 15.0, 1e9), (4, 40.0, 1e13), (10, 30.0, 1e12)])
def test_detailed_balance_gives_boltzmann(oracle, nlev, T, nH2):
    sp = cloud.co_like_rotor(nlev)
    # radiation = Planck at T on every line, with the SE's own line frequencies (E_u - E_l)
    nu = np.array([sp.e_levs_hz[sp.r_trans[1, k] - 1] - sp.e_levs_hz[sp.r_trans[0, k] - 1] for k in range(sp.r_num)])
    J0 = np.array([planck(T, v) for v in nu])
    rho = oracle.get_rhovec(sp, ac_for(oracle, sp, nH2, T, J0, np.zeros(sp.r_num)), T)
    pops = levels_from_rho(sp, rho)
    g = sp.lev_g.astype(float)
    boltz = g * np.exp(-H_OVER_K * (sp.e_levs_hz - sp.e_levs_hz[0]) / T)
    boltz /= boltz.sum()
    # 1.4744995e-50 vs 1/6.781962e49 agree to ~1e-7
    np.testing.assert_allclose(pops, boltz, rtol=2e-6)


def test_collision_limit_is_lte_and_unaligned(oracle):
    sp = cloud.co_like_rotor(6)
    T = 22.0
    rng = np.random.default_rng(5)
    J0 = 10 ** rng.uniform(-19, -17, sp.r_num)
    J2 = J0 * rng.uniform(-0.1, 0.1, sp.r_num)
    rho = oracle.get_rhovec(sp, ac_for(oracle, sp, 1e22, T, J0, J2), T)
    pops = levels_from_rho(sp, rho)
    g = sp.lev_g.astype(float)
    boltz = g * np.exp(-H_OVER_K * (sp.e_levs_hz - sp.e_levs_hz[0]) / T)
    boltz /= boltz.sum()
    np.testing.assert_allclose(pops, boltz, rtol=1e-6)
    for i in range(1, sp.e_num):
        assert abs(rho[sp.count_array[i] + 1]) < 1e-8 * abs(rho[sp.count_array[i]])


def test_alignment_follows_anisotropy_sign(oracle):
    # with J2 -> -J2 the linear response of every K=2 component changes sign
    sp = cloud.co_like_rotor(4)
    T = 20.0
    J0 = np.array([3e-18, 2e-18, 1e-18])
    J2 = 1e-3 * J0
    r_p = oracle.get_rhovec(sp, ac_for(oracle, sp, 1e8, T, J0, J2), T)
    r_m = oracle.get_rhovec(sp, ac_for(oracle, sp, 1e8, T, J0, -J2), T)
    for i in range(1, sp.e_num):
        a, b = r_p[sp.count_array[i] + 1], r_m[sp.count_array[i] + 1]
        assert a != 0 and a == pytest.approx(-b, rel=1e-2)


@pytest.mark.parametrize("nlev", [4, 10])
def test_solve_equals_dense_solve_of_assembled_matrix(oracle, nlev):
    sp = cloud.co_like_rotor(nlev)
    rng = np.random.default_rng(11)
    J0 = 10 ** rng.uniform(-18, -16, sp.r_num)
    J2 = J0 * rng.uniform(-0.05, 0.05, sp.r_num)
    T = 18.0
    AC = ac_for(oracle, sp, 5e10, T, J0, J2)
    A, Am, Bm, Cm = oracle.setup_stateq(sp, AC, T)
    # the class matrices add up to the full matrix
    np.testing.assert_allclose(Am + Bm + Cm, A, rtol=1e-12, atol=1e-12 * np.abs(A).max())
    A = A.copy()
    A[0, :] = 0
    for i in range(sp.e_num):
        A[0, sp.count_array[i]] = math.sqrt(2 * sp.j_lev[i] + 1)
    b = np.zeros(sp.c_dim)
    b[0] = 1
    ref = np.linalg.solve(A, b)
    rho = oracle.get_rhovec(sp, AC, T)
    np.testing.assert_allclose(rho, ref, rtol=1e-9, atol=1e-13)
    # residual of the extended-precision solve
    assert np.abs(A @ rho - b).max() < 1e-15 * np.abs(A).max()


def test_rate_columns_conserve(oracle):
    # without the normalisation row every column of the rate matrix sums to zero
    # over the K=0 rows weighted by sqrt(2J+1) (population conservation)
    sp = cloud.co_like_rotor(5)
    J0 = np.array([3e-18, 2e-18, 1e-18, 5e-19])
    T = 30.0
    A, *_ = oracle.setup_stateq(sp, ac_for(oracle, sp, 1e10, T, J0, 0.02 * J0), T)
    w = np.zeros(sp.c_dim)
    for i in range(sp.e_num):
        w[sp.count_array[i]] = math.sqrt(2 * sp.j_lev[i] + 1)
    col = w @ A
    assert np.abs(col).max() < 1e-10 * np.abs(A).max()


def test_singular_system_is_reported(oracle):
    """A level with no transition leaves its rows of the SE matrix empty: the solve reports
    the singular matrix (em_coefs.f90:170-178 stops the reference there) rather than
    returning numbers."""
    ct = np.array([1.0, 10.0, 100.0, 1e3, 1e4])
    sp = cloud.Species(e_num=3, j_lev=np.array([0, 1, 2], dtype=np.int32), e_levs_cm=np.array([0.0, 3.845, 11.535]),
                       mass_mol=28.0, r_trans=np.array([[1], [2]], dtype=np.int32), t_Aji=np.array([7.2e-8]),
                       nu_ghz=np.array([115.27]), c_trans=np.array([[1], [2]], dtype=np.int32), col_temps=ct,
                       t_Cji=np.asfortranarray(np.full((5, 1, 1), 1e-11)), kmax=2)
    AC = oracle.make_ac(sp, 1e12, 20.0, np.array([1e-17]), np.array([1e-19]))
    with pytest.raises(RuntimeError):
        oracle.get_rhovec(sp, AC, 20.0)
