"""The oracle's statistical equilibrium and emission coefficients against an
independent model in the magnetic-sublevel basis.

The reference works in the coherence-free limit (only Q = 0 multipoles), where the
density matrix of a level is diagonal in M along the field.  There the
irreducible-tensor equations of setup_stateq (pop_eqs.f90:361-653) and the Landi
(1984) coefficients of comp_emcoefs (em_coefs.f90:1-113) are equivalent to plain
rate equations and line-emission sums over the Zeeman components M_u -> M_l:

* a component with p = M_u - M_l has strength (2J+1) (J_u J_l 1; -M_u M_l p)^2 of its
  line's A, B_ul (from the upper level) or B_lu (from the lower level), so that every
  sublevel's total rates are the level's;
* it sees the radiation of its own polarisation, J_pi = J^0_0 - sqrt(2) J^2_0 for
  p = 0 and J_sigma = J^0_0 + J^2_0 / sqrt(2) for p = +-1 (the dipole patterns
  (3/2) sin^2 and (3/4)(1 + cos^2) averaged over the field
  J^2_0 = sum w I (3 mu^2 - 1) / (2 sqrt 2));
* isotropic collisions move a sublevel's population evenly to the other level's
  sublevels (pop_eqs.f90:467-471: rank-0 transfer only; every rank relaxes);
* a pi component emits (3/2) sin^2 theta with Q = +I, a sigma component
  (3/4)(1 + cos^2 theta) with Q = -(3/4) sin^2 theta (the reference's sign: its
  T^2_0(Q) = -3 sin^2 theta / (2 sqrt 2), em_coefs.f90:57-58).

The sublevel populations are converted to rho^K_0 with
rho^K_0 = sum_M (-1)^(J-M) sqrt(2K+1) (J J K; M -M 0) N_M, the 3j symbols taken
exactly from sympy, so nothing here shares code or coefficients with the oracle
(which uses its own Racah sums and the L&L04 7.20 constants).  A wrong 3j/6j
argument, phase, sqrt(2J+1) factor, sign of the J^2_0 coupling, stimulated-emission
term or T^2_0 factor fails these checks.
"""
import functools
import math

import numpy as np
import pytest
from sympy.physics.wigner import wigner_3j

from portal_amd import cloud

A_TO_B = 6.781962e49   # c^2 / 2h as the reference writes it (pop_eqs.f90:419, em_coefs.f90:79)
H_K = 4.79924e-11      # h / k_B as the reference writes it (pop_eqs.f90:535)
HV = 5.273e-35         # h / 4 pi as comp_emcoefs writes it (em_coefs.f90:99)
# CODATA 2018, for the checks against physical law
H_SI, K_SI, C_SI = 6.62607015e-34, 1.380649e-23, 2.99792458e8


@functools.lru_cache(maxsize=None)
def w3(j1, j2, j3, m1, m2, m3):
    return float(wigner_3j(j1, j2, j3, m1, m2, m3))


def strength(Ju, Jl, Mu, Ml):
    p = Mu - Ml
    return w3(Ju, Jl, 1, -Mu, Ml, p) ** 2 if abs(p) <= 1 else 0.0


def rho_from_sublevels(J, N, kmax):
    """rho^K_0, K = 0, 2, .., min(2J, kmax), of sublevel populations N[M + J]."""
    return [sum((-1) ** (J - M) * math.sqrt(2 * K + 1) * w3(J, J, K, M, -M, 0) * N[M + J] for M in range(-J, J + 1))
            for K in range(0, min(2 * J, kmax) + 1, 2)]


def make_species(J, E_cm, rad, kmax=6):
    r_trans = np.array(rad, dtype=np.int32).T.copy()
    n = len(J)
    pairs = [(l, u) for u in range(2, n + 1) for l in range(1, u)]
    ct = np.array([1.0, 10.0, 100.0, 1e3, 1e4])
    R = r_trans.shape[1]
    return cloud.Species(e_num=n, j_lev=np.array(J, dtype=np.int32), e_levs_cm=np.array(E_cm, float),
                         mass_mol=28.0, r_trans=r_trans, t_Aji=np.full(R, 1e-6), nu_ghz=np.full(R, 100.0),
                         c_trans=np.array(pairs, dtype=np.int32).T.copy(), col_temps=ct,
                         t_Cji=np.asfortranarray(np.full((len(ct), len(pairs), 1), 1e-11)), kmax=kmax)


def sublevel_se(sp, AC, T, depol=None):
    """Steady state of the sublevel rate equations, normalised to total population 1,
    returned as rho^K_0 in the SE vector's layout (count_array slots).  depol: None, or
    {level index: {K: D}}, an isotropic elastic process that relaxes rank K of that level
    at rate D and leaves its other ranks alone: the operator -sum_K D v_K v_K^T on the
    level's sublevels, v_K(M) = (-1)^(J-M) sqrt(2K+1) (J J K; M -M 0) (orthonormal, and
    rho^K_0 = v_K . N)."""
    J = [int(j) for j in sp.j_lev]
    states = [(i, M) for i in range(sp.e_num) for M in range(-J[i], J[i] + 1)]
    at = {s: k for k, s in enumerate(states)}
    n = len(states)
    R = np.zeros((n, n))  # R[to, from]
    for q in range(sp.n_trans):
        l, u = int(sp.all_trans[0, q]) - 1, int(sp.all_trans[1, q]) - 1
        A, C21, nu, J0, J2 = AC[:, q]
        gl, gu = 2 * J[l] + 1, 2 * J[u] + 1
        B21 = A_TO_B * A / nu ** 3
        B12 = B21 * gu / gl
        C12 = C21 * gu * math.exp(-H_K * nu / T) / gl
        J_pi, J_sigma = J0 - math.sqrt(2.0) * J2, J0 + J2 / math.sqrt(2.0)
        for Ml in range(-J[l], J[l] + 1):
            for Mu in range(-J[u], J[u] + 1):
                s = strength(J[u], J[l], Mu, Ml)
                Jp = J_pi if Mu == Ml else J_sigma
                a, b = at[(l, Ml)], at[(u, Mu)]
                R[b, a] += B12 * gl * s * Jp + C12 / gu
                R[a, b] += A * gu * s + B21 * gu * s * Jp + C21 / gl
    M = R - np.diag(R.sum(axis=0))
    for i, ranks in (depol or {}).items():
        idx = [at[(i, m)] for m in range(-J[i], J[i] + 1)]
        for K, D in ranks.items():
            v = np.array([(-1) ** (J[i] - m) * math.sqrt(2 * K + 1) * w3(J[i], J[i], K, m, -m, 0)
                          for m in range(-J[i], J[i] + 1)])
            M[np.ix_(idx, idx)] -= D * np.outer(v, v)
    M[0, :] = 1.0
    rhs = np.zeros(n)
    rhs[0] = 1.0
    N = np.linalg.solve(M, rhs)
    rho = np.zeros(sp.c_dim)
    for i in range(sp.e_num):
        v = rho_from_sublevels(J[i], [N[at[(i, m)]] for m in range(-J[i], J[i] + 1)], sp.kmax)
        rho[sp.count_array[i]: sp.count_array[i] + len(v)] = v
    return rho


def random_AC(sp, rad, E_cm, rng, aniso=(-0.5, 0.3), weak=False):
    AC = np.zeros((5, sp.n_trans), order="F")
    rads = {tuple(r) for r in rad}
    for q in range(sp.n_trans):
        l, u = int(sp.all_trans[0, q]), int(sp.all_trans[1, q])
        nu = (E_cm[u - 1] - E_cm[l - 1]) * cloud.CM1_TO_HZ
        on = (l, u) in rads
        J0 = (10 ** rng.uniform(-18, -15) if not weak else 1e-26) if on else 0.0
        # J_pi, J_sigma >= 0 needs -J0/sqrt(2) <= ... ; aniso within (-1/sqrt2, 1/sqrt2)
        AC[:, q] = (10 ** rng.uniform(-7, -4) if on else 0.0, 10 ** rng.uniform(-8, -5), nu, J0,
                    J0 * rng.uniform(*aniso))
    return AC


# ladders with Delta J = 0 and +-1 lines, J up to 3 so that kmax = 6 truncates nothing
LADDERS = [
    ([0, 1], [(1, 2)]),
    ([1, 0, 1], [(1, 2), (2, 3), (1, 3)]),
    ([0, 1, 2, 1, 3], [(1, 2), (2, 3), (1, 4), (2, 4), (3, 5), (3, 4)]),
    ([2, 3, 2, 1], [(1, 2), (1, 3), (2, 3), (3, 4), (1, 4)]),
]


@pytest.mark.parametrize("case", range(len(LADDERS)))
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_se_equals_sublevel_rate_equations(oracle, case, seed):
    """setup_stateq + get_rhovec against the sublevel rate equations: every rank K = 0..6
    of every level, with stimulated emission, Delta J = 0 lines and collisions."""
    J, rad = LADDERS[case]
    rng = np.random.default_rng(100 * case + seed)
    E = np.concatenate([[0.0], np.sort(rng.uniform(1.0, 60.0, len(J) - 1))])
    sp = make_species(J, E, rad)
    AC = random_AC(sp, rad, E, rng)
    T = float(rng.uniform(10, 80))
    got = oracle.get_rhovec(sp, AC, T)
    ref = sublevel_se(sp, AC, T)
    # every component of every level on the scale of that level's rho^0_0
    for i in range(sp.e_num):
        c0 = sp.count_array[i]
        n = len(range(0, min(2 * J[i], sp.kmax) + 1, 2))
        scale = abs(ref[c0])
        np.testing.assert_allclose(got[c0:c0 + n], ref[c0:c0 + n], rtol=0, atol=1e-10 * scale)
    # the alignment is not trivially small in these cases
    assert max(abs(ref[sp.count_array[i] + 1] / ref[sp.count_array[i]]) for i in range(sp.e_num) if J[i] > 0) > 1e-3


def test_se_truncated_at_kmax_keeps_low_ranks_close(oracle):
    """kmax = 2 drops rank 4 and 6 of J >= 2 levels; the ranks kept stay close to the
    untruncated sublevel solution (the dropped ranks feed back only at second order)."""
    J, rad = LADDERS[3]
    rng = np.random.default_rng(7)
    E = np.concatenate([[0.0], np.sort(rng.uniform(1.0, 60.0, len(J) - 1))])
    sp6 = make_species(J, E, rad, kmax=6)
    sp2 = make_species(J, E, rad, kmax=2)
    AC = random_AC(sp6, rad, E, rng, aniso=(-0.05, 0.05))
    ref = sublevel_se(sp6, AC, 30.0)
    got = oracle.get_rhovec(sp2, AC, 30.0)
    for i in range(sp2.e_num):
        a, b = sp6.count_array[i], sp2.count_array[i]
        assert abs(got[b] - ref[a]) < 1e-4 * ref[a]
        assert abs(got[b + 1] - ref[a + 1]) < 2e-2 * abs(ref[a + 1]) + 1e-9 * ref[a]


def test_two_level_weak_radiation_limit(oracle):
    """J = 0 -> 1, no collisions, B J << A: rho^2_0 / rho^0_0 of the upper level is
    w^(2)_{10} J^2_0 / J^0_0 with w^(2)_{10} = 1 (Landi Degl'Innocenti & Landolfi 2004,
    the two-level atom with an unpolarisable lower level)."""
    sp = make_species([0, 1], [0.0, 3.0], [(1, 2)])
    nu = 3.0 * cloud.CM1_TO_HZ
    for ratio in (0.1, -0.3, 0.5):
        AC = np.zeros((5, 1), order="F")
        AC[:, 0] = (1e-6, 0.0, nu, 1e-26, ratio * 1e-26)
        rho = oracle.get_rhovec(sp, AC, 20.0)
        assert rho[2] / rho[1] == pytest.approx(ratio, rel=1e-8)


# --------------------------------------------------------------------------------------
# emission and absorption coefficients (comp_emcoefs, em_coefs.f90:1-113)

def sublevel_coefficients(Jl, Ju, Nl, Nu, th, nu, A):
    """eta_I, eta_Q, eps_I, eps_Q of the Zeeman components of one line."""
    gl, gu = 2 * Jl + 1, 2 * Ju + 1
    B21 = A_TO_B * A / nu ** 3
    B12 = B21 * gu / gl
    s2, c2 = math.sin(th) ** 2, math.cos(th) ** 2
    fI = {0: 1.5 * s2, 1: 0.75 * (1 + c2), -1: 0.75 * (1 + c2)}
    fQ = {0: 1.5 * s2, 1: -0.75 * s2, -1: -0.75 * s2}
    eI = eQ = aI = aQ = 0.0
    for Mu in range(-Ju, Ju + 1):
        for Ml in range(-Jl, Jl + 1):
            p = Mu - Ml
            if abs(p) > 1:
                continue
            s = strength(Ju, Jl, Mu, Ml)
            up = Nu[Mu + Ju] * gu * s
            lo = Nl[Ml + Jl] * gl * s
            eI += A * up * fI[p]
            eQ += A * up * fQ[p]
            aI += B12 * lo * fI[p] - B21 * up * fI[p]
            aQ += B12 * lo * fQ[p] - B21 * up * fQ[p]
    h = HV * nu
    return h * aI, h * aQ, h * eI, h * eQ


def sym_pops(J, rng):
    half = rng.uniform(0.2, 1.0, J + 1)
    return np.concatenate([half[:0:-1], half])  # N(M) = N(-M): no orientation


@pytest.mark.parametrize("Jl,Ju", [(0, 1), (1, 0), (1, 1), (1, 2), (2, 1), (2, 2), (2, 3), (3, 2), (3, 3)])
def test_emcoefs_equal_zeeman_component_sums(oracle, Jl, Ju):
    rng = np.random.default_rng(10 * Jl + Ju)
    for _ in range(4):
        Nl, Nu = sym_pops(Jl, rng), 0.3 * sym_pops(Ju, rng)
        th = float(rng.uniform(0.05, math.pi - 0.05))
        nu, A = 2.3e11, 6.9e-7
        rl = rho_from_sublevels(Jl, Nl, 2) + [0.0]
        ru = rho_from_sublevels(Ju, Nu, 2) + [0.0]
        got = oracle.comp_emcoefs(rl[0], ru[0], rl[1], ru[1], Jl, Ju, th, nu, A)
        ref = sublevel_coefficients(Jl, Ju, Nl, Nu, th, nu, A)
        sI, sE = abs(ref[0]) + HV * nu * A_TO_B * A / nu ** 3 * 1e-3, abs(ref[2])
        assert got[0] == pytest.approx(ref[0], rel=1e-12, abs=1e-13 * sI)
        assert got[1] == pytest.approx(ref[1], rel=1e-11, abs=1e-13 * sI)
        assert got[2] == pytest.approx(ref[2], rel=1e-12)
        assert got[3] == pytest.approx(ref[3], rel=1e-11, abs=1e-13 * sE)


def w2_closed_form(J, Jp):
    """w^(2)_{J J'} (Landi Degl'Innocenti & Landolfi 2004, Table 10.1)."""
    if Jp == J - 1:
        return math.sqrt((J + 1) * (2 * J + 3) / (10 * J * (2 * J - 1)))
    if Jp == J:
        return -math.sqrt((2 * J - 1) * (2 * J + 3) / (10 * J * (J + 1)))
    return math.sqrt(J * (2 * J - 1) / (10 * (J + 1) * (2 * J + 3)))


@pytest.mark.parametrize("Jl,Ju", [(1, 2), (2, 1), (1, 1), (2, 2), (2, 3), (3, 2), (4, 5), (9, 10)])
def test_emcoefs_w2_closed_form(oracle, Jl, Ju):
    """At theta = 0, eps_I / eps_I(rho^2 = 0) = 1 + w^(2)_{J_u J_l} sigma / sqrt(2), and the
    same for eta with the lower level (stimulated emission off: rho_u = 0 in eta)."""
    nu, A, th = 2.3e11, 1e-6, 0.0
    s = 0.1
    e0 = oracle.comp_emcoefs(1.0, 1.0, 0.0, 0.0, Jl, Ju, th, nu, A)
    e1 = oracle.comp_emcoefs(1.0, 1.0, 0.0, s, Jl, Ju, th, nu, A)
    w_u = (e1[2] / e0[2] - 1.0) * math.sqrt(2.0) / s
    assert w_u == pytest.approx(w2_closed_form(Ju, Jl), rel=1e-12)
    # an upper level of 1e-300 (0/0 would make eta NaN, as in the reference)
    a0 = oracle.comp_emcoefs(1.0, 1e-300, 0.0, 0.0, Jl, Ju, th, nu, A)
    a1 = oracle.comp_emcoefs(1.0, 1e-300, s, 0.0, Jl, Ju, th, nu, A)
    w_l = (a1[0] / a0[0] - 1.0) * math.sqrt(2.0) / s
    assert w_l == pytest.approx(w2_closed_form(Jl, Ju), rel=1e-12)


def test_emcoefs_pi_and_sigma_light(oracle):
    """J_u = 1 -> J_l = 0 seen at 90 degrees: only M_u = 0 populated emits pure pi light
    (Q/I = +1), only M_u = +-1 pure sigma light (Q/I = -1); along the field Q = 0.
    The same for absorption from a J_l = 1 level towards J_u = 0."""
    nu, A = 1.15e11, 7.2e-8
    for Nu, sign in (([0.0, 1.0, 0.0], 1.0), ([0.5, 0.0, 0.5], -1.0)):
        ru = rho_from_sublevels(1, Nu, 2)
        aI, aQ, eI, eQ = oracle.comp_emcoefs(1.0, ru[0], 0.0, ru[1], 0, 1, math.pi / 2, nu, A)
        assert eQ / eI == pytest.approx(sign, rel=1e-14)
        # along the field: no Q (and no pi light at all: eps_I = 0 there for M_u = 0 alone)
        _, _, eI0, eQ0 = oracle.comp_emcoefs(1.0, ru[0], 0.0, ru[1], 0, 1, 0.0, nu, A)
        assert abs(eQ0) < 1e-15 * abs(eI)
        assert abs(eI0) < 1e-15 * abs(eI) if sign > 0 else eI0 == pytest.approx(2 * eI, rel=1e-14)
        # the same populations in the lower level of a 1 -> 0 line, upper level empty
        aI, aQ, _, _ = oracle.comp_emcoefs(ru[0], 1e-300, ru[1], 0.0, 1, 0, math.pi / 2, nu, A)
        assert aQ / aI == pytest.approx(sign, rel=1e-14)


@pytest.mark.parametrize("Jl,Ju,T", [(0, 1, 20.0), (3, 4, 45.0), (2, 2, 8.0)])
def test_emcoefs_kirchhoff_and_einstein_relations(oracle, Jl, Ju, T):
    """Unpolarised Boltzmann populations: eps_I / eta_I is Planck's B_nu(T) (CODATA), the
    emission is h nu / 4 pi A N_u, and Q vanishes."""
    nu, A = 3.45e11, 2.5e-6
    gl, gu = 2 * Jl + 1, 2 * Ju + 1
    Nl = 0.6
    Nu = Nl * gu / gl * math.exp(-H_SI * nu / (K_SI * T))
    aI, aQ, eI, eQ = oracle.comp_emcoefs(Nl / math.sqrt(gl), Nu / math.sqrt(gu), 0.0, 0.0, Jl, Ju, 0.7, nu, A)
    planck = 2 * H_SI * nu ** 3 / C_SI ** 2 / math.expm1(H_SI * nu / (K_SI * T))
    # the reference's c^2/2h (6.781962e49) and h/4pi (5.273e-35) are rounded constants
    assert eI / aI == pytest.approx(planck, rel=1e-6)
    assert eI == pytest.approx(H_SI / (4 * math.pi) * nu * A * Nu, rel=5e-5)
    assert aQ == 0.0 and eQ == 0.0


# --------------------------------------------------------------------------------------
# collisional depolarisation (depol.f90:242-359; the slot pop_eqs.f90:642-646)

def _depol_arrays(sp, J, rng, scale):
    """Random rank-dependent rates: the sublevel model's dict and the oracle's D0
    (e_num x (kmax - 1), column K - 2)."""
    D0 = np.zeros((sp.e_num, sp.kmax - 1), order="F")
    ranks = {}
    for i in range(sp.e_num):
        for K in range(2, min(2 * J[i], sp.kmax) + 1, 2):
            d = scale * 10 ** rng.uniform(-1, 1)
            D0[i, K - 2] = d
            ranks.setdefault(i, {})[K] = d
    return D0, ranks


@pytest.mark.parametrize("case", range(len(LADDERS)))
def test_depolarisation_relaxes_each_rank(oracle, case):
    """setup_stateq with D^(K) on the diagonal of the K >= 2 rows against the sublevel
    equations with a process that relaxes exactly rank K at rate D^(K): every rank of
    every level, with rates of each rank chosen independently (a D placed in the wrong
    row, with the wrong sign or on the K = 0 row fails)."""
    J, rad = LADDERS[case]
    rng = np.random.default_rng(900 + case)
    E = np.concatenate([[0.0], np.sort(rng.uniform(1.0, 60.0, len(J) - 1))])
    sp = make_species(J, E, rad)
    AC = random_AC(sp, rad, E, rng)
    T = float(rng.uniform(10, 80))
    base = sublevel_se(sp, AC, T)
    A_scale = float(np.max(AC[0]))
    D0, ranks = _depol_arrays(sp, J, rng, A_scale)
    got = oracle.get_rhovec(sp, AC, T, D0=D0)
    ref = sublevel_se(sp, AC, T, depol=ranks)
    for i in range(sp.e_num):
        c0 = sp.count_array[i]
        n = len(range(0, min(2 * J[i], sp.kmax) + 1, 2))
        np.testing.assert_allclose(got[c0:c0 + n], ref[c0:c0 + n], rtol=0, atol=1e-10 * abs(ref[c0]))
    # the rates matter here: the alignment moved
    moved = max(abs(ref[sp.count_array[i] + 1] - base[sp.count_array[i] + 1]) / ref[sp.count_array[i]]
                for i in range(sp.e_num) if J[i] > 0)
    assert moved > 1e-4
    # D0 = 0 is the reference's own case (main_sub.f90:96): bit for bit the plain solve
    np.testing.assert_array_equal(oracle.get_rhovec(sp, AC, T, D0=0 * D0), oracle.get_rhovec(sp, AC, T))


def test_strong_depolarisation_removes_alignment(oracle):
    """rho^2_0 / rho^0_0 falls as 1/D once D dominates the radiative rates of a level."""
    J, rad = LADDERS[2]
    rng = np.random.default_rng(5)
    E = np.concatenate([[0.0], np.sort(rng.uniform(1.0, 60.0, len(J) - 1))])
    sp = make_species(J, E, rad, kmax=2)
    AC = random_AC(sp, rad, E, rng)
    al = []
    for D in (1e2, 1e3, 1e4):
        D0 = np.zeros((sp.e_num, 1), order="F")
        D0[[i for i in range(sp.e_num) if J[i] > 0], 0] = D
        rho = oracle.get_rhovec(sp, AC, 30.0, D0=D0)
        al.append(np.array([rho[sp.count_array[i] + 1] / rho[sp.count_array[i]] for i in range(sp.e_num) if J[i] > 0]))
    np.testing.assert_allclose(al[1] * 1e3, al[0] * 1e2, rtol=2e-2)
    np.testing.assert_allclose(al[2] * 1e4, al[1] * 1e3, rtol=2e-3)


def _rotor_with_rates(nlev, rng, c_par=1):
    sp = cloud.co_like_rotor(nlev)
    assert sp.kmax == 6
    ct = sp.col_temps
    C = np.zeros((len(ct), sp.c_num, c_par), order="F")
    for c in range(sp.c_num):
        for p in range(c_par):
            C[:, c, p] = 10 ** rng.uniform(-12, -10) * (ct / ct[0]) ** rng.uniform(-0.5, 0.5)
    sp.t_Cji = C
    return sp


def test_depolarisation_table_closed_forms(oracle):
    """de_polarization's Morris (1985) B11 sum: J = 1 has only m = 2, (1 1 2; 0 0 0)^2 =
    2/15, {1 1 2; 1 1 0} = 1/3, {1 1 2; 1 1 2} = 1/30, so D^(2)(1) = 9 * 5 * (2/15) *
    (1/3 - 1/30) k_2 = (9/5) k_2 with k_2 the J = 2 <-> 0 rate; J = 2, 3 against the sum
    with sympy's exact 3j/6j; zero for J = 0, odd K and K > 2J; the two-partner mix."""
    from sympy.physics.wigner import wigner_6j
    rng = np.random.default_rng(11)
    for c_par in (1, 2):
        sp = _rotor_with_rates(5, rng, c_par)
        D = oracle.depol_table(sp)
        assert D.shape == (sp.c_temp, sp.e_num, sp.kmax - 1)

        def k(Jm):
            col = [c for c in range(sp.c_num) if set(sp.c_trans[:, c]) == {1, Jm + 1}][0]
            t = sp.t_Cji[:, col, :]
            return t[:, 0] if c_par == 1 else 0.25 * t[:, 0] + 0.75 * t[:, 1]

        np.testing.assert_allclose(D[:, 1, 0], 1.8 * k(2), rtol=1e-13)
        for Jl in (2, 3):
            for K in range(2, 2 * Jl + 1, 2):
                exp = np.zeros(sp.c_temp)
                for m in range(2, 2 * Jl + 1, 2):
                    if m > 4:
                        continue  # the 5-level rotor has no J = 6 level: no such transition
                    w = float(wigner_3j(Jl, Jl, m, 0, 0, 0)) ** 2
                    f = float(wigner_6j(Jl, Jl, m, Jl, Jl, 0) - wigner_6j(Jl, Jl, m, Jl, Jl, K))
                    exp += (2 * Jl + 1) ** 2 * (2 * m + 1) * w * f * k(m)
                np.testing.assert_allclose(D[:, Jl, K - 2], exp, rtol=1e-12, atol=1e-30)
        assert not D[:, 0, :].any()           # J = 0
        assert not D[:, :, 1::2].any()         # odd K
        assert not D[:, 1, 2:].any()           # J = 1: K = 4, 6 > 2J
        assert np.all(D[:, 1:, 0] > 0)         # rank 2 of every J > 0 level relaxes


def test_depolarisation_temperature_interpolation(oracle):
    """make_depolarization_temp's linear interpolation in T, bracketed as make_AC brackets
    the collision rates (so the first interval's line is extrapolated below the table),
    times n_H2 [m^-3] 1e-6; NaN above the table."""
    rng = np.random.default_rng(12)
    sp = _rotor_with_rates(4, rng)
    D = oracle.depol_table(sp)
    ct = sp.col_temps
    nH2 = 3e11
    for T in (0.5 * (ct[1] + ct[2]), ct[3], 0.3 * ct[0], ct[0]):
        got = oracle.depol_at(sp, D, T, nH2)
        for i in range(sp.e_num):
            for K in range(sp.kmax - 1):
                f = D[:, i, K]
                v = np.interp(T, ct, f) if T >= ct[0] else f[0] + (T - ct[0]) / (ct[1] - ct[0]) * (f[1] - f[0])
                ref = v * nH2 * 1e-6
                assert got[i, K] == pytest.approx(ref, rel=1e-14, abs=1e-300)
    assert np.isnan(oracle.depol_at(sp, D, ct[-1] * 2, nH2)).all()
