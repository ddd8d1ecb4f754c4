"""The oracle's emission factor for thin cells.

ray_in (inverse_rt.f90:314, :337) adds S (1 - e^-tau) per cell with 1 - e^-tau formed
as written, which in FP64 is zero below tau ~ 1.1e-16.  The FP64 GPU path follows that
form (its parity target is the reference); the mixed path evaluates the factor without
cancellation, and its checker is the oracle with ``thin_exact``, -expm1(-tau).  These
tests pin that variant to the closed form of a uniform ray and show where, and only
where, the two forms part.
"""
import math

import numpy as np
import pytest

from portal_amd import cloud

NU0, AJI, BIJ, G1, G2 = 2.882e12, 1e-2, 1e5, 49, 51
B, NVEL, DX, M, NCELL = 300.0, 8, 1e13, 10, 40


def _uniform_ray(oracle, tau_c, s_over_iin, thin_exact):
    """A ray of NCELL identical cells at rest (the own cell included), whose
    line-centre optical depth per cell is tau_c -> (oracle I_out, closed form, S)."""
    spi = math.sqrt(math.pi)
    cnu = 299792458.0 / NU0
    c1, c2 = 2.2102191e-42 * NU0, 5.272859e-35 * NU0
    ratio = 0.5  # p2 / p1
    kv1 = c1 * BIJ * (1.0 - (1.0 * G1) * ratio / (1.0 * G2))  # kv per unit p1
    p1 = tau_c / (kv1 * cnu / (spi * B) * DX * M)
    p2 = ratio * p1
    kv = c1 * BIJ * (p1 - (1.0 * G1) * p2 / (1.0 * G2))
    S = c2 * AJI * p2 / kv
    I_in = S / s_over_iin
    got = oracle.ray_in(B, 0.0, p1, p2, np.full(NCELL, p1), np.full(NCELL, p2), np.full(NCELL, B),
                        np.zeros(NCELL), M * np.arange(1, NCELL + 1), NVEL, DX, NU0, AJI, BIJ, G1, G2, I_in,
                        thin_exact=thin_exact)
    # closed form: per channel I_out = I_in e^-T + S (1 - e^-T), T = NCELL tau(channel)
    want = 0.0
    d_v = 2.0 * B / NVEL
    for nv in range(-NVEL, NVEL + 1):
        a = (2.0 * B * nv / NVEL) / B
        phi0 = math.exp(-a * a) / (spi * B)
        T = NCELL * kv * (cnu * phi0) * DX * M
        want += phi0 * (I_in * math.exp(-T) - S * math.expm1(-T)) * d_v
    return got, want, I_in


@pytest.mark.parametrize("tau_c", [1e-22, 1e-19, 1e-16, 1e-12, 1e-6, 1e-3, 0.5, 4.0])
@pytest.mark.parametrize("s_over_iin", [1e25, 1e3])
def test_thin_exact_ray_in_is_the_closed_form(oracle, tau_c, s_over_iin):
    got, want, _ = _uniform_ray(oracle, tau_c, s_over_iin, True)
    assert got == pytest.approx(want, rel=1e-12)


def test_reference_form_loses_thin_cell_emission(oracle):
    # thick cells: the reference's form is the closed form too
    for tau_c in (0.5, 4.0):
        got, want, _ = _uniform_ray(oracle, tau_c, 1e25, False)
        assert got == pytest.approx(want, rel=1e-12)
    # tau = 1e-22 per cell and S = 1e25 I_in: the emission (~4e4 I_in) is lost and the
    # ray carries its background only
    got, want, I_in = _uniform_ray(oracle, 1e-22, 1e25, False)
    assert want > 1e3 * I_in
    assert got == pytest.approx(I_in, rel=1e-12)


def test_forms_part_only_in_the_wien_tail_of_the_background(oracle):
    """On LAMDA's CO size (40 lines, 20 K gas) the two forms agree to rounding on the low
    lines, whose source functions stay within ~1e10 of the 2.73 K background, and part on
    lines in its Wien tail, where thin cells' emission is comparable to the background
    (by up to ~20% at these nodes, lines 20-24; the FP64 error passes 1e-6 from line 16)."""
    sp = cloud.co_like_rotor(41, kmax=2)
    g = cloud.make_cloud(sp, 1500, seed=71)
    xth = cloud.legendre_all(4, 8)
    nodes = np.nonzero(g.ab_all[:g.bulk_node_num] > 0)[0][[3, 200, 700]] + 1
    ref = oracle.inverse_nodes(g, sp, xth, 16, cloud.DX_INWARD, nodes, nthreads=4)[0]
    thin = oracle.inverse_nodes(g, sp, xth, 16, cloud.DX_INWARD, nodes, nthreads=4, thin_exact=True)[0]
    rel = np.abs(thin / ref - 1).max(axis=1)
    assert rel[:11].max() < 1e-12
    assert rel[16:].max() > 1e-2


def _prop_mp(eI, eQ, pI, pQ, ds, I, Q, U):
    """polarize_prop's own formula (raytrace.f90:1-56) in 60-digit arithmetic."""
    import mpmath as mp
    mp.mp.dps = 60
    eI, eQ, pI, pQ, ds, I, Q, U = (mp.mpf(x) for x in (eI, eQ, pI, pQ, ds, I, Q, U))
    shQ, chQ, et = mp.sinh(eQ * ds), mp.cosh(eQ * ds), mp.exp(-(eI * ds))
    fac = 1 / (eI * eI - eQ * eQ)
    o11 = fac * eI * (1 - (chQ + (eQ / eI) * shQ) * et)
    o12 = -(fac * eQ * (1 - (chQ + (eI / eQ) * shQ) * et))
    return (float(o11 * pI + o12 * pQ + et * (chQ * I - shQ * Q)),
            float(o12 * pI + o11 * pQ + et * (-(shQ * I) + chQ * Q)), float(U * et))


@pytest.mark.parametrize("t", [1e-22, 1e-16, 1e-12, 1e-6, 1e-2, 0.5, 3.0])
@pytest.mark.parametrize("r", [1e-3, 0.3, -0.5, 0.9])
def test_thin_exact_propagator_is_the_formula_evaluated_exactly(oracle, t, r):
    ds = 2.0e14
    eI = t / ds
    eQ = r * eI
    S = 1e25  # source function in units of the incoming I: thin cells' emission matters
    pI, pQ = S * eI, 0.2 * S * eQ
    I, Q, U = 1.0, 0.01, 0.02
    want = _prop_mp(eI, eQ, pI, pQ, ds, I, Q, U)
    got = oracle.polarize_prop(eI, eQ, pI, pQ, ds, I, Q, U, thin_exact=True)
    assert got[0] == pytest.approx(want[0], rel=1e-12)
    assert abs(got[1] - want[1]) <= 1e-12 * abs(want[0])
    assert got[2] == pytest.approx(want[2], rel=1e-15)
    if t >= 0.5:  # thick: the reference's form is exact to rounding too
        ref = oracle.polarize_prop(eI, eQ, pI, pQ, ds, I, Q, U)
        assert ref[0] == pytest.approx(want[0], rel=1e-12)
    if t == 1e-22:  # the reference's form drops the emission (S t = 1e3 here)
        ref = oracle.polarize_prop(eI, eQ, pI, pQ, ds, I, Q, U)
        assert abs(ref[0] - want[0]) > 0.5 * abs(want[0])
