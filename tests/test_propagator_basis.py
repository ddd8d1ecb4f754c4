"""The FP32 Stokes transfer's propagator against the reference's closed form.

The mixed-precision image kernel applies polarize_prop (raytrace.f90:1-56) in the
basis X = I + Q, Y = I - Q, where it is two scalar transfer problems with absorption
eta_I +- eta_Q.  Here that form, written out in numpy, is checked against the golden
polarize_prop vectors of the compiled reference: in float64 (with an exact (1-e^-x)/x)
it must agree to the reference's own rounding (the identity), in float32, with the
kernel's series, within single-precision error on the scale of the result (tau e_32
for optically thick segments).
"""
import os

import numpy as np


def decoupled(eI, eQ, pI, pQ, ds, I, Q, U, dt=np.float64):
    eI, eQ, pI, pQ, ds, I, Q, U = (dt(v) for v in (eI, eQ, pI, pQ, ds, I, Q, U))
    one, half = dt(1), dt(0.5)

    def phi(x):  # (1 - e^-x)/x; float32: the kernel's series below |x| = 1e-2
        if dt is np.float64:
            return -np.expm1(-x) / x
        return (one - x * (half - x / dt(6))) if abs(x) < dt(1e-2) else (one - np.exp(-x)) / x

    m, p, t = (eI + eQ) * ds, (eI - eQ) * ds, eI * ds
    X = (I + Q) * np.exp(-m) + (pI + pQ) * ds * phi(m)
    Y = (I - Q) * np.exp(-p) + (pI - pQ) * ds * phi(p)
    return half * (X + Y), half * (X - Y), U * np.exp(-t)


def test_decoupled_basis_is_the_closed_form(golden_dir):
    z = np.load(os.path.join(golden_dir, "image_pieces.npz"))
    rows = z["polprop"]
    assert len(rows) >= 20
    for row in rows:
        ref = row[8:]
        got = decoupled(*row[:8])
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12 * scale)


def test_decoupled_basis_in_float32(golden_dir):
    z = np.load(os.path.join(golden_dir, "image_pieces.npz"))
    worst = 0.0
    with np.errstate(over="ignore", under="ignore"):
        for row in z["polprop"]:
            ref = row[8:]
            # intensities in units of the incoming I, as the kernel carries them
            s = abs(row[5]) if row[5] != 0 else np.abs(ref).max()
            args = list(row[:8])
            args[2] /= s
            args[3] /= s
            args[5] /= s
            args[6] /= s
            args[7] /= s
            got = np.array(decoupled(*args, dt=np.float32), dtype=np.float64) * s
            worst = max(worst, np.abs(got - ref).max() / np.abs(ref).max())
    assert worst < 1e-5
