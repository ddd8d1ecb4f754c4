"""The driver-side steps around the accelerated loops, as the Fortran drivers do them.

These stay host code in the reference's drivers; they are mirrored here so the
Python interface, the tests and the benchmark can run a whole pass:

* ``rescale_pol_pops``  para_finish.f90:222-240 (LIME populations, SE alignment ratio)
* ``image_lines``       para_finish.f90:173-185 (nu0_all, Aji_all, kd_all lookup,
                        including the reference's nu0_all overwrite: every imaged
                        line gets the last line's frequency)
* ``channel_velocities`` para_finish.f90:124-126
* ``cube``              para_finish.f90:451-484 (x, y, v cube times I_to_T)
"""
from __future__ import annotations

import numpy as np


def rescale_pol_pops(pol_pops, lev_p, lev_g):
    """rho00(i) <- lev_p(i)/sqrt(g_i); rho20(i) <- (rho20/rho00)_SE * that, i >= 2."""
    pp = np.array(pol_pops, dtype=np.float64, order="F", copy=True)
    L = lev_p.shape[0]
    with np.errstate(divide="ignore", invalid="ignore"):
        pp[:, 0] = lev_p[0] / np.sqrt(1.0 * lev_g[0])
        for i in range(1, L):
            rat = pp[:, 2 * i + 1] / pp[:, 2 * i]
            pp[:, 2 * i + 1] = rat * lev_p[i] / np.sqrt(1.0 * lev_g[i])
            pp[:, 2 * i] = lev_p[i] / np.sqrt(1.0 * lev_g[i])
    return pp


def image_lines(sp, l1, l2, nu0_overwrite=True, aji_by_level=False):
    """[(l1, l2, nu0, Aji, kd)] for the imaged lines (1-based levels).

    The defaults reproduce the image driver: the frequency loop of para_finish.f90:173-175
    leaves every line with the last line's frequency.  nu0_overwrite=False gives each
    line its own frequency instead.  aji_by_level=True takes A as the reference's pix = 0
    branch does, t_Aji(l1(j)), indexing the transition list by the lower level
    (para_finish.f90:303; the grid branch uses the line's own A, 422).
    """
    nu0_last = sp.e_levs_hz[l2[-1] - 1] - sp.e_levs_hz[l1[-1] - 1]
    out = []
    for a, b in zip(l1, l2):
        A = kd = 0.0
        for j in range(sp.r_num):
            if sp.r_trans[1, j] == b and sp.r_trans[0, j] == a:
                A, kd = sp.t_Aji[j], sp.kd_v[j]
                break
        if aji_by_level:
            A = sp.t_Aji[a - 1]
        nu0 = nu0_last if nu0_overwrite else sp.e_levs_hz[b - 1] - sp.e_levs_hz[a - 1]
        out.append((a, b, nu0, A, kd))
    return out


def channel_velocities(nvel, dvel, v0=0.0):
    return np.array([i * dvel + v0 for i in range(-nvel, nvel + 1)])


def cube(all_X, line, im_pix, nu0=None):
    """(2im_pix+1, 2im_pix+1, C) cube of one line; multiplied by I_to_T when nu0 is given."""
    n_side = 2 * im_pix + 1
    a = all_X[:, line, :].reshape(n_side, n_side, -1)
    if nu0 is not None:
        a = a * (3.255e39 / (nu0 * nu0))
    return a
