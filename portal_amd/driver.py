"""The driver-side steps around the accelerated loops, as the Fortran drivers do them.

These stay host code in the reference's drivers; they are mirrored here so the
Python interface, the tests and the benchmark can run a whole pass:

* ``rescale_pol_pops``  para_finish.f90:222-240 (LIME populations, SE alignment ratio)
* ``image_lines``       para_finish.f90:173-185 (nu0_all, Aji_all, kd_all lookup,
                        including the reference's nu0_all overwrite: every imaged
                        line gets the last line's frequency)
* ``channel_velocities`` para_finish.f90:124-126
* ``cube``              para_finish.f90:451-484 (x, y, v cube times I_to_T)
* ``ang_scale``         para_finish.f90:188 (pixel size in degrees)
* ``write_fits``        writeimage (fits_writing.f90:1-158): a FITS primary array of
                        the float32 cube with the reference's header keywords, as
                        para_finish.f90:497-511 calls it
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


def ang_scale(im_size, tx, distance, im_pix):
    """Pixel size in degrees (para_finish.f90:188); distance in the units of tx."""
    import math
    return -180.0 * math.asin(2.0 * im_size * tx / distance) / (math.pi * (im_pix * 2 + 1))


def _card(key, value=None, raw=None):
    if key == "END":
        return "END".ljust(80)
    if raw is None:
        if isinstance(value, bool):
            raw = ("T" if value else "F").rjust(20)
        elif isinstance(value, str):
            raw = ("'" + value.ljust(8) + "'").ljust(20)
        elif isinstance(value, int):
            raw = str(value).rjust(20)
        else:  # a real*4 keyword with 8 decimals (ftpkye(..., 8, ...))
            raw = f"{float(np.float32(value)):.8E}".rjust(20)
    return (key.ljust(8) + "= " + raw).ljust(80)


def write_fits(path, cube, restfreq, cdelt1, crpix1, cdelt2, crpix2, cdelt3, crpix3, crval1=0.0, crval2=0.0,
               crval3=0.0):
    """writeimage (fits_writing.f90:1-158): cube is (npix, npix, nvel) indexed [x, y, v]
    as the reference's array(npix, npix, nvel); it is stored as big-endian float32 with
    x varying fastest, after the keywords the reference writes, in its order.  The
    image driver passes cdelt1 = ang_scale, cdelt2 = -ang_scale, crpix1 = crpix2 =
    im_pix + 1, cdelt3 = dvel, crpix3 = nvel + 1 (para_finish.f90:497-511)."""
    npix, npix2, nvel = cube.shape
    assert npix == npix2
    cards = [_card("SIMPLE", True), _card("BITPIX", -32), _card("NAXIS", 3), _card("NAXIS1", npix),
             _card("NAXIS2", npix), _card("NAXIS3", nvel), _card("EXTEND", True),
             _card("OBJECT", "PORTAL"), _card("EPOCH", 2.0e3), _card("LONPOLE", 1.8e2), _card("EQUINOX", 2.0e3),
             _card("SPECSYS", "LSRK"), _card("RESTFREQ", restfreq), _card("VELREF", 257),
             _card("CTYPE1", "RA---SIN"), _card("CDELT1", cdelt1), _card("CRPIX1", crpix1),
             _card("CRVAL1", crval1), _card("CUNIT1", "DEG"), _card("CTYPE2", "DEC--SIN"),
             _card("CDELT2", cdelt2), _card("CRPIX2", crpix2), _card("CRVAL2", crval2), _card("CUNIT2", "DEG"),
             _card("CTYPE3", "VELO-LSR"), _card("CDELT3", cdelt3), _card("CRPIX3", crpix3),
             _card("CRVAL3", crval3), _card("CUNIT3", "M/S"), _card("BSCALE", 1.0), _card("BZERO", 0.0),
             _card("BUNIT", "K"), _card("END")]
    head = "".join(cards).encode("ascii")
    head += b" " * (-len(head) % 2880)
    data = np.asarray(cube, dtype=">f4").ravel(order="F").tobytes()
    data += b"\0" * (-len(data) % 2880)
    with open(path, "wb") as f:
        f.write(head)
        f.write(data)
