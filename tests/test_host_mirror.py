"""Host-side mirrors of the driver steps and the synthetic model (CPU only)."""
import math

import numpy as np
import pytest

from portal_amd import cloud, driver


def test_model_follows_eat_lime_conventions():
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 300, seed=5)
    nb = g.bulk_node_num
    assert nb == 301  # N_pop + 1 (eat_lime.f90:209-224)
    q = nb - 1  # the extra bulk node has no properties
    assert g.ab_all[q] == 0 and g.h2_dens[q] == 0 and g.b_all[q] == 0 and np.all(g.lev_p[:, q] == 0)
    s = nb  # first sink node: defaults of eat_lime.f90:227-245
    assert g.h2_dens[s] == 1.0 and g.ab_all[s] == 1e-3 and g.temp_all[0, s] == 0
    assert g.b_all[s] == 200.0
    fr = g.lev_p[:, s] / (g.ab_all[s] * g.h2_dens[s])
    ref = np.array([1.0] + [math.exp(-e * cloud.SINK_POP_FAC) for e in sp.e_levs_cm[1:]])
    np.testing.assert_allclose(fr, ref / ref.sum(), rtol=1e-15)
    # thermal broadening in quadrature (eat_lime.f90:250-258)
    i = 7
    assert g.b_all[i] == math.sqrt(200.0 ** 2 + 2 * 8314.462 * g.temp_all[0, i] / sp.mass_mol)
    # positions normalised by the largest radius
    r = np.sqrt((g.XYZ ** 2).sum(axis=0))
    assert abs(r.max() - 1.0) < 1e-15
    # neighbour lists sorted ascending, symmetric, 1-based
    for n in range(0, g.node_num, 37):
        nb_ = g.nbr_idx[g.nbr_ptr[n]:g.nbr_ptr[n + 1]]
        assert np.all(np.diff(nb_) > 0) and nb_.min() >= 1
        for m in nb_[:3]:
            assert n + 1 in g.nbr_idx[g.nbr_ptr[m - 1]:g.nbr_ptr[m]]
    # del_ne table holds the same lists
    dn = g.del_ne()
    assert dn.flags.f_contiguous
    assert list(dn[10, : g.del_num_ne[10]]) == list(g.nbr_idx[g.nbr_ptr[10]:g.nbr_ptr[11]])


def test_species_derivations():
    sp = cloud.co_like_rotor(4)
    # CO-like Einstein A values and the single-precision cm^-1 -> Hz factor (main_para.f90:120)
    np.testing.assert_allclose(sp.t_Aji, [7.19e-8, 6.90e-7, 2.50e-6], rtol=5e-3)
    assert cloud.CM1_TO_HZ == 29979246592.0
    assert sp.c_dim == 1 + 2 + 3 + 4 and sp.n_trans == 6
    assert list(sp.count_array) == [0, 1, 3, 6]
    assert sp.all_trans[:, :3].tolist() == [[1, 1, 1], [2, 3, 4]]
    sp10 = cloud.co_like_rotor(10)
    assert sp10.c_dim == 34 and sp10.n_trans == 45


def test_rescale_pol_pops():
    # para_finish.f90:222-240: LIME populations with the SE's alignment ratio
    L = 3
    pp = np.zeros((2, 2 * L), order="F")
    pp[0] = [0.5, 0.0, 0.2, 0.01, 0.1, -0.002]
    lev = np.array([[2.0, 0.0], [3.0, 0.0], [4.0, 0.0]])
    g = np.array([1, 3, 5])
    out = driver.rescale_pol_pops(pp, lev, g)
    assert out[0, 0] == 2.0
    assert out[0, 2] == 3.0 / math.sqrt(3) and out[0, 3] == (0.01 / 0.2) * 3.0 / math.sqrt(3)
    assert out[0, 5] == (-0.002 / 0.1) * 4.0 / math.sqrt(5)


def test_image_lines_keep_the_nu0_overwrite():
    # para_finish.f90:173-175 assigns the whole array in the loop: every line gets the last line's frequency
    sp = cloud.co_like_rotor(4)
    lines = driver.image_lines(sp, [1, 2], [2, 3])
    nu_last = sp.e_levs_hz[2] - sp.e_levs_hz[1]
    assert lines[0][2] == nu_last and lines[1][2] == nu_last
    assert lines[0][3] == sp.t_Aji[0] and lines[1][3] == sp.t_Aji[1]
    v = driver.channel_velocities(2, 10.0, 1.0)
    assert list(v) == [-19.0, -9.0, 1.0, 11.0, 21.0]
    # the two policies: each line its own frequency; A indexed by the lower level as the
    # pix = 0 branch does (para_finish.f90:303)
    own = driver.image_lines(sp, [1, 3], [2, 4], nu0_overwrite=False)
    assert own[0][2] == sp.e_levs_hz[1] - sp.e_levs_hz[0] and own[1][2] == sp.e_levs_hz[3] - sp.e_levs_hz[2]
    assert own[1][3] == sp.t_Aji[2]
    bylev = driver.image_lines(sp, [1, 3], [2, 4], aji_by_level=True)
    assert bylev[0][3] == sp.t_Aji[0] and bylev[1][3] == sp.t_Aji[2]
    shifted = driver.image_lines(sp, [2], [3], aji_by_level=True)
    assert shifted[0][3] == sp.t_Aji[1]


def test_box_delaunay_equals_whole_set():
    # the parallel form used for large clouds (boxes with a halo, one worker process per
    # box) gives the whole-set triangulation's neighbour lists edge for edge
    from portal_amd import cloud
    rng = np.random.default_rng(17)
    n = 60000
    r = rng.random(n) ** (1 / 3)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    u = rng.normal(size=(n // 8, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    pos = np.concatenate([v * r[:, None], u]) * 1.5e15
    p1, i1 = cloud.delaunay_neighbours(pos, workers=1)
    p2, i2 = cloud.delaunay_neighbours(pos, workers=4)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(i1, i2)


def test_rescale_of_a_self_consistent_state_is_a_uniform_scaling(oracle):
    """para_finish.f90:222-240 on a state the lambda closure made self-consistent
    (lev_p = ab n_H2 sqrt(2J+1) rho^0_0 of the same SE solution) scales every column of
    a node by ab n_H2: the alignment ratios are kept and rho^0_0 comes back from lev_p."""
    sp = cloud.co_like_rotor(5)
    g = cloud.make_cloud(sp, 300, seed=8)
    xth = cloud.legendre_all(4, 8)
    pp, _, _ = oracle.se_pass(g, sp, xth, 8, cloud.DX_INWARD, nthreads=8)
    idx = np.where(g.ab_all[: g.bulk_node_num] > 0)[0]
    nab = g.ab_all * g.h2_dens
    sg = np.sqrt(2.0 * sp.j_lev + 1.0)
    lev = np.zeros_like(g.lev_p)
    lev[:, idx] = (pp[idx][:, 0::2] * sg[None, :] * nab[idx, None]).T
    out = driver.rescale_pol_pops(pp, lev, sp.lev_g)
    np.testing.assert_allclose(out[idx], pp[idx] * nab[idx, None], rtol=2e-15, atol=0)
    assert np.abs(pp[idx][:, 3::2]).max() > 0  # alignment present


def test_cube_orientation_and_brightness_temperature(oracle):
    """para_finish.f90:451-484: a cloud that emits only where X > 0.3 and Y < -0.2, seen
    face-on (no rotation), lights only the cube's pixels with x > 0 and y < 0 along its
    first and second axes; and I_to_T = 3.255e39 / nu^2 is c^2 / 2k (Rayleigh-Jeans)."""
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 3000, seed=9)
    lit = (g.XYZ[0] > 0.3) & (g.XYZ[1] < -0.2)
    lit[g.bulk_node_num - 1:] = False
    # elsewhere 1e-30 of the populations (zero would make the emission coefficients 0/0)
    lev = np.where(lit[None, :], g.lev_p, 1e-30 * g.lev_p)
    pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
    pp[:, 0::2] = (lev / np.sqrt(sp.lev_g[:, None].astype(float))).T
    # some alignment (the reference's closed-form propagator needs eta_Q != 0)
    rat = np.random.default_rng(2).uniform(-1e-3, 1e-3, (g.node_num, sp.e_num - 1))
    pp[:, 3::2] = rat * pp[:, 2::2]
    lines = driver.image_lines(sp, [1], [2])
    vo = driver.channel_velocities(4, 250.0)
    I, Q, U, T = oracle.image_pass(g, sp, pp, lines, 0.0, 0.0, 4, vo, 8, 0.7, cloud.DX_IMAGE, n_depth_rt=4096,
                                   nthreads=8)
    nu0 = lines[0][2]
    c = driver.cube(I, 0, 8, nu0=nu0)
    assert c.shape == (17, 17, 9)
    x = (np.arange(17) - 8) * 0.7 / 8
    on = np.abs(c).max(axis=2)
    peak = on.max()
    assert on[np.ix_(x > 0.4, x < -0.3)].min() > 1e-3 * peak  # inside the lit region
    assert on[x < 0, :].max() < 1e-20 * peak and on[:, x > 0].max() < 1e-20 * peak
    np.testing.assert_allclose(c * nu0 ** 2 / 3.255e39, I[:, 0, :].reshape(17, 17, 9), rtol=1e-15)
    assert 3.255e39 == pytest.approx(2.99792458e8 ** 2 / (2 * 1.380649e-23), rel=1e-4)


def test_fits_file_of_a_cube(tmp_path):
    """writeimage's layout (fits_writing.f90:1-158): 2880-byte blocks, the keywords in the
    reference's order with real*4 values, big-endian float32 data with x fastest."""
    rng = np.random.default_rng(4)
    cube = rng.normal(size=(5, 5, 3))
    asc = driver.ang_scale(0.7, 1.5e15, 3.1e18, 2)
    path = str(tmp_path / "inc01-trans01-im_i.fits")
    driver.write_fits(path, cube, 1.15e11, asc, 3, -asc, 3, 250.0, 2)
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode() for i in range(0, 2880, 80)]
    keys = [c[:8].strip() for c in cards if c.strip()]
    assert keys[:7] == ["SIMPLE", "BITPIX", "NAXIS", "NAXIS1", "NAXIS2", "NAXIS3", "EXTEND"]
    assert keys[7:] == ["OBJECT", "EPOCH", "LONPOLE", "EQUINOX", "SPECSYS", "RESTFREQ", "VELREF", "CTYPE1", "CDELT1",
                        "CRPIX1", "CRVAL1", "CUNIT1", "CTYPE2", "CDELT2", "CRPIX2", "CRVAL2", "CUNIT2", "CTYPE3",
                        "CDELT3", "CRPIX3", "CRVAL3", "CUNIT3", "BSCALE", "BZERO", "BUNIT", "END"]
    val = {c[:8].strip(): c[10:].strip() for c in cards if c[8:10] == "= "}
    assert val["BITPIX"] == "-32" and val["NAXIS3"] == "3" and val["VELREF"] == "257"
    assert np.float32(float(val["CDELT1"])) == np.float32(asc) and np.float32(float(val["CDELT2"])) == np.float32(-asc)
    assert float(val["CRPIX3"]) == 2.0 and val["CTYPE3"] == "'VELO-LSR'"
    data = np.frombuffer(raw[2880:2880 + 4 * cube.size], dtype=">f4").reshape((3, 5, 5))  # (v, y, x) in C order
    np.testing.assert_array_equal(data.transpose(2, 1, 0), cube.astype(np.float32))
    assert asc < 0  # RA increases to the left
