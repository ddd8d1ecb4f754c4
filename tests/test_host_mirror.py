"""Host-side mirrors of the driver steps and the synthetic model (CPU only)."""
import math

import numpy as np

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
