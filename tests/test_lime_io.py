"""The LIME/LAMDA reader against the reference's own input stage (CPU only).

tests/golden/lime_model/ holds a small model in LIME's vtk/pop, LAMDA and dust-table
formats (200 interior nodes and their sinks, a 5-level rotor with two collision
partners, a gas-to-dust ratio on half of the nodes); tests/golden/eat_lime.npz holds
what the reference's eat_lime_dim + eat_lime_no_delaunay (eat_lime.f90:1-313, compiled
from its sources) fill from those files.  The reader must give the same arrays bit for
bit, including the extra bulk node, the sink defaults, the neighbour table, the dust
opacities and the second collision partner's block as the reference parses it.
"""
import json
import os

import numpy as np
import pytest

from portal_amd import cloud, lime_io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _compare(G, S, ref):
    """The reader's Grid/Species against the reference's arrays, bit for bit."""
    same = np.testing.assert_array_equal
    same(G.XYZ, ref["XYZ"])
    same(G.vel_all, ref["vel_all"])
    same(G.temp_all, ref["temp_all"])
    same(G.lev_p, ref["lev_p"])
    same(G.h2_dens, ref["h2_dens"])
    same(G.b_all, ref["b_all"])
    same(G.ab_all, ref["ab_all"])
    same(G.dust_p, ref["dust_p"])
    assert G.tx == float(ref["rmax"])
    assert G.bulk_node_num == int(ref["bulk_node_num"])
    same(S.e_levs_cm, ref["e_levs"])
    same(S.t_Aji, ref["t_Aji"])
    same(S.nu_ghz, ref["nu"])
    same(S.t_Cji, ref["t_Cji"])
    same(S.col_temps, ref["col_temps"])
    same(S.j_lev, ref["j_lev"])
    same(S.r_trans, ref["r_trans"])
    same(S.c_trans, ref["c_trans"])
    same(S.kd_v, ref["kd_v"])
    assert S.mass_mol == float(ref["mass_mol"])
    same(np.diff(G.nbr_ptr), ref["del_num_ne"])
    dn = G.del_ne()
    same(dn, ref["del_ne"][:, : dn.shape[1]])
    assert not np.asarray(ref["del_ne"])[:, dn.shape[1]:].any()
    assert dn.shape[1] == int(ref["maxne"])


def test_reader_equals_the_reference_input_stage():
    z = np.load(os.path.join(GOLDEN, "eat_lime.npz"))
    files = {k: os.path.join(GOLDEN, v) for k, v in json.loads(str(z["files"])).items()}
    G, S, info = lime_io.read_model(files["pop"], files["grid"], files["mol"], files["dust"])
    _compare(G, S, z)
    # what the checks exercise: the extra bulk node, sinks, dust, two partners
    nb = G.bulk_node_num
    assert info["n_pop"] == nb - 1 and G.node_num > nb
    assert G.h2_dens[nb - 1] == 0 and G.b_all[nb - 1] == 0 and not G.lev_p[:, nb - 1].any()
    assert np.all(G.h2_dens[nb:] == 1.0) and np.all(G.ab_all[nb:] == 1e-3)
    assert (G.dust_p > 0).sum() > 50 and np.all(S.kd_v > 0)
    assert info["c_par"] == 2 and S.t_Cji[:, :, 1].any()


def test_reader_against_the_compiled_reference_on_a_new_model(tmp_path):
    """The same comparison live, on files this test writes (skipped where the reference
    sources are not mounted)."""
    from oracle.bind import LIB_REF, Reference
    if not os.path.exists(LIB_REF):
        pytest.skip("reference not built (make -C oracle ref)")
    from scipy.spatial import Delaunay
    sp = cloud.co_like_rotor(6)
    g = cloud.make_cloud(sp, 300, seed=19, n_exp=-1.8)
    pos = (g.XYZ * g.tx).T.copy()
    lam = np.logspace(0, 4, 67)
    gd = np.full(g.bulk_node_num, 150.0)
    paths = lime_io.write_model(str(tmp_path), g, sp, Delaunay(pos).simplices, pos,
                                dust_table=np.array([lam, 3e2 * lam ** -1.2]), gas_to_dust=gd)
    ref = Reference().eat_lime(paths["pop"], paths["grid"], paths["mol"], paths["dust"])
    G, S, _ = lime_io.read_model(paths["pop"], paths["grid"], paths["mol"], paths["dust"])
    _compare(G, S, ref)


def test_write_read_round_trip(tmp_path):
    """A generated cloud written in LIME's formats reads back as the same model."""
    from scipy.spatial import Delaunay
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 400, seed=23)
    pos = (g.XYZ * g.tx).T.copy()
    paths = lime_io.write_model(str(tmp_path), g, sp, Delaunay(pos).simplices, pos)
    G, S, info = lime_io.read_model(paths["pop"], paths["grid"], paths["mol"], None, bvec=g.bvec_all[:, 0],
                                    gvec=g.gvec)
    assert G.node_num == g.node_num and G.bulk_node_num == g.bulk_node_num
    np.testing.assert_allclose(G.XYZ, g.XYZ, rtol=0, atol=4e-16)
    np.testing.assert_array_equal(G.nbr_ptr, g.nbr_ptr)
    np.testing.assert_array_equal(G.nbr_idx, g.nbr_idx)
    np.testing.assert_allclose(G.b_all, g.b_all, rtol=1e-15)
    np.testing.assert_allclose(G.lev_p, g.lev_p, rtol=1e-15, atol=0)
    np.testing.assert_array_equal(G.vel_all, g.vel_all)
    np.testing.assert_array_equal(G.h2_dens, g.h2_dens)
    np.testing.assert_array_equal(S.e_levs_cm, sp.e_levs_cm)
    np.testing.assert_array_equal(S.t_Cji, sp.t_Cji)
    assert S.n_trans == sp.n_trans and S.c_dim == sp.c_dim
    np.testing.assert_array_equal(G.bvec_all, g.bvec_all)


def test_sink_defaults_and_norm2():
    """The sink levels follow exp(-E float(0.527)) normalised in order; the largest radius
    is NORM2's scaled sum of squares, which here differs from the plain formula."""
    import math
    z = np.load(os.path.join(GOLDEN, "eat_lime.npz"))
    e = z["e_levs"]
    w = [1.0] + [math.exp(-x * float(np.float32(0.527))) for x in e[1:]]
    s = 1.0
    for x in w[1:]:
        s = s + x
    nb = int(z["bulk_node_num"])
    np.testing.assert_array_equal(z["lev_p"][:, nb], np.array(w) / s * 1e-3 * 1.0)
    assert lime_io._norm2([3.0, 4.0, 12.0]) == 13.0
    assert lime_io._norm2([0.0, 0.0, 0.0]) == 0.0
