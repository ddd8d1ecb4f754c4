"""The Fortran binding (portal_amd/fortran) drives the library like a PORTAL driver would."""
import os
import subprocess

import numpy as np
import pytest

from portal_amd import cloud
from portal_amd.api import Portal

pytestmark = pytest.mark.gpu
DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "portal_amd", "portal_f90_demo")


def dump(path, g, sp, xth, nvel):
    del_ne = g.del_ne()
    with open(path, "wb") as f:
        hdr = np.array([g.node_num, g.bulk_node_num, sp.e_num, sp.r_num, sp.c_num, sp.c_temp, sp.c_par, sp.kmax,
                        del_ne.shape[1], xth.shape[1], nvel], dtype=np.int32)
        f.write(hdr.tobytes())
        f.write(np.array([g.tx, cloud.DX_INWARD]).tobytes())
        for a in (g.XYZ, g.vel_all, g.temp_all, g.b_all, g.h2_dens, g.ab_all, g.dust_p, g.lev_p):
            f.write(np.asfortranarray(a, dtype=np.float64).tobytes(order="F"))
        f.write(np.asfortranarray(del_ne, dtype=np.int32).tobytes(order="F"))
        f.write(np.asarray(g.del_num_ne, dtype=np.int32).tobytes())
        f.write(np.asarray(sp.j_lev, dtype=np.int32).tobytes())
        f.write(np.asarray(sp.e_levs_hz, dtype=np.float64).tobytes())
        f.write(np.asfortranarray(sp.r_trans, dtype=np.int32).tobytes(order="F"))
        for a in (sp.t_Aji, sp.nu_hz, sp.t_Bij, sp.kd_v):
            f.write(np.asarray(a, dtype=np.float64).tobytes())
        f.write(np.asfortranarray(sp.c_trans, dtype=np.int32).tobytes(order="F"))
        f.write(np.asfortranarray(sp.t_Cji, dtype=np.float64).tobytes(order="F"))
        f.write(np.asarray(sp.col_temps, dtype=np.float64).tobytes())
        f.write(np.asfortranarray(xth, dtype=np.float64).tobytes(order="F"))
        f.write(np.asarray(g.gvec, dtype=np.float64).tobytes())
        f.write(np.asfortranarray(g.bvec_all, dtype=np.float64).tobytes(order="F"))


def test_fortran_driver_matches_python_interface(tmp_path):
    if not os.path.exists(DEMO):
        pytest.skip("Fortran demo not built (no Fortran compiler when the library was built)")
    sp = cloud.co_like_rotor(4)
    g = cloud.make_cloud(sp, 1200, seed=21)
    xth = cloud.legendre_all(4, 8)
    fin, fout, fdat = str(tmp_path / "model.bin"), str(tmp_path / "pops.bin"), str(tmp_path / "pol_pops.dat")
    dump(fin, g, sp, xth, 16)
    r = subprocess.run([DEMO, fin, fout, fdat], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    pf = np.fromfile(fout, dtype=np.float64).reshape((g.node_num, 2 * sp.e_num), order="F")
    pt = Portal(0)
    pt.upload_grid(g)
    pt.upload_species(sp)
    pt.set_levpop(g.lev_p)
    pp, _ = pt.se_pass(xth, 16, cloud.DX_INWARD)
    pt.close()
    np.testing.assert_array_equal(pf, pp)
    # pol_pops.dat as main_para.f90:271-275 writes it, read back as para_finish.f90:226-228
    # does (one list-directed record of 2L doubles per node): the values survive exactly
    rows = [ln.split() for ln in open(fdat) if ln.strip()]
    assert len(rows) == g.node_num and all(len(r) == 2 * sp.e_num for r in rows)
    back = np.array([[float(v.replace("D", "E")) for v in r] for r in rows])
    np.testing.assert_array_equal(back, pp)
