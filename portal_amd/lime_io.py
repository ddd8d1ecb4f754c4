"""LIME model files and LAMDA molecular data, read as the reference's input stage reads them.

The reference's drivers read a model with eat_lime_dim and eat_lime_no_delaunay
(eat_lime.f90:1-313, with adj_list_no_delaunay and dust_to_opacities,
miscellaneous.f90:643-700, 909-1015) and convert units in main_para.f90:119-120 /
para_finish.f90:170-171.  In a drop-in those steps stay Fortran; this module does the
same in Python, with the same semantics, so that the Python interface can be fed real
LIME output and the tests can show that it gives the reference's arrays bit for bit:

* grid.vtk: 4 header lines, ``POINTS N <type>``, N lines of x y z [m], one separator
  line, ``CELLS ntet size``, then ntet lines ``4 a b c d`` (0-based vertex ids).  The
  neighbour list of a node is the sorted set of the other vertices of its tetrahedra.
* .pop: one header line, then per internal node x y z vx vy vz n_H2[m^-3] T_kin T_dust
  b[m/s] abundance gas-to-dust conv pops(1:L).  The positions come from the vtk.  The
  loop counts the failed read at the end of the file, so bulk_node_num is the number of
  data lines + 1, and that extra node keeps zero properties (eat_lime.f90:209-224).
  Nodes after it are LIME's sink nodes and get fixed defaults (eat_lime.f90:227-245).
* b gains thermal broadening in quadrature, 2 R T / m with R = 8314.462 J/(kmol K); level
  densities are fractions x abundance x n_H2; positions are divided by the largest radius
  (returned as tx); dust_p = n_H2 x 2.32E-27 / gas-to-dust.
* LAMDA: the standard layout, parsed by line counts.  With two collision partners the
  reference reads the second partner's number of temperatures as its number of
  transitions (eat_lime.f90:183-195); that is kept.
* dust table: 67 rows of wavelength [micron] and opacity [cm^2/g], interpolated in
  log-log at each line's wavelength (2.998E14 / nu) and converted to m^2/kg.

Single-precision literals of the reference (0.527, 2.32E-27, 2.998E14, 1E-1) are used as
the doubles they become.  write_model writes a generated cloud in these formats.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import cloud


def _f32(x: float) -> float:
    return float(np.float32(x))


SINK_POP_FAC = _f32(0.527)     # eat_lime.f90:237
DUST_MASS = _f32(2.32e-27)     # eat_lime.f90:307
NU_TO_MICRON = _f32(2.998e14)  # miscellaneous.f90:667
CGS_TO_SI = _f32(1e-1)         # miscellaneous.f90:669
N_DUST_TABLE = 67              # eat_lime.f90:84


class _Records:
    """Fortran list-directed reads over a text file: each read starts at a new record
    and takes as many values as it needs from that record and the ones after it (blank
    records supply none); what is left of its last record is dropped."""

    def __init__(self, path: str):
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.pos = 0

    def skip(self, n: int = 1):
        if self.pos + n > len(self.lines):
            raise EOFError
        self.pos += n

    def read(self, n: int) -> list[str]:
        out: list[str] = []
        while len(out) < n:
            if self.pos >= len(self.lines):
                raise EOFError
            toks = self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
            out.extend(toks[: n - len(out)])
        return out


def _num(t: str) -> float:
    return float(t.replace("D", "E").replace("d", "e"))


def read_dims(grid: str, mol: str) -> dict:
    """eat_lime_dim (eat_lime.f90:1-69)."""
    g = _Records(grid)
    g.skip(4)
    node_num = int(g.read(2)[1])
    g.skip(node_num + 1)
    del_num = int(g.read(2)[1])
    m = _Records(mol)
    m.skip(5)
    e_num = int(m.read(1)[0])
    m.skip(1 + e_num + 1)
    r_num = int(m.read(1)[0])
    m.skip(1 + r_num + 1)
    c_par = int(m.read(1)[0])
    m.skip(3)
    c_num = int(m.read(1)[0])
    m.skip(1)
    c_temp = int(m.read(1)[0])
    return dict(node_num=node_num, del_num=del_num, e_num=e_num, r_num=r_num, c_num=c_num, c_temp=c_temp,
                c_par=c_par)


def _neighbours(grid: str, node_num: int, del_num: int):
    """adj_list_no_delaunay (miscellaneous.f90:909-1015): CSR of the sorted unique
    neighbours (1-based) and the reference's zero-padded del_ne table."""
    r = _Records(grid)
    r.skip(5 + node_num + 2)
    lists = [[] for _ in range(node_num)]
    others = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    for _ in range(del_num):
        t = [int(x) + 1 for x in r.read(5)[1:]]
        for v in range(4):
            lists[t[v] - 1].extend(t[o] for o in others[v])
    uniq = [sorted(set(l)) for l in lists]
    cnt = np.array([len(u) for u in uniq], dtype=np.int32)
    ptr = np.zeros(node_num + 1, dtype=np.int32)
    ptr[1:] = np.cumsum(cnt)
    idx = np.array([x for u in uniq for x in u], dtype=np.int32)
    return ptr, idx


def _dust_opacities(nu_hz: np.ndarray, table: np.ndarray) -> np.ndarray:
    """dust_to_opacities (miscellaneous.f90:643-700) on the log10 table (2 x n): the
    opacity [m^2/kg] at each line's wavelength.  A wavelength on a table node or below
    the table leaves the reference's kd_m at its previous line's value (undefined for
    the first line: NaN here)."""
    lt = table
    n = lt.shape[1]
    out = np.zeros(len(nu_hz))
    kd_m = math.nan
    for i, nu in enumerate(nu_hz):
        mu0 = math.log10(NU_TO_MICRON / nu)
        for j in range(n - 1):
            if lt[0, j] < mu0 < lt[0, j + 1]:
                mu_res = (mu0 - lt[0, j]) / (lt[0, j + 1] - lt[0, j])
                kd_m = lt[1, j] + mu_res * (lt[1, j + 1] - lt[1, j])
                break
        if mu0 > lt[0, n - 1]:
            mu_res = (mu0 - lt[0, n - 1]) / (lt[0, n - 1] - lt[0, n - 2])
            kd_m = lt[1, n - 1] + mu_res * (lt[1, n - 1] - lt[1, n - 2])
        out[i] = 10.0 ** kd_m * CGS_TO_SI
    return out


def read_model(pop: str, grid: str, mol: str, dust: str | None = None, kmax: int = 6, bvec=None, gvec=None):
    """(Grid, Species, info) from LIME/LAMDA files, as eat_lime_no_delaunay and the
    drivers' unit conversions leave them.  bvec (a field direction for every node, or
    one for all) and gvec are the driver's (main_para.f90:149-176); unit vectors
    along z and x by default.  info: mass_mol, maxne, del_num, the raw .pop count, and
    the dust opacities' table."""
    d = read_dims(grid, mol)
    N, e_num, r_num, c_num, c_temp, c_par = (d[k] for k in ("node_num", "e_num", "r_num", "c_num", "c_temp",
                                                              "c_par"))
    g = _Records(grid)
    g.skip(5)
    XYZ = np.zeros((3, N), order="F")
    for i in range(N):
        XYZ[:, i] = [_num(t) for t in g.read(3)]

    m = _Records(mol)
    m.skip(3)
    mass = _num(m.read(1)[0])
    m.skip(3)
    e_cm = np.zeros(e_num)
    j_lev = np.zeros(e_num, dtype=np.int32)
    for i in range(e_num):
        t = m.read(3)
        e_cm[i] = _num(t[1])
        j_lev[i] = int(math.floor((_num(t[2]) - 1.0) / 2.0 + 0.5))  # nint, half away from zero
    m.skip(3)
    r_trans = np.zeros((2, r_num), dtype=np.int32)
    A = np.zeros(r_num)
    nu_ghz = np.zeros(r_num)
    for i in range(r_num):
        t = m.read(5)
        r_trans[1, i], r_trans[0, i] = int(t[1]), int(t[2])
        A[i], nu_ghz[i] = _num(t[3]), _num(t[4])
    m.skip(9)
    col_temps = np.array([_num(t) for t in m.read(c_temp)])
    m.skip(1)
    c_trans = np.zeros((2, c_num), dtype=np.int32)
    t_Cji = np.zeros((c_temp, c_num, c_par), order="F")
    for i in range(c_num):
        t = m.read(3 + c_temp)
        c_trans[1, i], c_trans[0, i] = int(t[1]), int(t[2])
        t_Cji[:, i, 0] = [_num(x) for x in t[3:]]
    if c_par == 2:
        m.skip(5)
        c_num_2 = int(m.read(1)[0])  # the second partner's number of temperatures (eat_lime.f90:188)
        m.skip(3)
        for i in range(c_num_2):
            t = m.read(3 + c_temp)
            t_Cji[:, i, 1] = [_num(x) for x in t[3:]]

    vel = np.zeros((3, N), order="F")
    temp = np.zeros((2, N), order="F")
    h2 = np.zeros(N)
    b = np.zeros(N)
    ab = np.zeros(N)
    dust_ii = np.zeros(N)
    lev = np.zeros((e_num, N), order="F")
    p = _Records(pop)
    p.skip(1)
    i = 0
    while True:
        try:
            t = [_num(x) for x in p.read(13 + e_num)]
        except EOFError:
            break
        vel[:, i] = t[3:6]
        h2[i], temp[0, i], temp[1, i], b[i], ab[i], dust_ii[i] = t[6:12]
        lev[:, i] = t[13:]
        i += 1
    n_pop = i
    bulk = n_pop + 1  # the read that meets the end of the file is counted too
    for j in range(bulk, N):
        h2[j], b[j], ab[j] = 1.0, 200.0, 1e-3
        lp = [1.0]
        s_p = 1.0
        for k in range(1, e_num):
            lp.append(1.0 * math.exp(-e_cm[k] * SINK_POP_FAC))
            s_p = s_p + lp[k]
        lev[:, j] = [x / s_p for x in lp]
    b_therm = 2.0 * 8314.462 * temp[0] / mass
    b = np.sqrt(b * b + b_therm)
    lev = np.asfortranarray(lev * ab[None, :] * h2[None, :])
    rmax = 0.0
    for i in range(N):
        r1 = _norm2(XYZ[:, i])
        if r1 > rmax:
            rmax = r1
    XYZ = np.asfortranarray(XYZ / rmax)
    ptr, idx = _neighbours(grid, N, d["del_num"])

    kd = np.zeros(r_num)
    table = None
    if dust is not None:
        tr = _Records(dust)
        table = np.zeros((2, N_DUST_TABLE))
        for i in range(N_DUST_TABLE):
            table[:, i] = [_num(x) for x in tr.read(2)]
        # the C library's log10, as the compiled reference calls it (numpy's vector
        # log10 can differ in the last bit)
        table = np.vectorize(math.log10, otypes=[float])(table)
        kd = _dust_opacities(nu_ghz * 1e9, table)
    dust_p = np.where(dust_ii != 0.0, h2 * DUST_MASS / np.where(dust_ii != 0.0, dust_ii, 1.0), 0.0)

    sp = cloud.Species(e_num=e_num, j_lev=j_lev, e_levs_cm=e_cm, mass_mol=mass, r_trans=r_trans, t_Aji=A,
                       nu_ghz=nu_ghz, c_trans=c_trans, col_temps=col_temps, t_Cji=t_Cji, kmax=kmax, kd_v=kd)
    if bvec is None:
        bvec = np.array([0.0, 0.0, 1.0])
    bvec = np.asarray(bvec, dtype=np.float64)
    bvec_all = np.asfortranarray(np.repeat(bvec[:, None], N, axis=1) if bvec.ndim == 1 else bvec)
    grid_ = cloud.Grid(node_num=N, bulk_node_num=bulk, XYZ=XYZ, vel_all=vel, temp_all=temp, b_all=b, h2_dens=h2,
                       ab_all=ab, dust_p=dust_p, lev_p=lev, nbr_ptr=ptr, nbr_idx=idx, tx=float(rmax),
                       bvec_all=bvec_all, gvec=np.array([1.0, 0.0, 0.0]) if gvec is None else np.asarray(gvec))
    info = dict(mass_mol=mass, maxne=int(np.diff(ptr).max()), del_num=d["del_num"], n_pop=n_pop, dust_table=table,
                c_par=c_par)
    return grid_, sp, info


def _norm2(v) -> float:
    """Fortran NORM2 as the reference's build evaluates it (amdflang's runtime): with m the
    running largest |x|, accumulate s = sum of (|x|/m)^2 over the other elements
    (rescaling s by (m_old/m_new)^2 and adding that ratio when m grows), then m sqrt(1 + s).
    It can differ from sqrt(x^2 + y^2 + z^2) in the last bit, which the normalised
    positions inherit."""
    mx, s = 0.0, 0.0
    for x in v:
        ax = abs(float(x))
        if mx == 0.0:
            mx = ax
        elif ax > mx:
            t = mx / ax
            tsq = t * t
            s = s * tsq
            s = s + tsq
            mx = ax
        else:
            t = ax / mx
            s = s + t * t
    return mx * math.sqrt(1.0 + s)


def write_model(directory: str, g, sp, tets: np.ndarray, positions_m: np.ndarray, name: str = "model",
                dust_table: np.ndarray | None = None, gas_to_dust: np.ndarray | None = None) -> dict:
    """Write a model in LIME's and LAMDA's formats: grid.vtk (positions [m] of every
    node, the tetrahedra), <name>.pop (the interior nodes, bulk_node_num - 1 of them),
    <name>.dat (LAMDA, one collision partner unless t_Cji holds two) and, with a table,
    dust.tab.  b is written without its thermal part and the level populations as
    fractions, so that read_model returns g's arrays.  Returns the four paths."""
    os.makedirs(directory, exist_ok=True)
    N = g.node_num
    n_pop = g.bulk_node_num - 1
    f17 = lambda x: f"{float(x):.17g}"  # noqa: E731
    grid = os.path.join(directory, "grid.vtk")
    with open(grid, "w") as f:
        f.write("# vtk DataFile Version 3.0\nLIME grid\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {N} float\n")
        for i in range(N):
            f.write(" ".join(f17(x) for x in positions_m[i]) + "\n")
        f.write("\n")
        f.write(f"CELLS {len(tets)} {5 * len(tets)}\n")
        for t in tets:
            f.write(f"4 {t[0]} {t[1]} {t[2]} {t[3]}\n")
    pop = os.path.join(directory, f"{name}.pop")
    mass = sp.mass_mol
    with open(pop, "w") as f:
        f.write("# x y z vx vy vz nH2 Tk Td b ab gas_to_dust conv pops\n")
        for i in range(n_pop):
            bt = 2.0 * 8314.462 * g.temp_all[0, i] / mass
            b_turb = math.sqrt(max(g.b_all[i] ** 2 - bt, 0.0))
            frac = g.lev_p[:, i] / (g.ab_all[i] * g.h2_dens[i]) if g.ab_all[i] * g.h2_dens[i] > 0 else g.lev_p[:, i]
            gd = 0.0 if gas_to_dust is None else gas_to_dust[i]
            vals = list(positions_m[i]) + list(g.vel_all[:, i]) + [g.h2_dens[i], g.temp_all[0, i], g.temp_all[1, i],
                                                                   b_turb, g.ab_all[i], gd, 0.0] + list(frac)
            f.write(" ".join(f17(x) for x in vals) + "\n")
    mol = os.path.join(directory, f"{name}.dat")
    c_par = sp.t_Cji.shape[2]
    with open(mol, "w") as f:
        f.write(f"!MOLECULE\n{name}\n!MOLECULAR WEIGHT\n{f17(mass)}\n!NUMBER OF ENERGY LEVELS\n{sp.e_num}\n")
        f.write("!LEVEL + ENERGIES(cm^-1) + WEIGHT + J\n")
        for i in range(sp.e_num):
            f.write(f"{i + 1} {f17(sp.e_levs_cm[i])} {2 * int(sp.j_lev[i]) + 1}.0 {int(sp.j_lev[i])}\n")
        f.write(f"!NUMBER OF RADIATIVE TRANSITIONS\n{sp.r_num}\n!TRANS + UP + LOW + EINSTEINA(s^-1) + FREQ(GHz)\n")
        for i in range(sp.r_num):
            f.write(f"{i + 1} {sp.r_trans[1, i]} {sp.r_trans[0, i]} {f17(sp.t_Aji[i])} {f17(sp.nu_ghz[i])}\n")
        for p in range(c_par):
            if p == 0:
                f.write(f"!NUMBER OF COLL PARTNERS\n{c_par}\n")
            f.write(f"!COLLISIONS BETWEEN\n{p + 2} partner {p + 1}\n!NUMBER OF COLL TRANS\n{sp.c_num}\n")
            f.write(f"!NUMBER OF COLL TEMPS\n{sp.c_temp}\n!COLL TEMPS\n")
            f.write(" ".join(f17(x) for x in sp.col_temps) + "\n!TRANS + UP + LOW + COLLRATES(cm^3 s^-1)\n")
            for i in range(sp.c_num):
                f.write(f"{i + 1} {sp.c_trans[1, i]} {sp.c_trans[0, i]} " +
                        " ".join(f17(x) for x in sp.t_Cji[:, i, p]) + "\n")
    dust = None
    if dust_table is not None:
        dust = os.path.join(directory, "dust.tab")
        with open(dust, "w") as f:
            for row in dust_table.T:
                f.write(f"{f17(row[0])} {f17(row[1])}\n")
    return dict(pop=pop, grid=grid, mol=mol, dust=dust)
