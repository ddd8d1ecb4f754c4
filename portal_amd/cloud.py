"""Synthetic LIME-style clouds and a CO-like linear-rotor species.

The arrays produced here are exactly what PORTAL's Fortran driver holds after it
has read a LIME model (``eat_lime_no_delaunay``, eat_lime.f90:71-313) and applied
its unit conversions (main_para.f90:119-135, 178-198).  Nothing is read from disk:
the generator builds the model in memory from a seed, so the same seed always gives
the same grid, adjacency, populations and field directions.

Conventions kept from the reference (they fix the kernels' inputs):

* node ``bulk_node_num = N_pop + 1`` is looped over but has no properties: the
  reference's read loop increments its counter before the failing read
  (eat_lime.f90:209-224).  It is a sink-surface point whose arrays stay zero.
* sink (surface) nodes ``N_pop+2..N`` get the defaults of eat_lime.f90:227-245.
* ``b_all`` gets thermal broadening added in quadrature (eat_lime.f90:250-258).
* ``lev_p`` is multiplied by abundance and n_H2 (eat_lime.f90:261-263).
* positions are normalised by the largest radius (eat_lime.f90:268-279) and
  ``tx`` is that radius in metres.
* neighbour lists are the Delaunay neighbours, unique and sorted ascending,
  1-based (adj_list_no_delaunay + Unique, miscellaneous.f90:910-1015,
  unique.f90:43-59); they are held both as CSR and as the reference's
  ``del_ne(N, maxne)`` column-major table.
* single-precision literals promoted to double are reproduced with
  ``float(np.float32(...))`` (main_para.f90:119-120, 135; eat_lime.f90:237).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

AU_M = 1.495978707e11          # astronomical unit [m]
F32 = lambda x: float(np.float32(x))  # noqa: E731  single-precision literal promoted to double

# Fortran single-precision literals of the reference, as the doubles it actually uses.
CM1_TO_HZ = F32(29.9792458e9)   # main_para.f90:120, para_finish.f90:171
DX_INWARD = F32(1e-6)           # main_para.f90:135
DX_IMAGE = F32(1e-4)            # para_finish.f90:201
SINK_POP_FAC = F32(0.527)       # eat_lime.f90:237
A_TO_B_DRIVER = 1.6179542e57    # c^3/(8 pi h), main_para.f90:178


@dataclass
class Species:
    """Level/transition tables of a molecule after the driver's conversions."""
    e_num: int
    j_lev: np.ndarray           # (L,) int32
    e_levs_cm: np.ndarray       # (L,) cm^-1 as read
    mass_mol: float
    r_trans: np.ndarray         # (2, R) int32, 1-based (lower, upper)  eat_lime.f90:167
    t_Aji: np.ndarray           # (R,)
    nu_ghz: np.ndarray          # (R,)
    c_trans: np.ndarray         # (2, Cn) int32, 1-based (lower, upper)  eat_lime.f90:181
    col_temps: np.ndarray       # (c_temp,)
    t_Cji: np.ndarray           # (c_temp, Cn, c_par) Fortran order, cm^3 s^-1
    kmax: int = 6
    kd_v: np.ndarray | None = None  # (R,) dust opacities at the line frequencies

    # derived (main_para.f90)
    lev_g: np.ndarray = field(init=False)
    nu_hz: np.ndarray = field(init=False)
    e_levs_hz: np.ndarray = field(init=False)
    t_Bij: np.ndarray = field(init=False)
    all_trans: np.ndarray = field(init=False)
    n_trans: int = field(init=False)
    count_array: np.ndarray = field(init=False)
    c_dim: int = field(init=False)

    def __post_init__(self):
        self.j_lev = np.ascontiguousarray(self.j_lev, dtype=np.int32)
        self.lev_g = (2 * self.j_lev + 1).astype(np.int32)
        # 1E9 is exact in single precision
        self.nu_hz = self.nu_ghz * 1e9
        self.e_levs_hz = self.e_levs_cm * CM1_TO_HZ
        g1 = self.lev_g[self.r_trans[0] - 1].astype(np.float64)
        g2 = self.lev_g[self.r_trans[1] - 1].astype(np.float64)
        self.t_Bij = (g2 / g1) * self.t_Aji * A_TO_B_DRIVER / (self.nu_hz ** 3.0)
        if self.kd_v is None:
            self.kd_v = np.zeros(self.r_num)
        self.all_trans, self.n_trans = all_transitions(self.e_num, self.r_trans, self.c_trans, self.e_levs_hz)
        self.count_array, self.c_dim = setup_count_array(self.j_lev, self.kmax)

    @property
    def r_num(self) -> int:
        return int(self.r_trans.shape[1])

    @property
    def c_num(self) -> int:
        return int(self.c_trans.shape[1])

    @property
    def c_temp(self) -> int:
        return int(self.col_temps.shape[0])

    @property
    def c_par(self) -> int:
        return int(self.t_Cji.shape[2])


def all_transitions(nL, lA, lC, E):
    """Unique (lower, upper) level pairs in lexicographic order (miscellaneous.f90:394-460)."""
    T = np.zeros((nL, nL), dtype=bool)
    for lst in (lA, lC):
        for i in range(lst.shape[1]):
            j, k = int(lst[0, i]), int(lst[1, i])
            if E[k - 1] > E[j - 1]:
                T[j - 1, k - 1] = True
            else:
                T[k - 1, j - 1] = True
    pairs = [(i + 1, j + 1) for i in range(nL) for j in range(nL) if T[i, j]]
    out = np.zeros((2, lA.shape[1] + lC.shape[1]), dtype=np.int32)
    for n, (i, j) in enumerate(pairs):
        out[0, n] = i
        out[1, n] = j
    return out, len(pairs)


def setup_count_array(j_lev, kmax):
    """0-based start slot of each level's (K=0,2,..,min(2J,kmax)) block (miscellaneous.f90:206-239)."""
    ca = np.zeros(len(j_lev), dtype=np.int32)
    n = 0
    for i, j in enumerate(j_lev):
        ca[i] = n
        n += len(range(0, min(2 * int(j), kmax) + 1, 2))
    return ca, n


def co_like_rotor(n_levels: int = 4, b_rot_ghz: float = 57.636, dipole_debye: float = 0.11,
                  mass: float = 28.0, k_down: float = 1e-11, kmax: int = 6,
                  col_temps=(1.0, 10.0, 100.0, 1e3, 1e4)) -> Species:
    """Linear rotor J=0..n_levels-1 with synthetic, temperature-flat collision rates.

    E_J = (B/c) J(J+1) [cm^-1]; lines J->J-1 at 2BJ; A from the dipole formula
    64 pi^4 nu^3 mu^2 / (3 h c^3) * J/(2J+1) (cgs); downward collision rate
    k_down / (1 + 0.1 dJ) cm^3/s for every pair.
    """
    h, c = 6.62607015e-27, 2.99792458e10
    mu = dipole_debye * 1e-18
    J = np.arange(n_levels)
    b_cm = b_rot_ghz * 1e9 / c
    e_cm = b_cm * J * (J + 1)
    r_trans = np.array([[j, j + 1] for j in range(1, n_levels)], dtype=np.int32).T.copy()
    Ju = J[1:]
    nu_ghz = 2.0 * b_rot_ghz * Ju
    nu = nu_ghz * 1e9
    A = 64 * math.pi ** 4 * nu ** 3 * mu ** 2 / (3 * h * c ** 3) * Ju / (2 * Ju + 1)
    pairs = [(l, u) for u in range(2, n_levels + 1) for l in range(1, u)]
    c_trans = np.array(pairs, dtype=np.int32).T.copy()
    ct = np.asarray(col_temps, dtype=np.float64)
    rates = np.array([k_down / (1.0 + 0.1 * (u - l)) for (l, u) in pairs])
    t_Cji = np.asfortranarray(np.repeat(rates[None, :], len(ct), axis=0)[:, :, None])
    return Species(e_num=n_levels, j_lev=J.astype(np.int32), e_levs_cm=e_cm, mass_mol=mass,
                   r_trans=r_trans, t_Aji=A, nu_ghz=nu_ghz, c_trans=c_trans,
                   col_temps=ct, t_Cji=t_Cji, kmax=kmax)


@dataclass
class Grid:
    """Per-node arrays of one cloud, laid out as the Fortran driver holds them."""
    node_num: int
    bulk_node_num: int
    XYZ: np.ndarray          # (3,N) normalised positions, F-order
    vel_all: np.ndarray      # (3,N) m/s
    temp_all: np.ndarray     # (2,N) (T_kin, T_dust)
    b_all: np.ndarray        # (N,) Doppler width incl. thermal, m/s
    h2_dens: np.ndarray      # (N,) m^-3
    ab_all: np.ndarray       # (N,)
    dust_p: np.ndarray       # (N,)
    lev_p: np.ndarray        # (L,N) level densities (fraction * ab * n_H2)
    nbr_ptr: np.ndarray      # (N+1,) int32 CSR row pointers
    nbr_idx: np.ndarray      # (nnz,) int32, 1-based neighbour ids, sorted per row
    tx: float                # normalisation radius [m]
    bvec_all: np.ndarray     # (3,N) unit field directions
    gvec: np.ndarray         # (3,) gauge direction

    @property
    def del_num_ne(self) -> np.ndarray:
        return np.diff(self.nbr_ptr).astype(np.int32)

    @property
    def maxne(self) -> int:
        return int(self.del_num_ne.max())

    def del_ne(self) -> np.ndarray:
        """The reference's ``del_ne(node_num, maxne)`` table (column-major, zero padded)."""
        N, m = self.node_num, self.maxne
        out = np.zeros((N, m), dtype=np.int32, order="F")
        cnt = self.del_num_ne
        rows = np.repeat(np.arange(N), cnt)
        cols = np.arange(len(self.nbr_idx)) - np.repeat(self.nbr_ptr[:-1], cnt)
        out[rows, cols] = self.nbr_idx
        return out


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def _simplex_edges(s: np.ndarray):
    """Directed edges (a, b) of the tetrahedra s (every ordered pair of distinct corners)."""
    s = s.astype(np.int64)
    a = np.concatenate([s[:, i] for i in range(4) for j in range(4) if i != j])
    b = np.concatenate([s[:, j] for i in range(4) for j in range(4) if i != j])
    return a, b


def _block_edges(args):
    """Delaunay edges from the points of one box (lo <= p < hi per axis) to their
    neighbours, triangulating the box grown by a halo of width h on every side."""
    from scipy.spatial import Delaunay
    pos, lo, hi, h, k = args
    inb = np.all((pos >= lo - h) & (pos < hi + h), axis=1)
    sel = np.where(inb)[0]
    a, b = _simplex_edges(Delaunay(pos[sel]).simplices)
    a, b = sel[a], sel[b]
    pa = pos[a]
    own = np.all((pa >= lo) & (pa < hi), axis=1)
    return k, a[own], b[own]


def _kd_boxes(pos: np.ndarray, splits):
    """Boxes of equal point count: the points are halved (or cut in s parts) along x,
    then each part along y, then along z, at point quantiles."""
    boxes = [(np.full(3, -np.inf), np.full(3, np.inf), np.arange(pos.shape[0]))]
    for axis, parts in enumerate(splits):
        nxt = []
        for lo, hi, idx in boxes:
            v = np.sort(pos[idx, axis])
            cuts = [-np.inf] + [v[(k * len(v)) // parts] for k in range(1, parts)] + [np.inf]
            for k in range(parts):
                l2, h2 = lo.copy(), hi.copy()
                l2[axis], h2[axis] = cuts[k], cuts[k + 1]
                m = (pos[idx, axis] >= cuts[k]) & (pos[idx, axis] < cuts[k + 1])
                nxt.append((l2, h2, idx[m]))
        boxes = nxt
    return [(lo, hi) for lo, hi, _ in boxes]


def delaunay_neighbours(pos: np.ndarray, workers: int | None = None, log=None):
    """Sorted unique Delaunay neighbours of every point, as CSR with 1-based ids.

    Up to 300,000 points one triangulation of the whole set.  Above that the set is cut
    into 32 boxes of equal point count (4 x 4 x 2 along x, y, z), each triangulated with
    a halo of 6 mean point spacings in worker processes (fork), and a point takes its
    neighbours from its own box's triangulation (the edge lists are then made
    symmetric).  A Delaunay edge of a point depends only on the points within its
    tetrahedra's circumspheres, which at these densities stay inside the halo; the box
    form is checked edge for edge against the whole-set triangulation in the tests.
    """
    from scipy.spatial import Delaunay
    n = pos.shape[0]
    if workers is None:
        workers = 1 if n <= 300000 else 16
    if workers <= 1:
        a, b = _simplex_edges(Delaunay(pos).simplices)
    else:
        import multiprocessing as mp
        boxes = _kd_boxes(pos, (4, 4, 2))
        ext = np.ptp(pos, axis=0)
        h = 6.0 * (float(np.prod(ext)) * (math.pi / 6.0) / n) ** (1.0 / 3.0)
        jobs = [(pos, lo, hi, h, k) for k, (lo, hi) in enumerate(boxes)]
        parts = []
        with mp.get_context("fork").Pool(workers) as pool:
            for k, a_k, b_k in pool.imap_unordered(_block_edges, jobs):
                parts.append((a_k, b_k))
                if log:
                    log(f"delaunay: box {k + 1} of {len(jobs)} triangulated ({len(parts)} done)")
        a = np.concatenate([p[0] for p in parts] + [p[1] for p in parts])
        b = np.concatenate([p[1] for p in parts] + [p[0] for p in parts])
    key = np.unique(a * n + b)
    a = key // n
    b = key % n
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, a + 1, 1)
    ptr = np.cumsum(ptr)
    return ptr.astype(np.int32), (b + 1).astype(np.int32)


def make_cloud(species: Species, n_int: int, seed: int, n_sink: int | None = None,
               r_out_au: float = 1e4, r0_au: float = 1e3, n0_cm3: float = 1e6, n_exp: float = -1.5,
               n_floor_cm3: float = 1e3, t0: float = 20.0, t_exp: float = -0.4, v0_ms: float = -1000.0,
               v_exp: float = -0.5, doppler_ms: float = 200.0, abundance: float = 1e-4,
               bvec=None, log=None) -> Grid:
    """Seeded cloud: n_int points uniform in a sphere of radius r_out_au plus sink points on its surface.

    n_H2 = n0 (r/r0)^n_exp floored at n_floor; T = t0 (r/r0)^t_exp; infall v = v0 (r/r0)^v_exp r_hat;
    microturbulent Doppler width doppler_ms; LTE starting populations; dust off; one uniform field
    direction for every node (seeded unless given) and a seeded gauge vector.
    """
    rng = np.random.default_rng(seed)
    if n_sink is None:
        n_sink = max(256, n_int // 8)
    R = r_out_au * AU_M
    r = R * rng.random(n_int) ** (1.0 / 3.0)
    pin = _unit(rng, n_int) * r[:, None]
    psk = _unit(rng, n_sink) * R
    pos = np.concatenate([pin, psk])
    N = pos.shape[0]
    ptr, idx = delaunay_neighbours(pos, log=log)

    L = species.e_num
    vel = np.zeros((3, N))
    temp = np.zeros((2, N))
    b = np.zeros(N)
    h2 = np.zeros(N)
    ab = np.zeros(N)
    lev = np.zeros((L, N))

    ra = r / AU_M / r0_au
    nh = np.maximum(n0_cm3 * ra ** n_exp, n_floor_cm3)
    T = t0 * ra ** t_exp
    vmag = v0_ms * ra ** v_exp
    vel[:, :n_int] = (pin / r[:, None] * vmag[:, None]).T
    temp[0, :n_int] = T
    temp[1, :n_int] = T
    b[:n_int] = doppler_ms
    h2[:n_int] = nh * 1e6
    ab[:n_int] = abundance
    hck = 1.4387769  # hc/k [cm K]
    g = species.lev_g.astype(np.float64)
    boltz = g[:, None] * np.exp(-species.e_levs_cm[:, None] * hck / T[None, :])
    lev[:, :n_int] = boltz / boltz.sum(axis=0)

    # node n_int+1 (1-based) keeps zero properties; sinks from n_int+2 get the defaults
    sk = slice(n_int + 1, N)
    h2[sk] = 1.0
    b[sk] = 200.0
    ab[sk] = 1e-3
    sp = np.ones(L)
    for k in range(1, L):
        sp[k] = 1.0 * np.exp(-species.e_levs_cm[k] * SINK_POP_FAC)
    sp = sp / sp.sum()
    lev[:, sk] = sp[:, None]

    b_therm = 2.0 * 8314.462 * temp[0] / species.mass_mol
    b = np.sqrt(b ** 2.0 + b_therm)
    lev = lev * (ab * h2)[None, :]

    rmax = np.sqrt((pos ** 2).sum(axis=1)).max()
    XYZ = np.asfortranarray((pos / rmax).T)

    if bvec is None:
        bvec = _unit(rng, 1)[0]
    bvec = np.asarray(bvec, dtype=np.float64)
    bvec = bvec / np.sqrt((bvec ** 2).sum())
    gvec = rng.random(3)
    gvec = gvec / np.sqrt((gvec ** 2.0).sum())
    bvec_all = np.asfortranarray(np.repeat(bvec[:, None], N, axis=1))

    return Grid(node_num=N, bulk_node_num=n_int + 1, XYZ=XYZ, vel_all=np.asfortranarray(vel),
                temp_all=np.asfortranarray(temp), b_all=b, h2_dens=h2, ab_all=ab,
                dust_p=np.zeros(N), lev_p=np.asfortranarray(lev), nbr_ptr=ptr, nbr_idx=idx,
                tx=float(rmax), bvec_all=bvec_all, gvec=gvec)


def legendre_all(nth: int, nph: int) -> np.ndarray:
    """Direction set (mu, phi, weight) as columns of a (3, n_sphere) array (miscellaneous.f90:1100-1137).

    Gauss-Legendre mu_i in ascending order; ceil(nph*sin(acos(mu_i))) uniform phi's per mu_i,
    each with weight w_i * (2 pi / np) / (4 pi).
    """
    x, w = np.polynomial.legendre.leggauss(nth)
    cols = []
    for i in range(nth):
        npp = int(math.ceil(nph * math.sin(math.acos(x[i]))))
        wph = 2.0 * math.pi / npp
        ww = wph * w[i]
        for j in range(npp):
            cols.append((x[i], j * wph, ww / (4.0 * math.pi)))
    return np.asfortranarray(np.array(cols).T)
