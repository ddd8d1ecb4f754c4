"""ctypes bindings to the CPU checkers (test infrastructure only).

* ``Oracle``: the C restatement in ``liboracle.so`` (portal_oracle.c).
* ``Reference``: the reference's own Fortran routines compiled from its sources
  into ``_ref/libportal_ref.so`` (only present where the reference is mounted and
  ``make -C oracle ref`` has run).  Fortran passes every argument by reference and
  expects column-major arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "liboracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libportal_ref.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


class OrcGrid(C.Structure):
    _fields_ = [("node_num", C.c_int), ("bulk_node_num", C.c_int), ("e_num", C.c_int),
                ("XYZ", dp), ("vel", dp), ("temp", dp), ("b_all", dp), ("h2_dens", dp),
                ("ab_all", dp), ("dust_p", dp), ("lev_p", dp), ("nbr_ptr", ip), ("nbr_idx", ip),
                ("tx", C.c_double)]


class OrcSpecies(C.Structure):
    _fields_ = [("e_num", C.c_int), ("r_num", C.c_int), ("c_num", C.c_int), ("c_temp", C.c_int),
                ("c_par", C.c_int), ("n_trans", C.c_int), ("kmax", C.c_int), ("c_dim", C.c_int),
                ("j_lev", ip), ("lev_g", ip), ("r_trans", ip), ("c_trans", ip), ("all_trans", ip),
                ("count_array", ip), ("e_levs_hz", dp), ("t_Aji", dp), ("nu_hz", dp), ("t_Bij", dp),
                ("kd_v", dp), ("t_Cji", dp), ("col_temps", dp)]


class OrcSeParams(C.Structure):
    _fields_ = [("n_sphere", C.c_int), ("xth", dp), ("nvel", C.c_int), ("dx", C.c_double),
                ("gvec", dp), ("bvec_all", dp), ("n_depth", C.c_int), ("extra_info", C.c_int),
                ("star_mod", C.c_int), ("nth", C.c_int), ("nph", C.c_int), ("R_star", C.c_double),
                ("T_star", C.c_double), ("X_star", dp), ("D0_all", dp)]


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))


class Oracle:
    def __init__(self, path: str = LIB_ORACLE):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_wig3jj.restype = C.c_double
        L.orc_wig3jj.argtypes = [C.c_int] * 6
        L.orc_wig6jj.restype = C.c_double
        L.orc_wig6jj.argtypes = [C.c_int] * 6
        L.orc_black_body.restype = C.c_double
        L.orc_black_body.argtypes = [C.c_double, C.c_double]
        L.orc_se_pass.restype = C.c_int
        L.orc_solve_from_J.restype = C.c_int
        L.orc_image_pass.restype = C.c_int
        L.orc_inverse_rt.restype = C.c_int
        L.orc_trace_out.restype = C.c_int
        L.orc_trace_out_raytrace.restype = C.c_int
        self._keep = []

    # -- structures -------------------------------------------------------
    def grid(self, g, lev_p=None):
        arrs = dict(XYZ=_f64(g.XYZ), vel=_f64(g.vel_all), temp=_f64(g.temp_all), b_all=_f64(g.b_all),
                    h2=_f64(g.h2_dens), ab=_f64(g.ab_all), dust=_f64(g.dust_p),
                    lev=_f64(g.lev_p if lev_p is None else lev_p), ptr=_i32(g.nbr_ptr), idx=_i32(g.nbr_idx))
        s = OrcGrid(g.node_num, g.bulk_node_num, arrs["lev"].shape[0], _d(arrs["XYZ"]), _d(arrs["vel"]),
                    _d(arrs["temp"]), _d(arrs["b_all"]), _d(arrs["h2"]), _d(arrs["ab"]), _d(arrs["dust"]),
                    _d(arrs["lev"]), _i(arrs["ptr"]), _i(arrs["idx"]), g.tx)
        self._keep.append(arrs)
        return s

    def species(self, sp):
        arrs = dict(j=_i32(sp.j_lev), g=_i32(sp.lev_g), rt=_i32(sp.r_trans), ct=_i32(sp.c_trans),
                    at=_i32(sp.all_trans), ca=_i32(sp.count_array), e=_f64(sp.e_levs_hz), A=_f64(sp.t_Aji),
                    nu=_f64(sp.nu_hz), B=_f64(sp.t_Bij), kd=_f64(sp.kd_v), C=_f64(sp.t_Cji), T=_f64(sp.col_temps))
        s = OrcSpecies(sp.e_num, sp.r_num, sp.c_num, sp.c_temp, sp.c_par, sp.n_trans, sp.kmax, sp.c_dim,
                       _i(arrs["j"]), _i(arrs["g"]), _i(arrs["rt"]), _i(arrs["ct"]), _i(arrs["at"]), _i(arrs["ca"]),
                       _d(arrs["e"]), _d(arrs["A"]), _d(arrs["nu"]), _d(arrs["B"]), _d(arrs["kd"]), _d(arrs["C"]),
                       _d(arrs["T"]))
        self._keep.append(arrs)
        return s

    def params(self, xth, nvel, dx, gvec, bvec_all, n_depth, extra_info=0, star=None, depol=None):
        """star: None, or dict(nth, nph, R_star, T_star, X_star) for star mode; depol: None,
        or the D0_all table of depol_table (collisional depolarisation)."""
        arrs = dict(xth=_f64(xth), g=_f64(gvec), b=_f64(bvec_all),
                    xs=_f64(star["X_star"] if star else np.zeros(3)))
        if depol is not None:
            arrs["d"] = _f64(depol)
        s = OrcSeParams(arrs["xth"].shape[1], _d(arrs["xth"]), nvel, dx, _d(arrs["g"]), _d(arrs["b"]),
                        n_depth, extra_info, 1 if star else 0, star["nth"] if star else 0,
                        star["nph"] if star else 0, star["R_star"] if star else 0.0,
                        star["T_star"] if star else 0.0, _d(arrs["xs"]), _d(arrs["d"]) if depol is not None else None)
        self._keep.append(arrs)
        return s

    def star_all(self, nth, nph, kvec, gvec, bvec, dOm):
        """(n_star, xth 3 x n_sphere) of orc_star_all."""
        xth = np.zeros((3, nth * nph), order="F")
        ns, nv = C.c_int(), C.c_int()
        rc = self.lib.orc_star_all(nth, nph, _d(_f64(kvec)), _d(_f64(gvec)), _d(_f64(bvec)), C.c_double(dOm),
                                   C.byref(ns), C.byref(nv), _d(xth))
        if rc:
            raise RuntimeError("oracle star_all: rays do not fit")
        return ns.value, np.asfortranarray(xth[:, : nv.value])

    # -- wrappers ---------------------------------------------------------
    def wig3jj(self, *a):
        return self.lib.orc_wig3jj(*[int(x) for x in a])

    def wig6jj(self, *a):
        return self.lib.orc_wig6jj(*[int(x) for x in a])

    def se_pass(self, g, sp, xth, nvel, dx, n_depth=4096, node_lo=1, node_hi=None, lev_p=None,
                extra_info=0, nthreads=0, bvec_all=None, gvec=None, star=None, depol=None):
        """Nodes node_lo..node_hi of the SE pass.  Returns (pol_pops N x 2L, J0 R x N, J2 R x N[, rates, JJ])."""
        if node_hi is None:
            node_hi = g.bulk_node_num
        G = self.grid(g, lev_p)
        S = self.species(sp)
        P = self.params(xth, nvel, dx, g.gvec if gvec is None else gvec,
                        g.bvec_all if bvec_all is None else bvec_all, n_depth, extra_info, star, depol)
        N, L, R = g.node_num, sp.e_num, sp.r_num
        pp = np.zeros((N, 2 * L), order="F")
        J0 = np.zeros((R, N), order="F")
        J2 = np.zeros((R, N), order="F")
        rates = np.zeros((N, 3, 2 * L), order="F")
        jj = np.zeros((N, R), order="F")
        rc = self.lib.orc_se_pass(C.byref(G), C.byref(S), C.byref(P), node_lo, node_hi, _d(pp), _d(rates),
                                  _d(jj), _d(J0), _d(J2), nthreads)
        if rc:
            raise RuntimeError(f"oracle se_pass failed ({rc})")
        if extra_info:
            return pp, J0, J2, rates, jj
        return pp, J0, J2

    def inverse_nodes(self, g, sp, xth, nvel, dx, nodes, n_depth=4096, nthreads=0, lev_p=None, star=None,
                      thin_exact=False):
        """inverse_rt for the given 1-based nodes -> J0, J2 (R x len(nodes)).  With
        ``thin_exact`` the emission factor 1 - e^-tau is evaluated as -expm1(-tau) (the
        mixed-precision path's definition) instead of as the reference writes it."""
        G = self.grid(g, lev_p)
        S = self.species(sp)
        P = self.params(xth, nvel, dx, g.gvec, g.bvec_all, n_depth, 0, star)
        nd = np.ascontiguousarray(nodes, dtype=np.int32)
        J0 = np.zeros((sp.r_num, len(nd)), order="F")
        J2 = np.zeros((sp.r_num, len(nd)), order="F")
        self.lib.orc_set_thin_exact(1 if thin_exact else 0)
        try:
            rc = self.lib.orc_inverse_nodes(C.byref(G), C.byref(S), C.byref(P), _i(nd), len(nd), _d(J0), _d(J2),
                                            nthreads)
        finally:
            self.lib.orc_set_thin_exact(0)
        if rc:
            raise RuntimeError(f"oracle inverse_nodes failed ({rc})")
        return J0, J2

    def polarize_prop(self, eI, eQ, pI, pQ, ds, I, Q, U, thin_exact=False):
        """One Stokes segment (raytrace.f90:1-56) -> (I, Q, U)."""
        a = [C.c_double(v) for v in (I, Q, U)]
        self.lib.orc_set_thin_exact(1 if thin_exact else 0)
        try:
            self.lib.orc_polarize_prop(*[C.c_double(v) for v in (eI, eQ, pI, pQ, ds)], *[C.byref(v) for v in a])
        finally:
            self.lib.orc_set_thin_exact(0)
        return tuple(v.value for v in a)

    def ray_in(self, temp_node, vel_node, p1_node, p2_node, dens_1, dens_2, temp_list, vel_list, length_list,
               nvel, dx, nu0, Aji, Bij, g1, g2, I_in, thin_exact=False):
        """ray_in (inverse_rt.f90:217-344) without dust over one gathered path: index 0
        of the per-cell arrays is the node's own cell -> the profile-weighted I_out."""
        nc = len(dens_1)
        z = np.zeros(nc)
        L = self.lib
        L.orc_ray_in.restype = C.c_double
        self.lib.orc_set_thin_exact(1 if thin_exact else 0)
        try:
            return L.orc_ray_in(C.c_double(temp_node), C.c_double(vel_node), C.c_double(p1_node),
                                C.c_double(p2_node), C.c_double(0.0), C.c_double(0.0), C.c_int(nc),
                                _d(_f64(dens_1)), _d(_f64(dens_2)), _d(z), C.c_double(0.0), _d(z),
                                _d(_f64(temp_list)), _d(_f64(vel_list)), _i(_i32(length_list)), C.c_int(nvel),
                                C.c_double(dx), C.c_double(nu0), C.c_double(Aji), C.c_double(Bij), C.c_int(g1),
                                C.c_int(g2), C.c_double(I_in))
        finally:
            self.lib.orc_set_thin_exact(0)

    def solve_from_J(self, g, sp, J0, J2, node_lo=1, node_hi=None):
        if node_hi is None:
            node_hi = g.bulk_node_num
        G = self.grid(g)
        S = self.species(sp)
        pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
        J0 = _f64(J0)
        J2 = _f64(J2)
        rc = self.lib.orc_solve_from_J(C.byref(G), C.byref(S), node_lo, node_hi, _d(J0), _d(J2), _d(pp))
        if rc:
            raise RuntimeError(f"oracle solve_from_J failed ({rc})")
        return pp

    def setup_stateq(self, sp, AC, Temp):
        S = self.species(sp)
        cd = sp.c_dim
        mats = [np.zeros((cd, cd), order="F") for _ in range(4)]
        AC = _f64(AC)
        self.lib.orc_setup_stateq(C.byref(S), _d(AC), C.c_double(Temp), *[_d(m) for m in mats])
        return mats

    def make_ac(self, sp, nH2, Temp, J0, J2):
        S = self.species(sp)
        AC = np.zeros((5, sp.n_trans), order="F")
        J0 = _f64(J0)
        J2 = _f64(J2)
        self.lib.orc_make_ac(C.byref(S), C.c_double(nH2), C.c_double(Temp), _d(J0), _d(J2), _d(AC))
        return AC

    def depol_table(self, sp):
        """D0_all (c_temp x e_num x (kmax - 1)) of orc_depol_table (depol.f90:242-317)."""
        S = self.species(sp)
        D = np.zeros((sp.c_temp, sp.e_num, max(sp.kmax - 1, 1)), order="F")
        if self.lib.orc_depol_table(C.byref(S), _d(D)):
            raise RuntimeError("no J = 0 level")
        return D

    def depol_at(self, sp, D0_all, Temp, nH2):
        """D0 (e_num x (kmax - 1)) of orc_depol_at (depol.f90:319-359)."""
        S = self.species(sp)
        D = np.zeros((sp.e_num, max(sp.kmax - 1, 1)), order="F")
        D0_all = _f64(D0_all)
        self.lib.orc_depol_at(C.byref(S), _d(D0_all), C.c_double(Temp), C.c_double(nH2), _d(D))
        return D

    def get_rhovec(self, sp, AC, Temp, extra_info=0, D0=None):
        """D0: None, or the depolarisation rates (e_num x (kmax - 1)) on the K >= 2 diagonal."""
        S = self.species(sp)
        rho = np.zeros(sp.c_dim)
        rates = np.zeros((3, sp.c_dim), order="F")
        AC = _f64(AC)
        D = _f64(D0) if D0 is not None else None
        rc = self.lib.orc_get_rhovec_d(C.byref(S), _d(AC), C.c_double(Temp), _d(D) if D is not None else None,
                                       _d(rho), extra_info, _d(rates))
        if rc:
            raise RuntimeError("singular SE matrix")
        return (rho, rates) if extra_info else rho

    def comp_emcoefs(self, r1_0, r2_0, r1_2, r2_2, j1, j2, th, nu0, Aji):
        """(eta_I, eta_Q, eps_I, eps_Q) of one line (em_coefs.f90:1-113)."""
        out = [C.c_double() for _ in range(4)]
        self.lib.orc_comp_emcoefs(C.c_double(r1_0), C.c_double(r2_0), C.c_double(r1_2), C.c_double(r2_2),
                                  int(j1), int(j2), C.c_double(th), C.c_double(nu0), C.c_double(Aji),
                                  *[C.byref(o) for o in out])
        return tuple(o.value for o in out)

    def coef(self, name, j, k, j2, k2):
        a = C.c_double()
        b = C.c_double()
        fn = getattr(self.lib, f"orc_{name}")
        if name == "t_spont_const":
            fn(j, k, j2, k2, C.byref(a))
            return a.value
        fn(j, k, j2, k2, C.byref(a), C.byref(b))
        return a.value, b.value

    def image_pass(self, g, sp, pol_pops, lines, th_inc, ph_az, nvel, vel_out, im_pix, im_size, dx,
                   alpha=0.15, n_depth_rt=4096, pix_lo=1, pix_hi=None, nthreads=0, bvec_all=None,
                   with_dust=False, star=None, node_pixels=False, thin_exact=False):
        """One inclination of the image pass.  lines = list of (l1, l2, nu0, Aji, kd).
        thin_exact: polarize_prop without the cancellation for thin cells (the
        mixed-precision path's definition) instead of as the reference writes it.
        star = dict(R_star=, T_star=, X_star=) (unit-sphere units, X_star before rotation)
        puts the star behind the pixels it covers (raytrace.f90:103-114).  node_pixels: one
        pixel per bulk node at its rotated (x, y), the reference's pix = 0 mode."""
        n_rt = g.bulk_node_num if node_pixels else (2 * im_pix + 1) ** 2
        if pix_hi is None:
            pix_hi = n_rt
        G = self.grid(g)
        S = self.species(sp)
        nl = len(lines)
        l1 = _i32([x[0] for x in lines])
        l2 = _i32([x[1] for x in lines])
        nu = _f64([x[2] for x in lines])
        A = _f64([x[3] for x in lines])
        kd = _f64([x[4] for x in lines])
        Cn = 2 * nvel + 1
        outs = [np.zeros((n_rt, nl, Cn), order="F") for _ in range(4)]
        dust = np.zeros((4, n_rt, nl, Cn)) if with_dust else None
        pp = _f64(pol_pops)
        vo = _f64(vel_out)
        bv = _f64(g.bvec_all if bvec_all is None else bvec_all)
        self.lib.orc_set_thin_exact(1 if thin_exact else 0)
        try:
            rc = self.lib.orc_image_pass(C.byref(G), C.byref(S), nl, _i(l1), _i(l2), _d(nu), _d(A), _d(kd),
                                         C.c_double(th_inc), C.c_double(ph_az), _d(bv), _d(pp), nvel, _d(vo), im_pix,
                                         C.c_double(im_size), C.c_double(dx), C.c_double(alpha), n_depth_rt, pix_lo,
                                         pix_hi, 1 if node_pixels else 0, 1 if star else 0,
                                         C.c_double(star["R_star"] if star else 0.0),
                                         C.c_double(star["T_star"] if star else 0.0),
                                         _d(_f64(star["X_star"])) if star else None,
                                         *[_d(o) for o in outs], _d(dust) if with_dust else None, nthreads)
        finally:
            self.lib.orc_set_thin_exact(0)
        if rc:
            raise RuntimeError(f"oracle image_pass failed ({rc})")
        if with_dust:
            return outs, dust
        return outs


class Reference:
    """The reference's own routines, compiled from its Fortran sources (never shipped)."""

    def __init__(self, path: str = LIB_REF):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)

    @staticmethod
    def _r(x, t=C.c_double):
        return C.byref(t(x))

    def eat_lime(self, pop, grid, mol, dust, max_ne_temp=300):
        """eat_lime_dim + eat_lime_no_delaunay (eat_lime.f90:1-313) on the four files: the
        arrays they fill, by name.  Fortran character(len=100) arguments get 100-character
        blank-padded buffers and their hidden lengths after the other arguments."""
        def s100(x):
            b = x.encode()
            assert len(b) <= 100
            return C.create_string_buffer(b.ljust(100), 100)
        fp, fg, fm, fd = (s100(x) for x in (pop, grid, mol, dust))
        L100 = C.c_size_t(100)
        dims = [C.c_int(0) for _ in range(7)]
        self.lib.eat_lime_dim_(fg, fm, *[C.byref(d) for d in dims], L100, L100)
        N, ndel, E, R, Cn, Ct, Cp = (d.value for d in dims)
        out = dict(node_num=N, del_num=ndel, e_num=E, r_num=R, c_num=Cn, c_temp=Ct, c_par=Cp)
        arr = dict(XYZ=np.zeros((3, N), order="F"), vel_all=np.zeros((3, N), order="F"),
                   temp_all=np.zeros((2, N), order="F"), lev_p=np.zeros((E, N), order="F"), e_levs=np.zeros(E),
                   t_Aji=np.zeros(R), nu=np.zeros(R), t_Cji=np.zeros((Ct, Cn, Cp), order="F"),
                   col_temps=np.zeros(Ct), h2_dens=np.zeros(N), b_all=np.zeros(N), ab_all=np.zeros(N),
                   j_lev=np.zeros(E, dtype=np.int32), r_trans=np.zeros((2, R), dtype=np.int32, order="F"),
                   c_trans=np.zeros((2, Cn), dtype=np.int32, order="F"), del_num_ne=np.zeros(N, dtype=np.int32),
                   del_ne=np.zeros((N, max_ne_temp), dtype=np.int32, order="F"), kd_v=np.zeros(R),
                   dust_p=np.zeros(N))
        maxne, bulk = C.c_int(0), C.c_int(0)
        rmax, dii, mass = C.c_double(0), C.c_double(0), C.c_double(0)
        ci = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
        p = lambda k: arr[k].ctypes.data_as(C.c_void_p)  # noqa: E731
        self.lib.eat_lime_no_delaunay_(
            fp, fg, fm, fd, ci(N), ci(ndel), ci(E), ci(R), ci(Cn), ci(Ct), p("XYZ"), p("vel_all"), p("temp_all"),
            p("lev_p"), p("e_levs"), p("t_Aji"), p("nu"), p("t_Cji"), p("col_temps"), ci(Cp), p("h2_dens"),
            p("b_all"), p("ab_all"), p("j_lev"), p("r_trans"), p("c_trans"), ci(max_ne_temp), C.byref(maxne),
            p("del_num_ne"), p("del_ne"), C.byref(rmax), C.byref(bulk), p("kd_v"), C.byref(dii), p("dust_p"),
            C.byref(mass), L100, L100, L100, L100)
        out.update(arr)
        out.update(maxne=maxne.value, rmax=rmax.value, bulk_node_num=bulk.value, mass_mol=mass.value)
        return out

    def legendre_all(self, nth, nph):
        xth = np.zeros((3, nth * nph), order="F")
        n = C.c_int(0)
        self.lib.legendre_all_(C.byref(C.c_int(nth)), C.byref(C.c_int(nph)), C.byref(n), _d(xth))
        return np.asfortranarray(xth[:, : n.value])

    def find_dir(self, bvec, gvec, th, ph):
        b = _f64(bvec)
        g = _f64(gvec)
        d = np.zeros(3)
        self.lib.find_dir_(_d(b), _d(g), self._r(th), self._r(ph), _d(d))
        return d

    def black_body(self, T, nu):
        out = C.c_double(0.0)
        self.lib.black_body_(self._r(T), self._r(nu), C.byref(out))
        return out.value

    def inverse_rt(self, node, g, sp, del_ne, xth, nvel, dx, bvec, gvec, n_depth, lev_p=None):
        """inverse_rt.f90:1-137 for one node (1-based) -> (J0, J2)."""
        N = g.node_num
        R = sp.r_num
        ne = _i32(del_ne)
        nnum = _i32(g.del_num_ne)
        X = _f64(g.XYZ)
        lp = _f64(g.lev_p if lev_p is None else lev_p)
        J0 = np.zeros(R)
        J2 = np.zeros(R)
        xt = _f64(xth)
        b = _f64(bvec)
        gv = _f64(gvec)
        rt = _i32(sp.r_trans)
        ci = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
        cd = lambda v: C.byref(C.c_double(float(v)))  # noqa: E731
        Xs = np.zeros(3)
        node_c = C.c_int(int(node))
        self.lib.inverse_rt_(
            _i(ne), _i(nnum), ci(N), ci(ne.shape[1]), C.byref(node_c), _d(X), ci(nvel), ci(xt.shape[1]), ci(0),
            _d(b), _d(gv), cd(dx), cd(g.tx), ci(R), _i(rt), _d(_f64(sp.nu_hz)), ci(sp.e_num), _d(lp),
            _i(_i32(sp.lev_g)), _d(_f64(g.b_all)), _d(_f64(g.vel_all)), _d(_f64(sp.t_Aji)), _d(_f64(sp.t_Bij)),
            _d(xt), _d(_f64(g.dust_p)), _d(_f64(sp.kd_v)), _d(_f64(g.temp_all[1])), _d(J0), _d(J2), ci(0), ci(0),
            cd(0.0), cd(0.0), ci(0), ci(0), _d(Xs), ci(n_depth))
        return J0, J2

    def inverse_nodes(self, g, sp, xth, nvel, dx, nodes, n_depth=4096, nthreads=0, lev_p=None):
        """inverse_rt.f90:1-137 for each node in `nodes` (1-based), as the reference's node
        loop (main_para.f90:210-257) runs it: one call per node, nodes spread over
        nthreads host threads (0: all cores).  The arrays are converted once; the Fortran
        routine keeps its locals on the stack, so concurrent calls are independent.
        Returns J0, J2 (r_num x len(nodes))."""
        from concurrent.futures import ThreadPoolExecutor
        N, R = g.node_num, sp.r_num
        ne = _i32(g.del_ne())
        keep = dict(ne=ne, nnum=_i32(g.del_num_ne), X=_f64(g.XYZ), lp=_f64(g.lev_p if lev_p is None else lev_p),
                    xt=_f64(xth), gv=_f64(g.gvec), rt=_i32(sp.r_trans), nu=_f64(sp.nu_hz), lg=_i32(sp.lev_g),
                    b=_f64(g.b_all), v=_f64(g.vel_all), A=_f64(sp.t_Aji), B=_f64(sp.t_Bij), dp=_f64(g.dust_p),
                    kd=_f64(sp.kd_v), td=_f64(g.temp_all[1]), bv=_f64(g.bvec_all))
        ints = {k: C.c_int(int(v)) for k, v in dict(N=N, maxne=ne.shape[1], nvel=nvel, ns=keep["xt"].shape[1],
                                                      zero=0, R=R, L=sp.e_num, nd=n_depth).items()}
        dbl = {k: C.c_double(float(v)) for k, v in dict(dx=dx, tx=g.tx, z=0.0).items()}
        nodes = [int(n) for n in nodes]
        J0 = np.zeros((R, len(nodes)), order="F")
        J2 = np.zeros((R, len(nodes)), order="F")
        Xs = np.zeros(3)
        k_ = keep

        def one(t):
            node = C.c_int(nodes[t])
            j0 = np.zeros(R)
            j2 = np.zeros(R)
            bvec = np.ascontiguousarray(k_["bv"][:, nodes[t] - 1])
            self.lib.inverse_rt_(
                _i(k_["ne"]), _i(k_["nnum"]), C.byref(ints["N"]), C.byref(ints["maxne"]), C.byref(node), _d(k_["X"]),
                C.byref(ints["nvel"]), C.byref(ints["ns"]), C.byref(ints["zero"]), _d(bvec), _d(k_["gv"]),
                C.byref(dbl["dx"]), C.byref(dbl["tx"]), C.byref(ints["R"]), _i(k_["rt"]), _d(k_["nu"]),
                C.byref(ints["L"]), _d(k_["lp"]), _i(k_["lg"]), _d(k_["b"]), _d(k_["v"]), _d(k_["A"]), _d(k_["B"]),
                _d(k_["xt"]), _d(k_["dp"]), _d(k_["kd"]), _d(k_["td"]), _d(j0), _d(j2), C.byref(ints["zero"]),
                C.byref(ints["zero"]), C.byref(dbl["z"]), C.byref(dbl["z"]), C.byref(ints["zero"]),
                C.byref(ints["zero"]), _d(Xs), C.byref(ints["nd"]))
            J0[:, t] = j0
            J2[:, t] = j2

        with ThreadPoolExecutor(max_workers=nthreads or os.cpu_count() or 1) as ex:
            list(ex.map(one, range(len(nodes))))
        return J0, J2

    def star_all(self, nth, nph, kvec, gvec, bvec, dOm):
        """star_all (miscellaneous.f90:1139-1226) -> (n_star, xth 3 x n_sphere)."""
        xth = np.zeros((3, nth * nph), order="F")
        ns, nv = C.c_int(0), C.c_int(0)
        self.lib.star_all_(C.byref(C.c_int(nth)), C.byref(C.c_int(nph)), _d(_f64(kvec)), _d(_f64(gvec)),
                           _d(_f64(bvec)), self._r(dOm), C.byref(ns), C.byref(nv), _d(xth))
        return ns.value, np.asfortranarray(xth[:, : nv.value])

    def circle_points(self, n, th_s, dth_s):
        x = np.zeros((3, n), order="F")
        self.lib.circle_points_(C.byref(C.c_int(n)), self._r(th_s), self._r(dth_s), _d(x))
        return x

    def find_angles(self, bvec, gvec, dir_):
        th, ph = C.c_double(0.0), C.c_double(0.0)
        self.lib.find_angles_(_d(_f64(bvec)), _d(_f64(gvec)), _d(_f64(dir_)), C.byref(th), C.byref(ph))
        return th.value, ph.value

    def inverse_rt_star(self, node, g, sp, del_ne, xth_s, n_star, nvel, dx, bvec, gvec, n_depth, X_star, R_star,
                        T_star):
        """inverse_rt.f90:1-137 in star mode for one node, with its star_all directions."""
        N = g.node_num
        R = sp.r_num
        ne = _i32(del_ne)
        nnum = _i32(g.del_num_ne)
        X = _f64(g.XYZ)
        lp = _f64(g.lev_p)
        J0 = np.zeros(R)
        J2 = np.zeros(R)
        xt = _f64(xth_s)
        ci = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
        cd = lambda v: C.byref(C.c_double(float(v)))  # noqa: E731
        node_c = C.c_int(int(node))
        self.lib.inverse_rt_(
            _i(ne), _i(nnum), ci(N), ci(ne.shape[1]), C.byref(node_c), _d(X), ci(nvel), ci(xt.shape[1]), ci(n_star),
            _d(_f64(bvec)), _d(_f64(gvec)), cd(dx), cd(g.tx), ci(R), _i(_i32(sp.r_trans)), _d(_f64(sp.nu_hz)),
            ci(sp.e_num), _d(lp), _i(_i32(sp.lev_g)), _d(_f64(g.b_all)), _d(_f64(g.vel_all)), _d(_f64(sp.t_Aji)),
            _d(_f64(sp.t_Bij)), _d(xt), _d(_f64(g.dust_p)), _d(_f64(sp.kd_v)), _d(_f64(g.temp_all[1])), _d(J0),
            _d(J2), ci(0), ci(1), cd(R_star), cd(T_star), ci(0), ci(0), _d(_f64(X_star)), ci(n_depth))
        return J0, J2

    def trace_out(self, node, g, del_ne, dir_, dx, n_depth):
        ne = _i32(del_ne)
        dl = np.zeros((2, n_depth + 1), dtype=np.int32, order="F")
        nc = C.c_int(0)
        Xs = np.zeros(3)
        d = _f64(dir_)
        self.lib.trace_out_(C.byref(C.c_int(int(node))), C.byref(C.c_int(g.node_num)), _d(_f64(g.XYZ)), _i(ne),
                            C.byref(C.c_int(ne.shape[1])), _i(_i32(g.del_num_ne)), _d(d), C.byref(C.c_double(dx)),
                            _i(dl), C.byref(nc), C.byref(C.c_int(n_depth + 1)), C.byref(C.c_int(0)), _d(Xs),
                            C.byref(C.c_double(0.0)))
        return dl[:, : nc.value].copy()

    def make_ac(self, sp, nH2, Temp, J0, J2):
        AC = np.zeros((5, sp.n_trans), order="F")
        ci = lambda v: C.byref(C.c_int(int(v)))  # noqa: E731
        self.lib.make_ac_(_d(_f64(sp.t_Aji)), _d(_f64(sp.t_Cji)), _i(_i32(sp.r_trans)), _i(_i32(sp.c_trans)),
                          ci(sp.r_num), ci(sp.c_num), ci(sp.c_temp), ci(sp.c_par), ci(sp.e_num),
                          _d(_f64(sp.e_levs_hz)), C.byref(C.c_double(nH2)), C.byref(C.c_double(Temp)),
                          _d(_f64(sp.col_temps)), ci(sp.n_trans), _i(_i32(sp.all_trans[:, : sp.n_trans])),
                          _d(_f64(J0)), _d(_f64(J2)), _d(AC))
        return AC

    def rotate_all_save(self, g, th, ph, bvec_all, X_star=None):
        N = g.node_num
        vn = np.zeros((3, N), order="F")
        bn = np.zeros((3, N), order="F")
        xn = np.zeros((3, N), order="F")
        xs = _f64(np.zeros(3) if X_star is None else X_star)
        xsn = np.zeros(3)
        self.lib.rotate_all_save_(C.byref(C.c_int(N)), C.byref(C.c_double(th)), C.byref(C.c_double(ph)),
                                  _d(_f64(g.vel_all)), _d(_f64(bvec_all)), _d(_f64(g.XYZ)), _d(xs), _d(vn), _d(bn),
                                  _d(xn), _d(xsn))
        if X_star is not None:
            return vn, bn, xn, xsn
        return vn, bn, xn

    def trace_out_raytrace(self, x, y, g, X_inc, del_ne, dx, n_depth_rt, nu0_all, star=None):
        """star = dict(R_star=, T_star=, Xs=) with Xs the rotated star position."""
        ne = _i32(del_ne)
        dl = np.zeros((2, n_depth_rt), dtype=np.int32, order="F")
        nc = C.c_int(0)
        nlev = len(nu0_all)
        I_in = np.zeros(nlev)
        d = np.array([0.0, 0.0, 1.0])
        xs = _f64(star["Xs"]) if star else np.zeros(3)
        sm, Rs, Ts = (1, star["R_star"], star["T_star"]) if star else (0, 0.0, 0.0)
        self.lib.trace_out_raytrace_(C.byref(C.c_double(x)), C.byref(C.c_double(y)), C.byref(C.c_int(g.node_num)),
                                     _d(_f64(X_inc)), C.byref(C.c_int(sm)), C.byref(C.c_double(Rs)), _d(xs),
                                     C.byref(C.c_double(Ts)), _i(ne), C.byref(C.c_int(ne.shape[1])),
                                     _i(_i32(g.del_num_ne)), _d(d), C.byref(C.c_double(dx)),
                                     C.byref(C.c_int(n_depth_rt)), _i(dl), C.byref(nc), C.byref(C.c_int(nlev)),
                                     _d(_f64(nu0_all)), _d(I_in))
        return dl[:, : nc.value].copy(), I_in

    def polarize_prop(self, eI, eQ, pI, pQ, ds, I, Q, U):
        a = [C.c_double(v) for v in (I, Q, U)]
        self.lib.polarize_prop_(*[C.byref(C.c_double(v)) for v in (eI, eQ, pI, pQ, ds)], *[C.byref(v) for v in a])
        return tuple(v.value for v in a)

    def rot_stokes(self, Q, U, b, back=False):
        qn = C.c_double()
        un = C.c_double()
        fn = self.lib.rot_stokes_back_ if back else self.lib.rot_stokes_
        fn(C.byref(C.c_double(Q)), C.byref(C.c_double(U)), C.byref(qn), C.byref(un), _d(_f64(b)))
        return qn.value, un.value

    def ray_in_raytrace(self, vel_obs, nc, pol, b_list, bvec_list, dust, alpha, kd, bb_list, vel_list, length_list,
                        dx, nu0, I_in):
        outs = [C.c_double() for _ in range(8)]
        self.lib.ray_in_raytrace_(C.byref(C.c_double(vel_obs)), C.byref(C.c_int(nc)), _d(_f64(pol)),
                                  _d(_f64(b_list)), _d(_f64(bvec_list)), _d(_f64(dust)), C.byref(C.c_double(alpha)),
                                  C.byref(C.c_double(kd)), _d(_f64(bb_list)), _d(_f64(vel_list)),
                                  _i(_i32(length_list)), C.byref(C.c_double(dx)), C.byref(C.c_double(nu0)),
                                  C.byref(C.c_double(I_in)), *[C.byref(o) for o in outs])
        return np.array([o.value for o in outs])
