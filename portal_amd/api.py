"""Python mirror of the drop-in boundary: one object per device context.

Method names follow the reference's loops they replace (see include/portal_hip.h):
``se_pass`` is the node loop of main_para.f90:210-257 (trace_node per node),
``inverse_rt`` the radiation-field part of it, ``se_solve`` the SE part,
``image_pass`` one inclination of para_finish.f90:382-449.  Arrays are numpy,
column-major as in the reference; errors raise ``PortalError`` carrying the
library's status code (the reference would ``stop``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


class PortalError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{L.STATUS.get(code, code)}: {msg}")
        self.code = code


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))


def _d(a):
    return a.ctypes.data_as(L.dp)


def _i(a):
    return a.ctypes.data_as(L.ip)


class _Core:
    def __init__(self, device: int = 0):
        self.lib = L.load()
        self.ctx = C.c_void_p()
        rc = self.lib.portal_ctx_create(C.byref(self.ctx), device)
        if rc:
            raise PortalError(rc, "no usable GPU (portal_ctx_create)")
        self._keep = []
        self.grid = None
        self.species = None

    def close(self):
        if self.ctx:
            self.lib.portal_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise PortalError(rc, self.lib.portal_last_error(self.ctx).decode())

    # -- inputs ------------------------------------------------------------------
    def upload_grid(self, g):
        del_ne = g.del_ne()
        self._check(self.lib.portal_upload_grid(
            self.ctx, g.node_num, g.bulk_node_num, _d(_f64(g.XYZ)), _d(_f64(g.vel_all)), _d(_f64(g.temp_all)),
            _d(_f64(g.b_all)), _d(_f64(g.h2_dens)), _d(_f64(g.ab_all)), _d(_f64(g.dust_p)), _i(del_ne),
            _i(_i32(g.del_num_ne)), del_ne.shape[1], C.c_double(g.tx)))
        self.grid = g

    def upload_species(self, sp):
        self._check(self.lib.portal_upload_species(
            self.ctx, sp.e_num, _i(_i32(sp.j_lev)), _d(_f64(sp.e_levs_hz)), sp.r_num, _i(_i32(sp.r_trans)),
            _d(_f64(sp.t_Aji)), _d(_f64(sp.nu_hz)), _d(_f64(sp.t_Bij)), _d(_f64(sp.kd_v)), sp.c_num,
            _i(_i32(sp.c_trans)), sp.c_temp, sp.c_par, _d(_f64(sp.t_Cji)), _d(_f64(sp.col_temps)), sp.kmax))
        self.species = sp

    def set_levpop(self, lev_p):
        a = _f64(lev_p)
        self._check(self.lib.portal_set_levpop(self.ctx, _d(a)))

    def get_levpop(self):
        out = np.zeros((self.species.e_num, self.grid.node_num), order="F")
        self._check(self.lib.portal_get_levpop(self.ctx, _d(out)))
        return out

    def species_info(self):
        nt, cd = C.c_int(), C.c_int()
        self._check(self.lib.portal_species_info(self.ctx, C.byref(nt), C.byref(cd)))
        return nt.value, cd.value

    # collisional depolarisation (depol.f90:242-359), off by default as in the reference
    def depol_table(self):
        """D0_all (c_temp x e_num x (kmax - 1)) of de_polarization for the uploaded species."""
        sp = self.species
        D = np.zeros((sp.c_temp, sp.e_num, max(sp.kmax - 1, 1)), order="F")
        self._check(self.lib.portal_depol_table(self.ctx, D.ctypes.data_as(C.c_void_p)))
        return D

    def set_depol(self, D0_all):
        """Subtract D^(K)(T) n_H2 1e-6 on the K >= 2 diagonal in every later SE solve (None: off)."""
        if D0_all is None:
            self._check(self.lib.portal_set_depol(self.ctx, None))
            return
        D = np.asfortranarray(D0_all, dtype=np.float64)
        sp = self.species
        if D.shape != (sp.c_temp, sp.e_num, sp.kmax - 1):
            raise ValueError(f"depolarisation table shape {D.shape}, expected {(sp.c_temp, sp.e_num, sp.kmax - 1)}")
        self._check(self.lib.portal_set_depol(self.ctx, D.ctypes.data_as(C.c_void_p)))

    # star mode (main_sub.f90:56-79): None, or dict(nth, nph, R_star, T_star, X_star) used by
    # every later pass; the directions then come from star_all instead of xth
    star = None

    def _se_params(self, xth, nvel, dx, n_depth, extra_info=0, precision=L.PORTAL_FP64, gvec=None, bvec_all=None):
        g = self.grid
        st = self.star
        arrs = (_f64(xth), _f64(g.gvec if gvec is None else gvec), _f64(g.bvec_all if bvec_all is None else bvec_all),
                _f64(st["X_star"] if st else np.zeros(3)))
        self._keep.append(arrs)
        return L.SeParams(arrs[0].shape[1], _d(arrs[0]), nvel, dx, _d(arrs[1]), _d(arrs[2]), n_depth, extra_info,
                          precision, 1 if st else 0, st["nth"] if st else 0, st["nph"] if st else 0,
                          st["R_star"] if st else 0.0, st["T_star"] if st else 0.0, _d(arrs[3]))

    # -- passes ------------------------------------------------------------------
    def se_pass(self, xth, nvel, dx, n_depth=4096, extra_info=0, precision=L.PORTAL_FP64, gvec=None, bvec_all=None):
        """trace_node for every bulk node.  Returns pol_pops (N x 2L)[, pol_rates, JJ_20], stats."""
        g, sp = self.grid, self.species
        p = self._se_params(xth, nvel, dx, n_depth, extra_info, precision, gvec, bvec_all)
        pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
        rates = np.zeros((g.node_num, 3, 2 * sp.e_num), order="F")
        jj = np.zeros((g.node_num, sp.r_num), order="F")
        st = L.Stats()
        self._check(self.lib.portal_se_pass(self.ctx, C.byref(p), _d(pp), _d(rates), _d(jj), C.byref(st)))
        if extra_info:
            return pp, rates, jj, st.as_dict()
        return pp, st.as_dict()

    def inverse_rt(self, xth, nvel, dx, n_depth=4096, precision=L.PORTAL_FP64, gvec=None, bvec_all=None):
        g, sp = self.grid, self.species
        p = self._se_params(xth, nvel, dx, n_depth, 0, precision, gvec, bvec_all)
        J0 = np.zeros((sp.r_num, g.node_num), order="F")
        J2 = np.zeros((sp.r_num, g.node_num), order="F")
        st = L.Stats()
        self._check(self.lib.portal_inverse_rt(self.ctx, C.byref(p), _d(J0), _d(J2), C.byref(st)))
        return J0, J2, st.as_dict()

    def inverse_rt_device(self, xth, nvel, dx, d_J0, d_J2, n_depth=4096, precision=L.PORTAL_FP64):
        """portal_inverse_rt_device: J0/J2 into caller-owned device buffers (addresses of
        r_num x node_num doubles), allocating nothing; needs an earlier inverse_rt or
        se_pass with the same directions on this context."""
        p = self._se_params(xth, nvel, dx, n_depth, 0, precision)
        self._check(self.lib.portal_inverse_rt_device(self.ctx, C.byref(p), C.c_void_p(d_J0), C.c_void_p(d_J2)))

    def se_solve(self, J0, J2, rates=False):
        g, sp = self.grid, self.species
        pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
        rr = np.zeros((g.node_num, 3, 2 * sp.e_num), order="F")
        st = L.Stats()
        a0, a2 = _f64(J0), _f64(J2)
        self._check(self.lib.portal_se_solve(self.ctx, _d(a0), _d(a2), _d(pp), _d(rr) if rates else None,
                                             C.byref(st)))
        return (pp, rr) if rates else pp

    def lambda_iterate(self, xth, nvel, dx, max_iter, tol_pop=1e-6, tol_align=1e-8, n_depth=4096,
                       precision=L.PORTAL_FP64, surrogate=0, fallback_tol=1e-8, rebuild_paths=0, outputs=True,
                       ng_accel=0, ali=0):
        g, sp = self.grid, self.species
        p = self._se_params(xth, nvel, dx, n_depth, 0, precision)
        it = L.IterParams(max_iter, tol_pop, tol_align, surrogate, fallback_tol, rebuild_paths, ng_accel, ali)
        st = L.Stats()
        if not outputs:
            self._check(self.lib.portal_lambda_iterate(self.ctx, C.byref(p), C.byref(it), None, None, C.byref(st)))
            return st.as_dict()
        pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
        lev = np.zeros((sp.e_num, g.node_num), order="F")
        self._check(self.lib.portal_lambda_iterate(self.ctx, C.byref(p), C.byref(it), _d(pp), _d(lev), C.byref(st)))
        return pp, lev, st.as_dict()

    def exp_peak(self, iters=4096):
        """Measured v_exp_f32 throughput of the device (exponentials per second)."""
        out = C.c_double()
        self._check(self.lib.portal_debug_exp_peak(self.ctx, int(iters), C.byref(out)))
        return out.value

    def iter_trace(self, max_iter=1000):
        """(population change, alignment change) of each iteration of the last lambda call."""
        out = np.zeros(2 * max_iter)
        n = self.lib.portal_debug_iter_trace(self.ctx, _d(out), max_iter)
        return out[: 2 * n].reshape(n, 2)

    def coop_unit(self, cells=0):
        """Test hook: cells per staged unit of the cooperative kernel (0: the default)."""
        self._check(self.lib.portal_debug_coop_unit(self.ctx, int(cells)))

    def ng_schedule(self, first=3, gap=3):
        """Test hook: first iteration and least gap of the Ng extrapolations."""
        self._check(self.lib.portal_debug_ng_schedule(self.ctx, int(first), int(gap)))

    def walk_chunks(self, scale):
        """Test hook: size the single-pass walk's chunk store at scale x its sampled estimate
        (a small scale forces the second-walk fallback, scale <= 0 the two-pass walk)."""
        self._check(self.lib.portal_debug_walk_chunks(self.ctx, C.c_double(scale)))

    def force_rt_kernel(self, precision, variant, waves=4):
        """Test hook: fix the inward-integration kernel (0 global, 1 staged, 2 cooperative,
        3 cooperative with each ray's cells back to front) and its waves per workgroup for
        one precision; variant -1 returns to autotuning."""
        self._check(self.lib.portal_debug_force_rt(self.ctx, int(precision), int(variant), int(waves)))

    def force_batch_rt(self, form=-1):
        """Test hook: the cooperative form of a batch this context leads (-1 timed; bit 0 one
        workgroup per item, else persistent workgroups; bit 1 rays back to front)."""
        self._check(self.lib.portal_debug_force_batch_rt(self.ctx, int(form)))

    def surrogate_perturb(self, stride, factor=1e3):
        """Test hook: every stride-th work item's surrogate prediction gets its first level's
        population scaled by factor, renormalised (stride 0: off)."""
        self._check(self.lib.portal_debug_surrogate_perturb(self.ctx, int(stride), C.c_double(factor)))

    def fallback_map(self):
        """Per work item of the last surrogate run: 0 kept, 1 failed the validity tests,
        2 reverted by the residual test."""
        out = np.zeros(max(self.grid.node_num, 1), dtype=np.int32)  # work items <= nodes
        m = self.lib.portal_debug_fallback_map(self.ctx, out.ctypes.data_as(C.POINTER(C.c_int)), len(out))
        if m < 0:
            self._check(m)
        return out[:m]

    def force_se(self, form=-1, waves=0):
        """Test hook: waves per system of this context's SE solve (1, 2, 4; 0 resets to 1);
        for a batch this context leads, the SE form (-1 autotuned, 0 one persistent launch,
        1 a launch per class of system sizes) and the classes' waves (0 autotuned)."""
        self._check(self.lib.portal_debug_force_se(self.ctx, int(form), int(waves)))

    def get_pol_pops(self):
        """pol_pops of the last SE solve (N x 2L) without recomputing."""
        g, sp = self.grid, self.species
        pp = np.zeros((g.node_num, 2 * sp.e_num), order="F")
        self._check(self.lib.portal_get_pol_pops(self.ctx, _d(pp)))
        return pp

    def image_pass(self, pol_pops, lines, th_inc, ph_az, nvel, vel_out, im_pix, im_size, dx, alpha=0.15,
                   n_depth_rt=4096, precision=L.PORTAL_FP64, bvec_all=None, with_dust=False, star=None,
                   node_pixels=False):
        """One inclination.  lines = [(l1, l2, nu0, Aji, kd)].  Returns I, Q, U, TAU (n_rt x nlines x C).
        star (default: the context's star mode) = dict(R_star, T_star, X_star) puts the star
        behind the pixels it covers (raytrace.f90:103-114).  node_pixels: one pixel per bulk
        node at its rotated (x, y) (the reference's pix = 0 mode); outputs are then
        bulk_node_num x nlines x C."""
        g = self.grid
        nl = len(lines)
        st = self.star if star is None else star
        arrs = (_i32([x[0] for x in lines]), _i32([x[1] for x in lines]), _f64([x[2] for x in lines]),
                _f64([x[3] for x in lines]), _f64([x[4] for x in lines]), _f64(vel_out),
                _f64(st["X_star"] if st else np.zeros(3)))
        self._keep.append(arrs)
        p = L.ImgParams(nl, _i(arrs[0]), _i(arrs[1]), _d(arrs[2]), _d(arrs[3]), _d(arrs[4]), th_inc, ph_az, nvel,
                        _d(arrs[5]), im_pix, im_size, dx, alpha, n_depth_rt, precision, 1 if st else 0,
                        st["R_star"] if st else 0.0, st["T_star"] if st else 0.0, _d(arrs[6]),
                        1 if node_pixels else 0)
        n_rt = g.bulk_node_num if node_pixels else (2 * im_pix + 1) ** 2
        Cn = 2 * nvel + 1
        outs = [np.zeros((n_rt, nl, Cn), order="F") for _ in range(4)]
        dust = np.zeros((4, n_rt, nl, Cn)) if with_dust else None
        bv = _f64(g.bvec_all if bvec_all is None else bvec_all)
        pp = _f64(pol_pops)
        st = L.Stats()
        self._check(self.lib.portal_image_pass(self.ctx, C.byref(p), _d(pp), _d(bv), *[_d(o) for o in outs],
                                               _d(dust) if with_dust else None, C.byref(st)))
        if with_dust:
            return outs, dust, st.as_dict()
        return outs, st.as_dict()


class SurrogateMixin:
    """Surrogate-network calls of the C ABI (portal_mlp_*, portal_surrogate_features)."""

    def mlp_create(self, n_in, width=256, depth=4, n_out=1, seed=0):
        self._check(self.lib.portal_mlp_create(self.ctx, n_in, width, depth, n_out, seed))
        self.mlp_dims = (n_in, width, depth, n_out)
        return int(self.lib.portal_mlp_num_params(self.ctx))

    def mlp_set_params(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self._check(self.lib.portal_mlp_set_params(self.ctx, None if p is None else p.ctypes.data_as(C.c_void_p),
                                                   None if n is None else n.ctypes.data_as(C.c_void_p)))

    def mlp_get_params(self):
        n_in, _, _, n_out = self.mlp_dims
        p = np.zeros(int(self.lib.portal_mlp_num_params(self.ctx)), dtype=np.float32)
        nm = np.zeros(2 * n_in + 2 * n_out, dtype=np.float32)
        self._check(self.lib.portal_mlp_get_params(self.ctx, p.ctypes.data_as(C.c_void_p), nm.ctypes.data_as(C.c_void_p)))
        return p, nm

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], self.mlp_dims[3]), dtype=np.float32)
        self._check(self.lib.portal_mlp_forward(self.ctx, x.shape[0], x.ctypes.data_as(C.c_void_p),
                                                out.ctypes.data_as(C.c_void_p)))
        return out

    def mlp_train_step(self, x, t, lr, want_grads=False):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(int(self.lib.portal_mlp_num_params(self.ctx)), dtype=np.float32) if want_grads else None
        self._check(self.lib.portal_mlp_train_step(self.ctx, x.shape[0], x.ctypes.data_as(C.c_void_p),
                                                   t.ctypes.data_as(C.c_void_p), C.c_float(lr), C.byref(loss),
                                                   None if g is None else g.ctypes.data_as(C.c_void_p)))
        return (loss.value, g) if want_grads else loss.value

    def mlp_train(self, x, t, epochs, batch, lr, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        ms = C.c_double()
        self._check(self.lib.portal_mlp_train(self.ctx, x.shape[0], x.ctypes.data_as(C.c_void_p),
                                              t.ctypes.data_as(C.c_void_p), epochs, batch, C.c_float(lr), seed,
                                              hist.ctypes.data_as(C.c_void_p), C.byref(ms)))
        return hist, ms.value

    def surrogate_features(self):
        sp = self.species
        n_in = 2 + 2 * sp.r_num + sp.e_num + int((np.asarray(sp.j_lev) > 0).sum())
        nw_max = self.grid.bulk_node_num
        f = np.zeros((nw_max, n_in), dtype=np.float32)
        nodes = np.zeros(nw_max, dtype=np.int32)
        nw = C.c_int()
        self._check(self.lib.portal_surrogate_features(self.ctx, n_in, f.ctypes.data_as(C.c_void_p),
                                                       nodes.ctypes.data_as(C.c_void_p), C.byref(nw)))
        return f[: nw.value], nodes[: nw.value]



class Portal(SurrogateMixin, _Core):
    """One device context: grid/species uploads, SE/image passes, lambda iteration, surrogate network."""


def batch_lambda_iterate(portals, xths, nvels, dx, max_iter, tol_pop=1e-6, tol_align=1e-8, n_depth=4096,
                         precision=L.PORTAL_MIXED, rebuild_paths=0, ng_accel=0, surrogate=0, fallback_tol=1e-8):
    """portal_batch_lambda_iterate over several Portal contexts (one model each).  With
    surrogate = 1, every model whose context holds a network gets its prediction after the
    first pass (the others iterate plainly)."""
    n = len(portals)
    lib = portals[0].lib
    arr = (C.c_void_p * n)(*[p.ctx.value for p in portals])
    ps = (L.SeParams * n)(*[p._se_params(x, nv, dx, n_depth, 0, precision) for p, x, nv in zip(portals, xths, nvels)])
    it = L.IterParams(max_iter, tol_pop, tol_align, surrogate, fallback_tol, rebuild_paths, ng_accel)
    per = (L.Stats * n)()
    tot = L.Stats()
    rc = lib.portal_batch_lambda_iterate(arr, n, ps, C.byref(it), per, C.byref(tot))
    if rc:
        raise PortalError(rc, lib.portal_last_error(portals[0].ctx).decode())
    return [s.as_dict() for s in per], tot.as_dict()
