// Image pass of one inclination (FP64): the three OpenMP loops of
// para_finish.f90:382-449.
//
//   rotate_kernel     rotate_all_save (miscellaneous.f90:844-907) on XYZ, v, B
//   emcoef_kernel     comp_emcoefs (em_coefs.f90:1-113) per bulk node and imaged line,
//                     plus the per-cell rotation angles of rot_stokes
//                     (miscellaneous.f90:463-551) and of the dust branch
//   start_kernel      each pixel's entry point (raytrace.f90:98-118: the far side of the
//                     unit sphere, or in star mode the star's near side for pixels on
//                     its disc, flagged in start.w) and the nearest node to it, the
//                     brute-force O(N) search of trace_out_raytrace
//                     (raytrace.f90:126-134): node
//                     positions are staged through LDS and every lane scans them for
//                     its pixel, first minimal distance wins (strict <)
//   march_kernel      the fixed-step march of trace_out_raytrace (raytrace.f90:141-167)
//                     with check_neigh (miscellaneous.f90:1-46), one lane per pixel;
//                     the photon position is advanced by repeated addition exactly as
//                     the reference does, so each cell change lands on the same step
//   stokes_kernel     ray_in_raytrace (raytrace.f90:179-323) with polarize_prop
//                     (raytrace.f90:1-56), one lane per (pixel, line, channel)
// Compiled with -ffp-contract=off.
#include "common.h"
#include "image.h"

namespace portal {

__global__ void __launch_bounds__(256) rotate_kernel(ImgArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const double *R = a.Rtot;  // column-major 3x3
    double v[3], b[3], x[3];
    for (int j = 0; j < 3; j++) {
        double sv = 0.0, sb = 0.0, sx = 0.0;
        for (int k = 0; k < 3; k++) {
            sv = sv + R[j + 3 * k] * a.vel[3 * (size_t)i + k];
            sb = sb + R[j + 3 * k] * a.bvec[3 * (size_t)i + k];
            sx = sx + R[j + 3 * k] * a.XYZ[3 * (size_t)i + k];
        }
        v[j] = sv;
        b[j] = sb;
        x[j] = sx;
    }
    a.pos_r[i] = make_double4(x[0], x[1], x[2], 0.0);
    a.vz_r[i] = v[2];
    // rot_stokes angle: chi = (bx, by, 0)/|.|, alpha = acos(chi_y), sign flipped if chi_x > 0
    const double n = sqrt(b[0] * b[0] + b[1] * b[1] + 0.0 * 0.0);
    const double chi1 = b[0] / n, chi2 = b[1] / n;
    double alpha = acos(chi2);
    if (chi1 > 0.0) alpha = -alpha;
    // dust angle: phi = acos(by / sqrt(bx^2 + by^2)), sign flipped if bx < 0
    double phi = acos(b[1] / sqrt(b[0] * b[0] + b[1] * b[1]));
    if (b[0] < 0.0) phi = -phi;
    CellGeo cg;
    cg.c2a = cos(2.0 * alpha);
    cg.s2a = sin(2.0 * alpha);
    cg.mu2 = 1.0 - b[2] * b[2];
    cg.c2p = cos(2.0 * phi);
    cg.s2p = sin(2.0 * phi);
    cg.bz = b[2];
    a.geo[i] = cg;
}

__device__ static double wig6_1122(int j1, int j2, const double *tab, int jmax) {
    return tab[j1 * (jmax + 1) + j2];
}

__global__ void __launch_bounds__(256) emcoef_kernel(ImgArgs a) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)a.N * a.nlines) return;
    const int node = (int)(id / a.nlines), j = (int)(id % a.nlines);
    double4 out = make_double4(0.0, 0.0, 0.0, 0.0);
    const size_t N = a.N;
    if (node < a.N_b && a.pol[node] != 0.0) {
        const int lo = a.l1[j], up = a.l2[j];
        const double r1_0 = a.pol[node + N * (2 * lo)], r2_0 = a.pol[node + N * (2 * up)];
        const double r1_2 = a.pol[node + N * (2 * lo + 1)], r2_2 = a.pol[node + N * (2 * up + 1)];
        const int j1 = a.jlev[lo], j2 = a.jlev[up];
        const double th = acos(a.geo[node].bz);
        const double nu0 = a.nu0[j], Aji = a.Aji[j];
        const double fas = pow(-1.0, 1.0 + j1 + j2);
        const double fac1 = sqrt(3.0 * (2.0 * j1 + 1.0));
        const double fac2 = sqrt(3.0 * (2.0 * j2 + 1.0));
        const double w6_1 = wig6_1122(j1, j2, a.w6tab, a.jmax);
        const double w6_2 = wig6_1122(j2, j1, a.w6tab, a.jmax);
        const double wjj2_1 = fas * fac1 * w6_1;
        const double wjj2_2 = fas * fac2 * w6_2;
        const double g1 = 2.0 * j1 + 1.0, g2 = 2.0 * j2 + 1.0;
        const double ct = cos(th), sn = sin(th);
        const double tki = (3.0 * (ct * ct) - 1.0) / (2.0 * sqrt(2.0));
        const double tkq = -(3.0 * (sn * sn) / (2.0 * sqrt(2.0)));
        const double na_I = 1.0 + wjj2_1 * (r1_2 / r1_0) * tki;
        const double na_Q = wjj2_1 * (r1_2 / r1_0) * tkq;
        const double ns_I = 1.0 + wjj2_2 * (r2_2 / r2_0) * tki;
        const double ns_Q = wjj2_2 * (r2_2 / r2_0) * tkq;
        const double a_to_b = 6.781962e49;
        const double b_to_a = pow(nu0, 3.0) / a_to_b;
        const double Bji = Aji / b_to_a;
        double es_I = ns_I, es_Q = ns_Q;
        double nsI = sqrt(g1) * r1_0 * na_I - sqrt(g2) * r2_0 * ns_I * g1 / g2;
        double nsQ = sqrt(g1) * r1_0 * na_Q - sqrt(g2) * r2_0 * ns_Q * g1 / g2;
        const double hv_fac = 5.273e-35 * nu0;
        es_I = hv_fac * Aji * es_I * sqrt(g2) * r2_0;
        es_Q = hv_fac * Aji * es_Q * sqrt(g2) * r2_0;
        nsI = hv_fac * (g2 / g1) * Bji * nsI;
        nsQ = hv_fac * (g2 / g1) * Bji * nsQ;
        out = make_double4(nsI, nsQ, es_I, es_Q);
    }
    a.polco[id] = out;
}

// nearest node to every pixel's start point (strict <, first index wins)
__global__ void __launch_bounds__(256) start_kernel(ImgArgs a) {
    __shared__ double4 tile[1024];
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = pix < a.n_rt;
    double sx = 0, sy = 0, sz = 0;
    if (active) {
        if (a.node_pixels) {  // para_finish.f90:296, the node's rotated position
            const double4 p = a.pos_r[pix];
            sx = p.x;
            sy = p.y;
        } else {  // para_finish.f90:369-376
            const int ii = pix / a.n_side - a.im_pix, jj = pix % a.n_side - a.im_pix;
            sx = ii * a.im_size / a.im_pix;
            sy = jj * a.im_size / a.im_pix;
        }
        sz = -1.0 * sqrt(1.0 - sx * sx - sy * sy);
        double on_star = 0.0;
        if (a.star) {
            const double xc = sx - a.Xs0, yc = sy - a.Xs1;
            if (sqrt(xc * xc + yc * yc) < a.R_star) {
                sz = a.Xs2 + a.R_star;
                on_star = 1.0;
            }
        }
        a.start[pix] = make_double4(sx, sy, sz, on_star);
    }
    double rsave = 10.0;
    int node = -1;
    for (int base = 0; base < a.N; base += 1024) {
        const int cnt = min(1024, a.N - base);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) tile[t] = a.pos_r[base + t];
        __syncthreads();
        if (active)
            for (int t = 0; t < cnt; t++) {
                const double4 p = tile[t];
                const double dx = sx - p.x, dy = sy - p.y, dz = sz - p.z;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r < rsave) {
                    rsave = r;
                    node = base + t;
                }
            }
    }
    if (active) a.start_node[pix] = node;
}

__global__ void __launch_bounds__(256) march_kernel(ImgArgs a) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= a.n_rt) return;
    const double4 st = a.start[pix];
    const double xp0 = st.x, xp1 = st.y;
    double xp2 = st.z;
    int node = a.start_node[pix];
    int *cells = a.rec_cell + (size_t)pix * a.n_depth_rt;
    int *steps = a.rec_step + (size_t)pix * a.n_depth_rt;
    int nc = 0, k_cur = 0;
    bool in_sim = true, over = false;
    const double dx = a.dx;
    if (node < 0 || !(xp2 == xp2)) {  // no start node, or a start point off the sphere
        a.rec_n[pix] = -1;
        atomicOr(a.flags, FLAG_START);
        return;
    }
    // the march crosses the unit sphere in at most 2/dx steps; stop well after that
    const long long max_steps = (long long)(2.0 / dx) + 16;
    while (in_sim) {
        if (k_cur > max_steps) {
            over = true;
            break;
        }
        // X_phot = X_phot + dx*dir with dir = (0,0,1): x and y are unchanged
        xp2 = xp2 + dx * 1.0;
        k_cur++;
        if (sqrt(xp0 * xp0 + xp1 * xp1 + xp2 * xp2) > 1.0) {
            nc++;
            in_sim = false;
            if (nc <= a.n_depth_rt) {
                steps[nc - 1] = k_cur;
                cells[nc - 1] = node;
            } else {
                over = true;
            }
        }
        const double4 pn = a.pos_r[node];
        const double a0 = xp0 - pn.x, a1 = xp1 - pn.y, a2 = xp2 - pn.z;
        double dis0 = a0 * a0 + a1 * a1 + a2 * a2;
        int nn = node;
        const int qe = a.nbr_ptr[node + 1];
        for (int q = a.nbr_ptr[node]; q < qe; q++) {
            const int nb = a.nbr_idx[q];
            const double4 pb = a.pos_r[nb];
            const double b0 = xp0 - pb.x, b1 = xp1 - pb.y, b2 = xp2 - pb.z;
            const double dis = b0 * b0 + b1 * b1 + b2 * b2;
            if (dis < dis0) {
                dis0 = dis;
                nn = nb;
            }
        }
        if (nn != node) {
            nc++;
            if (nc <= a.n_depth_rt) {
                steps[nc - 1] = k_cur;
                cells[nc - 1] = nn;
            } else {
                over = true;
            }
            node = nn;
        }
        if (over) break;
    }
    a.rec_n[pix] = over ? -1 : nc;
    if (over) atomicOr(a.flags, FLAG_DEPTH_RT);
}

__device__ __forceinline__ void polarize_prop_d(double eta_I, double eta_Q, double eps_I, double eps_Q, double ds,
                                                double &I, double &Q, double &U) {
    const double I0 = I, Q0 = Q;
    const double shQ = sinh(eta_Q * ds);
    const double chQ = cosh(eta_Q * ds);
    const double e_tau = exp(-(eta_I * ds));
    const double fac = 1.0 / (eta_I * eta_I - eta_Q * eta_Q);
    const double o11 = fac * eta_I * (1.0 - (chQ + (eta_Q / eta_I) * shQ) * e_tau);
    const double o12 = -(fac * eta_Q * (1.0 - (chQ + (eta_I / eta_Q) * shQ) * e_tau));
    I = o11 * eps_I + o12 * eps_Q + e_tau * (chQ * I0 - shQ * Q0);
    Q = o12 * eps_I + o11 * eps_Q + e_tau * (-(shQ * I0) + chQ * Q0);
    U = U * e_tau;
}

__device__ __forceinline__ double black_body_d(double T, double nu0) {
    if (T == 0.0) return 0.0;
    const double c1 = 1.4744995e-50, c2 = 4.79924e-11 / T;
    return c1 * pow(nu0, 3.0) * (1.0 / (exp(c2 * nu0) - 1.0));
}

__global__ void __launch_bounds__(256) stokes_kernel(ImgArgs a) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)a.n_rt * a.nlines * a.C;
    if (id >= total) return;
    const int k = (int)(id % a.C);
    const long long pl = id / a.C;
    const int j = (int)(pl % a.nlines);
    const int pix = (int)(pl / a.nlines);
    const int nc = a.rec_n[pix];
    const int *cells = a.rec_cell + (size_t)pix * a.n_depth_rt;
    const int *steps = a.rec_step + (size_t)pix * a.n_depth_rt;
    const double nu0 = a.nu0[j];
    const double cnu = (double)2.99792E8f / nu0;
    // B_nu(T_star) behind a pixel on the star's disc, else the CMB (raytrace.f90:98-114)
    const double I_in = a.start[pix].w != 0.0 ? a.Istar[j] : a.Iin[j];
    const double vel_obs = a.vel_out[k];
    const double spi = sqrt(acos(-1.0));
    const double kd = a.kd[j], alpha = a.alpha;
    double I_out = I_in, Q_in = 0.0, U_in = 0.0, Q_out = 0.0, U_out = 0.0, tau = 0.0, tau_dust_tot = 0.0;
    double I_dust = I_in, sig0 = 0.0, sig2 = 0.0, Qd = 0.0, Ud = 0.0;
    const double dxl = a.dl_scale;
    for (int i = 0; i < nc - 1; i++) {
        const int c = cells[i];
        const double b = a.b_all[c];
        const double dv = vel_obs - a.vz_r[c];
        const double v_corr = (1.0 / (b * spi)) * exp((-(dv * dv)) / (b * b));
        const CellGeo cg = a.geo[c];
        Q_out = Q_in * cg.c2a + U_in * cg.s2a;
        U_out = U_in * cg.c2a - Q_in * cg.s2a;
        const double4 pc = a.polco[(size_t)c * a.nlines + j];
        const double eta_I = cnu * v_corr * pc.x;
        const double eta_Q = cnu * v_corr * pc.y;
        const double eps_I = cnu * v_corr * pc.z;
        const double eps_Q = cnu * v_corr * pc.w;
        const double delta_x = dxl * (double)(steps[i + 1] - steps[i]);
        if (delta_x * eta_I > 1e-30) {
            polarize_prop_d(eta_I, eta_Q, eps_I, eps_Q, delta_x, I_out, Q_out, U_out);
            tau = tau + eta_I * delta_x;
        }
        const double tau_dust = kd * a.dust[c] * delta_x;
        if (tau_dust > 1e-30) {
            const double I_bb = black_body_d(a.tdust[c], nu0);
            const double e_dust = exp(-tau_dust);
            const double sI = 0.5 * alpha * (cg.mu2 - 2.0 / 3.0);
            const double sQ = alpha * cg.mu2;
            I_dust = I_dust * e_dust + I_bb * (1.0 - e_dust);
            sig0 = sig0 * e_dust + I_bb * (1.0 - e_dust);
            sig2 = sig2 * e_dust + sI * I_bb * (1.0 - e_dust);
            Qd = Qd * e_dust + sQ * I_bb * (1.0 - e_dust) * cg.c2p;
            Ud = Ud * e_dust + sQ * I_bb * (1.0 - e_dust) * cg.s2p;
            tau_dust_tot = tau_dust_tot + tau_dust;
        }
        Q_in = Q_out * cg.c2a - U_out * cg.s2a;
        U_in = U_out * cg.c2a + Q_out * cg.s2a;
    }
    const double I_corr = a.Iin[j];
    I_out = I_out - I_corr;
    Q_out = Q_in;
    U_out = U_in;
    const size_t o = (size_t)pix + (size_t)a.n_rt * ((size_t)j + (size_t)a.nlines * k);
    a.I[o] = I_out;
    a.Q[o] = Q_out;
    a.U[o] = U_out;
    a.TAU[o] = tau;
    if (a.dust_out) {
        I_dust = I_dust - I_corr;
        double Q_dust, U_dust, tdust;
        if (I_dust == 0.0) {
            Q_dust = 0.0;
            U_dust = 0.0;
            tdust = 0.0;
        } else {
            Q_dust = I_dust * Qd / (sig0 - sig2);
            U_dust = I_dust * Ud / (sig0 - sig2);
            tdust = tau_dust_tot;
        }
        const size_t cube = (size_t)a.n_rt * a.nlines * a.C;
        a.dust_out[o] = I_dust;
        a.dust_out[cube + o] = Q_dust;
        a.dust_out[2 * cube + o] = U_dust;
        a.dust_out[3 * cube + o] = tdust;
    }
}

// (1 - e^-x)/x, by its series below |x| = 0.1 (relative error < 1e-9 there, and the
// direct form loses at most 6e-7 to the rounding of e^-x above)
__device__ __forceinline__ float phi1_f(float x, float ex) {
    const float series =
        fmaf(x, fmaf(x, fmaf(x, fmaf(x, fmaf(x, -1.3888889e-3f, 8.3333333e-3f), -4.1666668e-2f), 0.16666667f), -0.5f),
             1.0f);
    return fabsf(x) < 0.1f ? series : __fdividef(1.0f - ex, x);
}

// (1 - e^-t (1 + t)) / t^2, by its series below t = 1 (terms to t^10), directly above
__device__ __forceinline__ float g_f(float t, float et) {
    float s = 2.2964459e-8f;  // 1/43545600
    s = fmaf(s, t, -2.5052108e-7f);
    s = fmaf(s, t, 2.4801587e-6f);
    s = fmaf(s, t, -2.2045855e-5f);
    s = fmaf(s, t, 1.7361111e-4f);
    s = fmaf(s, t, -1.1904762e-3f);
    s = fmaf(s, t, 6.9444444e-3f);
    s = fmaf(s, t, -3.3333335e-2f);
    s = fmaf(s, t, 0.125f);
    s = fmaf(s, t, -0.33333334f);
    s = fmaf(s, t, 0.5f);
    return t < 1.0f ? s : __fdividef(fmaf(-et, 1.0f + t, 1.0f), t * t);
}

// Mixed-precision Stokes transfer: the same cell walk, profile, rotations and skip
// test as stokes_kernel, evaluated in FP32.  Over a segment with t = eta_I ds,
// q = eta_Q ds, a = eps_I ds, b = eps_Q ds, the 2x2 propagator of polarize_prop
// (raytrace.f90:1-56) is, in exact arithmetic,
//   I' = I C - Q S + a A + b D,   Q' = Q C - I S + a D + b A,   U' = U e^-t,
//   C = e^-t cosh q,  S = e^-t sinh q,
//   A = (phi(t+q) + phi(t-q)) / 2,  D = (phi(t+q) - phi(t-q)) / 2,  phi(x) = (1 - e^-x)/x.
// Q is small beside I (a few 1e-4 of it in these clouds), so it is updated from these
// terms, each formed without cancellation, rather than as the difference of the
// transported I + Q and I - Q (which would leave Q an FP32 rounding of I, 1e-3 of Q
// itself).  With r = q/t (|r| < 1: eta_Q is a fraction of eta_I),
//   A = (phi(t) - e^-t (q r c + r^2 (1 + q^2 s))) / (1 - r^2),
//   D = q (-g(t) + r e^-t (q s + r c)) / (1 - r^2),
// where s = (sinh q / q - 1)/q^2, c = (cosh q - 1)/q^2 and g(t) = (1 - e^-t (1 + t))/t^2,
// each by its series where small arguments would cancel.  Near total polarisation
// (|r| >= 0.999) the transported I + Q and I - Q are used instead.  Intensities are
// carried in units of u = max(I_in, 2^-100), the pixel's background intensity, so that
// emissivities of weak lines stay in the FP32 normal range; the floor keeps the unit
// finite for lines whose background lies below FP32's range (above ~3 THz), and the
// coefficients are scaled in FP64 before they are rounded.  The optional dust branch
// stays FP64.
__global__ void __launch_bounds__(256) stokes_f32_kernel(ImgArgs a) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)a.n_rt * a.nlines * a.C;
    if (id >= total) return;
    const int k = (int)(id % a.C);
    const long long pl = id / a.C;
    const int j = (int)(pl % a.nlines);
    const int pix = (int)(pl / a.nlines);
    const int nc = a.rec_n[pix];
    const int *cells = a.rec_cell + (size_t)pix * a.n_depth_rt;
    const int *steps = a.rec_step + (size_t)pix * a.n_depth_rt;
    const double nu0 = a.nu0[j];
    const double I_cmb = a.Iin[j];
    const double I_in = a.start[pix].w != 0.0 ? a.Istar[j] : I_cmb;
    // eta scale c/nu (float literal as in raytrace.f90), eps additionally / u
    const double u = fmax(I_in, 0x1p-100);
    const double cnu = (double)2.99792E8f / nu0;
    const double cnu_e = cnu / u;
    const float vel_obs = (float)a.vel_out[k];
    const float rspi = 0.56418958f;  // 1/sqrt(pi)
    const double kd = a.kd[j], alpha = a.alpha;
    const float I0 = (float)(I_in / u);  // 1 unless the floor applies
    float I = I0, Q_in = 0.0f, U_in = 0.0f, Q_out = 0.0f, U_out = 0.0f, tau = 0.0f;
    double I_dust = I_in, sig0 = 0.0, sig2 = 0.0, Qd = 0.0, Ud = 0.0, tau_dust_tot = 0.0;
    const double dxl = a.dl_scale;
    for (int i = 0; i < nc - 1; i++) {
        const int c = cells[i];
        const float ib = (float)(1.0 / a.b_all[c]);
        const float x = (vel_obs - (float)a.vz_r[c]) * ib;
        const float v_corr = ib * rspi * __expf(-(x * x));
        const CellGeo cg = a.geo[c];
        const float c2a = (float)cg.c2a, s2a = (float)cg.s2a;
        Q_out = fmaf(Q_in, c2a, U_in * s2a);
        U_out = fmaf(U_in, c2a, -(Q_in * s2a));
        const double4 pc = a.polco[(size_t)c * a.nlines + j];
        const float eta_I = v_corr * (float)(cnu * pc.x);
        const float eta_Q = v_corr * (float)(cnu * pc.y);
        const float eps_I = v_corr * (float)(cnu_e * pc.z);
        const float eps_Q = v_corr * (float)(cnu_e * pc.w);
        const double delta_x = dxl * (double)(steps[i + 1] - steps[i]);
        const float ds = (float)delta_x;
        if (ds * eta_I > 1e-30f) {
            const float t = eta_I * ds, q = eta_Q * ds, sa = eps_I * ds, sb = eps_Q * ds;
            const float et = __expf(-t);
            const float r = __fdividef(q, t);
            if (fabsf(r) < 0.999f) {
                const float q2 = q * q;
                float sn, cn;
                if (fabsf(q) < 1.0f) {
                    sn = fmaf(q2, fmaf(q2, fmaf(q2, 2.7557319e-6f, 1.9841270e-4f), 8.3333333e-3f), 0.16666667f);
                    cn = fmaf(q2, fmaf(q2, fmaf(q2, 2.4801587e-5f, 1.3888889e-3f), 4.1666668e-2f), 0.5f);
                } else {
                    const float eq = __expf(q), emq = __expf(-q);
                    sn = __fdividef(__fdividef(eq - emq, 2.0f * q) - 1.0f, q2);
                    cn = __fdividef(fmaf(eq + emq, 0.5f, -1.0f), q2);
                }
                const float s1 = q2 * sn, c1 = q2 * cn;
                const float Cc = et * (1.0f + c1), Sc = et * q * (1.0f + s1);
                const float om = fmaf(-r, r, 1.0f);
                const float A = __fdividef(fmaf(-et, fmaf(q * r, cn, r * r * (1.0f + s1)), phi1_f(t, et)), om);
                const float D = __fdividef(q * fmaf(r * et, fmaf(q, sn, r * cn), -g_f(t, et)), om);
                const float In = fmaf(I, Cc, fmaf(-Q_out, Sc, fmaf(sa, A, sb * D)));
                Q_out = fmaf(Q_out, Cc, fmaf(-I, Sc, fmaf(sa, D, sb * A)));
                I = In;
            } else {
                // near total polarisation: the transported I + Q and I - Q
                const float m = t + q, p = t - q;
                const float em = __expf(-m), ep = __expf(-p);
                const float X = fmaf(I + Q_out, em, (sa + sb) * phi1_f(m, em));
                const float Y = fmaf(I - Q_out, ep, (sa - sb) * phi1_f(p, ep));
                I = 0.5f * (X + Y);
                Q_out = 0.5f * (X - Y);
            }
            U_out = U_out * et;
            tau = tau + t;
        }
        const double tau_dust = kd * a.dust[c] * delta_x;
        if (tau_dust > 1e-30) {
            const double I_bb = black_body_d(a.tdust[c], nu0);
            const double e_dust = exp(-tau_dust);
            const double sI = 0.5 * alpha * (cg.mu2 - 2.0 / 3.0);
            const double sQ = alpha * cg.mu2;
            I_dust = I_dust * e_dust + I_bb * (1.0 - e_dust);
            sig0 = sig0 * e_dust + I_bb * (1.0 - e_dust);
            sig2 = sig2 * e_dust + sI * I_bb * (1.0 - e_dust);
            Qd = Qd * e_dust + sQ * I_bb * (1.0 - e_dust) * cg.c2p;
            Ud = Ud * e_dust + sQ * I_bb * (1.0 - e_dust) * cg.s2p;
            tau_dust_tot = tau_dust_tot + tau_dust;
        }
        Q_in = fmaf(Q_out, c2a, -(U_out * s2a));
        U_in = fmaf(U_out, c2a, Q_out * s2a);
    }
    const size_t o = (size_t)pix + (size_t)a.n_rt * ((size_t)j + (size_t)a.nlines * k);
    // I_out - B_nu(2.73 K) with I_out = I u; the second term is 0 off the star
    a.I[o] = (double)(I - I0) * u + (I_in - I_cmb);
    a.Q[o] = (double)Q_in * u;
    a.U[o] = (double)U_in * u;
    a.TAU[o] = (double)tau;
    if (a.dust_out) {
        I_dust = I_dust - I_cmb;
        double Q_dust, U_dust, tdust;
        if (I_dust == 0.0) {
            Q_dust = 0.0;
            U_dust = 0.0;
            tdust = 0.0;
        } else {
            Q_dust = I_dust * Qd / (sig0 - sig2);
            U_dust = I_dust * Ud / (sig0 - sig2);
            tdust = tau_dust_tot;
        }
        const size_t cube = (size_t)a.n_rt * a.nlines * a.C;
        a.dust_out[o] = I_dust;
        a.dust_out[cube + o] = Q_dust;
        a.dust_out[2 * cube + o] = U_dust;
        a.dust_out[3 * cube + o] = tdust;
    }
}

hipError_t launch_image(const ImgArgs &a, hipStream_t s, float *ms_geom, float *ms_rt) {
    hipEvent_t e0, e1, e2;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventCreate(&e2);
    hipEventRecord(e0, s);
    hipLaunchKernelGGL(rotate_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a);
    const long long ne = (long long)a.N * a.nlines;
    hipLaunchKernelGGL(emcoef_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(start_kernel, dim3((a.n_rt + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(march_kernel, dim3((a.n_rt + 255) / 256), dim3(256), 0, s, a);
    hipEventRecord(e1, s);
    const long long nt = (long long)a.n_rt * a.nlines * a.C;
    if (a.mixed)
        hipLaunchKernelGGL(stokes_f32_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(stokes_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, a);
    hipEventRecord(e2, s);
    hipError_t err = hipEventSynchronize(e2);
    if (ms_geom) hipEventElapsedTime(ms_geom, e0, e1);
    if (ms_rt) hipEventElapsedTime(ms_rt, e1, e2);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    return err == hipSuccess ? hipGetLastError() : err;
}

}  // namespace portal
