// Lambda-iteration closure (a build extension; the reference makes one SE pass).
//
// After an SE pass the level densities the next inward pass sees are
//   lev_p(i, node) <- ab * n_H2 * sqrt(2 J_i + 1) * rho^0_0(i)
// which follows from the SE normalisation row sum_i sqrt(2J_i+1) rho^0_0(i) = 1
// (em_coefs.f90:158-164) and eat_lime.f90:261-263.  The kernel also reduces the
// largest relative population change and the largest change of rho^2_0/rho^0_0,
// the convergence measures of the iteration.  Levels that hold less than 1e-10 of a
// node's population (sqrt(2J+1) rho^0_0 < 1e-10, the SE normalisation summing these to
// 1) are left out of both measures: their relative change is rounding noise in the
// lines' J and their lines carry nothing at the images' tolerance.
#include "common.h"

namespace portal {

struct LambdaArgs {
    int n_work, N, L;
    const int *work;
    const double *ab, *h2;
    const int *j_lev;
    const double *pol;       // N x 2L, this pass
    const double *pol_prev;  // N x 2L, previous pass (or null)
    double *levp;            // L x N, updated in place
    unsigned long long *maxes;  // [0] pop change, [1] alignment change (double bits)
    double *pol_save;        // when set, receives this pass's pol (the next pass's pol_prev)
    unsigned long long *node_res;  // when set, per work item the largest relative population change
                                   // over its levels (double bits, ORed in with atomicMax)
};

__global__ void __launch_bounds__(256) lambda_update_kernel(LambdaArgs a) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dpop = 0.0, dal = 0.0;
    if (id < (long long)a.n_work * a.L) {
        const int w = (int)(id / a.L), i = (int)(id % a.L);
        const int node = a.work[w];
        const size_t N = a.N;
        const double r0 = a.pol[node + N * (2 * i)];
        const double r2 = a.pol[node + N * (2 * i + 1)];
        const double nv = a.ab[node] * a.h2[node] * sqrt(2.0 * a.j_lev[i] + 1.0) * r0;
        const double old = a.levp[(size_t)node * a.L + i];
        if (old > 0.0) dpop = fabs(nv - old) / old;
        else if (nv != 0.0) dpop = 1.0;
        if (a.pol_prev) {
            const double p0 = a.pol_prev[node + N * (2 * i)], p2 = a.pol_prev[node + N * (2 * i + 1)];
            const double q = r0 != 0.0 ? r2 / r0 : 0.0, qp = p0 != 0.0 ? p2 / p0 : 0.0;
            dal = fabs(q - qp);
        } else {
            dal = 1.0;
        }
        if (!(dpop == dpop)) dpop = 1.0;  // NaN counts as not converged
        if (!(dal == dal)) dal = 1.0;
        if (sqrt(2.0 * a.j_lev[i] + 1.0) * r0 < 1e-10) dpop = dal = 0.0;  // false for NaN
        a.levp[(size_t)node * a.L + i] = nv;
        if (a.node_res) atomicMax(a.node_res + w, (unsigned long long)__double_as_longlong(dpop));
        if (a.pol_save) {
            a.pol_save[node + N * (2 * i)] = r0;
            a.pol_save[node + N * (2 * i + 1)] = r2;
        }
    }
    // block max then one atomic per block (non-negative doubles order like their bits)
    for (int o = 32; o > 0; o >>= 1) {
        dpop = fmax(dpop, __shfl_xor(dpop, o, 64));
        dal = fmax(dal, __shfl_xor(dal, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(a.maxes, (unsigned long long)__double_as_longlong(dpop));
        atomicMax(a.maxes + 1, (unsigned long long)__double_as_longlong(dal));
    }
}

hipError_t launch_lambda_update(int n_work, int N, int L, const int *work, const double *ab, const double *h2,
                                const int *j_lev, const double *pol, const double *pol_prev, double *levp,
                                unsigned long long *maxes, hipStream_t s, double *pol_save,
                                unsigned long long *node_res) {
    LambdaArgs a{n_work, N, L, work, ab, h2, j_lev, pol, pol_prev, levp, maxes, pol_save, node_res};
    const long long n = (long long)n_work * L;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(lambda_update_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Ng acceleration of the iteration (Ng 1974, in the form of Auer 1987, "Acceleration
// of convergence", Numerical Radiative Transfer, 101-109): from the last four level
// density iterates x0 (newest) .. x3,
//   d0 = x0 - x1, d1 = d0 - (x1 - x2), d2 = d0 - (x2 - x3),
//   [d1.d1 d1.d2; d1.d2 d2.d2] (a, b) = (d0.d1, d0.d2),
//   x* = (1 - a - b) x0 + a x1 + b x2,
// with the dot products weighted by 1/x0^2 (relative changes).  Only solved nodes
// enter; an entry whose extrapolation is not positive keeps x0.
struct NgArgs {
    int n_work, L;
    const int *work;
    const double *x0, *x1, *x2, *x3;
    double *partial;  // NG_BLOCKS x 5
};

constexpr int NG_BLOCKS = 512;

__global__ void __launch_bounds__(256) ng_dots_kernel(NgArgs a) {
    const long long n = (long long)a.n_work * a.L;
    double s[5] = {0, 0, 0, 0, 0};
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int w = (int)(e / a.L), i = (int)(e % a.L);
        const size_t k = (size_t)a.work[w] * a.L + i;
        const double x0 = a.x0[k], x1 = a.x1[k], x2 = a.x2[k], x3 = a.x3[k];
        const double wt = x0 > 0.0 ? 1.0 / (x0 * x0) : 0.0;
        const double d0 = x0 - x1, d1 = d0 - (x1 - x2), d2 = d0 - (x2 - x3);
        s[0] += wt * d1 * d1;
        s[1] += wt * d1 * d2;
        s[2] += wt * d2 * d2;
        s[3] += wt * d0 * d1;
        s[4] += wt * d0 * d2;
    }
    __shared__ double red[5][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = 0; j < 5; j++) {
        double v = s[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[j][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int j = threadIdx.x;
        a.partial[blockIdx.x * 5 + j] = ((red[j][0] + red[j][1]) + red[j][2]) + red[j][3];
    }
}

// fixed-order sum of the block partials: out[5]
__global__ void __launch_bounds__(64) ng_sum_kernel(const double *partial, int nb, double *out) {
    const int j = threadIdx.x;
    if (j >= 5) return;
    double v = 0.0;
    for (int b = 0; b < nb; b++) v += partial[b * 5 + j];
    out[j] = v;
}

__global__ void __launch_bounds__(256) ng_apply_kernel(int n_work, int L, const int *work, double *x0, const double *x1,
                                                       const double *x2, double ca, double cb) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n_work * L) return;
    const int w = (int)(e / L), i = (int)(e % L);
    const size_t k = (size_t)work[w] * L + i;
    const double v = (1.0 - ca - cb) * x0[k] + ca * x1[k] + cb * x2[k];
    if (v > 0.0 && v < 1e300) x0[k] = v;
}

hipError_t launch_ng_dots(int n_work, int L, const int *work, const double *x0, const double *x1, const double *x2,
                          const double *x3, double *partial, double *out5, hipStream_t s) {
    NgArgs a{n_work, L, work, x0, x1, x2, x3, partial};
    hipLaunchKernelGGL(ng_dots_kernel, dim3(NG_BLOCKS), dim3(256), 0, s, a);
    hipLaunchKernelGGL(ng_sum_kernel, dim3(1), dim3(64), 0, s, (const double *)partial, NG_BLOCKS, out5);
    return hipGetLastError();
}

hipError_t launch_ng_apply(int n_work, int L, const int *work, double *x0, const double *x1, const double *x2,
                           double ca, double cb, hipStream_t s) {
    const long long n = (long long)n_work * L;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(ng_apply_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_work, L, work, x0, x1,
                       x2, ca, cb);
    return hipGetLastError();
}

size_t ng_partial_doubles() { return (size_t)NG_BLOCKS * 5 + 8; }

// Anderson acceleration (Anderson 1965; the "type II" form of Walker & Ni 2011,
// SIAM J. Numer. Anal. 49, 1715) of the map G: x (level densities a pass starts
// from) -> G(x) (level densities its SE solution closes to).  With f = G(x) - x and
// the window of the AA_M previous pairs, differences to the newest pair
//   dF_j = f_k - f_{k-j},  dG_j = G_k - G_{k-j},
// gamma minimises |f_k - sum_j gamma_j dF_j| in the 1/G_k^2-weighted norm, and
//   x_{k+1} = G_k - sum_j gamma_j dG_j.
// The pairs are kept compact (solved nodes x levels, in work-list order).

__global__ void __launch_bounds__(256) aa_gather_kernel(int n_work, int L, const int *work, const double *levp,
                                                        double *dst) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n_work * L) return;
    const int w = (int)(e / L), i = (int)(e % L);
    dst[e] = levp[(size_t)work[w] * L + i];
}

struct AaDotArgs {
    long long n;
    int mk;
    const double *x0, *g0;
    const double *xj[AA_M_MAX], *gj[AA_M_MAX];
    double *partial;  // NG_BLOCKS x AA_NSUM
};

__global__ void __launch_bounds__(256) aa_dots_kernel(AaDotArgs a) {
    double s[AA_NSUM];
#pragma unroll
    for (int t = 0; t < AA_NSUM; t++) s[t] = 0.0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < a.n; e += (long long)gridDim.x * blockDim.x) {
        const double g0 = a.g0[e], f0 = g0 - a.x0[e];
        const double wt = g0 > 0.0 ? 1.0 / (g0 * g0) : 0.0;
        double df[AA_M_MAX];
#pragma unroll
        for (int j = 0; j < AA_M_MAX; j++) df[j] = j < a.mk ? f0 - (a.gj[j][e] - a.xj[j][e]) : 0.0;
        int t = 0;
#pragma unroll
        for (int i = 0; i < AA_M_MAX; i++) {
#pragma unroll
            for (int j = i; j < AA_M_MAX; j++) s[t++] += wt * df[i] * df[j];
            s[t++] += wt * df[i] * f0;
        }
    }
    __shared__ double red[AA_NSUM][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < AA_NSUM; j++) {
        double v = s[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[j][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < AA_NSUM) {
        const int j = threadIdx.x;
        a.partial[blockIdx.x * AA_NSUM + j] = ((red[j][0] + red[j][1]) + red[j][2]) + red[j][3];
    }
}

__global__ void __launch_bounds__(64) aa_sum_kernel(const double *partial, int nb, double *out) {
    const int j = threadIdx.x;
    if (j >= AA_NSUM) return;
    double v = 0.0;
    for (int b = 0; b < nb; b++) v += partial[b * AA_NSUM + j];
    out[j] = v;
}

struct AaApplyArgs {
    int n_work, L, mk;
    const int *work;
    const double *g0;
    const double *gj[AA_M_MAX];
    double gamma[AA_M_MAX];
    double *levp;
};

__global__ void __launch_bounds__(256) aa_apply_kernel(AaApplyArgs a) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)a.n_work * a.L) return;
    const double g0 = a.g0[e];
    double v = g0;
#pragma unroll
    for (int j = 0; j < AA_M_MAX; j++)
        if (j < a.mk) v -= a.gamma[j] * (g0 - a.gj[j][e]);
    const int w = (int)(e / a.L), i = (int)(e % a.L);
    a.levp[(size_t)a.work[w] * a.L + i] = (v > 0.0 && v < 1e300) ? v : g0;
}

hipError_t launch_aa_gather(int n_work, int L, const int *work, const double *levp, double *dst, hipStream_t s) {
    const long long n = (long long)n_work * L;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(aa_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_work, L, work, levp,
                       dst);
    return hipGetLastError();
}

hipError_t launch_aa_dots(long long n, int mk, const double *x0, const double *g0, const double *const *xj,
                          const double *const *gj, double *partial, double *out, hipStream_t s) {
    AaDotArgs a{};
    a.n = n;
    a.mk = mk;
    a.x0 = x0;
    a.g0 = g0;
    for (int j = 0; j < AA_M_MAX; j++) {
        a.xj[j] = j < mk ? xj[j] : x0;
        a.gj[j] = j < mk ? gj[j] : g0;
    }
    a.partial = partial;
    hipLaunchKernelGGL(aa_dots_kernel, dim3(NG_BLOCKS), dim3(256), 0, s, a);
    hipLaunchKernelGGL(aa_sum_kernel, dim3(1), dim3(64), 0, s, (const double *)partial, NG_BLOCKS, out);
    return hipGetLastError();
}

hipError_t launch_aa_apply(int n_work, int L, int mk, const int *work, const double *g0, const double *const *gj,
                           const double *gamma, double *levp, hipStream_t s) {
    AaApplyArgs a{};
    a.n_work = n_work;
    a.L = L;
    a.mk = mk;
    a.work = work;
    a.g0 = g0;
    for (int j = 0; j < AA_M_MAX; j++) {
        a.gj[j] = j < mk ? gj[j] : g0;
        a.gamma[j] = j < mk ? gamma[j] : 0.0;
    }
    a.levp = levp;
    const long long n = (long long)n_work * L;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(aa_apply_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

size_t aa_partial_doubles() { return (size_t)NG_BLOCKS * AA_NSUM + 32; }

}  // namespace portal

namespace portal {

// Surrogate input features of every work node, from the last inward pass and the SE
// solution rho_1 computed from it:
//   log10 n_H2[cm^-3], log10 T_kin; per line log10(J^0_0 / I_cmb), 100 J^2_0/J^0_0;
//   per level log10(sqrt(2J+1) rho^0_0); per level with J > 0 1e3 rho^2_0/rho^0_0.
// The outputs are corrections to rho_1: per level log10(rho^0_0 / rho_1^0_0), per level
// with J > 0 1e3 (rho^2_0/rho^0_0 - rho_1^2_0/rho_1^0_0).
__device__ __forceinline__ double surr_logpop(double sg_r0) { return log10(fmax(sg_r0, 1e-12)); }

__global__ void __launch_bounds__(256) surr_features_kernel(SurrArgs a) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n_work) return;
    const int node = a.work[w];
    const size_t N = a.N;
    float *f = a.feats + (size_t)w * a.n_in;
    f[0] = (float)log10(a.h2[node] * 1e-6);
    f[1] = (float)log10(a.tkin[node]);
    for (int k = 0; k < a.R; k++) {
        const double j0 = a.J0[(size_t)node * a.R + k], j2 = a.J2[(size_t)node * a.R + k];
        f[2 + 2 * k] = (float)log10(j0 / a.Iin[k]);
        f[3 + 2 * k] = (float)(100.0 * j2 / j0);
    }
    int q = 2 + 2 * a.R;
    for (int i = 0; i < a.L; i++) {
        const double r0 = a.pol[node + N * (2 * i)];
        f[q++] = (float)surr_logpop(sqrt(2.0 * a.j_lev[i] + 1.0) * r0);
    }
    for (int i = 0; i < a.L; i++) {
        if (a.j_lev[i] == 0) continue;
        const double r0 = a.pol[node + N * (2 * i)], r2 = a.pol[node + N * (2 * i + 1)];
        f[q++] = (float)(r0 != 0.0 ? 1e3 * r2 / r0 : 0.0);
    }
}

// Decode the network output into populations and alignment; reject non-finite or
// badly normalised predictions, and ones that invert a line that rho_1 does not invert
// (those nodes keep the plain iterate).  A level whose sqrt(2J+1) rho_1^0_0 is at or
// below the features' floor (1e-12) keeps rho_1: its features and training targets were
// floored, so the network's output says nothing about it.
__device__ __forceinline__ double surr_decode(double x1, float o) {
    return x1 > 1e-12 ? exp10(surr_logpop(x1) + (double)o) : x1;
}

__global__ void __launch_bounds__(256) surr_apply_kernel(SurrArgs a) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n_work) return;
    const int node = a.work[w];
    const float *o = a.out + (size_t)w * a.n_out;
    const size_t N = a.N;
    // sqrt(2J+1) rho^0_0 of level i (proportional to its population): first solution, prediction
    auto first = [&](int i) { return sqrt(2.0 * a.j_lev[i] + 1.0) * a.pol[node + N * (2 * i)]; };
    auto pred = [&](int i) { return surr_decode(first(i), o[i]); };
    double s = 0.0;
    bool ok = true;
    for (int i = 0; i < a.L; i++) {
        const double p = pred(i);
        ok = ok && isfinite(p);
        s += p;
    }
    ok = ok && s > 0.5 && s < 2.0;
    int q = a.L;
    for (int i = 0; i < a.L && ok; i++) {
        if (a.j_lev[i] == 0) continue;
        const double r0 = a.pol[node + N * (2 * i)], r2 = a.pol[node + N * (2 * i + 1)];
        const double rat = (r0 != 0.0 ? r2 / r0 : 0.0) + 1e-3 * (double)o[q++];
        ok = ok && isfinite(rat) && fabs(rat) < 0.5;
    }
    // line k between levels l1 (lower) and l2 (upper) is inverted where n1 g2 < g1 n2
    for (int k = 0; k < a.R && ok; k++) {
        const int lo = a.l1[k], up = a.l2[k];
        const bool inv_pred = pred(lo) * a.g2[k] < a.g1[k] * pred(up);
        const bool inv_first = first(lo) * a.g2[k] < a.g1[k] * first(up);
        ok = !inv_pred || inv_first;
    }
    if (a.flags) a.flags[w] = ok ? 0 : 1;
    if (!ok) {
        atomicAdd(a.n_fallback, 1);
        return;
    }
    // test hook: every perturb_stride-th item gets level 1's population scaled by
    // perturb_factor, renormalised (a wrong prediction that passes the checks above)
    const bool pert = a.perturb_stride > 0 && w % a.perturb_stride == 0;
    double p0 = 0.0;
    if (pert) {
        p0 = pred(0);
        s += (a.perturb_factor - 1.0) * p0;
    }
    const double nab = a.ab[node] * a.h2[node];
    q = a.L;
    for (int i = 0; i < a.L; i++) {
        const double sg = sqrt(2.0 * a.j_lev[i] + 1.0);
        const double r0_1 = a.pol[node + N * (2 * i)], r2_1 = a.pol[node + N * (2 * i + 1)];
        const double p = surr_decode(sg * r0_1, o[i]) * (pert && i == 0 ? a.perturb_factor : 1.0) / s;
        const double r0 = p / sg;
        const double rat = a.j_lev[i] != 0 ? (r0_1 != 0.0 ? r2_1 / r0_1 : 0.0) + 1e-3 * (double)o[q++] : 0.0;
        a.levp[(size_t)node * a.L + i] = nab * p;
        a.pol[node + N * (2 * i)] = r0;
        a.pol[node + N * (2 * i + 1)] = rat * r0;
    }
}

// The residual test of the surrogate iterate, one lambda iteration after it was applied:
// an accepted item whose largest relative level-population change in that iteration
// (res_s) exceeds both tol and the change its plain first iterate made (res_p) reverts
// to that plain iterate (plain_lev / plain_pol); flags[w] = 2 marks it.  A NaN change
// counts as exceeding.
__global__ void __launch_bounds__(256) surr_revert_kernel(SurrRevertArgs a) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n_work || a.flags[w] != 0) return;
    const double rs = __longlong_as_double((long long)a.res_s[w]);
    const double rp = __longlong_as_double((long long)a.res_p[w]);
    if (rs <= fmax(a.tol, rp)) return;  // false for NaN
    a.flags[w] = 2;
    atomicAdd(a.n_fallback, 1);
    const int node = a.work[w];
    const size_t N = a.N;
    for (int i = 0; i < a.L; i++) {
        a.levp[(size_t)node * a.L + i] = a.plain_lev[(size_t)node * a.L + i];
        a.pol[node + N * (2 * i)] = a.plain_pol[node + N * (2 * i)];
        a.pol[node + N * (2 * i + 1)] = a.plain_pol[node + N * (2 * i + 1)];
    }
}

hipError_t launch_surr_revert(const SurrRevertArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    hipLaunchKernelGGL(surr_revert_kernel, dim3((a.n_work + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_surr_features(const SurrArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    hipLaunchKernelGGL(surr_features_kernel, dim3((a.n_work + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_surr_apply(const SurrArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    hipLaunchKernelGGL(surr_apply_kernel, dim3((a.n_work + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace portal

namespace portal {

// zero fill used in place of hipMemsetAsync inside captured graphs
__global__ void fill_zero_kernel(unsigned int *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 4;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (unsigned int *)p, n);
    return hipGetLastError();
}

}  // namespace portal
