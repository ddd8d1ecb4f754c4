// Inward ray integration over stored paths, shared by the FP64 (rt_f64.hip) and
// mixed-precision (rt_mixed.hip) builds.  Restates ray_in (inverse_rt.f90:217-344)
// and the J accumulation of inverse_rt (inverse_rt.f90:126-129).
//
// One workgroup per node; wave w takes rays w, w+W, ...  Per ray the wave first
// stages the path into its own LDS slice: per cell the projected velocity, 1/b and
// the segment length, and per (cell, line) the line coefficients K (tau per unit
// profile and length) and S (source function).  Lanes then own (line, channel)
// pairs p = p0 + lane + 64 s and integrate them front to back,
//   I_out = sum_i S_i (1 - e^-tau_i) prod_{j<i} e^-tau_j + I_in prod_j e^-tau_j,
// the reference's back-to-front recurrence summed in the other order, with its skip
// of segments whose tau is not > 1e-30 (NaN included) and no skip for the node's own
// cell.  GROUP slots are integrated together so that their exponential chains
// overlap; the cell loop is wave-uniform (cell data are LDS broadcasts) and stops
// when every lane's transmission is exactly zero.  Per pair the lane accumulates
// sum_rays w I and sum_rays w c_in I; the channel sums with the profile weight
// phi0 d_v and the sum over waves are done once per node in a fixed order.
#pragma once
#include "common.h"

namespace portal {

__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous
    // range of (Morton-ordered) nodes so that its L2 holds one region of the cloud.
    const int per = (nblocks + 7) >> 3;
    const int xcd = bid & 7, slot = bid >> 3;
    const int full = nblocks - (per - 1) * 8;
    return xcd < full ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
}

template <typename Real>
struct Arith;

// Staged quantities and the segment update.  Per cell the staging stores vb = s v.dir/b
// and inv_b = s/b, per (cell, line) L and S, so that a segment's profile argument is
// u = fma(v0 + v_node, inv_b, -vb) and its optical depth one exponential of L - u^2.
//   FP64:  s = 1, L = log(K dl), tau = exp(L - u^2); a segment is skipped unless
//          tau > 1e-30 (false for NaN), the reference's test.
//   mixed: s = sqrt(log2 e), L = log2(K dl log2 e), so exp2(L - u^2) = tau log2 e
//          ("tau2") and exp2(-tau2) = e^-tau with no further scaling.  A (cell, line)
//          with K dl not > 0 (zero, negative or NaN: skipped by the reference) is
//          staged as L = -inf, S = 0, which makes its update an exact no-op, so the
//          mixed update needs no skip test; a finite tau below 1e-30 changes nothing
//          at FP32 resolution.
template <>
struct Arith<double> {
    typedef double2 R2;
    static constexpr bool kahan = false;
    static constexpr bool skip_test = true;
    static constexpr double vscale = 1.0;
    static constexpr double tr_floor = 0.0;  // FP64: a channel stops only at exactly zero transmission
    __device__ static double ex(double x) { return exp(x); }
    __device__ static double one_minus(double tau, double et) { return 1.0 - et; }
    __device__ static R2 stage(R2 k, double dl, bool) { return double2{log(k.x * dl), k.y}; }
    __device__ static bool cell_ok(double, double) { return true; }
    __device__ static double tau_of(double L, double u) { return exp(L - u * u); }
    __device__ static double att(double tau) { return exp(-tau); }
    __device__ static double om(double tau, double et) { return 1.0 - et; }
    __device__ static double thr() { return 1e-30; }
    // exponent of the intensity unit: FP64 carries intensities in W m^-2 Hz^-1 sr^-1
    __device__ static int uexp(double) { return 0; }
    __device__ static R2 ks(const RtArgs &a, size_t i) { return *reinterpret_cast<const double2 *>(a.KS + 2 * i); }
    __device__ static R2 ksd(const RtArgs &a, size_t i) { return *reinterpret_cast<const double2 *>(a.KSD + 2 * i); }
    // dust of a segment: (alpha dl, B_nu(T_dust)) in the units of tau_of
    __device__ static R2 stage_dust(R2 d, double dl, bool) { return double2{d.x * dl, d.y}; }
    __device__ static double rcp(double x) { return 1.0 / x; }
};

template <>
struct Arith<float> {
    typedef float2 R2;
    static constexpr bool kahan = true;
    static constexpr bool skip_test = false;
    static constexpr float vscale = 1.2011224087864498f;  // sqrt(log2 e)
    // mixed: a channel whose transmission fell below 1e-9 stops (the dropped tail is
    // below 1e-9 of the remaining source function)
    static constexpr float tr_floor = 1e-9f;
    __device__ static float ex(float x) { return __expf(x); }
    // 1 - exp(-tau) without cancellation for small tau: the smaller of two upper bounds,
    // the cubic Taylor sum (above 1 - e^-tau for every tau > 0, exact to tau^4/24) and
    // 1 + 2^-22 - e^-tau (the offset exceeds the error of et = __expf(-tau), at most
    // 1.5e-7 absolute); relative error below 5e-6 at any tau, with no branch or select
    __device__ static float one_minus(float tau, float et) {
        const float cub = tau * fmaf(-tau, fmaf(-tau, 0.16666667f, 0.5f), 1.0f);
        return __builtin_fminf(cub, 1.00000024f - et);
    }
    __device__ static R2 stage(R2 k, float dl, bool cell) {
        const float kdl = k.x * dl;
        const bool pos = cell && kdl > 0.0f;
        return float2{pos ? __log2f(kdl * 1.44269504f) : -INFINITY, pos ? k.y : 0.0f};
    }
    // a cell whose profile argument is not finite (b = 0 gives 1/b = inf and, with
    // v.dir = 0, a NaN product; the reference's tau is then NaN and skipped) is staged
    // with zero profile terms and L = -inf, an exact no-op
    __device__ static bool cell_ok(float vb, float inv_b) { return __builtin_isfinite(vb) && __builtin_isfinite(inv_b); }
    // v_exp_f32 directly: a result below the FP32 normal range is below 1e-30
    __device__ static float tau_of(float L, float u) { return __builtin_amdgcn_exp2f(fmaf(-u, u, L)); }
    __device__ static float att(float tau2) { return __builtin_amdgcn_exp2f(-tau2); }
    // 1 - e^-tau from tau2 = tau log2 e: tau (1 - tau/2 + tau^2/6) below tau = 1e-2
    // (relative error < 5e-8), 1 - e^-tau above
    __device__ static float om(float tau2, float et) {
        const float series = tau2 * (0.693147182f - tau2 * (0.240226507f - tau2 * 0.0555041087f));
        const float direct = 1.0f - et;
        const int m = -(int)(tau2 < 1.44269504e-2f);
        return __int_as_float((__float_as_int(series) & m) | (__float_as_int(direct) & ~m));
    }
    __device__ static float thr() { return 1.44269504e-30f; }  // 1e-30 in units of tau2
    // FP32 carries a line's intensities in units of 2^line_iexp(I_in) (S is staged so)
    __device__ static int uexp(double iin) { return line_iexp(iin); }
    __device__ static R2 ks(const RtArgs &a, size_t i) { return *reinterpret_cast<const float2 *>(a.KS32 + 2 * i); }
    __device__ static R2 ksd(const RtArgs &a, size_t i) { return *reinterpret_cast<const float2 *>(a.KSD32 + 2 * i); }
    __device__ static R2 stage_dust(R2 d, float dl, bool cell) {
        const bool on = cell && d.x * dl > 0.0f;
        return float2{on ? d.x * dl * 1.44269504f : 0.0f, on ? d.y : 0.0f};
    }
    __device__ static float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};

template <typename Real>
struct __attribute__((aligned(8))) CellRec {
    Real vb, inv_b, dl, pad;  // v.dir / b, 1/b, segment length
};

template <typename Real>
__host__ __device__ constexpr size_t stage_bytes_per_cell(int R) {
    return sizeof(CellRec<Real>) + (size_t)R * sizeof(typename Arith<Real>::R2);
}

template <typename Real, int MAXS, int GROUP>
__device__ __forceinline__ void rt_staged_node(const RtArgs &a, const int bid, const int cap, char *smem) {
    typedef Arith<Real> A;
    typedef typename A::R2 R2;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = blockDim.x >> 6;
    const int node = a.work[bid];
    const double bd = a.velb[node].w;
    const Real b = (Real)bd;
    const Real inv_bn = (Real)(1.0 / bd);
    const int C = a.C, nvel = a.nvel, R = a.R;
    const int npairs = R * C;
    const Real dls = (Real)a.dl_scale;
    const size_t wave_bytes = (size_t)cap * stage_bytes_per_cell<Real>(R);
    CellRec<Real> *cr = reinterpret_cast<CellRec<Real> *>(smem + wave * wave_bytes);
    R2 *ks = reinterpret_cast<R2 *>(cr + cap);
    // after the ray loop the staging area is reused for the reduction
    double *red0 = reinterpret_cast<double *>(smem);
    double *red2 = red0 + 64 * MAXS;
    double *jacc = red2 + 64 * MAXS;

    double jl0 = 0.0, jl2 = 0.0;  // thread t < R accumulates line t
    for (int p0 = 0; p0 < npairs; p0 += 64 * MAXS) {
        Real S0[MAXS], C0[MAXS], S2[MAXS], C2[MAXS];
#pragma unroll
        for (int s = 0; s < MAXS; s++) S0[s] = C0[s] = S2[s] = C2[s] = (Real)0;
        for (int r = wave; r < a.D; r += W) {
            const Real mu = (Real)a.xth[3 * r], w = (Real)a.xth[3 * r + 2];
            const Real wc = w * (((Real)3 * (mu * mu) - (Real)1) / (Real)2);
            const long long off = a.offsets[(long long)bid * a.D + r];
            const int nc = (int)(a.offsets[(long long)bid * a.D + r + 1] - off);
            const PathEntry *P = a.paths + off;
            for (int i = lane; i < nc; i += 64) {
                const PathEntry e = P[i];
                CellRec<Real> c;
                const double ib = a.inv_b[e.cell];
                c.vb = (Real)(e.vray * ib) * (Real)A::vscale;
                c.inv_b = (Real)ib * (Real)A::vscale;
                const bool ok = A::cell_ok(c.vb, c.inv_b);
                if (!ok) c.vb = c.inv_b = (Real)0;
                c.dl = (Real)e.len * dls;
                c.pad = (Real)e.vray;
                cr[i] = c;
                for (int l = 0; l < R; l++) ks[i * R + l] = A::stage(A::ks(a, (size_t)e.cell * R + l), c.dl, ok);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const Real vel_node = cr[0].pad;
            const Real dl0 = cr[0].dl;
#pragma unroll
            for (int g0 = 0; g0 < MAXS; g0 += GROUP) {
                if (p0 + 64 * g0 >= npairs) break;  // wave-uniform
                Real Tr[GROUP], Em[GROUP], v0s[GROUP];
                int lg[GROUP];
                bool on[GROUP];
#pragma unroll
                for (int q = 0; q < GROUP; q++) {
                    const int p = p0 + lane + 64 * (g0 + q);
                    on[q] = p < npairs;
                    const int l = on[q] ? p / C : 0;
                    const int nv = on[q] ? p - l * C - nvel : 0;
                    lg[q] = l;
                    const Real v0 = ((Real)2 * b) * (Real)nv / (Real)nvel;
                    const Real aa = v0 * inv_bn;
                    const Real e0 = A::ex(-(aa * aa));
                    const R2 k0 = A::ks(a, (size_t)node * R + l);  // the node's own cell, no skip
                    const Real tau0 = k0.x * e0 * dl0;
                    Tr[q] = A::ex(-tau0);
                    Em[q] = k0.y * A::one_minus(tau0, Tr[q]);
                    v0s[q] = v0 + vel_node;
                    if (!on[q]) Tr[q] = (Real)0;
                }
                for (int i = 1; i < nc; i++) {
                    bool live = false;
#pragma unroll
                    for (int q = 0; q < GROUP; q++) {
                        if (!(Tr[q] > (Real)A::tr_floor)) Tr[q] = (Real)0;
                        live = live || (Tr[q] != (Real)0);
                    }
                    if (!__any(live)) break;
                    const CellRec<Real> c = cr[i];
#pragma unroll
                    for (int q = 0; q < GROUP; q++) {
                        const R2 k = ks[i * R + lg[q]];
                        const Real u = fma(v0s[q], c.inv_b, -c.vb);
                        const Real tau = A::tau_of(k.x, u);
                        if (tau > A::thr() && Tr[q] != (Real)0) {
                            const Real et = A::att(tau);
                            Em[q] = Em[q] + Tr[q] * k.y * A::om(tau, et);
                            Tr[q] = Tr[q] * et;
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < GROUP; q++) {
                    if (!on[q]) continue;
                    const double iin = a.I_in[lg[q]];
                    const Real Iout = Em[q] + Tr[q] * (Real)ldexp(iin, -A::uexp(iin));
                    const int s = g0 + q;
                    if (A::kahan) {
                        Real y = w * Iout - C0[s];
                        Real t = S0[s] + y;
                        C0[s] = (t - S0[s]) - y;
                        S0[s] = t;
                        y = wc * Iout - C2[s];
                        t = S2[s] + y;
                        C2[s] = (t - S2[s]) - y;
                        S2[s] = t;
                    } else {
                        S0[s] = S0[s] + w * Iout;
                        S2[s] = S2[s] + wc * Iout;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();  // every wave is done with its staging area
        for (int ww = 0; ww < W; ww++) {
            if (wave == ww) {
#pragma unroll
                for (int s = 0; s < MAXS; s++) {
                    const int q = lane + 64 * s;
                    const double v0 = (double)S0[s] - (double)C0[s];
                    const double v2 = (double)S2[s] - (double)C2[s];
                    red0[q] = (ww == 0 ? 0.0 : red0[q]) + v0;
                    red2[q] = (ww == 0 ? 0.0 : red2[q]) + v2;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x < R) {
            const int l = threadIdx.x;
            const double spi = 1.7724538509055160273;
            const double d_v = 2.0 * bd / nvel;
            for (int c = 0; c < C; c++) {
                const int p = l * C + c;
                if (p < p0 || p >= p0 + 64 * MAXS) continue;
                const double v0 = (2.0 * bd) * (c - nvel) / nvel;
                const double aa = v0 / bd;
                const double phi0 = (1.0 / (spi * bd)) * exp(-(aa * aa));
                jl0 = jl0 + phi0 * d_v * red0[p - p0];
                jl2 = jl2 + phi0 * d_v * red2[p - p0];
            }
        }
        __syncthreads();  // the staging area is reused by the next pair chunk
    }
    (void)jacc;
    if (threadIdx.x < R) {
        const int ue = A::uexp(a.I_in[threadIdx.x]);  // back from the intensity unit (exact)
        a.J0[(size_t)node * R + threadIdx.x] = ldexp(jl0, ue);
        a.J2[(size_t)node * R + threadIdx.x] = ldexp(jl2 / 1.4142135623730951, ue);
    }
}

template <typename Real, int MAXS, int GROUP>
__global__ void __launch_bounds__(256) rt_staged_kernel(RtArgs a, int cap) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    rt_staged_node<Real, MAXS, GROUP>(a, rt_item(a, xcd_block(blockIdx.x, gridDim.x)), cap, smem);
}

// Persistent form for batches of independent models: every workgroup takes the
// next (model, node) item from a global counter until the list is exhausted.  The
// list is ordered by decreasing estimated cost, so the long nodes start first.
template <typename Real, int MAXS, int GROUP>
__global__ void __launch_bounds__(256) rt_persistent_kernel(const RtArgs *models, const int2 *items, int n_items,
                                                            int *counter, int cap) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= n_items) break;
        const int2 mi = items[it];
        rt_staged_node<Real, MAXS, GROUP>(models[mi.x], mi.y, cap, smem);
    }
}

// LDS bytes of one workgroup for path capacity `cap`
template <typename Real, int MAXS>
size_t staged_lds_bytes(int cap, int R, int W) {
    const size_t stage = (size_t)W * cap * stage_bytes_per_cell<Real>(R);
    const size_t red = sizeof(double) * (2 * 64 * MAXS + 128);
    return stage > red ? stage : red;
}

template <typename Real>
hipError_t launch_staged(const RtArgs &a, int cap, hipStream_t s) {
    const int W = a.waves;
    const int npairs = a.R * a.C;
    if (npairs <= 64 * 4) {
        const size_t lds = staged_lds_bytes<Real, 4>(cap, a.R, W);
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        hipLaunchKernelGGL((rt_staged_kernel<Real, 4, 4>), dim3(a.n_work), dim3(64 * W), lds, s, a, cap);
    } else if (npairs <= 64 * 12) {
        const size_t lds = staged_lds_bytes<Real, 12>(cap, a.R, W);
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        hipLaunchKernelGGL((rt_staged_kernel<Real, 12, 4>), dim3(a.n_work), dim3(64 * W), lds, s, a, cap);
    } else {
        const size_t lds = staged_lds_bytes<Real, 20>(cap, a.R, W);
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        hipLaunchKernelGGL((rt_staged_kernel<Real, 20, 4>), dim3(a.n_work), dim3(64 * W), lds, s, a, cap);
    }
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_persistent(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap, int R_max,
                             int waves, int n_blocks, hipStream_t s) {
    const size_t lds = staged_lds_bytes<Real, 20>(cap, R_max, waves);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL((rt_persistent_kernel<Real, 20, 4>), dim3(n_blocks), dim3(64 * waves), lds, s, models, items,
                       n_items, counter, cap);
    return hipGetLastError();
}

}  // namespace portal
