// Inward ray integration in FP64 (the parity path).
//
// Restates ray_in (inverse_rt.f90:217-344) and the J accumulation of inverse_rt
// (inverse_rt.f90:126-129) over the stored inward paths.  The reference integrates
// each channel back to front; here each channel is integrated front to back
// (I_out = sum_i S_i (1 - e^-tau_i) prod_{j<i} e^-tau_j + I_in prod_j e^-tau_j),
// the same sum in another order, with the reference's skip of segments whose
// tau is not > 1e-30 (NaN included) and no skip for the central cell.  A channel
// stops once its transmission is exactly zero, where every later term is zero.
//
// Work mapping: one workgroup per node, each wave takes every W-th ray, each lane
// owns (line, channel) pairs p = lane + 64*s.  Per pair the lane accumulates
// A0 = sum_rays w I and A2 = sum_rays w c_in I; the channel sum with the profile
// weight phi0*d_v and the sum over waves are done once per node in a fixed order,
// so J is deterministic.  Compiled with -ffp-contract=off.
#include "common.h"
#include "rt_body.h"
#include "rt_coop.h"

namespace portal {

__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous
    // range of (Morton-ordered) nodes so its L2 holds one region of the cloud.
    const int per = (nblocks + 7) >> 3;
    const int xcd = bid & 7, slot = bid >> 3;
    const int full = nblocks - (per - 1) * 8;  // XCDs that get `per` blocks
    return xcd < full ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
}

__global__ void __launch_bounds__(256) line_prep_kernel(PrepArgs a) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)a.N * a.R) return;
    const int n = (int)(id / a.R), l = (int)(id % a.R);
    const double p1 = a.levp[(size_t)n * a.L + a.l1[l]];
    const double p2 = a.levp[(size_t)n * a.L + a.l2[l]];
    const double kv = a.c1B[l] * (p1 - a.g1[l] * p2 / a.g2[l]);
    const double ev = a.c2A[l] * p2;
    const double b = a.velb[n].w;
    const double K = kv * (a.cnu_spi[l] * (1.0 / b));
    const double S = ev / kv;
    if (a.KS) {
        a.KS[2 * id] = K;
        a.KS[2 * id + 1] = S;
    }
    // FP32: S in the line's intensity unit 2^line_iexp(I_in)
    const int ue = a.KS32 || a.KSD32 ? line_iexp(a.I_in[l]) : 0;
    if (a.KS32) {
        a.KS32[2 * id] = (float)K;
        a.KS32[2 * id + 1] = (float)ldexp(S, -ue);
    }
    if (a.KSD || a.KSD32) {
        // dust: alpha = kd dust_p, j = alpha B_nu(T_dust) (inverse_rt.f90:292-297,
        // black_body miscellaneous.f90:263-285)
        const double alpha = a.kd[l] * a.dust[n];
        const double T = a.tdust[n], nu0 = a.nu[l];
        const double Bb = T == 0.0 ? 0.0 : 1.4744995e-50 * pow(nu0, 3.0) * pow(exp(4.79924e-11 / T * nu0) - 1.0, -1.0);
        if (a.KSD) {
            a.KSD[2 * id] = alpha;
            a.KSD[2 * id + 1] = Bb;
        }
        if (a.KSD32) {
            a.KSD32[2 * id] = (float)alpha;
            a.KSD32[2 * id + 1] = (float)ldexp(Bb, -ue);
        }
    }
}

hipError_t launch_line_prep(const PrepArgs &a, hipStream_t s) {
    const long long n = (long long)a.N * a.R;
    hipLaunchKernelGGL(line_prep_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <int MAXS>
__global__ void __launch_bounds__(256) rt_f64_kernel(RtArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
    const int bid = rt_item(a, xcd_remap(blockIdx.x, gridDim.x));
    const int node = a.work[bid];
    const double b = a.velb[node].w;
    const double spi = 1.7724538509055160273;  // sqrt(pi)
    const int C = a.C, nvel = a.nvel, R = a.R;
    const int npairs = R * C;
    const double d_v = 2.0 * b / nvel;
    const double dls = a.dl_scale;
    __shared__ double red0[64 * MAXS], red2[64 * MAXS];
    __shared__ double jacc[2 * 64];

    if (threadIdx.x < 2 * 64) jacc[threadIdx.x] = 0.0;
    for (int p0 = 0; p0 < npairs; p0 += 64 * MAXS) {
        double A0[MAXS], A2[MAXS];
#pragma unroll
        for (int s = 0; s < MAXS; s++) A0[s] = A2[s] = 0.0;
        for (int r = wave; r < a.D; r += W) {
            const double mu = a.xth[3 * r], w = a.xth[3 * r + 2];
            const double c_in = (3.0 * (mu * mu) - 1.0) / 2.0;
            const long long off = a.offsets[(long long)bid * a.D + r];
            const int nc = (int)(a.offsets[(long long)bid * a.D + r + 1] - off);
            const PathEntry *P = a.paths + off;
            const double vel_node = P[0].vray;
            const double len0 = (double)P[0].len;
#pragma unroll
            for (int s = 0; s < MAXS; s++) {
                const int p = p0 + lane + 64 * s;
                if (p < npairs) {
                    const int l = p / C, nv = p - l * C - nvel;
                    const double v0 = (2.0 * b) * nv / nvel;
                    const double aa = v0 / b;
                    const double e0 = exp(-(aa * aa));
                    const double K0 = a.KS[2 * ((size_t)node * R + l)];
                    const double S0 = a.KS[2 * ((size_t)node * R + l) + 1];
                    const double tau0 = K0 * e0 * (len0 * dls);
                    double Tr = exp(-tau0);
                    double Em = S0 * (1.0 - Tr);
                    const double v0s = v0 + vel_node;
                    for (int i = 1; i < nc; i++) {
                        if (Tr == 0.0) break;
                        const PathEntry e = P[i];
                        const double2 ks = *reinterpret_cast<const double2 *>(a.KS + 2 * ((size_t)e.cell * R + l));
                        const double u = (v0s - e.vray) * a.inv_b[e.cell];
                        const double tau = ks.x * exp(-(u * u)) * ((double)e.len * dls);
                        if (tau > 1e-30) {
                            const double et = exp(-tau);
                            Em = Em + Tr * ks.y * (1.0 - et);
                            Tr = Tr * et;
                        }
                    }
                    const double Iout = Em + Tr * a.I_in[l];
                    A0[s] = A0[s] + w * Iout;
                    A2[s] = A2[s] + (w * c_in) * Iout;
                }
            }
        }
        // sum the waves' accumulators in a fixed order
        for (int ww = 0; ww < W; ww++) {
            if (wave == ww) {
#pragma unroll
                for (int s = 0; s < MAXS; s++) {
                    const int q = lane + 64 * s;
                    red0[q] = (ww == 0 ? 0.0 : red0[q]) + A0[s];
                    red2[q] = (ww == 0 ? 0.0 : red2[q]) + A2[s];
                }
            }
            __syncthreads();
        }
        // channel sums with the profile weight, one thread per line, channels in order
        if (threadIdx.x < R) {
            const int l = threadIdx.x;
            double j0 = jacc[l], j2 = jacc[64 + l];
            for (int c = 0; c < C; c++) {
                const int p = l * C + c;
                if (p < p0 || p >= p0 + 64 * MAXS) continue;
                const int nv = c - nvel;
                const double v0 = (2.0 * b) * nv / nvel;
                const double aa = v0 / b;
                const double phi0 = (1.0 / (spi * b)) * exp(-(aa * aa));
                j0 = j0 + phi0 * d_v * red0[p - p0];
                j2 = j2 + phi0 * d_v * red2[p - p0];
            }
            jacc[l] = j0;
            jacc[64 + l] = j2;
        }
        __syncthreads();
    }
    if (threadIdx.x < R) {
        a.J0[(size_t)node * R + threadIdx.x] = jacc[threadIdx.x];
        a.J2[(size_t)node * R + threadIdx.x] = jacc[64 + threadIdx.x] / sqrt(2.0);
    }
}

hipError_t launch_rt_f64(const RtArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    if (a.R > 64) return hipErrorInvalidValue;
    // dust, star and the own-cell operator: cooperative kernel only
    if ((a.dust || a.xthn || a.LAM) && a.variant != 2 && a.variant != 3) return hipErrorInvalidValue;
    if (a.variant == 1) return launch_staged<double>(a, a.max_nc, s);
    if (a.variant == 2 || a.variant == 3) return launch_coop<double>(a, a.max_nc, s);
    const int W = a.waves;
    const int npairs = a.R * a.C;
    if (npairs <= 64 * 4)
        hipLaunchKernelGGL(rt_f64_kernel<4>, dim3(a.n_work), dim3(64 * W), 0, s, a);
    else if (npairs <= 64 * 10)
        hipLaunchKernelGGL(rt_f64_kernel<10>, dim3(a.n_work), dim3(64 * W), 0, s, a);
    else
        hipLaunchKernelGGL(rt_f64_kernel<20>, dim3(a.n_work), dim3(64 * W), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rt_persistent_f64(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap,
                                    int R_max, int waves, int n_blocks, hipStream_t s) {
    return launch_persistent<double>(models, items, n_items, counter, cap, R_max, waves, n_blocks, s);
}

size_t rt_persistent_lds_f64(int cap, int R_max, int waves) { return staged_lds_bytes<double, 20>(cap, R_max, waves); }

hipError_t launch_coop_persistent_f64(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap,
                                      int slots, int waves, size_t lds, int n_blocks, bool dust, bool btf,
                                      hipStream_t s) {
    return launch_coop_persistent<double>(models, items, n_items, counter, cap, slots, waves, lds, n_blocks, dust,
                                          btf, PROF_DIRECT, s);
}

}  // namespace portal
