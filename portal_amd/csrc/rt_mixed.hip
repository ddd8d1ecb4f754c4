// Inward ray integration in mixed precision.
//
// Same sum as rt_f64.hip (ray_in, inverse_rt.f90:217-344; J accumulation
// inverse_rt.f90:126-129), evaluated in FP32: the Doppler profile and the
// attenuation use the hardware exponential (v_exp_f32), 1 - e^-tau uses a series
// below tau = 1e-2 so thin cells keep full relative accuracy, and the per-pair sums
// over rays are Kahan-compensated FP32.  The line coefficients are formed in FP64
// (line_prep_kernel) and rounded once, because the reference's constants (h nu/c,
// h nu/4pi) underflow FP32; S and the background are carried in the line's
// intensity unit 2^line_iexp(I_in).  Geometry (the stored paths) is the FP64 walk's.
// The channel sum and the sum over waves are done in FP64.
#include "common.h"
#include "rt_body.h"
#include "rt_coop.h"

namespace portal {

__device__ __forceinline__ int xcd_remap_m(int bid, int nblocks) {
    const int per = (nblocks + 7) >> 3;
    const int xcd = bid & 7, slot = bid >> 3;
    const int full = nblocks - (per - 1) * 8;
    return xcd < full ? xcd * per + slot : full * per + (xcd - full) * (per - 1) + slot;
}

__device__ __forceinline__ float one_minus_exp_neg(float tau, float et) {
    // 1 - exp(-tau) without cancellation for small tau
    return tau < 1e-2f ? tau * (1.0f - tau * (0.5f - tau * (0.16666667f - tau * 0.041666668f))) : 1.0f - et;
}

template <int MAXS>
__global__ void __launch_bounds__(256) rt_mixed_kernel(RtArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
    const int bid = rt_item(a, xcd_remap_m(blockIdx.x, gridDim.x));
    const int node = a.work[bid];
    const double bd = a.velb[node].w;
    const float b = (float)bd;
    const float inv_bn = (float)(1.0 / bd);
    const int C = a.C, nvel = a.nvel, R = a.R;
    const int npairs = R * C;
    const float dls = (float)a.dl_scale;
    __shared__ double red0[64 * MAXS], red2[64 * MAXS];
    __shared__ double jacc[2 * 64];

    if (threadIdx.x < 2 * 64) jacc[threadIdx.x] = 0.0;
    for (int p0 = 0; p0 < npairs; p0 += 64 * MAXS) {
        float S0s[MAXS], C0s[MAXS], S2s[MAXS], C2s[MAXS];
#pragma unroll
        for (int s = 0; s < MAXS; s++) S0s[s] = C0s[s] = S2s[s] = C2s[s] = 0.0f;
        for (int r = wave; r < a.D; r += W) {
            const float mu = (float)a.xth[3 * r], w = (float)a.xth[3 * r + 2];
            const float wc = w * ((3.0f * (mu * mu) - 1.0f) * 0.5f);
            const long long off = a.offsets[(long long)bid * a.D + r];
            const int nc = (int)(a.offsets[(long long)bid * a.D + r + 1] - off);
            const PathEntry *P = a.paths + off;
            const float vel_node = (float)P[0].vray;
            const float len0 = (float)P[0].len * dls;
#pragma unroll
            for (int s = 0; s < MAXS; s++) {
                const int p = p0 + lane + 64 * s;
                if (p < npairs) {
                    const int l = p / C, nv = p - l * C - nvel;
                    const float v0 = (2.0f * b) * (float)nv / (float)nvel;
                    const float aa = v0 * inv_bn;
                    const float e0 = __expf(-(aa * aa));
                    const float2 ks0 = *reinterpret_cast<const float2 *>(a.KS32 + 2 * ((size_t)node * R + l));
                    const float tau0 = ks0.x * e0 * len0;
                    float Tr = __expf(-tau0);
                    float Em = ks0.y * one_minus_exp_neg(tau0, Tr);
                    const float v0s = v0 + vel_node;
                    for (int i = 1; i < nc; i++) {
                        if (Tr == 0.0f) break;
                        const PathEntry e = P[i];
                        const float2 ks = *reinterpret_cast<const float2 *>(a.KS32 + 2 * ((size_t)e.cell * R + l));
                        const float u = (v0s - (float)e.vray) * (float)a.inv_b[e.cell];
                        const float tau = ks.x * __expf(-(u * u)) * ((float)e.len * dls);
                        if (tau > 1e-30f) {
                            const float et = __expf(-tau);
                            Em = Em + Tr * ks.y * one_minus_exp_neg(tau, et);
                            Tr = Tr * et;
                        }
                    }
                    const double iin = a.I_in[l];
                    const float Iout = Em + Tr * (float)ldexp(iin, -line_iexp(iin));
                    // Kahan-compensated sums over rays
                    float y = w * Iout - C0s[s];
                    float t = S0s[s] + y;
                    C0s[s] = (t - S0s[s]) - y;
                    S0s[s] = t;
                    y = wc * Iout - C2s[s];
                    t = S2s[s] + y;
                    C2s[s] = (t - S2s[s]) - y;
                    S2s[s] = t;
                }
            }
        }
        for (int ww = 0; ww < W; ww++) {
            if (wave == ww) {
#pragma unroll
                for (int s = 0; s < MAXS; s++) {
                    const int q = lane + 64 * s;
                    const double v0 = (double)S0s[s] - (double)C0s[s];
                    const double v2 = (double)S2s[s] - (double)C2s[s];
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
            double j0 = jacc[l], j2 = jacc[64 + l];
            for (int c = 0; c < C; c++) {
                const int p = l * C + c;
                if (p < p0 || p >= p0 + 64 * MAXS) continue;
                const double v0 = (2.0 * bd) * (c - nvel) / nvel;
                const double aa = v0 / bd;
                const double phi0 = (1.0 / (spi * bd)) * exp(-(aa * aa));
                j0 = j0 + phi0 * d_v * red0[p - p0];
                j2 = j2 + phi0 * d_v * red2[p - p0];
            }
            jacc[l] = j0;
            jacc[64 + l] = j2;
        }
        __syncthreads();
    }
    if (threadIdx.x < R) {
        // back from the line's intensity unit (exact)
        const int ue = line_iexp(a.I_in[threadIdx.x]);
        a.J0[(size_t)node * R + threadIdx.x] = ldexp(jacc[threadIdx.x], ue);
        a.J2[(size_t)node * R + threadIdx.x] = ldexp(jacc[64 + threadIdx.x] / 1.4142135623730951, ue);
    }
}

hipError_t launch_rt_mixed(const RtArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    if (a.R > 64) return hipErrorInvalidValue;
    // dust, star and the own-cell operator: cooperative kernel only
    if ((a.dust || a.xthn || a.LAM) && a.variant != 2 && a.variant != 3) return hipErrorInvalidValue;
    if (a.variant == 1) return launch_staged<float>(a, a.max_nc, s);
    if (a.variant == 2 || a.variant == 3) return launch_coop<float>(a, a.max_nc, s);
    const int W = a.waves;
    const int npairs = a.R * a.C;
    if (npairs <= 64 * 4)
        hipLaunchKernelGGL(rt_mixed_kernel<4>, dim3(a.n_work), dim3(64 * W), 0, s, a);
    else if (npairs <= 64 * 10)
        hipLaunchKernelGGL(rt_mixed_kernel<10>, dim3(a.n_work), dim3(64 * W), 0, s, a);
    else
        hipLaunchKernelGGL(rt_mixed_kernel<20>, dim3(a.n_work), dim3(64 * W), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rt_persistent_f64(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap,
                                    int R_max, int waves, int n_blocks, hipStream_t s);
size_t rt_persistent_lds_f64(int cap, int R_max, int waves);

hipError_t launch_rt_persistent(bool mixed, const RtArgs *models, const int2 *items, int n_items, int *counter,
                                int cap, int R_max, int waves, int n_blocks, hipStream_t s) {
    if (!mixed) return launch_rt_persistent_f64(models, items, n_items, counter, cap, R_max, waves, n_blocks, s);
    return launch_persistent<float>(models, items, n_items, counter, cap, R_max, waves, n_blocks, s);
}

hipError_t launch_coop_persistent_f64(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap,
                                      int slots, int waves, size_t lds, int n_blocks, bool dust, bool btf,
                                      hipStream_t s);

hipError_t launch_rt_coop_persistent(bool mixed, const RtArgs *models, const int2 *items, int n_items, int *counter,
                                     int cap, int slots, int waves, size_t lds, int n_blocks, bool dust, bool btf,
                                     int prof, hipStream_t s) {
    if (!mixed)
        return launch_coop_persistent_f64(models, items, n_items, counter, cap, slots, waves, lds, n_blocks, dust,
                                          btf, s);
    return launch_coop_persistent<float>(models, items, n_items, counter, cap, slots, waves, lds, n_blocks, dust,
                                         btf, prof, s);
}

int coop_profile_form_host(int R, int C, int W, int slots, double b_ratio, int nvel) {
    return coop_profile_form(R, C, W, slots, b_ratio, nvel);
}

int coop_slots_host(int R, int C, int W) { return coop_slots(R, C, W); }

size_t rt_coop_persistent_lds(bool mixed, int cap, int R_max, int C, bool dust, int waves) {
    return mixed ? coop_lds_bytes<float>(cap, R_max, R_max * C, dust, false, waves)
                 : coop_lds_bytes<double>(cap, R_max, R_max * C, dust, false, waves);
}

size_t rt_persistent_lds(bool mixed, int cap, int R_max, int waves) {
    return mixed ? staged_lds_bytes<float, 20>(cap, R_max, waves) : rt_persistent_lds_f64(cap, R_max, waves);
}

}  // namespace portal
