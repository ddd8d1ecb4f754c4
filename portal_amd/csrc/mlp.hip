// Surrogate network: a dense MLP (width x depth hidden layers, SiLU) trained and run
// with hand-written BF16 MFMA kernels and FP32 accumulation.
//
// There is no counterpart in the reference.  The network maps local node features
// (log n_H2, log T, per-line J^0_0 and J^2_0/J^0_0 of the first pass) to the
// populations and alignment of the converged iteration; the lambda iteration uses
// it as its initial iterate (ctx.cpp) and the FP64 SE solve judges and corrects it.
//
// GEMM: C[M x N] = A[M x K] B[K x N] with bf16 operands read through arbitrary
// strides (so the transposed products of back-propagation need no copies).  A block of
// four waves owns a TM x TN tile (64 x 64, or 128 x 128 when that still gives a few
// blocks per CU), each wave a (TM/2) x (TN/2) quarter of 16 x 16 fragments; K is staged
// through LDS 32 at a time, the next slice's global loads in flight while the MFMAs
// consume the current one.  An operand whose k (or whose m / n) is the unit stride is
// read 8 elements (16 bytes) per load; other strides fall back to element loads.
// v_mfma_f32_16x16x32_bf16: each lane feeds 8 consecutive k of one row of A and one
// column of B and receives 4 rows of one column of the 16x16 result.  Epilogues fuse
// bias, SiLU, the SiLU derivative and split-K partial stores.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "common.h"

namespace portal {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// epilogues: split-K partial store; + bias; + bias, keeping Z and H = SiLU(Z) (training);
// dZ = dH SiLU'(Z); + bias, H = SiLU(Z) only (inference, Z is not kept)
enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2, EPI_DSILU = 3, EPI_BIAS_SILU_H = 4 };

struct GemmArgs {
    int M, N, K, ksplit, epi;
    const bf16 *A;
    long long sam, sak;
    const bf16 *B;
    long long sbk, sbn;
    const float *bias;
    float *C;
    long long ldc;
    bf16 *H;
    long long ldh;
    const float *Z;
    long long ldz;
};

constexpr int BK = 32, PAD = 8;

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

// how an operand's TR x BK slice (TR rows of m or n, BK of k) is fetched: 8 consecutive k
// of one row per load (KVEC), 8 consecutive rows at one k (RVEC), or element by element
enum { LD_ELEM = 0, LD_KVEC = 1, LD_RVEC = 2 };

template <int TR>
struct Stage {
    static constexpr int NV = TR * BK / 8 / 256;  // 8-element groups per thread
    uint4 v[NV];
};

// load thread tid's groups of the slice rows r0.., k0.. of X (row stride sr, k stride sk)
template <int TR>
__device__ __forceinline__ void stage_load(Stage<TR> &st, int mode, const bf16 *X, long long sr, long long sk,
                                           int r0, int rlim, int k0, int klim, int tid) {
#pragma unroll
    for (int q = 0; q < Stage<TR>::NV; q++) {
        const int e = tid + 256 * q;
        int r, c;
        if (mode == LD_RVEC) {
            r = (e % (TR / 8)) * 8;
            c = e / (TR / 8);
        } else {
            r = e / (BK / 8);
            c = (e % (BK / 8)) * 8;
        }
        const int rr = r0 + r, kk = k0 + c;
        union {
            uint4 u;
            bf16 h[8];
        } x;
        const bool full = mode == LD_RVEC ? (rr + 7 < rlim && kk < klim) : (rr < rlim && kk + 7 < klim);
        if (mode != LD_ELEM && full) {
            x.u = *reinterpret_cast<const uint4 *>(X + (long long)rr * sr + (long long)kk * sk);
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int ri = mode == LD_RVEC ? rr + i : rr, ki = mode == LD_RVEC ? kk : kk + i;
                x.h[i] = (ri < rlim && ki < klim) ? X[(long long)ri * sr + (long long)ki * sk] : (bf16)0.0f;
            }
        }
        st.v[q] = x.u;
    }
}

// store the groups into S[row][k] (k contiguous, row stride BK + PAD)
template <int TR>
__device__ __forceinline__ void stage_store(const Stage<TR> &st, int mode, bf16 (*S)[BK + PAD], int tid) {
#pragma unroll
    for (int q = 0; q < Stage<TR>::NV; q++) {
        const int e = tid + 256 * q;
        if (mode == LD_RVEC) {
            const int r = (e % (TR / 8)) * 8, c = e / (TR / 8);
            union {
                uint4 u;
                bf16 h[8];
            } x;
            x.u = st.v[q];
#pragma unroll
            for (int i = 0; i < 8; i++) S[r + i][c] = x.h[i];
        } else {
            const int r = e / (BK / 8), c = (e % (BK / 8)) * 8;
            *reinterpret_cast<uint4 *>(&S[r][c]) = st.v[q];
        }
    }
}

template <int TM, int TN>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g, int amode, int bmode) {
    constexpr int FI = TM / 32, FJ = TN / 32;  // 16x16 fragments per wave along m and n
    __shared__ __attribute__((aligned(16))) bf16 As[TM][BK + PAD];
    __shared__ __attribute__((aligned(16))) bf16 Bs[TN][BK + PAD];  // transposed: Bs[n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(g.K, kbeg + kchunk);
    f32x4 acc[FI][FJ];
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    Stage<TM> sa;
    Stage<TN> sb;
    if (kbeg < kend) {
        stage_load<TM>(sa, amode, g.A, g.sam, g.sak, m0, g.M, kbeg, kend, tid);
        stage_load<TN>(sb, bmode, g.B, g.sbn, g.sbk, n0, g.N, kbeg, kend, tid);
    }
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        stage_store<TM>(sa, amode, As, tid);
        stage_store<TN>(sb, bmode, Bs, tid);
        __syncthreads();
        if (k0 + BK < kend) {
            stage_load<TM>(sa, amode, g.A, g.sam, g.sak, m0, g.M, k0 + BK, kend, tid);
            stage_load<TN>(sb, bmode, g.B, g.sbn, g.sbk, n0, g.N, k0 + BK, kend, tid);
        }
        const int kk = (lane >> 4) * 8;
        bf16x8 b[FJ];
#pragma unroll
        for (int j = 0; j < FJ; j++)
            b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn * (TN / 2) + j * 16 + (lane & 15)][kk]);
#pragma unroll
        for (int i = 0; i < FI; i++) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm * (TM / 2) + i * 16 + (lane & 15)][kk]);
#pragma unroll
            for (int j = 0; j < FJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int m = m0 + wm * (TM / 2) + i * 16 + (lane >> 4) * 4 + t;
                const int n = n0 + wn * (TN / 2) + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][t];
                switch (g.epi) {
                    case EPI_STORE:
                        g.C[((long long)blockIdx.z * g.M + m) * g.ldc + n] = v;
                        break;
                    case EPI_BIAS:
                        g.C[(long long)m * g.ldc + n] = v + g.bias[n];
                        break;
                    case EPI_BIAS_SILU: {
                        const float z = v + g.bias[n];
                        g.C[(long long)m * g.ldc + n] = z;
                        g.H[(long long)m * g.ldh + n] = (bf16)silu(z);
                        break;
                    }
                    case EPI_BIAS_SILU_H:
                        g.H[(long long)m * g.ldh + n] = (bf16)silu(v + g.bias[n]);
                        break;
                    default: {  // EPI_DSILU: dZ = dH * silu'(Z)
                        const float z = g.Z[(long long)m * g.ldz + n];
                        const float s = 1.0f / (1.0f + __expf(-z));
                        const float dz = v * s * (1.0f + z * (1.0f - s));
                        g.H[(long long)m * g.ldh + n] = (bf16)dz;
                        if (g.C) g.C[(long long)m * g.ldc + n] = dz;
                        break;
                    }
                }
            }
}

// the 16-byte form for an operand with k stride sk and row (m or n) stride sr: its base
// must be 16-byte aligned and the stride that is not 1 a multiple of 8 elements
static int load_mode(const bf16 *X, long long sk, long long sr) {
    const bool aligned = (reinterpret_cast<uintptr_t>(X) & 15) == 0;
    if (aligned && sk == 1 && sr % 8 == 0) return LD_KVEC;
    if (aligned && sr == 1 && sk % 8 == 0) return LD_RVEC;
    return LD_ELEM;
}

static hipError_t gemm(const GemmArgs &g, hipStream_t s) {
    const int am = load_mode(g.A, g.sak, g.sam);
    const int bm = load_mode(g.B, g.sbk, g.sbn);
    // the large tile only where it still leaves two blocks per CU
    const long long big = (long long)((g.N + 127) / 128) * ((g.M + 127) / 128) * g.ksplit;
    if (big >= 512) {
        dim3 grid((g.N + 127) / 128, (g.M + 127) / 128, g.ksplit);
        hipLaunchKernelGGL((gemm_bf16_kernel<128, 128>), grid, dim3(256), 0, s, g, am, bm);
    } else {
        dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, g.ksplit);
        hipLaunchKernelGGL((gemm_bf16_kernel<64, 64>), grid, dim3(256), 0, s, g, am, bm);
    }
    return hipGetLastError();
}

__global__ void split_reduce_kernel(const float *P, float *out, int S, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < S; z++) s += P[(long long)z * n + i];
    out[i] = s;
}

// column sums of an M x N fp32 matrix (bias gradients) in two fixed-order stages:
// block (x, y) sums columns 64x..64x+63 over row block y (each wave a quarter of its
// rows, lane = column), then colsum_final_kernel adds the COLSUM_RB row blocks in order
constexpr int COLSUM_RB = 16;

__global__ void __launch_bounds__(256) colsum_kernel(const float *X, float *part, int M, int N) {
    __shared__ float sh[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane;
    const int rows = (M + COLSUM_RB - 1) / COLSUM_RB;
    const int r0 = blockIdx.y * rows, r1 = min(M, r0 + rows);
    float s = 0.0f;
    if (n < N)
        for (int m = r0 + w; m < r1; m += 4) s += X[(long long)m * N + n];
    sh[w][lane] = s;
    __syncthreads();
    if (w == 0 && n < N) part[(long long)blockIdx.y * N + n] = ((sh[0][lane] + sh[1][lane]) + sh[2][lane]) + sh[3][lane];
}

__global__ void colsum_final_kernel(const float *part, float *out, int N) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int y = 0; y < COLSUM_RB; y++) s += part[(long long)y * N + n];
    out[n] = s;
}

__global__ void to_bf16_kernel(const float *x, bf16 *y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

__global__ void normalise_kernel(const float *x, const float *mu, const float *sd, bf16 *y, int B, int d) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)B * d) return;
    const int k = (int)(i % d);
    y[i] = (bf16)((x[i] - mu[k]) / sd[k]);
}

// d(loss)/dY for loss = mean over batch and outputs of (Y - T)^2 (targets normalised)
__global__ void mse_grad_kernel(const float *Y, const float *T, const float *mu, const float *sd, float *dY,
                                bf16 *dY16, float *part, int B, int d) {
    __shared__ float sh[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < (long long)B * d) {
        const int k = (int)(i % d);
        const float t = (T[i] - mu[k]) / sd[k];
        const float r = Y[i] - t;
        l = r * r;
        const float gr = 2.0f * r / (float)((long long)B * d);
        dY[i] = gr;
        dY16[i] = (bf16)gr;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// acc += (sum of the block partials) / ny, in a fixed order (one block of 256)
__global__ void __launch_bounds__(256) loss_acc_kernel(const float *part, int nbl, long long ny, double *acc) {
    __shared__ double sh[256];
    double v = 0.0;
    for (int b = threadIdx.x; b < nbl; b += 256) v += (double)part[b];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += sh[0] / (double)ny;
}

__global__ void adam_kernel(float *w, const float *g, float *m, float *v, bf16 *w16, long long n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = (bf16)wi;
}

__global__ void denorm_kernel(const float *Yn, const float *mu, const float *sd, float *Y, int B, int d) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)B * d) return;
    const int k = (int)(i % d);
    Y[i] = Yn[i] * sd[k] + mu[k];
}

struct Mlp {
    int n_in = 0, width = 0, depth = 0, n_out = 0, nl = 0;
    std::vector<int> din, dout;
    std::vector<long long> woff, boff;  // offsets into the flat parameter vector
    long long n_params = 0;
    float *p32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
    bf16 *p16 = nullptr;
    float *norm = nullptr;  // in_mu, in_sd, out_mu, out_sd
    int cap = 0;
    bf16 *X16 = nullptr;
    std::vector<float *> Z;   // per layer fp32 pre-activation [cap x dout]
    std::vector<bf16 *> H;    // per layer bf16 activation [cap x dout]
    float *dZ32 = nullptr;    // [cap x max(dout)]
    bf16 *dZa = nullptr, *dZb = nullptr;
    float *partial = nullptr;
    long long partial_n = 0;
    float *Xd = nullptr, *Td = nullptr, *Yd = nullptr;
    float *loss_part = nullptr;
    double *loss_acc = nullptr;  // per-epoch loss sum on device
    long long step = 0;
};

constexpr int KSPLIT_MAX = 32;

static void mlp_free_buffers(Mlp &m) {
    auto f = [](auto *&p) {
        if (p) hipFree((void *)p);
        p = nullptr;
    };
    f(m.X16);
    for (auto *&z : m.Z) f(z);
    for (auto *&h : m.H) f(h);
    f(m.dZ32);
    f(m.dZa);
    f(m.dZb);
    f(m.partial);
    f(m.Xd);
    f(m.Td);
    f(m.Yd);
    f(m.loss_part);
    f(m.loss_acc);
    m.cap = 0;
}

static hipError_t mlp_reserve(Mlp &m, int B) {
    if (B <= m.cap) return hipSuccess;
    mlp_free_buffers(m);
    hipError_t e;
    const int mx = std::max(m.width, std::max(m.n_out, m.n_in));
    if ((e = hipMalloc((void **)&m.X16, sizeof(bf16) * (size_t)B * m.n_in))) return e;
    m.Z.assign(m.nl, nullptr);
    m.H.assign(m.nl, nullptr);
    for (int l = 0; l < m.nl; l++) {
        if ((e = hipMalloc((void **)&m.Z[l], sizeof(float) * (size_t)B * m.dout[l]))) return e;
        if ((e = hipMalloc((void **)&m.H[l], sizeof(bf16) * (size_t)B * m.dout[l]))) return e;
    }
    if ((e = hipMalloc((void **)&m.dZ32, sizeof(float) * (size_t)B * mx))) return e;
    if ((e = hipMalloc((void **)&m.dZa, sizeof(bf16) * (size_t)B * mx))) return e;
    if ((e = hipMalloc((void **)&m.dZb, sizeof(bf16) * (size_t)B * mx))) return e;
    m.partial_n = (long long)KSPLIT_MAX * mx * mx;
    if ((e = hipMalloc((void **)&m.partial, sizeof(float) * m.partial_n))) return e;
    if ((e = hipMalloc((void **)&m.Xd, sizeof(float) * (size_t)B * m.n_in))) return e;
    if ((e = hipMalloc((void **)&m.Td, sizeof(float) * (size_t)B * m.n_out))) return e;
    if ((e = hipMalloc((void **)&m.Yd, sizeof(float) * (size_t)B * m.n_out))) return e;
    if ((e = hipMalloc((void **)&m.loss_part, sizeof(float) * ((size_t)B * m.n_out / 256 + 2)))) return e;
    if (!m.loss_acc && (e = hipMalloc((void **)&m.loss_acc, sizeof(double)))) return e;
    m.cap = B;
    return hipSuccess;
}

static unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

// forward on m.Xd (fp32 features, B x n_in): normalised output in m.Z[nl-1]; keep_z
// keeps every layer's pre-activation for back-propagation (training)
static hipError_t mlp_forward_dev(Mlp &m, int B, hipStream_t s, bool keep_z = true) {
    const float *mu = m.norm, *sd = m.norm + m.n_in;
    hipLaunchKernelGGL(normalise_kernel, dim3(blocks((long long)B * m.n_in)), dim3(256), 0, s, m.Xd, mu, sd, m.X16, B,
                       m.n_in);
    for (int l = 0; l < m.nl; l++) {
        GemmArgs g{};
        g.M = B;
        g.N = m.dout[l];
        g.K = m.din[l];
        g.ksplit = 1;
        g.A = l == 0 ? m.X16 : m.H[l - 1];
        g.sam = m.din[l];
        g.sak = 1;
        g.B = m.p16 + m.woff[l];
        g.sbk = m.dout[l];
        g.sbn = 1;
        g.bias = m.p32 + m.boff[l];
        g.C = m.Z[l];
        g.ldc = m.dout[l];
        g.H = m.H[l];
        g.ldh = m.dout[l];
        g.epi = (l < m.nl - 1) ? (keep_z ? EPI_BIAS_SILU : EPI_BIAS_SILU_H) : EPI_BIAS;
        hipError_t e = gemm(g, s);
        if (e) return e;
    }
    return hipGetLastError();
}

// gradients of the mean squared error w.r.t. every parameter into m.g32; returns loss
static hipError_t mlp_backward_dev(Mlp &m, int B, hipStream_t s, float *loss_out) {
    const int L = m.nl;
    const float *omu = m.norm + 2 * m.n_in, *osd = m.norm + 2 * m.n_in + m.n_out;
    const long long ny = (long long)B * m.n_out;
    const unsigned nbl = blocks(ny);
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nbl), dim3(256), 0, s, m.Z[L - 1], m.Td, omu, osd, m.dZ32, m.dZa,
                       m.loss_part, B, m.n_out);
    bf16 *dcur = m.dZa, *dnext = m.dZb;
    for (int l = L - 1; l >= 0; l--) {
        const int din = m.din[l], dout = m.dout[l];
        // dW = H_{l-1}^T dZ_l, split over the batch
        const int S = std::min(KSPLIT_MAX, std::max(1, B / 256));
        GemmArgs g{};
        g.M = din;
        g.N = dout;
        g.K = B;
        g.ksplit = S;
        g.A = l == 0 ? m.X16 : m.H[l - 1];
        g.sam = 1;
        g.sak = din;
        g.B = dcur;
        g.sbk = dout;
        g.sbn = 1;
        g.C = m.partial;
        g.ldc = dout;
        g.epi = EPI_STORE;
        hipError_t e = gemm(g, s);
        if (e) return e;
        const long long nw = (long long)din * dout;
        hipLaunchKernelGGL(split_reduce_kernel, dim3(blocks(nw)), dim3(256), 0, s, m.partial, m.g32 + m.woff[l], S,
                           nw);
        // bias gradient; the split-K buffer is free again once split_reduce has run
        hipLaunchKernelGGL(colsum_kernel, dim3((dout + 63) / 64, COLSUM_RB), dim3(256), 0, s, (const float *)m.dZ32,
                           m.partial, B, dout);
        hipLaunchKernelGGL(colsum_final_kernel, dim3((dout + 255) / 256), dim3(256), 0, s, (const float *)m.partial,
                           m.g32 + m.boff[l], dout);
        if (l > 0) {
            // dZ_{l-1} = (dZ_l W_l^T) * silu'(Z_{l-1})
            GemmArgs h{};
            h.M = B;
            h.N = din;
            h.K = dout;
            h.ksplit = 1;
            h.A = dcur;
            h.sam = dout;
            h.sak = 1;
            h.B = m.p16 + m.woff[l];
            h.sbk = 1;
            h.sbn = dout;
            h.Z = m.Z[l - 1];
            h.ldz = din;
            h.H = dnext;
            h.ldh = din;
            h.C = m.dZ32;
            h.ldc = din;
            h.epi = EPI_DSILU;
            e = gemm(h, s);
            if (e) return e;
            std::swap(dcur, dnext);
        }
    }
    if (loss_out) {
        std::vector<float> part(nbl);
        hipError_t e = hipMemcpyAsync(part.data(), m.loss_part, sizeof(float) * nbl, hipMemcpyDeviceToHost, s);
        if (e) return e;
        e = hipStreamSynchronize(s);
        if (e) return e;
        double tot = 0;
        for (float v : part) tot += v;
        *loss_out = (float)(tot / (double)ny);
    }
    return hipGetLastError();
}

static hipError_t mlp_adam(Mlp &m, float lr, hipStream_t s) {
    m.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)m.step), c2 = 1.0f - std::pow(b2, (float)m.step);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks(m.n_params)), dim3(256), 0, s, m.p32, m.g32, m.m32, m.v32, m.p16,
                       m.n_params, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

}  // namespace portal

using namespace portal;

namespace portal {
// used by ctx.cpp: run the network on device features (B x n_in, fp32) -> denormalised outputs
hipError_t mlp_infer_device(Ctx &c, int B, const float *feats, float *out) {
    Mlp &m = *(Mlp *)c.mlp;
    hipError_t e = mlp_reserve(m, B);
    if (e) return e;
    if ((e = hipMemcpyAsync(m.Xd, feats, sizeof(float) * (size_t)B * m.n_in, hipMemcpyDeviceToDevice, c.stream)))
        return e;
    if ((e = mlp_forward_dev(m, B, c.stream, false))) return e;
    hipLaunchKernelGGL(denorm_kernel, dim3(blocks((long long)B * m.n_out)), dim3(256), 0, c.stream,
                       m.Z[m.nl - 1], m.norm + 2 * m.n_in, m.norm + 2 * m.n_in + m.n_out, out, B, m.n_out);
    return hipGetLastError();
}
int mlp_dims(const Ctx &c, int *n_in, int *n_out) {
    if (!c.mlp) return 0;
    const Mlp &m = *(const Mlp *)c.mlp;
    *n_in = m.n_in;
    *n_out = m.n_out;
    return 1;
}
}  // namespace portal

void portal_mlp_free(Ctx &c) {
    if (!c.mlp) return;
    Mlp *m = (Mlp *)c.mlp;
    mlp_free_buffers(*m);
    auto f = [](auto *&p) {
        if (p) hipFree((void *)p);
        p = nullptr;
    };
    f(m->p32);
    f(m->g32);
    f(m->m32);
    f(m->v32);
    f(m->p16);
    f(m->norm);
    delete m;
    c.mlp = nullptr;
}

extern "C" {

int portal_mlp_create(portal_ctx *x, int n_in, int width, int depth, int n_out, unsigned long long seed) {
    if (!x || n_in <= 0 || width <= 0 || depth <= 0 || n_out <= 0 || width > 4096) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    portal_mlp_free(c);
    Mlp *m = new Mlp();
    c.mlp = m;
    m->n_in = n_in;
    m->width = width;
    m->depth = depth;
    m->n_out = n_out;
    m->nl = depth + 1;
    long long off = 0;
    for (int l = 0; l < m->nl; l++) {
        const int a = l == 0 ? n_in : width, b = l == m->nl - 1 ? n_out : width;
        m->din.push_back(a);
        m->dout.push_back(b);
        m->woff.push_back(off);
        off += (long long)a * b;
        m->boff.push_back(off);
        off += b;
    }
    m->n_params = off;
    // He-uniform initialisation, zero biases
    std::vector<float> p(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < m->nl; l++) {
        const float lim = std::sqrt(6.0f / (float)m->din[l]) * (l == m->nl - 1 ? 0.5f : 1.0f);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (long long i = 0; i < (long long)m->din[l] * m->dout[l]; i++) p[m->woff[l] + i] = U(rng);
    }
    std::vector<float> norm(2 * n_in + 2 * n_out);
    for (int i = 0; i < n_in; i++) norm[n_in + i] = 1.0f;
    for (int i = 0; i < n_out; i++) norm[2 * n_in + n_out + i] = 1.0f;
    HIP_TRY(&c, hipMalloc((void **)&m->p32, sizeof(float) * off));
    HIP_TRY(&c, hipMalloc((void **)&m->g32, sizeof(float) * off));
    HIP_TRY(&c, hipMalloc((void **)&m->m32, sizeof(float) * off));
    HIP_TRY(&c, hipMalloc((void **)&m->v32, sizeof(float) * off));
    HIP_TRY(&c, hipMalloc((void **)&m->p16, sizeof(bf16) * off));
    HIP_TRY(&c, hipMalloc((void **)&m->norm, sizeof(float) * norm.size()));
    HIP_TRY(&c, hipMemcpy(m->p32, p.data(), sizeof(float) * off, hipMemcpyHostToDevice));
    HIP_TRY(&c, hipMemset(m->m32, 0, sizeof(float) * off));
    HIP_TRY(&c, hipMemset(m->v32, 0, sizeof(float) * off));
    HIP_TRY(&c, hipMemset(m->g32, 0, sizeof(float) * off));
    HIP_TRY(&c, hipMemcpy(m->norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(off)), dim3(256), 0, c.stream, m->p32, m->p16, off);
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    return PORTAL_OK;
}

int portal_mlp_destroy(portal_ctx *x) {
    if (!x) return PORTAL_ERR_ARG;
    hipSetDevice(x->c.device);
    portal_mlp_free(x->c);
    return PORTAL_OK;
}

long long portal_mlp_num_params(const portal_ctx *x) {
    if (!x || !x->c.mlp) return -1;
    return ((const Mlp *)x->c.mlp)->n_params;
}

int portal_mlp_set_params(portal_ctx *x, const float *params, const float *norm) {
    if (!x || !x->c.mlp) return PORTAL_ERR_STATE;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    Mlp &m = *(Mlp *)c.mlp;
    if (params) {
        HIP_TRY(&c, hipMemcpy(m.p32, params, sizeof(float) * m.n_params, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(m.n_params)), dim3(256), 0, c.stream, m.p32, m.p16, m.n_params);
        HIP_TRY(&c, hipMemset(m.m32, 0, sizeof(float) * m.n_params));
        HIP_TRY(&c, hipMemset(m.v32, 0, sizeof(float) * m.n_params));
        m.step = 0;
    }
    if (norm)
        HIP_TRY(&c, hipMemcpy(m.norm, norm, sizeof(float) * (2 * m.n_in + 2 * m.n_out), hipMemcpyHostToDevice));
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    return PORTAL_OK;
}

int portal_mlp_get_params(portal_ctx *x, float *params, float *norm) {
    if (!x || !x->c.mlp) return PORTAL_ERR_STATE;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    Mlp &m = *(Mlp *)c.mlp;
    if (params) HIP_TRY(&c, hipMemcpy(params, m.p32, sizeof(float) * m.n_params, hipMemcpyDeviceToHost));
    if (norm)
        HIP_TRY(&c, hipMemcpy(norm, m.norm, sizeof(float) * (2 * m.n_in + 2 * m.n_out), hipMemcpyDeviceToHost));
    return PORTAL_OK;
}

int portal_mlp_forward(portal_ctx *x, int batch, const float *feats, float *out) {
    if (!x || !x->c.mlp || batch <= 0 || !feats || !out) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    Mlp &m = *(Mlp *)c.mlp;
    HIP_TRY(&c, mlp_reserve(m, batch));
    HIP_TRY(&c, hipMemcpyAsync(m.Xd, feats, sizeof(float) * (size_t)batch * m.n_in, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(&c, mlp_forward_dev(m, batch, c.stream, false));
    hipLaunchKernelGGL(denorm_kernel, dim3(blocks((long long)batch * m.n_out)), dim3(256), 0, c.stream,
                       m.Z[m.nl - 1], m.norm + 2 * m.n_in, m.norm + 2 * m.n_in + m.n_out, m.Yd, batch, m.n_out);
    HIP_TRY(&c, hipMemcpyAsync(out, m.Yd, sizeof(float) * (size_t)batch * m.n_out, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    return PORTAL_OK;
}

// One optimisation step on a batch (features and targets in physical units; the
// stored normalisation is applied on device).  grads (n_params, may be NULL)
// receives the gradient before the update; lr = 0 leaves the parameters unchanged.
int portal_mlp_train_step(portal_ctx *x, int batch, const float *feats, const float *targets, float lr, float *loss,
                          float *grads) {
    if (!x || !x->c.mlp || batch <= 0 || !feats || !targets) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    Mlp &m = *(Mlp *)c.mlp;
    HIP_TRY(&c, mlp_reserve(m, batch));
    HIP_TRY(&c, hipMemcpyAsync(m.Xd, feats, sizeof(float) * (size_t)batch * m.n_in, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(&c, hipMemcpyAsync(m.Td, targets, sizeof(float) * (size_t)batch * m.n_out, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(&c, mlp_forward_dev(m, batch, c.stream));
    float l = 0;
    HIP_TRY(&c, mlp_backward_dev(m, batch, c.stream, &l));
    if (grads) HIP_TRY(&c, hipMemcpy(grads, m.g32, sizeof(float) * m.n_params, hipMemcpyDeviceToHost));
    if (lr != 0.0f) HIP_TRY(&c, mlp_adam(m, lr, c.stream));
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    if (loss) *loss = l;
    return PORTAL_OK;
}

// Train over a device-resident copy of a data set: `epochs` passes of shuffled
// mini-batches.  loss_hist (epochs, may be NULL) receives each epoch's mean loss.
int portal_mlp_train(portal_ctx *x, int n, const float *feats, const float *targets, int epochs, int batch, float lr,
                     unsigned long long seed, float *loss_hist, double *ms) {
    if (!x || !x->c.mlp || n <= 0 || batch <= 0 || epochs <= 0 || !feats || !targets) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    Mlp &m = *(Mlp *)c.mlp;
    batch = std::min(batch, n);
    HIP_TRY(&c, mlp_reserve(m, batch));
    float *Xall = nullptr, *Tall = nullptr;
    HIP_TRY(&c, hipMalloc((void **)&Xall, sizeof(float) * (size_t)n * m.n_in));
    HIP_TRY(&c, hipMalloc((void **)&Tall, sizeof(float) * (size_t)n * m.n_out));
    HIP_TRY(&c, hipMemcpy(Xall, feats, sizeof(float) * (size_t)n * m.n_in, hipMemcpyHostToDevice));
    HIP_TRY(&c, hipMemcpy(Tall, targets, sizeof(float) * (size_t)n * m.n_out, hipMemcpyHostToDevice));
    std::mt19937_64 rng(seed);
    const int nb = n / batch;
    std::vector<int> order(nb);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, c.stream);
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = 0; i < nb; i++) order[i] = i;
        std::shuffle(order.begin(), order.end(), rng);
        // the loss is summed on the device and read once per epoch, so the steps queue
        // back to back with no host synchronisation
        HIP_TRY(&c, hipMemsetAsync(m.loss_acc, 0, sizeof(double), c.stream));
        const long long ny = (long long)batch * m.n_out;
        for (int bi = 0; bi < nb; bi++) {
            const size_t o = (size_t)order[bi] * batch;
            HIP_TRY(&c, hipMemcpyAsync(m.Xd, Xall + o * m.n_in, sizeof(float) * (size_t)batch * m.n_in,
                                       hipMemcpyDeviceToDevice, c.stream));
            HIP_TRY(&c, hipMemcpyAsync(m.Td, Tall + o * m.n_out, sizeof(float) * (size_t)batch * m.n_out,
                                       hipMemcpyDeviceToDevice, c.stream));
            HIP_TRY(&c, mlp_forward_dev(m, batch, c.stream));
            HIP_TRY(&c, mlp_backward_dev(m, batch, c.stream, nullptr));
            if (loss_hist)
                hipLaunchKernelGGL(loss_acc_kernel, dim3(1), dim3(256), 0, c.stream, (const float *)m.loss_part,
                                   (int)blocks(ny), ny, m.loss_acc);
            HIP_TRY(&c, mlp_adam(m, lr, c.stream));
        }
        if (loss_hist) {
            double tot = 0;
            HIP_TRY(&c, hipMemcpyAsync(&tot, m.loss_acc, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            HIP_TRY(&c, hipStreamSynchronize(c.stream));
            loss_hist[ep] = (float)(tot / std::max(1, nb));
        }
    }
    hipEventRecord(e1, c.stream);
    hipEventSynchronize(e1);
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(Xall);
    hipFree(Tall);
    return PORTAL_OK;
}

}  // extern "C"
