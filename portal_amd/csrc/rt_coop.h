// Inward ray integration, cooperative form: the waves of a workgroup share one ray
// at a time and split the (line, channel) pairs between them.  Restates ray_in
// (inverse_rt.f90:217-344) and the J accumulation of inverse_rt
// (inverse_rt.f90:126-129) with the same per-segment arithmetic as rt_body.h.
//
// Per node, rays are taken in order 0..D-1.  All threads stage ray r+1 (per cell
// v.dir/b, 1/b, segment length; per (cell, line) log(K dl) and S, and with dust
// alpha dl and B_nu(T_dust)) into one half of a
// double-buffered LDS area while the waves integrate ray r from the other half; one
// workgroup barrier per ray.  Thread t = 64 w + lane owns SLOTS consecutive channels
// of one line: line t mod R, channels (t div R) SLOTS + q, for t < R G with
// G = ceil(C / SLOTS).  A lane reads one staged (L, S) per cell for all its slots
// (the lanes of a wave read R neighbouring entries, conflict-free), every wave holds
// every line, so the waves' work per ray stays balanced, and a lane holds SLOTS
// accumulators (not one per pair of the node, as in the per-wave form), which keeps
// the register count, and with it occupancy, independent of R C.
//
// The segment update is branch-free: tau and e^-tau are evaluated for every lane.  In
// FP64 the update is selected when tau > 1e-30 (false for NaN), which is the
// reference's skip; in mixed precision the staging makes skipped segments exact
// no-ops (Arith<float>::stage, fold_seg), so no test is needed.  Without dust the FP32
// record of a (cell, line) carries the source function folded into the terms of
// 1 - e^-tau (SegF), which leaves nine VALU instructions and two v_exp_f32 per
// (cell, line, channel).
// The check for exhausted transmission (every lane's Tr exactly zero in FP64, at or
// below the floor in mixed precision) runs every fourth cell; in FP64, cells visited
// after a lane's Tr reached zero add exactly zero.  Per pair the sum over rays is in
// ray order; the channel sum with phi0 d_v is done per line in channel order.
#pragma once
#include <type_traits>

#include "rt_body.h"

namespace portal {

// Per (cell, line) record of the FP32 kernel without dust.  With tau2 = tau log2 e,
//   S (1 - e^-tau) = min(tau2 (S a0 - tau2 (S a1 - tau2 S a2)), S c - S e^-tau),
// a0..a2 the cubic's coefficients in units of tau2 (ln 2, ln^2 2 / 2, ln^3 2 / 6) and
// c = 1 + 2^-22.  The cubic is the Taylor sum to tau^3, which lies above 1 - e^-tau
// for every tau > 0 (the remainder's next term is negative) and is exact to tau^4/24;
// S c - S e^-tau lies above S (1 - e^-tau) too, because the offset 2^-22 S exceeds the
// rounding of e^-tau (1 ulp) and of S c.  The smaller of two upper bounds is within
// min(tau^4/24, 2.4e-7) of the value, relative error below 7e-6 at any tau, where the
// select at tau = 1e-2 of the cubic or 1 - e^-tau left up to 9e-6.  A skipped segment
// (K dl not > 0, or a cell whose profile argument is not finite) is staged with
// S = 0 and L = -inf (tau2 = 2^(L - u^2)) or, for the shared profile (fold_profile), the
// factor K = 0 (tau2 = K 2^(-u^2)), which makes its update an exact no-op.
struct __attribute__((aligned(8))) SegF {
    float L, S, Sc, s0, s1, s2;  // L: log2 of K dl log2 e, or K dl log2 e itself (KF)
};

template <bool KF>
__device__ __forceinline__ SegF fold_seg(float2 k, float dl, bool cell) {
    const float kdl = k.x * dl;
    const bool pos = cell && kdl > 0.0f;
    const float S = pos ? k.y : 0.0f;
    SegF g;
    if (KF)
        g.L = pos ? kdl * 1.44269504f : 0.0f;
    else
        g.L = pos ? __log2f(kdl * 1.44269504f) : -INFINITY;
    g.S = S;
    g.Sc = S * 1.00000024f;  // 1 + 2^-22
    g.s0 = S * 0.693147182f;
    g.s1 = S * 0.240226507f;
    g.s2 = S * 0.0555041087f;
    return g;
}

// I <- I e + x in I's own register.  Written as fmaf(I, e, x), the compiler's FMAC form
// ties the result to x's register, and the loop over cells then copies every slot's
// intensity back into place (six v_mov per five-slot cell).  e comes from v_exp_f32 at
// least two dependent instructions earlier (x is formed from it), so no trans-forwarding
// wait state falls inside the asm.
__device__ __forceinline__ void bt_update(float &I, float e, float x) {
    asm("v_fma_f32 %0, %0, %1, %2" : "+v"(I) : "v"(e), "v"(x));
}

// Per cell record of the cooperative kernel: v.dir/b and 1/b in the profile's units, and
// for the FP32 folded record the terms of the profile recurrence over a lane's
// consecutive channels.  With delta = (2 b_node / nvel) / b_cell the step of the profile
// argument u from one channel to the next, tau2(u + delta) = tau2(u) rho(u) with
// rho(u) = 2^(-delta (2 u + delta)), and rho(u + delta) = rho(u) cc, cc = 2^(-2 delta^2);
// nd = -delta.
template <typename Real>
struct __attribute__((aligned(16))) CoopCell {
    Real vb, inv_b, nd, cc;
};
// per staged unit, the ray's first path entry (the node's own cell): its v.dir and
// segment length
template <typename Real>
struct __attribute__((aligned(16))) CoopHdr {
    Real vel_node, dl0;
};

// How the FP32 folded update forms the profile terms tau2_q of a lane's SLOTS
// consecutive channels (u_q = u_0 + q delta), fold_profile:
//   PROF_DIRECT: tau2_q = 2^(L - u_q^2), one exponential per slot;
//   PROF_REC: one exponential and the recurrence of CoopCell (two products per further
//     slot), taken only where (SLOTS - 1) delta <= 1 for every cell of the grid
//     (coop_profile_rec).  Then a slot whose term the recurrence loses because tau2_0
//     fell below FP32's range holds less than 2^-100 for line-centre optical depths below
//     1e12 per cell, and the bound on rho's exponent acts only where
//     |u_0| > 32 / delta >= 32, so that every slot's term is zero in FP32 anyway;
//   PROF_SHARE: the profile 2^(-u^2) depends on the channel, not the line, and the lanes
//     of a wave that hold one channel group differ only in their line.  Each lane
//     evaluates 2^(-u^2) for one channel of its wave's window (the window's channel
//     cw0 + lane) into the wave's 64-entry LDS row, then reads its slots' values and
//     multiplies them by its line's K dl log2 e: one exponential per lane and cell
//     instead of one per slot, and the lanes of the same channel see the same value.  It
//     needs every wave's window to fit 64 channels (coop_profile_share).  A term lost
//     because 2^(-u^2) fell below FP32's range is below 2^-86 for K dl log2 e < 2^40.
constexpr int PROF_DIRECT = 0, PROF_REC = 1, PROF_SHARE = 2;

template <int SLOTS, int PROF>
__device__ __forceinline__ void fold_profile(const float (&v0s)[SLOTS], const CoopCell<float> &c, float L,
                                             float (&t2)[SLOTS], float vg, float *gw, int lane, int go) {
    if constexpr (PROF == PROF_SHARE) {
        const float ug = fmaf(vg, c.inv_b, -c.vb);
        gw[lane] = __builtin_amdgcn_exp2f(-(ug * ug));
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int q = 0; q < SLOTS; q++) t2[q] = L * gw[go + q];
        __builtin_amdgcn_wave_barrier();
    } else if constexpr (PROF == PROF_REC && SLOTS > 1) {
        const float u0 = fmaf(v0s[0], c.inv_b, -c.vb);
        t2[0] = __builtin_amdgcn_exp2f(fmaf(-u0, u0, L));
        float rho = __builtin_amdgcn_exp2f(fminf(c.nd * fmaf(2.0f, u0, -c.nd), 64.0f));
#pragma unroll
        for (int q = 1; q < SLOTS; q++) {
            t2[q] = t2[q - 1] * rho;
            if (q + 1 < SLOTS) rho = rho * c.cc;
        }
    } else {
#pragma unroll
        for (int q = 0; q < SLOTS; q++) {
            const float uu = fmaf(v0s[q], c.inv_b, -c.vb);
            t2[q] = __builtin_amdgcn_exp2f(fmaf(-uu, uu, L));
        }
    }
}

// the staged record of a (cell, line): SegF for FP32 without dust, else (L, S) of Arith
template <typename Real, bool DUST>
struct CoopSeg {
    typedef typename Arith<Real>::R2 T;
    static constexpr bool fold = false;
};
template <>
struct CoopSeg<float, false> {
    typedef SegF T;
    static constexpr bool fold = true;
};

template <typename Real>
__host__ __device__ constexpr size_t coop_seg_bytes(bool dust) {
    return dust || !std::is_same<Real, float>::value ? sizeof(typename Arith<Real>::R2) : sizeof(SegF);
}

// one staging buffer: the unit's header, cap cell records, cap x R per-line records, and
// with dust cap x R (alpha dl, B_nu(T_dust))
template <typename Real>
__host__ __device__ constexpr size_t coop_buf_bytes(int cap, int R, bool dust) {
    return (sizeof(CoopHdr<Real>) +
            (size_t)cap * (sizeof(CoopCell<Real>) + (size_t)R * coop_seg_bytes<Real>(dust) +
                           (dust ? (size_t)R * sizeof(typename Arith<Real>::R2) : 0)) +
            15) &
           ~(size_t)15;
}

// the shared profile's LDS rows (fold_profile): 64 floats per wave, for the FP32 folded
// record only
template <typename Real>
__host__ __device__ constexpr size_t coop_share_bytes(bool dust, int waves) {
    return CoopSeg<Real, false>::fold && !dust ? (size_t)waves * 64 * sizeof(float) : 0;
}

// Cells per staged unit: COOP_UNIT, or fewer where the two staging buffers of a
// workgroup (and the shared profile's rows) would not leave room for 7 workgroups per CU
// (the FP32 kernel's occupancy).  LDS is 160 KiB per CU, allocated to a workgroup in
// blocks of 512 bytes, so 7 workgroups have 45 blocks each: at 46 the CU holds 6.  Only
// the folded FP32 record (24 bytes per line) reaches that bound at the configurations'
// line counts (47 cells at 9 lines and 4 waves).
template <typename Real>
__host__ __device__ constexpr int coop_unit_cells(int R, bool dust, int waves) {
    const size_t per_wg =
        (160 * 1024) / 7 / 512 * 512 - 2 * sizeof(CoopHdr<Real>) - coop_share_bytes<Real>(dust, waves);
    const size_t per_cell = 2 * (sizeof(CoopCell<Real>) + (size_t)R * coop_seg_bytes<Real>(dust) +
                                 (dust ? (size_t)R * sizeof(typename Arith<Real>::R2) : 0));
    const int fit = (int)(per_wg / per_cell);
    return fit >= COOP_UNIT ? COOP_UNIT : (fit < 8 ? 8 : fit);
}

// the two staging buffers and the shared profile's rows, or the reduction of J0, J2
// (and the own-cell operator) per (line, channel), whichever is larger
template <typename Real>
size_t coop_lds_bytes(int cap, int R, int npairs, bool dust, bool lam, int waves) {
    const size_t stage = 2 * coop_buf_bytes<Real>(cap, R, dust) + coop_share_bytes<Real>(dust, waves);
    const size_t red = sizeof(double) * (lam ? 3 : 2) * (size_t)npairs;
    return stage > red ? stage : red;
}

// All threads of the workgroup stage cells c0 .. c0+cnt-1 of ray r of work item bid
// into buf (cnt <= cap); dv is the channel spacing 2 b_node / nvel.
template <typename Real, bool DUST, bool KF>
__device__ __forceinline__ void coop_stage(const RtArgs &a, const int bid, const int r, const int c0, const int cnt,
                                           const int cap, const Real dv, const int4 pos, char *buf) {
    typedef Arith<Real> A;
    typedef typename A::R2 R2;
    typedef typename CoopSeg<Real, DUST>::T Seg;
    const int R = a.R;
    const long long off = a.offsets[(long long)bid * a.D + r];
    const PathEntry *P = a.paths + off + c0;
    CoopHdr<Real> *hdr = reinterpret_cast<CoopHdr<Real> *>(buf);
    CoopCell<Real> *cr = reinterpret_cast<CoopCell<Real> *>(hdr + 1);
    Seg *ks = reinterpret_cast<Seg *>(cr + cap);
    R2 *kds = reinterpret_cast<R2 *>(ks + (size_t)cap * R);
    const Real dls = (Real)a.dl_scale;
    const int n = cnt * R;
    if (threadIdx.x == 0) {
        const PathEntry e0 = a.paths[off];
        CoopHdr<Real> h;
        h.vel_node = (Real)e0.vray;
        h.dl0 = (Real)e0.len * dls;
        *hdr = h;
    }
    // item j = i R + l: the thread's first item is (pos.x, pos.y), and a step of the block
    // size moves it by (pos.z, pos.w) with a carry, so no item divides by R
    int i = pos.x, l = pos.y;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const PathEntry e = P[i];
        const double ib = a.inv_b[e.cell];
        const Real dl = (Real)e.len * dls;
        const Real vb = (Real)(e.vray * ib) * (Real)A::vscale, sib = (Real)ib * (Real)A::vscale;
        const bool ok = A::cell_ok(vb, sib);
        if constexpr (CoopSeg<Real, DUST>::fold)
            ks[j] = fold_seg<KF>(A::ks(a, (size_t)e.cell * R + l), dl, ok);
        else
            ks[j] = A::stage(A::ks(a, (size_t)e.cell * R + l), dl, ok);
        if (DUST) kds[j] = A::stage_dust(A::ksd(a, (size_t)e.cell * R + l), dl, ok);
        if (l == 0) {
            CoopCell<Real> c;
            c.vb = ok ? vb : (Real)0;
            c.inv_b = ok ? sib : (Real)0;
            c.nd = (Real)0;
            c.cc = (Real)1;
            if constexpr (CoopSeg<Real, DUST>::fold) {
                const float nd = -(dv * c.inv_b);
                c.nd = nd;
                c.cc = __builtin_amdgcn_exp2f(-2.0f * (nd * nd));
            }
            cr[i] = c;
        }
        i += pos.z;
        l += pos.w;
        if (l >= R) {
            l -= R;
            i++;
        }
    }
}

// nonzero exactly when x is not +0 (transmissions are never negative)
__device__ __forceinline__ unsigned nonzero_bits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ unsigned nonzero_bits(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (unsigned)(b >> 32) | (unsigned)b;
}

// BTF: the cells of a ray are taken back to front, from the far end to the node, as the
// reference's ray_in does (I <- I e^-tau + S (1 - e^-tau)); the units of a ray are then
// staged last to first.  The update needs one product fewer than front to back (no
// transmission to carry) but cannot leave an exhausted ray early.
// PROF: how the FP32 folded update forms its profile terms (fold_profile).
template <typename Real, int SLOTS, bool DUST, bool LAM, bool BTF = false, int PROF = PROF_DIRECT>
__device__ __forceinline__ void rt_coop_node(const RtArgs &a, const int bid, const int cap, char *smem) {
    typedef Arith<Real> A;
    typedef typename A::R2 R2;
    typedef typename CoopSeg<Real, DUST>::T Seg;
    constexpr bool FOLD = CoopSeg<Real, DUST>::fold;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = blockDim.x >> 6;
    const int node = a.work[bid];
    const double bd = a.velb[node].w;
    const Real b = (Real)bd;
    const Real inv_bn = (Real)(1.0 / bd);
    const int C = a.C, nvel = a.nvel, R = a.R, D = a.D;
    const Real dv = ((Real)2 * b) / (Real)nvel;  // channel spacing
    const Real da = (Real)2 / (Real)nvel;        // its profile argument at the node
    // channel nv's velocity offset 2 b nv / nvel and profile argument v / b at the node:
    // FP64 in the operation order of the other inward kernels (which give the same bits),
    // FP32 as products with the spacing (no division per channel and ray)
    auto chan_v = [&](int nv) -> Real {
        if constexpr (std::is_same<Real, float>::value)
            return (Real)nv * dv;
        else
            return ((Real)2 * b) * (Real)nv / (Real)nvel;
    };
    auto chan_a = [&](int nv) -> Real {
        if constexpr (std::is_same<Real, float>::value)
            return (Real)nv * da;
        else
            return (((Real)2 * b) * (Real)nv / (Real)nvel) * inv_bn;
    };
    // this thread's first staging item (cell, line) and the step of one block
    const int4 spos = make_int4((int)threadIdx.x / a.R, (int)threadIdx.x % a.R, (int)blockDim.x / a.R,
                                (int)blockDim.x % a.R);
    const int npairs = R * C;
    const size_t bb = coop_buf_bytes<Real>(cap, R, DUST);

    const int G = (C + SLOTS - 1) / SLOTS;  // lanes per line
    const int t = wave * 64 + lane;
    const bool lane_on = t < G * R;
    const int lo = lane_on ? t % R : 0;      // this lane's line
    const int cb = (t / R) * SLOTS;          // its first channel
    // the shared profile (PROF_SHARE): the wave's window starts at channel cw0, lane l
    // evaluates channel cw0 + l into the wave's row gw, and this lane's slots are the
    // row's entries go .. go + SLOTS - 1
    const int cw0 = ((64 * wave) / R) * SLOTS;
    const int go = lane_on ? cb - cw0 : 0;
    float *gw = reinterpret_cast<float *>(smem + 2 * coop_buf_bytes<Real>(cap, R, DUST)) + 64 * wave;
    Real vg = (Real)0;  // the velocity of channel cw0 + lane, relative as v0s
    // slot q of this lane holds channel cb + q when on(q); the channel's velocity offset
    // and profile weight at the node are recomputed per ray, which keeps registers free
    auto on = [&](int q) { return lane_on && cb + q < C; };
    Real S0[SLOTS], C0[SLOTS], S2[SLOTS], C2[SLOTS], Lw[LAM ? SLOTS : 1];
    const R2 k0 = A::ks(a, (size_t)node * R + lo);  // the node's own cell, no skip
    R2 d0;
    if (DUST) d0 = A::ksd(a, (size_t)node * R + lo);
    // backgrounds in the line's intensity unit, the unit of the staged S
    const int ue = A::uexp(a.I_in[lo]);
    const Real iin_bg = (Real)ldexp(a.I_in[lo], -ue);
    const Real iin_star = a.xthn ? (Real)ldexp(a.I_star[lo], -ue) : iin_bg;
    const int n_star = a.xthn ? a.nstar[bid] : 0;
#pragma unroll
    for (int q = 0; q < SLOTS; q++) S0[q] = C0[q] = S2[q] = C2[q] = (Real)0;
    if (LAM)
#pragma unroll
        for (int q = 0; q < SLOTS; q++) Lw[q] = (Real)0;

    // Rays are staged in units of at most cap cells, so that the staging area does not
    // grow with the longest path; unit u of ray r holds its cells u cap .. u cap + cap - 1.
    auto ray_cells = [&](int r) {
        const long long o = (long long)bid * D + r;
        return (int)(a.offsets[o + 1] - a.offsets[o]);
    };
    auto stage_unit = [&](int r, int u, char *buf) {
        const int nc = ray_cells(r);
        const int c0 = u * cap;
        coop_stage<Real, DUST, PROF == PROF_SHARE>(a, bid, r, c0, min(cap, nc - c0), cap, dv, spos, buf);
    };
    // per-ray state, carried across the units of a ray.  Front to back: Em the emission
    // reaching the node so far and Tr the transmission to the current cell.  Back to
    // front: Em the intensity entering the current cell (Tr unused).
    Real Tr[SLOTS], Em[SLOTS], v0s[SLOTS];
    Real w = (Real)0, wc = (Real)0, iin = (Real)0;
    bool ray_done = false;
    auto units = [&](int r) { return max(1, (ray_cells(r) + cap - 1) / cap); };
    stage_unit(0, BTF ? units(0) - 1 : 0, smem);
    __syncthreads();
    int r = 0, u = BTF ? units(0) - 1 : 0, half = 0;
    for (;;) {
        const char *cur = smem + half * bb;
        const int nc = ray_cells(r);
        const int nu = max(1, (nc + cap - 1) / cap);
        int rn = r, un;
        if constexpr (BTF) {
            un = u - 1;
            if (un < 0) {
                rn = r + 1;
                un = rn < D ? units(rn) - 1 : 0;
            }
        } else {
            un = u + 1;
            if (un >= nu) {
                rn = r + 1;
                un = 0;
            }
        }
        if (rn < D) stage_unit(rn, un, smem + (half ^ 1) * bb);
        const CoopHdr<Real> *hdr = reinterpret_cast<const CoopHdr<Real> *>(cur);
        const CoopCell<Real> *cr = reinterpret_cast<const CoopCell<Real> *>(hdr + 1);
        const Seg *ks = reinterpret_cast<const Seg *>(cr + cap);
        const R2 *kds = reinterpret_cast<const R2 *>(ks + (size_t)cap * R);
        const int c0 = u * cap, cnt = min(cap, nc - c0);
        if (BTF && u == nu - 1) {
            // the ray's first unit back to front: its direction, background and channels
            const double *xr = a.xthn ? a.xthn + 3 * ((size_t)bid * D + r) : a.xth + 3 * r;
            const Real mu = (Real)xr[0];
            w = (Real)xr[2];
            iin = r < n_star ? iin_star : iin_bg;
            wc = w * (((Real)3 * (mu * mu) - (Real)1) / (Real)2);
            const Real vel_node = hdr->vel_node;
#pragma unroll
            for (int q = 0; q < SLOTS; q++) {
                const int nv = on(q) ? cb + q - nvel : 0;
                v0s[q] = chan_v(nv) + vel_node;
                Em[q] = iin;
            }
            if (PROF == PROF_SHARE) vg = chan_v(cw0 + lane - nvel) + vel_node;
        }
        if (!BTF && u == 0) {
            // star mode: the node's own directions; rays towards the star start with B_nu(T_star)
            const double *xr = a.xthn ? a.xthn + 3 * ((size_t)bid * D + r) : a.xth + 3 * r;
            const Real mu = (Real)xr[0];
            w = (Real)xr[2];
            iin = r < n_star ? iin_star : iin_bg;
            wc = w * (((Real)3 * (mu * mu) - (Real)1) / (Real)2);
            const Real vel_node = hdr->vel_node;
            const Real dl0 = hdr->dl0;
            ray_done = false;
            if (PROF == PROF_SHARE) vg = chan_v(cw0 + lane - nvel) + vel_node;
#pragma unroll
            for (int q = 0; q < SLOTS; q++) {
                const int nv = on(q) ? cb + q - nvel : 0;
                const Real v0 = chan_v(nv);
                const Real aa = chan_a(nv);
                const Real e0 = A::ex(-(aa * aa));
                if (DUST) {
                    // S = (e phi + j)/(k phi + alpha), tau = (k phi + alpha) dl (inverse_rt.f90:289-297)
                    const Real kp = k0.x * e0;
                    const Real tau0 = (kp + d0.x) * dl0;
                    const Real Sc = (k0.y * kp + d0.y * d0.x) / (kp + d0.x);
                    Tr[q] = A::ex(-tau0);
                    const Real om0 = A::one_minus(tau0, Tr[q]);
                    Em[q] = Sc * om0;
                    if (LAM) Lw[q] = Lw[q] + w * om0;
                } else {
                    const Real tau0 = k0.x * e0 * dl0;
                    Tr[q] = A::ex(-tau0);
                    const Real om0 = A::one_minus(tau0, Tr[q]);
                    Em[q] = k0.y * om0;
                    // the own cell's share of J per unit of its source function
                    if (LAM) Lw[q] = Lw[q] + w * om0;
                }
                v0s[q] = v0 + vel_node;
                if (!on(q)) Tr[q] = (Real)0;
            }
        }
        if constexpr (BTF) {
            // cells c0 + cnt - 1 down to max(c0, 1) (local index j = i - c0); the staged
            // data are read at the top of each cell (reading a cell ahead made the compiler
            // copy the read-ahead registers every cell; the other resident waves hide the
            // LDS latency); the bounds are uniform, which keeps the loop test scalar
            const int jb = __builtin_amdgcn_readfirstlane(c0 == 0 ? 1 : 0);
            const int jt = __builtin_amdgcn_readfirstlane(cnt - 1);
            const char *kl = reinterpret_cast<const char *>(ks + lo);
            const int kstride = R * (int)sizeof(Seg);
            for (int j = jt; j >= jb; j--) {
                const CoopCell<Real> c = cr[j];
                const Seg k = *reinterpret_cast<const Seg *>(kl + j * kstride);
                R2 d;
                if (DUST) d = kds[(size_t)j * R + lo];
                if constexpr (FOLD) {
                    // every slot's profile term, then its attenuation, then the updates:
                    // independent exponentials in a row rather than a dependent pair per slot
                    float t2v[SLOTS], etv[SLOTS];
                    fold_profile<SLOTS, PROF>(v0s, c, k.L, t2v, vg, gw, lane, go);
#pragma unroll
                    for (int q = 0; q < SLOTS; q++) etv[q] = __builtin_amdgcn_exp2f(-t2v[q]);
#pragma unroll
                    for (int q = 0; q < SLOTS; q++) {
                        const float tau2 = t2v[q], et = etv[q];
                        const float cub = tau2 * fmaf(-tau2, fmaf(-tau2, k.s2, k.s1), k.s0);
                        const float dir = fmaf(-k.S, et, k.Sc);
                        bt_update(Em[q], et, __builtin_fminf(cub, dir));
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < SLOTS; q++) {
                        const Real uu = fma(v0s[q], c.inv_b, -c.vb);
                        Real tau = A::tau_of(k.x, uu), S = k.y;
                        if (DUST) {
                            const Real tl = tau;
                            tau = tl + d.x;
                            S = (k.y * tl + d.y * d.x) * A::rcp(tau);
                            if (!A::skip_test && !(tau > (Real)0)) S = (Real)0;
                        }
                        const Real et = A::att(tau);
                        // the reference's Iv * e_tau + Sv * (1 - e_tau) (inverse_rt.f90:314)
                        const Real iv = Em[q] * et + S * A::om(tau, et);
                        if constexpr (A::skip_test)
                            Em[q] = tau > A::thr() ? iv : Em[q];
                        else
                            Em[q] = iv;
                    }
                }
            }
        } else if (!ray_done) {
            // cells max(c0, 1) .. c0 + cnt - 1 of the ray (local index j = i - c0), their
            // staged data read at the top of each cell; uniform bounds keep the test scalar
            const int jb = __builtin_amdgcn_readfirstlane(c0 == 0 ? 1 : 0);
            const int je = __builtin_amdgcn_readfirstlane(cnt);
            // this lane's record of cell j is at byte j R sizeof(Seg) from kl: a uniform
            // offset on one address register
            const char *kl = reinterpret_cast<const char *>(ks + lo);
            const int kstride = R * (int)sizeof(Seg);
            // a lane with one channel reads a cell ahead: it has too little work per cell to
            // hide the LDS latency otherwise (with more slots the read-ahead registers cost
            // copies every cell)
            constexpr bool AHEAD = SLOTS == 1;
            CoopCell<Real> cn{};
            Seg kn{};
            R2 dn{};
            if constexpr (AHEAD) {
                const int j0 = jb < je ? jb : 0;
                cn = cr[j0];
                kn = *reinterpret_cast<const Seg *>(kl + j0 * kstride);
                if (DUST) dn = kds[(size_t)j0 * R + lo];
            }
            for (int j = jb; j < je; j++) {
                if (((c0 + j) & 3) == 0) {
                    // mixed: the wave leaves the ray once every channel's transmission is at
                    // or below the floor (the dropped tail is below the floor times the
                    // remaining source function; a channel under the floor still adds its
                    // share until then); FP64: transmission is never negative, so only
                    // exact zeros stop
                    bool live;
                    if constexpr (A::tr_floor > 0) {
                        Real m = Tr[0];
#pragma unroll
                        for (int q = 1; q < SLOTS; q++) m = fmaxf(m, Tr[q]);
                        live = m > (Real)A::tr_floor;
                    } else {
                        unsigned bits = 0;
#pragma unroll
                        for (int q = 0; q < SLOTS; q++) bits |= nonzero_bits(Tr[q]);
                        live = bits != 0u;
                    }
                    if (!__any(live)) {
                        ray_done = true;
                        break;
                    }
                }
                CoopCell<Real> c;
                Seg k;
                R2 d{};
                if constexpr (AHEAD) {
                    c = cn;
                    k = kn;
                    d = dn;
                    const int jn = j + 1 < je ? j + 1 : j;
                    cn = cr[jn];
                    kn = *reinterpret_cast<const Seg *>(kl + jn * kstride);
                    if (DUST) dn = kds[(size_t)jn * R + lo];
                } else {
                    c = cr[j];
                    k = *reinterpret_cast<const Seg *>(kl + j * kstride);
                    if (DUST) d = kds[(size_t)j * R + lo];
                }
                if constexpr (FOLD) {
                    float t2v[SLOTS];
                    fold_profile<SLOTS, PROF>(v0s, c, k.L, t2v, vg, gw, lane, go);
#pragma unroll
                    for (int q = 0; q < SLOTS; q++) {
                        const float tau2 = t2v[q];
                        const float et = __builtin_amdgcn_exp2f(-tau2);
                        const float cub = tau2 * fmaf(-tau2, fmaf(-tau2, k.s2, k.s1), k.s0);
                        const float dir = fmaf(-k.S, et, k.Sc);
                        Em[q] = fmaf(Tr[q], __builtin_fminf(cub, dir), Em[q]);
                        Tr[q] = Tr[q] * et;
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < SLOTS; q++) {
                        const Real uu = fma(v0s[q], c.inv_b, -c.vb);
                        Real tau = A::tau_of(k.x, uu), S = k.y;
                        if (DUST) {
                            // tau = tau_line + alpha dl; S the opacity-weighted mean of the line's
                            // and the dust's source functions (inverse_rt.f90:273-287)
                            const Real tl = tau;
                            tau = tl + d.x;
                            S = (k.y * tl + d.y * d.x) * A::rcp(tau);
                            if (!A::skip_test && !(tau > (Real)0)) S = (Real)0;
                        }
                        const Real et = A::att(tau);
                        const Real em = Em[q] + Tr[q] * S * A::om(tau, et);
                        const Real tr = Tr[q] * et;
                        if constexpr (A::skip_test) {
                            const bool ok = tau > A::thr();
                            Em[q] = ok ? em : Em[q];
                            Tr[q] = ok ? tr : Tr[q];
                        } else {
                            Em[q] = em;
                            Tr[q] = tr;
                        }
                    }
                }
            }
        }
        if (BTF ? u == 0 : u == nu - 1) {
            // back to front: the intensity leaving the node's own cell, kept apart from Em
            // (which the next ray sets afresh), so that no register copy carries it back
            Real Io[SLOTS];
            if constexpr (BTF) {
                // the node's own cell last, with no skip (inverse_rt.f90:318-337)
                const Real dl0 = hdr->dl0;
#pragma unroll
                for (int q = 0; q < SLOTS; q++) {
                    const int nv = on(q) ? cb + q - nvel : 0;
                    const Real aa = chan_a(nv);
                    const Real e0 = A::ex(-(aa * aa));
                    Real tau0, Sc;
                    if (DUST) {
                        const Real kp = k0.x * e0;
                        tau0 = (kp + d0.x) * dl0;
                        Sc = (k0.y * kp + d0.y * d0.x) / (kp + d0.x);
                    } else {
                        tau0 = k0.x * e0 * dl0;
                        Sc = k0.y;
                    }
                    const Real t0 = A::ex(-tau0);
                    const Real om0 = A::one_minus(tau0, t0);
                    Io[q] = Em[q] * t0 + Sc * om0;
                    if (LAM) Lw[q] = Lw[q] + w * om0;
                }
            }
#pragma unroll
            for (int q = 0; q < SLOTS; q++) {
                // front to back the background arrives through the ray's transmission
                const Real Iout = BTF ? Io[q] : Em[q] + Tr[q] * iin;
                if (A::kahan) {
                    Real y = w * Iout - C0[q];
                    Real t2 = S0[q] + y;
                    C0[q] = (t2 - S0[q]) - y;
                    S0[q] = t2;
                    y = wc * Iout - C2[q];
                    t2 = S2[q] + y;
                    C2[q] = (t2 - S2[q]) - y;
                    S2[q] = t2;
                } else {
                    S0[q] = S0[q] + w * Iout;
                    S2[q] = S2[q] + wc * Iout;
                }
            }
        }
        __syncthreads();  // the next unit staged; this unit's buffer free for the one after
        if (rn >= D) break;
        r = rn;
        u = un;
        half ^= 1;
    }

    // every lane weights its own channels' sums with phi0 d_v (inverse_rt.f90:258,340);
    // one thread per line then adds them in channel order
    double *red0 = reinterpret_cast<double *>(smem);
    double *red2 = red0 + npairs;
    double *red3 = red2 + npairs;
    {
        const double spi = 1.7724538509055160273;
        const double d_v = 2.0 * bd / nvel;
#pragma unroll
        for (int q = 0; q < SLOTS; q++) {
            const int p = lo * C + cb + q;
            if (on(q)) {
                const double vv = (2.0 * bd) * (cb + q - nvel) / nvel;
                const double aa = vv / bd;
                const double wgt = (1.0 / (spi * bd)) * exp(-(aa * aa)) * d_v;
                red0[p] = wgt * ((double)S0[q] - (double)C0[q]);
                red2[p] = wgt * ((double)S2[q] - (double)C2[q]);
                if (LAM) red3[p] = wgt * (double)Lw[q];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int l = threadIdx.x;
        double j0 = 0.0, j2 = 0.0, lam = 0.0;
        for (int c = 0; c < C; c++) {
            j0 = j0 + red0[l * C + c];
            j2 = j2 + red2[l * C + c];
            if (LAM) lam = lam + red3[l * C + c];
        }
        // back from the intensity unit (exact); the own-cell operator is per unit of S
        const int ul = A::uexp(a.I_in[l]);
        a.J0[(size_t)node * R + l] = ldexp(j0, ul);
        a.J2[(size_t)node * R + l] = ldexp(j2 / 1.4142135623730951, ul);
        if (LAM) a.LAM[(size_t)node * R + l] = lam;
    }
}

template <typename Real, int SLOTS, bool DUST, bool LAM, bool BTF, int PROF>
__global__ void __launch_bounds__(512) rt_coop_kernel(RtArgs a, int cap) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    rt_coop_node<Real, SLOTS, DUST, LAM, BTF, PROF>(a, rt_item(a, xcd_block(blockIdx.x, gridDim.x)), cap, smem);
}

// The FP32 forms with up to 6 slots are compiled for 7 waves per SIMD (at most 72
// VGPRs).  The FP32 update is bound partly by the latency of its dependent exponentials,
// which more resident waves hide: 723 ms per production pass against 781 ms at the 6
// waves the kernel's natural 80 VGPRs give (one box, same inputs; 12 dwords spill).
// With more slots, or in the persistent form, the budget would spill far more.
template <int SLOTS, bool DUST, bool LAM, bool BTF, int PROF>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(7, 8)))
rt_coop_kernel_f32(RtArgs a, int cap) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    rt_coop_node<float, SLOTS, DUST, LAM, BTF, PROF>(a, rt_item(a, xcd_block(blockIdx.x, gridDim.x)), cap, smem);
}

// The profile recurrence (fold_profile) is taken where (SLOTS - 1) delta <= 1 for every
// cell: delta = (2 b_node / nvel) sqrt(log2 e) / b_cell <= 2 b_ratio sqrt(log2 e) / nvel,
// b_ratio the grid's largest over its smallest positive Doppler width (0: unknown).
inline bool coop_profile_rec(double b_ratio, int nvel, int slots) {
    return slots > 1 && b_ratio > 0.0 && nvel > 0 &&
           (slots - 1) * 2.0 * b_ratio * 1.2011224087864498 / nvel <= 1.0;
}

// The shared profile needs every wave's window, the channels of the groups its lanes
// hold (lane t = 64 w + l holds group t / R), to fit the wave's 64 lanes.
inline bool coop_profile_share(int R, int C, int W, int slots) {
    if (slots < 2 || R < 2) return false;
    const int G = (C + slots - 1) / slots, n_on = G * R;
    for (int w = 0; w < W; w++) {
        const int t0 = 64 * w, t1 = std::min(64 * w + 63, n_on - 1);
        if (t0 > t1) break;
        if ((t1 / R - t0 / R + 1) * slots > 64) return false;
    }
    return true;
}

// the profile form of a launch: shared where it fits, else the recurrence where it is
// exact enough, else one exponential per slot
// (with fewer than three slots neither saves an exponential's worth of work: the shared
// row's LDS round trip cost a production-sized pass of the 4-level cloud 11% at two slots)
inline int coop_profile_form(int R, int C, int W, int slots, double b_ratio, int nvel) {
    if (slots < 3) return PROF_DIRECT;
    if (coop_profile_share(R, C, W, slots)) return PROF_SHARE;
    if (coop_profile_rec(b_ratio, nvel, slots)) return PROF_REC;
    return PROF_DIRECT;
}

// variant 3 of RtArgs takes the cells of a ray back to front
template <typename Real, int SLOTS, bool DUST, bool LAM>
void launch_coop_k(dim3 g, dim3 b, size_t lds, hipStream_t s, const RtArgs &a, int cap) {
    auto go = [&](auto btf_c, auto prof_c) {
        constexpr bool BTF = decltype(btf_c)::value;
        constexpr int PROF = CoopSeg<Real, DUST>::fold && SLOTS > 1 ? decltype(prof_c)::value : PROF_DIRECT;
        if constexpr (std::is_same<Real, float>::value && SLOTS <= 6)
            hipLaunchKernelGGL((rt_coop_kernel_f32<SLOTS, DUST, LAM, BTF, PROF>), g, b, lds, s, a, cap);
        else
            hipLaunchKernelGGL((rt_coop_kernel<Real, SLOTS, DUST, LAM, BTF, PROF>), g, b, lds, s, a, cap);
    };
    auto by_prof = [&](auto btf_c) {
        switch (coop_profile_form(a.R, a.C, a.waves, SLOTS, a.b_ratio, a.nvel)) {
            case PROF_SHARE: go(btf_c, std::integral_constant<int, PROF_SHARE>{}); break;
            case PROF_REC: go(btf_c, std::integral_constant<int, PROF_REC>{}); break;
            default: go(btf_c, std::integral_constant<int, PROF_DIRECT>{}); break;
        }
    };
    if (a.variant == 3)
        by_prof(std::true_type{});
    else
        by_prof(std::false_type{});
}

template <typename Real, int SLOTS>
hipError_t launch_coop_s(const RtArgs &a, int cap, size_t lds, hipStream_t s) {
    const dim3 g(a.n_work), b(64 * a.waves);
    const bool lam = a.LAM != nullptr;
    if (a.dust && lam) return hipErrorInvalidValue;  // the own-cell operator covers the line alone
    if (a.dust)
        launch_coop_k<Real, SLOTS, true, false>(g, b, lds, s, a, cap);
    else if (lam)
        launch_coop_k<Real, SLOTS, false, true>(g, b, lds, s, a, cap);
    else
        launch_coop_k<Real, SLOTS, false, false>(g, b, lds, s, a, cap);
    return hipGetLastError();
}

// Persistent form over the (model, work item) items of a batch of models whose lines
// and channels need the same slot count: workgroups take items from a counter until the
// list is exhausted (every workgroup reaches the exit test after each item).  DUST: the
// class's models carry dust (every model of a class agrees).
template <typename Real, int SLOTS, bool DUST, bool BTF, int PROF>
__device__ __forceinline__ void coop_persistent_body(const RtArgs *__restrict__ models,
                                                     const int2 *__restrict__ items, int n_items, int *counter,
                                                     int cap) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        // the item is uniform over the workgroup: scalar registers hold it and its model
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const int2 mi = items[it];
        const int m = __builtin_amdgcn_readfirstlane(mi.x), w = __builtin_amdgcn_readfirstlane(mi.y);
        const RtArgs am = models[m];
        rt_coop_node<Real, SLOTS, DUST, false, BTF, PROF>(am, w, cap, smem);
        __syncthreads();
    }
}

template <typename Real, int SLOTS, bool DUST, bool BTF, int PROF>
__global__ void __launch_bounds__(512) rt_coop_persistent_kernel(const RtArgs *__restrict__ models,
                                                                 const int2 *__restrict__ items, int n_items,
                                                                 int *counter, int cap) {
    coop_persistent_body<Real, SLOTS, DUST, BTF, PROF>(models, items, n_items, counter, cap);
}

// Slots per lane: the fewest consecutive channels per lane with which every line's
// channels fit in its share of the 64 W lanes, ceil(C / floor(64 W / R)), rounded up to
// an instantiated count (0: more pairs than one workgroup holds).
inline int coop_slots(int R, int C, int W) {
    const int per_line = (64 * W) / R;
    if (per_line < 1) return 0;
    const int need = (C + per_line - 1) / per_line;
    if (need <= 6) return need;
    if (need <= 8) return 8;
    if (need <= 10) return 10;
    if (need <= 12) return 12;
    return 0;
}

// Flat form over the same items: one workgroup per item, in the items' (cost) order,
// the hardware's dispatch doing the balancing.  It needs no loop state, so it keeps the
// register budget of rt_coop_kernel.
template <typename Real, int SLOTS, bool DUST, bool BTF, int PROF>
__device__ __forceinline__ void coop_flat_body(const RtArgs *__restrict__ models, const int2 *__restrict__ items,
                                               int cap) {
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int2 mi = items[blockIdx.x];
    const int m = __builtin_amdgcn_readfirstlane(mi.x), w = __builtin_amdgcn_readfirstlane(mi.y);
    rt_coop_node<Real, SLOTS, DUST, false, BTF, PROF>(models[m], w, cap, smem);
}

template <typename Real, int SLOTS, bool DUST, bool BTF, int PROF>
__global__ void __launch_bounds__(512) rt_coop_flat_kernel(const RtArgs *__restrict__ models,
                                                           const int2 *__restrict__ items, int cap) {
    coop_flat_body<Real, SLOTS, DUST, BTF, PROF>(models, items, cap);
}

template <int SLOTS, bool DUST, bool BTF, int PROF>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(7, 8)))
rt_coop_flat_kernel_f32(const RtArgs *__restrict__ models, const int2 *__restrict__ items, int cap) {
    coop_flat_body<float, SLOTS, DUST, BTF, PROF>(models, items, cap);
}

// n_blocks > 0: the persistent form on n_blocks workgroups; n_blocks == 0: the flat form.
// btf: rays back to front; prof: the profile form (fold_profile), one that every model of
// the launch allows.
template <typename Real, bool DUST>
hipError_t launch_coop_persistent_d(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap,
                                    int slots, int waves, size_t lds, int n_blocks, bool btf, int prof,
                                    hipStream_t s) {
    if (lds > 160 * 1024 || waves < 1 || waves > 8) return hipErrorInvalidValue;
    if (n_items <= 0) return hipSuccess;
    auto launch = [&](auto slots_c, auto btf_c, auto prof_c) {
        constexpr int S = decltype(slots_c)::value;
        constexpr bool BTF = decltype(btf_c)::value;
        constexpr int PROF = CoopSeg<Real, DUST>::fold && S > 1 ? decltype(prof_c)::value : PROF_DIRECT;
        if (n_blocks == 0) {
            const dim3 g(n_items), b(64 * waves);
            if constexpr (std::is_same<Real, float>::value && S <= 6)
                hipLaunchKernelGGL((rt_coop_flat_kernel_f32<S, DUST, BTF, PROF>), g, b, lds, s, models, items, cap);
            else
                hipLaunchKernelGGL((rt_coop_flat_kernel<Real, S, DUST, BTF, PROF>), g, b, lds, s, models, items,
                                   cap);
        } else {
            const dim3 g(n_blocks), b(64 * waves);
            hipLaunchKernelGGL((rt_coop_persistent_kernel<Real, S, DUST, BTF, PROF>), g, b, lds, s, models, items,
                               n_items, counter, cap);
        }
    };
    auto by_prof = [&](auto slots_c, auto btf_c) {
        switch (prof) {
            case PROF_SHARE: launch(slots_c, btf_c, std::integral_constant<int, PROF_SHARE>{}); break;
            case PROF_REC: launch(slots_c, btf_c, std::integral_constant<int, PROF_REC>{}); break;
            default: launch(slots_c, btf_c, std::integral_constant<int, PROF_DIRECT>{}); break;
        }
    };
    auto by_order = [&](auto slots_c) {
        if (btf)
            by_prof(slots_c, std::true_type{});
        else
            by_prof(slots_c, std::false_type{});
    };
    switch (slots) {
        case 1: by_order(std::integral_constant<int, 1>{}); break;
        case 2: by_order(std::integral_constant<int, 2>{}); break;
        case 3: by_order(std::integral_constant<int, 3>{}); break;
        case 4: by_order(std::integral_constant<int, 4>{}); break;
        case 5: by_order(std::integral_constant<int, 5>{}); break;
        case 6: by_order(std::integral_constant<int, 6>{}); break;
        case 8: by_order(std::integral_constant<int, 8>{}); break;
        case 10: by_order(std::integral_constant<int, 10>{}); break;
        case 12: by_order(std::integral_constant<int, 12>{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// n_blocks > 0: the persistent form on n_blocks workgroups; n_blocks == 0: the flat form
template <typename Real>
hipError_t launch_coop_persistent(const RtArgs *models, const int2 *items, int n_items, int *counter, int cap,
                                  int slots, int waves, size_t lds, int n_blocks, bool dust, bool btf, int prof,
                                  hipStream_t s) {
    return dust ? launch_coop_persistent_d<Real, true>(models, items, n_items, counter, cap, slots, waves, lds,
                                                       n_blocks, btf, prof, s)
                : launch_coop_persistent_d<Real, false>(models, items, n_items, counter, cap, slots, waves, lds,
                                                        n_blocks, btf, prof, s);
}

template <typename Real>
hipError_t launch_coop(const RtArgs &a, int cap, hipStream_t s) {
    const int W = a.waves;
    if (W < 1 || W > 8 || a.R < 1) return hipErrorInvalidValue;
    const int unit = a.unit > 0 ? a.unit : coop_unit_cells<Real>(a.R, a.dust != 0, W);
    cap = cap < unit ? (cap < 1 ? 1 : cap) : unit;  // cells per staged unit
    const int npairs = a.R * a.C;
    const size_t lds = coop_lds_bytes<Real>(cap, a.R, npairs, a.dust != 0, a.LAM != nullptr, W);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    const int per_line = (64 * W) / a.R;
    if (per_line < 1) return hipErrorInvalidValue;
    const int need = (a.C + per_line - 1) / per_line;
    switch (need) {
        case 1: return launch_coop_s<Real, 1>(a, cap, lds, s);
        case 2: return launch_coop_s<Real, 2>(a, cap, lds, s);
        case 3: return launch_coop_s<Real, 3>(a, cap, lds, s);
        case 4: return launch_coop_s<Real, 4>(a, cap, lds, s);
        case 5: return launch_coop_s<Real, 5>(a, cap, lds, s);
        case 6: return launch_coop_s<Real, 6>(a, cap, lds, s);
        case 7: case 8: return launch_coop_s<Real, 8>(a, cap, lds, s);
        case 9: case 10: return launch_coop_s<Real, 10>(a, cap, lds, s);
        case 11: case 12: return launch_coop_s<Real, 12>(a, cap, lds, s);
        default: return hipErrorInvalidValue;  // more pairs than one workgroup holds
    }
}

}  // namespace portal
