// Statistical-equilibrium assembly and solve, one workgroup of 1, 2 or 4 wavefronts per node (FP64).
//
// Restates make_AC (miscellaneous.f90:288-392), setup_stateq (pop_eqs.f90:361-653)
// and get_rhovec (em_coefs.f90:115-201) with the packing of trace_node
// (main_sub.f90:112-137).  The Wigner-coupled coefficients t_A, t_S, t_E, r_A, r_S
// (pop_eqs.f90:1-359) depend only on (J,K,J',K'); the host folds them into a term
// list per matrix entry, so a node's matrix is  A[e] = sum_k coef_k * x[q_k]  with
// ten node scalars per transition  x = (B21 J0, B21 J2, B12 J0, B12 J2, A, C21, C12,
// B21 J0*, B12 J0*, A*).  The last three serve the entries that couple the K = 0 slots
// (the population block); they equal B21 J0, B12 J0 and A unless the accelerated
// lambda iteration is on.  Then J0* = J0 - lam S and A* = A (1 - lam), with lam the
// node's own-cell share of J0 per unit source function and S = A n_u / (B12 n_l -
// B21 n_u) from the level densities that produced J0: the own cell's emission is
// moved into the rates with the populations being solved for (Rybicki & Hummer 1991,
// A&A 245, 171, eq. 2-4), and at a fixed point J0* + lam S = J0, so the solution is
// that of the plain iteration.
// Row 1 is replaced by sqrt(2J+1) at every level's K=0 slot and A rho = e1 is
// solved by LU with partial pivoting (first maximal pivot, as DGETRF) in LDS,
// followed by one step of iterative refinement with a compensated
// residual.  The reference takes column 1 of DGETRI's explicit inverse, which is the
// same vector up to rounding.  Compiled with -ffp-contract=off.
#include "common.h"

namespace portal {

__device__ __forceinline__ void two_prod(double a, double b, double &p, double &e) {
    p = a * b;
    e = fma(a, b, -p);
}

// lane l's x to the whole wave (l wave-uniform)
__device__ __forceinline__ double lane_bcast(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Collisional depolarisation: subtract D^(K)(T) n_H2 1e-6 from the diagonal of every
// K >= 2 row of M (K = 2 sl for level i's slot sl), the table interpolated with the
// collision rates' bracket (ib, below); the threads of the workgroup split the levels.
__device__ __forceinline__ void add_depol(const SeArgs &a, double *M, int cd, int ib, bool below, double T,
                                          double nH2, int tid, int nt) {
    const int L = a.L;
    for (int i = tid; i < L; i += nt) {
        const int c0 = a.count_array[i], c1 = i + 1 < L ? a.count_array[i + 1] : cd;
        for (int sl = 1; sl < c1 - c0; sl++) {
            const double *f = a.depol + (size_t)a.c_temp * (i + (size_t)L * (2 * sl - 2));
            double D;
            if (ib >= 0) {
                const double ddT = (T - a.col_temps[ib - 1]) / (a.col_temps[ib] - a.col_temps[ib - 1]);
                D = f[ib - 1] + ddT * (f[ib] - f[ib - 1]);
            } else {
                D = below ? f[0] : __builtin_nan("");
            }
            const int d = c0 + sl;
            M[d + (size_t)cd * d] = M[d + (size_t)cd * d] - D * nH2 * 1e-6;
        }
    }
}

// A system is solved by W wavefronts together (64W threads, "waves per system"): the
// elementwise loops stride over all of them, the trailing update of each LU step splits
// its columns between the waves (column j to wave j mod W), and every wave runs the
// pivot search on its own, so the factors and the solution do not depend on W.
// The substitutions and the refinement of se_node for more than 128 unknowns, on all
// NT threads with the vectors in LDS (one barrier per step): thread t owns rows t, t+NT,
// .., and every row sees the same operations in the same order as in the one-wave form:
// y_i = b_perm(i); forward y_i -= l_ik y_k for k < i in order; backward y_i /= u_ii at
// k = i and y_i -= u_ik y_k for k > i in decreasing k; the residual r = e1 - M0 v summed
// over j in order with TwoProduct compensation.
template <int NT>
__device__ __forceinline__ void se_solve_coop(const double *M, const double *M0, double *v, double *y, double *b,
                                              const int *perm, const int cd, const int tid) {
    for (int it = 0; it < 2; it++) {
        for (int i = tid; i < cd; i += NT) {
            if (it == 0) {
                b[i] = i == 0 ? 1.0 : 0.0;
            } else {
                double s0 = i == 0 ? 1.0 : 0.0, c0 = 0.0;
                for (int j = 0; j < cd; j++) {
                    double p, e;
                    two_prod(-M0[i + (size_t)cd * j], v[j], p, e);
                    const double t = s0 + p;
                    const double z = t - s0;
                    c0 += (s0 - (t - z)) + (p - z) + e;
                    s0 = t;
                }
                b[i] = s0 + c0;
            }
        }
        __syncthreads();
        for (int i = tid; i < cd; i += NT) y[i] = b[perm[i]];
        __syncthreads();
        for (int k = 0; k < cd; k++) {
            const double yk = y[k];
            for (int i = k + 1 + tid; i < cd; i += NT) y[i] = y[i] - M[i + (size_t)cd * k] * yk;
            __syncthreads();
        }
        for (int k = cd - 1; k >= 0; k--) {
            if (tid == 0) y[k] = y[k] / M[k + (size_t)cd * k];
            __syncthreads();
            const double yk = y[k];
            for (int i = tid; i < k; i += NT) y[i] = y[i] - M[i + (size_t)cd * k] * yk;
            __syncthreads();
        }
        for (int i = tid; i < cd; i += NT) v[i] = it == 0 ? y[i] : v[i] + y[i];
        __syncthreads();
    }
}

// MODE 0: everything in LDS.  MODE 1 (slab form): the copy of the assembled matrix
// that the refinement residual reads, and the transition scalars, live in global memory
// at slab (the workgroup's own, cd^2 + 10 nt doubles) rather than in LDS.  MODE 2 (more
// than 128 unknowns): the factors too are in the slab (2 cd^2 + 10 nt doubles), LDS holds
// three vectors, and the substitutions run on all waves (se_solve_coop) instead of in one
// wave's registers; the arithmetic of every row is the same.
template <int W, int MODE>
__device__ __forceinline__ void se_node(const SeArgs &a, const int wi, double *sm, double *slab) {
    constexpr int NT = 64 * W;
    constexpr bool G = MODE >= 1;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int node = a.work[wi];
    const int cd = a.c_dim, nt = a.n_trans, L = a.L, R = a.R;
    const size_t cd2 = (size_t)cd * cd;
    double *x = MODE == 0 ? sm : MODE == 1 ? slab + cd2 : slab + 2 * cd2;  // 10*nt scalars
    double *M = MODE == 0 ? x + 10 * nt : MODE == 1 ? sm : slab;          // LU in place, column-major
    double *M0 = MODE == 0 ? M + cd2 : MODE == 1 ? slab : slab + cd2;     // assembled matrix (residual)
    double *v = MODE == 0 ? M0 + cd2 : MODE == 1 ? M + cd2 : sm;          // solution
    double *y = v + cd;                      // rhs / work
    double *bv = y + cd;                     // MODE 2: right-hand side
    int *perm = (int *)(MODE == 2 ? bv + cd : y + cd);  // row permutation of the factorisation
    __shared__ int s_info;

    const double T = a.tkin[node];
    const double nH2 = a.h2[node];
    // collision-rate interpolation bracket exactly as make_AC's loop
    int ib = -1;
    for (int i = 1; i < a.c_temp; i++)
        if (a.col_temps[i] > T) { ib = i; break; }
    const bool below = T <= a.col_temps[0];
    bool temp_ok = (ib >= 0) || (below && a.c_temp == 1);
    if (tid == 0) s_info = 0;
    if (!temp_ok && tid == 0) atomicOr(a.flags, FLAG_TEMP);

    for (int t = tid; t < nt; t += NT) {
        const int rad = a.t_rad[t], col = a.t_col[t];
        const double A = a.t_A[t], B21 = a.t_B21[t], nu = a.t_nu[t], gul = a.t_gul[t];
        const double J0v = rad >= 0 ? a.J0[(size_t)node * R + rad] : 0.0;
        const double J2v = rad >= 0 ? a.J2[(size_t)node * R + rad] : 0.0;
        double Cr = 0.0;
        if (col >= 0) {
            const int nT = a.c_temp, nC = a.c_num;
            const double *Ct = a.t_Cji;
            if (ib >= 0) {
                const double ddT = (T - a.col_temps[ib - 1]) / (a.col_temps[ib] - a.col_temps[ib - 1]);
                if (a.c_par == 1) {
                    const double f1 = Ct[(ib - 1) + (size_t)nT * col], f2 = Ct[ib + (size_t)nT * col];
                    Cr = f1 + ddT * (f2 - f1);
                } else {
                    const double f1 = Ct[(ib - 1) + (size_t)nT * col], f2 = Ct[ib + (size_t)nT * col];
                    const double f3 = Ct[(ib - 1) + (size_t)nT * (col + nC)], f4 = Ct[ib + (size_t)nT * (col + nC)];
                    Cr = 0.25 * (f1 + ddT * (f2 - f1)) + 0.75 * (f3 + ddT * (f4 - f3));
                }
            } else if (below) {
                Cr = a.c_par == 1 ? Ct[(size_t)nT * col] : 0.25 * Ct[(size_t)nT * col] + 0.75 * Ct[(size_t)nT * (col + nC)];
            } else {
                Cr = __builtin_nan("");
            }
            Cr = Cr * nH2 * 1e-6;
        }
        const double B12 = B21 * gul;
        const double C12 = Cr * gul * exp(-4.79924e-11 * nu / T);
        double Jp = J0v, Ap = A;
        if (a.lam && rad >= 0) {
            const double lam = a.lam[(size_t)node * R + rad];
            const double nl = a.levp[(size_t)node * L + a.t_lo[t]], nu_ = a.levp[(size_t)node * L + a.t_up[t]];
            const double den = B12 * nl - B21 * nu_;
            const double S = A * nu_ / den;
            // no operator for an inverted or empty line, or a lam outside [0, 1)
            if (den > 0.0 && nu_ >= 0.0 && isfinite(S) && lam >= 0.0 && lam < 1.0) {
                Jp = fmax(J0v - lam * S, 0.0);
                Ap = A * (1.0 - lam);
            }
        }
        x[10 * t + 0] = B21 * J0v;
        x[10 * t + 1] = B21 * J2v;
        x[10 * t + 2] = B12 * J0v;
        x[10 * t + 3] = B12 * J2v;
        x[10 * t + 4] = A;
        x[10 * t + 5] = Cr;
        x[10 * t + 6] = C12;
        x[10 * t + 7] = B21 * Jp;
        x[10 * t + 8] = B12 * Jp;
        x[10 * t + 9] = Ap;
    }
    for (int i = tid; i < cd * cd; i += NT) M[i] = 0.0;
    __syncthreads();
    for (int e = tid; e < a.n_ent; e += NT) {
        double s = 0.0;
        for (int k = a.ent_ptr[e]; k < a.ent_ptr[e + 1]; k++) s = s + a.term_coef[k] * x[a.term_q[k]];
        M[a.ent_idx[e]] = s;
    }
    __syncthreads();
    // first row <- normalisation sum_i sqrt(2J_i+1) rho^0_0(i) = 1; collisional
    // depolarisation on the diagonal of the K >= 2 rows (pop_eqs.f90:642-646; never row 0),
    // interpolated in T as the collision rates are, times n_H2 1e-6
    // (make_depolarization_temp, depol.f90:319-359)
    for (int j = tid; j < cd; j += NT) M[(size_t)cd * j] = 0.0;
    if (a.depol) add_depol(a, M, cd, ib, below, T, nH2, tid, NT);
    __syncthreads();
    for (int i = tid; i < L; i += NT) M[(size_t)cd * a.count_array[i]] = sqrt(2.0 * a.j_lev[i] + 1.0);
    __syncthreads();
    for (int i = tid; i < cd * cd; i += NT) M0[i] = M[i];
    __syncthreads();

    for (int i = tid; i < cd; i += NT) perm[i] = i;
    __syncthreads();

    // LU with partial pivoting, three barriers per column.  A: every wave finds the
    // pivot in column k and the rows k and bi are swapped in every column but k.
    // B: column k is scaled, taking the swap of its entries k and bi into account.
    // C: the trailing update, a lane owning rows k+1+lane, k+65+lane of its wave's
    // columns; M[k][k] receives the pivot.  perm records the swaps for the solve.
    for (int k = 0; k < cd; k++) {
        double best = -1.0, bv = 0.0;
        int bi = cd;
        for (int i = k + lane; i < cd; i += 64) {
            const double mv = M[i + (size_t)cd * k], vv = fabs(mv);
            if (vv > best) { best = vv; bi = i; bv = mv; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            const double ov = __shfl_xor(bv, o, 64);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; bv = ov; }
        }
        if (!(best > 0.0)) {
            if (tid == 0) s_info = k + 1;
            break;
        }
        const double pv = bv;
        if (bi != k) {
            if (tid == 0) {
                const int t = perm[k];
                perm[k] = perm[bi];
                perm[bi] = t;
            }
            for (int j = tid; j < cd; j += NT) {
                if (j == k) continue;
                const double t = M[k + (size_t)cd * j];
                M[k + (size_t)cd * j] = M[bi + (size_t)cd * j];
                M[bi + (size_t)cd * j] = t;
            }
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < cd; i += NT) {
            const double mv = (i == bi) ? M[k + (size_t)cd * k] : M[i + (size_t)cd * k];
            M[i + (size_t)cd * k] = mv / pv;
        }
        __syncthreads();
        if (tid == 0) M[k + (size_t)cd * k] = pv;
        for (int i = k + 1 + lane; i < cd; i += 64) {
            const double li = M[i + (size_t)cd * k];
            int j = k + 1 + wv;
            // eight columns' loads in flight before their stores (rows i and k differ)
            for (; j + 7 * W < cd; j += 8 * W) {
                double u[8], m[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    u[q] = M[k + (size_t)cd * (j + q * W)];
                    m[q] = M[i + (size_t)cd * (j + q * W)];
                }
#pragma unroll
                for (int q = 0; q < 8; q++) M[i + (size_t)cd * (j + q * W)] = m[q] - li * u[q];
            }
            for (; j < cd; j += W) M[i + (size_t)cd * j] = M[i + (size_t)cd * j] - li * M[k + (size_t)cd * j];
        }
        __syncthreads();
    }
    __syncthreads();
    const int info = s_info;
    if (info) {
        if (tid == 0) atomicOr(a.flags, FLAG_LU);
        return;
    }
    // Solve for e1, then one refinement step with a compensated residual, in wave 0's
    // registers with no barrier: lane l holds rows l and l + 64 (cd <= 128), the
    // permutation is a gather by perm, and each substitution step broadcasts one
    // element by readlane.
    if constexpr (MODE == 2) {
        se_solve_coop<NT>(M, M0, v, y, bv, perm, cd, tid);
    } else if (wv == 0) {
        const int i0 = lane, i1 = lane + 64;
        const bool h0 = i0 < cd, h1 = i1 < cd;
        const int p0 = h0 ? perm[i0] : 0, p1 = h1 ? perm[i1] : 0;
        double b0 = (i0 == 0) ? 1.0 : 0.0, b1 = 0.0, v0 = 0.0, v1 = 0.0;
        for (int it = 0; it < 2; it++) {
            if (it == 1) {
                // r = e1 - M0 v
                double s0 = (i0 == 0) ? 1.0 : 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
                for (int j0 = 0; j0 < cd; j0 += 8) {
                    double m0[8], m1[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        m0[q] = (h0 && j0 + q < cd) ? M0[i0 + (size_t)cd * (j0 + q)] : 0.0;
                        m1[q] = (h1 && j0 + q < cd) ? M0[i1 + (size_t)cd * (j0 + q)] : 0.0;
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int j = j0 + q;
                        if (j >= cd) break;
                        const double vj = lane_bcast(j < 64 ? v0 : v1, j & 63);
                        if (h0) {
                            double p, e;
                            two_prod(-m0[q], vj, p, e);
                            const double t = s0 + p;
                            const double z = t - s0;
                            c0 += (s0 - (t - z)) + (p - z) + e;
                            s0 = t;
                        }
                        if (h1) {
                            double p, e;
                            two_prod(-m1[q], vj, p, e);
                            const double t = s1 + p;
                            const double z = t - s1;
                            c1 += (s1 - (t - z)) + (p - z) + e;
                            s1 = t;
                        }
                    }
                }
                b0 = s0 + c0;
                b1 = s1 + c1;
            }
            const double g00 = __shfl(b0, p0 & 63, 64), g01 = __shfl(b1, p0 & 63, 64);
            const double g10 = __shfl(b0, p1 & 63, 64), g11 = __shfl(b1, p1 & 63, 64);
            double y0 = p0 < 64 ? g00 : g01, y1 = p1 < 64 ? g10 : g11;
            // the factors' entries are loaded eight steps at a time
            for (int k0 = 0; k0 < cd; k0 += 8) {
                double l0[8], l1[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    l0[q] = (h0 && k0 + q < cd) ? M[i0 + (size_t)cd * (k0 + q)] : 0.0;
                    l1[q] = (h1 && k0 + q < cd) ? M[i1 + (size_t)cd * (k0 + q)] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int k = k0 + q;
                    if (k >= cd) break;
                    const double yk = lane_bcast(k < 64 ? y0 : y1, k & 63);
                    if (h0 && i0 > k) y0 = y0 - l0[q] * yk;
                    if (h1 && i1 > k) y1 = y1 - l1[q] * yk;
                }
            }
            for (int k0 = (cd - 1) & ~7; k0 >= 0; k0 -= 8) {
                double l0[8], l1[8], dg[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int kk = k0 + q;
                    l0[q] = (h0 && kk < cd) ? M[i0 + (size_t)cd * kk] : 0.0;
                    l1[q] = (h1 && kk < cd) ? M[i1 + (size_t)cd * kk] : 0.0;
                    dg[q] = kk < cd ? M[kk + (size_t)cd * kk] : 1.0;
                }
#pragma unroll
                for (int q = 7; q >= 0; q--) {
                    const int k = k0 + q;
                    if (k >= cd) continue;
                    const double dk = dg[q];
                    if (i0 == k) y0 = y0 / dk;
                    if (i1 == k) y1 = y1 / dk;
                    const double yk = lane_bcast(k < 64 ? y0 : y1, k & 63);
                    if (i0 < k) y0 = y0 - l0[q] * yk;
                    if (h1 && i1 < k) y1 = y1 - l1[q] * yk;
                }
            }
            if (it == 0) {
                v0 = y0;
                v1 = y1;
            } else {
                v0 = v0 + y0;
                v1 = v1 + y1;
            }
        }
        if (h0) v[i0] = v0;
        if (h1) v[i1] = v1;
    }
    __syncthreads();
    // pack (rho00, rho20) per level, NaN -> 0 (main_sub.f90:112-137)
    const int N = a.N;
    for (int i = tid; i < L; i += NT) {
        const int ca = a.count_array[i];
        double r0 = v[ca], r2 = a.j_lev[i] != 0 ? v[ca + 1] : 0.0;
        if (isnan(r0)) { r0 = 0.0; r2 = 0.0; }
        a.pol[node + (size_t)N * (2 * i)] = r0;
        a.pol[node + (size_t)N * (2 * i + 1)] = r2;
    }
    if (a.extra_info && a.rates) {
        // |spontaneous|, |radiative|, |collisional| rate matrices (before the row-1
        // replacement) times rho, em_coefs.f90:188-199
        for (int cls = 0; cls < 3; cls++) {
            __syncthreads();
            for (int i = tid; i < cd * cd; i += NT) M[i] = 0.0;
            __syncthreads();
            for (int e = tid; e < a.n_ent; e += NT) {
                double s = 0.0;
                for (int k = a.ent_ptr[e]; k < a.ent_ptr[e + 1]; k++)
                    if (a.term_class[k] == cls) s = s + a.term_coef[k] * x[a.term_q[k]];
                M[a.ent_idx[e]] = s;
            }
            __syncthreads();
            if (cls == 2 && a.depol) {  // the collisional matrix carries -D^(K) too (pop_eqs.f90:645)
                add_depol(a, M, cd, ib, below, T, nH2, tid, NT);
                __syncthreads();
            }
            for (int i = tid; i < cd; i += NT) {
                double s = 0.0;
                for (int j = 0; j < cd; j++) s = s + fabs(M[i + (size_t)cd * j]) * v[j];
                y[i] = s;
            }
            __syncthreads();
            for (int i = tid; i < L; i += NT) {
                const int ca = a.count_array[i];
                const bool nanr = isnan(v[ca]);
                a.rates[node + (size_t)N * (cls + 3 * (size_t)(2 * i))] = y[ca];
                a.rates[node + (size_t)N * (cls + 3 * (size_t)(2 * i + 1))] =
                    (a.j_lev[i] != 0 && !nanr) ? y[ca + 1] : 0.0;
            }
        }
    }
}

template <int W>
__global__ void __launch_bounds__(64 * W) se_kernel(SeArgs a) {
    extern __shared__ double sm[];
    if ((int)blockIdx.x >= a.n_work) return;
    se_node<W, 0>(a, blockIdx.x, sm, nullptr);
}

// Persistent form over the (model, node) items of a batch of models, one wave per
// system, with the LDS of the largest system.
__global__ void __launch_bounds__(64) se_persistent_kernel(const SeArgs *models, const int2 *items, int n_items,
                                                           int *counter) {
    extern __shared__ double sm[];
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= n_items) break;
        const int2 mi = items[it];
        se_node<1, 0>(models[mi.x], mi.y, sm, nullptr);
        __syncthreads();
    }
}

// Persistent form for a class of models with one system size: W waves per system, the
// residual's matrix copy and the transition scalars in the workgroup's slab of `slab`
// doubles, so that LDS holds only the factors and two vectors (MODE 1); above 128
// unknowns the factors too (MODE 2, see se_node).
template <int W, int MODE>
__global__ void __launch_bounds__(64 * W) se_slab_kernel(const SeArgs *models, const int2 *items, int n_items,
                                                        int *counter, double *slabs, int slab) {
    extern __shared__ double sm[];
    __shared__ int s_item;
    double *mine = slabs + (size_t)blockIdx.x * slab;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= n_items) break;
        const int2 mi = items[it];
        se_node<W, MODE>(models[mi.x], mi.y, sm, mine);
        __syncthreads();
    }
}

// LDS of one system: with m0_in_lds the scalars, the factors, the matrix copy and two
// vectors; without (the slab form) the factors and two vectors, or above 128 unknowns
// three vectors
size_t se_lds_bytes(int c_dim, int n_trans, bool m0_in_lds, bool all_global) {
    const size_t cd = (size_t)c_dim;
    const bool g = all_global || c_dim > 128;
    const size_t dbl = m0_in_lds ? 10 * (size_t)n_trans + 2 * cd * cd + 2 * cd : g ? 3 * cd : cd * cd + 2 * cd;
    return sizeof(double) * dbl + sizeof(int) * cd + 16;
}

// doubles of one workgroup's slab: the matrix copy and the transition scalars, and above
// 128 unknowns the factors
size_t se_slab_doubles(int c_dim, int n_trans, bool all_global) {
    return (size_t)(all_global || c_dim > 128 ? 2 : 1) * c_dim * c_dim + 10 * (size_t)n_trans;
}

hipError_t launch_se_persistent(const SeArgs *models, const int2 *items, int n_items, int *counter, int cd_max,
                                int nt_max, int n_blocks, hipStream_t s) {
    const size_t lds = se_lds_bytes(cd_max, nt_max, true);
    if (lds > 160 * 1024 || cd_max > 128) return hipErrorInvalidValue;
    hipLaunchKernelGGL(se_persistent_kernel, dim3(n_blocks), dim3(64), lds, s, models, items, n_items, counter);
    return hipGetLastError();
}

template <int MODE>
static int slab_blocks(int waves, size_t lds) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (waves) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, se_slab_kernel<1, MODE>, 64, lds); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, se_slab_kernel<2, MODE>, 128, lds); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, se_slab_kernel<4, MODE>, 256, lds); break;
        default: break;
    }
    return e == hipSuccess ? nb : 0;
}

int se_slab_blocks_per_cu(int waves, int cd, int nt, bool all_global) {
    const size_t lds = se_lds_bytes(cd, nt, false, all_global);
    if (lds > 160 * 1024 || cd > SE_MAX_UNKNOWNS) return 0;
    // above 128 unknowns the factorisation works in the slab: a few systems per CU
    // already keep it busy, and the slabs stay within a few GB
    return all_global || cd > 128 ? std::min(4, slab_blocks<2>(waves, lds)) : slab_blocks<1>(waves, lds);
}

template <int MODE>
static hipError_t launch_slab(const SeArgs *models, const int2 *items, int n_items, int *counter, int waves,
                              int n_blocks, double *slabs, int slab, size_t lds, hipStream_t s) {
    switch (waves) {
        case 1:
            hipLaunchKernelGGL((se_slab_kernel<1, MODE>), dim3(n_blocks), dim3(64), lds, s, models, items, n_items,
                               counter, slabs, slab);
            break;
        case 2:
            hipLaunchKernelGGL((se_slab_kernel<2, MODE>), dim3(n_blocks), dim3(128), lds, s, models, items, n_items,
                               counter, slabs, slab);
            break;
        case 4:
            hipLaunchKernelGGL((se_slab_kernel<4, MODE>), dim3(n_blocks), dim3(256), lds, s, models, items, n_items,
                               counter, slabs, slab);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_se_slab(const SeArgs *models, const int2 *items, int n_items, int *counter, int cd, int nt,
                          int waves, int n_blocks, double *slabs, hipStream_t s, bool all_global) {
    const size_t lds = se_lds_bytes(cd, nt, false, all_global);
    if (lds > 160 * 1024 || cd > SE_MAX_UNKNOWNS || n_items < 0 || n_blocks < 1) return hipErrorInvalidValue;
    const int slab = (int)se_slab_doubles(cd, nt, all_global);
    return all_global || cd > 128 ? launch_slab<2>(models, items, n_items, counter, waves, n_blocks, slabs, slab, lds, s)
                    : launch_slab<1>(models, items, n_items, counter, waves, n_blocks, slabs, slab, lds, s);
}

hipError_t launch_se(const SeArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    const size_t lds = se_lds_bytes(a.c_dim, a.n_trans, true);
    if (lds > 160 * 1024 || a.c_dim > 128) return hipErrorInvalidValue;
    switch (a.waves <= 1 ? 1 : a.waves) {
        case 1: hipLaunchKernelGGL(se_kernel<1>, dim3(a.n_work), dim3(64), lds, s, a); break;
        case 2: hipLaunchKernelGGL(se_kernel<2>, dim3(a.n_work), dim3(128), lds, s, a); break;
        case 4: hipLaunchKernelGGL(se_kernel<4>, dim3(a.n_work), dim3(256), lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace portal
