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

// A system is solved by W wavefronts together (64W threads, "waves per system"): the
// elementwise loops stride over all of them, the trailing update of each LU step splits
// its columns between the waves (column j to wave j mod W), and every wave runs the
// pivot search on its own, so the factors and the solution do not depend on W.
template <int W>
__device__ __forceinline__ void se_node(const SeArgs &a, const int wi, double *sm) {
    constexpr int NT = 64 * W;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int node = a.work[wi];
    const int cd = a.c_dim, nt = a.n_trans, L = a.L, R = a.R;
    double *x = sm;                          // 10*nt scalars
    double *M = x + 10 * nt;                 // LU in place, column-major cd x cd
    double *M0 = M + cd * cd;                // assembled matrix (for the residual)
    double *v = M0 + cd * cd;                // solution
    double *y = v + cd;                      // rhs / work
    int *piv = (int *)(y + cd);              // pivots
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
    // first row <- normalisation sum_i sqrt(2J_i+1) rho^0_0(i) = 1
    for (int j = tid; j < cd; j += NT) M[(size_t)cd * j] = 0.0;
    __syncthreads();
    for (int i = tid; i < L; i += NT) M[(size_t)cd * a.count_array[i]] = sqrt(2.0 * a.j_lev[i] + 1.0);
    __syncthreads();
    for (int i = tid; i < cd * cd; i += NT) M0[i] = M[i];
    __syncthreads();

    // LU with partial pivoting; in the trailing update a lane owns rows k+1+lane,
    // k+65+lane, ... of its wave's columns
    for (int k = 0; k < cd; k++) {
        double best = -1.0;
        int bi = cd;
        for (int i = k + lane; i < cd; i += 64) {
            const double vv = fabs(M[i + (size_t)cd * k]);
            if (vv > best) { best = vv; bi = i; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (!(best > 0.0)) {
            if (tid == 0) s_info = k + 1;
            break;
        }
        if (tid == 0) piv[k] = bi;
        // every wave has read column k before any wave swaps rows
        if (W > 1) __syncthreads();
        if (bi != k)
            for (int j = tid; j < cd; j += NT) {
                const double t = M[k + (size_t)cd * j];
                M[k + (size_t)cd * j] = M[bi + (size_t)cd * j];
                M[bi + (size_t)cd * j] = t;
            }
        __syncthreads();
        const double pv = M[k + (size_t)cd * k];
        for (int i = k + 1 + tid; i < cd; i += NT) M[i + (size_t)cd * k] = M[i + (size_t)cd * k] / pv;
        __syncthreads();
        for (int i = k + 1 + lane; i < cd; i += 64) {
            const double li = M[i + (size_t)cd * k];
            for (int j = k + 1 + wv; j < cd; j += W)
                M[i + (size_t)cd * j] = M[i + (size_t)cd * j] - li * M[k + (size_t)cd * j];
        }
        __syncthreads();
    }
    __syncthreads();
    const int info = s_info;
    if (info) {
        if (tid == 0) atomicOr(a.flags, FLAG_LU);
        return;
    }
    // solve with rhs e1, then one refinement step
    for (int it = 0; it < 2; it++) {
        if (it == 0) {
            for (int i = tid; i < cd; i += NT) y[i] = (i == 0) ? 1.0 : 0.0;
        } else {
            // compensated residual r = e1 - M0 v
            for (int i = tid; i < cd; i += NT) {
                double s = (i == 0) ? 1.0 : 0.0, c = 0.0;
                for (int j = 0; j < cd; j++) {
                    double p, e;
                    two_prod(-M0[i + (size_t)cd * j], v[j], p, e);
                    const double t = s + p;
                    const double z = t - s;
                    c += (s - (t - z)) + (p - z) + e;
                    s = t;
                }
                y[i] = s + c;
            }
        }
        __syncthreads();
        if (tid == 0)
            for (int k = 0; k < cd; k++) {
                const int p = piv[k];
                if (p != k) { const double t = y[k]; y[k] = y[p]; y[p] = t; }
            }
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
        for (int i = tid; i < cd; i += NT) v[i] = (it == 0) ? y[i] : v[i] + y[i];
        __syncthreads();
    }
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
    se_node<W>(a, blockIdx.x, sm);
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
        se_node<1>(models[mi.x], mi.y, sm);
        __syncthreads();
    }
}

// One workgroup of W waves per (model, node) item: a class of models whose systems
// share one LDS size.
template <int W>
__global__ void __launch_bounds__(64 * W) se_items_kernel(const SeArgs *models, const int2 *items, int n_items) {
    extern __shared__ double sm[];
    if ((int)blockIdx.x >= n_items) return;
    const int2 mi = items[blockIdx.x];
    se_node<W>(models[mi.x], mi.y, sm);
}

size_t se_lds_bytes(int c_dim, int n_trans) {
    return sizeof(double) * (10 * (size_t)n_trans + 2 * (size_t)c_dim * c_dim + 2 * c_dim) + sizeof(int) * c_dim + 16;
}

hipError_t launch_se_persistent(const SeArgs *models, const int2 *items, int n_items, int *counter, int cd_max,
                                int nt_max, int n_blocks, hipStream_t s) {
    const size_t lds = se_lds_bytes(cd_max, nt_max);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(se_persistent_kernel, dim3(n_blocks), dim3(64), lds, s, models, items, n_items, counter);
    return hipGetLastError();
}

hipError_t launch_se_items(const SeArgs *models, const int2 *items, int n_items, int cd_max, int nt_max, int waves,
                           hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    const size_t lds = se_lds_bytes(cd_max, nt_max);
    if (lds > 160 * 1024 || n_items < 0) return hipErrorInvalidValue;
    switch (waves) {
        case 1: hipLaunchKernelGGL(se_items_kernel<1>, dim3(n_items), dim3(64), lds, s, models, items, n_items); break;
        case 2: hipLaunchKernelGGL(se_items_kernel<2>, dim3(n_items), dim3(128), lds, s, models, items, n_items); break;
        case 4: hipLaunchKernelGGL(se_items_kernel<4>, dim3(n_items), dim3(256), lds, s, models, items, n_items); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_se(const SeArgs &a, hipStream_t s) {
    if (a.n_work == 0) return hipSuccess;
    const size_t lds = se_lds_bytes(a.c_dim, a.n_trans);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    switch (a.waves <= 1 ? 1 : a.waves) {
        case 1: hipLaunchKernelGGL(se_kernel<1>, dim3(a.n_work), dim3(64), lds, s, a); break;
        case 2: hipLaunchKernelGGL(se_kernel<2>, dim3(a.n_work), dim3(128), lds, s, a); break;
        case 4: hipLaunchKernelGGL(se_kernel<4>, dim3(a.n_work), dim3(256), lds, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace portal
