// Inward Voronoi walk: the cells every inward ray of every node crosses.
//
// Restates trace_out (inverse_rt.f90:139-214) with check_neigh_fast and
// pos_quadratic (miscellaneous.f90:1017-1098) and the ray direction of find_dir
// (miscellaneous.f90:48-82, rotmat 146-181, two_mat_vec_prod 183-204).  The walk
// depends only on the geometry, so it runs once per (grid, direction set) and its
// result stays resident in HBM for every later pass and lambda iteration.
//
// One lane walks one (node, ray): the neighbour test is the reference's sequential
// loop (ties go to the later neighbour, a NaN resets the running minimum), which a
// lane reproduces exactly.  Compiled with -ffp-contract=off: every product is
// rounded as in the reference, so the nint() quantisation sees the same values.
#include "common.h"

namespace portal {

__device__ static void rotmat_d(const double nv[3], double th, double R[9]) {
    const double cth = cos(th), octh = 1.0 - cth, sth = sin(th);
    // column-major R[(i-1) + 3*(j-1)]
    R[0] = cth + octh * nv[0] * nv[0];
    R[4] = cth + octh * nv[1] * nv[1];
    R[8] = cth + octh * nv[2] * nv[2];
    R[3] = nv[0] * nv[1] * octh - nv[2] * sth;  // (1,2)
    R[1] = nv[1] * nv[0] * octh + nv[2] * sth;  // (2,1)
    R[6] = nv[0] * nv[2] * octh + nv[1] * sth;  // (1,3)
    R[2] = nv[2] * nv[0] * octh - nv[1] * sth;  // (3,1)
    R[7] = nv[1] * nv[2] * octh - nv[0] * sth;  // (2,3)
    R[5] = nv[2] * nv[1] * octh + nv[0] * sth;  // (3,2)
}

__device__ static void find_dir_d(const double b[3], const double g[3], double th, double ph, double dir[3]) {
    double xv[3];
    xv[0] = b[1] * g[2] - b[2] * g[1];
    xv[1] = b[2] * g[0] - b[0] * g[2];
    xv[2] = b[0] * g[1] - b[1] * g[0];
    const double xn = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    xv[0] = xv[0] / xn;
    xv[1] = xv[1] / xn;
    xv[2] = xv[2] / xn;
    double Rth[9], Rph[9];
    rotmat_d(xv, th, Rth);
    rotmat_d(b, ph, Rph);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double d = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) d = d + Rph[i + 3 * j] * Rth[j + 3 * k] * b[k];
        dir[i] = d;
    }
}

// Chunked output of one path: entries go to chunks of WALK_CHUNK taken from a global
// counter; a chunk records its (path, ordinal).  Past the capacity nothing is written and
// FLAG_CHUNKS is raised (the caller then writes the entries by a second walk).
struct ChunkWriter {
    long long chunk = -1;
    int slot = WALK_CHUNK, ord = -1;
    __device__ void put(const WalkArgs &a, long long id, const PathEntry &e) {
        if (slot == WALK_CHUNK) {
            ord++;
            slot = 0;
            const unsigned long long c = atomicAdd(a.chunk_next, 1ull);
            if ((long long)c < a.chunk_cap) {
                chunk = (long long)c;
                a.chunk_meta[c] = make_int2((int)id, ord);
            } else {
                chunk = -2;
                atomicOr(a.flags, FLAG_CHUNKS);
            }
        }
        if (chunk >= 0) a.chunks[chunk * WALK_CHUNK + slot] = e;
        slot++;
    }
};

template <int MODE>
__global__ void __launch_bounds__(256) walk_kernel(WalkArgs a) {
    constexpr bool WRITE = MODE == WALK_WRITE, CHUNKED = MODE == WALK_CHUNKED;
    const long long n_items = (a.n_work + a.sample - 1) / a.sample;
    const long long n_paths = n_items * a.D;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n_paths) return;
    // thread tid walks ray tid / n_items of work item sample * (tid % n_items): a wave's
    // 64 rays share a direction and start at nodes adjacent on the work list's Morton
    // curve, so they cross neighbouring cells; the path id stays w * D + r
    const int r = (int)(tid / n_items), w = (int)((tid % n_items) * a.sample);
    const long long id = (long long)w * a.D + r;
    const long long cidx = a.sample > 1 ? tid : id;  // a sampled count goes to its thread
    const int node = a.work[w];
    bool star_ray = false;
    ChunkWriter cw;
    if (a.xthn && r >= a.nvar[w]) {
        // star mode: a padding ray (weight 0) gets the node's own cell with zero length
        PathEntry e;
        e.cell = node;
        e.len = 0;
        e.vray = 0.0;
        if (WRITE) a.paths[a.offsets[id]] = e;
        if (CHUNKED) cw.put(a, id, e);
        if (!WRITE) {
            a.counts[cidx] = 1;
            atomicMax(a.flags + 1, 1);
        }
        return;
    }
    if (a.xthn) star_ray = r < a.nstar[w];
    const double *xr = a.xthn ? a.xthn + 3 * (size_t)id : a.xth + 3 * r;
    const double mu = xr[0], ph = xr[1];
    const double b[3] = {a.bvec[3 * (size_t)node], a.bvec[3 * (size_t)node + 1], a.bvec[3 * (size_t)node + 2]};
    const double g[3] = {a.gvec[0], a.gvec[1], a.gvec[2]};
    double dir[3];
    find_dir_d(b, g, acos(mu), ph, dir);

    const double4 p0 = a.pos[node];
    double xp0 = p0.x, xp1 = p0.y, xp2 = p0.z;
    int cur = node, nc = 0;
    double delta_x = 0.0;
    const double thr = (double)1e-10f;  // single-precision literal of miscellaneous.f90:1048
    const double nd = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const double qa = nd * nd;
    long long off = WRITE ? a.offsets[id] : 0;
    for (;;) {
        // distance to the unit sphere (pos_quadratic)
        const double npn = sqrt(xp0 * xp0 + xp1 * xp1 + xp2 * xp2);
        const double qb = 2.0 * (xp0 * dir[0] + xp1 * dir[1] + xp2 * dir[2]);
        const double qc = npn * npn - 1.0;
        const double dis = qb * qb - 4.0 * qa * qc;
        if (dis > 0) {
            double x1 = (-qb + sqrt(dis)) / (2.0 * qa);
            const double x2 = (-qb - sqrt(dis)) / (2.0 * qa);
            if (x2 > 0) x1 = x2;
            delta_x = x1;
        }
        int next = -1;
        const double4 pn = a.pos[cur];
        const double d0 = pn.x - xp0, d1 = pn.y - xp1, d2 = pn.z - xp2;
        const double t4 = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        const double q4 = t4 * t4;
        const double q5 = d0 * dir[0] + d1 * dir[1] + d2 * dir[2];
        const int q0 = a.nbr_ptr[cur], q_end = a.nbr_ptr[cur + 1];
        // a neighbour's position is loaded one neighbour, its index two neighbours, ahead of
        // its test, so that the gathers overlap the arithmetic (past the end of the list the
        // loads read the current node's own entry, which is never used)
        int nb_n = q0 < q_end ? a.nbr_idx[q0] : cur;
        double4 pb_n = a.pos[nb_n];
        int nb_nn = q0 + 1 < q_end ? a.nbr_idx[q0 + 1] : cur;
        for (int q = q0; q < q_end; q++) {
            const int nb = nb_n;
            const double4 pb = pb_n;
            nb_n = nb_nn;
            pb_n = a.pos[nb_n];
            nb_nn = q + 2 < q_end ? a.nbr_idx[q + 2] : cur;
            const double e0 = pb.x - xp0, e1 = pb.y - xp1, e2 = pb.z - xp2;
            const double u = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            const double q1 = u * u;
            const double q2 = e0 * dir[0] + e1 * dir[1] + e2 * dir[2];
            const double new_x = 0.5 * (q1 - q4) / (q2 - q5);
            if (new_x > delta_x || new_x <= thr) continue;
            delta_x = new_x;
            next = nb;
        }
        if (star_ray) {
            // check_neigh_fast's star test (miscellaneous.f90:1056-1069)
            const double s0 = xp0 - a.Xs0, s1 = xp1 - a.Xs1, s2 = xp2 - a.Xs2;
            const double sn = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
            const double sb = 2.0 * (s0 * dir[0] + s1 * dir[1] + s2 * dir[2]);
            const double sc = sn * sn - a.R_star * a.R_star;
            const double sdis = sb * sb - 4.0 * qa * sc;
            if (sdis > 0) {
                double x1 = (-sb + sqrt(sdis)) / (2.0 * qa);
                const double x2 = (-sb - sqrt(sdis)) / (2.0 * qa);
                if (x2 > 0) x1 = x2;
                if (x1 < delta_x && x1 > 0.0) {
                    delta_x = x1;
                    next = -1;
                }
            }
        }
        xp0 = xp0 + delta_x * dir[0];
        xp1 = xp1 + delta_x * dir[1];
        xp2 = xp2 + delta_x * dir[2];
        nc++;
        if (nc > a.n_depth) {
            atomicOr(a.flags, FLAG_DEPTH);
            break;
        }
        if (WRITE || CHUNKED) {
            PathEntry e;
            e.cell = cur;
            e.len = (int)round(delta_x / a.dx);
            const double4 v = a.velb[cur];
            e.vray = v.x * dir[0] + v.y * dir[1] + v.z * dir[2];
            if (WRITE) a.paths[off + nc - 1] = e;
            if (CHUNKED) cw.put(a, id, e);
        }
        cur = next;
        if (next < 0) break;
    }
    if (!WRITE) {
        a.counts[cidx] = nc > a.n_depth ? 0 : nc;
        atomicMax(a.flags + 1, nc);
    }
}

// one lane per (chunk, slot): entry ord*WALK_CHUNK + slot of its path, if the path has it
__global__ void __launch_bounds__(256) walk_compact_kernel(WalkArgs a, long long n_chunks) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long c = t / WALK_CHUNK;
    if (c >= n_chunks) return;
    const int slot = (int)(t % WALK_CHUNK);
    const int2 m = a.chunk_meta[c];
    const int k = m.y * WALK_CHUNK + slot;
    if (k < a.counts[m.x]) a.paths[a.offsets[m.x] + k] = a.chunks[t];
}

hipError_t launch_walk(const WalkArgs &a, int mode, hipStream_t s) {
    if (a.sample < 1 || (a.sample > 1 && mode != WALK_COUNT)) return hipErrorInvalidValue;
    const long long n = (long long)((a.n_work + a.sample - 1) / a.sample) * a.D;
    const int bs = 256;
    const unsigned nb = (unsigned)((n + bs - 1) / bs);
    if (nb == 0) return hipSuccess;
    if (mode == WALK_WRITE)
        hipLaunchKernelGGL(walk_kernel<WALK_WRITE>, dim3(nb), dim3(bs), 0, s, a);
    else if (mode == WALK_CHUNKED)
        hipLaunchKernelGGL(walk_kernel<WALK_CHUNKED>, dim3(nb), dim3(bs), 0, s, a);
    else
        hipLaunchKernelGGL(walk_kernel<WALK_COUNT>, dim3(nb), dim3(bs), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_walk_compact(const WalkArgs &a, long long n_chunks, hipStream_t s) {
    const long long n = n_chunks * WALK_CHUNK;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(walk_compact_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n_chunks);
    return hipGetLastError();
}

// ---- exclusive prefix sum of the per-path counts (int -> int64 offsets) ----------
// Three phases: per-block scan of 2048 elements, scan of block sums in one block,
// add block offsets.  Offsets has n+1 entries; offsets[n] is the total.
static constexpr int SCAN_BS = 256, SCAN_PER = 8, SCAN_CHUNK = SCAN_BS * SCAN_PER;

__global__ void __launch_bounds__(SCAN_BS) scan_local(const int *in, long long *out, long long *bsum, long long n) {
    __shared__ long long sh[SCAN_BS];
    const long long base = (long long)blockIdx.x * SCAN_CHUNK + (long long)threadIdx.x * SCAN_PER;
    long long v[SCAN_PER];
    long long acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; i++) {
        const long long k = base + i;
        v[i] = acc;
        acc += (k < n) ? in[k] : 0;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 1; o < SCAN_BS; o <<= 1) {
        const long long t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const long long excl = sh[threadIdx.x] - acc;
#pragma unroll
    for (int i = 0; i < SCAN_PER; i++) {
        const long long k = base + i;
        if (k < n) out[k] = v[i] + excl;
    }
    if (threadIdx.x == SCAN_BS - 1) bsum[blockIdx.x] = sh[SCAN_BS - 1];
}

__global__ void __launch_bounds__(1024) scan_blocks(long long *bsum, long long nb) {
    __shared__ long long sh[1024];
    const long long per = (nb + 1023) / 1024;
    const long long b0 = threadIdx.x * per;
    long long acc = 0;
    for (long long i = 0; i < per; i++) {
        const long long k = b0 + i;
        if (k < nb) acc += bsum[k];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = sh[threadIdx.x] - acc;
    for (long long i = 0; i < per; i++) {
        const long long k = b0 + i;
        if (k < nb) {
            const long long t = bsum[k];
            bsum[k] = run;
            run += t;
        }
    }
    if (threadIdx.x == 1023) bsum[nb] = sh[1023];
}

__global__ void __launch_bounds__(SCAN_BS) scan_add(long long *out, const long long *bsum, long long n) {
    const long long k = (long long)blockIdx.x * SCAN_CHUNK + threadIdx.x;
    const long long add = bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_PER; i++) {
        const long long kk = k + (long long)i * SCAN_BS;
        if (kk < n) out[kk] += add;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = bsum[gridDim.x];
}

hipError_t launch_scan_counts(const int *counts, long long *offsets, long long n, hipStream_t s,
                              long long *total_host) {
    if (n <= 0) {
        *total_host = 0;
        return n < 0 ? hipErrorInvalidValue : hipMemsetAsync(offsets, 0, sizeof(long long), s);
    }
    const long long nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    long long *bsum = nullptr;
    hipError_t e = hipMalloc((void **)&bsum, sizeof(long long) * (nb + 1));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(scan_local, dim3((unsigned)nb), dim3(SCAN_BS), 0, s, counts, offsets, bsum, n);
    hipLaunchKernelGGL(scan_blocks, dim3(1), dim3(1024), 0, s, bsum, nb);
    hipLaunchKernelGGL(scan_add, dim3((unsigned)nb), dim3(SCAN_BS), 0, s, offsets, bsum, n);
    e = hipMemcpyAsync(total_host, offsets + n, sizeof(long long), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(bsum);
    return e == hipSuccess ? hipGetLastError() : e;
}

}  // namespace portal
