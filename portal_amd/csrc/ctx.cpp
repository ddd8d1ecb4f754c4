// Host side of the C ABI: context, uploads, pass orchestration.
//
// The context keeps the grid, species tables, the inward paths and all per-pass
// buffers resident on the device; a call moves host inputs in once and the
// results out once.  Each pass is a short fixed sequence of kernel launches on
// the context's stream (line coefficients -> inward integration -> SE solve),
// timed with HIP events on that stream.
#include <algorithm>
#include <cmath>
#include <functional>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>

#include "common.h"
#include "image.h"

extern "C" double portal_debug_wig6jj(int a, int b, int c, int d, int e, int f);

// Device memory of the allocation-free inward pass (portal_inverse_rt_device): the
// direction table (3 x n_sphere doubles) and, in star mode, the star's intensity per
// line, staged here instead of in buffers allocated per call.  Every device holds its
// own copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for that device (g_scratch below), never through this
// name.  Calls that overlap on one device share it and race.
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 * 1024;
constexpr size_t NOALLOC_STAR_OFFSET = NOALLOC_SCRATCH_BYTES - 64 * sizeof(double);  // r_num <= 64 intensities
constexpr int NOALLOC_MAX_DIRS = (int)(NOALLOC_STAR_OFFSET / (3 * sizeof(double)));  // 10,901 directions
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace portal {
int build_species_tables(Species &s);
}

using namespace portal;

namespace {

template <typename T>
hipError_t upload(T *&d, const std::vector<T> &h) {
    if (d) hipFree(d);
    d = nullptr;
    if (h.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void **)&d, sizeof(T) * h.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

template <typename T>
void dfree(T *&p) {
    if (p) hipFree((void *)p);
    p = nullptr;
}

// runs f when its scope ends, on every return path
struct Defer {
    std::function<void()> f;
    ~Defer() {
        if (f) f();
    }
};

int fail(Ctx &c, int code, const std::string &msg) {
    c.err = msg;
    return code;
}

// the device address of `scratch` on each device, resolved once per device by the
// first context created there
constexpr int MAX_DEVICES = 64;
void *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// with the context's device current
int resolve_scratch(Ctx &c) {
    if (c.device >= MAX_DEVICES) return fail(c, PORTAL_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[c.device]) return PORTAL_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();  // leave no error behind for the context's other calls
        return fail(c, PORTAL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch[c.device] = ptr;
    return PORTAL_OK;
}

uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

double black_body(double T, double nu0) {  // miscellaneous.f90:263-285
    if (T == 0.0) return 0.0;
    const double c1 = 1.4744995e-50, c2 = 4.79924e-11 / T;
    return c1 * std::pow(nu0, 3.0) * std::pow(std::exp(c2 * nu0) - 1.0, -1.0);
}

struct Timer {
    hipEvent_t a, b;
    hipStream_t s;
    explicit Timer(hipStream_t st) : s(st) {
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, s);
    }
    double stop() {
        hipEventRecord(b, s);
        hipEventSynchronize(b);
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        return ms;
    }
    ~Timer() {
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
};

size_t node_buffer_key(const Ctx &c) {
    const size_t N = c.g.N, R = c.sp.r_num, L = c.sp.e_num;
    return N * std::max<size_t>(R, 1) * 1000 + L;
}

// the per-pass node buffers exist for the uploaded grid and species
bool node_buffers_ready(const Ctx &c) { return c.g.valid && c.sp.valid && c.cap_nodes == node_buffer_key(c); }

int ensure_node_buffers(Ctx &c) {
    const size_t N = c.g.N, R = c.sp.r_num, L = c.sp.e_num;
    if (!c.g.valid || !c.sp.valid) return fail(c, PORTAL_ERR_STATE, "grid and species must be uploaded first");
    const size_t need = node_buffer_key(c);
    if (c.cap_nodes == need) return PORTAL_OK;
    dfree(c.d_KS);
    dfree(c.d_KS32);
    dfree(c.d_KSD);
    dfree(c.d_KSD32);
    dfree(c.d_J0);
    dfree(c.d_J2);
    dfree(c.d_LAM);
    dfree(c.d_pol);
    dfree(c.d_rates);
    HIP_TRY(&c, hipMalloc((void **)&c.d_KS, sizeof(double) * 2 * N * R));
    HIP_TRY(&c, hipMalloc((void **)&c.d_KS32, sizeof(float) * 2 * N * R));
    HIP_TRY(&c, hipMalloc((void **)&c.d_J0, sizeof(double) * N * R));
    HIP_TRY(&c, hipMalloc((void **)&c.d_J2, sizeof(double) * N * R));
    HIP_TRY(&c, hipMalloc((void **)&c.d_LAM, sizeof(double) * N * R));
    HIP_TRY(&c, hipMalloc((void **)&c.d_pol, sizeof(double) * N * 2 * L));
    HIP_TRY(&c, hipMalloc((void **)&c.d_rates, sizeof(double) * N * 6 * L));
    if (!c.d_flags) HIP_TRY(&c, hipMalloc((void **)&c.d_flags, sizeof(int) * 4));
    if (!c.d_levp) {
        HIP_TRY(&c, hipMalloc((void **)&c.d_levp, sizeof(double) * N * L));
        HIP_TRY(&c, hipMemset(c.d_levp, 0, sizeof(double) * N * L));
    }
    c.cap_nodes = need;
    return PORTAL_OK;
}

bool same(const std::vector<double> &v, const double *p, size_t n) {
    return v.size() == n && std::memcmp(v.data(), p, n * sizeof(double)) == 0;
}

// The built inward paths are those of this direction set and field geometry.
bool paths_match(const Ctx &c, const portal_se_params *p) {
    const PathSet &P = c.paths;
    const bool star = p->star_mod == 1;
    const int N = c.g.N, D = star ? p->nth * p->nph : p->n_sphere;
    const bool same_star = !star ? P.star == 0
                                 : (P.star == 1 && P.nth == p->nth && P.nph == p->nph && P.R_star == p->R_star &&
                                    same(P.X_star, p->X_star, 3));
    return P.valid && same_star && P.D == D && P.dx == p->dx && P.n_depth == p->n_depth &&
           (star || same(P.xth, p->xth, 3 * (size_t)D)) && same(P.gvec, p->gvec, 3) &&
           same(P.bvec, p->bvec_all, 3 * (size_t)N);
}

// Build (or reuse) the inward paths for this direction set and field geometry.
int ensure_paths(Ctx &c, const portal_se_params *p, portal_stats *st) {
    PathSet &P = c.paths;
    const bool star = p->star_mod == 1;
    const int N = c.g.N, D = star ? p->nth * p->nph : p->n_sphere;
    if (paths_match(c, p)) return PORTAL_OK;
    P.valid = false;
    const int nw = (int)c.g.work.size();
    // the reference's node loop would do nothing; here a grid without a node to solve
    // is refused before any launch
    if (nw == 0) return fail(c, PORTAL_ERR_ARG, "no bulk node has a non-zero abundance: nothing to solve");
    double *d_xth = nullptr, *d_bvec = nullptr, *d_gvec = nullptr;
    int *d_nvar = nullptr;
    dfree(P.d_xthn);
    dfree(P.d_nstar);
    P.star = 0;
    if (star) {
        // trace_node's star branch: every node's own direction set (main_sub.f90:56-79)
        std::vector<double> xthn;
        std::vector<int> nvar, nstar;
        const int bad = build_star_dirs(c.g.work, c.g.XYZ.data(), p->bvec_all, p->gvec, p->nth, p->nph, p->X_star,
                                        p->R_star, xthn, nvar, nstar);
        if (bad >= 0)
            return fail(c, PORTAL_ERR_ARG, "star mode: the star_all rays of node " + std::to_string(c.g.work[bad] + 1) +
                                               " do not fit in nth*nph (is it inside the star?)");
        HIP_TRY(&c, upload(P.d_xthn, xthn));
        HIP_TRY(&c, upload(P.d_nstar, nstar));
        HIP_TRY(&c, upload(d_nvar, nvar));
        P.star = 1;
        P.nth = p->nth;
        P.nph = p->nph;
        P.R_star = p->R_star;
        P.X_star.assign(p->X_star, p->X_star + 3);
    }
    P.xth.assign(p->xth, p->xth + 3 * (size_t)p->n_sphere);
    P.gvec.assign(p->gvec, p->gvec + 3);
    P.bvec.assign(p->bvec_all, p->bvec_all + 3 * (size_t)N);
    HIP_TRY(&c, upload(d_xth, P.xth));
    HIP_TRY(&c, upload(d_gvec, P.gvec));
    HIP_TRY(&c, upload(d_bvec, P.bvec));
    dfree(P.counts);
    dfree(P.offsets);
    const long long n = (long long)nw * D;
    HIP_TRY(&c, hipMalloc((void **)&P.counts, sizeof(int) * std::max<long long>(n, 1)));
    HIP_TRY(&c, hipMalloc((void **)&P.offsets, sizeof(long long) * (n + 1)));
    HIP_TRY(&c, hipMemsetAsync(c.d_flags, 0, sizeof(int) * 4, c.stream));
    Timer tm(c.stream);
    WalkArgs a{};
    a.n_work = nw;
    a.D = D;
    a.N = N;
    a.work = c.g.d_work;
    a.pos = c.g.d_pos;
    a.nbr_ptr = c.g.d_nbr_ptr;
    a.nbr_idx = c.g.d_nbr_idx;
    a.xth = d_xth;
    a.bvec = d_bvec;
    a.gvec = d_gvec;
    a.velb = c.g.d_velb;
    a.dx = p->dx;
    a.n_depth = p->n_depth;
    a.counts = P.counts;
    a.flags = c.d_flags;
    a.sample = 1;
    if (star) {
        a.xthn = P.d_xthn;
        a.nvar = d_nvar;
        a.nstar = P.d_nstar;
        a.Xs0 = p->X_star[0];
        a.Xs1 = p->X_star[1];
        a.Xs2 = p->X_star[2];
        a.R_star = p->R_star;
    }
    // One walk writes the entries into chunks taken from a counter, and a compaction
    // lays them out contiguously per path.  The chunk store is sized from a sampled count
    // over 4,096 to 8,192 work items (+10% and a margin); if it still overflows, or cannot
    // be allocated, the entries are written by a second walk at the exact offsets.
    bool chunked = false;
    long long n_chunks = 0;
    unsigned long long *d_next = nullptr;
    // (at most 8,192 work items: the two-pass walk, which costs no more there)
    if (n > 0 && n < (1LL << 31) && c.walk_chunk_scale > 0 && nw > 8192) {
        a.sample = nw / 4096;
        const long long ns = (long long)((nw + a.sample - 1) / a.sample) * D;
        HIP_TRY(&c, launch_walk(a, WALK_COUNT, c.stream));
        std::vector<int> cs(ns);
        HIP_TRY(&c, hipMemcpyAsync(cs.data(), P.counts, sizeof(int) * ns, hipMemcpyDeviceToHost, c.stream));
        HIP_TRY(&c, hipStreamSynchronize(c.stream));
        a.sample = 1;
        double per = 0;
        for (int v : cs) per += (v + WALK_CHUNK - 1) / WALK_CHUNK;
        per = per / (double)std::max<long long>(ns, 1);
        const long long need = (long long)(c.walk_chunk_scale * per * (double)n) + 4096;
        if (need > P.cap_chunks) {
            dfree(P.chunks);
            dfree(P.chunk_meta);
            P.cap_chunks = 0;
            if (hipMalloc((void **)&P.chunks, sizeof(PathEntry) * WALK_CHUNK * (size_t)need) == hipSuccess &&
                hipMalloc((void **)&P.chunk_meta, sizeof(int2) * (size_t)need) == hipSuccess) {
                P.cap_chunks = need;
            } else {
                (void)hipGetLastError();
                dfree(P.chunks);
                dfree(P.chunk_meta);
            }
        }
        if (P.cap_chunks > 0) {
            HIP_TRY(&c, hipMalloc((void **)&d_next, sizeof(unsigned long long)));
            HIP_TRY(&c, hipMemsetAsync(d_next, 0, sizeof(unsigned long long), c.stream));
            HIP_TRY(&c, hipMemsetAsync(c.d_flags, 0, sizeof(int) * 4, c.stream));
            a.chunks = P.chunks;
            a.chunk_meta = P.chunk_meta;
            a.chunk_next = d_next;
            a.chunk_cap = P.cap_chunks;
            HIP_TRY(&c, launch_walk(a, WALK_CHUNKED, c.stream));
            chunked = true;
        }
    }
    if (!chunked) {
        HIP_TRY(&c, hipMemsetAsync(c.d_flags, 0, sizeof(int) * 4, c.stream));
        HIP_TRY(&c, launch_walk(a, WALK_COUNT, c.stream));
    }
    int fl[2] = {0, 0};
    unsigned long long used = 0;
    HIP_TRY(&c, hipMemcpyAsync(fl, c.d_flags, sizeof(fl), hipMemcpyDeviceToHost, c.stream));
    if (d_next) HIP_TRY(&c, hipMemcpyAsync(&used, d_next, sizeof(used), hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    dfree(d_next);
    const int flags = fl[0];
    P.max_nc = fl[1];
    if (flags & FLAG_DEPTH) {
        dfree(d_xth);
        dfree(d_gvec);
        dfree(d_bvec);
        dfree(d_nvar);
        return fail(c, PORTAL_ERR_DEPTH, "an inward ray crosses more than n_depth cells (inverse_rt.f90:206-210)");
    }
    if (chunked && (flags & FLAG_CHUNKS)) chunked = false;  // overflow: write by a second walk
    n_chunks = (long long)used;
    long long total = 0;
    if (n > 0) HIP_TRY(&c, launch_scan_counts(P.counts, P.offsets, n, c.stream, &total));
    if ((size_t)total > P.cap_entries) {
        dfree(P.entries);
        HIP_TRY(&c, hipMalloc((void **)&P.entries, sizeof(PathEntry) * std::max<long long>(total, 1)));
        P.cap_entries = (size_t)total;
    }
    a.offsets = P.offsets;
    a.paths = P.entries;
    if (chunked)
        HIP_TRY(&c, launch_walk_compact(a, n_chunks, c.stream));
    else
        HIP_TRY(&c, launch_walk(a, WALK_WRITE, c.stream));
    P.single_pass = chunked;
    const double ms = tm.stop();
    // the chunk store (about 1.4x the paths) is not kept between builds
    dfree(P.chunks);
    dfree(P.chunk_meta);
    P.cap_chunks = 0;
    dfree(d_xth);
    dfree(d_gvec);
    dfree(d_bvec);
    dfree(d_nvar);
    P.total = total;
    P.n_work = nw;
    P.D = D;
    P.dx = p->dx;
    P.n_depth = p->n_depth;
    P.valid = true;
    if (st) st->t_walk_ms += ms;
    return PORTAL_OK;
}

// Every pointer the inward-integration kernels dereference must be set.
int check_rt_args(Ctx &c, const RtArgs &a, bool mixed) {
    const void *need[] = {a.work, a.velb, a.inv_b, a.xth, a.offsets, a.paths, a.I_in, a.J0, a.J2,
                          mixed ? (const void *)a.KS32 : (const void *)a.KS};
    for (const void *p : need)
        if (!p) return fail(c, PORTAL_ERR_STATE, "inward integration launched with an unset buffer");
    if (a.dust && !(mixed ? (const void *)a.KSD32 : (const void *)a.KSD))
        return fail(c, PORTAL_ERR_STATE, "inward integration with dust launched without dust coefficients");
    if (a.n_work < 0 || a.D <= 0 || a.R <= 0 || a.R > 64 || a.C <= 0 || a.max_nc < 0)
        return fail(c, PORTAL_ERR_ARG, "inward integration launched with bad sizes");
    return PORTAL_OK;
}

// Autotuner of the inward-integration launch: on the first pass of a geometry the
// candidate (kernel variant, waves per workgroup) pairs are timed on a slice of the
// work list and the fastest is kept for every later pass with the same shape.
// inward-integration launch shapes the autotuner times: (variant, waves per workgroup);
// variant 3 is the cooperative kernel taking the cells of a ray back to front.  The
// staged and global kernels are timed only when no cooperative shape fits the launch.
constexpr int N_RT_CANDS = 11;
constexpr struct RtCand { int variant, waves; } RT_CANDS[N_RT_CANDS] = {
    {2, 2}, {2, 4}, {2, 5}, {2, 8}, {3, 2}, {3, 4}, {3, 5}, {3, 8}, {1, 4}, {1, 2}, {0, 4}};
static_assert(N_RT_CANDS <= RT_TUNE_SLOTS, "tune_ms holds every candidate");

// the autotuner's key: the paths and the launch's shape
int rt_tune_key(const Ctx &c, const RtArgs &ra) {
    return (int)((c.paths.total % 1000003) ^ (ra.C * 131 + ra.R * 7 + ra.D)) * 4 + ra.dust + (ra.xthn ? 2 : 0);
}

// a launch shape is chosen for these paths and this shape (forced, or timed earlier)
bool rt_tuned(const Ctx &c, const RtArgs &ra, bool mixed) {
    const int pi = mixed ? 1 : 0;
    return c.forced_variant[pi] >= 0 || c.tuned_key[pi] == rt_tune_key(c, ra);
}

int autotune_rt(Ctx &c, RtArgs ra, bool mixed) {
    const int pi = mixed ? 1 : 0;
    const int key = rt_tune_key(c, ra);
    if (c.forced_variant[pi] >= 0) {
        c.tuned_variant[pi] = c.forced_variant[pi];
        c.tuned_waves[pi] = c.forced_waves[pi];
        c.tuned_key[pi] = 0;
        return PORTAL_OK;
    }
    if (c.tuned_key[pi] == key) return PORTAL_OK;
    constexpr int NC = N_RT_CANDS;
    const RtCand *cands = RT_CANDS;
    // the sample: one launch over 16 runs of 1,024 consecutive items spread evenly over the
    // Morton-ordered work list (the whole list when it is shorter).  The runs keep the
    // locality of a full pass (neighbouring nodes' rays share cells in L2), the spread lets
    // the cloud's centre and its edge both weigh in, and 16,384 items are several times the
    // workgroups the device keeps resident, so the timing is one of throughput, not of one
    // workgroup's latency.  The first 2,048 items alone, or a sample of every 122nd item,
    // chose kernels that ran the 10^6-node cloud 10 to 25% slower than this one's choice.
    const int n_all = ra.n_work, run = 1024, n_runs = 16;
    const bool whole = n_all <= run * n_runs;
    auto launch_slice = [&](int n, bool sample) {
        RtArgs rs = ra;
        rs.n_work = n;
        if (sample) {
            rs.item_run = run;
            rs.item_stride = n_all / n_runs;
        }
        return mixed ? launch_rt_mixed(rs, c.stream) : launch_rt_f64(rs, c.stream);
    };
    double best = 1e30;
    int bv = 0, bw = 4;
    for (int k = 0; k < NC; k++) {
        // FP64 keeps each ray front to back: its candidates then give the same bits, so the
        // timed choice cannot change an FP64 result (back to front stays a forced option)
        if ((!mixed && cands[k].variant == 3) || (cands[k].variant < 2 && best < 1e30)) {
            c.tune_ms[pi][k] = -1.0;
            continue;
        }
        ra.variant = cands[k].variant;
        ra.waves = cands[k].waves;
        // an untimed launch first: a kernel's first launch also loads its code
        hipError_t e = launch_slice(std::min(n_all, 64), false);
        double ms = 0.0;
        if (e == hipSuccess) {
            Timer t(c.stream);
            e = whole ? launch_slice(n_all, false) : launch_slice(run * n_runs, true);
            ms = t.stop();
        }
        c.tune_ms[pi][k] = e == hipSuccess ? ms : -1.0;
        if (e != hipSuccess) {
            (void)hipGetLastError();  // an LDS budget the variant cannot meet: skip it
            continue;
        }
        if (ms < best) {
            best = ms;
            bv = cands[k].variant;
            bw = cands[k].waves;
        }
    }
    c.tuned_key[pi] = key;
    c.tuned_variant[pi] = bv;
    c.tuned_waves[pi] = bw;
    return PORTAL_OK;
}

// One inward pass over built paths: the line coefficients, J0/J2 cleared, the
// integration launch.  xth (3 x n_sphere) and I_star (star mode: r_num) are on the
// device; J0/J2 (r_num x node_num) are the outputs.  may_alloc = false is the
// allocation-free form: no buffer is created (the dust coefficients must exist), no
// launch shape is timed (one must have been chosen for these paths), no event is
// recorded (st must be null).
int inward_pass(Ctx &c, const portal_se_params *p, const double *xth, const double *I_star, double *J0, double *J2,
                bool may_alloc, portal_stats *st) {
    const Species &s = c.sp;
    PrepArgs pa{};
    pa.N = c.g.N;
    pa.R = s.r_num;
    pa.L = s.e_num;
    pa.levp = c.d_levp;
    pa.velb = c.g.d_velb;
    pa.l1 = s.d_l1;
    pa.l2 = s.d_l2;
    pa.c1B = s.d_c1B;
    pa.c2A = s.d_c2A;
    pa.g1 = s.d_g1;
    pa.g2 = s.d_g2;
    pa.cnu_spi = s.d_cnu_spi;
    const bool mixed = p->precision == PORTAL_MIXED;
    pa.KS = mixed ? nullptr : c.d_KS;
    pa.KS32 = mixed ? c.d_KS32 : nullptr;
    pa.I_in = s.d_Iin;
    // dust enters the inward pass when some node has dust and some line a dust opacity
    bool dust = false;
    if (c.g.has_dust)
        for (double kd : s.kd_v) dust = dust || kd != 0.0;
    if (dust) {
        if (!may_alloc && (!c.d_KSD || !c.d_KSD32))
            return fail(c, PORTAL_ERR_STATE, "dust coefficients not set up: run the allocating pass once first");
        if (!c.d_KSD) HIP_TRY(&c, hipMalloc((void **)&c.d_KSD, sizeof(double) * 2 * c.g.N * s.r_num));
        if (!c.d_KSD32) HIP_TRY(&c, hipMalloc((void **)&c.d_KSD32, sizeof(float) * 2 * c.g.N * s.r_num));
        pa.dust = c.g.d_dust;
        pa.tdust = c.g.d_tdust;
        pa.kd = s.d_kd;
        pa.nu = s.d_nu;
        pa.KSD = mixed ? nullptr : c.d_KSD;
        pa.KSD32 = mixed ? c.d_KSD32 : nullptr;
    }
    double t_prep = 0.0;
    {
        std::unique_ptr<Timer> tp(st ? new Timer(c.stream) : nullptr);
        HIP_TRY(&c, launch_line_prep(pa, c.stream));
        if (tp) t_prep = tp->stop();
    }
    HIP_TRY(&c, hipMemsetAsync(J0, 0, sizeof(double) * c.g.N * s.r_num, c.stream));
    HIP_TRY(&c, hipMemsetAsync(J2, 0, sizeof(double) * c.g.N * s.r_num, c.stream));
    RtArgs ra{};
    ra.n_work = (int)c.g.work.size();
    ra.D = c.paths.D;
    ra.R = s.r_num;
    ra.nvel = p->nvel;
    ra.C = 2 * p->nvel + 1;
    ra.work = c.g.d_work;
    ra.velb = c.g.d_velb;
    ra.inv_b = c.g.d_inv_b;
    ra.b_ratio = c.g.b_ratio;
    ra.KS = c.d_KS;
    ra.KS32 = c.d_KS32;
    ra.xth = xth;
    ra.offsets = c.paths.offsets;
    ra.paths = c.paths.entries;
    ra.I_in = s.d_Iin;
    ra.dl_scale = c.g.tx * p->dx;
    ra.J0 = J0;
    ra.J2 = J2;
    ra.max_nc = (c.paths.max_nc + 7) & ~7;
    ra.dust = dust ? 1 : 0;
    ra.KSD = c.d_KSD;
    ra.KSD32 = c.d_KSD32;
    if (c.paths.star) {
        if (!I_star) return fail(c, PORTAL_ERR_STATE, "star-mode paths without the star's intensities");
        ra.xthn = c.paths.d_xthn;
        ra.nstar = c.paths.d_nstar;
        ra.I_star = I_star;
    }
    int rc = check_rt_args(c, ra, mixed);
    if (rc) return rc;
    ra.unit = c.coop_unit;
    if (may_alloc)
        autotune_rt(c, ra, mixed);
    else if (!rt_tuned(c, ra, mixed))
        return fail(c, PORTAL_ERR_STATE, "no launch shape chosen for these paths: run the allocating pass once first");
    ra.variant = c.tuned_variant[mixed ? 1 : 0];
    ra.waves = c.tuned_waves[mixed ? 1 : 0];
    if (c.ali) {
        // the own-cell operator comes from the cooperative kernel: its fastest timed shape
        ra.LAM = c.d_LAM;
        if (ra.variant != 2 && ra.variant != 3) {
            const int pi = mixed ? 1 : 0;
            int best = -1;
            for (int k = 0; k < N_RT_CANDS; k++)
                if ((RT_CANDS[k].variant == 2 || RT_CANDS[k].variant == 3) && c.tune_ms[pi][k] > 0 &&
                    (best < 0 || c.tune_ms[pi][k] < c.tune_ms[pi][best]))
                    best = k;
            ra.variant = best < 0 ? 2 : RT_CANDS[best].variant;
            ra.waves = best < 0 ? 4 : RT_CANDS[best].waves;
        }
    }
    double t_rt = 0.0;
    hipError_t e;
    {
        std::unique_ptr<Timer> tr(st ? new Timer(c.stream) : nullptr);
        e = mixed ? launch_rt_mixed(ra, c.stream) : launch_rt_f64(ra, c.stream);
        if (tr) t_rt = tr->stop();
    }
    HIP_TRY(&c, e);
    if (st) {
        st->t_prep_ms += t_prep;
        st->t_rt_ms += t_rt;
        st->rt_kernel_avg_ms = t_rt;
        st->n_cross = c.paths.total;
        st->n_paths = (long long)c.paths.n_work * c.paths.D;
        st->rt_variant = ra.variant;
        st->rt_waves = ra.waves;
        st->rt_order = ra.variant == 3 ? 1 : 0;
        st->max_path = c.paths.max_nc;
    }
    return PORTAL_OK;
}

// B_nu(T_star) per line: the incoming intensity of rays towards the star (inverse_rt.f90:113-117)
void star_intensities(const Species &s, double T_star, double *out) {
    for (int k = 0; k < s.r_num; k++) out[k] = black_body(T_star, s.nu_hz[k]);
}

int run_inward(Ctx &c, const portal_se_params *p, portal_stats *st) {
    c.pol_valid = false;
    int rc = ensure_paths(c, p, st);
    if (rc) return rc;
    double *d_xth = nullptr, *d_Istar = nullptr;
    Defer release{[&] {
        dfree(d_xth);
        dfree(d_Istar);
    }};
    std::vector<double> xv(p->xth, p->xth + 3 * (size_t)p->n_sphere);
    HIP_TRY(&c, upload(d_xth, xv));
    if (c.paths.star) {
        std::vector<double> Is(c.sp.r_num);
        star_intensities(c.sp, p->T_star, Is.data());
        HIP_TRY(&c, upload(d_Istar, Is));
    }
    return inward_pass(c, p, d_xth, d_Istar, c.d_J0, c.d_J2, true, st);
}

int run_se(Ctx &c, int extra_info, portal_stats *st) {
    const Species &s = c.sp;
    HIP_TRY(&c, hipMemsetAsync(c.d_flags, 0, sizeof(int) * 4, c.stream));
    SeArgs a{};
    a.n_work = (int)c.g.work.size();
    a.N = c.g.N;
    a.L = s.e_num;
    a.R = s.r_num;
    a.c_dim = s.c_dim;
    a.n_trans = s.n_trans;
    a.n_ent = (int)s.se.ent_idx.size();
    a.work = c.g.d_work;
    a.J0 = c.d_J0;
    a.J2 = c.d_J2;
    a.tkin = c.g.d_tkin;
    a.h2 = c.g.d_h2;
    a.t_A = s.d_tA;
    a.t_B21 = s.d_tB21;
    a.t_nu = s.d_tnu;
    a.t_gul = s.d_tgul;
    a.t_rad = s.d_trad;
    a.t_col = s.d_tcol;
    a.t_Cji = s.d_Cji;
    a.col_temps = s.d_coltemps;
    a.c_temp = s.c_temp;
    a.c_par = s.c_par;
    a.c_num = s.c_num;
    a.ent_idx = s.se.d_ent_idx;
    a.ent_ptr = s.se.d_ent_ptr;
    a.term_q = s.se.d_term_q;
    a.term_class = s.se.d_term_class;
    a.term_coef = s.se.d_term_coef;
    a.count_array = s.d_count;
    a.j_lev = s.d_jlev;
    a.t_lo = s.d_tlo;
    a.t_up = s.d_tup;
    if (c.ali) {
        a.lam = c.d_LAM;
        a.levp = c.d_levp;
    }
    a.depol = s.d_depol;
    a.pol = c.d_pol;
    a.rates = extra_info ? c.d_rates : nullptr;
    a.flags = c.d_flags;
    a.extra_info = extra_info;
    a.waves = c.se_waves;
    HIP_TRY(&c, hipMemsetAsync(c.d_pol, 0, sizeof(double) * c.g.N * 2 * s.e_num, c.stream));
    if (extra_info) HIP_TRY(&c, hipMemsetAsync(c.d_rates, 0, sizeof(double) * c.g.N * 6 * s.e_num, c.stream));
    // a system whose matrix, copy and scalars do not fit one workgroup's LDS takes the
    // batch's slab form (copy and scalars in a global slab per workgroup) as a batch of one
    const bool slab_form = se_lds_bytes(a.c_dim, a.n_trans, true) > 160 * 1024 || c.se_force_slab > 0;
    const bool all_global = c.se_force_slab == 2;
    SeArgs *d_a = nullptr;
    int2 *d_items = nullptr;
    int *d_cnt = nullptr;
    double *d_slabs = nullptr;
    Defer slab_at_exit{[&] {
        dfree(d_a);
        dfree(d_items);
        dfree(d_cnt);
        dfree(d_slabs);
    }};
    int blocks = 0;
    if (slab_form) {
        const int nw = (int)c.g.work.size();
        int n_cu = 256;
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c.device);
        a.waves = c.se_forced_waves > 0 ? c.se_forced_waves : SE_SLAB_WAVES;
        const int bpc = se_slab_blocks_per_cu(a.waves, a.c_dim, a.n_trans, all_global);
        if (bpc < 1) return fail(c, PORTAL_ERR_ARG, "SE system too large for one workgroup's LDS");
        blocks = n_cu * bpc;
        std::vector<SeArgs> av(1, a);
        std::vector<int2> items(std::max(nw, 1));
        for (int w = 0; w < nw; w++) items[w] = make_int2(0, w);
        HIP_TRY(&c, upload(d_a, av));
        HIP_TRY(&c, upload(d_items, items));
        HIP_TRY(&c, hipMalloc((void **)&d_cnt, sizeof(int)));
        HIP_TRY(&c, hipMalloc((void **)&d_slabs,
                              sizeof(double) * (size_t)blocks * se_slab_doubles(a.c_dim, a.n_trans, all_global)));
        HIP_TRY(&c, launch_fill_zero(d_cnt, sizeof(int), c.stream));
    }
    Timer t(c.stream);
    if (slab_form)
        HIP_TRY(&c, launch_se_slab(d_a, d_items, (int)c.g.work.size(), d_cnt, a.c_dim, a.n_trans, a.waves, blocks,
                                   d_slabs, c.stream, all_global));
    else
        HIP_TRY(&c, launch_se(a, c.stream));
    const double ms = t.stop();
    int flags = 0;
    HIP_TRY(&c, hipMemcpy(&flags, c.d_flags, sizeof(int), hipMemcpyDeviceToHost));
    if (st) {
        st->t_se_ms += ms;
        st->se_form = 0;
        st->se_waves = a.waves;
    }
    if (flags & FLAG_TEMP)
        return fail(c, PORTAL_ERR_TEMP, "kinetic temperature above the collision-rate table (miscellaneous.f90:331-360)");
    if (flags & FLAG_LU) return fail(c, PORTAL_ERR_LU, "Density-matrix calculation: failed (singular SE matrix)");
    c.pol_valid = true;
    return PORTAL_OK;
}

// the arguments alone (no device state)
int check_se_args(Ctx &c, const portal_se_params *p) {
    if (!p || !p->xth || !p->gvec || !p->bvec_all) return fail(c, PORTAL_ERR_ARG, "null parameter");
    if (p->star_mod != 0 && p->star_mod != 1) return fail(c, PORTAL_ERR_ARG, "star_mod must be 0 or 1");
    if (p->star_mod == 1 && (p->nth <= 0 || p->nph <= 0 || !p->X_star || !(p->R_star > 0) || !(p->T_star >= 0)))
        return fail(c, PORTAL_ERR_ARG, "star mode needs nth, nph > 0, X_star, R_star > 0 and T_star >= 0");
    if (p->n_sphere <= 0 || p->nvel <= 0 || p->n_depth <= 0 || !(p->dx > 0))
        return fail(c, PORTAL_ERR_ARG, "n_sphere, nvel, n_depth and dx must be positive");
    if (p->precision != PORTAL_FP64 && p->precision != PORTAL_MIXED) return fail(c, PORTAL_ERR_ARG, "bad precision");
    return PORTAL_OK;
}

int check_se_params(Ctx &c, const portal_se_params *p) {
    const int rc = check_se_args(c, p);
    return rc ? rc : ensure_node_buffers(c);
}

}  // namespace

extern "C" {

int portal_ctx_create(portal_ctx **out, int device) {
    if (!out) return PORTAL_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PORTAL_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PORTAL_ERR_DEVICE;
    auto *x = new portal_ctx();
    x->c.device = device;
    if (hipStreamCreate(&x->c.stream) != hipSuccess) {
        delete x;
        return PORTAL_ERR_DEVICE;
    }
    // the allocation-free entry point's scratch; only that entry point fails without it
    // (it reports this message)
    if (resolve_scratch(x->c) != PORTAL_OK) x->c.scratch_err = x->c.err;
    *out = x;
    return PORTAL_OK;
}


int portal_ctx_destroy(portal_ctx *x) {
    if (!x) return PORTAL_OK;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    hipStreamSynchronize(c.stream);
    portal_mlp_free(c);
    Grid &g = c.g;
    dfree(g.d_pos);
    dfree(g.d_velb);
    dfree(g.d_inv_b);
    dfree(g.d_tdust);
    dfree(g.d_tkin);
    dfree(g.d_h2);
    dfree(g.d_dust);
    dfree(g.d_ab);
    dfree(g.d_nbr_ptr);
    dfree(g.d_nbr_idx);
    dfree(g.d_work);
    Species &s = c.sp;
    dfree(s.d_l1);
    dfree(s.d_l2);
    dfree(s.d_c1B);
    dfree(s.d_c2A);
    dfree(s.d_g1);
    dfree(s.d_g2);
    dfree(s.d_cnu_spi);
    dfree(s.d_kd);
    dfree(s.d_nu);
    dfree(s.d_Iin);
    dfree(s.d_tA);
    dfree(s.d_tB21);
    dfree(s.d_tnu);
    dfree(s.d_tgul);
    dfree(s.d_trad);
    dfree(s.d_tcol);
    dfree(s.d_tlo);
    dfree(s.d_tup);
    dfree(s.d_Cji);
    dfree(s.d_coltemps);
    dfree(s.d_count);
    dfree(s.d_jlev);
    dfree(s.d_depol);
    dfree(s.se.d_ent_idx);
    dfree(s.se.d_ent_ptr);
    dfree(s.se.d_term_q);
    dfree(s.se.d_term_class);
    dfree(s.se.d_term_coef);
    dfree(c.paths.entries);
    dfree(c.paths.offsets);
    dfree(c.paths.counts);
    dfree(c.paths.d_xthn);
    dfree(c.paths.d_nstar);
    dfree(c.paths.chunks);
    dfree(c.paths.chunk_meta);
    dfree(c.d_levp);
    dfree(c.d_KS);
    dfree(c.d_KS32);
    dfree(c.d_KSD);
    dfree(c.d_KSD32);
    dfree(c.d_dustco);
    dfree(c.d_J0);
    dfree(c.d_J2);
    dfree(c.d_LAM);
    dfree(c.d_pol);
    dfree(c.d_rates);
    dfree(c.d_flags);
    if (c.d_img) hipFree(c.d_img);
    hipStreamDestroy(c.stream);
    delete x;
    return PORTAL_OK;
}

const char *portal_last_error(const portal_ctx *x) { return x ? x->c.err.c_str() : "null context"; }

int portal_upload_grid(portal_ctx *x, int N, int N_b, const double *XYZ, const double *vel, const double *temp,
                       const double *b_all, const double *h2, const double *ab, const double *dust,
                       const int *del_ne, const int *del_num_ne, int maxne, double tx) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    c.pol_chain = false;
    c.pol_valid = false;
    hipSetDevice(c.device);
    if (N <= 0 || N_b < 0 || N_b > N || maxne <= 0 || !XYZ || !vel || !temp || !b_all || !h2 || !ab || !dust ||
        !del_ne || !del_num_ne || !(tx > 0))
        return fail(c, PORTAL_ERR_ARG, "bad grid arguments");
    Grid &g = c.g;
    g.valid = false;
    c.paths.valid = false;
    g.N = N;
    g.N_b = N_b;
    g.tx = tx;
    g.XYZ.assign(XYZ, XYZ + 3 * (size_t)N);
    g.vel.assign(vel, vel + 3 * (size_t)N);
    g.temp.assign(temp, temp + 2 * (size_t)N);
    g.b_all.assign(b_all, b_all + N);
    g.h2.assign(h2, h2 + N);
    g.ab.assign(ab, ab + N);
    g.dust.assign(dust, dust + N);
    g.has_dust = false;
    for (int i = 0; i < N; i++)
        if (dust[i] != 0.0) g.has_dust = true;
    // CSR adjacency from del_ne(N, maxne) (column-major, 1-based, zero padded)
    g.nbr_ptr.assign(N + 1, 0);
    g.nbr_idx.clear();
    for (int i = 0; i < N; i++) {
        const int k = del_num_ne[i];
        if (k < 0 || k > maxne) return fail(c, PORTAL_ERR_ARG, "del_num_ne out of range");
        for (int j = 0; j < k; j++) {
            const int nb = del_ne[i + (size_t)N * j];
            if (nb < 1 || nb > N) return fail(c, PORTAL_ERR_ARG, "del_ne entry out of range");
            g.nbr_idx.push_back(nb - 1);
        }
        g.nbr_ptr[i + 1] = (int)g.nbr_idx.size();
    }
    std::vector<double4> pos(N), velb(N);
    std::vector<double> invb(N), tdust(N), tkin(N);
    double b_lo = 0.0, b_hi = 0.0;
    for (int i = 0; i < N; i++) {
        pos[i] = make_double4(XYZ[3 * (size_t)i], XYZ[3 * (size_t)i + 1], XYZ[3 * (size_t)i + 2], 0.0);
        velb[i] = make_double4(vel[3 * (size_t)i], vel[3 * (size_t)i + 1], vel[3 * (size_t)i + 2], b_all[i]);
        invb[i] = 1.0 / b_all[i];
        tkin[i] = temp[2 * (size_t)i];
        tdust[i] = temp[2 * (size_t)i + 1];
        const double b = b_all[i];
        if (b > 0.0 && std::isfinite(b)) {
            b_lo = b_lo > 0.0 ? std::min(b_lo, b) : b;
            b_hi = std::max(b_hi, b);
        }
    }
    g.b_ratio = b_lo > 0.0 ? b_hi / b_lo : 0.0;
    // work list: nodes 1..bulk_node_num with ab_all > 0, ordered along a Morton curve
    std::vector<std::pair<uint64_t, int>> key;
    for (int i = 0; i < N_b; i++)
        if (ab[i] > 0.0) {
            auto q = [](double v) { return (uint32_t)std::min(2097151.0, std::max(0.0, (v + 1.0) * 0.5 * 2097151.0)); };
            const uint64_t m = spread3(q(XYZ[3 * (size_t)i])) | spread3(q(XYZ[3 * (size_t)i + 1])) << 1 |
                               spread3(q(XYZ[3 * (size_t)i + 2])) << 2;
            key.push_back({m, i});
        }
    std::sort(key.begin(), key.end());
    g.work.resize(key.size());
    for (size_t i = 0; i < key.size(); i++) g.work[i] = key[i].second;
    HIP_TRY(&c, upload(g.d_pos, pos));
    HIP_TRY(&c, upload(g.d_velb, velb));
    HIP_TRY(&c, upload(g.d_inv_b, invb));
    HIP_TRY(&c, upload(g.d_tdust, tdust));
    HIP_TRY(&c, upload(g.d_tkin, tkin));
    HIP_TRY(&c, upload(g.d_h2, g.h2));
    HIP_TRY(&c, upload(g.d_dust, g.dust));
    HIP_TRY(&c, upload(g.d_ab, g.ab));
    HIP_TRY(&c, upload(g.d_nbr_ptr, g.nbr_ptr));
    HIP_TRY(&c, upload(g.d_nbr_idx, g.nbr_idx));
    HIP_TRY(&c, upload(g.d_work, g.work));
    dfree(c.d_levp);
    c.cap_nodes = 0;
    g.valid = true;
    return PORTAL_OK;
}

int portal_upload_species(portal_ctx *x, int e_num, const int *j_lev, const double *e_levs_hz, int r_num,
                          const int *r_trans, const double *t_Aji, const double *nu_hz, const double *t_Bij,
                          const double *kd_v, int c_num, const int *c_trans, int c_temp, int c_par,
                          const double *t_Cji, const double *col_temps, int kmax) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    c.pol_chain = false;
    c.pol_valid = false;
    hipSetDevice(c.device);
    if (e_num <= 0 || r_num <= 0 || c_num < 0 || c_temp <= 0 || (c_par != 1 && c_par != 2) || kmax < 2 ||
        (kmax & 1) || !j_lev || !e_levs_hz || !r_trans || !t_Aji || !nu_hz || !t_Bij || !kd_v ||
        (c_num > 0 && (!c_trans || !t_Cji)) || !col_temps)
        return fail(c, PORTAL_ERR_ARG, "bad species arguments (kmax must be even and >= 2)");
    if (r_num > 64) return fail(c, PORTAL_ERR_ARG, "at most 64 radiative lines");
    Species &s = c.sp;
    s.valid = false;
    c.paths.valid = false;
    s.e_num = e_num;
    s.r_num = r_num;
    s.c_num = c_num;
    s.c_temp = c_temp;
    s.c_par = c_par;
    s.kmax = kmax;
    s.j_lev.assign(j_lev, j_lev + e_num);
    s.e_levs_hz.assign(e_levs_hz, e_levs_hz + e_num);
    s.r_trans.assign(r_trans, r_trans + 2 * (size_t)r_num);
    s.c_trans.assign(c_trans ? c_trans : r_trans, c_trans ? c_trans + 2 * (size_t)c_num : r_trans);
    s.t_Aji.assign(t_Aji, t_Aji + r_num);
    s.nu_hz.assign(nu_hz, nu_hz + r_num);
    s.t_Bij.assign(t_Bij, t_Bij + r_num);
    s.kd_v.assign(kd_v, kd_v + r_num);
    s.t_Cji.assign(t_Cji ? t_Cji : col_temps, t_Cji ? t_Cji + (size_t)c_temp * c_num * c_par : col_temps);
    s.col_temps.assign(col_temps, col_temps + c_temp);
    for (int i = 0; i < 2 * r_num; i++)
        if (s.r_trans[i] < 1 || s.r_trans[i] > e_num) return fail(c, PORTAL_ERR_ARG, "r_trans out of range");
    for (int i = 0; i < 2 * c_num; i++)
        if (s.c_trans[i] < 1 || s.c_trans[i] > e_num) return fail(c, PORTAL_ERR_ARG, "c_trans out of range");
    {
        // unknowns of the SE system (setup_count_array): K = 0, 2, .., min(2J, kmax) per level
        long long unknowns = 0;
        for (int i = 0; i < e_num; i++) unknowns += std::min(2 * std::max(s.j_lev[i], 0), kmax) / 2 + 1;
        if (unknowns > SE_MAX_UNKNOWNS)
            return fail(c, PORTAL_ERR_ARG, "SE system too large (more than 512 unknowns)");
    }
    build_species_tables(s);
    const int cd = s.c_dim;
    // larger systems take the slab form, whose LDS is a few vectors
    if (cd > SE_MAX_UNKNOWNS || se_lds_bytes(cd, s.n_trans, false) > 150 * 1024)
        return fail(c, PORTAL_ERR_ARG, "SE system too large (more than 512 unknowns)");
    // per radiative line
    std::vector<int> l1(r_num), l2(r_num);
    std::vector<double> c1B(r_num), c2A(r_num), g1(r_num), g2(r_num), cnu(r_num), Iin(r_num);
    const double spi = std::sqrt(std::acos(-1.0));
    for (int k = 0; k < r_num; k++) {
        l1[k] = s.r_trans[2 * k] - 1;
        l2[k] = s.r_trans[2 * k + 1] - 1;
        const double nu0 = s.nu_hz[k];
        c1B[k] = 2.2102191e-42 * nu0 * s.t_Bij[k];
        c2A[k] = 5.272859e-35 * nu0 * s.t_Aji[k];
        g1[k] = 1.0 * (2 * s.j_lev[l1[k]] + 1);
        g2[k] = 1.0 * (2 * s.j_lev[l2[k]] + 1);
        cnu[k] = (299792458.0 / nu0) / spi;
        Iin[k] = black_body(2.73, nu0);
    }
    // per unique transition
    const int nt = s.n_trans;
    std::vector<double> tA(nt), tB21(nt), tnu(nt), tgul(nt);
    std::vector<int> tlo(nt), tup(nt);
    for (int t = 0; t < nt; t++) {
        const int lo = s.all_trans[2 * t] - 1, up = s.all_trans[2 * t + 1] - 1;
        tlo[t] = lo;
        tup[t] = up;
        const double nu0 = s.e_levs_hz[up] - s.e_levs_hz[lo];
        const double A = s.trans_rad[t] >= 0 ? s.t_Aji[s.trans_rad[t]] : 0.0;
        tA[t] = A;
        tnu[t] = nu0;
        tB21[t] = 6.781962e49 * A / std::pow(nu0, 3.0);
        tgul[t] = (2.0 * s.j_lev[up] + 1.0) / (2.0 * s.j_lev[lo] + 1.0);
    }
    HIP_TRY(&c, upload(s.d_l1, l1));
    HIP_TRY(&c, upload(s.d_l2, l2));
    HIP_TRY(&c, upload(s.d_c1B, c1B));
    HIP_TRY(&c, upload(s.d_c2A, c2A));
    HIP_TRY(&c, upload(s.d_g1, g1));
    HIP_TRY(&c, upload(s.d_g2, g2));
    HIP_TRY(&c, upload(s.d_cnu_spi, cnu));
    HIP_TRY(&c, upload(s.d_kd, s.kd_v));
    HIP_TRY(&c, upload(s.d_nu, s.nu_hz));
    HIP_TRY(&c, upload(s.d_Iin, Iin));
    HIP_TRY(&c, upload(s.d_tA, tA));
    HIP_TRY(&c, upload(s.d_tB21, tB21));
    HIP_TRY(&c, upload(s.d_tnu, tnu));
    HIP_TRY(&c, upload(s.d_tgul, tgul));
    HIP_TRY(&c, upload(s.d_trad, s.trans_rad));
    HIP_TRY(&c, upload(s.d_tcol, s.trans_col));
    HIP_TRY(&c, upload(s.d_tlo, tlo));
    HIP_TRY(&c, upload(s.d_tup, tup));
    HIP_TRY(&c, upload(s.d_Cji, s.t_Cji));
    HIP_TRY(&c, upload(s.d_coltemps, s.col_temps));
    HIP_TRY(&c, upload(s.d_count, s.count_array));
    HIP_TRY(&c, upload(s.d_jlev, s.j_lev));
    HIP_TRY(&c, upload(s.se.d_ent_idx, s.se.ent_idx));
    HIP_TRY(&c, upload(s.se.d_ent_ptr, s.se.ent_ptr));
    HIP_TRY(&c, upload(s.se.d_term_q, s.se.term_q));
    HIP_TRY(&c, upload(s.se.d_term_class, s.se.term_class));
    HIP_TRY(&c, upload(s.se.d_term_coef, s.se.term_coef));
    dfree(c.d_levp);
    c.cap_nodes = 0;
    s.valid = true;
    return PORTAL_OK;
}

// Collisional depolarisation (depol.f90:242-359).  The reference computes D^(K) in
// de_polarization but trace_node passes D0 = 0 (main_sub.f90:96); here the table is an
// option of the uploaded species.
int portal_depol_table(const portal_ctx *x, double *D0_all) {
    if (!x || !D0_all) return PORTAL_ERR_ARG;
    const Species &s = x->c.sp;
    if (!s.valid) return PORTAL_ERR_STATE;
    std::vector<double> t;
    if (depol_table(s, t)) return PORTAL_ERR_ARG;
    std::copy(t.begin(), t.end(), D0_all);
    return PORTAL_OK;
}

int portal_set_depol(portal_ctx *x, const double *D0_all) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    Species &s = c.sp;
    if (!s.valid) return fail(c, PORTAL_ERR_STATE, "no species (portal_upload_species)");
    dfree(s.d_depol);
    if (!D0_all) return PORTAL_OK;
    if (s.kmax < 2) return fail(c, PORTAL_ERR_ARG, "depolarisation needs kmax >= 2");
    const size_t n = (size_t)s.c_temp * s.e_num * (s.kmax - 1);
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(D0_all[i])) return fail(c, PORTAL_ERR_ARG, "depolarisation table holds a non-finite value");
    std::vector<double> v(D0_all, D0_all + n);
    HIP_TRY(&c, upload(s.d_depol, v));
    return PORTAL_OK;
}

int portal_species_info(const portal_ctx *x, int *n_trans, int *c_dim) {
    if (!x || !x->c.sp.valid) return PORTAL_ERR_STATE;
    if (n_trans) *n_trans = x->c.sp.n_trans;
    if (c_dim) *c_dim = x->c.sp.c_dim;
    return PORTAL_OK;
}

int portal_set_levpop(portal_ctx *x, const double *lev_p) {
    if (!x || !lev_p) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    c.pol_chain = false;
    hipSetDevice(c.device);
    int rc = ensure_node_buffers(c);
    if (rc) return rc;
    HIP_TRY(&c, hipMemcpy(c.d_levp, lev_p, sizeof(double) * c.g.N * c.sp.e_num, hipMemcpyHostToDevice));
    return PORTAL_OK;
}

int portal_get_levpop(portal_ctx *x, double *lev_p) {
    if (!x || !lev_p) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    int rc = ensure_node_buffers(c);
    if (rc) return rc;
    HIP_TRY(&c, hipMemcpy(lev_p, c.d_levp, sizeof(double) * c.g.N * c.sp.e_num, hipMemcpyDeviceToHost));
    return PORTAL_OK;
}

int portal_inverse_rt(portal_ctx *x, const portal_se_params *p, double *J0, double *J2, portal_stats *st) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    int rc = check_se_params(c, p);
    if (rc) return rc;
    if (st) std::memset(st, 0, sizeof(*st));
    Timer tt(c.stream);
    rc = run_inward(c, p, st);
    if (rc) return rc;
    const size_t n = (size_t)c.g.N * c.sp.r_num;
    if (J0) HIP_TRY(&c, hipMemcpyAsync(J0, c.d_J0, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    if (J2) HIP_TRY(&c, hipMemcpyAsync(J2, c.d_J2, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    const double ms = tt.stop();
    if (st) st->t_total_ms = ms;
    return PORTAL_OK;
}

int portal_inverse_rt_device(portal_ctx *x, const portal_se_params *p, double *d_J0, double *d_J2) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    if (!d_J0 || !d_J2) return fail(c, PORTAL_ERR_ARG, "J0 and J2 must be device buffers of r_num x node_num");
    if (hipSetDevice(c.device) != hipSuccess) return fail(c, PORTAL_ERR_DEVICE, "hipSetDevice failed");
    int rc = check_se_args(c, p);
    if (rc) return rc;
    if (c.sp.valid && c.sp.r_num > 64) return fail(c, PORTAL_ERR_ARG, "more than 64 lines");
    if (p->n_sphere > NOALLOC_MAX_DIRS)
        return fail(c, PORTAL_ERR_ARG, "n_sphere = " + std::to_string(p->n_sphere) +
                                           " exceeds the allocation-free pass's scratch (at most " +
                                           std::to_string(NOALLOC_MAX_DIRS) + " directions)");
    if (!node_buffers_ready(c))
        return fail(c, PORTAL_ERR_STATE, "no pass buffers for this grid and species: call portal_inverse_rt "
                                         "(or portal_se_pass) once first");
    if (!paths_match(c, p))
        return fail(c, PORTAL_ERR_STATE, "no inward paths for these directions and this field: call "
                                         "portal_inverse_rt (or portal_se_pass) once with them first");
    unsigned char *scr = static_cast<unsigned char *>(g_scratch[c.device]);
    if (!scr) return fail(c, PORTAL_ERR_DEVICE, "scratch address not resolved for this device: " + c.scratch_err);
    c.pol_valid = false;
    double *d_xth = reinterpret_cast<double *>(scr);
    double *d_Istar = nullptr;
    // host copies into the scratch; the stack buffer outlives the copy (synchronised below)
    double Is[64];
    HIP_TRY(&c, hipMemcpyAsync(d_xth, p->xth, sizeof(double) * 3 * (size_t)p->n_sphere, hipMemcpyHostToDevice,
                               c.stream));
    if (c.paths.star) {
        star_intensities(c.sp, p->T_star, Is);
        d_Istar = reinterpret_cast<double *>(scr + NOALLOC_STAR_OFFSET);
        HIP_TRY(&c, hipMemcpyAsync(d_Istar, Is, sizeof(double) * c.sp.r_num, hipMemcpyHostToDevice, c.stream));
    }
    rc = inward_pass(c, p, d_xth, d_Istar, d_J0, d_J2, false, nullptr);
    const hipError_t e = hipStreamSynchronize(c.stream);
    if (rc) return rc;
    HIP_TRY(&c, e);
    return PORTAL_OK;
}

int portal_se_solve(portal_ctx *x, const double *J0, const double *J2, double *pol_pops, double *pol_rates,
                    portal_stats *st) {
    if (!x || !J0 || !J2) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    c.pol_chain = false;
    hipSetDevice(c.device);
    int rc = ensure_node_buffers(c);
    if (rc) return rc;
    if (st) std::memset(st, 0, sizeof(*st));
    const size_t n = (size_t)c.g.N * c.sp.r_num;
    HIP_TRY(&c, hipMemcpy(c.d_J0, J0, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_TRY(&c, hipMemcpy(c.d_J2, J2, sizeof(double) * n, hipMemcpyHostToDevice));
    rc = run_se(c, pol_rates != nullptr, st);
    if (rc) return rc;
    const size_t L = c.sp.e_num, N = c.g.N;
    if (pol_pops) HIP_TRY(&c, hipMemcpy(pol_pops, c.d_pol, sizeof(double) * N * 2 * L, hipMemcpyDeviceToHost));
    if (pol_rates) HIP_TRY(&c, hipMemcpy(pol_rates, c.d_rates, sizeof(double) * N * 6 * L, hipMemcpyDeviceToHost));
    return PORTAL_OK;
}

int portal_se_pass(portal_ctx *x, const portal_se_params *p, double *pol_pops, double *pol_rates, double *JJ_20,
                   portal_stats *st) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    c.pol_chain = false;
    hipSetDevice(c.device);
    int rc = check_se_params(c, p);
    if (rc) return rc;
    if (st) std::memset(st, 0, sizeof(*st));
    Timer tt(c.stream);
    rc = run_inward(c, p, st);
    if (rc) return rc;
    const int extra = p->extra_info == 1;
    rc = run_se(c, extra, st);
    if (rc) return rc;
    if (st) st->t_total_ms = tt.stop();
    const size_t L = c.sp.e_num, N = c.g.N, R = c.sp.r_num;
    if (pol_pops) HIP_TRY(&c, hipMemcpy(pol_pops, c.d_pol, sizeof(double) * N * 2 * L, hipMemcpyDeviceToHost));
    if (extra && pol_rates)
        HIP_TRY(&c, hipMemcpy(pol_rates, c.d_rates, sizeof(double) * N * 6 * L, hipMemcpyDeviceToHost));
    if (extra && JJ_20) {
        // JJ_20 = J2/J0 for solved nodes (main_sub.f90:139), node_num x r_num
        std::vector<double> J0(N * R), J2(N * R);
        HIP_TRY(&c, hipMemcpy(J0.data(), c.d_J0, sizeof(double) * N * R, hipMemcpyDeviceToHost));
        HIP_TRY(&c, hipMemcpy(J2.data(), c.d_J2, sizeof(double) * N * R, hipMemcpyDeviceToHost));
        std::fill(JJ_20, JJ_20 + N * R, 0.0);
        for (int node : c.g.work)
            for (size_t k = 0; k < R; k++) JJ_20[node + N * k] = J2[k + R * node] / J0[k + R * node];
    }
    return PORTAL_OK;
}

static int surrogate_step(Ctx &c, int *n_fallback, int *d_item_flags, double *t_mlp_ms) {
    int n_in = 0, n_out = 0;
    if (!mlp_dims(c, &n_in, &n_out)) return fail(c, PORTAL_ERR_STATE, "no surrogate network (portal_mlp_create)");
    const Species &s = c.sp;
    int nj = 0;
    for (int i = 0; i < s.e_num; i++) nj += s.j_lev[i] != 0;
    if (n_in != 2 + 2 * s.r_num + s.e_num + nj || n_out != s.e_num + nj)
        return fail(c, PORTAL_ERR_ARG, "surrogate network shape does not match the species");
    const int nw = (int)c.g.work.size();
    float *d_f = nullptr, *d_o = nullptr;
    int *d_n = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    Defer release{[&] {
        dfree(d_f);
        dfree(d_o);
        dfree(d_n);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }};
    HIP_TRY(&c, hipEventCreate(&ev[0]));
    HIP_TRY(&c, hipEventCreate(&ev[1]));
    HIP_TRY(&c, hipMalloc((void **)&d_f, sizeof(float) * (size_t)nw * n_in + 16));
    HIP_TRY(&c, hipMalloc((void **)&d_o, sizeof(float) * (size_t)nw * n_out + 16));
    HIP_TRY(&c, hipMalloc((void **)&d_n, sizeof(int)));
    HIP_TRY(&c, hipMemsetAsync(d_n, 0, sizeof(int), c.stream));
    SurrArgs a{};
    a.n_work = nw;
    a.N = c.g.N;
    a.L = s.e_num;
    a.R = s.r_num;
    a.n_in = n_in;
    a.n_out = n_out;
    a.work = c.g.d_work;
    a.h2 = c.g.d_h2;
    a.tkin = c.g.d_tkin;
    a.ab = c.g.d_ab;
    a.J0 = c.d_J0;
    a.J2 = c.d_J2;
    a.Iin = s.d_Iin;
    a.j_lev = s.d_jlev;
    a.l1 = s.d_l1;
    a.l2 = s.d_l2;
    a.g1 = s.d_g1;
    a.g2 = s.d_g2;
    a.feats = d_f;
    a.out = d_o;
    a.levp = c.d_levp;
    a.pol = c.d_pol;
    a.n_fallback = d_n;
    a.flags = d_item_flags;
    a.perturb_stride = c.surr_perturb_stride;
    a.perturb_factor = c.surr_perturb_factor;
    HIP_TRY(&c, launch_surr_features(a, c.stream));
    HIP_TRY(&c, hipEventRecord(ev[0], c.stream));
    HIP_TRY(&c, mlp_infer_device(c, nw, d_f, d_o));
    HIP_TRY(&c, hipEventRecord(ev[1], c.stream));
    HIP_TRY(&c, launch_surr_apply(a, c.stream));
    HIP_TRY(&c, hipMemcpyAsync(n_fallback, d_n, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    float ms = 0.0f;
    HIP_TRY(&c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    if (t_mlp_ms) *t_mlp_ms += ms;
    return PORTAL_OK;
}

// the autotuner's timings of the inward-integration candidates (variant/waves: cooperative
// 2, 4, 5, 8, cooperative back to front 2, 4, 5, 8, staged 4, staged 2, global 4) for one
// precision, -1 for a candidate that could not run or was not timed; out holds n doubles,
// returns the number of candidates
int portal_debug_tune_ms(const portal_ctx *x, int precision, double *out, int n) {
    if (!x || !out || n < 0 || precision < 0 || precision > 1) return PORTAL_ERR_ARG;
    for (int k = 0; k < n && k < N_RT_CANDS; k++) out[k] = x->c.tune_ms[precision][k];
    return N_RT_CANDS;
}

// Test hooks of the surrogate's residual fallback: every stride-th work item's
// prediction gets its first level's population scaled by factor (stride 0: off); and
// the per-item outcome of the last surrogate run (0 kept, 1 failed the validity tests,
// 2 reverted by the residual test), out holding n_work ints (returns n_work).
int portal_debug_surrogate_perturb(portal_ctx *x, int stride, double factor) {
    if (!x || stride < 0 || !(factor > 0)) return PORTAL_ERR_ARG;
    x->c.surr_perturb_stride = stride;
    x->c.surr_perturb_factor = factor;
    return PORTAL_OK;
}

int portal_debug_fallback_map(const portal_ctx *x, int *out, int n) {
    if (!x || !out || n < 0) return PORTAL_ERR_ARG;
    const int m = (int)std::min<size_t>(x->c.fallback_map.size(), (size_t)n);
    for (int i = 0; i < m; i++) out[i] = x->c.fallback_map[i];
    return (int)x->c.fallback_map.size();
}

// the population and alignment changes of each iteration of the last lambda call
// (2 doubles per iteration, at most max_iter iterations); returns the iteration count
int portal_debug_iter_trace(const portal_ctx *x, double *out, int max_iter) {
    if (!x || !out || max_iter < 0) return PORTAL_ERR_ARG;
    const int n = (int)std::min<size_t>(x->c.iter_trace.size() / 2, (size_t)max_iter);
    for (int i = 0; i < 2 * n; i++) out[i] = x->c.iter_trace[i];
    return n;
}

// Test hook: the Ng schedule of the single-model iteration, the first iteration
// (0-based, at least 3) and the least gap between extrapolations (at least 3).
int portal_debug_ng_schedule(portal_ctx *x, int first, int gap) {
    if (!x || first < 3 || gap < 3) return PORTAL_ERR_ARG;
    x->c.ng_first = first;
    x->c.ng_gap = gap;
    return PORTAL_OK;
}

// Test hook: the chunk store of the single-pass walk is sized at scale x the sampled
// estimate (default 1.1); a small scale forces the overflow fallback, scale <= 0 the
// two-pass walk.  The next path build uses it.
int portal_debug_walk_chunks(portal_ctx *x, double scale) {
    if (!x) return PORTAL_ERR_ARG;
    x->c.walk_chunk_scale = scale;
    x->c.paths.valid = false;
    return PORTAL_OK;
}

// Test hook: cells per staged unit of the cooperative kernel (0: the default, 64).
int portal_debug_coop_unit(portal_ctx *x, int cells) {
    if (!x || cells < 0) return PORTAL_ERR_ARG;
    x->c.coop_unit = cells;
    x->c.tuned_key[0] = x->c.tuned_key[1] = 0;
    return PORTAL_OK;
}

// Test hook: run the inward integration of one precision with a fixed kernel variant
// (0 global, 1 staged, 2 cooperative, 3 cooperative back to front) and wave count;
// variant -1 returns to autotuning.
int portal_debug_force_rt(portal_ctx *x, int precision, int variant, int waves) {
    if (!x || precision < 0 || precision > 1 || variant < -1 || variant > 3 || waves < 1 || waves > 8)
        return PORTAL_ERR_ARG;
    x->c.forced_variant[precision] = variant;
    x->c.forced_waves[precision] = waves;
    x->c.tuned_key[precision] = 0;
    return PORTAL_OK;
}

// Test hook: the cooperative form of the inward integration in a batch this context
// leads: -1 timed, else bit 0 one workgroup per item (else persistent workgroups), bit 1
// each ray's cells back to front.
int portal_debug_force_batch_rt(portal_ctx *x, int form) {
    if (!x || form < -1 || form > 3) return PORTAL_ERR_ARG;
    x->c.batch_rt_form = form;
    return PORTAL_OK;
}

// Test hook: waves per system of this context's SE solve (1, 2 or 4), and for a batch
// that this context leads, the SE form (-1 autotuned, 0 one persistent launch, 1 a
// launch per class of system sizes) and the classes' waves (0 autotuned).  Form 2 puts
// this context's own solve in the slab form that larger systems take, form 3 in the form
// of systems above 128 unknowns (factors in the slab, substitutions on all waves).
int portal_debug_force_se(portal_ctx *x, int form, int waves) {
    if (!x || form < -1 || form > 3 || !(waves == 0 || waves == 1 || waves == 2 || waves == 4)) return PORTAL_ERR_ARG;
    x->c.se_waves = waves > 0 ? waves : 1;
    x->c.se_forced_form = form >= 2 ? -1 : form;
    x->c.se_forced_waves = waves;
    x->c.se_force_slab = form >= 2 ? form - 1 : 0;
    return PORTAL_OK;
}

int portal_get_pol_pops(portal_ctx *x, double *pol_pops) {
    if (!x || !pol_pops) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    int rc = ensure_node_buffers(c);
    if (rc) return rc;
    HIP_TRY(&c, hipMemcpy(pol_pops, c.d_pol, sizeof(double) * c.g.N * 2 * c.sp.e_num, hipMemcpyDeviceToHost));
    return PORTAL_OK;
}

int portal_surrogate_features(portal_ctx *x, int n_in, float *feats, int *nodes, int *n_work) {
    if (!x || !feats || !n_work) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    if (!c.g.valid || !c.sp.valid || !c.d_J0 || !c.pol_valid)
        return fail(c, PORTAL_ERR_STATE, "no SE pass has run");
    int nj = 0;
    for (int i = 0; i < c.sp.e_num; i++) nj += c.sp.j_lev[i] != 0;
    if (n_in != 2 + 2 * c.sp.r_num + c.sp.e_num + nj)
        return fail(c, PORTAL_ERR_ARG, "n_in must be 2 + 2 r_num + e_num + (levels with J > 0)");
    const int nw = (int)c.g.work.size();
    float *d_f = nullptr;
    HIP_TRY(&c, hipMalloc((void **)&d_f, sizeof(float) * (size_t)nw * n_in + 16));
    SurrArgs a{};
    a.n_work = nw;
    a.N = c.g.N;
    a.L = c.sp.e_num;
    a.R = c.sp.r_num;
    a.n_in = n_in;
    a.work = c.g.d_work;
    a.h2 = c.g.d_h2;
    a.tkin = c.g.d_tkin;
    a.J0 = c.d_J0;
    a.J2 = c.d_J2;
    a.Iin = c.sp.d_Iin;
    a.j_lev = c.sp.d_jlev;
    a.pol = c.d_pol;
    a.feats = d_f;
    HIP_TRY(&c, launch_surr_features(a, c.stream));
    HIP_TRY(&c, hipMemcpyAsync(feats, d_f, sizeof(float) * (size_t)nw * n_in, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(&c, hipStreamSynchronize(c.stream));
    hipFree(d_f);
    if (nodes)
        for (int i = 0; i < nw; i++) nodes[i] = c.g.work[i] + 1;
    *n_work = nw;
    return PORTAL_OK;
}

int portal_lambda_iterate(portal_ctx *x, const portal_se_params *p, const portal_iter_params *it, double *pol_pops,
                          double *lev_p_out, portal_stats *st) {
    if (!x || !it) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    int rc = check_se_params(c, p);
    if (rc) return rc;
    if (it->max_iter <= 0) return fail(c, PORTAL_ERR_ARG, "max_iter must be positive");
    if (it->surrogate && !c.mlp) return fail(c, PORTAL_ERR_STATE, "no surrogate network (portal_mlp_create)");
    // acceleration: 1 Ng (the three previous level-density iterates), 2 Anderson (the
    // window of (x, G(x)) pairs, compact over solved nodes)
    if (it->ng_accel < 0 || it->ng_accel > 2) return fail(c, PORTAL_ERR_ARG, "ng_accel must be 0, 1 or 2");
    if (it->ali < 0 || it->ali > 1) return fail(c, PORTAL_ERR_ARG, "ali must be 0 or 1");
    if (it->surrogate && !(it->fallback_tol >= 0.0))
        return fail(c, PORTAL_ERR_ARG, "fallback_tol must be >= 0 (+inf turns the residual test off)");
    if (st) std::memset(st, 0, sizeof(*st));
    if (it->rebuild_paths) c.paths.valid = false;
    const size_t L = c.sp.e_num, N = c.g.N;
    const int nw = (int)c.g.work.size();
    double *d_prev = nullptr, *d_hist = nullptr, *d_ngp = nullptr, *d_aa = nullptr;
    unsigned long long *d_max = nullptr;
    // the surrogate's residual test: per item the changes of the plain first iterate and
    // of the iteration after the prediction, the plain first iterate itself, item flags
    unsigned long long *d_res = nullptr;
    double *d_plain_lev = nullptr, *d_plain_pol = nullptr;
    int *d_iflags = nullptr, *d_nrev = nullptr;
    Defer release{[&] {
        dfree(d_prev);
        dfree(d_max);
        dfree(d_hist);
        dfree(d_ngp);
        dfree(d_aa);
        dfree(d_res);
        dfree(d_plain_lev);
        dfree(d_plain_pol);
        dfree(d_iflags);
        dfree(d_nrev);
    }};
    HIP_TRY(&c, hipMalloc((void **)&d_prev, sizeof(double) * N * 2 * L));
    HIP_TRY(&c, hipMalloc((void **)&d_max, sizeof(unsigned long long) * 2));
    if (it->surrogate) {
        HIP_TRY(&c, hipMalloc((void **)&d_res, sizeof(unsigned long long) * 2 * std::max(nw, 1)));
        HIP_TRY(&c, hipMalloc((void **)&d_plain_lev, sizeof(double) * N * L));
        HIP_TRY(&c, hipMalloc((void **)&d_plain_pol, sizeof(double) * N * 2 * L));
        HIP_TRY(&c, hipMalloc((void **)&d_iflags, sizeof(int) * std::max(nw, 1)));
        HIP_TRY(&c, hipMalloc((void **)&d_nrev, sizeof(int)));
        HIP_TRY(&c, hipMemsetAsync(d_res, 0, sizeof(unsigned long long) * 2 * std::max(nw, 1), c.stream));
        HIP_TRY(&c, hipMemsetAsync(d_iflags, 0, sizeof(int) * std::max(nw, 1), c.stream));
        HIP_TRY(&c, hipMemsetAsync(d_nrev, 0, sizeof(int), c.stream));
    }
    unsigned long long *res_p = d_res, *res_s = d_res ? d_res + std::max(nw, 1) : nullptr;
    const bool ng = it->ng_accel == 1, aa = it->ng_accel == 2;
    // the own-cell operator covers the line alone: with dust in the inward pass the
    // iteration stays plain
    bool dust_on = false;
    if (c.g.has_dust)
        for (double kd : c.sp.kd_v) dust_on = dust_on || kd != 0.0;
    struct AliScope {
        Ctx &c;
        ~AliScope() { c.ali = false; }
    } ali_scope{c};
    c.ali = it->ali == 1 && !dust_on;
    if (ng) {
        HIP_TRY(&c, hipMalloc((void **)&d_hist, sizeof(double) * N * L * 3));
        HIP_TRY(&c, hipMalloc((void **)&d_ngp, sizeof(double) * ng_partial_doubles()));
    }
    const long long n_aa = (long long)c.g.work.size() * (long long)L;
    constexpr int AW = AA_M_MAX + 1;  // pairs kept
    if (aa) {
        HIP_TRY(&c, hipMalloc((void **)&d_aa, sizeof(double) * (size_t)n_aa * 2 * AW));
        HIP_TRY(&c, hipMalloc((void **)&d_ngp, sizeof(double) * aa_partial_doubles()));
    }
    auto aa_x = [&](int slot) { return d_aa + (size_t)n_aa * slot; };
    auto aa_g = [&](int slot) { return d_aa + (size_t)n_aa * (AW + slot); };
    int aa_cnt = 0, aa_pos = 0;
    double *hist[3] = {d_hist, d_hist ? d_hist + N * L : nullptr, d_hist ? d_hist + 2 * N * L : nullptr};
    int nhist = 0, n_ng = 0, last_ng = -1, conv = 0;
    Timer tt(c.stream);
    int iters = 0, nfb = 0;
    double t_mlp = 0.0;
    double dpop = 0, dal = 0;
    // continuing an earlier call: its last SE solution is this call's first comparison
    const bool chained = c.pol_chain && !it->surrogate;
    if (chained)
        HIP_TRY(&c, hipMemcpyAsync(d_prev, c.d_pol, sizeof(double) * N * 2 * L, hipMemcpyDeviceToDevice, c.stream));
    c.pol_chain = false;
    c.iter_trace.clear();
    for (int k = 0; k < it->max_iter; k++) {
        if (aa) HIP_TRY(&c, launch_aa_gather(nw, (int)L, c.g.d_work, c.d_levp, aa_x(aa_pos % AW), c.stream));
        rc = run_inward(c, p, st);
        if (!rc) rc = run_se(c, 0, st);
        if (rc) break;
        HIP_TRY(&c, hipMemsetAsync(d_max, 0, sizeof(unsigned long long) * 2, c.stream));
        // with the surrogate, the per-item population changes of the first two iterations
        unsigned long long *node_res = !it->surrogate ? nullptr : k == 0 ? res_p : k == 1 ? res_s : nullptr;
        HIP_TRY(&c, launch_lambda_update((int)c.g.work.size(), (int)N, (int)L, c.g.d_work, c.g.d_ab, c.g.d_h2,
                                         c.sp.d_jlev, c.d_pol, (k > 0 || chained) ? d_prev : nullptr, c.d_levp, d_max,
                                         c.stream, nullptr, node_res));
        if (k == 0 && it->surrogate) {
            // keep the plain first iterate, then the network's prediction of the converged
            // state replaces it wherever it passes the validity checks
            HIP_TRY(&c, hipMemcpyAsync(d_plain_lev, c.d_levp, sizeof(double) * N * L, hipMemcpyDeviceToDevice,
                                       c.stream));
            HIP_TRY(&c, hipMemcpyAsync(d_plain_pol, c.d_pol, sizeof(double) * N * 2 * L, hipMemcpyDeviceToDevice,
                                       c.stream));
            rc = surrogate_step(c, &nfb, d_iflags, &t_mlp);
            if (rc) break;
        }
        if (k == 1 && it->surrogate) {
            // the residual test: an item whose change in this first iteration after the
            // prediction exceeds fallback_tol and the change of its plain first iterate
            // goes back to that plain iterate
            SurrRevertArgs ra{};
            ra.n_work = nw;
            ra.N = (int)N;
            ra.L = (int)L;
            ra.work = c.g.d_work;
            ra.res_s = res_s;
            ra.res_p = res_p;
            ra.tol = it->fallback_tol;
            ra.plain_lev = d_plain_lev;
            ra.plain_pol = d_plain_pol;
            ra.levp = c.d_levp;
            ra.pol = c.d_pol;
            ra.flags = d_iflags;
            ra.n_fallback = d_nrev;
            HIP_TRY(&c, launch_surr_revert(ra, c.stream));
            int nrev = 0;
            HIP_TRY(&c, hipMemcpyAsync(&nrev, d_nrev, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            HIP_TRY(&c, hipStreamSynchronize(c.stream));
            nfb += nrev;
        }
        unsigned long long mx[2];
        HIP_TRY(&c, hipMemcpyAsync(mx, d_max, sizeof(mx), hipMemcpyDeviceToHost, c.stream));
        HIP_TRY(&c, hipMemcpyAsync(d_prev, c.d_pol, sizeof(double) * N * 2 * L, hipMemcpyDeviceToDevice, c.stream));
        HIP_TRY(&c, hipStreamSynchronize(c.stream));
        std::memcpy(&dpop, &mx[0], 8);
        std::memcpy(&dal, &mx[1], 8);
        c.iter_trace.push_back(dpop);
        c.iter_trace.push_back(dal);
        iters = k + 1;
        if (aa && !(k == 0 && it->surrogate)) {
            // the pair (x_k, G(x_k)); the surrogate's replacement of the first iterate is not one
            HIP_TRY(&c, launch_aa_gather(nw, (int)L, c.g.d_work, c.d_levp, aa_g(aa_pos % AW), c.stream));
            aa_pos++;
            aa_cnt = std::min(aa_cnt + 1, AW);
        }
        if ((k > 0 || chained) && dpop <= it->tol_pop && dal <= it->tol_align) {
            conv = 1;
            break;
        }
        if (aa && aa_cnt >= 2) {
            const int mk = aa_cnt - 1, s0 = (aa_pos - 1) % AW;
            const double *xj[AA_M_MAX], *gj[AA_M_MAX];
            for (int j = 0; j < mk; j++) {
                const int sj = (aa_pos - 1 - (j + 1)) % AW;
                xj[j] = aa_x(sj);
                gj[j] = aa_g(sj);
            }
            double sums[AA_NSUM];
            double *d_out = d_ngp + aa_partial_doubles() - 32;
            HIP_TRY(&c, launch_aa_dots(n_aa, mk, aa_x(s0), aa_g(s0), xj, gj, d_ngp, d_out, c.stream));
            HIP_TRY(&c, hipMemcpyAsync(sums, d_out, sizeof(sums), hipMemcpyDeviceToHost, c.stream));
            HIP_TRY(&c, hipStreamSynchronize(c.stream));
            // normal equations A gamma = b (A: dF.dF, b: dF.f), solved with partial pivoting
            double A[AA_M_MAX][AA_M_MAX + 1];
            int t = 0;
            for (int i = 0; i < AA_M_MAX; i++) {
                for (int j = i; j < AA_M_MAX; j++) A[i][j] = A[j][i] = sums[t++];
                A[i][AA_M_MAX] = sums[t++];
            }
            double gam[AA_M_MAX] = {0, 0, 0, 0};
            bool ok = true;
            double dmax = 0;
            for (int i = 0; i < mk; i++) dmax = std::max(dmax, A[i][i]);
            ok = std::isfinite(dmax) && dmax > 0;
            for (int i = 0; i < mk && ok; i++) A[i][i] += 1e-12 * dmax;
            double M[AA_M_MAX][AA_M_MAX + 1];
            for (int i = 0; i < mk; i++) {
                for (int j = 0; j < mk; j++) M[i][j] = A[i][j];
                M[i][mk] = A[i][AA_M_MAX];
            }
            for (int col = 0; col < mk && ok; col++) {
                int piv = col;
                for (int r = col + 1; r < mk; r++)
                    if (std::fabs(M[r][col]) > std::fabs(M[piv][col])) piv = r;
                if (!(std::fabs(M[piv][col]) > 1e-300)) {
                    ok = false;
                    break;
                }
                if (piv != col)
                    for (int j = 0; j <= mk; j++) std::swap(M[col][j], M[piv][j]);
                for (int r = col + 1; r < mk; r++) {
                    const double f = M[r][col] / M[col][col];
                    for (int j = col; j <= mk; j++) M[r][j] -= f * M[col][j];
                }
            }
            for (int i = mk - 1; i >= 0 && ok; i--) {
                double v = M[i][mk];
                for (int j = i + 1; j < mk; j++) v -= M[i][j] * gam[j];
                gam[i] = v / M[i][i];
                ok = std::isfinite(gam[i]) && std::fabs(gam[i]) < 1e3;
            }
            if (ok) {
                HIP_TRY(&c, launch_aa_apply(nw, (int)L, mk, c.g.d_work, aa_g(s0), gj, gam, c.d_levp, c.stream));
                n_ng++;
                last_ng = k;
            } else {
                aa_cnt = 1;  // an ill-conditioned window restarts from the newest pair
            }
        }
        if (ng) {
            // history x1 = hist[0] (newest) .. x3 = hist[2]; x0 is the level densities just formed
            if (nhist == 3 && k >= c.ng_first && (last_ng < 0 || k - last_ng >= c.ng_gap)) {
                double sums[5];
                HIP_TRY(&c, launch_ng_dots(nw, (int)L, c.g.d_work, c.d_levp, hist[0], hist[1], hist[2], d_ngp,
                                           d_ngp + ng_partial_doubles() - 8, c.stream));
                HIP_TRY(&c, hipMemcpyAsync(sums, d_ngp + ng_partial_doubles() - 8, sizeof(sums),
                                           hipMemcpyDeviceToHost, c.stream));
                HIP_TRY(&c, hipStreamSynchronize(c.stream));
                const double a11 = sums[0], a12 = sums[1], a22 = sums[2], b1 = sums[3], b2 = sums[4];
                const double det = a11 * a22 - a12 * a12;
                if (std::isfinite(det) && det > 1e-14 * a11 * a22 && a11 > 0 && a22 > 0) {
                    const double ca = (b1 * a22 - b2 * a12) / det, cb = (b2 * a11 - b1 * a12) / det;
                    if (std::isfinite(ca) && std::isfinite(cb)) {
                        HIP_TRY(&c, launch_ng_apply(nw, (int)L, c.g.d_work, c.d_levp, hist[0], hist[1], ca, cb,
                                                    c.stream));
                        n_ng++;
                        last_ng = k;
                    }
                }
                nhist = 0;  // the extrapolated iterate starts a new history
            }
            double *tail = hist[2];
            hist[2] = hist[1];
            hist[1] = hist[0];
            hist[0] = tail;
            HIP_TRY(&c, hipMemcpyAsync(hist[0], c.d_levp, sizeof(double) * N * L, hipMemcpyDeviceToDevice, c.stream));
            nhist = std::min(nhist + 1, 3);
        }
    }
    const double ms = tt.stop();
    if (rc) return rc;
    if (it->surrogate) {
        c.fallback_map.assign(std::max(nw, 1), 0);
        HIP_TRY(&c, hipMemcpy(c.fallback_map.data(), d_iflags, sizeof(int) * std::max(nw, 1), hipMemcpyDeviceToHost));
        c.fallback_map.resize(nw);
    }
    // the Ng history is not carried over, and an extrapolation leaves d_levp ahead of d_pol
    c.pol_chain = n_ng == 0 || last_ng < iters - 1;
    if (st) {
        st->t_total_ms = ms;
        st->iterations = iters;
        st->max_rel_change = dpop;
        st->fallbacks = nfb;
        st->t_mlp_ms = t_mlp;
        st->ng_steps = n_ng;
        st->converged = conv;
        st->max_align_change = dal;
    }
    if (pol_pops) HIP_TRY(&c, hipMemcpy(pol_pops, c.d_pol, sizeof(double) * N * 2 * L, hipMemcpyDeviceToHost));
    if (lev_p_out) HIP_TRY(&c, hipMemcpy(lev_p_out, c.d_levp, sizeof(double) * N * L, hipMemcpyDeviceToHost));
    return PORTAL_OK;
}

int portal_image_pass(portal_ctx *x, const portal_img_params *p, const double *pol_pops, const double *bvec_all,
                      double *I, double *Q, double *U, double *TAU, double *dust, portal_stats *st) {
    if (!x) return PORTAL_ERR_ARG;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    if (!c.g.valid || !c.sp.valid) return fail(c, PORTAL_ERR_STATE, "grid and species must be uploaded first");
    if (!p || !pol_pops || !bvec_all || !I || !Q || !U || !TAU || p->nlevels <= 0 || p->nvel < 0 || p->im_pix < 0 ||
        !p->l1 || !p->l2 || !p->nu0_all || !p->Aji_all || !p->kd_all || !p->vel_out || p->n_depth_rt <= 0 ||
        !(p->dx > 0))
        return fail(c, PORTAL_ERR_ARG, "bad image-pass arguments");
    if (p->precision != PORTAL_FP64 && p->precision != PORTAL_MIXED) return fail(c, PORTAL_ERR_ARG, "bad precision");
    if (p->star_mod != 0 && p->star_mod != 1) return fail(c, PORTAL_ERR_ARG, "star_mod must be 0 or 1");
    if (p->star_mod == 1 &&
        (!p->X_star || !(p->R_star > 0.0) || !std::isfinite(p->R_star) || !(p->T_star >= 0.0) ||
         !std::isfinite(p->T_star) || !std::isfinite(p->X_star[0]) || !std::isfinite(p->X_star[1]) ||
         !std::isfinite(p->X_star[2])))
        return fail(c, PORTAL_ERR_ARG, "star mode needs a finite X_star, R_star > 0 and T_star >= 0");
    if (p->node_pixels != 0 && p->node_pixels != 1) return fail(c, PORTAL_ERR_ARG, "node_pixels must be 0 or 1");
    if (p->node_pixels == 1 && c.g.N_b <= 0) return fail(c, PORTAL_ERR_ARG, "node_pixels needs bulk nodes");
    if (!p->node_pixels && p->im_pix > 0) {
        // the corner pixel must start on the model sphere (raytrace.f90:98 takes sqrt(1-x^2-y^2))
        const double xc = p->im_pix * p->im_size / p->im_pix;
        if (!(1.0 - xc * xc - xc * xc >= 0.0))
            return fail(c, PORTAL_ERR_ARG, "image corners lie outside the model sphere (im_size > 1/sqrt(2))");
    }
    const Species &s = c.sp;
    const int L = s.e_num;
    for (int j = 0; j < p->nlevels; j++)
        if (p->l1[j] < 1 || p->l1[j] > L || p->l2[j] < 1 || p->l2[j] > L)
            return fail(c, PORTAL_ERR_ARG, "imaged line levels out of range");
    if (st) std::memset(st, 0, sizeof(*st));
    const int N = c.g.N, nl = p->nlevels, C = 2 * p->nvel + 1, n_side = 2 * p->im_pix + 1;
    const int n_rt = p->node_pixels ? c.g.N_b : n_side * n_side;
    // R_tot = R_z(ph_az) R_x(th_inc) (miscellaneous.f90:870-880)
    auto rotm = [](const double nv[3], double th, double *R) {
        const double cth = std::cos(th), octh = 1.0 - cth, sth = std::sin(th);
        R[0] = cth + octh * nv[0] * nv[0];
        R[4] = cth + octh * nv[1] * nv[1];
        R[8] = cth + octh * nv[2] * nv[2];
        R[3] = nv[0] * nv[1] * octh - nv[2] * sth;
        R[1] = nv[1] * nv[0] * octh + nv[2] * sth;
        R[6] = nv[0] * nv[2] * octh + nv[1] * sth;
        R[2] = nv[2] * nv[0] * octh - nv[1] * sth;
        R[7] = nv[1] * nv[2] * octh - nv[0] * sth;
        R[5] = nv[2] * nv[1] * octh + nv[0] * sth;
    };
    const double xv[3] = {1, 0, 0}, zv[3] = {0, 0, 1};
    double Rth[9], Rph[9];
    std::vector<double> Rt(9);
    rotm(xv, p->th_inc, Rth);
    rotm(zv, p->ph_az, Rph);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double sum = 0.0;
            for (int k = 0; k < 3; k++) sum += Rph[i + 3 * k] * Rth[k + 3 * j];
            Rt[i + 3 * j] = sum;
        }
    int jmax = 0;
    for (int i = 0; i < L; i++) jmax = std::max(jmax, s.j_lev[i]);
    std::vector<double> w6((jmax + 1) * (jmax + 1));
    for (int a = 0; a <= jmax; a++)
        for (int b = 0; b <= jmax; b++) w6[a * (jmax + 1) + b] = portal_debug_wig6jj(2, 2, 4, 2 * a, 2 * a, 2 * b);
    std::vector<int> l1(nl), l2(nl);
    std::vector<double> Iin(2 * (size_t)nl);  // B_nu(2.73 K), then B_nu(T_star)
    for (int j = 0; j < nl; j++) {
        l1[j] = p->l1[j] - 1;
        l2[j] = p->l2[j] - 1;
        Iin[j] = black_body(2.73, p->nu0_all[j]);
        Iin[nl + j] = p->star_mod == 1 ? black_body(p->T_star, p->nu0_all[j]) : 0.0;
    }
    // the star's position in the rotated frame, summed as rotate_all_save does
    // (miscellaneous.f90:900-905)
    double Xs[3] = {0.0, 0.0, 0.0};
    if (p->star_mod == 1)
        for (int j = 0; j < 3; j++) {
            double sum = 0.0;
            for (int k = 0; k < 3; k++) sum = sum + Rt[j + 3 * k] * p->X_star[k];
            Xs[j] = sum;
        }
    // device scratch, one allocation
    const size_t nrec = (size_t)n_rt * p->n_depth_rt;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
    const size_t o_R = take(9 * 8), o_xyz = take(24 * (size_t)N), o_vel = take(24 * (size_t)N),
                 o_bv = take(24 * (size_t)N), o_pol = take(16 * (size_t)N * L), o_jl = take(4 * (size_t)L),
                 o_l1 = take(4 * (size_t)nl), o_l2 = take(4 * (size_t)nl), o_nu = take(8 * (size_t)nl),
                 o_A = take(8 * (size_t)nl), o_kd = take(8 * (size_t)nl), o_Iin = take(16 * (size_t)nl),
                 o_vo = take(8 * (size_t)C), o_w6 = take(8 * w6.size()), o_pos = take(32 * (size_t)N),
                 o_vz = take(8 * (size_t)N), o_geo = take(sizeof(CellGeo) * (size_t)N),
                 o_pc = take(32 * (size_t)N * nl), o_st = take(32 * (size_t)n_rt), o_sn = take(4 * (size_t)n_rt),
                 o_rc = take(4 * nrec), o_rs = take(4 * nrec), o_rn = take(4 * (size_t)n_rt),
                 o_out = take(8 * (size_t)n_rt * nl * C * (dust ? 8 : 4)), o_b = take(8 * (size_t)N),
                 o_du = take(8 * (size_t)N), o_td = take(8 * (size_t)N);
    if (off > c.img_bytes) {
        if (c.d_img) hipFree(c.d_img);
        c.d_img = nullptr;
        c.img_bytes = 0;
        HIP_TRY(&c, hipMalloc(&c.d_img, off));
        c.img_bytes = off;
    }
    char *B = (char *)c.d_img;
    auto H2D = [&](size_t o, const void *src, size_t bytes) {
        return hipMemcpyAsync(B + o, src, bytes, hipMemcpyHostToDevice, c.stream);
    };
    std::vector<double> tdust(N), dustv(N);
    for (int i = 0; i < N; i++) tdust[i] = c.g.temp[2 * (size_t)i + 1];
    HIP_TRY(&c, H2D(o_R, Rt.data(), 72));
    HIP_TRY(&c, H2D(o_xyz, c.g.XYZ.data(), 24 * (size_t)N));
    HIP_TRY(&c, H2D(o_vel, c.g.vel.data(), 24 * (size_t)N));
    HIP_TRY(&c, H2D(o_bv, bvec_all, 24 * (size_t)N));
    HIP_TRY(&c, H2D(o_pol, pol_pops, 16 * (size_t)N * L));
    HIP_TRY(&c, H2D(o_jl, s.j_lev.data(), 4 * (size_t)L));
    HIP_TRY(&c, H2D(o_l1, l1.data(), 4 * (size_t)nl));
    HIP_TRY(&c, H2D(o_l2, l2.data(), 4 * (size_t)nl));
    HIP_TRY(&c, H2D(o_nu, p->nu0_all, 8 * (size_t)nl));
    HIP_TRY(&c, H2D(o_A, p->Aji_all, 8 * (size_t)nl));
    HIP_TRY(&c, H2D(o_kd, p->kd_all, 8 * (size_t)nl));
    HIP_TRY(&c, H2D(o_Iin, Iin.data(), 16 * (size_t)nl));
    HIP_TRY(&c, H2D(o_vo, p->vel_out, 8 * (size_t)C));
    HIP_TRY(&c, H2D(o_w6, w6.data(), 8 * w6.size()));
    HIP_TRY(&c, H2D(o_b, c.g.b_all.data(), 8 * (size_t)N));
    HIP_TRY(&c, H2D(o_du, c.g.dust.data(), 8 * (size_t)N));
    HIP_TRY(&c, H2D(o_td, tdust.data(), 8 * (size_t)N));
    if (!c.d_flags) HIP_TRY(&c, hipMalloc((void **)&c.d_flags, sizeof(int) * 4));
    HIP_TRY(&c, hipMemsetAsync(c.d_flags, 0, sizeof(int) * 4, c.stream));
    ImgArgs a{};
    a.N = N;
    a.N_b = c.g.N_b;
    a.nlines = nl;
    a.C = C;
    a.n_rt = n_rt;
    a.n_side = n_side;
    a.im_pix = p->im_pix;
    a.n_depth_rt = p->n_depth_rt;
    a.jmax = jmax;
    a.im_size = p->im_size;
    a.dx = p->dx;
    a.dl_scale = c.g.tx * p->dx;
    a.alpha = p->alpha;
    a.Rtot = (const double *)(B + o_R);
    a.XYZ = (const double *)(B + o_xyz);
    a.vel = (const double *)(B + o_vel);
    a.bvec = (const double *)(B + o_bv);
    a.b_all = (const double *)(B + o_b);
    a.dust = (const double *)(B + o_du);
    a.tdust = (const double *)(B + o_td);
    a.pol = (const double *)(B + o_pol);
    a.jlev = (const int *)(B + o_jl);
    a.nbr_ptr = c.g.d_nbr_ptr;
    a.nbr_idx = c.g.d_nbr_idx;
    a.l1 = (const int *)(B + o_l1);
    a.l2 = (const int *)(B + o_l2);
    a.nu0 = (const double *)(B + o_nu);
    a.Aji = (const double *)(B + o_A);
    a.kd = (const double *)(B + o_kd);
    a.Iin = (const double *)(B + o_Iin);
    a.vel_out = (const double *)(B + o_vo);
    a.w6tab = (const double *)(B + o_w6);
    a.pos_r = (double4 *)(B + o_pos);
    a.vz_r = (double *)(B + o_vz);
    a.geo = (CellGeo *)(B + o_geo);
    a.polco = (double4 *)(B + o_pc);
    a.start = (double4 *)(B + o_st);
    a.start_node = (int *)(B + o_sn);
    a.rec_cell = (int *)(B + o_rc);
    a.rec_step = (int *)(B + o_rs);
    a.rec_n = (int *)(B + o_rn);
    a.flags = c.d_flags;
    const size_t cube = (size_t)n_rt * nl * C;
    double *outb = (double *)(B + o_out);
    a.I = outb;
    a.Q = outb + cube;
    a.U = outb + 2 * cube;
    a.TAU = outb + 3 * cube;
    a.dust_out = dust ? outb + 4 * cube : nullptr;
    a.mixed = p->precision == PORTAL_MIXED ? 1 : 0;
    a.star = p->star_mod;
    a.Xs0 = Xs[0];
    a.Xs1 = Xs[1];
    a.Xs2 = Xs[2];
    a.R_star = p->R_star;
    a.Istar = a.Iin + nl;
    a.node_pixels = p->node_pixels;
    float ms_g = 0, ms_r = 0;
    HIP_TRY(&c, launch_image(a, c.stream, &ms_g, &ms_r));
    int flags = 0;
    HIP_TRY(&c, hipMemcpy(&flags, c.d_flags, sizeof(int), hipMemcpyDeviceToHost));
    if (flags & FLAG_START)
        return fail(c, PORTAL_ERR_ARG, "a pixel's start point lies off the model sphere (raytrace.f90:98)");
    if (flags & FLAG_DEPTH_RT)
        return fail(c, PORTAL_ERR_DEPTH_RT, "an image ray crosses more than n_depth_rt cells (raytrace.f90:169-173)");
    HIP_TRY(&c, hipMemcpy(I, a.I, 8 * cube, hipMemcpyDeviceToHost));
    HIP_TRY(&c, hipMemcpy(Q, a.Q, 8 * cube, hipMemcpyDeviceToHost));
    HIP_TRY(&c, hipMemcpy(U, a.U, 8 * cube, hipMemcpyDeviceToHost));
    HIP_TRY(&c, hipMemcpy(TAU, a.TAU, 8 * cube, hipMemcpyDeviceToHost));
    if (dust) HIP_TRY(&c, hipMemcpy(dust, a.dust_out, 32 * cube, hipMemcpyDeviceToHost));
    if (st) {
        st->t_img_ms = ms_g + ms_r;
        st->t_walk_ms = ms_g;
        st->t_rt_ms = ms_r;
        st->rt_kernel_avg_ms = ms_r;
        st->t_total_ms = ms_g + ms_r;
        // cells visited by the pixels' rays (their march records), and the longest ray
        std::vector<int> rn(n_rt);
        HIP_TRY(&c, hipMemcpy(rn.data(), a.rec_n, sizeof(int) * (size_t)n_rt, hipMemcpyDeviceToHost));
        long long tot = 0;
        int mx = 0;
        for (int v : rn)
            if (v > 0) {
                tot += v;
                mx = std::max(mx, v);
            }
        st->n_cross = tot;
        st->n_paths = n_rt;
        st->max_path = mx;
    }
    return PORTAL_OK;
}


// ---------------------------------------------------------------------------------
// Batches of independent models (one context each, same device): every iteration is
// one hipGraph replay of {reset, line preparation per model, the inward integration of
// all (model, node) items (one cooperative launch per slot-and-dust class, or one staged
// persistent launch), the SE solves (one persistent launch per system size, or one over
// all items), closure per model}.
int portal_batch_lambda_iterate(portal_ctx **xs, int n, const portal_se_params *params, const portal_iter_params *it,
                                portal_stats *per_model, portal_stats *total) {
    if (!xs || n <= 0 || !params || !it || it->max_iter <= 0) return PORTAL_ERR_ARG;
    Ctx &c0 = xs[0]->c;
    hipSetDevice(c0.device);
    if (it->surrogate && !(it->fallback_tol >= 0.0)) return fail(c0, PORTAL_ERR_ARG, "fallback_tol must be >= 0");
    if (it->ng_accel < 0 || it->ng_accel > 1)
        return fail(c0, it->ng_accel == 2 ? PORTAL_ERR_UNSUPPORTED : PORTAL_ERR_ARG,
                    "a batch takes ng_accel 0 or 1 (Anderson acceleration is for single models)");
    if (it->ali) return fail(c0, PORTAL_ERR_UNSUPPORTED, "the own-cell operator is not available for batches");
    const bool mixed = params[0].precision == PORTAL_MIXED;
    // a model whose last SE solution formed its current level densities continues that
    // iteration: its first pass here is compared with that solution
    std::vector<char> chained(n, 0);
    for (int m = 0; m < n; m++) {
        if (!xs[m] || xs[m]->c.device != c0.device) return fail(c0, PORTAL_ERR_ARG, "batch contexts must share one device");
        if ((params[m].precision == PORTAL_MIXED) != mixed) return fail(c0, PORTAL_ERR_ARG, "mixed precisions in one batch");
        Ctx &c = xs[m]->c;
        int rc = check_se_params(c, &params[m]);
        if (rc) return fail(c0, rc, "model " + std::to_string(m) + ": " + c.err);
        chained[m] = c.pol_chain && !it->surrogate ? 1 : 0;
        c.pol_chain = false;
    }
    // with the surrogate, every model that holds a network gets its prediction after the
    // first pass and the residual test after the second, as a single model does; a model
    // without a network iterates plainly
    std::vector<char> use_surr(n, 0);
    for (int m = 0; m < n; m++) {
        Ctx &c = xs[m]->c;
        int ni = 0, no = 0;
        use_surr[m] = it->surrogate && mlp_dims(c, &ni, &no) ? 1 : 0;
        c.fallback_map.clear();
    }
    if (total) std::memset(total, 0, sizeof(*total));
    if (per_model) std::memset(per_model, 0, sizeof(portal_stats) * n);
    hipStream_t s = c0.stream;
    Timer ttot(s);
    // paths of every model (each on its own stream, synchronised)
    std::vector<RtArgs> rts(n);
    std::vector<SeArgs> ses(n);
    std::vector<std::pair<double, int2>> rt_items, se_items;
    std::vector<double *> d_xths(n, nullptr), d_istars(n, nullptr);
    std::vector<char> dusty(n, 0);
    auto free_xths = [&]() {
        for (auto *&p : d_xths) dfree(p);
        for (auto *&p : d_istars) dfree(p);
    };
    Defer xths_at_exit{[&] { free_xths(); }};
    // line coefficients of model m, with its dust coefficients when it has dust
    auto prep_of = [&](int m) {
        Ctx &c = xs[m]->c;
        PrepArgs pa{c.g.N, c.sp.r_num, c.sp.e_num, c.d_levp, c.g.d_velb, c.sp.d_l1, c.sp.d_l2, c.sp.d_c1B, c.sp.d_c2A,
                    nullptr, c.sp.d_g1, c.sp.d_g2, c.sp.d_cnu_spi, mixed ? nullptr : c.d_KS, mixed ? c.d_KS32 : nullptr};
        pa.I_in = c.sp.d_Iin;
        if (dusty[m]) {
            pa.dust = c.g.d_dust;
            pa.tdust = c.g.d_tdust;
            pa.kd = c.sp.d_kd;
            pa.nu = c.sp.d_nu;
            pa.KSD = mixed ? nullptr : c.d_KSD;
            pa.KSD32 = mixed ? c.d_KSD32 : nullptr;
        }
        return pa;
    };
    int cap = 8, R_max = 1, cd_max = 1, nt_max = 1;
    for (int m = 0; m < n; m++) {
        Ctx &c = xs[m]->c;
        if (it->rebuild_paths) c.paths.valid = false;
        portal_stats st{};
        int rc = ensure_paths(c, &params[m], &st);
        if (rc) {
            free_xths();
            return fail(c0, rc, "model " + std::to_string(m) + ": " + c.err);
        }
        if (total) total->t_walk_ms += st.t_walk_ms;
        const Species &sp = c.sp;
        const int D = c.paths.D, nw = (int)c.g.work.size();
        RtArgs &ra = rts[m];
        ra = RtArgs{};
        ra.n_work = nw;
        ra.D = D;
        ra.R = sp.r_num;
        ra.nvel = params[m].nvel;
        ra.C = 2 * params[m].nvel + 1;
        ra.work = c.g.d_work;
        ra.velb = c.g.d_velb;
        ra.inv_b = c.g.d_inv_b;
        ra.b_ratio = c.g.b_ratio;
        ra.KS = c.d_KS;
        ra.KS32 = c.d_KS32;
        ra.offsets = c.paths.offsets;
        ra.paths = c.paths.entries;
        ra.I_in = sp.d_Iin;
        ra.dl_scale = c.g.tx * params[m].dx;
        ra.J0 = c.d_J0;
        ra.J2 = c.d_J2;
        ra.variant = 1;
        ra.max_nc = (c.paths.max_nc + 7) & ~7;
        {
            std::vector<double> xv(params[m].xth, params[m].xth + 3 * (size_t)params[m].n_sphere);
            HIP_TRY(&c0, upload(d_xths[m], xv));
        }
        ra.xth = d_xths[m];
        // dust, as in the single-model pass: some node with dust and some line with a dust
        // opacity (inverse_rt.f90:273-287)
        bool dust = false;
        if (c.g.has_dust)
            for (double kd : sp.kd_v) dust = dust || kd != 0.0;
        if (dust) {
            if (!c.d_KSD) HIP_TRY(&c0, hipMalloc((void **)&c.d_KSD, sizeof(double) * 2 * c.g.N * sp.r_num));
            if (!c.d_KSD32) HIP_TRY(&c0, hipMalloc((void **)&c.d_KSD32, sizeof(float) * 2 * c.g.N * sp.r_num));
        }
        dusty[m] = dust ? 1 : 0;
        ra.dust = dust ? 1 : 0;
        ra.KSD = c.d_KSD;
        ra.KSD32 = c.d_KSD32;
        if (c.paths.star) {
            // rays towards the star start with B_nu(T_star) (inverse_rt.f90:113-117)
            std::vector<double> Is(sp.r_num);
            for (int k = 0; k < sp.r_num; k++) Is[k] = black_body(params[m].T_star, sp.nu_hz[k]);
            HIP_TRY(&c0, upload(d_istars[m], Is));
            ra.xthn = c.paths.d_xthn;
            ra.nstar = c.paths.d_nstar;
            ra.I_star = d_istars[m];
        }
        rc = check_rt_args(c, ra, mixed);
        if (rc) {
            free_xths();
            return fail(c0, rc, "model " + std::to_string(m) + ": " + c.err);
        }
        cap = std::max(cap, (c.paths.max_nc + 7) & ~7);
        R_max = std::max(R_max, sp.r_num);
        cd_max = std::max(cd_max, sp.c_dim);
        nt_max = std::max(nt_max, sp.n_trans);
        SeArgs &a = ses[m];
        a = SeArgs{};
        a.n_work = nw;
        a.N = c.g.N;
        a.L = sp.e_num;
        a.R = sp.r_num;
        a.c_dim = sp.c_dim;
        a.n_trans = sp.n_trans;
        a.n_ent = (int)sp.se.ent_idx.size();
        a.work = c.g.d_work;
        a.J0 = c.d_J0;
        a.J2 = c.d_J2;
        a.tkin = c.g.d_tkin;
        a.h2 = c.g.d_h2;
        a.t_A = sp.d_tA;
        a.t_B21 = sp.d_tB21;
        a.t_nu = sp.d_tnu;
        a.t_gul = sp.d_tgul;
        a.t_rad = sp.d_trad;
        a.t_col = sp.d_tcol;
        a.t_Cji = sp.d_Cji;
        a.col_temps = sp.d_coltemps;
        a.c_temp = sp.c_temp;
        a.c_par = sp.c_par;
        a.c_num = sp.c_num;
        a.ent_idx = sp.se.d_ent_idx;
        a.ent_ptr = sp.se.d_ent_ptr;
        a.term_q = sp.se.d_term_q;
        a.term_class = sp.se.d_term_class;
        a.term_coef = sp.se.d_term_coef;
        a.count_array = sp.d_count;
        a.j_lev = sp.d_jlev;
        a.depol = sp.d_depol;
        a.pol = c.d_pol;
        a.rates = nullptr;
        a.flags = c0.d_flags;
        // estimated cost of each node: stored crossings x lines x channels
        std::vector<long long> off((size_t)nw * D + 1);
        HIP_TRY(&c0, hipMemcpy(off.data(), c.paths.offsets, sizeof(long long) * off.size(), hipMemcpyDeviceToHost));
        for (int w = 0; w < nw; w++) {
            const double cost = (double)(off[(size_t)(w + 1) * D] - off[(size_t)w * D]) * ra.R * ra.C;
            rt_items.push_back({cost, make_int2(m, w)});
            se_items.push_back({(double)sp.c_dim, make_int2(m, w)});
        }
        HIP_TRY(&c0, hipMemset(c.d_J0, 0, sizeof(double) * c.g.N * sp.r_num));
        HIP_TRY(&c0, hipMemset(c.d_J2, 0, sizeof(double) * c.g.N * sp.r_num));
        // a continued model keeps its last SE solution: it is the first comparison
        if (!chained[m]) HIP_TRY(&c0, hipMemset(c.d_pol, 0, sizeof(double) * c.g.N * 2 * sp.e_num));
    }
    auto by_cost = [](const std::pair<double, int2> &a, const std::pair<double, int2> &b) { return a.first > b.first; };
    std::stable_sort(rt_items.begin(), rt_items.end(), by_cost);
    std::stable_sort(se_items.begin(), se_items.end(), by_cost);
    std::vector<int2> ri(rt_items.size()), si(se_items.size());
    for (size_t i = 0; i < ri.size(); i++) ri[i] = rt_items[i].second;
    for (size_t i = 0; i < si.size(); i++) si[i] = se_items[i].second;
    const int n_items = (int)ri.size();
    // the cooperative form: the items split by the slot count of their model's lines and
    // channels, each class one persistent launch (in cost order within the class); the
    // staged per-wave form serves batches that some class does not fit
    constexpr int COOP_W = 4;
    struct CoopClass {
        int slots = 0, cap = 8, R_max = 1, C_max = 1;
        bool dust = false;
        int prof = 2;  // the FP32 profile form every model of the class allows
        std::vector<int2> items;
        size_t lds = 0;
        int2 *d_items = nullptr;
    };
    std::vector<CoopClass> classes;
    bool coop = true;
    // the timed choices of the lead context's last call are kept while the models, their
    // paths and shapes and the precision stay the same (the test hooks time afresh)
    unsigned long long bkey = 1469598103934665603ull;
    auto mix = [&](unsigned long long v) { bkey = (bkey ^ v) * 1099511628211ull; };
    mix((unsigned long long)n);
    mix(mixed ? 1 : 0);
    for (int m = 0; m < n; m++) {
        mix((unsigned long long)(uintptr_t)xs[m]);
        mix((unsigned long long)xs[m]->c.paths.total);
        mix((unsigned long long)rts[m].R << 32 | (unsigned)rts[m].C);
        mix((unsigned long long)ses[m].c_dim << 32 | (unsigned)ses[m].n_trans);
        mix((unsigned long long)dusty[m]);
    }
    const bool reuse = c0.batch_key == bkey && c0.batch_rt_form < 0 && c0.se_forced_form < 0 && c0.se_forced_waves == 0;
    {
        std::vector<int> cls_of(n, -1);
        for (int m = 0; m < n && coop; m++) {
            const int sl = coop_slots_host(rts[m].R, rts[m].C, COOP_W);
            if (sl == 0) {
                coop = false;
                break;
            }
            for (size_t k = 0; k < classes.size(); k++)
                if (classes[k].slots == sl && classes[k].dust == (dusty[m] != 0)) cls_of[m] = (int)k;
            if (cls_of[m] < 0) {
                cls_of[m] = (int)classes.size();
                classes.emplace_back();
                classes.back().slots = sl;
                classes.back().dust = dusty[m] != 0;
            }
            CoopClass &k = classes[cls_of[m]];
            k.prof = std::min(k.prof, coop_profile_form_host(rts[m].R, rts[m].C, COOP_W, sl, rts[m].b_ratio,
                                                             rts[m].nvel));
            k.cap = std::max(k.cap, std::min(rts[m].max_nc, COOP_UNIT));
            k.R_max = std::max(k.R_max, rts[m].R);
            k.C_max = std::max(k.C_max, rts[m].C);
        }
        if (coop) {
            for (const int2 &itm : ri) classes[cls_of[itm.x]].items.push_back(itm);
            for (auto &k : classes) {
                k.lds = rt_coop_persistent_lds(mixed, k.cap, k.R_max, k.C_max, k.dust, COOP_W);
                coop = coop && k.lds <= 160 * 1024;
            }
        }
    }
    if (!coop)
        for (int m = 0; m < n; m++)
            if (dusty[m] || rts[m].xthn) {
                free_xths();
                return fail(c0, PORTAL_ERR_UNSUPPORTED, "model " + std::to_string(m) +
                                                            ": dust or star mode in a batch needs the cooperative "
                                                            "kernel, and some model's lines and channels do not fit it");
            }
    RtArgs *d_rt = nullptr;
    SeArgs *d_se = nullptr;
    int2 *d_ri = nullptr, *d_si = nullptr;
    int *d_cnt = nullptr;
    unsigned long long *d_max = nullptr;
    std::vector<double *> prev(n, nullptr);
    auto free_classes = [&]() {
        for (auto &k : classes) dfree(k.d_items);
    };
    Defer buffers_at_exit{[&] {
        free_classes();
        dfree(d_rt);
        dfree(d_se);
        dfree(d_ri);
        dfree(d_si);
        dfree(d_cnt);
        dfree(d_max);
        for (auto *&p : prev) dfree(p);
    }};
    HIP_TRY(&c0, upload(d_rt, rts));
    HIP_TRY(&c0, upload(d_se, ses));
    HIP_TRY(&c0, upload(d_ri, ri));
    HIP_TRY(&c0, upload(d_si, si));
    if (coop)
        for (auto &k : classes) HIP_TRY(&c0, upload(k.d_items, k.items));
    // [0] staged items, [1] SE items, [2..] one per cooperative class; the staged
    // form's timing clears four
    const int n_cnt = std::max(4, 2 + (int)classes.size());
    HIP_TRY(&c0, hipMalloc((void **)&d_cnt, sizeof(int) * n_cnt));
    HIP_TRY(&c0, hipMalloc((void **)&d_max, sizeof(unsigned long long) * 2 * n));
    // per model with the surrogate: the per-item change of the current iteration (res_cur,
    // written by the captured population update), that of the plain first iterate (res_p),
    // the plain first iterate itself, per-item outcomes and the revert count
    struct SurrBufs {
        unsigned long long *res_cur = nullptr, *res_p = nullptr;
        double *plain_lev = nullptr, *plain_pol = nullptr;
        int *iflags = nullptr, *nrev = nullptr;
    };
    std::vector<SurrBufs> sb(n);
    Defer free_sb{[&] {
        for (SurrBufs &b : sb) {
            dfree(b.res_cur);
            dfree(b.res_p);
            dfree(b.plain_lev);
            dfree(b.plain_pol);
            dfree(b.iflags);
            dfree(b.nrev);
        }
    }};
    for (int m = 0; m < n; m++) {
        if (!use_surr[m]) continue;
        Ctx &c = xs[m]->c;
        const size_t nw = std::max<size_t>(c.g.work.size(), 1), NL = (size_t)c.g.N * c.sp.e_num;
        HIP_TRY(&c0, hipMalloc((void **)&sb[m].res_cur, sizeof(unsigned long long) * nw));
        HIP_TRY(&c0, hipMalloc((void **)&sb[m].res_p, sizeof(unsigned long long) * nw));
        HIP_TRY(&c0, hipMalloc((void **)&sb[m].plain_lev, sizeof(double) * NL));
        HIP_TRY(&c0, hipMalloc((void **)&sb[m].plain_pol, sizeof(double) * 2 * NL));
        HIP_TRY(&c0, hipMalloc((void **)&sb[m].iflags, sizeof(int) * nw));
        HIP_TRY(&c0, hipMalloc((void **)&sb[m].nrev, sizeof(int)));
        HIP_TRY(&c0, hipMemset(sb[m].iflags, 0, sizeof(int) * nw));
        HIP_TRY(&c0, hipMemset(sb[m].nrev, 0, sizeof(int)));
    }
    for (int m = 0; m < n; m++) {
        Ctx &c = xs[m]->c;
        HIP_TRY(&c0, hipMalloc((void **)&prev[m], sizeof(double) * c.g.N * 2 * c.sp.e_num));
        if (chained[m])
            HIP_TRY(&c0, hipMemcpy(prev[m], c.d_pol, sizeof(double) * c.g.N * 2 * c.sp.e_num, hipMemcpyDeviceToDevice));
        else
            HIP_TRY(&c0, hipMemset(prev[m], 0, sizeof(double) * c.g.N * 2 * c.sp.e_num));
    }
    HIP_TRY(&c0, hipMemset(c0.d_flags, 0, sizeof(int) * 4));
    int n_cu = 256;
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c0.device);
    // the cooperative form: persistent workgroups or one workgroup per item, each with the
    // rays front to back or back to front, timed on a sample of the largest class's items:
    // 16 runs of 1,024 consecutive items spread over its cost-ordered list (all of it when
    // shorter), for the locality, spread and size of the single-model autotuner's sample
    bool coop_flat = false, coop_btf = false;
    int2 *d_tune_items = nullptr;
    Defer free_tune_items{[&] { dfree(d_tune_items); }};
    if (coop && reuse) {
        coop_flat = c0.batch_flat;
        coop_btf = c0.batch_btf;
    } else if (coop) {
        size_t kb = 0;
        for (size_t k = 1; k < classes.size(); k++)
            if (classes[k].items.size() > classes[kb].items.size()) kb = k;
        const CoopClass &cl = classes[kb];
        std::vector<int2> smp;
        {
            const size_t n_cl = cl.items.size(), run = 1024, n_runs = 16;
            if (n_cl <= run * n_runs) {
                smp = cl.items;
            } else {
                for (size_t g = 0; g < n_runs; g++) {
                    const size_t b = n_cl * g / n_runs;
                    smp.insert(smp.end(), cl.items.begin() + b, cl.items.begin() + b + run);
                }
            }
        }
        HIP_TRY(&c0, upload(d_tune_items, smp));
        const int ns = (int)smp.size();
        for (int m = 0; m < n; m++) {
            Ctx &c = xs[m]->c;
            HIP_TRY(&c0, launch_line_prep(prep_of(m), s));
        }
        // f: bit 0 the flat form, bit 1 back to front (mixed precision only, as for a single
        // model: the FP64 forms give the same bits)
        double ms_form[4] = {0, 0, 0, 0};
        int fb = c0.batch_rt_form >= 0 ? c0.batch_rt_form : 0;
        for (int f = 0; f < (mixed ? 4 : 2) && c0.batch_rt_form < 0; f++) {
            HIP_TRY(&c0, launch_fill_zero(d_cnt, sizeof(int) * n_cnt, s));
            const int bpc = std::max(1, std::min(8, (int)((160 * 1024) / std::max<size_t>(cl.lds, 1))));
            Timer t(s);
            hipError_t e = launch_rt_coop_persistent(mixed, d_rt, d_tune_items, ns, d_cnt + 2, cl.cap, cl.slots,
                                                     COOP_W, cl.lds, (f & 1) ? 0 : n_cu * bpc, cl.dust,
                                                     (f & 2) != 0, cl.prof, s);
            ms_form[f] = t.stop();
            if (e != hipSuccess) {
                (void)hipGetLastError();
                ms_form[f] = 1e30;
            }
            if (ms_form[f] < ms_form[fb]) fb = f;
        }
        coop_flat = (fb & 1) != 0;
        coop_btf = (fb & 2) != 0;
    }
    // autotune the staged persistent integration's waves per workgroup on a slice of the
    // items (when the cooperative form serves the batch, it runs with COOP_W waves)
    int best_w = coop ? COOP_W : 4;
    double best_ms = 1e30;
    for (int w : {2, 4}) {
        if (coop) break;
        const size_t lds = rt_persistent_lds(mixed, cap, R_max, w);
        if (lds > 160 * 1024) continue;
        const int bpc = std::max(1, std::min(8 / (w / 2 > 0 ? w / 2 : 1), (int)((160 * 1024) / lds)));
        for (int m = 0; m < n; m++) {
            Ctx &c = xs[m]->c;
            HIP_TRY(&c0, launch_line_prep(prep_of(m), s));
        }
        HIP_TRY(&c0, launch_fill_zero(d_cnt, sizeof(int) * 4, s));
        Timer t(s);
        hipError_t e = launch_rt_persistent(mixed, d_rt, d_ri, std::min(n_items, 8192), d_cnt, cap, R_max, w,
                                            n_cu * bpc, s);
        const double ms = t.stop();
        if (e != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best_w = w;
        }
    }
    const size_t rt_lds = rt_persistent_lds(mixed, cap, R_max, best_w);
    if (!coop && rt_lds > 160 * 1024) return fail(c0, PORTAL_ERR_ARG, "batch paths too long for LDS staging");
    const int rt_blocks = n_cu * std::max(1, std::min(8 / std::max(1, best_w / 2), (int)((160 * 1024) / rt_lds)));
    const size_t se_lds = se_lds_bytes(cd_max, nt_max, true);
    const int se_blocks = n_cu * std::max(1, std::min(16, (int)((160 * 1024) / se_lds)));
    // The SE solves: one persistent launch whose LDS fits the largest system (form 0), or
    // a persistent launch per class of models with one system size (form 1), each with
    // its own waves per system and LDS; there the residual's copy of the matrix lives in
    // a global slab per workgroup, which leaves room for more systems per CU.  The
    // autotuner times both forms on every k-th item.
    struct SeClass {
        int cd = 0, nt = 1, waves = 1, blocks = 0;
        std::vector<int2> items, sample;
        int2 *d_items = nullptr, *d_sample = nullptr;
    };
    std::vector<SeClass> se_cls;
    int se_form = 1, se_waves_or = 0;
    int *d_se_cnt = nullptr;
    double *d_slabs = nullptr;
    auto free_se_cls = [&]() {
        for (auto &k : se_cls) {
            dfree(k.d_items);
            dfree(k.d_sample);
        }
        dfree(d_se_cnt);
        dfree(d_slabs);
    };
    Defer se_at_exit{[&] { free_se_cls(); }};
    {
        std::vector<int> cls_of(n, -1);
        for (int m = 0; m < n; m++) {
            for (size_t k = 0; k < se_cls.size(); k++)
                if (se_cls[k].cd == ses[m].c_dim) cls_of[m] = (int)k;
            if (cls_of[m] < 0) {
                cls_of[m] = (int)se_cls.size();
                se_cls.emplace_back();
                se_cls.back().cd = ses[m].c_dim;
            }
            se_cls[cls_of[m]].nt = std::max(se_cls[cls_of[m]].nt, ses[m].n_trans);
        }
        for (const int2 &itm : si) se_cls[cls_of[itm.x]].items.push_back(itm);
        const size_t stride = std::max<size_t>(1, si.size() / 65536);
        std::vector<int2> gsample;
        for (size_t i = 0; i < si.size(); i += stride) gsample.push_back(si[i]);
        size_t slab_doubles = 1;
        for (auto &k : se_cls) {
            for (size_t i = 0; i < k.items.size(); i += stride) k.sample.push_back(k.items[i]);
            HIP_TRY(&c0, upload(k.d_items, k.items));
            HIP_TRY(&c0, upload(k.d_sample, k.sample));
            for (int w : {1, 2, 4})
                slab_doubles = std::max(slab_doubles, (size_t)n_cu * se_slab_blocks_per_cu(w, k.cd, k.nt) *
                                                          se_slab_doubles(k.cd, k.nt));
        }
        HIP_TRY(&c0, hipMalloc((void **)&d_se_cnt, sizeof(int) * std::max<size_t>(1, se_cls.size())));
        HIP_TRY(&c0, hipMalloc((void **)&d_slabs, sizeof(double) * slab_doubles));
        const int forced_form = c0.se_forced_form, forced_w = c0.se_forced_waves;
        double ms_cls = 0.0;
        for (auto &k : se_cls) {
            double best = 1e30;
            if (reuse) {
                for (const auto &cw : c0.batch_se_waves)
                    if (cw.first == k.cd) {
                        k.waves = cw.second;
                        k.blocks = n_cu * se_slab_blocks_per_cu(cw.second, k.cd, k.nt);
                    }
                if (k.blocks > 0) {
                    se_waves_or |= k.waves;
                    continue;
                }
            }
            for (int w : {1, 2, 4}) {
                if (forced_w > 0 && w != forced_w) continue;
                const int bpc = se_slab_blocks_per_cu(w, k.cd, k.nt);
                if (bpc < 1) continue;
                if (forced_form == 0 && k.blocks > 0) break;  // the classes are not used: any runnable choice
                HIP_TRY(&c0, launch_fill_zero(d_se_cnt, sizeof(int), s));
                Timer t(s);
                hipError_t e = launch_se_slab(d_se, k.d_sample, (int)k.sample.size(), d_se_cnt, k.cd, k.nt, w,
                                              n_cu * bpc, d_slabs, s);
                const double ms = t.stop();
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    continue;
                }
                if (ms < best || k.blocks == 0) {
                    best = ms;
                    k.waves = w;
                    k.blocks = n_cu * bpc;
                }
            }
            if (k.blocks == 0 && forced_form != 0) {
                free_se_cls();
                free_xths();
                return fail(c0, PORTAL_ERR_ARG, "SE system of dimension " + std::to_string(k.cd) +
                                                    " does not fit one workgroup");
            }
            ms_cls += best;
            se_waves_or |= k.waves;
        }
        double ms_one = 1e30;
        if (forced_form != 1 && se_lds <= 160 * 1024 && !reuse) {
            int2 *d_gs = nullptr;
            HIP_TRY(&c0, upload(d_gs, gsample));
            HIP_TRY(&c0, launch_fill_zero(d_cnt, sizeof(int) * n_cnt, s));
            Timer t(s);
            hipError_t e = launch_se_persistent(d_se, d_gs, (int)gsample.size(), d_cnt + 1, cd_max, nt_max, se_blocks, s);
            ms_one = t.stop();
            if (e != hipSuccess) {
                (void)hipGetLastError();
                ms_one = 1e30;
            }
            dfree(d_gs);
        }
        se_form = forced_form >= 0 ? forced_form : reuse ? c0.batch_se_form : (ms_one <= ms_cls ? 0 : 1);
        if (se_form == 0 && se_lds > 160 * 1024) {
            free_se_cls();
            free_xths();
            return fail(c0, PORTAL_ERR_ARG, "SE system too large for one workgroup's LDS");
        }
        if (se_form == 0) se_waves_or = 1;
        // the timing launches solved systems of partial fields: their flags do not count
        HIP_TRY(&c0, hipMemset(c0.d_flags, 0, sizeof(int) * 4));
        if (c0.batch_rt_form < 0 && forced_form < 0 && forced_w == 0) {
            c0.batch_key = bkey;
            c0.batch_flat = coop_flat;
            c0.batch_btf = coop_btf;
            c0.batch_se_form = se_form;
            c0.batch_se_waves.clear();
            for (const auto &k : se_cls) c0.batch_se_waves.push_back({k.cd, k.waves});
        }
    }
    // one iteration captured as a graph over the models still iterating: a model leaves
    // once it has converged, as its single-model iteration stops there (the work lists
    // are compacted in place, in their cost order, and the graph captured again)
    std::vector<char> live(n, 1);
    std::vector<int> n_live_cls(classes.size()), n_live_se(se_cls.size());
    int n_live_ri = n_items, n_live_si = (int)si.size();
    auto compact = [&](const std::vector<int2> &all, int2 *d, int &cnt) -> hipError_t {
        std::vector<int2> v;
        v.reserve(all.size());
        for (const int2 &x : all)
            if (live[x.x]) v.push_back(x);
        cnt = (int)v.size();
        return v.empty() ? hipSuccess : hipMemcpy(d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice);
    };
    auto compact_all = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (coop)
            for (size_t k = 0; k < classes.size() && e == hipSuccess; k++)
                e = compact(classes[k].items, classes[k].d_items, n_live_cls[k]);
        else
            e = compact(ri, d_ri, n_live_ri);
        if (e != hipSuccess) return e;
        if (se_form == 0) return compact(si, d_si, n_live_si);
        for (size_t k = 0; k < se_cls.size() && e == hipSuccess; k++)
            e = compact(se_cls[k].items, se_cls[k].d_items, n_live_se[k]);
        return e;
    };
    for (size_t k = 0; k < classes.size(); k++) n_live_cls[k] = (int)classes[k].items.size();
    for (size_t k = 0; k < se_cls.size(); k++) n_live_se[k] = (int)se_cls[k].items.size();
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Defer graph_at_exit{[&] {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    }};
    auto capture = [&]() -> hipError_t {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        launch_fill_zero(d_cnt, sizeof(int) * n_cnt, s);
        launch_fill_zero(d_se_cnt, sizeof(int) * std::max<size_t>(1, se_cls.size()), s);
        launch_fill_zero(d_max, sizeof(unsigned long long) * 2 * n, s);
        for (int m = 0; m < n; m++)
            if (live[m]) launch_line_prep(prep_of(m), s);
        if (coop) {
            for (size_t k = 0; k < classes.size(); k++) {
                const CoopClass &cl = classes[k];
                if (n_live_cls[k] == 0) continue;
                const int bpc = std::max(1, std::min(8, (int)((160 * 1024) / std::max<size_t>(cl.lds, 1))));
                launch_rt_coop_persistent(mixed, d_rt, cl.d_items, n_live_cls[k], d_cnt + 2 + k, cl.cap, cl.slots,
                                          COOP_W, cl.lds, coop_flat ? 0 : n_cu * bpc, cl.dust, coop_btf, cl.prof, s);
            }
        } else if (n_live_ri > 0) {
            launch_rt_persistent(mixed, d_rt, d_ri, n_live_ri, d_cnt, cap, R_max, best_w, rt_blocks, s);
        }
        if (se_form == 0) {
            if (n_live_si > 0)
                launch_se_persistent(d_se, d_si, n_live_si, d_cnt + 1, cd_max, nt_max, se_blocks, s);
        } else {
            for (size_t k = 0; k < se_cls.size(); k++) {
                const SeClass &cl = se_cls[k];
                if (n_live_se[k] == 0) continue;
                launch_se_slab(d_se, cl.d_items, n_live_se[k], d_se_cnt + k, cl.cd, cl.nt, cl.waves, cl.blocks,
                               d_slabs, s);
            }
        }
        for (int m = 0; m < n; m++) {
            if (!live[m]) continue;
            Ctx &c = xs[m]->c;
            if (use_surr[m])
                launch_fill_zero(sb[m].res_cur, sizeof(unsigned long long) * std::max<size_t>(c.g.work.size(), 1), s);
            launch_lambda_update((int)c.g.work.size(), c.g.N, c.sp.e_num, c.g.d_work, c.g.d_ab, c.g.d_h2,
                                 c.sp.d_jlev, c.d_pol, prev[m], c.d_levp, d_max + 2 * m, s, prev[m], sb[m].res_cur);
        }
        e = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess) return e;
        return hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    };
    {
        hipError_t ce = capture();
        if (ce != hipSuccess) {
            c0.err = std::string("graph capture failed: ") + hipGetErrorString(ce);
            return PORTAL_ERR_DEVICE;
        }
    }
    std::vector<unsigned long long> mx(2 * n);
    std::vector<int> conv_at(n, 0);
    std::vector<double> last(n, 0.0), last_al(n, 0.0);
    int iters = 0;
    double t_iter = 0;
    // Ng acceleration per model between replays, on the same schedule as the
    // single-model iteration: from the fourth iterate on, every third
    const bool ng = it->ng_accel == 1;
    std::vector<double *> hist(3 * (size_t)n, nullptr);
    double *d_ngp = nullptr, *d_ngs = nullptr;
    auto free_ng = [&]() {
        for (auto *&p : hist) dfree(p);
        dfree(d_ngp);
        dfree(d_ngs);
    };
    Defer ng_at_exit{[&] { free_ng(); }};
    if (ng) {
        for (int m = 0; m < n; m++)
            for (int h = 0; h < 3; h++)
                HIP_TRY(&c0, hipMalloc((void **)&hist[3 * m + h],
                                       sizeof(double) * xs[m]->c.g.N * xs[m]->c.sp.e_num));
        HIP_TRY(&c0, hipMalloc((void **)&d_ngp, sizeof(double) * ng_partial_doubles()));
        HIP_TRY(&c0, hipMalloc((void **)&d_ngs, sizeof(double) * 8 * n));
    }
    int nhist = 0, n_ng = 0;
    std::vector<int> last_ng(n, -1);  // per model, the iteration of its last extrapolation
    std::vector<double> sums(8 * (size_t)n);
    std::vector<int> nfb(n, 0);
    double t_mlp = 0.0;
    for (int k = 0; k < it->max_iter; k++) {
        Timer ti(s);
        HIP_TRY(&c0, hipGraphLaunch(exec, s));
        HIP_TRY(&c0, hipMemcpyAsync(mx.data(), d_max, sizeof(unsigned long long) * 2 * n, hipMemcpyDeviceToHost, s));
        t_iter += ti.stop();
        iters = k + 1;
        if (it->surrogate && k <= 1) {
            HIP_TRY(&c0, hipStreamSynchronize(s));
            for (int m = 0; m < n; m++) {
                if (!use_surr[m] || !live[m]) continue;
                Ctx &c = xs[m]->c;
                const int nw = (int)c.g.work.size();
                const size_t NL = (size_t)c.g.N * c.sp.e_num;
                if (k == 0) {
                    // keep the plain first iterate and its per-item change, then the
                    // network's prediction replaces it wherever it passes the validity tests
                    HIP_TRY(&c0, hipMemcpyAsync(sb[m].res_p, sb[m].res_cur, sizeof(unsigned long long) * std::max(nw, 1),
                                                hipMemcpyDeviceToDevice, s));
                    HIP_TRY(&c0, hipMemcpyAsync(sb[m].plain_lev, c.d_levp, sizeof(double) * NL, hipMemcpyDeviceToDevice, s));
                    HIP_TRY(&c0, hipMemcpyAsync(sb[m].plain_pol, c.d_pol, sizeof(double) * 2 * NL, hipMemcpyDeviceToDevice, s));
                    HIP_TRY(&c0, hipStreamSynchronize(s));
                    int f = 0;
                    int rc = surrogate_step(c, &f, sb[m].iflags, &t_mlp);
                    if (rc) return fail(c0, rc, "model " + std::to_string(m) + ": " + c.err);
                    nfb[m] += f;
                } else {
                    // the residual test of the single-model iteration
                    SurrRevertArgs ra{};
                    ra.n_work = nw;
                    ra.N = c.g.N;
                    ra.L = c.sp.e_num;
                    ra.work = c.g.d_work;
                    ra.res_s = sb[m].res_cur;
                    ra.res_p = sb[m].res_p;
                    ra.tol = it->fallback_tol;
                    ra.plain_lev = sb[m].plain_lev;
                    ra.plain_pol = sb[m].plain_pol;
                    ra.levp = c.d_levp;
                    ra.pol = c.d_pol;
                    ra.flags = sb[m].iflags;
                    ra.n_fallback = sb[m].nrev;
                    HIP_TRY(&c0, launch_surr_revert(ra, s));
                    int nrev = 0;
                    HIP_TRY(&c0, hipMemcpyAsync(&nrev, sb[m].nrev, sizeof(int), hipMemcpyDeviceToHost, s));
                    HIP_TRY(&c0, hipStreamSynchronize(s));
                    nfb[m] += nrev;
                }
                // the next pass is compared with the state it starts from
                HIP_TRY(&c0, hipMemcpyAsync(prev[m], c.d_pol, sizeof(double) * 2 * NL, hipMemcpyDeviceToDevice, s));
            }
        }
        bool all = true, left = false;
        for (int m = 0; m < n; m++) {
            if (!live[m]) continue;
            double dp, da;
            std::memcpy(&dp, &mx[2 * m], 8);
            std::memcpy(&da, &mx[2 * m + 1], 8);
            last[m] = dp;
            last_al[m] = da;
            // the first pass has no previous alignment to compare with, unless it continues
            if ((k > 0 || chained[m]) && dp <= it->tol_pop && da <= it->tol_align) {
                conv_at[m] = k + 1;
                live[m] = 0;
                left = true;
            }
            all = all && conv_at[m];
        }
        if (all) break;
        if (left && k + 1 < it->max_iter) {
            HIP_TRY(&c0, hipStreamSynchronize(s));
            HIP_TRY(&c0, compact_all());
            hipError_t ce = capture();
            if (ce != hipSuccess) {
                c0.err = std::string("graph capture failed: ") + hipGetErrorString(ce);
                return PORTAL_ERR_DEVICE;
            }
        }
        if (ng) {
            // history x1 = hist[3m] (newest) .. x3 = hist[3m+2]; x0 the level densities just formed
            bool applied = false;
            if (nhist == 3 && k >= 3) {
                for (int m = 0; m < n; m++) {
                    if (!live[m]) continue;
                    Ctx &c = xs[m]->c;
                    HIP_TRY(&c0, launch_ng_dots((int)c.g.work.size(), c.sp.e_num, c.g.d_work, c.d_levp, hist[3 * m],
                                                hist[3 * m + 1], hist[3 * m + 2], d_ngp, d_ngs + 8 * m, s));
                }
                HIP_TRY(&c0, hipMemcpyAsync(sums.data(), d_ngs, sizeof(double) * 8 * n, hipMemcpyDeviceToHost, s));
                HIP_TRY(&c0, hipStreamSynchronize(s));
                for (int m = 0; m < n; m++) {
                    if (!live[m]) continue;
                    Ctx &c = xs[m]->c;
                    const double *q = &sums[8 * (size_t)m];
                    const double a11 = q[0], a12 = q[1], a22 = q[2], b1 = q[3], b2 = q[4];
                    const double det = a11 * a22 - a12 * a12;
                    if (!(std::isfinite(det) && det > 1e-14 * a11 * a22 && a11 > 0 && a22 > 0)) continue;
                    const double ca = (b1 * a22 - b2 * a12) / det, cb = (b2 * a11 - b1 * a12) / det;
                    if (!(std::isfinite(ca) && std::isfinite(cb))) continue;
                    HIP_TRY(&c0, launch_ng_apply((int)c.g.work.size(), c.sp.e_num, c.g.d_work, c.d_levp, hist[3 * m],
                                                 hist[3 * m + 1], ca, cb, s));
                    applied = true;
                    last_ng[m] = k;
                }
                if (applied) n_ng++;
                nhist = 0;  // the extrapolated iterate starts a new history
            }
            for (int m = 0; m < n; m++) {
                if (!live[m]) continue;
                Ctx &c = xs[m]->c;
                std::swap(hist[3 * m + 2], hist[3 * m + 1]);
                std::swap(hist[3 * m + 1], hist[3 * m]);
                HIP_TRY(&c0, hipMemcpyAsync(hist[3 * m], c.d_levp, sizeof(double) * c.g.N * c.sp.e_num,
                                            hipMemcpyDeviceToDevice, s));
            }
            nhist = std::min(nhist + 1, 3);
        }
    }
    free_ng();
    int flags = 0;
    HIP_TRY(&c0, hipMemcpy(&flags, c0.d_flags, sizeof(int), hipMemcpyDeviceToHost));
    hipGraphExecDestroy(exec);
    exec = nullptr;
    hipGraphDestroy(graph);
    graph = nullptr;
    dfree(d_rt);
    dfree(d_se);
    dfree(d_ri);
    dfree(d_si);
    dfree(d_cnt);
    dfree(d_max);
    for (auto *&p : prev) dfree(p);
    free_xths();
    free_classes();
    free_se_cls();
    const double ms = ttot.stop();
    if (flags & FLAG_TEMP) return fail(c0, PORTAL_ERR_TEMP, "kinetic temperature above the collision-rate table");
    if (flags & FLAG_LU) return fail(c0, PORTAL_ERR_LU, "Density-matrix calculation: failed (singular SE matrix)");
    long long ncross = 0;
    bool all_conv = true;
    for (int m = 0; m < n; m++) {
        Ctx &c = xs[m]->c;
        ncross += c.paths.total;
        // d_pol formed d_levp in the model's last pass, unless an extrapolation followed it
        c.pol_chain = last_ng[m] < (conv_at[m] ? conv_at[m] : iters) - 1;
        all_conv = all_conv && conv_at[m];
        if (use_surr[m]) {
            const int nw = (int)c.g.work.size();
            c.fallback_map.assign(std::max(nw, 1), 0);
            HIP_TRY(&c0, hipMemcpy(c.fallback_map.data(), sb[m].iflags, sizeof(int) * std::max(nw, 1),
                                   hipMemcpyDeviceToHost));
            c.fallback_map.resize(nw);
        }
        if (per_model) {
            per_model[m].fallbacks = nfb[m];
            per_model[m].converged = conv_at[m] ? 1 : 0;
            per_model[m].iterations = conv_at[m] ? conv_at[m] : iters;
            per_model[m].max_rel_change = last[m];
            per_model[m].max_align_change = last_al[m];
            per_model[m].n_cross = c.paths.total;
            per_model[m].n_paths = (long long)c.paths.n_work * c.paths.D;
        }
    }
    if (total) {
        total->t_total_ms = ms;
        total->t_rt_ms = t_iter;
        total->iterations = iters;
        total->n_cross = ncross;
        total->rt_waves = best_w;
        total->rt_variant = coop ? (coop_flat ? 3 : 2) : 1;  // 3: the cooperative form, one workgroup per item
        total->rt_order = coop_btf ? 1 : 0;
        total->ng_steps = n_ng;
        total->se_form = se_form;
        total->se_waves = se_waves_or;
        total->max_path = cap;
        total->converged = all_conv ? 1 : 0;
        for (int m = 0; m < n; m++) total->fallbacks += nfb[m];
        total->t_mlp_ms = t_mlp;
    }
    return PORTAL_OK;
}

}  // extern "C"
