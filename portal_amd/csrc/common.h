// Internal declarations shared by the host API and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "portal_hip.h"

namespace portal {

// One stored crossing of an inward ray: the cell, its nint(delta/dx) length
// (inverse_rt.f90:196) and the cell velocity projected on the ray (inverse_rt.f90:103).
struct __attribute__((aligned(16))) PathEntry {
    int cell;     // 0-based node id
    int len;      // length in units of dx
    double vray;  // v(cell) . dir  [m/s]
};

// Geometry-invariant inward paths of every (work node, ray).
struct PathSet {
    PathEntry *entries = nullptr;   // device
    long long *offsets = nullptr;   // device, n_work*D + 1
    int *counts = nullptr;          // device, n_work*D
    long long total = 0;
    int n_work = 0, D = 0;
    int max_nc = 0;
    size_t cap_entries = 0;
    // key of the geometry the paths were built for
    std::vector<double> xth, gvec, bvec;
    double dx = 0;
    int n_depth = 0;
    bool valid = false;
    // star mode: per work item D rays (mu, phi, weight), rays in use, rays towards the star
    int star = 0, nth = 0, nph = 0;
    double R_star = 0;
    std::vector<double> X_star;
    double *d_xthn = nullptr;
    int *d_nstar = nullptr;
    // chunk store of the single-pass walk (kept for the next build)
    PathEntry *chunks = nullptr;
    int2 *chunk_meta = nullptr;
    long long cap_chunks = 0;
    bool single_pass = false;  // the last build used the single-pass walk
};

// Statistical-equilibrium coefficient tables of one species (host-built, FP64).
struct SeTables {
    int c_dim = 0, n_trans = 0;
    // per matrix entry (CSR over the c_dim*c_dim entries that have terms)
    std::vector<int> ent_idx;       // entry = row + c_dim*col
    std::vector<int> ent_ptr;       // n_ent+1
    std::vector<int> term_q;        // scalar index 7*t + kind
    std::vector<double> term_coef;
    std::vector<int> term_class;    // 0 spontaneous, 1 radiative, 2 collisional
    int *d_ent_idx = nullptr, *d_ent_ptr = nullptr, *d_term_q = nullptr, *d_term_class = nullptr;
    double *d_term_coef = nullptr;
};

// Species after the driver's conversions, plus derived lists.
struct Species {
    int e_num = 0, r_num = 0, c_num = 0, c_temp = 0, c_par = 0, kmax = 0;
    std::vector<int> j_lev, r_trans, c_trans, all_trans, count_array;
    std::vector<double> e_levs_hz, t_Aji, nu_hz, t_Bij, kd_v, t_Cji, col_temps;
    int n_trans = 0, c_dim = 0;
    SeTables se;
    // per unique transition: radiative line index (-1 if none), collisional index (-1 if none)
    std::vector<int> trans_rad, trans_col;
    // device copies
    int *d_l1 = nullptr, *d_l2 = nullptr;                 // per line, 0-based levels
    double *d_c1B = nullptr, *d_c2A = nullptr, *d_g1 = nullptr, *d_g2 = nullptr;
    double *d_cnu_spi = nullptr, *d_Iin = nullptr;
    double *d_kd = nullptr, *d_nu = nullptr;              // per line: dust opacity, frequency
    double *d_tA = nullptr, *d_tB21 = nullptr, *d_tnu = nullptr, *d_tgul = nullptr;
    int *d_trad = nullptr, *d_tcol = nullptr;
    int *d_tlo = nullptr, *d_tup = nullptr;               // per transition, 0-based levels
    double *d_Cji = nullptr, *d_coltemps = nullptr;
    int *d_count = nullptr, *d_jlev = nullptr;
    // collisional depolarisation table (c_temp x e_num x (kmax - 1)), or null for none
    double *d_depol = nullptr;
    bool valid = false;
};
int depol_table(const Species &s, std::vector<double> &out);

struct Grid {
    int N = 0, N_b = 0;
    double tx = 0;
    std::vector<double> XYZ, vel, temp, b_all, h2, ab, dust;
    std::vector<int> nbr_ptr, nbr_idx;  // CSR, 0-based ids
    bool has_dust = false;
    double b_ratio = 0;  // largest over smallest positive finite Doppler width
    // device
    double4 *d_pos = nullptr;    // x, y, z, 0
    double4 *d_velb = nullptr;   // vx, vy, vz, b
    double *d_inv_b = nullptr;   // 1/b
    double *d_tdust = nullptr;   // T_dust
    double *d_tkin = nullptr;
    double *d_h2 = nullptr;
    double *d_dust = nullptr;
    double *d_ab = nullptr;
    int *d_nbr_ptr = nullptr, *d_nbr_idx = nullptr;
    // work list: bulk nodes with ab_all > 0 (main_para.f90:211-213), Morton-sorted
    std::vector<int> work;
    int *d_work = nullptr;
    bool valid = false;
};

// autotuner candidates of the inward integration whose timings a context keeps
constexpr int RT_TUNE_SLOTS = 12;

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Grid g;
    Species sp;
    PathSet paths;
    // device lev_p (L x N) and per-iteration buffers
    double *d_levp = nullptr;
    double *d_KS = nullptr;     // N x R x {K, S}  (FP64)
    float *d_KS32 = nullptr;    // N x R x {K, S}  (FP32, mixed path)
    double *d_KSD = nullptr;    // N x R x {alpha_dust, B_nu(T_dust)}, when the grid has dust
    float *d_KSD32 = nullptr;
    double *d_dustco = nullptr; // N x R x {Ka, j} when dust is present
    double *d_J0 = nullptr, *d_J2 = nullptr;  // R x N
    double *d_LAM = nullptr;    // R x N, own-cell operator of the accelerated lambda iteration
    bool ali = false;           // the lambda iteration in progress uses it
    double *d_pol = nullptr;    // N x 2L
    double *d_rates = nullptr;  // N x 3 x 2L
    int *d_flags = nullptr;     // error flags
    size_t cap_nodes = 0;
    // scratch of the image pass
    void *d_img = nullptr;
    size_t img_bytes = 0;
    // MLP surrogate state (owned by mlp.hip)
    void *mlp = nullptr;
    // autotuned inward-integration launch (per precision): variant, waves per block
    int tuned_key[2] = {0, 0};
    int tuned_variant[2] = {0, 0}, tuned_waves[2] = {4, 4};
    double tune_ms[2][RT_TUNE_SLOTS] = {{0}};
    int forced_variant[2] = {-1, -1}, forced_waves[2] = {4, 4};  // test hook: bypass the autotuner
    double walk_chunk_scale = 1.1;  // chunk store of the single-pass walk / sampled estimate
    int ng_first = 3, ng_gap = 3;   // Ng schedule: first iteration, least gap between extrapolations
    int coop_unit = 0;              // cells per staged unit of the cooperative kernel (0: COOP_UNIT)
    std::string scratch_err;        // why the allocation-free entry point's scratch is unavailable
    int surr_perturb_stride = 0;    // test hook of the surrogate fallback (0: off)
    double surr_perturb_factor = 1.0;
    std::vector<int> fallback_map;  // per work item of the last surrogate run: 0 kept, 1 invalid, 2 residual
    int se_waves = 1;               // waves per system of the single-model SE solve
    int se_force_slab = 0;          // test hook: the single-model solve in the slab form (1), with the factors in the slab too (2)
    // test hook for a batch led by this context: SE form (-1 autotuned, 0 one persistent
    // launch, 1 a launch per class) and the waves per system of the classes (0 autotuned)
    int se_forced_form = -1, se_forced_waves = 0;
    // test hook: the cooperative form of a batch this context leads (-1 timed; bit 0 one
    // workgroup per item, bit 1 rays back to front)
    int batch_rt_form = -1;
    // a batch this context leads: the timed choices of its last call, kept for the next call
    // over the same models with the same paths and shapes (batch_key)
    unsigned long long batch_key = 0;
    bool batch_flat = false, batch_btf = false;
    int batch_se_form = 1;
    std::vector<std::pair<int, int>> batch_se_waves;  // (c_dim, waves per system)
    // d_pol holds the SE solution from which the current d_levp was formed (set by the
    // lambda iteration; cleared by anything that changes the state otherwise), so that
    // a lambda iteration continued by a further call measures its first change exactly
    bool pol_chain = false;
    bool pol_valid = false;  // d_pol is the SE solution of the J in d_J0/d_J2
    std::vector<double> iter_trace;  // per iteration of the last lambda call: population change, alignment change
};

// error flags (bits in d_flags[0])
enum : int { FLAG_DEPTH = 1, FLAG_LU = 2, FLAG_TEMP = 4, FLAG_DEPTH_RT = 8, FLAG_START = 16, FLAG_CHUNKS = 32 };

#define HIP_TRY(ctx, expr)                                                          \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            return PORTAL_ERR_DEVICE;                                               \
        }                                                                           \
    } while (0)

// ---- launchers implemented in the kernel files ------------------------------------
struct WalkArgs {
    int n_work, D, N;
    const int *work;
    const double4 *pos;
    const int *nbr_ptr, *nbr_idx;
    const double *xth;    // 3 x D
    const double *bvec;   // 3 x N
    const double *gvec;   // 3
    const double4 *velb;
    double dx;
    int n_depth;
    int *counts;
    const long long *offsets;
    PathEntry *paths;
    int *flags;
    // star mode (xthn null: off): per work item D rays (mu, phi, w) and the counts of rays
    // in use and towards the star; rays towards the star end on its surface
    const double *xthn;
    const int *nvar, *nstar;
    double Xs0, Xs1, Xs2, R_star;
    // single-pass form: entries go to fixed-size chunks taken from a counter, each chunk
    // tagged with its (path, ordinal); a compaction then writes the contiguous layout
    PathEntry *chunks;
    int2 *chunk_meta;
    unsigned long long *chunk_next;
    long long chunk_cap;
    // sampling form (count mode): walk only every sample-th work item
    int sample;
};
// walk modes: count (counts and the longest path), write (entries at offsets), chunked
// (counts and entries in chunks, one pass)
enum WalkMode : int { WALK_COUNT = 0, WALK_WRITE = 1, WALK_CHUNKED = 2 };
constexpr int WALK_CHUNK = 16;  // entries per chunk
constexpr int COOP_UNIT = 64;   // cells per staged unit of the cooperative inward integration
hipError_t launch_walk(const WalkArgs &a, int mode, hipStream_t s);
hipError_t launch_walk_compact(const WalkArgs &a, long long n_chunks, hipStream_t s);
int build_star_dirs(const std::vector<int> &work, const double *XYZ, const double *bvec_all, const double *gvec,
                    int nth, int nph, const double *X_star, double R_star, std::vector<double> &xthn,
                    std::vector<int> &nvar, std::vector<int> &nstar);
hipError_t launch_scan_counts(const int *counts, long long *offsets, long long n, hipStream_t s,
                              long long *total_host);

struct PrepArgs {
    int N, R, L;
    const double *levp;      // L x N
    const double4 *velb;
    const int *l1, *l2;      // 0-based level of each line (device)
    const double *c1B;       // per line: 2.2102191e-42*nu * B_lu
    const double *c2A;       // per line: 5.272859e-35*nu * A
    const double *gratio;    // per line: g_l / g_u   computed as (1.0*g1)/(1.0*g2) split below
    const double *g1, *g2;
    const double *cnu_spi;   // per line: c/nu / sqrt(pi)
    double *KS;
    float *KS32;
    // dust (KSD/KSD32 null: none): per node dust_p and T_dust, per line kd and nu
    const double *dust, *tdust, *kd, *nu;
    double *KSD;
    float *KSD32;
    const double *I_in;      // per line background, which sets the unit of S in KS32/KSD32
};
hipError_t launch_line_prep(const PrepArgs &a, hipStream_t s);

// The FP32 inward pass carries a line's intensities (source functions, backgrounds,
// and so J before the FP64 channel sum) in units of 2^e, e = line_iexp(I_in): the
// power of two at or below the line's background B_nu(2.73 K), but not below 2^-100
// and not above 1 (W m^-2 Hz^-1 sr^-1).  The backgrounds of lines above ~3 THz lie
// below FP32's range in those SI units; in units of 2^e they are 1 (or, under the
// 2^-100 floor, still normal FP32 numbers for any background above 1e-68), while a
// source function of 1e-9, far above any thermal gas's, stays below 2^72.  A power of
// two makes the scaling and its undoing exact.
__host__ __device__ inline int line_iexp(double iin) {
    if (!(iin > 0.0) || !(iin < 1e300)) return -100;
    const int e = ilogb(iin);
    return e < -100 ? -100 : (e > 0 ? 0 : e);
}

struct RtArgs {
    int n_work, D, R, nvel, C;
    const int *work;
    const double4 *velb;
    const double *inv_b;
    const double *KS;       // N x R x 2
    const float *KS32;
    const double *dustco;   // N x R x 2 or null
    const double *xth;      // 3 x D
    const long long *offsets;
    const PathEntry *paths;
    const double *I_in;     // per line
    double dl_scale;        // tx*dx
    double *J0, *J2;        // R x N
    int waves;              // waves per block
    int max_nc;             // longest stored path (LDS staging capacity)
    int variant;            // 0: global-memory kernel, 1: LDS-staged per-wave rays, 2: cooperative
    const double *KSD;      // N x R x {alpha_dust, B_nu(T_dust)} when dust is on (cooperative kernel only)
    const float *KSD32;
    int dust;
    // star mode (cooperative kernel only): per work item D rays (mu, phi, w), rays towards
    // the star (their background is I_star per line instead of I_in)
    const double *xthn;
    const int *nstar;
    const double *I_star;
    // accelerated lambda iteration (cooperative kernel only): per (node, line) the share of
    // the node's own cell in its J^0_0 per unit source function, sum over rays and channels
    // of w phi0 d_v (1 - e^-tau_own) (R x N), or null
    double *LAM;
    int unit;               // cooperative kernel: cells per staged unit (0: COOP_UNIT)
    double b_ratio;         // largest over smallest positive Doppler width of the grid (0: unknown)
    // the autotuner's sample of the work list: with item_run > 0, block b integrates work
    // item (b / item_run) * item_stride + b % item_run, runs of consecutive items spread
    // over the list (0: block b integrates item b)
    int item_run, item_stride;
};

// the work item of block b (RtArgs.item_run, item_stride)
__device__ __forceinline__ int rt_item(const RtArgs &a, int b) {
    return a.item_run > 0 ? (b / a.item_run) * a.item_stride + b % a.item_run : b;
}
hipError_t launch_rt_f64(const RtArgs &a, hipStream_t s);
hipError_t launch_rt_mixed(const RtArgs &a, hipStream_t s);

struct SeArgs {
    int n_work, N, L, R, c_dim, n_trans, n_ent;
    const int *work;
    const double *J0, *J2;      // R x N
    const double *tkin, *h2;    // N
    // per transition
    const double *t_A, *t_B21, *t_nu, *t_gul;   // A, B21 (a_to_b*A/nu^3), nu (levels), g_u/g_l
    const int *t_rad, *t_col;                   // radiative line / collisional index or -1
    const double *t_Cji;                        // c_temp x c_num x c_par
    const double *col_temps;
    int c_temp, c_par, c_num;
    const int *ent_idx, *ent_ptr, *term_q, *term_class;
    const double *term_coef;
    const int *count_array, *j_lev;             // L
    // accelerated lambda iteration, or null: lam (R x N) from the inward pass, the level
    // densities (L x N) that produced J, and each transition's lower/upper level (0-based)
    const double *lam, *levp;
    const int *t_lo, *t_up;
    // collisional depolarisation D^(K) (c_temp x L x (kmax - 1) on the collision
    // temperatures), or null: -D^(K)(T) n_H2 1e-6 on the diagonal of each K >= 2 row
    const double *depol;
    double *pol;                                // N x 2L
    double *rates;                              // N x 3 x 2L or null
    int *flags;
    int extra_info;
    int waves;
};
hipError_t launch_se(const SeArgs &a, hipStream_t s);

// surrogate features / decoding (lambda.hip) and network inference (mlp.hip)
struct SurrArgs {
    int n_work, N, L, R, n_in, n_out;
    const int *work;
    const double *h2, *tkin, *ab, *J0, *J2, *Iin;
    const int *j_lev;
    const int *l1, *l2;      // per line: lower and upper level (0-based)
    const double *g1, *g2;   // per line: their statistical weights
    float *feats;            // n_work x n_in
    const float *out;        // n_work x n_out (network outputs, physical units)
    double *levp, *pol;      // updated for accepted nodes
    int *n_fallback;
    int *flags;              // n_work: 1 where the prediction failed the validity tests, else 0 (may be null)
    int perturb_stride;      // test hook (0: off): items w with w % stride == 0 get a wrong prediction
    double perturb_factor;
};
struct SurrRevertArgs {
    int n_work, N, L;
    const int *work;
    const unsigned long long *res_s, *res_p;  // per item, double bits
    double tol;
    const double *plain_lev, *plain_pol;
    double *levp, *pol;
    int *flags, *n_fallback;
};
hipError_t launch_surr_features(const SurrArgs &a, hipStream_t s);
hipError_t launch_surr_apply(const SurrArgs &a, hipStream_t s);
hipError_t launch_surr_revert(const SurrRevertArgs &a, hipStream_t s);
hipError_t mlp_infer_device(Ctx &c, int B, const float *feats, float *out);
int mlp_dims(const Ctx &c, int *n_in, int *n_out);

hipError_t launch_lambda_update(int n_work, int N, int L, const int *work, const double *ab, const double *h2,
                                const int *j_lev, const double *pol, const double *pol_prev, double *levp,
                                unsigned long long *maxes, hipStream_t s, double *pol_save = nullptr,
                                unsigned long long *node_res = nullptr);
// batched (persistent) launches over (model, node) items
hipError_t launch_rt_persistent(bool mixed, const RtArgs *models, const int2 *items, int n_items, int *counter,
                                int cap, int R_max, int waves, int n_blocks, hipStream_t s);
hipError_t launch_se_persistent(const SeArgs *models, const int2 *items, int n_items, int *counter, int cd_max,
                                int nt_max, int n_blocks, hipStream_t s);
// persistent workgroups of `waves` (1, 2 or 4) waves over the items of one system size
// (cd, nt), each with a slab of cd*cd doubles at slabs + blockIdx.x*cd*cd
hipError_t launch_se_slab(const SeArgs *models, const int2 *items, int n_items, int *counter, int cd, int nt,
                          int waves, int n_blocks, double *slabs, hipStream_t s, bool all_global = false);
int se_slab_blocks_per_cu(int waves, int cd, int nt, bool all_global = false);
size_t se_lds_bytes(int c_dim, int n_trans, bool m0_in_lds = true, bool all_global = false);
size_t se_slab_doubles(int c_dim, int n_trans, bool all_global = false);
constexpr int SE_SLAB_WAVES = 4;  // waves per system of a single model's slab-form solve
constexpr int SE_MAX_UNKNOWNS = 512;  // largest SE system (c_dim) the solver takes
size_t rt_persistent_lds(bool mixed, int cap, int R_max, int waves);
// the cooperative form of the persistent integration, one launch per (slot count, dust);
// btf: each ray's cells back to front; prof: the FP32 profile form, at most what
// coop_profile_form_host gives for every model of the launch (2 shared, 1 recurrence,
// 0 one exponential per slot)
hipError_t launch_rt_coop_persistent(bool mixed, const RtArgs *models, const int2 *items, int n_items, int *counter,
                                     int cap, int slots, int waves, size_t lds, int n_blocks, bool dust, bool btf,
                                     int prof, hipStream_t s);
int coop_profile_form_host(int R, int C, int W, int slots, double b_ratio, int nvel);
size_t rt_coop_persistent_lds(bool mixed, int cap, int R_max, int C, bool dust, int waves);
int coop_slots_host(int R, int C, int W);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_ng_dots(int n_work, int L, const int *work, const double *x0, const double *x1, const double *x2,
                          const double *x3, double *partial, double *out5, hipStream_t s);
hipError_t launch_ng_apply(int n_work, int L, const int *work, double *x0, const double *x1, const double *x2,
                           double ca, double cb, hipStream_t s);
size_t ng_partial_doubles();
constexpr int AA_M_MAX = 4;                               // Anderson window
constexpr int AA_NSUM = AA_M_MAX * (AA_M_MAX + 1) / 2 + AA_M_MAX;  // dF.dF (upper) and dF.f sums
hipError_t launch_aa_gather(int n_work, int L, const int *work, const double *levp, double *dst, hipStream_t s);
hipError_t launch_aa_dots(long long n, int mk, const double *x0, const double *g0, const double *const *xj,
                          const double *const *gj, double *partial, double *out, hipStream_t s);
hipError_t launch_aa_apply(int n_work, int L, int mk, const int *work, const double *g0, const double *const *gj,
                           const double *gamma, double *levp, hipStream_t s);
size_t aa_partial_doubles();

}  // namespace portal

struct portal_ctx {
    portal::Ctx c;
};
void portal_mlp_free(portal::Ctx &c);
