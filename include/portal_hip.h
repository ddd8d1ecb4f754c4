/*
 * portal_hip.h -- C ABI of the MI355X (gfx950) implementation of PORTAL's hot path.
 *
 * One call replaces one OpenMP loop of the reference's drivers:
 *
 *   portal_se_pass      the node loop of main_para.f90:210-257, i.e. trace_node
 *                       (main_sub.f90:1-141) for every bulk node: inverse_rt
 *                       (inverse_rt.f90:1-137) + make_AC (miscellaneous.f90:288-392)
 *                       + get_rhovec (em_coefs.f90:115-201) + packing.
 *   portal_inverse_rt   inverse_rt alone for every bulk node (J^0_0, J^2_0).
 *   portal_se_solve     make_AC + get_rhovec + packing from given J's.
 *   portal_image_pass   the three loops of one inclination, para_finish.f90:382-449
 *                       (rotate_all_save, trace_out_raytrace, comp_emcoefs,
 *                       ray_in_raytrace).
 *   portal_lambda_iterate  a build extension with no reference counterpart: the
 *                       SE pass repeated with lev_p <- ab n_H2 sqrt(2J+1) rho^0_0.
 *
 * Conventions (those of the reference's Fortran interface):
 *   - arrays are column-major as declared in the reference; ids are 1-based;
 *   - all host memory belongs to the caller; device memory belongs to the context;
 *   - every function returns PORTAL_OK (0) or a negative status code in place of
 *     the reference's `stop`; portal_last_error() gives the message;
 *   - one context per device, called from one host thread at a time.
 *
 * Fortran binds these through ISO_C_BINDING (portal_amd/fortran/portal_hip_mod.f90).
 */
#ifndef PORTAL_HIP_H
#define PORTAL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PORTAL_OK 0
#define PORTAL_ERR_ARG (-1)          /* bad argument or unsupported size */
#define PORTAL_ERR_DEVICE (-2)       /* HIP runtime error */
#define PORTAL_ERR_DEPTH (-3)        /* node_count > n_depth, inverse_rt.f90:206-210 */
#define PORTAL_ERR_DEPTH_RT (-4)     /* node_count > n_depth_rt, raytrace.f90:169-173 */
#define PORTAL_ERR_LU (-5)           /* singular SE matrix, em_coefs.f90:170-178 */
#define PORTAL_ERR_TEMP (-6)         /* T_kin above the collision table, miscellaneous.f90:331-360 */
#define PORTAL_ERR_STATE (-7)        /* grid/species not uploaded */
#define PORTAL_ERR_NOMEM (-8)        /* device allocation failed */
#define PORTAL_ERR_UNSUPPORTED (-9)  /* a batch option that is for single models */

#define PORTAL_FP64 0   /* every stage in FP64 */
#define PORTAL_MIXED 1  /* FP32 ray transfer + compensated FP32 J sums; FP64 SE */

typedef struct portal_ctx portal_ctx;

typedef struct {
    int n_sphere;            /* directions (legendre_all, miscellaneous.f90:1100-1137) */
    const double *xth;       /* 3 x n_sphere: (mu, phi, weight) */
    int nvel;                /* channels = 2*nvel+1 (inverse_rt.f90:275) */
    double dx;               /* length quantum, the reference passes float 1E-6 (main_para.f90:135) */
    const double *gvec;      /* 3: gauge vector (main_para.f90:172-176) */
    const double *bvec_all;  /* 3 x node_num: unit field directions */
    int n_depth;             /* maximum cells per inward ray */
    int extra_info;          /* 1: also fill pol_rates and JJ_20 */
    int precision;           /* PORTAL_FP64 or PORTAL_MIXED */
    int star_mod;            /* 1: star mode (main_sub.f90:56-79): each node's directions are
                                star_all(nth, nph) about the star instead of xth; rays towards the
                                star end on its surface and start with B_nu(T_star) */
    int nth, nph;            /* star mode: the quadrature of star_all (main_para.f90 nth, nph) */
    double R_star, T_star;   /* star mode: radius (normalised like XYZ) and temperature [K] */
    const double *X_star;    /* star mode: position (3, normalised like XYZ) */
} portal_se_params;

typedef struct {
    int max_iter;            /* lambda iterations at most */
    double tol_pop;          /* stop when max |d(sqrt(g) rho00)| / (sqrt(g) rho00) <= tol_pop */
    double tol_align;        /* ... and max |d(rho20/rho00)| <= tol_align */
    int surrogate;           /* 1: the trained MLP supplies the first iterate */
    double fallback_tol;     /* the surrogate's residual test (>= 0; +inf turns it off): one lambda
                                iteration after the prediction, a node whose largest relative
                                level-population change in that iteration exceeds both fallback_tol
                                and the change its plain first iterate made (from the starting
                                populations) reverts to that plain iterate.  Predictions that fail
                                the validity tests (non-finite, populations summing outside
                                [0.5, 2], |rho2/rho0| >= 0.5) are rejected when applied.
                                stats.fallbacks counts both kinds */
    int rebuild_paths;       /* 1: rebuild the inward paths even if the geometry is unchanged */
    int ng_accel;            /* 0: plain iteration; 1: Ng acceleration of the level densities from the
                                last four iterates, every third iteration from the fourth on;
                                2: Anderson acceleration over a window of 4, every iteration */
    int ali;                 /* 1: accelerated lambda iteration: each node's own-cell share of J^0_0
                                (from the inward pass) moves into the K = 0 rates of the SE solve
                                (Rybicki & Hummer 1991); same fixed point, fewer iterations.
                                Ignored when dust enters the inward pass */
} portal_iter_params;

typedef struct {
    int nlevels;             /* imaged lines */
    const int *l1, *l2;      /* lower / upper level of each imaged line (1-based) */
    const double *nu0_all;   /* line frequencies [Hz] as the driver holds them */
    const double *Aji_all;
    const double *kd_all;
    double th_inc, ph_az;    /* one inclination */
    int nvel;                /* channels = 2*nvel+1 */
    const double *vel_out;   /* 2*nvel+1 channel velocities */
    int im_pix;              /* image is (2*im_pix+1)^2 pixels */
    double im_size;
    double dx;               /* march step, the reference passes float 1E-4 (para_finish.f90:201) */
    double alpha;            /* dust polarisation (para_finish.f90:78) */
    int n_depth_rt;
    int precision;           /* PORTAL_FP64 or PORTAL_MIXED */
    int star_mod;            /* 1: star mode (raytrace.f90:103-114): a pixel within R_star of the
                                rotated star's centre starts on the star's near side,
                                z = Z_star + R_star, with B_nu(T_star) as incoming intensity; the
                                final background subtraction stays B_nu(2.73 K) */
    double R_star, T_star;   /* star mode: radius (normalised like XYZ) and temperature [K] */
    const double *X_star;    /* star mode: position (3, normalised like XYZ, before rotation) */
    int node_pixels;         /* 1: the reference's pix = 0 mode (para_finish.f90:245-330), one pixel
                                per bulk node at its rotated (x, y), outputs bulk_node_num x nlevels
                                x (2nvel+1), im_pix and im_size unused; 0: the (2 im_pix+1)^2 grid */
} portal_img_params;

typedef struct {
    double t_walk_ms;        /* inward path construction (done once per geometry); in the image
                                pass the start search and march */
    double t_prep_ms;        /* per-iteration line coefficients */
    double t_rt_ms;          /* inward ray integration */
    double t_se_ms;          /* SE assembly + solve */
    double t_img_ms;         /* image pass */
    double t_total_ms;       /* device time of the call */
    double rt_kernel_avg_ms; /* average duration of one inward-integration launch */
    long long n_cross;       /* ray-cell crossings integrated per pass (image pass: cells the
                                pixels' rays visit) */
    long long n_paths;       /* (node, ray) paths (image pass: pixels) */
    int iterations;
    int fallbacks;           /* nodes where the surrogate iterate was rejected */
    int lu_failures;
    double max_rel_change;   /* last iteration's population change */
    int rt_variant;          /* inward-integration kernel chosen by the autotuner: 0 global,
                                1 staged, 2 cooperative, 3 cooperative with each ray's cells
                                back to front; in a batch, 1 the staged persistent form, 2 the
                                cooperative persistent form, 3 the cooperative form with one
                                workgroup per item */
    int rt_waves;            /* its waves per workgroup */
    int max_path;            /* longest inward path (cells; image pass: longest pixel ray) */
    int ng_steps;            /* Ng extrapolations applied by the lambda iteration */
    int converged;           /* 1: the lambda iteration met both tolerances */
    double max_align_change; /* last iteration's largest change of rho^2_0/rho^0_0 */
    int se_form;             /* SE solve: 0 one launch (one persistent launch in a batch),
                                1 a batch's launch per class of system sizes */
    int se_waves;            /* waves per SE system; for a launch per class, the bitwise
                                OR of the classes' wave counts (1, 2, 4) */
    double t_mlp_ms;         /* surrogate network inference in the surrogate step (input
                                normalisation and the BF16 GEMMs), 0 without it */
    int rt_order;            /* 1: the inward integration took each ray's cells back to front
                                (as ray_in does), 0: front to back, leaving exhausted rays */
} portal_stats;

/* Context -------------------------------------------------------------------- */
int portal_ctx_create(portal_ctx **ctx, int device);
int portal_ctx_destroy(portal_ctx *ctx);
const char *portal_last_error(const portal_ctx *ctx);

/* Grid: the arrays eat_lime_no_delaunay (eat_lime.f90:71-313) hands to the node
 * loop, after main_para.f90:119-135.  XYZ/vel_all 3 x N, temp_all 2 x N
 * (T_kin, T_dust), del_ne N x maxne (1-based, sorted, zero padded). */
int portal_upload_grid(portal_ctx *ctx, int node_num, int bulk_node_num, const double *XYZ,
                       const double *vel_all, const double *temp_all, const double *b_all,
                       const double *h2_dens, const double *ab_all, const double *dust_p,
                       const int *del_ne, const int *del_num_ne, int maxne, double tx);

/* Species: LAMDA tables after main_para.f90:119-120 and 178-185; the transition
 * list and state index of main_para.f90:190-198 are built here.  r_trans/c_trans
 * are 2 x R / 2 x Cn (lower, upper); t_Cji is c_temp x c_num x c_par. */
int portal_upload_species(portal_ctx *ctx, int e_num, const int *j_lev, const double *e_levs_hz,
                          int r_num, const int *r_trans, const double *t_Aji, const double *nu_hz,
                          const double *t_Bij, const double *kd_v, int c_num, const int *c_trans,
                          int c_temp, int c_par, const double *t_Cji, const double *col_temps,
                          int kmax);
int portal_species_info(const portal_ctx *ctx, int *n_trans, int *c_dim);

/* Collisional depolarisation D^(K) (depol.f90:242-359; the slot pop_eqs.f90:642-646),
 * which the reference computes but switches off (trace_node passes D0 = 0,
 * main_sub.f90:96).  portal_depol_table fills D0_all (c_temp x e_num x (kmax-1),
 * column-major, the reference's D0(nT, nlev, 2:kmax)) from the uploaded species as
 * de_polarization does (depol.f90:242-317: Morris 1985 B11, sudden approximation; k_m is
 * the rate between the J = 0 level and the level with J = m, where the reference's code
 * reads level index m).  portal_set_depol(ctx, D0_all) makes every later SE solve of the
 * context subtract D^(K)(T) n_H2 1e-6 (make_depolarization_temp, depol.f90:319-359,
 * bracketed in T as make_AC brackets the collision rates) on the diagonal of each K >= 2
 * row; NULL switches it off again (the reference's behaviour).  A new species upload
 * clears it. */
int portal_depol_table(const portal_ctx *ctx, double *D0_all);
int portal_set_depol(portal_ctx *ctx, const double *D0_all);

/* Level densities lev_p (e_num x node_num), the populations the inward rays see
 * (inverse_rt.f90:121-124). */
int portal_set_levpop(portal_ctx *ctx, const double *lev_p);
int portal_get_levpop(portal_ctx *ctx, double *lev_p);

/* SE pass: pol_pops node_num x 2L (main_para.f90:202, written 271-275);
 * pol_rates node_num x 3 x 2L and JJ_20 node_num x R when extra_info (may be NULL). */
int portal_se_pass(portal_ctx *ctx, const portal_se_params *p, double *pol_pops, double *pol_rates,
                   double *JJ_20, portal_stats *stats);
/* J0, J2: r_num x node_num (zero for nodes that are not solved). */
int portal_inverse_rt(portal_ctx *ctx, const portal_se_params *p, double *J0, double *J2,
                      portal_stats *stats);
int portal_se_solve(portal_ctx *ctx, const double *J0, const double *J2, double *pol_pops,
                    double *pol_rates, portal_stats *stats);

/* No counterpart in the reference: portal_inverse_rt for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  d_J0, d_J2 are
 * device buffers of r_num x node_num that the caller owns; the direction table (and
 * in star mode the star's intensities) go to a static device array of 256 KiB per
 * device instead of buffers allocated per call, so n_sphere is at most 10,901
 * (PORTAL_ERR_ARG beyond).  Nothing is allocated: the pass buffers, the inward paths
 * for exactly these directions and field, and the launch shape must come from an
 * earlier portal_inverse_rt or portal_se_pass on this context (PORTAL_ERR_STATE
 * otherwise).  The results equal portal_inverse_rt's bit for bit.  Returns when the
 * outputs are written.  Calls that overlap on one device share the static array and
 * race. */
int portal_inverse_rt_device(portal_ctx *ctx, const portal_se_params *p, double *d_J0, double *d_J2);

/* pol_pops (node_num x 2L) of the last SE solve on the device, without recomputing. */
int portal_get_pol_pops(portal_ctx *ctx, double *pol_pops);

/* Lambda iteration (no reference counterpart).  lev_p_out (e_num x node_num, may be
 * NULL) receives the final level densities. */
int portal_lambda_iterate(portal_ctx *ctx, const portal_se_params *p, const portal_iter_params *it,
                          double *pol_pops, double *lev_p_out, portal_stats *stats);

/* Batch of independent models (no reference counterpart): one context per model,
 * all on one device, all with the same precision.  Every iteration is one replay of
 * a captured hipGraph holding each model's line preparation, the inward integration
 * over all (model, node) items ordered by cost (one launch per class of models with the
 * same slot count; persistent workgroups or one workgroup per item, whichever timed
 * faster on the batch), the SE solves (one persistent launch per system size with
 * 1, 2 or 4 waves per system, or one persistent launch over all items, whichever timed
 * faster; the same bits either way) and each model's closure.  Models with dust
 * (dust classes of the cooperative kernel) or in star mode (their per-node direction
 * sets) are integrated as in the single-model pass.  it->ng_accel = 1 applies Ng
 * acceleration per model between replays; Anderson acceleration and the own-cell
 * operator are for single models (PORTAL_ERR_UNSUPPORTED).  it->surrogate = 1 gives
 * every model whose context holds a network (portal_mlp_create + weights) the
 * network's prediction after the first replay and the residual test (fallback_tol)
 * after the second, as portal_lambda_iterate does; models without one iterate
 * plainly.  Iterates until every model meets the tolerances (or max_iter); a model
 * that meets them leaves the iteration there, as portal_lambda_iterate stops, so its
 * results are those of its single-model call.  per_model
 * (n, may be NULL) gets each model's iteration count at convergence, whether it
 * converged and its surrogate fallbacks, total the device time of the whole batch.  A
 * further call continues each model's iteration (not with the surrogate, which starts
 * afresh).  Results stay in each context (portal_get_pol_pops / portal_get_levpop). */
int portal_batch_lambda_iterate(portal_ctx **ctxs, int n_models, const portal_se_params *params,
                                const portal_iter_params *it, portal_stats *per_model, portal_stats *total);

/* Image pass for one inclination.  pol_pops is the rescaled table of
 * para_finish.f90:222-240 (node_num x 2L).  Outputs are n_rt x nlevels x (2nvel+1)
 * (all_I(:,:,inc,:) of para_finish.f90:356-359), n_rt = (2 im_pix+1)^2, or
 * bulk_node_num with node_pixels (para_finish.f90:246-249); dust (4 such cubes) may be
 * NULL.  A pixel whose start point lies off the model sphere returns PORTAL_ERR_ARG
 * (the reference's march would not end there, raytrace.f90:98). */
int portal_image_pass(portal_ctx *ctx, const portal_img_params *p, const double *pol_pops,
                      const double *bvec_all, double *I, double *Q, double *U, double *TAU,
                      double *dust, portal_stats *stats);

/* Surrogate network (no counterpart in the reference).  A dense MLP, depth hidden
 * layers of `width` with SiLU, BF16 weights and activations, FP32 accumulation and
 * FP32 master weights (Adam).  Inputs of a node: log10 n_H2[cm^-3], log10 T_kin;
 * per radiative line log10(J^0_0/B_nu(2.73 K)) and 100 J^2_0/J^0_0; and its SE
 * solution rho_1 for that J: per level log10(sqrt(2J_i+1) rho^0_0(i)), per level with
 * J > 0 1e3 rho^2_0/rho^0_0 (n_in = 2 + 2 r_num + e_num + nj, nj = levels with J > 0).
 * Outputs, the converged solution as corrections to rho_1: per level
 * log10(rho^0_0(i)/rho_1^0_0(i)), per level with J > 0 1e3 (rho^2_0/rho^0_0 -
 * rho_1^2_0/rho_1^0_0) (n_out = e_num + nj).
 * feats/targets/out are batch x n_in / batch x n_out row-major floats; norm holds the
 * input means, input standard deviations, output means, output standard deviations.
 * portal_lambda_iterate with surrogate = 1 uses it for its first iterate. */
int portal_mlp_create(portal_ctx *ctx, int n_in, int width, int depth, int n_out, unsigned long long seed);
int portal_mlp_destroy(portal_ctx *ctx);
long long portal_mlp_num_params(const portal_ctx *ctx);
int portal_mlp_set_params(portal_ctx *ctx, const float *params, const float *norm);
int portal_mlp_get_params(portal_ctx *ctx, float *params, float *norm);
int portal_mlp_forward(portal_ctx *ctx, int batch, const float *feats, float *out);
int portal_mlp_train_step(portal_ctx *ctx, int batch, const float *feats, const float *targets, float lr,
                          float *loss, float *grads);
int portal_mlp_train(portal_ctx *ctx, int n, const float *feats, const float *targets, int epochs, int batch,
                     float lr, unsigned long long seed, float *loss_hist, double *ms);
/* Network inputs of every solved node from the last SE pass (portal_se_pass or a
 * lambda iteration); nodes (1-based, may be NULL) gives the node of each row. */
int portal_surrogate_features(portal_ctx *ctx, int n_in, float *feats, int *nodes, int *n_work);

#ifdef __cplusplus
}
#endif
#endif
