/*
 * portal_oracle.h -- CPU FP64 restatement of PORTAL's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product (portal_amd/libportal_hip.so) never
 * links or calls it.
 *
 * Every function cites the reference routine it restates (file:line under the
 * reference's src/).  Arrays follow the reference's layouts: column-major,
 * node and level ids 1-based where the reference stores ids.  Adjacency is CSR
 * (row pointers + 1-based sorted neighbour ids); that is the same content as the
 * reference's del_ne(N,maxne)/del_num_ne pair.
 */
#ifndef PORTAL_ORACLE_H
#define PORTAL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int node_num, bulk_node_num, e_num;
    const double *XYZ;      /* 3 x N, normalised */
    const double *vel;      /* 3 x N */
    const double *temp;     /* 2 x N (T_kin, T_dust) */
    const double *b_all;    /* N */
    const double *h2_dens;  /* N (m^-3) */
    const double *ab_all;   /* N */
    const double *dust_p;   /* N */
    const double *lev_p;    /* L x N */
    const int *nbr_ptr;     /* N+1 */
    const int *nbr_idx;     /* 1-based */
    double tx;
} orc_grid;

typedef struct {
    int e_num, r_num, c_num, c_temp, c_par, n_trans, kmax, c_dim;
    const int *j_lev, *lev_g;
    const int *r_trans;     /* 2 x R (lower, upper), 1-based */
    const int *c_trans;     /* 2 x Cn */
    const int *all_trans;   /* 2 x n_trans */
    const int *count_array; /* L, 0-based slot offsets */
    const double *e_levs_hz, *t_Aji, *nu_hz, *t_Bij, *kd_v;
    const double *t_Cji;    /* c_temp x Cn x c_par */
    const double *col_temps;
} orc_species;

typedef struct {
    int n_sphere;
    const double *xth;      /* 3 x n_sphere (mu, phi, w) */
    int nvel;
    double dx;              /* normalised step quantum (1E-6 as float) */
    const double *gvec;     /* 3 */
    const double *bvec_all; /* 3 x N */
    int n_depth;
    int extra_info;
    /* star mode (star_mod = 1, main_sub.f90:56-79): a star of radius R_star and
     * temperature T_star at X_star (normalised coordinates); each node's directions are
     * star_all(nth, nph, ...) instead of xth */
    int star_mod, nth, nph;
    double R_star, T_star;
    const double *X_star;
    /* collisional depolarisation: the D0_all table of orc_depol_table (c_temp x e_num x
     * (kmax - 1)), or NULL for none (main_sub.f90:96 passes D0 = 0) */
    const double *D0_all;
} orc_se_params;

/* Wigner symbols, arguments are 2j / 2m like WIGXJPF's fwig3jj/fwig6jj. */
double orc_wig3jj(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3);
double orc_wig6jj(int tj1, int tj2, int tj3, int tj4, int tj5, int tj6);

/* quadrature / geometry */
int  orc_legendre_nodes(int n, double *x, double *w);
void orc_find_dir(const double *bvec, const double *gvec, double th, double ph, double *dir);
void orc_pos_quadratic(double a, double b, double c, double *x1, int *comp);
void orc_check_neigh_fast(double *X_phot, const double *dir, const double *X_node, int neigh_num,
                          const int *neigh, const double *X_all, double *delta_x, int *next);
void orc_find_angles(const double *bvec, const double *gvec, const double *dir, double *th, double *ph);
void orc_circle_points(int n, double th_s, double dth_s, double *x);
int  orc_star_all(int nth, int nph, const double *kvec, const double *gvec, const double *bvec, double dOm_star,
                  int *n_star, int *n_sphere, double *xth);
int  orc_trace_out_s(int node, const orc_grid *g, const double *dir, double dx, int n_depth, int *cells,
                     int *lens, int star_mode, const double *X_star, double R_star);
int  orc_trace_out(int node, const orc_grid *g, const double *dir, double dx, int n_depth,
                   int *cells, int *lens);
double orc_black_body(double T, double nu0);
double orc_ray_in(double temp_node, double vel_node, double p1_node, double p2_node, double dust_node,
                  double bb_node, int node_count, const double *dens_1, const double *dens_2,
                  const double *dust_dens, double kd, const double *bb_list, const double *temp_list,
                  const double *vel_list, const int *length_list, int nvel, double dx, double nu0,
                  double Aji, double Bij, int g1, int g2, double I_in);
int  orc_inverse_rt(int node, const orc_grid *g, const orc_species *s, const orc_se_params *p,
                    const double *bvec, double *J0, double *J2);
/* 1: ray_in's emission factor 1 - e^-tau as -expm1(-tau), and polarize_prop in the
 * basis I +- Q with -expm1 (no cancellation for thin cells); 0 (default): both as the
 * reference writes them.  Process-wide; set it between calls. */
void orc_set_thin_exact(int on);

/* statistical equilibrium */
void orc_t_abs_const(int j, int k, int jl, int kl, double *t0, double *t2);
void orc_t_stim_const(int j, int k, int ju, int ku, double *t0, double *t2);
void orc_t_spont_const(int j, int k, int ju, int ku, double *t0);
void orc_r_abs_const(int j, int k, int ju, int kp, double *r0, double *r2);
void orc_r_stim_const(int j, int k, int jl, int kp, double *r0, double *r2);
void orc_make_ac(const orc_species *s, double nH2, double Temp, const double *J0, const double *J2, double *AC);
void orc_setup_stateq(const orc_species *s, const double *AC, double Temp, double *Amat,
                      double *A_mat, double *B_mat, double *C_mat);
int  orc_get_rhovec(const orc_species *s, const double *AC, double Temp, double *rhovec,
                    int extra_info, double *rates);
void orc_setup_stateq_d(const orc_species *s, const double *AC, double Temp, const double *D0, double *Amat,
                        double *A_mat, double *B_mat, double *C_mat);
int  orc_get_rhovec_d(const orc_species *s, const double *AC, double Temp, const double *D0, double *rhovec,
                      int extra_info, double *rates);
int  orc_depol_table(const orc_species *s, double *D0_all);
void orc_depol_at(const orc_species *s, const double *D0_all, double Temp, double nH2, double *D0);
int  orc_trace_node(int node, const orc_grid *g, const orc_species *s, const orc_se_params *p,
                    double *node_pops, double *node_rates, double *JJ_20, double *J0, double *J2);
int  orc_se_pass(const orc_grid *g, const orc_species *s, const orc_se_params *p, int node_lo,
                 int node_hi, double *pol_pops, double *pol_rates, double *JJ_20,
                 double *J0_all, double *J2_all, int nthreads);
int  orc_inverse_nodes(const orc_grid *g, const orc_species *s, const orc_se_params *p, const int *nodes,
                       int n_nodes, double *J0, double *J2, int nthreads);
int  orc_solve_from_J(const orc_grid *g, const orc_species *s, int node_lo, int node_hi,
                      const double *J0_all, const double *J2_all, double *pol_pops);

/* image pass */
void orc_rotate_all_save(int N, double th_inc, double ph_az, const double *vel, const double *bvec,
                         const double *XYZ, double *vel_new, double *bvec_new, double *XYZ_new,
                         const double *XYZ_star, double *XYZ_star_new);
int  orc_trace_out_raytrace(double x_co, double y_co, int N, const double *X_inc, const int *nbr_ptr,
                            const int *nbr_idx, double dx, int n_depth_rt, int star_mod, double R_star,
                            const double *Xs, int *cells, int *steps, int *node_count, int *star_hit);
void orc_comp_emcoefs(double r1_0, double r2_0, double r1_2, double r2_2, int j1, int j2, double th,
                      double nu0, double Aji, double *nsI, double *nsQ, double *esI, double *esQ);
void orc_polarize_prop(double eta_I, double eta_Q, double eps_I, double eps_Q, double ds,
                       double *I, double *Q, double *U);
void orc_rot_stokes(double Q, double U, double *Qn, double *Un, const double *b);
void orc_rot_stokes_back(double Q, double U, double *Qn, double *Un, const double *b);
void orc_ray_in_raytrace(double vel_obs, int node_count, const double *pol, const double *b_list,
                         const double *bvec_list, const double *dust, double alpha, double kd,
                         const double *bb_list, const double *vel_list, const int *length_list,
                         double dx, double nu0, double I_in, double *out8);
int  orc_image_pass(const orc_grid *g, const orc_species *s, int nlevels, const int *l1, const int *l2,
                    const double *nu0_all, const double *Aji_all, const double *kd_all,
                    double th_inc, double ph_az, const double *bvec_all, const double *pol_pops,
                    int nvel, const double *vel_out, int im_pix, double im_size, double dx,
                    double alpha, int n_depth_rt, int pix_lo, int pix_hi, int node_pixels, int star_mod,
                    double R_star, double T_star, const double *X_star,
                    double *out_I, double *out_Q, double *out_U, double *out_TAU,
                    double *dust_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
