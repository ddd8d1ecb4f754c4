/*
 * portal_oracle.c -- CPU FP64 restatement of PORTAL's hot path (test infrastructure only).
 *
 * This is the checker the HIP library is compared against.  It is never linked
 * into the product.  It restates the reference's algorithm routine by routine,
 * keeping the reference's evaluation order, its single-precision literals
 * (promoted to double exactly as a Fortran compiler does) and its quirks, so
 * that its results agree with the reference to rounding level.  Each function
 * names the reference routine (file:line under src/) it follows.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (see oracle/Makefile).  Contraction
 * into FMA is disabled so that every product is rounded as in the reference.
 *
 * Parity status per function (how it is pinned, see tests/):
 *   wig3jj / wig6jj            exact values (sympy) on the whole config domain
 *   legendre nodes / find_dir  compiled reference (_ref) + closed forms
 *   check_neigh_fast / trace_out / ray_in / inverse_rt / black_body
 *                              compiled reference (_ref), bit-level and <=1e-12
 *   ray_in with thin_exact     the closed form of a uniform ray, S + (I_in - S) e^-tau,
 *                              at tau down to 1e-22 with S / I_in up to 1e25; equal to
 *                              the reference form (<=1e-12) where no cell is thin
 *   polarize_prop, thin_exact  the reference's own formula evaluated in 60-digit
 *                              arithmetic (mpmath) at t from 1e-22 to 3, q/t to 0.9
 *   make_ac                    compiled reference (_ref)
 *   t_xxx_const, r_xxx_const   exact rational re-evaluation (sympy) + selection rules
 *   setup_stateq / get_rhovec  an independent rate-equation model in the magnetic-
 *                              sublevel basis (every rank K <= 6, Delta J = 0/+-1 lines,
 *                              stimulated emission, collisions) to 1e-10; the two-level
 *                              J = 0 -> 1 weak-radiation limit rho20/rho00 = J20/J00;
 *                              normalisation, detailed balance => Boltzmann, collision
 *                              limit => LTE
 *   rotate_all_save / trace_out_raytrace / polarize_prop / rot_stokes /
 *   ray_in_raytrace            compiled reference (_ref) + expm closed form
 *   comp_emcoefs               the Zeeman-component emission/absorption sums of the same
 *                              sublevel model to 1e-12; w^(2) against the closed forms of
 *                              Landi Degl'Innocenti & Landolfi (2004) Table 10.1; pure pi /
 *                              sigma light gives Q/I = +1 / -1; Kirchhoff's law (CODATA)
 *   depol_table / depol_at / setup_stateq_d (collisional depolarisation, an option the
 *                              reference never runs: D0 = 0 at main_sub.f90:96)
 *                              the sublevel model with a process relaxing exactly rank K
 *                              at rate D^(K), every rank of every level, to 1e-10;
 *                              D^(2)(J=1) = (9/5) k_2 by hand; J = 2, 3 with exact 3j/6j;
 *                              zeros for J = 0, odd K, K > 2J; make_AC's bracket in T
 *
 * The same file also runs under AddressSanitizer + UBSan (oracle/sanitize_main.c,
 * make -C oracle sanitize), which the CPU tests require to be clean.
 */
#include "portal_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define F32(x) ((double)(float)(x))

/* ------------------------------------------------------------------------- */
/* Wigner 3j / 6j (Racah's single-sum formulas, long double).  WIGXJPF, the    */
/* library the reference calls (pop_eqs.f90:48-61 etc.), returns the exact     */
/* value rounded to double; these sums have at most ~2K+1 terms for the        */
/* arguments PORTAL uses, so long double keeps them at double accuracy.        */
/* ------------------------------------------------------------------------- */
static long double lfact(int n) {
    static long double tab[512];
    static int init = 0;
    if (!init) {
        tab[0] = 1.0L;
        for (int i = 1; i < 512; i++) tab[i] = tab[i - 1] * (long double)i;
        init = 1;
    }
    return (n < 0 || n >= 512) ? NAN : tab[n];
}

static int tri_ok(int a, int b, int c) { /* twice-j arguments */
    if ((a + b + c) & 1) return 0;
    if (c > a + b || c < abs(a - b)) return 0;
    return 1;
}

static long double tri_delta(int a, int b, int c) { /* Delta(a b c), twice-j args */
    return lfact((a + b - c) / 2) * lfact((a - b + c) / 2) * lfact((-a + b + c) / 2) /
           lfact((a + b + c) / 2 + 1);
}

double orc_wig3jj(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3) {
    if (tm1 + tm2 + tm3 != 0) return 0.0;
    if (!tri_ok(tj1, tj2, tj3)) return 0.0;
    if (abs(tm1) > tj1 || abs(tm2) > tj2 || abs(tm3) > tj3) return 0.0;
    if (((tj1 + tm1) & 1) || ((tj2 + tm2) & 1) || ((tj3 + tm3) & 1)) return 0.0;
    int a = (tj1 + tj2 - tj3) / 2, b = (tj1 - tm1) / 2, c = (tj2 + tm2) / 2;
    int d = (tj3 - tj2 + tm1) / 2, e = (tj3 - tj1 - tm2) / 2;
    int kmin = 0;
    if (-d > kmin) kmin = -d;
    if (-e > kmin) kmin = -e;
    int kmax = a;
    if (b < kmax) kmax = b;
    if (c < kmax) kmax = c;
    long double s = 0.0L;
    for (int k = kmin; k <= kmax; k++) {
        long double t = 1.0L / (lfact(k) * lfact(d + k) * lfact(e + k) * lfact(a - k) *
                                lfact(b - k) * lfact(c - k));
        s += (k & 1) ? -t : t;
    }
    long double pre = sqrtl(tri_delta(tj1, tj2, tj3) * lfact((tj1 + tm1) / 2) * lfact((tj1 - tm1) / 2) *
                            lfact((tj2 + tm2) / 2) * lfact((tj2 - tm2) / 2) *
                            lfact((tj3 + tm3) / 2) * lfact((tj3 - tm3) / 2));
    int ph = (tj1 - tj2 - tm3) / 2;
    long double r = pre * s;
    if (ph & 1) r = -r;
    return (double)r;
}

double orc_wig6jj(int a, int b, int c, int d, int e, int f) {
    if (!tri_ok(a, b, c) || !tri_ok(a, e, f) || !tri_ok(d, b, f) || !tri_ok(d, e, c)) return 0.0;
    int a1 = (a + b + c) / 2, a2 = (a + e + f) / 2, a3 = (d + b + f) / 2, a4 = (d + e + c) / 2;
    int b1 = (a + b + d + e) / 2, b2 = (b + c + e + f) / 2, b3 = (c + a + f + d) / 2;
    int tmin = a1;
    if (a2 > tmin) tmin = a2;
    if (a3 > tmin) tmin = a3;
    if (a4 > tmin) tmin = a4;
    int tmax = b1;
    if (b2 < tmax) tmax = b2;
    if (b3 < tmax) tmax = b3;
    long double s = 0.0L;
    for (int t = tmin; t <= tmax; t++) {
        long double v = lfact(t + 1) / (lfact(t - a1) * lfact(t - a2) * lfact(t - a3) * lfact(t - a4) *
                                        lfact(b1 - t) * lfact(b2 - t) * lfact(b3 - t));
        s += (t & 1) ? -v : v;
    }
    long double pre = sqrtl(tri_delta(a, b, c) * tri_delta(a, e, f) * tri_delta(d, b, f) * tri_delta(d, e, c));
    return (double)(pre * s);
}

/* ------------------------------------------------------------------------- */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------- */

/* Gauss-Legendre nodes in ascending order with weights (cdgqf kind=1,
 * legendre_rule.f90:1-83); Newton iteration on P_n in long double. */
int orc_legendre_nodes(int n, double *x, double *w) {
    for (int i = 0; i < n; i++) {
        long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double pp = 0;
        for (int it = 0; it < 100; it++) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= n; j++) {
                long double p3 = p2;
                p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0L);
            long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)(2.0L / ((1.0L - z * z) * pp * pp));
    }
    return n;
}

/* rotmat, miscellaneous.f90:146-181 (Rot column-major 3x3) */
static void rotmat(const double *nv, double th, double *R) {
    double cth = cos(th), octh = 1.0 - cth, sth = sin(th);
#define RM(i, j) R[(i - 1) + 3 * (j - 1)]
    for (int i = 1; i <= 3; i++) RM(i, i) = cth + octh * nv[i - 1] * nv[i - 1];
    RM(1, 2) = nv[0] * nv[1] * octh - nv[2] * sth;
    RM(2, 1) = nv[1] * nv[0] * octh + nv[2] * sth;
    RM(1, 3) = nv[0] * nv[2] * octh + nv[1] * sth;
    RM(3, 1) = nv[2] * nv[0] * octh - nv[1] * sth;
    RM(2, 3) = nv[1] * nv[2] * octh - nv[0] * sth;
    RM(3, 2) = nv[2] * nv[1] * octh + nv[0] * sth;
#undef RM
}

/* find_dir, miscellaneous.f90:48-82 (+ two_mat_vec_prod 183-204) */
void orc_find_dir(const double *bvec, const double *gvec, double th, double ph, double *dir) {
    double xv[3], Rth[9], Rph[9];
    xv[0] = bvec[1] * gvec[2] - bvec[2] * gvec[1];
    xv[1] = bvec[2] * gvec[0] - bvec[0] * gvec[2];
    xv[2] = bvec[0] * gvec[1] - bvec[1] * gvec[0];
    double xn = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    for (int i = 0; i < 3; i++) xv[i] = xv[i] / xn;
    rotmat(xv, th, Rth);
    rotmat(bvec, ph, Rph);
    for (int i = 0; i < 3; i++) {
        double d = 0.0;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) d = d + Rph[i + 3 * j] * Rth[j + 3 * k] * bvec[k];
        dir[i] = d;
    }
}

static double norm2v(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

/* pos_quadratic, miscellaneous.f90:1078-1098 */
void orc_pos_quadratic(double a, double b, double c, double *x1, int *comp) {
    double dis = b * b - 4.0 * a * c;
    if (dis > 0) {
        *comp = 0;
        double r1 = (-b + sqrt(dis)) / (2.0 * a);
        double r2 = (-b - sqrt(dis)) / (2.0 * a);
        if (r2 > 0) r1 = r2;
        *x1 = r1;
    } else {
        *comp = 1;
    }
}

/* check_neigh_fast, miscellaneous.f90:1017-1076, with its star branch (1056-1069):
 * a star sphere met before the next cell boundary ends the ray there */
static void check_neigh_fast_s(double *Xp, const double *dir, const double *Xn, int neigh_num, const int *neigh,
                               const double *X_all, double *delta_x, int *next, int star_mode, const double *Xs,
                               double R_star);

void orc_check_neigh_fast(double *Xp, const double *dir, const double *Xn, int neigh_num, const int *neigh,
                          const double *X_all, double *delta_x, int *next) {
    check_neigh_fast_s(Xp, dir, Xn, neigh_num, neigh, X_all, delta_x, next, 0, NULL, 0.0);
}

static void check_neigh_fast_s(double *Xp, const double *dir, const double *Xn, int neigh_num, const int *neigh,
                               const double *X_all, double *delta_x, int *next, int star_mode, const double *Xs,
                               double R_star) {
    int comp;
    double nd = norm2v(dir[0], dir[1], dir[2]);
    double np_ = norm2v(Xp[0], Xp[1], Xp[2]);
    orc_pos_quadratic(nd * nd, 2.0 * (Xp[0] * dir[0] + Xp[1] * dir[1] + Xp[2] * dir[2]), np_ * np_ - 1.0,
                      delta_x, &comp);
    *next = 0;
    double d0 = Xn[0] - Xp[0], d1 = Xn[1] - Xp[1], d2 = Xn[2] - Xp[2];
    double t = norm2v(d0, d1, d2);
    double q4 = t * t;
    double q5 = d0 * dir[0] + d1 * dir[1] + d2 * dir[2];
    const double thr = F32(1E-10);
    for (int i = 0; i < neigh_num; i++) {
        const double *Xb = X_all + 3 * (size_t)(neigh[i] - 1);
        double e0 = Xb[0] - Xp[0], e1 = Xb[1] - Xp[1], e2 = Xb[2] - Xp[2];
        double u = norm2v(e0, e1, e2);
        double q1 = u * u;
        double q2 = e0 * dir[0] + e1 * dir[1] + e2 * dir[2];
        double new_x = 0.5 * (q1 - q4) / (q2 - q5);
        if (new_x > *delta_x || new_x <= thr) continue;
        *delta_x = new_x;
        *next = neigh[i];
    }
    if (star_mode == 1) {
        double s0 = Xp[0] - Xs[0], s1 = Xp[1] - Xs[1], s2 = Xp[2] - Xs[2];
        double sn = norm2v(s0, s1, s2);
        double new_x = 0.0;
        orc_pos_quadratic(nd * nd, 2.0 * (s0 * dir[0] + s1 * dir[1] + s2 * dir[2]), sn * sn - pow(R_star, 2.0),
                          &new_x, &comp);
        if (comp != 1 && new_x < *delta_x && new_x > 0.0) {
            *delta_x = new_x;
            *next = 0;
        }
    }
    for (int k = 0; k < 3; k++) Xp[k] = Xp[k] + *delta_x * dir[k];
}

/* trace_out, inverse_rt.f90:139-214.  cells[0..nc-1] are the cells crossed
 * (cells[0] = node), lens[] the cumulative nint(delta/dx) lengths.  Returns
 * node_count, or -1 when it exceeds n_depth (the reference stops there). */
int orc_trace_out(int node, const orc_grid *g, const double *dir, double dx, int n_depth, int *cells, int *lens) {
    return orc_trace_out_s(node, g, dir, dx, n_depth, cells, lens, 0, NULL, 0.0);
}

/* trace_out with the star test of check_neigh_fast on (star_mode = 1) */
int orc_trace_out_s(int node, const orc_grid *g, const double *dir, double dx, int n_depth, int *cells, int *lens,
                    int star_mode, const double *X_star, double R_star) {
    double Xp[3];
    const double *X = g->XYZ;
    for (int k = 0; k < 3; k++) Xp[k] = X[3 * (size_t)(node - 1) + k];
    int cur = node, nc = 0;
    double delta_x = 0.0;
    cells[0] = node;
    for (;;) {
        int nn = g->nbr_ptr[cur] - g->nbr_ptr[cur - 1];
        int nxt;
        check_neigh_fast_s(Xp, dir, X + 3 * (size_t)(cur - 1), nn, g->nbr_idx + g->nbr_ptr[cur - 1], X, &delta_x,
                           &nxt, star_mode, X_star, R_star);
        nc++;
        if (nc > n_depth) return -1;
        lens[nc - 1] = (int)lround(delta_x / dx);
        if (nc < n_depth) cells[nc] = nxt;
        cur = nxt;
        if (nxt == 0) break;
    }
    for (int i = 1; i < nc; i++) lens[i] = lens[i - 1] + lens[i];
    return nc;
}

/* find_angles, miscellaneous.f90:86-134: polar angle to B and azimuth about B (from
 * the x axis B x g) of a direction */
void orc_find_angles(const double *bvec, const double *gvec, const double *dir, double *th, double *ph) {
    double xv[3], cbd[3];
    xv[0] = bvec[1] * gvec[2] - bvec[2] * gvec[1];
    xv[1] = bvec[2] * gvec[0] - bvec[0] * gvec[2];
    xv[2] = bvec[0] * gvec[1] - bvec[1] * gvec[0];
    double xn = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    for (int i = 0; i < 3; i++) xv[i] = xv[i] / xn;
    double bd = bvec[0] * dir[0] + bvec[1] * dir[1] + bvec[2] * dir[2];
    if (bd > 1.0) *th = acos(1.0);
    else if (bd < -1.0) *th = acos(-1.0);
    else *th = acos(bd);
    cbd[0] = bvec[1] * dir[2] - bvec[2] * dir[1];
    cbd[1] = bvec[2] * dir[0] - bvec[0] * dir[2];
    cbd[2] = bvec[0] * dir[1] - bvec[1] * dir[0];
    *ph = atan2(xv[0] * dir[0] + xv[1] * dir[1] + xv[2] * dir[2], xv[0] * cbd[0] + xv[1] * cbd[1] + xv[2] * cbd[2]);
}

/* circle_points, miscellaneous.f90:1228-1265: n points (cos th, phi, w) on the star's
 * disc of angular radius th_s; x is 3 x n.  One point sits at (cos 1, cos 1) as the
 * reference writes it. */
void orc_circle_points(int n, double th_s, double dth_s, double *x) {
    const double pi = acos(-1.0);
    if (n == 1) {
        x[0] = x[1] = cos(1.0);
        x[2] = 2.0 * pi * dth_s;
    } else if (n < 10) {
        double wph = 2.0 * pi / n, ww = wph * dth_s, xx = cos(th_s / 2.0);
        for (int i = 1; i <= n; i++) {
            x[3 * (i - 1)] = xx;
            x[3 * (i - 1) + 1] = (i - 1) * wph;
            x[3 * (i - 1) + 2] = ww;
        }
    } else {
        double ga = pi * (3.0 - sqrt(5.0));  /* golden angle (Vogel) */
        double ww = dth_s * 2.0 * pi / n;
        for (int i = 1; i <= n; i++) {
            x[3 * (i - 1)] = cos(th_s * sqrt(1.0 * i / n));
            x[3 * (i - 1) + 1] = i * ga;
            x[3 * (i - 1) + 2] = ww;
        }
    }
}

/* star_all, miscellaneous.f90:1139-1226: the direction set of a node that sees a star
 * of solid angle dOm_star in direction kvec: n_star rays on the star's disc, then
 * nth rings (with ceil(nph sin th) azimuths each) over the rest of the sphere, all
 * re-expressed as (cos th, phi) about B.  xth is 3 x nth nph.  Returns 0, or -1 when
 * the rays do not fit (the reference stops there). */
int orc_star_all(int nth, int nph, const double *kvec, const double *gvec, const double *bvec, double dOm_star,
                 int *n_star, int *n_sphere, double *xth) {
    const double pi = acos(-1.0);
    double dth_star = dOm_star / (2.0 * pi);
    double th_star = acos(1 - dth_star);
    double dOm_max = 4.0 * pi / (nth * nph);
    int ns = (int)ceil(dOm_star / dOm_max);
    if (ns > nth * nph) return -1;
    orc_circle_points(ns, th_star, dth_star, xth);
    for (int i = 0; i < ns; i++) xth[3 * i + 2] = xth[3 * i + 2] / (4.0 * pi);
    int n = ns;
    double dth_rest = (pi - th_star) / nth;
    double th_i = th_star + dth_rest / 2.0;
    for (int i = 1; i <= nth; i++) {
        double th = (i - 1) * dth_rest + th_i, wth;
        if (i == 1) {
            double th_u = (i - 0.5) * dth_rest + th_i;
            wth = cos(th_star) - cos(th_u);
        } else if (i == nth) {
            double th_l = (i - 1.5) * dth_rest + th_i;
            wth = cos(th_l) + 1.0;
        } else {
            double th_u = (i - 0.5) * dth_rest + th_i;
            double th_l = (i - 1.5) * dth_rest + th_i;
            wth = cos(th_l) - cos(th_u);
        }
        int np_ = (int)ceil(nph * sin(th));
        double wph = 2.0 * pi / np_;
        double xx = cos(th);
        for (int j = 1; j <= np_; j++) {
            n++;
            if (n > nph * nth) return -1;
            xth[3 * (n - 1)] = xx;
            xth[3 * (n - 1) + 1] = (j - 1) * wph;
            xth[3 * (n - 1) + 2] = wth * wph / (4.0 * pi);
        }
    }
    for (int i = 0; i < n; i++) {
        double dir[3], th, ph;
        orc_find_dir(kvec, gvec, acos(xth[3 * i]), xth[3 * i + 1], dir);
        orc_find_angles(bvec, gvec, dir, &th, &ph);
        xth[3 * i] = cos(th);
        xth[3 * i + 1] = ph;
    }
    *n_star = ns;
    *n_sphere = n;
    return 0;
}

/* black_body, miscellaneous.f90:263-285 */
double orc_black_body(double T, double nu0) {
    if (T == 0.0) return 0.0;
    double c1 = 1.4744995e-50;
    double c2 = 4.79924e-11 / T;
    return c1 * pow(nu0, 3.0) * pow(exp(c2 * nu0) - 1.0, -1.0);
}

/* The emission factor 1 - e^-tau of ray_in.  The reference forms it as written
 * (inverse_rt.f90:314, :337), which in FP64 is exactly zero below tau ~ 1.1e-16 and
 * coarse above: a thin cell's emission S tau is lost.  It matters only where S tau
 * is comparable to what the ray already carries, i.e. S / I_in > 1e16 (lines far in
 * the Wien tail of the 2.73 K background).  orc_set_thin_exact(1) evaluates the
 * factor as -expm1(-tau), the definition the mixed-precision path follows. */
static int thin_exact = 0;
void orc_set_thin_exact(int on) { thin_exact = on != 0; }
static double emit_factor(double tau, double e_tau) { return thin_exact ? -expm1(-tau) : 1.0 - e_tau; }

/* ray_in, inverse_rt.f90:217-344.  Arrays are the per-cell gathers of the ray
 * path, index 0 = the central node's own cell. */
double orc_ray_in(double temp_node, double vel_node, double p1_node, double p2_node, double dust_node,
                  double bb_node, int node_count, const double *dens_1, const double *dens_2,
                  const double *dust_dens, double kd, const double *bb_list, const double *temp_list,
                  const double *vel_list, const int *length_list, int nvel, double dx, double nu0,
                  double Aji, double Bij, int g1, int g2, double I_in) {
    double c1 = 2.2102191e-42 * nu0;
    double c2 = 5.272859e-35 * nu0;
    double cnu = 299792458.0 / nu0;
    double I_tot = 0.0;
    double d_v = 2.0 * temp_node / nvel;
    double spi = sqrt(acos(-1.0));
    for (int nv = -nvel; nv <= nvel; nv++) {
        double v0 = (2.0 * temp_node) * nv / nvel;
        double a = v0 / temp_node;
        double phi_v0 = (1.0 / (spi * temp_node)) * exp(-(a * a));
        v0 = v0 + vel_node;
        double Iv = I_in;
        for (int i = node_count; i >= 2; i--) {
            double p1 = dens_1[i - 1], p2 = dens_2[i - 1];
            double kv = c1 * Bij * (p1 - (1.0 * g1) * p2 / (1.0 * g2));
            double ev = c2 * Aji * p2;
            double I_bb = orc_black_body(bb_list[i - 1], nu0);
            double alpha_v = kd * dust_dens[i - 1];
            double j_v = alpha_v * I_bb;
            double u = (v0 - vel_list[i - 1]) / temp_list[i - 1];
            double phi_v = cnu * (1.0 / (spi * temp_list[i - 1])) * exp(-(u * u));
            double tau_v = (kv * phi_v + alpha_v) * dx * (double)(length_list[i - 1] - length_list[i - 2]);
            if (tau_v > 1e-30) {
                double Sv = (ev * phi_v + j_v) / (kv * phi_v + alpha_v);
                double e_tau = exp(-tau_v);
                Iv = Iv * e_tau + Sv * emit_factor(tau_v, e_tau);
            }
        }
        double kv = c1 * Bij * (p1_node - (1.0 * g1) * p2_node / (1.0 * g2));
        double ev = c2 * Aji * p2_node;
        double I_bb = orc_black_body(bb_node, nu0);
        double alpha_v = kd * dust_node;
        double j_v = alpha_v * I_bb;
        double phi_v = phi_v0 * cnu;
        double Sv = (ev * phi_v + j_v) / (kv * phi_v + alpha_v);
        double tau_v = (kv * phi_v + alpha_v) * dx * (double)length_list[0];
        double e_tau = exp(-tau_v);
        double I_out = Iv * e_tau + Sv * emit_factor(tau_v, e_tau);
        I_tot = I_tot + phi_v0 * I_out * d_v;
    }
    return I_tot;
}

/* inverse_rt, inverse_rt.f90:1-137 (star_mode 0).  J0/J2 of every radiative
 * line at `node` (1-based).  Returns 0, or -1 on path overflow. */
static int inverse_rt_dirs(int node, const orc_grid *g, const orc_species *s, const orc_se_params *p,
                           const double *bvec, int n_sphere, const double *xth, int n_star, double *J0, double *J2);

int orc_inverse_rt(int node, const orc_grid *g, const orc_species *s, const orc_se_params *p,
                   const double *bvec, double *J0, double *J2) {
    if (p->star_mod != 1) return inverse_rt_dirs(node, g, s, p, bvec, p->n_sphere, p->xth, 0, J0, J2);
    /* trace_node's star branch, main_sub.f90:56-79 */
    const double pi = acos(-1.0);
    const double *Xn = g->XYZ + 3 * (size_t)(node - 1);
    double dv[3] = {p->X_star[0] - Xn[0], p->X_star[1] - Xn[1], p->X_star[2] - Xn[2]};
    double dn = norm2v(dv[0], dv[1], dv[2]);
    double dOm = pi * pow(p->R_star / dn, 2.0);
    for (int k = 0; k < 3; k++) dv[k] = dv[k] / dn;
    double *xs = (double *)calloc(3 * (size_t)p->nth * p->nph, sizeof(double));
    int n_star = 0, n_var = 0;
    int rc = orc_star_all(p->nth, p->nph, dv, p->gvec, bvec, dOm, &n_star, &n_var, xs);
    if (!rc) rc = inverse_rt_dirs(node, g, s, p, bvec, n_var, xs, n_star, J0, J2);
    free(xs);
    return rc ? -1 : 0;
}

static int inverse_rt_dirs(int node, const orc_grid *g, const orc_species *s, const orc_se_params *p,
                           const double *bvec, int n_sphere, const double *xth, int n_star, double *J0, double *J2) {
    const int R = s->r_num, L = s->e_num;
    int nd = p->n_depth;
    int *cells = (int *)malloc(sizeof(int) * (nd + 1));
    int *lens = (int *)malloc(sizeof(int) * (nd + 1));
    double *d1 = (double *)malloc(sizeof(double) * (nd + 1) * 7);
    double *d2 = d1 + (nd + 1), *dd = d2 + (nd + 1), *bb = dd + (nd + 1), *tl = bb + (nd + 1),
           *vr = tl + (nd + 1);
    int rc = 0;
    for (int k = 0; k < R; k++) J0[k] = J2[k] = 0.0;
    const double *V = g->vel;
    for (int i = 0; i < n_sphere; i++) {
        double mu = xth[3 * i], ph = xth[3 * i + 1], w = xth[3 * i + 2];
        double c_in = (3.0 * (mu * mu) - 1.0) / 2.0;
        double dir[3];
        const int star_ray = p->star_mod == 1 && i < n_star;
        orc_find_dir(bvec, p->gvec, acos(mu), ph, dir);
        int nc = orc_trace_out_s(node, g, dir, p->dx, nd, cells, lens, star_ray, p->X_star, p->R_star);
        if (nc < 0) { rc = -1; break; }
        const double *vn = V + 3 * (size_t)(node - 1);
        double vel_node = vn[0] * dir[0] + vn[1] * dir[1] + vn[2] * dir[2];
        for (int l = 0; l < nc; l++) {
            const double *vc = V + 3 * (size_t)(cells[l] - 1);
            vr[l] = vc[0] * dir[0] + vc[1] * dir[1] + vc[2] * dir[2];
            int c = cells[l] - 1;
            dd[l] = g->dust_p[c];
            bb[l] = g->temp[2 * (size_t)c + 1];
            tl[l] = g->b_all[c];
        }
        for (int k = 0; k < R; k++) {
            double nu0 = s->nu_hz[k];
            int l1 = s->r_trans[2 * k], l2 = s->r_trans[2 * k + 1];
            double I_in = orc_black_body(star_ray ? p->T_star : 2.73, nu0);
            for (int l = 0; l < nc; l++) {
                int c = cells[l] - 1;
                d1[l] = g->lev_p[(size_t)c * L + (l1 - 1)];
                d2[l] = g->lev_p[(size_t)c * L + (l2 - 1)];
            }
            int n0 = node - 1;
            double I_out = orc_ray_in(g->b_all[n0], vel_node, g->lev_p[(size_t)n0 * L + l1 - 1],
                                      g->lev_p[(size_t)n0 * L + l2 - 1], g->dust_p[n0], g->temp[2 * (size_t)n0 + 1],
                                      nc, d1, d2, dd, s->kd_v[k], bb, tl, vr, lens, p->nvel, g->tx * p->dx, nu0,
                                      s->t_Aji[k], s->t_Bij[k], s->lev_g[l1 - 1], s->lev_g[l2 - 1], I_in);
            J0[k] = J0[k] + I_out * w;
            J2[k] = J2[k] + I_out * w * c_in / sqrt(2.0);
        }
    }
    free(cells);
    free(lens);
    free(d1);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Statistical equilibrium                                                   */
/* ------------------------------------------------------------------------- */
#define W3(a, b, c, d, e, f) orc_wig3jj(2 * (a), 2 * (b), 2 * (c), 2 * (d), 2 * (e), 2 * (f))

/* t_abs_const, pop_eqs.f90:1-78 */
void orc_t_abs_const(int j, int k, int jl, int kl, double *t0, double *t2) {
    int dj = jl - j, dk = kl - k;
    *t0 = 0.0;
    *t2 = 0.0;
    if (abs(dj) <= 1) {
        double ljl = 2.0 * jl + 1.0;
        double fac1 = sqrt(3.0 * (2.0 * k + 1.0) * (2.0 * kl + 1.0));
        double fas = pow(-1.0, 1.0 - dj);
        double fac = ljl * fac1 * fas;
        if (dk == 0 || abs(dk) == 2) {
            for (int m = -j; m <= j; m++)
                for (int q = -1; q <= 1; q++) {
                    int ml = q + m;
                    double w1 = W3(j, j, k, m, -m, 0);
                    double w2 = W3(jl, jl, kl, ml, -ml, 0);
                    double w3 = W3(j, jl, 1, -m, ml, -q);
                    double w40 = W3(1, 1, 0, q, -q, 0);
                    double w42 = W3(1, 1, 2, q, -q, 0);
                    double ww = w1 * w2 * w3 * w3;
                    if (dk == 0) *t0 = *t0 + ww * w40;
                    *t2 = *t2 + ww * w42;
                }
            *t0 = *t0 * fac;
            *t2 = *t2 * sqrt(5.0) * fac;
        }
    }
}

/* t_stim_const, pop_eqs.f90:80-156 */
void orc_t_stim_const(int j, int k, int ju, int ku, double *t0, double *t2) {
    int dj = ju - j, dk = ku - k;
    *t0 = 0.0;
    *t2 = 0.0;
    if (abs(dj) <= 1) {
        double lju = 2.0 * ju + 1.0;
        double fac1 = sqrt(3.0 * (2.0 * k + 1.0) * (2.0 * ku + 1.0));
        double fas = pow(-1.0, 1.0 - dj);
        double fac = lju * fac1 * fas;
        if (dk == 0 || abs(dk) == 2) {
            for (int m = -j; m <= j; m++)
                for (int q = -1; q <= 1; q++) {
                    int mu = m - q;
                    double w1 = W3(j, j, k, m, -m, 0);
                    double w2 = W3(ju, ju, ku, mu, -mu, 0);
                    double w3 = W3(ju, j, 1, -mu, m, -q);
                    double w40 = W3(1, 1, 0, q, -q, 0);
                    double w42 = W3(1, 1, 2, q, -q, 0);
                    double ww = w1 * w2 * w3 * w3;
                    if (dk == 0) *t0 = *t0 + ww * w40;
                    *t2 = *t2 + ww * w42;
                }
            *t0 = *t0 * fac;
            *t2 = *t2 * sqrt(5.0) * fac;
        }
    }
}

/* t_spont_const, pop_eqs.f90:158-201 */
void orc_t_spont_const(int j, int k, int ju, int ku, double *t0) {
    int dj = ju - j, dk = ku - k;
    *t0 = 0.0;
    if (abs(dj) <= 1 && dk == 0) {
        double lju = 2.0 * ju + 1.0;
        double fas = pow(-1.0, 1.0 + 1.0 * j + 1.0 * ju + 1.0 * k);
        double w1 = orc_wig6jj(2 * ju, 2 * ju, 2 * k, 2 * j, 2 * j, 2);
        *t0 = lju * fas * w1;
    }
}

/* r_abs_const, pop_eqs.f90:204-281 */
void orc_r_abs_const(int j, int k, int ju, int kp, double *r0, double *r2) {
    int dj = ju - j, dk = kp - k;
    *r0 = 0.0;
    *r2 = 0.0;
    if (abs(dj) <= 1) {
        double lj = 2.0 * j + 1.0;
        double fac1 = sqrt(3.0 * (2.0 * k + 1.0) * (2.0 * kp + 1.0));
        if (dk == 0 || abs(dk) == 2) {
            for (int m = -j; m <= j; m++)
                for (int q = -1; q <= 1; q++) {
                    int mu = m - q;
                    double fas = pow(-1.0, 1.0 + q);
                    double fac = lj * fac1 * fas;
                    double w1 = W3(j, j, k, m, -m, 0);
                    double w2 = W3(j, j, kp, m, -m, 0);
                    double w3 = W3(ju, j, 1, -mu, m, -q);
                    double w40 = W3(1, 1, 0, q, -q, 0);
                    double w42 = W3(1, 1, 2, q, -q, 0);
                    double ww = w1 * w2 * w3 * w3;
                    if (dk == 0) *r0 = *r0 + fac * ww * w40;
                    *r2 = *r2 + fac * ww * w42;
                }
            *r2 = *r2 * sqrt(5.0);
        }
    }
}

/* r_stim_const, pop_eqs.f90:283-359 */
void orc_r_stim_const(int j, int k, int jl, int kp, double *r0, double *r2) {
    int dj = jl - j, dk = kp - k;
    *r0 = 0.0;
    *r2 = 0.0;
    if (abs(dj) <= 1) {
        double lj = 2.0 * j + 1.0;
        double fac1 = sqrt(3.0 * (2.0 * k + 1.0) * (2.0 * kp + 1.0));
        if (dk == 0 || abs(dk) == 2) {
            for (int m = -j; m <= j; m++)
                for (int q = -1; q <= 1; q++) {
                    int ml = m + q;
                    double fas = pow(-1.0, 1.0 - q);
                    double fac = lj * fac1 * fas;
                    double w1 = W3(j, j, k, m, -m, 0);
                    double w2 = W3(j, j, kp, m, -m, 0);
                    double w3 = W3(j, jl, 1, -m, ml, -q);
                    double w40 = W3(1, 1, 0, q, -q, 0);
                    double w42 = W3(1, 1, 2, q, -q, 0);
                    double ww = w1 * w2 * w3 * w3;
                    if (dk == 0) *r0 = *r0 + fac * ww * w40;
                    *r2 = *r2 + fac * ww * w42;
                }
            *r2 = *r2 * sqrt(5.0);
        }
    }
}

/* make_AC, miscellaneous.f90:288-392.  AC is 5 x n_trans column-major:
 * (A, C_down [s^-1], nu = E_u - E_l, J0, J2).  Above the top of the
 * collision-temperature table the reference leaves C undefined; here it is NaN. */
void orc_make_ac(const orc_species *s, double nH2, double Temp, const double *J0, const double *J2, double *AC) {
    int nC = s->c_num, nT = s->c_temp, nP = s->c_par;
    const double *C = s->t_Cji, *ct = s->col_temps;
#define CC(t, c, p) C[(t) + (size_t)nT * ((c) + (size_t)nC * (p))]
    double *C_r = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1));
    for (int j = 0; j < nC; j++) C_r[j] = NAN;
    if (Temp <= ct[0] && nP == 1)
        for (int j = 0; j < nC; j++) C_r[j] = CC(0, j, 0);
    if (Temp <= ct[0] && nP == 2)
        for (int j = 0; j < nC; j++) C_r[j] = 0.25 * CC(0, j, 0) + 0.75 * CC(0, j, 1);
    for (int i = 1; i < nT; i++) {
        if (ct[i] > Temp) {
            double ddT = (Temp - ct[i - 1]) / (ct[i] - ct[i - 1]);
            for (int j = 0; j < nC; j++) {
                if (nP == 1) {
                    double f1 = CC(i - 1, j, 0), f2 = CC(i, j, 0);
                    C_r[j] = f1 + ddT * (f2 - f1);
                } else if (nP == 2) {
                    double f1 = CC(i - 1, j, 0), f2 = CC(i, j, 0), f3 = CC(i - 1, j, 1), f4 = CC(i, j, 1);
                    double ff1 = f1 + ddT * (f2 - f1), ff2 = f3 + ddT * (f4 - f3);
                    C_r[j] = 0.25 * ff1 + 0.75 * ff2;
                }
            }
            break;
        }
    }
#undef CC
    for (int j = 0; j < nC; j++) C_r[j] = C_r[j] * nH2 * 1e-6;
    for (int i = 0; i < 5 * s->n_trans; i++) AC[i] = 0.0;
    for (int i = 0; i < s->n_trans; i++) {
        int lo = s->all_trans[2 * i], up = s->all_trans[2 * i + 1];
        AC[5 * i + 2] = s->e_levs_hz[up - 1] - s->e_levs_hz[lo - 1];
        for (int j = 0; j < s->r_num; j++)
            if (s->r_trans[2 * j] == lo && s->r_trans[2 * j + 1] == up) {
                AC[5 * i + 0] = s->t_Aji[j];
                AC[5 * i + 3] = J0[j];
                AC[5 * i + 4] = J2[j];
                break;
            }
        for (int j = 0; j < nC; j++)
            if (s->c_trans[2 * j] == lo && s->c_trans[2 * j + 1] == up) {
                AC[5 * i + 1] = C_r[j];
                break;
            }
    }
    free(C_r);
}

/* Collisional depolarisation (depol.f90:242-359).  The reference computes D^(K) but
 * trace_node passes D0 = 0 (main_sub.f90:96), so these are the option's restatement,
 * with the index policy below where the reference's code cannot be followed.
 *
 * orc_depol_table, de_polarization (depol.f90:242-317): Morris (1985) Eq. B11 in the
 * sudden approximation,
 *   D^(K)(J) = (2J+1)^2 sum_{m=1}^{2J} (-1)^m (2m+1) (J J m; 0 0 0)^2
 *              [{J J m; J J 0} - {J J m; J J K}] k_m(T),
 * for even K = 2..min(2J, kmax) at every tabulated collision temperature.  k_m is the
 * tabulated rate coefficient of the collisional transition between the J = 0 level and
 * the level whose J is m, with make_AC's partner mix (0.25 / 0.75 for two partners), and
 * 0 where the species has no such transition.  The reference reads C_j0 at level index
 * m, the level with J = m - 1 when j_lev(i) = i - 1 (and past its array for m >= e_num);
 * this takes the level with J = m.  D0_all is c_temp x e_num x (kmax - 1) column-major,
 * the reference's D0(nT, nlev, 2:kmax); odd K, K > 2J and levels with J = 0 hold 0.
 * Returns 0, or -1 without a J = 0 level. */
int orc_depol_table(const orc_species *s, double *D0_all) {
    const int nT = s->c_temp, L = s->e_num, nk = s->kmax - 1, nC = s->c_num;
    if (nk <= 0) return 0;
    for (size_t i = 0; i < (size_t)nT * L * nk; i++) D0_all[i] = 0.0;
    int l0 = -1;
    for (int i = 0; i < L; i++)
        if (s->j_lev[i] == 0) { l0 = i; break; }
    if (l0 < 0) return -1;
#define CC(t, c, p) s->t_Cji[(t) + (size_t)nT * ((c) + (size_t)nC * (p))]
    for (int i = 0; i < L; i++) {
        const int J = s->j_lev[i];
        const double lj2 = (2.0 * J + 1.0) * (2.0 * J + 1.0);
        for (int K = 2; K <= s->kmax && K <= 2 * J; K += 2)
            for (int m = 1; m <= 2 * J; m++) {
                const double w1 = orc_wig3jj(2 * J, 2 * J, 2 * m, 0, 0, 0);
                if (w1 == 0.0) continue;
                const double fac = ((m & 1) ? -1.0 : 1.0) * (2.0 * m + 1.0);
                const double fac2 = orc_wig6jj(2 * J, 2 * J, 2 * m, 2 * J, 2 * J, 0) -
                                    orc_wig6jj(2 * J, 2 * J, 2 * m, 2 * J, 2 * J, 2 * K);
                int lm = -1, col = -1;
                for (int q = 0; q < L; q++)
                    if (q != l0 && s->j_lev[q] == m) { lm = q; break; }
                if (lm >= 0)
                    for (int c = 0; c < nC; c++) {
                        const int a = s->c_trans[2 * c] - 1, b = s->c_trans[2 * c + 1] - 1;
                        if ((a == l0 && b == lm) || (a == lm && b == l0)) { col = c; break; }
                    }
                if (col < 0) continue;
                for (int t = 0; t < nT; t++) {
                    const double k = s->c_par == 2 ? 0.25 * CC(t, col, 0) + 0.75 * CC(t, col, 1) : CC(t, col, 0);
                    D0_all[t + (size_t)nT * (i + (size_t)L * (K - 2))] += lj2 * fac * w1 * w1 * fac2 * k;
                }
            }
    }
#undef CC
    return 0;
}

/* make_depolarization_temp, depol.f90:319-359: D0 (e_num x (kmax - 1)) at Temp, times
 * n_H2 [m^-3] * 1e-6.  The reference's bracket loop has no exit, so it takes the last
 * interval whose upper temperature exceeds Temp, and leaves D0 unset above the table;
 * this brackets as make_AC does: the first interval with col_temps(i) > Temp, so below
 * col_temps(1) the first interval's line extrapolated, and NaN above the table. */
void orc_depol_at(const orc_species *s, const double *D0_all, double Temp, double nH2, double *D0) {
    const int nT = s->c_temp, L = s->e_num, nk = s->kmax - 1;
    const double *ct = s->col_temps;
    for (int e = 0; e < L * nk; e++) {
        const double *f = D0_all + (size_t)nT * e;
        double v = NAN;
        if (Temp <= ct[0]) v = f[0];
        for (int i = 1; i < nT; i++)
            if (ct[i] > Temp) {
                const double ddT = (Temp - ct[i - 1]) / (ct[i] - ct[i - 1]);
                v = f[i - 1] + ddT * (f[i] - f[i - 1]);
                break;
            }
        D0[e] = v * nH2 * 1e-6;
    }
}

/* setup_stateq, pop_eqs.f90:361-653.  Matrices are c_dim x c_dim column-major.
 * D0 (e_num x (kmax - 1), the depolarisation rates of make_depolarization_temp, or NULL
 * for none, as trace_node passes) enters the diagonal of every K >= 2 row, as
 * pop_eqs.f90:642-646 does.  The reference's col_dep is not reset between entries, so a
 * nonzero D0 there would also enter every later entry's Amat sum (and is uninitialised
 * before the first K >= 2 diagonal); here it enters the diagonal only, the relaxation
 * term -D^(K) rho^K_0 of Landi Degl'Innocenti & Landolfi (2004) Eq. 7.101.  With D0 = 0
 * the two agree. */
void orc_setup_stateq_d(const orc_species *s, const double *AC, double Temp, const double *D0, double *Amat,
                        double *A_mat, double *B_mat, double *C_mat) {
    const int nlev = s->e_num, nt = s->n_trans, cd = s->c_dim, kmax = s->kmax;
    const double a_to_b = 6.781962e49;
    size_t nn = (size_t)cd * cd;
    memset(Amat, 0, nn * sizeof(double));
    memset(A_mat, 0, nn * sizeof(double));
    memset(B_mat, 0, nn * sizeof(double));
    memset(C_mat, 0, nn * sizeof(double));
#define M(X, r, c) X[((r) - 1) + (size_t)cd * ((c) - 1)]
    int *up = (int *)malloc(sizeof(int) * nt), *dn = (int *)malloc(sizeof(int) * nt);
    const double **acu = (const double **)malloc(sizeof(double *) * nt);
    const double **acd = (const double **)malloc(sizeof(double *) * nt);
    double col_dep = 0.0;
    for (int i1 = 1; i1 <= nlev; i1++) {
        int cu = 0, cdn = 0;
        for (int i2 = 1; i2 <= nlev; i2++)
            for (int q = 0; q < nt; q++) {
                if (s->all_trans[2 * q] == i1 && s->all_trans[2 * q + 1] == i2) { up[cu] = i2; acu[cu++] = AC + 5 * q; }
                if (s->all_trans[2 * q] == i2 && s->all_trans[2 * q + 1] == i1) { dn[cdn] = i2; acd[cdn++] = AC + 5 * q; }
            }
        int j1 = s->j_lev[i1 - 1];
        double g1 = 2.0 * j1 + 1.0;
        for (int t = 0; t < cu; t++) {
            int i2 = up[t], j2 = s->j_lev[i2 - 1];
            double g2 = 2.0 * j2 + 1.0;
            double A21 = acu[t][0], C21 = acu[t][1], nu0 = acu[t][2], JR0 = acu[t][3], JR2 = acu[t][4];
            double B21 = a_to_b * A21 / pow(nu0, 3.0);
            int c1 = s->count_array[i1 - 1];
            for (int k1 = 0; k1 <= (2 * j1 < kmax ? 2 * j1 : kmax); k1 += 2) {
                c1++;
                int c2 = s->count_array[i2 - 1];
                for (int k2 = 0; k2 <= (2 * j2 < kmax ? 2 * j2 : kmax); k2 += 2) {
                    c2++;
                    double ts0, ts2, te0;
                    orc_t_stim_const(j1, k1, j2, k2, &ts0, &ts2);
                    orc_t_spont_const(j1, k1, j2, k2, &te0);
                    double t_rad = B21 * (ts0 * JR0 + ts2 * JR2) + A21 * te0;
                    double t_col = 0.0;
                    if (k1 == k2 && k1 == 0) t_col = sqrt(g2 / g1) * C21;
                    M(Amat, c1, c2) = t_rad + t_col;
                    M(A_mat, c1, c2) = M(A_mat, c1, c2) + A21 * te0;
                    M(B_mat, c1, c2) = M(B_mat, c1, c2) + B21 * (ts0 * JR0 + ts2 * JR2);
                    M(C_mat, c1, c2) = M(C_mat, c1, c2) + t_col;
                }
            }
        }
        for (int t = 0; t < cdn; t++) {
            int i2 = dn[t], j2 = s->j_lev[i2 - 1];
            double g2 = 2.0 * j2 + 1.0;
            double A21 = acd[t][0], C21 = acd[t][1], nu0 = acd[t][2], JR0 = acd[t][3], JR2 = acd[t][4];
            double B21 = a_to_b * A21 / pow(nu0, 3.0);
            double B12 = B21 * g1 / g2;
            double C12 = C21 * g1 * exp(-4.79924e-11 * nu0 / Temp) / g2;
            int c1 = s->count_array[i1 - 1];
            for (int k1 = 0; k1 <= (2 * j1 < kmax ? 2 * j1 : kmax); k1 += 2) {
                c1++;
                int c2 = s->count_array[i2 - 1];
                for (int k2 = 0; k2 <= (2 * j2 < kmax ? 2 * j2 : kmax); k2 += 2) {
                    c2++;
                    double ta0, ta2;
                    orc_t_abs_const(j1, k1, j2, k2, &ta0, &ta2);
                    double t_rad = B12 * (ta0 * JR0 + ta2 * JR2);
                    double t_col = 0.0;
                    if (k1 == k2 && k1 == 0) t_col = sqrt(g2 / g1) * C12;
                    M(Amat, c1, c2) = t_rad + t_col;
                    M(B_mat, c1, c2) = M(B_mat, c1, c2) + t_rad;
                    M(C_mat, c1, c2) = M(C_mat, c1, c2) + t_col;
                }
            }
        }
        int c1 = s->count_array[i1 - 1];
        for (int k1 = 0; k1 <= (2 * j1 < kmax ? 2 * j1 : kmax); k1 += 2) {
            c1++;
            int c2 = s->count_array[i1 - 1];
            for (int k2 = 0; k2 <= (2 * j1 < kmax ? 2 * j1 : kmax); k2 += 2) {
                c2++;
                double rad_abs = 0.0, col_abs = 0.0;
                for (int t = 0; t < cu; t++) {
                    int i3 = up[t], j3 = s->j_lev[i3 - 1];
                    double g3 = 2.0 * j3 + 1.0;
                    double A21 = acu[t][0], C21 = acu[t][1], nu0 = acu[t][2], JR0 = acu[t][3], JR2 = acu[t][4];
                    double B21 = a_to_b * A21 / pow(nu0, 3.0);
                    double B12 = B21 * g3 / g1;
                    double C12 = C21 * g3 * exp(-4.79924e-11 * nu0 / Temp) / g1;
                    double r0, r2;
                    orc_r_abs_const(j1, k1, j3, k2, &r0, &r2);
                    rad_abs = rad_abs + B12 * (JR0 * r0 + JR2 * r2);
                    M(B_mat, c1, c2) = M(B_mat, c1, c2) - B12 * (JR0 * r0 + JR2 * r2);
                    if (k1 == k2) {
                        col_abs = col_abs + C12;
                        M(C_mat, c1, c2) = M(C_mat, c1, c2) - C12;
                    }
                }
                double rad_stim = 0.0, col_stim = 0.0;
                for (int t = 0; t < cdn; t++) {
                    int i3 = dn[t], j3 = s->j_lev[i3 - 1];
                    double A21 = acd[t][0], C21 = acd[t][1], nu0 = acd[t][2], JR0 = acd[t][3], JR2 = acd[t][4];
                    double B21 = a_to_b * A21 / pow(nu0, 3.0);
                    double r0, r2;
                    orc_r_stim_const(j1, k1, j3, k2, &r0, &r2);
                    rad_stim = rad_stim + B21 * (JR0 * r0 + JR2 * r2);
                    M(B_mat, c1, c2) = M(B_mat, c1, c2) - B21 * (JR0 * r0 + JR2 * r2);
                    if (k1 == k2) {
                        rad_stim = rad_stim + A21;
                        col_stim = col_stim + C21;
                        M(A_mat, c1, c2) = M(A_mat, c1, c2) - A21;
                        M(C_mat, c1, c2) = M(C_mat, c1, c2) - C21;
                    }
                }
                col_dep = 0.0;
                if (k1 == k2 && k2 >= 2) {
                    col_dep = D0 ? D0[(i1 - 1) + (size_t)nlev * (k1 - 2)] : 0.0;
                    M(C_mat, c1, c2) = M(C_mat, c1, c2) - col_dep;
                }
                M(Amat, c1, c2) = -1.0 * (rad_abs + rad_stim + col_abs + col_stim + col_dep);
            }
        }
    }
#undef M
    free(up);
    free(dn);
    free(acu);
    free(acd);
}

void orc_setup_stateq(const orc_species *s, const double *AC, double Temp, double *Amat, double *A_mat,
                      double *B_mat, double *C_mat) {
    orc_setup_stateq_d(s, AC, Temp, NULL, Amat, A_mat, B_mat, C_mat);
}

/* Dense solve A x = b by LU with partial pivoting (first maximal pivot, as
 * DGETRF) in long double, plus one long-double refinement step.  A is n x n
 * column-major (double).  Returns 0, or the 1-based column of a zero pivot. */
static int solve_ld(int n, const double *A, const double *b, double *x) {
    long double *LU = (long double *)malloc(sizeof(long double) * n * n);
    long double *y = (long double *)malloc(sizeof(long double) * n * 2);
    long double *r = y + n;
    int *piv = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n * n; i++) LU[i] = A[i];
#define L_(i, j) LU[(i) + (size_t)n * (j)]
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k;
        long double mx = fabsl(L_(k, k));
        for (int i = k + 1; i < n; i++)
            if (fabsl(L_(i, k)) > mx) { mx = fabsl(L_(i, k)); p = i; }
        piv[k] = p;
        if (mx == 0.0L) { info = k + 1; break; }
        if (p != k)
            for (int j = 0; j < n; j++) { long double t = L_(k, j); L_(k, j) = L_(p, j); L_(p, j) = t; }
        for (int i = k + 1; i < n; i++) {
            L_(i, k) /= L_(k, k);
            for (int j = k + 1; j < n; j++) L_(i, j) -= L_(i, k) * L_(k, j);
        }
    }
    if (!info) {
        for (int i = 0; i < n; i++) y[i] = b[i];
        for (int it = 0; it < 2; it++) {
            /* rhs in y (long double) -> solution in r */
            for (int k = 0; k < n; k++) { long double t = y[k]; y[k] = y[piv[k]]; y[piv[k]] = t; }
            for (int i = 0; i < n; i++)
                for (int j = 0; j < i; j++) y[i] -= L_(i, j) * y[j];
            for (int i = n - 1; i >= 0; i--) {
                for (int j = i + 1; j < n; j++) y[i] -= L_(i, j) * y[j];
                y[i] /= L_(i, i);
            }
            if (it == 0) {
                for (int i = 0; i < n; i++) r[i] = y[i];
                /* residual b - A r in long double */
                for (int i = 0; i < n; i++) {
                    long double s = b[i];
                    for (int j = 0; j < n; j++) s -= (long double)A[i + (size_t)n * j] * r[j];
                    y[i] = s;
                }
            } else {
                for (int i = 0; i < n; i++) r[i] += y[i];
            }
        }
        for (int i = 0; i < n; i++) x[i] = (double)r[i];
    }
#undef L_
    free(LU);
    free(y);
    free(piv);
    return info;
}

/* get_rhovec, em_coefs.f90:115-201: row 1 <- sqrt(2J+1) at each level's K=0
 * slot, rho = A^{-1} e1.  The reference takes column 1 of DGETRI's inverse;
 * here the system is solved directly in extended precision, which gives the
 * exact answer that the reference's column approximates.  Returns 0 or the
 * LU failure code (the reference stops the run there). */
int orc_get_rhovec(const orc_species *s, const double *AC, double Temp, double *rhovec, int extra_info,
                   double *rates) {
    return orc_get_rhovec_d(s, AC, Temp, NULL, rhovec, extra_info, rates);
}

/* get_rhovec with the depolarisation rates D0 (e_num x (kmax - 1), or NULL) that the
 * reference's signature carries (em_coefs.f90:115,130). */
int orc_get_rhovec_d(const orc_species *s, const double *AC, double Temp, const double *D0, double *rhovec,
                     int extra_info, double *rates) {
    int cd = s->c_dim;
    size_t nn = (size_t)cd * cd;
    double *mats = (double *)malloc(sizeof(double) * nn * 4);
    double *Am = mats, *Rm = mats + nn, *Bm = mats + 2 * nn, *Cm = mats + 3 * nn;
    orc_setup_stateq_d(s, AC, Temp, D0, Am, Rm, Bm, Cm);
    for (int j = 0; j < cd; j++) Am[(size_t)cd * j] = 0.0;
    for (int i = 0; i < s->e_num; i++) Am[(size_t)cd * s->count_array[i]] = sqrt(2.0 * s->j_lev[i] + 1.0);
    double *b = (double *)calloc(cd, sizeof(double));
    b[0] = 1.0;
    int info = solve_ld(cd, Am, b, rhovec);
    if (!info && extra_info) {
        for (int i = 0; i < 3 * cd; i++) rates[i] = 0.0;
        for (int i = 0; i < cd; i++)
            for (int j = 0; j < cd; j++) {
                rates[3 * i + 0] += fabs(Rm[i + (size_t)cd * j]) * rhovec[j];
                rates[3 * i + 1] += fabs(Bm[i + (size_t)cd * j]) * rhovec[j];
                rates[3 * i + 2] += fabs(Cm[i + (size_t)cd * j]) * rhovec[j];
            }
    }
    free(b);
    free(mats);
    return info;
}

static int pack_node(const orc_species *s, const double *rho, const double *rates, int extra_info,
                     double *node_pops, double *node_rates) {
    /* trace_node packing, main_sub.f90:112-137 */
    for (int i = 0; i < s->e_num; i++) {
        int ca = s->count_array[i];
        node_pops[2 * i] = rho[ca];
        node_pops[2 * i + 1] = s->j_lev[i] != 0 ? rho[ca + 1] : 0.0;
        if (extra_info && node_rates) {
            for (int r = 0; r < 3; r++) {
                node_rates[r + 3 * (2 * i)] = rates[r + 3 * ca];
                node_rates[r + 3 * (2 * i + 1)] = s->j_lev[i] != 0 ? rates[r + 3 * (ca + 1)] : 0.0;
            }
        }
        if (isnan(node_pops[2 * i])) {
            node_pops[2 * i] = 0.0;
            node_pops[2 * i + 1] = 0.0;
            if (extra_info && node_rates)
                for (int r = 0; r < 3; r++) node_rates[r + 3 * (2 * i + 1)] = 0.0; /* as main_sub.f90:134-135 */
        }
    }
    return 0;
}

/* trace_node, main_sub.f90:1-141 (star_mod 0). */
int orc_trace_node(int node, const orc_grid *g, const orc_species *s, const orc_se_params *p, double *node_pops,
                   double *node_rates, double *JJ_20, double *J0, double *J2) {
    const double *bvec = p->bvec_all + 3 * (size_t)(node - 1);
    if (orc_inverse_rt(node, g, s, p, bvec, J0, J2)) return -1;
    double *AC = (double *)malloc(sizeof(double) * 5 * s->n_trans);
    double *rho = (double *)malloc(sizeof(double) * s->c_dim * 4);
    double *rates = rho + s->c_dim;
    orc_make_ac(s, g->h2_dens[node - 1], g->temp[2 * (size_t)(node - 1)], J0, J2, AC);
    double *D0 = NULL;
    if (p->D0_all && s->kmax >= 2) {
        D0 = (double *)malloc(sizeof(double) * s->e_num * (s->kmax - 1));
        orc_depol_at(s, p->D0_all, g->temp[2 * (size_t)(node - 1)], g->h2_dens[node - 1], D0);
    }
    int info = orc_get_rhovec_d(s, AC, g->temp[2 * (size_t)(node - 1)], D0, rho, p->extra_info, rates);
    free(D0);
    if (!info) pack_node(s, rho, rates, p->extra_info, node_pops, node_rates);
    if (p->extra_info && JJ_20)
        for (int k = 0; k < s->r_num; k++) JJ_20[k] = J2[k] / J0[k];
    free(AC);
    free(rho);
    return info ? -2 : 0;
}

/* The node loop of main_para.f90:210-257 for nodes node_lo..node_hi (1-based,
 * inclusive).  pol_pops is N x 2L column-major, pol_rates N x 3 x 2L,
 * JJ_20 N x R, J0_all/J2_all R x N (any may be NULL). */
int orc_se_pass(const orc_grid *g, const orc_species *s, const orc_se_params *p, int node_lo, int node_hi,
                double *pol_pops, double *pol_rates, double *JJ_20, double *J0_all, double *J2_all, int nthreads) {
    const int N = g->node_num, L = s->e_num, R = s->r_num;
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4) reduction(min : err)
    for (int node = node_lo; node <= node_hi; node++) {
        if (!(g->ab_all[node - 1] > 0.0)) continue;
        double *np_ = (double *)malloc(sizeof(double) * (2 * L + 6 * L + 3 * R));
        double *nr = np_ + 2 * L, *jj = nr + 6 * L, *J0 = jj + R, *J2 = J0 + R;
        int rc = orc_trace_node(node, g, s, p, np_, nr, jj, J0, J2);
        if (rc) err = rc < err ? rc : err;
        else {
            for (int c = 0; c < 2 * L; c++) {
                if (pol_pops) pol_pops[(node - 1) + (size_t)N * c] = np_[c];
                if (pol_rates && p->extra_info)
                    for (int r = 0; r < 3; r++) pol_rates[(node - 1) + (size_t)N * (r + 3 * (size_t)c)] = nr[r + 3 * c];
            }
            if (JJ_20 && p->extra_info)
                for (int k = 0; k < R; k++) JJ_20[(node - 1) + (size_t)N * k] = jj[k];
            for (int k = 0; k < R; k++) {
                if (J0_all) J0_all[k + (size_t)R * (node - 1)] = J0[k];
                if (J2_all) J2_all[k + (size_t)R * (node - 1)] = J2[k];
            }
        }
        free(np_);
    }
    return err;
}

/* inverse_rt for a list of nodes (1-based); J0/J2 are R x n_nodes. */
int orc_inverse_nodes(const orc_grid *g, const orc_species *s, const orc_se_params *p, const int *nodes,
                      int n_nodes, double *J0, double *J2, int nthreads) {
    const int R = s->r_num;
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(min : err)
    for (int i = 0; i < n_nodes; i++) {
        const double *bvec = p->bvec_all + 3 * (size_t)(nodes[i] - 1);
        int rc = orc_inverse_rt(nodes[i], g, s, p, bvec, J0 + (size_t)R * i, J2 + (size_t)R * i);
        if (rc) err = rc < err ? rc : err;
    }
    return err;
}

/* SE only: make_AC + get_rhovec + packing from given J0/J2 (R x N). */
int orc_solve_from_J(const orc_grid *g, const orc_species *s, int node_lo, int node_hi, const double *J0_all,
                     const double *J2_all, double *pol_pops) {
    const int N = g->node_num, L = s->e_num, R = s->r_num;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(min : err)
    for (int node = node_lo; node <= node_hi; node++) {
        if (!(g->ab_all[node - 1] > 0.0)) continue;
        double *AC = (double *)malloc(sizeof(double) * 5 * s->n_trans);
        double *rho = (double *)malloc(sizeof(double) * (s->c_dim + 2 * L));
        double *np_ = rho + s->c_dim;
        double T = g->temp[2 * (size_t)(node - 1)];
        orc_make_ac(s, g->h2_dens[node - 1], T, J0_all + (size_t)R * (node - 1), J2_all + (size_t)R * (node - 1), AC);
        int info = orc_get_rhovec(s, AC, T, rho, 0, NULL);
        if (info) err = -2;
        else {
            pack_node(s, rho, NULL, 0, np_, NULL);
            for (int c = 0; c < 2 * L; c++) pol_pops[(node - 1) + (size_t)N * c] = np_[c];
        }
        free(AC);
        free(rho);
    }
    return err;
}

/* ------------------------------------------------------------------------- */
/* Image pass                                                                */
/* ------------------------------------------------------------------------- */

/* rotate_all_save, miscellaneous.f90:844-907 */
void orc_rotate_all_save(int N, double th_inc, double ph_az, const double *vel, const double *bvec, const double *XYZ,
                         double *vel_new, double *bvec_new, double *XYZ_new, const double *XYZ_star,
                         double *XYZ_star_new) {
    double xv[3] = {1.0, 0.0, 0.0}, zv[3] = {0.0, 0.0, 1.0}, Rth[9], Rph[9], R[9];
    rotmat(xv, th_inc, Rth);
    rotmat(zv, ph_az, Rph);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += Rph[i + 3 * k] * Rth[k + 3 * j];
            R[i + 3 * j] = s;
        }
    for (int i = 0; i < N; i++)
        for (int j = 0; j < 3; j++) {
            double v = 0.0, b = 0.0, x = 0.0;
            for (int k = 0; k < 3; k++) {
                v = v + R[j + 3 * k] * vel[3 * (size_t)i + k];
                b = b + R[j + 3 * k] * bvec[3 * (size_t)i + k];
                x = x + R[j + 3 * k] * XYZ[3 * (size_t)i + k];
            }
            vel_new[3 * (size_t)i + j] = v;
            bvec_new[3 * (size_t)i + j] = b;
            XYZ_new[3 * (size_t)i + j] = x;
        }
    /* the star's position, rotated the same way (miscellaneous.f90:900-905) */
    if (XYZ_star && XYZ_star_new)
        for (int j = 0; j < 3; j++) {
            double x = 0.0;
            for (int k = 0; k < 3; k++) x = x + R[j + 3 * k] * XYZ_star[k];
            XYZ_star_new[j] = x;
        }
}

/* trace_out_raytrace, raytrace.f90:58-177 with check_neigh, miscellaneous.f90:1-46.
 * cells/steps are the records of depth_list (cell id 1-based, cumulative step count).
 * With star_mod 1 a pixel within R_star of the (rotated) star's centre Xs in the sky
 * plane starts on the star's near side, z = Xs_z + R_star (raytrace.f90:103-114), and
 * *star_hit is set so that the caller takes B(T_star) as its incoming intensity.
 * Returns 0, -1 on n_depth_rt overflow, or -2 for a start point off the sphere. */
int orc_trace_out_raytrace(double x_co, double y_co, int N, const double *X, const int *nbr_ptr, const int *nbr_idx,
                           double dx, int n_depth_rt, int star_mod, double R_star, const double *Xs, int *cells,
                           int *steps, int *node_count, int *star_hit) {
    double z_co = -1.0 * sqrt(1.0 - x_co * x_co - y_co * y_co);
    int hit = 0;
    if (star_mod == 1) {
        const double xc = x_co - Xs[0], yc = y_co - Xs[1];
        const double rc0 = sqrt(xc * xc + yc * yc);
        if (rc0 < R_star) {
            z_co = Xs[2] + R_star;
            hit = 1;
        }
    }
    if (star_hit) *star_hit = hit;
    /* a start point off the sphere (NaN z) would march forever in the reference */
    if (!(z_co == z_co)) {
        *node_count = 0;
        return -2;
    }
    double st[3] = {x_co, y_co, z_co};
    double rsave = 10.0;
    int node = 0;
    for (int i = 0; i < N; i++) {
        double r = norm2v(st[0] - X[3 * (size_t)i], st[1] - X[3 * (size_t)i + 1], st[2] - X[3 * (size_t)i + 2]);
        if (r < rsave) { rsave = r; node = i + 1; }
    }
    double Xp[3] = {st[0], st[1], st[2]};
    const double dir[3] = {0.0, 0.0, 1.0};
    int nc = 0, k_cur = 0, in_sim = 0, rc = 0;
    while (!in_sim) {
        for (int k = 0; k < 3; k++) Xp[k] = Xp[k] + dx * dir[k];
        k_cur++;
        if (norm2v(Xp[0], Xp[1], Xp[2]) > 1.0) {
            nc++;
            in_sim = 1;
            if (nc <= n_depth_rt) { steps[nc - 1] = k_cur; cells[nc - 1] = node; } else rc = -1;
        }
        /* check_neigh */
        const double *Xn = X + 3 * (size_t)(node - 1);
        double a0 = Xp[0] - Xn[0], a1 = Xp[1] - Xn[1], a2 = Xp[2] - Xn[2];
        double dis0 = a0 * a0 + a1 * a1 + a2 * a2;
        int nn = node;
        for (int q = nbr_ptr[node - 1]; q < nbr_ptr[node]; q++) {
            const double *Xb = X + 3 * (size_t)(nbr_idx[q] - 1);
            double b0 = Xp[0] - Xb[0], b1 = Xp[1] - Xb[1], b2 = Xp[2] - Xb[2];
            double dis = b0 * b0 + b1 * b1 + b2 * b2;
            if (dis < dis0) { dis0 = dis; nn = nbr_idx[q]; }
        }
        if (nn != node) {
            nc++;
            if (nc <= n_depth_rt) { steps[nc - 1] = k_cur; cells[nc - 1] = nn; } else rc = -1;
            node = nn;
        }
    }
    *node_count = nc;
    return rc;
}

/* comp_emcoefs, em_coefs.f90:1-113 */
void orc_comp_emcoefs(double r1_0, double r2_0, double r1_2, double r2_2, int j1, int j2, double th, double nu0,
                      double Aji, double *nsI, double *nsQ, double *esI, double *esQ) {
    double fas = pow(-1.0, 1.0 + j1 + j2);
    double fac1 = sqrt(3.0 * (2.0 * j1 + 1.0));
    double fac2 = sqrt(3.0 * (2.0 * j2 + 1.0));
    double w6_1 = orc_wig6jj(2, 2, 4, 2 * j1, 2 * j1, 2 * j2);
    double w6_2 = orc_wig6jj(2, 2, 4, 2 * j2, 2 * j2, 2 * j1);
    double wjj2_1 = fas * fac1 * w6_1;
    double wjj2_2 = fas * fac2 * w6_2;
    double g1 = 2.0 * j1 + 1.0, g2 = 2.0 * j2 + 1.0;
    double ct = cos(th), sn = sin(th);
    double tki = (3.0 * (ct * ct) - 1.0) / (2.0 * sqrt(2.0));
    double tkq = -(3.0 * (sn * sn) / (2.0 * sqrt(2.0)));
    double na_I = 1.0 + wjj2_1 * (r1_2 / r1_0) * tki;
    double na_Q = wjj2_1 * (r1_2 / r1_0) * tkq;
    double ns_I = 1.0 + wjj2_2 * (r2_2 / r2_0) * tki;
    double ns_Q = wjj2_2 * (r2_2 / r2_0) * tkq;
    double a_to_b = 6.781962e49;
    double b_to_a = pow(nu0, 3.0) / a_to_b;
    double Bji = Aji / b_to_a;
    double es_I = ns_I, es_Q = ns_Q;
    double nsI_tot = sqrt(g1) * r1_0 * na_I - sqrt(g2) * r2_0 * ns_I * g1 / g2;
    double nsQ_tot = sqrt(g1) * r1_0 * na_Q - sqrt(g2) * r2_0 * ns_Q * g1 / g2;
    double hv_fac = 5.273e-35 * nu0;
    es_I = hv_fac * Aji * es_I * sqrt(g2) * r2_0;
    es_Q = hv_fac * Aji * es_Q * sqrt(g2) * r2_0;
    nsI_tot = hv_fac * (g2 / g1) * Bji * nsI_tot;
    nsQ_tot = hv_fac * (g2 / g1) * Bji * nsQ_tot;
    *nsI = nsI_tot;
    *nsQ = nsQ_tot;
    *esI = es_I;
    *esQ = es_Q;
}

/* (1 - e^-x)/x without cancellation, 1 at x = 0 */
static double phi_exact(double x) { return x == 0.0 ? 1.0 : -expm1(-x) / x; }

/* polarize_prop, raytrace.f90:1-56.  With orc_set_thin_exact(1) the same propagator
 * is evaluated in the basis I +- Q, where it is diagonal: o11 +- o12 =
 * (1 - e^-(t +- q))/(eta_I +- eta_Q) and the attenuation e^-t (ch q I -+ sh q Q)
 * splits into e^-(t +- q)(I +- Q); 1 - e^-x as -expm1(-x).  The reference's form
 * loses a thin cell's emission to the cancellation in 1 - (ch q + ...) e^-t. */
void orc_polarize_prop(double eta_I, double eta_Q, double eps_I, double eps_Q, double ds, double *I, double *Q,
                       double *U) {
    double I0 = *I, Q0 = *Q;
    if (thin_exact) {
        const double m = (eta_I + eta_Q) * ds, p = (eta_I - eta_Q) * ds;
        const double X = (I0 + Q0) * exp(-m) + (eps_I + eps_Q) * ds * phi_exact(m);
        const double Y = (I0 - Q0) * exp(-p) + (eps_I - eps_Q) * ds * phi_exact(p);
        *I = 0.5 * (X + Y);
        *Q = 0.5 * (X - Y);
        *U = *U * exp(-(eta_I * ds));
        return;
    }
    double shQ = sinh(eta_Q * ds);
    double chQ = cosh(eta_Q * ds);
    double e_tau = exp(-(eta_I * ds));
    double fac = pow(eta_I * eta_I - eta_Q * eta_Q, -1.0);
    double o11 = fac * eta_I * (1.0 - (chQ + (eta_Q / eta_I) * shQ) * e_tau);
    double o12 = -(fac * eta_Q * (1.0 - (chQ + (eta_I / eta_Q) * shQ) * e_tau));
    *I = o11 * eps_I + o12 * eps_Q + e_tau * (chQ * I0 - shQ * Q0);
    *Q = o12 * eps_I + o11 * eps_Q + e_tau * (-(shQ * I0) + chQ * Q0);
    *U = *U * e_tau;
}

static void stokes_angle(const double *b, double *c2a, double *s2a) {
    double n = norm2v(b[0], b[1], 0.0);
    double chi1 = b[0] / n, chi2 = b[1] / n;
    double alpha = acos(chi2);
    if (chi1 > 0.0) alpha = -alpha;
    *c2a = cos(2.0 * alpha);
    *s2a = sin(2.0 * alpha);
}

/* rot_stokes, miscellaneous.f90:463-506 */
void orc_rot_stokes(double Q, double U, double *Qn, double *Un, const double *b) {
    double c, s;
    stokes_angle(b, &c, &s);
    *Qn = Q * c + U * s;
    *Un = U * c - Q * s;
}

/* rot_stokes_back, miscellaneous.f90:508-551 */
void orc_rot_stokes_back(double Q, double U, double *Qn, double *Un, const double *b) {
    double c, s;
    stokes_angle(b, &c, &s);
    *Qn = Q * c - U * s;
    *Un = U * c + Q * s;
}

/* ray_in_raytrace, raytrace.f90:179-323.  out8 = (I, Q, U, tau, I_dust,
 * Q_dust, U_dust, tau_dust).  Per-cell arrays are indexed by record 0..nc-1. */
void orc_ray_in_raytrace(double vel_obs, int node_count, const double *pol, const double *b_list,
                         const double *bvec_list, const double *dust, double alpha, double kd, const double *bb_list,
                         const double *vel_list, const int *length_list, double dx, double nu0, double I_in,
                         double *out8) {
    double cnu = F32(2.99792E8) / nu0;
    double I_out = I_in, Q_in = 0.0, U_in = 0.0, Q_out = 0.0, U_out = 0.0, tau = 0.0, tau_dust_tot = 0.0;
    double I_dust = I_in, sig0 = 0.0, sig2 = 0.0, Qd = 0.0, Ud = 0.0;
    double spi = sqrt(acos(-1.0));
    for (int i = 0; i < node_count - 1; i++) {
        double b = b_list[i];
        double dv = vel_obs - vel_list[i];
        double v_corr = (1.0 / (b * spi)) * exp((-(dv * dv)) / (b * b));
        orc_rot_stokes(Q_in, U_in, &Q_out, &U_out, bvec_list + 3 * i);
        double mu = bvec_list[3 * i + 2];
        double mu2 = 1.0 - mu * mu;
        double eta_I = cnu * v_corr * pol[4 * i + 0];
        double eta_Q = cnu * v_corr * pol[4 * i + 1];
        double eps_I = cnu * v_corr * pol[4 * i + 2];
        double eps_Q = cnu * v_corr * pol[4 * i + 3];
        double delta_x = dx * (double)(length_list[i + 1] - length_list[i]);
        if (delta_x * eta_I > 1e-30) {
            orc_polarize_prop(eta_I, eta_Q, eps_I, eps_Q, delta_x, &I_out, &Q_out, &U_out);
            tau = tau + eta_I * delta_x;
        }
        double tau_dust = kd * dust[i] * delta_x;
        if (tau_dust > 1e-30) {
            double I_bb = orc_black_body(bb_list[i], nu0);
            double e_dust = exp(-tau_dust);
            double sI = 0.5 * alpha * (mu2 - 2.0 / 3.0);
            double sQ = alpha * mu2;
            I_dust = I_dust * e_dust + I_bb * (1.0 - e_dust);
            sig0 = sig0 * e_dust + I_bb * (1.0 - e_dust);
            sig2 = sig2 * e_dust + sI * I_bb * (1.0 - e_dust);
            const double *bv = bvec_list + 3 * i;
            double phi = acos(bv[1] / sqrt(bv[0] * bv[0] + bv[1] * bv[1]));
            if (bv[0] < 0.0) phi = -phi;
            double qr = cos(2.0 * phi), ur = sin(2.0 * phi);
            Qd = Qd * e_dust + sQ * I_bb * (1.0 - e_dust) * qr;
            Ud = Ud * e_dust + sQ * I_bb * (1.0 - e_dust) * ur;
            tau_dust_tot = tau_dust_tot + tau_dust;
        }
        orc_rot_stokes_back(Q_out, U_out, &Q_in, &U_in, bvec_list + 3 * i);
    }
    double I_corr = orc_black_body(2.73, nu0);
    I_out = I_out - I_corr;
    Q_out = Q_in;
    U_out = U_in;
    I_dust = I_dust - I_corr;
    double Q_dust, U_dust, tau_dust;
    if (I_dust == 0.0) {
        Q_dust = 0.0;
        U_dust = 0.0;
        tau_dust = 0.0;
    } else {
        Q_dust = I_dust * Qd / (sig0 - sig2);
        U_dust = I_dust * Ud / (sig0 - sig2);
        tau_dust = tau_dust_tot;
    }
    out8[0] = I_out; out8[1] = Q_out; out8[2] = U_out; out8[3] = tau;
    out8[4] = I_dust; out8[5] = Q_dust; out8[6] = U_dust; out8[7] = tau_dust;
}

/* One inclination of the pix_mod=1 image pass, para_finish.f90:382-449:
 * rotate, trace every pixel, emission coefficients per node and line, Stokes
 * transfer per pixel x line x channel.  pol_pops (N x 2L) is the rescaled
 * table of para_finish.f90:222-240.  Outputs are n_rt x nlevels x (2nvel+1)
 * column-major (all_I(:,:,inc,:) of the reference); dust_out, when given, holds
 * the four dust cubes back to back.  Pixels pix_lo..pix_hi (1-based) only. */
int orc_image_pass(const orc_grid *g, const orc_species *s, int nlevels, const int *l1, const int *l2,
                   const double *nu0_all, const double *Aji_all, const double *kd_all, double th_inc, double ph_az,
                   const double *bvec_all, const double *pol_pops, int nvel, const double *vel_out, int im_pix,
                   double im_size, double dx, double alpha, int n_depth_rt, int pix_lo, int pix_hi, int node_pixels,
                   int star_mod, double R_star, double T_star, const double *X_star, double *out_I, double *out_Q,
                   double *out_U, double *out_TAU, double *dust_out, int nthreads) {
    const int N = g->node_num, L = s->e_num;
    /* node_pixels: the reference's pix = 0 mode (para_finish.f90:245-330), one pixel per
     * bulk node at its rotated (x, y); else the (2 im_pix + 1)^2 grid (369-376) */
    const int n_side = 2 * im_pix + 1, n_rt = node_pixels ? g->bulk_node_num : n_side * n_side, C = 2 * nvel + 1;
    double *vi = (double *)malloc(sizeof(double) * 9 * (size_t)N);
    double *bi = vi + 3 * (size_t)N, *xi = bi + 3 * (size_t)N;
    const double xs0[3] = {0.0, 0.0, 0.0};
    double xs[3] = {0.0, 0.0, 0.0};
    orc_rotate_all_save(N, th_inc, ph_az, g->vel, bvec_all, g->XYZ, vi, bi, xi, X_star ? X_star : xs0, xs);
    double *pr = (double *)calloc((size_t)nlevels * 4 * N, sizeof(double)); /* all_pol_rayt(nlevels,4,N) */
    for (int node = 1; node <= g->bulk_node_num; node++) {
        if (pol_pops[(node - 1)] != 0.0) {
            for (int j = 0; j < nlevels; j++) {
                int a = l1[j], b = l2[j];
                double *o = pr + (size_t)j + (size_t)nlevels * 4 * (node - 1);
                orc_comp_emcoefs(pol_pops[(node - 1) + (size_t)N * (2 * (a - 1))],
                                 pol_pops[(node - 1) + (size_t)N * (2 * (b - 1))],
                                 pol_pops[(node - 1) + (size_t)N * (2 * (a - 1) + 1)],
                                 pol_pops[(node - 1) + (size_t)N * (2 * (b - 1) + 1)], s->j_lev[a - 1], s->j_lev[b - 1],
                                 acos(bi[3 * (size_t)(node - 1) + 2]), nu0_all[j], Aji_all[j], o, o + nlevels,
                                 o + 2 * nlevels, o + 3 * nlevels);
            }
        }
    }
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 8) reduction(min : err)
    for (int pix = pix_lo; pix <= pix_hi; pix++) {
        double x, y;
        if (node_pixels) {
            x = xi[3 * (size_t)(pix - 1)];
            y = xi[3 * (size_t)(pix - 1) + 1];
        } else {
            int ii = (pix - 1) / n_side - im_pix, jj = (pix - 1) % n_side - im_pix;
            x = ii * im_size / im_pix;
            y = jj * im_size / im_pix;
        }
        int *cells = (int *)malloc(sizeof(int) * 2 * (size_t)n_depth_rt);
        int *steps = cells + n_depth_rt;
        int nc = 0;
        int hit = 0;
        if (orc_trace_out_raytrace(x, y, N, xi, g->nbr_ptr, g->nbr_idx, dx, n_depth_rt, star_mod, R_star, xs, cells,
                                   steps, &nc, &hit)) {
            err = -1;
            free(cells);
            continue;
        }
        double *buf = (double *)malloc(sizeof(double) * (size_t)(nc > 0 ? nc : 1) * 11);
        double *pl = buf, *bl = pl + 4 * (size_t)nc, *bv = bl + nc, *du = bv + 3 * (size_t)nc, *bb = du + nc,
               *vl = bb + nc;
        for (int l = 0; l < nc; l++) {
            int c = cells[l] - 1;
            bl[l] = g->b_all[c];
            for (int k = 0; k < 3; k++) bv[3 * l + k] = bi[3 * (size_t)c + k];
            du[l] = g->dust_p[c];
            bb[l] = g->temp[2 * (size_t)c + 1];
            vl[l] = vi[3 * (size_t)c + 2];
        }
        for (int j = 0; j < nlevels; j++) {
            for (int l = 0; l < nc; l++)
                for (int k = 0; k < 4; k++) pl[4 * l + k] = pr[(size_t)j + (size_t)nlevels * (k + 4 * (size_t)(cells[l] - 1))];
            double I_in = orc_black_body(hit ? T_star : 2.73, nu0_all[j]);
            for (int k = 0; k < C; k++) {
                double o8[8];
                orc_ray_in_raytrace(vel_out[k], nc, pl, bl, bv, du, alpha, kd_all[j], bb, vl, steps, g->tx * dx,
                                    nu0_all[j], I_in, o8);
                size_t idx = (size_t)(pix - 1) + (size_t)n_rt * ((size_t)j + (size_t)nlevels * k);
                out_I[idx] = o8[0];
                out_Q[idx] = o8[1];
                out_U[idx] = o8[2];
                out_TAU[idx] = o8[3];
                if (dust_out) {
                    size_t cube = (size_t)n_rt * nlevels * C;
                    for (int q = 0; q < 4; q++) dust_out[q * cube + idx] = o8[4 + q];
                }
            }
        }
        free(buf);
        free(cells);
    }
    free(pr);
    free(vi);
    (void)L;
    return err;
}
