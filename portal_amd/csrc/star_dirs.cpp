// Star mode: the per-node direction sets of trace_node's star branch
// (main_sub.f90:56-79), built on the host once per geometry.
//
// A node that sees a star of radius R_star at X_star gets star_all(nth, nph, ...)
// (miscellaneous.f90:1139-1226): n_star rays on the star's disc (circle_points,
// 1228-1265), then nth rings over the rest of the sphere with ceil(nph sin th)
// azimuths each, every direction re-expressed as (cos th, phi) about the node's field
// (find_dir 48-82, find_angles 86-134).  The sets are computed in FP64 on the host with
// the same operation order as the reference (built with -ffp-contract=off), so that the
// walk that follows sees the reference's directions bit for bit.
#include <cmath>
#include <vector>

#include "common.h"

namespace portal {

namespace {

void rotmat_h(const double *nv, double th, double *R) {
    const double cth = std::cos(th), octh = 1.0 - cth, sth = std::sin(th);
    for (int i = 0; i < 3; i++) R[i + 3 * i] = cth + octh * nv[i] * nv[i];
    R[0 + 3 * 1] = nv[0] * nv[1] * octh - nv[2] * sth;
    R[1 + 3 * 0] = nv[1] * nv[0] * octh + nv[2] * sth;
    R[0 + 3 * 2] = nv[0] * nv[2] * octh + nv[1] * sth;
    R[2 + 3 * 0] = nv[2] * nv[0] * octh - nv[1] * sth;
    R[1 + 3 * 2] = nv[1] * nv[2] * octh - nv[0] * sth;
    R[2 + 3 * 1] = nv[2] * nv[1] * octh + nv[0] * sth;
}

void find_dir_h(const double *b, const double *g, double th, double ph, double *dir) {
    double xv[3], Rth[9], Rph[9];
    xv[0] = b[1] * g[2] - b[2] * g[1];
    xv[1] = b[2] * g[0] - b[0] * g[2];
    xv[2] = b[0] * g[1] - b[1] * g[0];
    const double xn = std::sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    for (int i = 0; i < 3; i++) xv[i] = xv[i] / xn;
    rotmat_h(xv, th, Rth);
    rotmat_h(b, ph, Rph);
    for (int i = 0; i < 3; i++) {
        double d = 0.0;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) d = d + Rph[i + 3 * j] * Rth[j + 3 * k] * b[k];
        dir[i] = d;
    }
}

void find_angles_h(const double *b, const double *g, const double *dir, double *th, double *ph) {
    double xv[3], cbd[3];
    xv[0] = b[1] * g[2] - b[2] * g[1];
    xv[1] = b[2] * g[0] - b[0] * g[2];
    xv[2] = b[0] * g[1] - b[1] * g[0];
    const double xn = std::sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    for (int i = 0; i < 3; i++) xv[i] = xv[i] / xn;
    const double bd = b[0] * dir[0] + b[1] * dir[1] + b[2] * dir[2];
    if (bd > 1.0) *th = std::acos(1.0);
    else if (bd < -1.0) *th = std::acos(-1.0);
    else *th = std::acos(bd);
    cbd[0] = b[1] * dir[2] - b[2] * dir[1];
    cbd[1] = b[2] * dir[0] - b[0] * dir[2];
    cbd[2] = b[0] * dir[1] - b[1] * dir[0];
    *ph = std::atan2(xv[0] * dir[0] + xv[1] * dir[1] + xv[2] * dir[2],
                     xv[0] * cbd[0] + xv[1] * cbd[1] + xv[2] * cbd[2]);
}

// points on the star's disc; a single point is written at (cos 1, cos 1) as in the reference
void circle_points_h(int n, double th_s, double dth_s, double *x) {
    const double pi = std::acos(-1.0);
    if (n == 1) {
        x[0] = x[1] = std::cos(1.0);
        x[2] = 2.0 * pi * dth_s;
    } else if (n < 10) {
        const double wph = 2.0 * pi / n, ww = wph * dth_s, xx = std::cos(th_s / 2.0);
        for (int i = 1; i <= n; i++) {
            x[3 * (i - 1)] = xx;
            x[3 * (i - 1) + 1] = (i - 1) * wph;
            x[3 * (i - 1) + 2] = ww;
        }
    } else {
        const double ga = pi * (3.0 - std::sqrt(5.0));  // golden angle (Vogel's spiral)
        const double ww = dth_s * 2.0 * pi / n;
        for (int i = 1; i <= n; i++) {
            x[3 * (i - 1)] = std::cos(th_s * std::sqrt(1.0 * i / n));
            x[3 * (i - 1) + 1] = i * ga;
            x[3 * (i - 1) + 2] = ww;
        }
    }
}

// star_all into xth (3 x nth nph); false when the rays do not fit (the reference stops)
bool star_all_h(int nth, int nph, const double *kvec, const double *g, const double *b, double dOm, int *n_star,
                int *n_sphere, double *xth) {
    const double pi = std::acos(-1.0);
    const double dth_star = dOm / (2.0 * pi);
    const double th_star = std::acos(1 - dth_star);
    const double dOm_max = 4.0 * pi / (nth * nph);
    const int ns = (int)std::ceil(dOm / dOm_max);
    if (!(ns >= 1 && ns <= nth * nph)) return false;
    circle_points_h(ns, th_star, dth_star, xth);
    for (int i = 0; i < ns; i++) xth[3 * i + 2] = xth[3 * i + 2] / (4.0 * pi);
    int n = ns;
    const double dth_rest = (pi - th_star) / nth;
    const double th_i = th_star + dth_rest / 2.0;
    for (int i = 1; i <= nth; i++) {
        const double th = (i - 1) * dth_rest + th_i;
        double wth;
        if (i == 1) {
            wth = std::cos(th_star) - std::cos((i - 0.5) * dth_rest + th_i);
        } else if (i == nth) {
            wth = std::cos((i - 1.5) * dth_rest + th_i) + 1.0;
        } else {
            wth = std::cos((i - 1.5) * dth_rest + th_i) - std::cos((i - 0.5) * dth_rest + th_i);
        }
        const int np_ = (int)std::ceil(nph * std::sin(th));
        const double wph = 2.0 * pi / np_;
        const double xx = std::cos(th);
        for (int j = 1; j <= np_; j++) {
            n++;
            if (n > nph * nth) return false;
            xth[3 * (n - 1)] = xx;
            xth[3 * (n - 1) + 1] = (j - 1) * wph;
            xth[3 * (n - 1) + 2] = wth * wph / (4.0 * pi);
        }
    }
    for (int i = 0; i < n; i++) {
        double dir[3], th, ph;
        find_dir_h(kvec, g, std::acos(xth[3 * i]), xth[3 * i + 1], dir);
        find_angles_h(b, g, dir, &th, &ph);
        xth[3 * i] = std::cos(th);
        xth[3 * i + 1] = ph;
    }
    *n_star = ns;
    *n_sphere = n;
    return true;
}

}  // namespace

// Per work node (in work-list order) the star-mode directions, padded to D = nth nph
// rays of weight 0: xthn (n_work x D x 3), nvar (rays in use) and nstar (rays towards
// the star).  Returns the index of the first node whose rays do not fit, or -1.
int build_star_dirs(const std::vector<int> &work, const double *XYZ, const double *bvec_all, const double *gvec,
                    int nth, int nph, const double *X_star, double R_star, std::vector<double> &xthn,
                    std::vector<int> &nvar, std::vector<int> &nstar) {
    const double pi = std::acos(-1.0);
    const int D = nth * nph, nw = (int)work.size();
    xthn.assign((size_t)nw * D * 3, 0.0);
    nvar.assign(nw, 0);
    nstar.assign(nw, 0);
    for (int w = 0; w < nw; w++) {
        const int node = work[w];
        const double *X = XYZ + 3 * (size_t)node;
        double dv[3] = {X_star[0] - X[0], X_star[1] - X[1], X_star[2] - X[2]};
        const double dn = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        const double dOm = pi * std::pow(R_star / dn, 2.0);
        for (int k = 0; k < 3; k++) dv[k] = dv[k] / dn;
        int ns = 0, nv = 0;
        if (!star_all_h(nth, nph, dv, gvec, bvec_all + 3 * (size_t)node, dOm, &ns, &nv, &xthn[(size_t)w * D * 3]))
            return w;
        nvar[w] = nv;
        nstar[w] = ns;
    }
    return -1;
}

}  // namespace portal
