// Arguments of the image-pass kernels (image.hip).
#pragma once
#include "common.h"

namespace portal {

// per-node quantities of the rotated frame used by the Stokes transfer
struct CellGeo {
    double c2a, s2a;   // cos/sin(2 alpha) of rot_stokes
    double mu2;        // 1 - B_z^2
    double c2p, s2p;   // cos/sin(2 phi) of the dust branch
    double bz;         // rotated B_z
};

struct ImgArgs {
    int N, N_b, nlines, C, n_rt, n_side, im_pix, n_depth_rt, jmax;
    double im_size, dx, dl_scale, alpha;
    const double *Rtot;        // 3x3 column-major R_z(ph) R_x(th)
    const double *XYZ, *vel, *bvec;  // 3 x N
    const double *b_all, *dust, *tdust;
    const double *pol;         // N x 2L (rescaled populations)
    const int *jlev;
    const int *nbr_ptr, *nbr_idx;
    const int *l1, *l2;        // 0-based levels of each imaged line
    const double *nu0, *Aji, *kd, *Iin, *vel_out, *w6tab;
    // work
    double4 *pos_r;
    double *vz_r;
    CellGeo *geo;
    double4 *polco;            // N x nlines (eta_I, eta_Q, eps_I, eps_Q)
    double4 *start;
    int *start_node;
    int *rec_cell, *rec_step, *rec_n;
    int *flags;
    // outputs: n_rt x nlines x C
    double *I, *Q, *U, *TAU, *dust_out;
    int mixed;                 // 1: Stokes transfer in FP32 (stokes_f32_kernel)
    // star mode: pixels within R_star of the rotated star centre (Xs0, Xs1) in the sky
    // plane start at z = Xs2 + R_star with incoming intensity Istar[line]
    int star;
    double Xs0, Xs1, Xs2, R_star;
    const double *Istar;
    // 1: one pixel per bulk node at its rotated (x, y) (n_rt = N_b); 0: the pixel grid
    int node_pixels;
};

hipError_t launch_image(const ImgArgs &a, hipStream_t s, float *ms_geom, float *ms_rt);

}  // namespace portal
