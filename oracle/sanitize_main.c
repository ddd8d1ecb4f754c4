/*
 * sanitize_main.c -- runs the oracle (portal_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer (test infrastructure only; built by `make -C oracle
 * sanitize` into _san/oracle_san with gcc, never linked into the product).
 *
 * Reads one model from a binary file written by tests/test_oracle_sanitized.py (a
 * sequence of tagged arrays, in the order read below), runs the SE pass over every
 * bulk node and one image pass, and writes the outputs in the same format.  The test
 * compares them with the regular build's and fails on any sanitizer report (the run
 * exits non-zero on the first one).
 *
 * File format: per array an int32 element count and an int32 type (0 int32, 1 float64),
 * then the elements.
 */
#include <stdio.h>
#include <stdlib.h>

#include "portal_oracle.h"

/* every buffer the driver allocates, released before exit so that the leak check sees
   only the oracle's own allocations */
static void *owned[128];
static int n_owned;

static void *keep(void *p) {
    if (!p || n_owned == (int)(sizeof owned / sizeof owned[0])) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return owned[n_owned++] = p;
}

static void *read_array(FILE *f, int want_type, int *count) {
    int hdr[2];
    if (fread(hdr, sizeof(int), 2, f) != 2 || hdr[1] != want_type || hdr[0] < 0) {
        fprintf(stderr, "bad input record\n");
        exit(2);
    }
    const size_t sz = want_type ? sizeof(double) : sizeof(int);
    void *p = keep(malloc(sz * (size_t)(hdr[0] > 0 ? hdr[0] : 1)));
    if (fread(p, sz, (size_t)hdr[0], f) != (size_t)hdr[0]) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
    *count = hdr[0];
    return p;
}

static int *ri(FILE *f, int *n) { return (int *)read_array(f, 0, n); }
static double *rd(FILE *f, int *n) { return (double *)read_array(f, 1, n); }

static void write_array(FILE *f, const double *p, int n) {
    const int hdr[2] = {n, 1};
    fwrite(hdr, sizeof(int), 2, f);
    fwrite(p, sizeof(double), (size_t)n, f);
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int n;
    int *dims = ri(f, &n); /* node_num bulk e_num r_num c_num c_temp c_par n_trans kmax c_dim n_sphere nvel
                              n_depth nlevels im_pix img_nvel n_depth_rt depolarisation */
    if (n != 18) return 2;
    const int N = dims[0], L = dims[2];
    orc_grid g = {dims[0], dims[1], dims[2]};
    g.XYZ = rd(f, &n);
    g.vel = rd(f, &n);
    g.temp = rd(f, &n);
    g.b_all = rd(f, &n);
    g.h2_dens = rd(f, &n);
    g.ab_all = rd(f, &n);
    g.dust_p = rd(f, &n);
    g.lev_p = rd(f, &n);
    g.nbr_ptr = ri(f, &n);
    g.nbr_idx = ri(f, &n);
    double *tx = rd(f, &n);
    g.tx = tx[0];
    orc_species s = {dims[2], dims[3], dims[4], dims[5], dims[6], dims[7], dims[8], dims[9]};
    s.j_lev = ri(f, &n);
    s.lev_g = ri(f, &n);
    s.r_trans = ri(f, &n);
    s.c_trans = ri(f, &n);
    s.all_trans = ri(f, &n);
    s.count_array = ri(f, &n);
    s.e_levs_hz = rd(f, &n);
    s.t_Aji = rd(f, &n);
    s.nu_hz = rd(f, &n);
    s.t_Bij = rd(f, &n);
    s.kd_v = rd(f, &n);
    s.t_Cji = rd(f, &n);
    s.col_temps = rd(f, &n);
    orc_se_params p = {0};
    p.n_sphere = dims[10];
    p.xth = rd(f, &n);
    p.nvel = dims[11];
    double *dxs = rd(f, &n); /* inward dx, image dx, th_inc, ph_az, im_size, alpha */
    p.dx = dxs[0];
    p.gvec = rd(f, &n);
    p.bvec_all = rd(f, &n);
    p.n_depth = dims[12];
    p.extra_info = 1;
    if (dims[17]) {
        /* the depolarisation option: the table of orc_depol_table in every node's SE solve */
        double *D0_all = keep(calloc((size_t)s.c_temp * s.e_num * (s.kmax > 1 ? s.kmax - 1 : 1), sizeof(double)));
        if (orc_depol_table(&s, D0_all)) return 3;
        p.D0_all = D0_all;
    }
    const int nl = dims[13];
    int *l1 = ri(f, &n), *l2 = ri(f, &n);
    double *nu0 = rd(f, &n), *Aji = rd(f, &n), *kd = rd(f, &n), *vel_out = rd(f, &n);
    fclose(f);

    const int R = s.r_num, C = 2 * dims[15] + 1, n_rt = (2 * dims[14] + 1) * (2 * dims[14] + 1);
    double *pp = keep(calloc((size_t)N * 2 * L, sizeof(double)));
    double *rates = keep(calloc((size_t)N * 6 * L, sizeof(double)));
    double *jj = keep(calloc((size_t)N * R, sizeof(double)));
    double *J0 = keep(calloc((size_t)N * R, sizeof(double))), *J2 = keep(calloc((size_t)N * R, sizeof(double)));
    int rc = orc_se_pass(&g, &s, &p, 1, g.bulk_node_num, pp, rates, jj, J0, J2, 1);
    if (rc) {
        fprintf(stderr, "se_pass failed (%d)\n", rc);
        return 3;
    }
    const size_t cube = (size_t)n_rt * nl * C;
    double *out = keep(calloc(cube * 4, sizeof(double)));
    rc = orc_image_pass(&g, &s, nl, l1, l2, nu0, Aji, kd, dxs[2], dxs[3], p.bvec_all, pp, dims[15], vel_out,
                        dims[14], dxs[4], dxs[1], dxs[5], dims[16], 1, n_rt, 0, 0, 0.0, 0.0, NULL, out, out + cube,
                        out + 2 * cube, out + 3 * cube, NULL, 1);
    if (rc) {
        fprintf(stderr, "image_pass failed (%d)\n", rc);
        return 3;
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    write_array(o, pp, N * 2 * L);
    write_array(o, J0, N * R);
    write_array(o, J2, N * R);
    write_array(o, out, (int)(cube * 4));
    fclose(o);
    while (n_owned) free(owned[--n_owned]);
    return 0;
}
