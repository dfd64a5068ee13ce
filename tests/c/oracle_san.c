/* Driver for the AddressSanitizer / UndefinedBehaviorSanitizer build of the
 * CPU oracle (oracle/s2oracle.c): every public function at even and odd
 * bandwidths, all cutoffs kinds (bw, bw/4, 1, 0) and both data formats, on
 * buffers of exactly the documented sizes, so any read or write past them
 * (or any undefined arithmetic) stops the run.  Also checks that the
 * results are sane: the round trip of band-limited coefficients is the
 * identity (degrees below bw-1 at odd bw), within 1e-11.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "s2oracle.h"

static double* alloc(size_t n) {
    double* p = malloc(sizeof(double) * (n ? n : 1));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(3);
    }
    return p;
}

static int run(int bw, int cutoff, int fmt) {
    const int N = 2 * bw;
    const size_t nt = (size_t)so_reduced_naive_table_size(bw, cutoff) + so_reduced_sph_table_size(bw, cutoff);
    double *fwd = alloc(nt), *inv = alloc(nt), *w = alloc(4 * bw);
    long* off = malloc(sizeof(long) * bw);
    const double** tf = malloc(sizeof(double*) * bw);
    const double** ti = malloc(sizeof(double*) * bw);
    double *rc = alloc((size_t)bw * bw), *ic = alloc((size_t)bw * bw);
    double *rc2 = alloc((size_t)bw * bw), *ic2 = alloc((size_t)bw * bw);
    double *rd = alloc((size_t)N * N), *id = alloc((size_t)N * N);
    double *zr = alloc(bw), *zi = alloc(bw);
    so_tables(bw, cutoff, fwd, inv, off);
    for (int m = 0; m < bw; ++m) {
        tf[m] = fwd + off[m];
        ti[m] = inv + off[m];
    }
    so_weights(bw, w);
    so_make_coeffs(11u + (unsigned)bw, bw, fmt == 1 ? 2 : 1, rc, ic);
    if (bw % 2) /* the odd-bw inverse synthesises degree bw-1 only in part */
        for (int m = -(bw - 1); m < bw; ++m) {
            rc[so_index_of_coeff(m, bw - 1, bw)] = 0.0;
            ic[so_index_of_coeff(m, bw - 1, bw)] = 0.0;
        }
    so_inv_fst(rc, ic, rd, id, bw, ti, fmt, cutoff);
    so_fst(rd, id, rc2, ic2, bw, tf, w, fmt, cutoff);
    double err = 0.0, ref = 0.0;
    for (size_t k = 0; k < (size_t)bw * bw; ++k) {
        err = fmax(err, hypot(rc2[k] - rc[k], ic2[k] - ic[k]));
        ref = fmax(ref, hypot(rc[k], ic[k]));
    }
    if (cutoff > 0) so_fzt(rd, id, zr, zi, bw, tf[0], w, fmt);
    int bad = !(err <= 1e-11 * ref);
    printf("bw=%d cutoff=%d fmt=%d round trip %.2e%s\n", bw, cutoff, fmt, err / ref, bad ? "  FAILED" : "");
    free(fwd); free(inv); free(w); free(off); free(tf); free(ti);
    free(rc); free(ic); free(rc2); free(ic2); free(rd); free(id); free(zr); free(zi);
    return bad;
}

static int per_order(int bw) {
    const int N = 2 * bw;
    double *w = alloc(4 * bw), *data = alloc(N), *res = alloc(N), *sinv = alloc(N);
    so_weights(bw, w);
    so_uniform(5, N, data);
    for (int j = 0; j < N; ++j) sinv[j] = sin((2.0 * j + 1.0) * M_PI / (4.0 * bw));
    for (int m = 0; m < bw; m += (bw > 8 ? bw / 8 : 1)) {
        const int ts = so_table_size(m, bw);
        double *t = alloc(ts), *tt = alloc(ts), *pml = alloc((size_t)N * (bw - m)), *c = alloc(bw - m);
        so_cospml_table(bw, m, t);
        so_transpose_cospml_table(bw, m, t, tt);
        so_pml_table(bw, m, pml);
        so_dlt_semi(data, bw, m, c, t, w);
        so_inv_dlt_semi(c, bw, m, res, tt, sinv);
        so_dlt_naive(data, bw, m, w, c, pml);
        so_inv_dlt_naive(c, bw, m, res, pml);
        free(t); free(tt); free(pml); free(c);
    }
    double *pmm = alloc(N), *re = alloc(N), *im = alloc(N), *ore = alloc(N), *oim = alloc(N), *o2 = alloc(N);
    so_pmm_l2(bw / 2, sinv, N, pmm);
    so_uniform(6, N, re);
    so_uniform(7, N, im);
    so_dft(N, re, im, ore, oim, -1);
    so_redft10(N, re, o2);
    so_redft01(N, re, o2);
    free(w); free(data); free(res); free(sinv); free(pmm); free(re); free(im); free(ore); free(oim); free(o2);
    return 0;
}

static int conv(int bw) {
    const size_t n = 4 * (size_t)bw * bw;
    double *s = alloc(n), *z = alloc(n), *h = alloc(n), *or_ = alloc(n), *oi = alloc(n);
    so_uniform(8, n, s);
    so_uniform(9, n, h);
    for (size_t k = 0; k < n; ++k) z[k] = 0.0;
    so_conv(s, z, h, z, or_, oi, bw);
    double *a = alloc((size_t)bw * bw), *b = alloc((size_t)bw * bw), *f = alloc(bw), *g = alloc(bw);
    double *rr = alloc((size_t)bw * bw), *ri = alloc((size_t)bw * bw);
    so_uniform(10, (size_t)bw * bw, a);
    so_uniform(11, (size_t)bw * bw, b);
    so_uniform(12, bw, f);
    so_uniform(13, bw, g);
    so_trans_mult(a, b, f, g, rr, ri, bw);
    free(s); free(z); free(h); free(or_); free(oi); free(a); free(b); free(f); free(g); free(rr); free(ri);
    return 0;
}

int main(void) {
    int bad = 0;
    const int bws[] = {2, 3, 8, 9, 16, 17, 23, 32};
    for (size_t i = 0; i < sizeof(bws) / sizeof(bws[0]); ++i) {
        const int bw = bws[i];
        const int cuts[] = {bw, bw / 4, 1, 0};
        for (int c = 0; c < 4; ++c)
            for (int fmt = 0; fmt < 2; ++fmt) bad |= run(bw, cuts[c], fmt);
        bad |= per_order(bw);
        bad |= conv(bw);
    }
    for (int bw = 1; bw <= 64; ++bw)
        for (int m = 0; m < bw; ++m) {
            (void)so_table_size(m, bw);
            for (int i = 0; i <= bw; ++i) (void)so_transpose_row_size(i, m, bw);
        }
    printf(bad ? "oracle sanitizer run: FAILED\n" : "oracle sanitizer run: ok\n");
    return bad;
}
