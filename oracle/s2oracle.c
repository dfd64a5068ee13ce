/*
 * s2oracle.c -- CPU FP64 restatement of S2kit 1.1's memoised seminaive SHT.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP product path is tested against, and the
 * "port" CPU baseline.  Nothing in the product library links or calls it.
 *
 * Written from the mathematics and from the reference's documented behaviour
 * (file:line citations below point into the S2kit 1.1 sources).  FFTW's
 * transforms are restated from the FFTW 3 manual's definitions:
 *   split DFT (sign -1, unnormalised), REDFT10 (DCT-II), REDFT01 (DCT-III).
 * The FFT below is a plain mixed-radix recursive Cooley-Tukey with exact
 * (cos/sin per index) twiddles, valid for every length.
 *
 * Numerical conventions that the reference fixes and that this file keeps,
 * because callers hand tables and coefficients across the API:
 *   - grid f[j*N+k] = f(theta_j, phi_k), theta_j = pi(2j+1)/(4B), N = 2B;
 *   - coefficient order of IndexOfHarmonicCoeff (util.c:42-49);
 *   - zero-stripped cosine-series tables (cospml.c:145-153);
 *   - the odd-B truncation of Transpose_RowSize (cospml.c:270-288), which
 *     makes the reference's inverse drop terms when B is odd.
 */
#include "s2oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    double re, im;
} cx;

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) abort();
    return p;
}

/* ====================================================================== */
/* integer layout arithmetic                                              */
/* ====================================================================== */

/* Number of stored cosine coefficients of P_l^m (m even) or of
 * G_l^m = P_l^m / sin(theta) (m odd): the series has degree l (resp. l-1)
 * and only every other index is non-zero.  cospml.c:250-258. */
int so_row_size(int m, int l) {
    if (l < m) return 0;
    int degree = (m & 1) ? l - 1 : l;
    return degree / 2 + 1;
}

/* Start of row l inside order m's stripped table: rows are packed one after
 * another, so this is the prefix sum of row sizes from l'=m.  Closed form of
 * that sum (cospml.c:123-134): an odd order is the even order m-1 shifted by
 * one degree. */
int so_table_offset(int m, int l) {
    if (m & 1) {
        m -= 1;
        l -= 1;
    }
    int hl = l / 2, hm = m / 2;
    int off = hl * (hl + 1) - hm * (hm + 1);
    if (l & 1) off += hl + 1;
    return off;
}

/* Total stored entries of order m (cospml.c:39-59): sum of row sizes. */
int so_table_size(int m, int bw) {
    int s = 0;
    for (int l = m; l < bw; ++l) s += so_row_size(m, l);
    return s;
}

/* The reference's closed-form upper bound for a full table, with its "+bw"
 * slack; for bw above 512 it falls back to the exact sum (cospml.c:85-92). */
int so_sph_table_size(int bw) {
    if (bw <= 512) return ((4 * bw * bw * bw) + (6 * bw * bw) - (8 * bw)) / 24 + bw;
    return so_reduced_sph_table_size(bw, bw);
}

int so_reduced_sph_table_size(int bw, int m) {
    int s = 0;
    for (int i = 0; i < m; ++i) s += so_table_size(i, bw);
    return s;
}

int so_reduced_naive_table_size(int bw, int m) {
    int s = 0;
    for (int i = m; i < bw; ++i) s += bw - i;
    return 2 * bw * s;
}

/* Entries in row `row` (cosine index) of the transposed order-m table
 * (cospml.c:270-288).  Restated piecewise: an even order keeps (bw-m)/2
 * entries for every cosine index up to m, then one per remaining degree of
 * matching parity; an odd order is the even order m-1 seen one cosine index
 * higher, minus one for odd indices below m.  For odd bw the "(bw-m)/2"
 * branch is one short of the true count; kept, because the reference's
 * inverse is defined by it. */
int so_transpose_row_size(int row, int m, int bw) {
    if (row >= bw) return 0;
    if (!(m & 1)) {
        if (row <= m) return (bw - m) / 2;
        return (bw - row - 1) / 2 + 1;
    }
    if (row == bw - 1) return 0;
    int base = so_transpose_row_size(row + 1, m - 1, bw);
    return row >= m ? base : base - (row & 1);
}

/* util.c:42-49: positive orders first (m = 0..bw-1, l = m..bw-1), then the
 * negative orders m = -(bw-1) .. -1. */
int so_index_of_coeff(int m, int l, int bw) {
    int L = bw - 1;
    if (m >= 0) return m * (L + 1) - (m * (m - 1)) / 2 + (l - m);
    int am = m < 0 ? -m : m;
    return (L * (L + 3)) / 2 + 1 + ((L + m) * (L + m + 1)) / 2 + (l - am);
}

/* ====================================================================== */
/* FFT of any length (mixed radix, recursive, exact twiddles)              */
/* ====================================================================== */

typedef struct {
    int n;
    int sign;
    cx *tw;  /* tw[k] = exp(sign * 2 pi i k / n) */
    cx *tmp; /* n scratch values for the generic butterfly */
} fft_plan;

static void plan_init(fft_plan *p, int n, int sign) {
    p->n = n;
    p->sign = sign;
    p->tw = (cx *)xmalloc(sizeof(cx) * (size_t)n);
    p->tmp = (cx *)xmalloc(sizeof(cx) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        p->tw[k].re = cos(a);
        p->tw[k].im = sign * sin(a);
    }
}

static void plan_free(fft_plan *p) {
    free(p->tw);
    free(p->tmp);
}

static inline cx cmul(cx a, cx b) {
    cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* out[k] = sum_{j<n} in[j*istride] w^{jk},  w = exp(sign 2 pi i / n).
 * `step` = (top-level length)/n, the stride into the twiddle table. */
static void fft_rec(const fft_plan *P, const cx *in, int istride, cx *out, int n, int step) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor(n), q = n / p;
    for (int r = 0; r < p; ++r) fft_rec(P, in + (size_t)r * istride, istride * p, out + (size_t)r * q, q, step * p);
    const int N = P->n;
    if (p == 2) {
        for (int k = 0; k < q; ++k) {
            cx a = out[k], b = cmul(out[q + k], P->tw[(size_t)k * step]);
            out[k].re = a.re + b.re;
            out[k].im = a.im + b.im;
            out[q + k].re = a.re - b.re;
            out[q + k].im = a.im - b.im;
        }
        return;
    }
    if (p == 4) {
        const double s = (double)P->sign; /* w^(n/4) = sign * i */
        for (int k = 0; k < q; ++k) {
            cx a0 = out[k];
            cx a1 = cmul(out[q + k], P->tw[(size_t)k * step]);
            cx a2 = cmul(out[2 * q + k], P->tw[(size_t)2 * k * step]);
            cx a3 = cmul(out[3 * q + k], P->tw[(size_t)3 * k * step]);
            cx t0 = {a0.re + a2.re, a0.im + a2.im}, t1 = {a0.re - a2.re, a0.im - a2.im};
            cx t2 = {a1.re + a3.re, a1.im + a3.im}, t3 = {a1.re - a3.re, a1.im - a3.im};
            /* i*sign*t3 */
            cx u3 = {-s * t3.im, s * t3.re};
            out[k].re = t0.re + t2.re;
            out[k].im = t0.im + t2.im;
            out[q + k].re = t1.re + u3.re;
            out[q + k].im = t1.im + u3.im;
            out[2 * q + k].re = t0.re - t2.re;
            out[2 * q + k].im = t0.im - t2.im;
            out[3 * q + k].re = t1.re - u3.re;
            out[3 * q + k].im = t1.im - u3.im;
        }
        return;
    }
    cx *t = P->tmp;
    for (int k = 0; k < q; ++k) {
        for (int r = 0; r < p; ++r) t[r] = cmul(out[(size_t)r * q + k], P->tw[((size_t)r * k * step) % N]);
        for (int s = 0; s < p; ++s) {
            cx acc = {0.0, 0.0};
            for (int r = 0; r < p; ++r) {
                cx v = cmul(t[r], P->tw[(((size_t)r * s) % p) * q * step]);
                acc.re += v.re;
                acc.im += v.im;
            }
            out[k + (size_t)q * s] = acc;
        }
    }
}

static void fft_exec(const fft_plan *P, const cx *in, cx *out) { fft_rec(P, in, 1, out, P->n, 1); }

void so_dft(int n, const double *re, const double *im, double *ore, double *oim, int sign) {
    fft_plan P;
    plan_init(&P, n, sign < 0 ? -1 : 1);
    cx *a = (cx *)xmalloc(sizeof(cx) * (size_t)n), *b = (cx *)xmalloc(sizeof(cx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        a[j].re = re[j];
        a[j].im = im ? im[j] : 0.0;
    }
    fft_exec(&P, a, b);
    for (int k = 0; k < n; ++k) {
        ore[k] = b[k].re;
        oim[k] = b[k].im;
    }
    free(a);
    free(b);
    plan_free(&P);
}

/* ---- DCTs of complex sequences through a length-2n FFT ----------------- */
/* Both DCTs below are real-linear with real kernels, so a complex sequence
 * re + i*im yields DCT(re) + i*DCT(im). */

typedef struct {
    int n;
    fft_plan fwd2n; /* length 2n, sign -1 */
    fft_plan inv2n; /* length 2n, sign +1 */
    cx *half;       /* half[k] = exp(-i pi k / (2n)), k < n */
    cx *y, *z;      /* 2n scratch */
} dct_plan;

static void dct_init(dct_plan *D, int n) {
    D->n = n;
    plan_init(&D->fwd2n, 2 * n, -1);
    plan_init(&D->inv2n, 2 * n, +1);
    D->half = (cx *)xmalloc(sizeof(cx) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = M_PI * (double)k / (2.0 * n);
        D->half[k].re = cos(a);
        D->half[k].im = -sin(a);
    }
    D->y = (cx *)xmalloc(sizeof(cx) * (size_t)(2 * n));
    D->z = (cx *)xmalloc(sizeof(cx) * (size_t)(2 * n));
}

static void dct_free(dct_plan *D) {
    plan_free(&D->fwd2n);
    plan_free(&D->inv2n);
    free(D->half);
    free(D->y);
    free(D->z);
}

/* REDFT10: c[k] = 2 sum_j x[j] cos(pi (j+1/2) k / n).  With y = x followed by
 * x reversed, Y = DFT_2n(y) gives c[k] = exp(-i pi k/(2n)) Y[k]. */
static void dct2_cx(dct_plan *D, const cx *x, cx *c) {
    const int n = D->n;
    for (int j = 0; j < n; ++j) {
        D->y[j] = x[j];
        D->y[2 * n - 1 - j] = x[j];
    }
    fft_exec(&D->fwd2n, D->y, D->z);
    for (int k = 0; k < n; ++k) c[k] = cmul(D->half[k], D->z[k]);
}

/* REDFT01: y[k] = X[0] + 2 sum_{j>=1} X[j] cos(pi j (k+1/2) / n).  Writing the
 * cosine as two exponentials gives y = IDFT_2n(Z) (unnormalised, +i) with
 * Z[0] = X[0], Z[j] = X[j] e^{+i pi j/(2n)}, Z[2n-j] = X[j] e^{-i pi j/(2n)}. */
static void dct3_cx(dct_plan *D, const cx *X, cx *y) {
    const int n = D->n;
    D->y[0] = X[0];
    D->y[n].re = D->y[n].im = 0.0;
    for (int j = 1; j < n; ++j) {
        cx e = D->half[j]; /* e^{-i pi j / 2n} */
        cx ec = {e.re, -e.im};
        D->y[j] = cmul(X[j], ec);
        D->y[2 * n - j] = cmul(X[j], e);
    }
    fft_exec(&D->inv2n, D->y, D->z);
    for (int k = 0; k < n; ++k) y[k] = D->z[k];
}

void so_redft10(int n, const double *in, double *out) {
    dct_plan D;
    dct_init(&D, n);
    cx *x = (cx *)xmalloc(sizeof(cx) * (size_t)n), *c = (cx *)xmalloc(sizeof(cx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        x[j].re = in[j];
        x[j].im = 0.0;
    }
    dct2_cx(&D, x, c);
    for (int k = 0; k < n; ++k) out[k] = c[k].re;
    free(x);
    free(c);
    dct_free(&D);
}

void so_redft01(int n, const double *in, double *out) {
    dct_plan D;
    dct_init(&D, n);
    cx *x = (cx *)xmalloc(sizeof(cx) * (size_t)n), *c = (cx *)xmalloc(sizeof(cx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        x[j].re = in[j];
        x[j].im = 0.0;
    }
    dct3_cx(&D, x, c);
    for (int k = 0; k < n; ++k) out[k] = c[k].re;
    free(x);
    free(c);
    dct_free(&D);
}

/* ====================================================================== */
/* weights, Legendre seeds and recurrences                                 */
/* ====================================================================== */

/* Driscoll-Healy quadrature weights on theta_j = pi(2j+1)/(4B), j < 2B
 * (weights.c:11-15, 32-46).  w[2B + j] = w[j] sin(theta_j) serves the odd
 * orders, whose tables hold P/sin. */
void so_weights(int bw, double *w) {
    const double c = M_PI / (4.0 * bw);
    for (int i = 0; i < 2 * bw; ++i) {
        double k = 2.0 * i + 1.0, s = 0.0;
        for (int j = 0; j < bw; ++j) s += 1.0 / (2.0 * j + 1.0) * sin(k * (2.0 * j + 1.0) * c);
        s *= 2.0 * sin(k * c) / bw;
        w[i] = s;
        w[i + 2 * bw] = s * sin(k * c);
    }
}

/* L2-normalised P_m^m(cos a) = c_m sin^m a with the Condon-Shortley sign
 * (pmm.c:21-34). */
void so_pmm_l2(int m, const double *eval_args, int n, double *out) {
    double c = sqrt(m + 0.5);
    for (int i = 0; i < m; ++i) c *= sqrt((m - (i / 2.0)) / ((double)m - i));
    if (m) c *= pow(2.0, -m / 2.0);
    if (m & 1) c = -c;
    for (int i = 0; i < n; ++i) out[i] = c * pow(sin(eval_args[i]), m);
}

/* Three-term recurrence P_{l+1} = an(m,l) x P_l + cn(m,l) P_{l-1} for the
 * L2-normalised functions (l2_norms.c:16-38). */
double so_l2_an(int m, int l) {
    return sqrt(((2.0 * l + 3.0) / (2.0 * l + 1.0)) * ((l - m + 1.0) / (l + m + 1.0))) *
           ((2.0 * l + 1.0) / (l - m + 1.0));
}

double so_l2_cn(int m, int l) {
    if (l == 0) return 0.0;
    return -1.0 *
           sqrt(((2.0 * l + 3.0) / (2.0 * l - 1.0)) * ((l - m + 1.0) / (l + m + 1.0)) *
                (((double)l - m) / ((double)l + m))) *
           ((l + m) / (l - m + 1.0));
}

/* Samples of P_l^m (l = m..bw-1) at n nodes theta_i = (2i+1)pi/(2n); row
 * l-m of `rows` (n doubles each).  Seed 1/sqrt(2) for m = 0.  If div_sin,
 * the seed is divided by sin(theta) (G_l^m for odd m, cospml.c:190-192). */
static void legendre_rows(int bw, int m, int n, int div_sin, double *rows) {
    double *arg = (double *)xmalloc(sizeof(double) * (size_t)n);
    double *x = (double *)xmalloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        arg[i] = (2.0 * i + 1.0) * M_PI / (2.0 * n);
        x[i] = cos((2.0 * i + 1.0) * M_PI / (2.0 * n));
    }
    double *p0 = rows;
    if (m == 0)
        for (int i = 0; i < n; ++i) p0[i] = sqrt(0.5);
    else
        so_pmm_l2(m, arg, n, p0);
    if (div_sin)
        for (int i = 0; i < n; ++i) p0[i] /= sin(arg[i]);
    for (int l = m; l + 1 < bw; ++l) {
        const double an = so_l2_an(m, l), cn = so_l2_cn(m, l);
        const double *cur = rows + (size_t)(l - m) * n;
        const double *prv = (l > m) ? cur - n : NULL;
        double *nxt = rows + (size_t)(l - m + 1) * n;
        for (int i = 0; i < n; ++i) {
            double a = an * (cur[i] * x[i]);
            double b = prv ? cn * prv[i] : cn * 0.0;
            nxt[i] = a + b;
        }
    }
    free(arg);
    free(x);
}

/* Cosine-series table of order m (cospml.c:161-242): the P_l^m (or G_l^m)
 * sampled at bw Chebyshev nodes, REDFT10 of length bw, c[0] *= 1/sqrt2,
 * all *= 1/sqrt(bw), stored zero-stripped row by row. */
void so_cospml_table(int bw, int m, double *table) {
    const int nrow = bw - m;
    double *rows = (double *)xmalloc(sizeof(double) * (size_t)nrow * bw);
    legendre_rows(bw, m, bw, m & 1, rows);
    dct_plan D;
    dct_init(&D, bw);
    cx *x = (cx *)xmalloc(sizeof(cx) * (size_t)bw), *c = (cx *)xmalloc(sizeof(cx) * (size_t)bw);
    const double fudge = 1.0 / sqrt((double)bw);
    for (int l = m; l < bw; ++l) {
        const double *r = rows + (size_t)(l - m) * bw;
        for (int i = 0; i < bw; ++i) {
            x[i].re = r[i];
            x[i].im = 0.0;
        }
        dct2_cx(&D, x, c);
        c[0].re *= sqrt(0.5);
        const int tau = (l - m) & 1, rs = so_row_size(m, l), off = so_table_offset(m, l);
        for (int t = 0; t < rs; ++t) table[off + t] = c[2 * t + tau].re * fudge;
    }
    free(x);
    free(c);
    dct_free(&D);
    free(rows);
}

/* Transposed table (cospml.c:301-362), restated from what it holds: row i
 * (cosine index) lists, for rising degree l of matching parity, the cosine
 * coefficient i of row l.  Row i starts at degree l0(i) = m + (i - m) +
 * (m mod 2) for i > m, else m + (i mod 2) -- the same coefficient offsets
 * InvDLTSemi reads (seminaive.c:81-84) -- and has Transpose_RowSize entries. */
void so_transpose_cospml_table(int bw, int m, const double *t, double *tt) {
    if (m == bw - 1) { /* single row: the layouts coincide (cospml.c:309-312) */
        memcpy(tt, t, sizeof(double) * (size_t)so_table_size(m, bw));
        return;
    }
    double *o = tt;
    for (int i = 0; i < bw; ++i) {
        if (i == bw - 1 && (m & 1)) break;
        const int l0 = (i > m) ? i + (m & 1) : m + (i & 1);
        const int rs = so_transpose_row_size(i, m, bw);
        for (int k = 0; k < rs; ++k) {
            const int l = l0 + 2 * k;
            o[k] = t[so_table_offset(m, l) + i / 2];
        }
        o += rs;
    }
}

/* Naive table: P_l^m at the 2B grid latitudes, one row of 2B per degree
 * (pml.c:41-79). */
void so_pml_table(int bw, int m, double *table) { legendre_rows(bw, m, 2 * bw, 0, table); }

void so_tables(int bw, int cutoff, double *fwd, double *inv, long *off) {
    long pos = 0;
    for (int m = 0; m < bw; ++m) {
        off[m] = pos;
        if (m < cutoff) {
            if (fwd) so_cospml_table(bw, m, fwd + pos);
            if (inv) {
                double *tmp = fwd ? fwd + pos : (double *)xmalloc(sizeof(double) * (size_t)so_table_size(m, bw));
                if (!fwd) so_cospml_table(bw, m, tmp);
                so_transpose_cospml_table(bw, m, tmp, inv + pos);
                if (!fwd) free(tmp);
            }
            pos += so_table_size(m, bw);
        } else {
            if (fwd) so_pml_table(bw, m, fwd + pos);
            if (inv) so_pml_table(bw, m, inv + pos);
            pos += (long)2 * bw * (bw - m);
        }
    }
}

/* ====================================================================== */
/* per-order Legendre transforms                                          */
/* ====================================================================== */

/* Projection of one weighted, cosine-transformed row onto the order-m
 * table (seminaive.c:183-197): degree l uses cosine indices of parity
 * (l-m) mod 2, RowSize(m,l) of them. */
static void semi_project(const double *c, int stride, int bw, int m, const double *tab, double *out) {
    for (int l = m; l < bw; ++l) {
        const int tau = (l - m) & 1, rs = so_row_size(m, l);
        const double *row = tab + so_table_offset(m, l);
        double v = 0.0;
        for (int t = 0; t < rs; ++t) v += c[(size_t)(2 * t + tau) * stride] * row[t];
        out[l - m] = v;
    }
}

/* weights -> REDFT10 -> scaling (seminaive.c:161-176), on a complex row */
static void semi_cosines(dct_plan *D, const cx *row, int bw, int m, const double *w, cx *g, cx *c) {
    const int N = 2 * bw;
    const double *wp = (m & 1) ? w + N : w;
    for (int i = 0; i < N; ++i) {
        g[i].re = row[i].re * wp[i];
        g[i].im = row[i].im * wp[i];
    }
    dct2_cx(D, g, c);
    c[0].re *= sqrt(0.5);
    c[0].im *= sqrt(0.5);
    const double s = 1.0 / sqrt(2.0 * N);
    for (int i = 0; i < N; ++i) {
        c[i].re *= s;
        c[i].im *= s;
    }
}

void so_dlt_semi(const double *data, int bw, int m, double *result, const double *tab, const double *w) {
    const int N = 2 * bw;
    dct_plan D;
    dct_init(&D, N);
    cx *row = (cx *)xmalloc(sizeof(cx) * (size_t)N), *g = (cx *)xmalloc(sizeof(cx) * (size_t)N),
       *c = (cx *)xmalloc(sizeof(cx) * (size_t)N);
    for (int i = 0; i < N; ++i) {
        row[i].re = data[i];
        row[i].im = 0.0;
    }
    semi_cosines(&D, row, bw, m, w, g, c);
    semi_project(&c[0].re, 2, bw, m, tab, result);
    free(row);
    free(g);
    free(c);
    dct_free(&D);
}

/* Cosine coefficients of the inverse (seminaive.c:60-98): fcos[i] =
 * (1/(2 sqrt B)) sum_k a(l0(i)+2k) TT[i][k]; fcos[0] rescaled to sum/sqrt(N);
 * fcos[B-1] = 0 for odd m; fcos[B..N) = 0. `cstride` is the distance between
 * consecutive degrees in `coeffs`. */
static void inv_semi_cosines(const double *coeffs, int cstride, int bw, int m, const double *tt, double *fcos) {
    const int N = 2 * bw;
    const double scale = 0.5 / sqrt((double)bw);
    for (int i = 0; i < N; ++i) fcos[i] = 0.0;
    const double *tp = tt;
    for (int i = 0; i < bw; ++i) {
        if (i == bw - 1 && (m & 1)) {
            fcos[i] = 0.0;
            break;
        }
        const int first = (i > m) ? (i - m) + (m & 1) : (i & 1); /* offset of degree l0 - m */
        const int rs = so_transpose_row_size(i, m, bw);
        double v = 0.0;
        for (int k = 0; k < rs; ++k) v += coeffs[(size_t)(first + 2 * k) * cstride] * tp[k];
        fcos[i] = v * scale;
        tp += rs;
    }
    fcos[0] /= (sqrt((double)N) * scale);
}

void so_inv_dlt_semi(const double *coeffs, int bw, int m, double *result, const double *tt, const double *sinv) {
    const int N = 2 * bw;
    double *fc = (double *)xmalloc(sizeof(double) * (size_t)N);
    inv_semi_cosines(coeffs, 1, bw, m, tt, fc);
    so_redft01(N, fc, result);
    if (m & 1)
        for (int i = 0; i < N; ++i) result[i] *= sinv[i];
    free(fc);
}

void so_dlt_naive(const double *data, int bw, int m, const double *w, double *result, const double *pml) {
    const int N = 2 * bw;
    double *wd = (double *)xmalloc(sizeof(double) * (size_t)N);
    for (int j = 0; j < N; ++j) wd[j] = data[j] * w[j];
    for (int i = 0; i < bw - m; ++i) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += wd[j] * pml[(size_t)i * N + j];
        result[i] = s;
    }
    free(wd);
}

void so_inv_dlt_naive(const double *coeffs, int bw, int m, double *result, const double *pml) {
    const int N = 2 * bw;
    for (int j = 0; j < N; ++j) result[j] = 0.0;
    for (int i = 0; i < bw - m; ++i) {
        const double a = coeffs[i];
        if (a == 0.0) continue;
        for (int j = 0; j < N; ++j) result[j] += a * pml[(size_t)i * N + j];
    }
}

/* ====================================================================== */
/* spherical transforms                                                   */
/* ====================================================================== */

/* Forward DLT of one complex order row (real and imaginary parts are
 * independent transforms in the reference; both use the same table). */
static void fwd_order(dct_plan *D, const cx *row, int bw, int am, int cutoff, const double *tab, const double *w,
                      cx *g, cx *c, double *ore, double *oim, double *tmp) {
    const int N = 2 * bw;
    if (am < cutoff) {
        semi_cosines(D, row, bw, am, w, g, c);
        semi_project(&c[0].re, 2, bw, am, tab, ore);
        semi_project(&c[0].im, 2, bw, am, tab, oim);
    } else { /* DLTNaive: even weights for every order (naive.c:38-48) */
        for (int j = 0; j < N; ++j) tmp[j] = row[j].re;
        so_dlt_naive(tmp, bw, am, w, ore, tab);
        for (int j = 0; j < N; ++j) tmp[j] = row[j].im;
        so_dlt_naive(tmp, bw, am, w, oim, tab);
    }
}

void so_fst(const double *rdata, const double *idata, double *rc, double *ic, int bw, const double *const *tables,
            const double *w, int data_format, int cutoff) {
    const int N = 2 * bw;
    cx *F = (cx *)xmalloc(sizeof(cx) * (size_t)N * N); /* F[r*N + j], order-major */
    cx *a = (cx *)xmalloc(sizeof(cx) * (size_t)N), *b = (cx *)xmalloc(sizeof(cx) * (size_t)N);
    fft_plan P;
    plan_init(&P, N, -1);
    /* phi-FFT per ring, transposed store, scaled by sqrt(2 pi)/N (:81-90) */
    const double nc = sqrt(2.0 * M_PI) / N;
    for (int j = 0; j < N; ++j) {
        for (int k = 0; k < N; ++k) {
            a[k].re = rdata[(size_t)j * N + k];
            a[k].im = idata[(size_t)j * N + k];
        }
        fft_exec(&P, a, b);
        for (int r = 0; r < N; ++r) {
            F[(size_t)r * N + j].re = b[r].re * nc;
            F[(size_t)r * N + j].im = b[r].im * nc;
        }
    }
    plan_free(&P);
    dct_plan D;
    dct_init(&D, N);
    cx *g = (cx *)xmalloc(sizeof(cx) * (size_t)N), *c = (cx *)xmalloc(sizeof(cx) * (size_t)N);
    double *ore = (double *)xmalloc(sizeof(double) * (size_t)bw), *oim = (double *)xmalloc(sizeof(double) * (size_t)bw);
    double *tmp = (double *)xmalloc(sizeof(double) * (size_t)N);
    /* orders 0..bw-1 in row m (:96-120) */
    for (int m = 0; m < bw; ++m) {
        fwd_order(&D, F + (size_t)m * N, bw, m, cutoff, tables[m], w, g, c, ore, oim, tmp);
        const int base = so_index_of_coeff(m, m, bw);
        for (int l = m; l < bw; ++l) {
            rc[base + l - m] = ore[l - m];
            ic[base + l - m] = oim[l - m];
        }
    }
    if (data_format == 1) { /* REAL: f(l,-m) = (-1)^m conj f(l,m) (:131-145) */
        for (int m = 1; m < bw; ++m) {
            const double s = (m & 1) ? -1.0 : 1.0;
            for (int l = m; l < bw; ++l) {
                int p = so_index_of_coeff(m, l, bw), q = so_index_of_coeff(-m, l, bw);
                rc[q] = s * rc[p];
                ic[q] = -s * ic[p];
            }
        }
    } else { /* rows r = bw+1..N-1 hold orders -(N-r); (-1)^|m| (:153-201) */
        for (int r = bw + 1; r < N; ++r) {
            const int am = N - r;
            fwd_order(&D, F + (size_t)r * N, bw, am, cutoff, tables[am], w, g, c, ore, oim, tmp);
            const double s = (r & 1) ? -1.0 : 1.0;
            const int base = so_index_of_coeff(-am, am, bw);
            for (int l = am; l < bw; ++l) {
                rc[base + l - am] = s * ore[l - am];
                ic[base + l - am] = s * oim[l - am];
            }
        }
    }
    free(tmp);
    free(ore);
    free(oim);
    free(g);
    free(c);
    dct_free(&D);
    free(a);
    free(b);
    free(F);
}

/* Inverse DLT of one complex order (coefficients of degrees am..bw-1). */
static void inv_order(dct_plan *D, const double *cre, const double *cim, int bw, int am, int cutoff,
                      const double *tab, const double *sinv, double *fr, double *fi, cx *X, cx *y, cx *out) {
    const int N = 2 * bw;
    if (am < cutoff) {
        inv_semi_cosines(cre, 1, bw, am, tab, fr);
        inv_semi_cosines(cim, 1, bw, am, tab, fi);
        for (int i = 0; i < N; ++i) {
            X[i].re = fr[i];
            X[i].im = fi[i];
        }
        dct3_cx(D, X, y);
        if (am & 1)
            for (int i = 0; i < N; ++i) {
                y[i].re *= sinv[i];
                y[i].im *= sinv[i];
            }
        for (int i = 0; i < N; ++i) out[i] = y[i];
    } else {
        so_inv_dlt_naive(cre, bw, am, fr, tab);
        so_inv_dlt_naive(cim, bw, am, fi, tab);
        for (int i = 0; i < N; ++i) {
            out[i].re = fr[i];
            out[i].im = fi[i];
        }
    }
}

void so_inv_fst(const double *rc, const double *ic, double *rdata, double *idata, int bw,
                const double *const *tt, int data_format, int cutoff) {
    const int N = 2 * bw;
    double *sinv = (double *)xmalloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; ++i) sinv[i] = sin((2.0 * i + 1.0) * M_PI / (2.0 * N));
    cx *G = (cx *)xmalloc(sizeof(cx) * (size_t)N * N); /* G[r*N + j] */
    dct_plan D;
    dct_init(&D, N);
    double *fr = (double *)xmalloc(sizeof(double) * (size_t)N), *fi = (double *)xmalloc(sizeof(double) * (size_t)N);
    cx *X = (cx *)xmalloc(sizeof(cx) * (size_t)N), *y = (cx *)xmalloc(sizeof(cx) * (size_t)N);
    for (int m = 0; m < bw; ++m) {
        const int base = so_index_of_coeff(m, m, bw);
        inv_order(&D, rc + base, ic + base, bw, m, cutoff, tt[m], sinv, fr, fi, X, y, G + (size_t)m * N);
    }
    for (int j = 0; j < N; ++j) G[(size_t)bw * N + j].re = G[(size_t)bw * N + j].im = 0.0; /* (:283-284) */
    if (data_format == 0) {
        for (int r = bw + 1; r < N; ++r) {
            const int am = N - r;
            const int base = so_index_of_coeff(-am, am, bw);
            cx *row = G + (size_t)r * N;
            inv_order(&D, rc + base, ic + base, bw, am, cutoff, tt[am], sinv, fr, fi, X, y, row);
            if (r & 1)
                for (int j = 0; j < N; ++j) {
                    row[j].re = -row[j].re;
                    row[j].im = -row[j].im;
                }
        }
    } else { /* REAL: row N-m = conj(row m) (:333-341) */
        for (int r = bw + 1; r < N; ++r)
            for (int j = 0; j < N; ++j) {
                G[(size_t)r * N + j].re = G[(size_t)(N - r) * N + j].re;
                G[(size_t)r * N + j].im = -G[(size_t)(N - r) * N + j].im;
            }
    }
    /* x 1/sqrt(2 pi), then the +i DFT along phi per ring (:344-350) */
    const double nc = 1.0 / sqrt(2.0 * M_PI);
    fft_plan P;
    plan_init(&P, N, +1);
    for (int j = 0; j < N; ++j) {
        for (int r = 0; r < N; ++r) {
            X[r].re = G[(size_t)r * N + j].re * nc;
            X[r].im = G[(size_t)r * N + j].im * nc;
        }
        fft_exec(&P, X, y);
        for (int k = 0; k < N; ++k) {
            rdata[(size_t)j * N + k] = y[k].re;
            idata[(size_t)j * N + k] = y[k].im;
        }
    }
    plan_free(&P);
    free(X);
    free(y);
    free(fr);
    free(fi);
    dct_free(&D);
    free(G);
    free(sinv);
}

/* Zonal transform (FST_semi_memo.c:374-407): the m = 0 bin of each ring as a
 * direct sum, scaled by sqrt(2 pi)/N, then DLTSemi at m = 0.  REAL zeroes
 * ires; this restatement zeroes the documented bw entries (the reference
 * writes 2*bw doubles there, past the documented length). */
void so_fzt(const double *rdata, const double *idata, double *rres, double *ires, int bw, const double *tab0,
            const double *w, int data_format) {
    const int N = 2 * bw;
    const double ds = sqrt(2.0 * M_PI) / N;
    double *r0 = (double *)xmalloc(sizeof(double) * (size_t)N), *i0 = (double *)xmalloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < N; ++i) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < N; ++j) {
            sr += rdata[(size_t)i * N + j];
            si += idata[(size_t)i * N + j];
        }
        r0[i] = sr * ds;
        i0[i] = si * ds;
    }
    so_dlt_semi(r0, bw, 0, rres, tab0, w);
    if (data_format == 0)
        so_dlt_semi(i0, bw, 0, ires, tab0, w);
    else
        for (int i = 0; i < bw; ++i) ires[i] = 0.0;
    free(r0);
    free(i0);
}

/* Filter x signal in coefficient space with sqrt(4 pi / (2l+1))
 * (util.c:68-103).  The complex product keeps the reference's imaginary
 * part x*v - y*u (util.c:26), which differs from the true product only
 * when the filter has an imaginary part. */
void so_trans_mult(const double *rdc, const double *idc, const double *rfc, const double *ifc, double *rres,
                   double *ires, int bw) {
    for (int m = -(bw - 1); m < bw; ++m) {
        const int am = m < 0 ? -m : m;
        for (int l = am; l < bw; ++l) {
            const int p = so_index_of_coeff(m, l, bw);
            const double x = rfc[l], y = ifc[l], u = rdc[p], v = idc[p];
            double re = x * u - y * v;
            double im = x * v - y * u;
            re *= sqrt(4.0 * M_PI / (2.0 * l + 1.0));
            im *= sqrt(4.0 * M_PI / (2.0 * l + 1.0));
            rres[p] = re;
            ires[p] = im;
        }
    }
}

void so_conv(const double *rdata, const double *idata, const double *rfilter, const double *ifilter, double *rres,
             double *ires, int bw) {
    const int cutoff = bw;
    const long tsz = so_reduced_sph_table_size(bw, bw);
    double *fwd = (double *)xmalloc(sizeof(double) * (size_t)tsz), *inv = (double *)xmalloc(sizeof(double) * (size_t)tsz);
    long *off = (long *)xmalloc(sizeof(long) * (size_t)bw);
    so_tables(bw, cutoff, fwd, inv, off);
    const double **tf = (const double **)xmalloc(sizeof(double *) * (size_t)bw);
    const double **ti = (const double **)xmalloc(sizeof(double *) * (size_t)bw);
    for (int m = 0; m < bw; ++m) {
        tf[m] = fwd + off[m];
        ti[m] = inv + off[m];
    }
    double *w = (double *)xmalloc(sizeof(double) * (size_t)4 * bw);
    so_weights(bw, w);
    const size_t B2 = (size_t)bw * bw;
    double *fr = (double *)xmalloc(sizeof(double) * B2), *fi = (double *)xmalloc(sizeof(double) * B2);
    double *tr = (double *)xmalloc(sizeof(double) * B2), *tii = (double *)xmalloc(sizeof(double) * B2);
    double *hr = (double *)xmalloc(sizeof(double) * (size_t)bw), *hi = (double *)xmalloc(sizeof(double) * (size_t)bw);
    so_fst(rdata, idata, fr, fi, bw, tf, w, 1, cutoff);
    so_fzt(rfilter, ifilter, hr, hi, bw, tf[0], w, 1);
    so_trans_mult(fr, fi, hr, hi, tr, tii, bw);
    so_inv_fst(tr, tii, rres, ires, bw, ti, 1, cutoff);
    free(fr);
    free(fi);
    free(tr);
    free(tii);
    free(hr);
    free(hi);
    free(w);
    free(tf);
    free(ti);
    free(off);
    free(fwd);
    free(inv);
}

/* ====================================================================== */
/* deterministic inputs                                                   */
/* ====================================================================== */

static unsigned long long splitmix64(unsigned long long *s) {
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double unif(unsigned long long *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

void so_uniform(unsigned long long seed, long n, double *out) {
    unsigned long long s = seed;
    for (long i = 0; i < n; ++i) out[i] = unif(&s);
}

void so_make_coeffs(unsigned long long seed, int bw, int kind, double *rc, double *ic) {
    unsigned long long s = seed;
    if (kind == 2) { /* Hermitian, as test_s2_semi_memo.c:156-172 builds it */
        for (int m = 0; m < bw; ++m)
            for (int l = m; l < bw; ++l) {
                double x = 2.0 * (unif(&s) - 0.5), y = 2.0 * (unif(&s) - 0.5);
                int p = so_index_of_coeff(m, l, bw), q = so_index_of_coeff(-m, l, bw);
                rc[p] = x;
                ic[p] = y;
                rc[q] = ((m & 1) ? -1.0 : 1.0) * x;
                ic[q] = ((m & 1) ? 1.0 : -1.0) * y;
            }
        for (int l = 0; l < bw; ++l) ic[l] = 0.0;
        return;
    }
    for (int m = -(bw - 1); m < bw; ++m) {
        const int am = m < 0 ? -m : m;
        for (int l = am; l < bw; ++l) {
            const int p = so_index_of_coeff(m, l, bw);
            if (kind == 0) {
                rc[p] = 2.0 * unif(&s) - 1.0;
                ic[p] = 2.0 * unif(&s) - 1.0;
            } else { /* Box-Muller, variance (l+1)^-2 split over re and im */
                double u1 = unif(&s), u2 = unif(&s);
                double r = sqrt(-2.0 * log(1.0 - u1));
                double sd = sqrt(0.5) / (l + 1.0);
                rc[p] = sd * r * cos(2.0 * M_PI * u2);
                ic[p] = sd * r * sin(2.0 * M_PI * u2);
            }
        }
    }
}
