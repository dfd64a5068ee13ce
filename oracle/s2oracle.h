/*
 * s2oracle.h -- CPU FP64 restatement of S2kit's memoised seminaive spherical
 * harmonic transform.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP product path (s2kit_amd/csrc) is tested
 * against this library, which is also the "port" CPU baseline timed by bench.py.  It is
 * never linked into, loaded by, or called from the product library.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The FFTW 3 transforms the reference calls (split DFT, REDFT10, REDFT01) are
 * restated from the public FFTW 3 manual's definitions with an FFT of our own.
 */
#ifndef S2ORACLE_H
#define S2ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- integer layout arithmetic (cospml.c, util.c) ---------------------- */
int so_table_size(int m, int bw);                 /* cospml.c:39-59   */
int so_sph_table_size(int bw);                    /* cospml.c:85-92   */
int so_reduced_sph_table_size(int bw, int m);     /* cospml.c:107-115 */
int so_reduced_naive_table_size(int bw, int m);   /* cospml.c:434-440 */
int so_table_offset(int m, int l);                /* cospml.c:123-134 */
int so_row_size(int m, int l);                    /* cospml.c:250-258 */
int so_transpose_row_size(int row, int m, int bw);/* cospml.c:270-288 */
int so_index_of_coeff(int m, int l, int bw);      /* util.c:42-49     */

/* ---- FFTW-defined primitives (public FFTW 3 manual definitions) -------- */
/* out[k] = sum_j in[j] exp(sign * 2 pi i j k / n), unnormalised, any n >= 1 */
void so_dft(int n, const double *re, const double *im, double *ore, double *oim, int sign);
/* REDFT10: Y_k = 2 sum_j X_j cos(pi (j+1/2) k / n) */
void so_redft10(int n, const double *in, double *out);
/* REDFT01: Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2) / n) */
void so_redft01(int n, const double *in, double *out);

/* ---- tables and weights ------------------------------------------------ */
void so_weights(int bw, double *w);                                    /* weights.c:32-46 */
void so_pmm_l2(int m, const double *eval_args, int n, double *out);    /* pmm.c:21-34 */
double so_l2_an(int m, int l);                                         /* l2_norms.c:16-22 */
double so_l2_cn(int m, int l);                                         /* l2_norms.c:28-38 */
void so_cospml_table(int bw, int m, double *table);                    /* cospml.c:161-242 */
void so_transpose_cospml_table(int bw, int m, const double *t, double *tt); /* cospml.c:301-362 */
void so_pml_table(int bw, int m, double *table);                       /* pml.c:41-79 */
/* Mixed seminaive (m < cutoff) + naive (m >= cutoff) tables, forward and
 * transposed, packed as SemiNaive_Naive_Pml_Table / Transpose_... do
 * (cospml.c:451-518).  off[m] receives the start of order m's table (bw
 * entries); fwd and inv must hold so_reduced_naive_table_size(bw,cutoff) +
 * so_reduced_sph_table_size(bw,cutoff) doubles.  Either space may be NULL. */
void so_tables(int bw, int cutoff, double *fwd, double *inv, long *off);

/* ---- per-order Legendre transforms ------------------------------------ */
void so_dlt_semi(const double *data, int bw, int m, double *result,
                 const double *cos_pml_table, const double *weights);   /* seminaive.c:153-198 */
void so_inv_dlt_semi(const double *coeffs, int bw, int m, double *result,
                     const double *trans_cos_pml_table, const double *sin_values); /* seminaive.c:56-115 */
void so_dlt_naive(const double *data, int bw, int m, const double *weights,
                  double *result, const double *pml_table);            /* naive.c:34-59 */
void so_inv_dlt_naive(const double *coeffs, int bw, int m, double *result,
                      const double *pml_table);                        /* naive.c:78-95 */

/* ---- spherical transforms (FST_semi_memo.c) --------------------------- */
/* tables[m] = start of order m's (seminaive or naive) table. data_format:
 * 0 = COMPLEX, 1 = REAL (util.h:10-13). */
void so_fst(const double *rdata, const double *idata, double *rcoeffs, double *icoeffs,
            int bw, const double *const *tables, const double *weights,
            int data_format, int cutoff);                               /* :68-202 */
void so_inv_fst(const double *rcoeffs, const double *icoeffs, double *rdata, double *idata,
                int bw, const double *const *ttables, int data_format, int cutoff); /* :228-351 */
void so_fzt(const double *rdata, const double *idata, double *rres, double *ires,
            int bw, const double *cos_pml_table0, const double *weights,
            int data_format);                                           /* :374-407 */
void so_trans_mult(const double *rdc, const double *idc, const double *rfc, const double *ifc,
                   double *rres, double *ires, int bw);                 /* util.c:68-103 */
/* ConvOn2SphereSemiMemo (:439-518): REAL, cutoff = bw, tables built inside. */
void so_conv(const double *rdata, const double *idata, const double *rfilter,
             const double *ifilter, double *rres, double *ires, int bw);

/* ---- deterministic input generator shared by tests and bench ----------- */
/* splitmix64 stream -> 53-bit uniforms in [0,1) */
void so_uniform(unsigned long long seed, long n, double *out);
/* coefficients: kind 0 = uniform[-1,1] re/im for every (m,l);
 * kind 1 = complex Gaussian, variance (l+1)^-2 split evenly over re and im;
 * kind 2 = like 0, but Hermitian (test_s2_semi_memo.c:156-172), so the grid is real. */
void so_make_coeffs(unsigned long long seed, int bw, int kind, double *rc, double *ic);

#ifdef __cplusplus
}
#endif
#endif
