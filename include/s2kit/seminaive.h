/* Drop-in for S2kit's include/s2kit/seminaive.h (seminaive.h:6-8): the
 * per-order seminaive Legendre transforms, on host arrays, computed on the
 * current HIP device in FP64.  The FFTW plan and workspace arguments are
 * accepted and ignored (NULL is fine); errors behave as in FST_semi_memo.h. */
#ifndef S2KIT_AMD_SEMINAIVE_H
#define S2KIT_AMD_SEMINAIVE_H

#ifndef S2KIT_AMD_FFTW_PLAN
#define S2KIT_AMD_FFTW_PLAN
typedef struct fftw_plan_s* fftw_plan;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces DLTSemi (seminaive.h:6, seminaive.c:153-198): data [2bw] ->
 * result [bw-m]; cos_pml_table from GenerateCosPmlTable(bw, m); weights from
 * GenerateWeightsForDLT (the odd-order half is used for odd m). */
void DLTSemi(double* data, const int bw, const int m, double* result, double* workspace, double* cos_pml_table,
             double* weights, fftw_plan* plan);

/* Replaces InvDLTSemi (seminaive.h:8, seminaive.c:56-115): coeffs [bw-m] ->
 * result [2bw]; trans_cos_pml_table from TransposeCosPmlTable; sin_values =
 * sin(theta_j) at the 2bw grid latitudes (used for odd m). */
void InvDLTSemi(double* coeffs, const int bw, const int m, double* result, double* trans_cos_pml_table,
                double* sin_values, double* workspace, fftw_plan* plan);

#ifdef __cplusplus
}
#endif
#endif
