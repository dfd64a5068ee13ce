/* Drop-in for S2kit's include/s2kit/naive.h (naive.h:4-6): the per-order
 * naive Legendre transforms on host arrays, computed on the current HIP
 * device in FP64; errors behave as in FST_semi_memo.h. */
#ifndef S2KIT_AMD_NAIVE_H
#define S2KIT_AMD_NAIVE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces InvDLTNaive (naive.h:4, naive.c:78-95): coeffs [bw-m] -> result
 * [2bw]; pml_table from GeneratePmlTable(bw, m). */
void InvDLTNaive(double* coeffs, const int bw, const int m, double* result, double* pml_table);

/* Replaces DLTNaive (naive.h:6, naive.c:34-59): data [2bw] -> result [bw-m];
 * weights from GenerateWeightsForDLT (the even half); workspace unused. */
void DLTNaive(double* data, const int bw, const int m, double* weights, double* result, double* pml_table,
              double* workspace);

#ifdef __cplusplus
}
#endif
#endif
