/* Drop-in for S2kit's include/s2kit/cospml.h (cospml.h:6-30).
 * Table layouts are the reference's, byte for byte: callers build tables
 * with these functions and hand them to FSTSemiMemo / InvFSTSemiMemo. */
#ifndef S2KIT_AMD_COSPML_H
#define S2KIT_AMD_COSPML_H

#ifdef __cplusplus
extern "C" {
#endif

/* Layout arithmetic (host integer code), same results as the reference: */
int TableSize(const int m, const int bw);                          /* cospml.c:39-59   */
int Spharmonic_TableSize(const int bw);                            /* cospml.c:85-92   */
int Reduced_SpharmonicTableSize(const int bw, const int m);        /* cospml.c:107-115 */
int Reduced_Naive_TableSize(const int bw, const int m);            /* cospml.c:434-440 */
int TableOffset(int m, int l);                                     /* cospml.c:123-134 */
int RowSize(const int m, const int l);                             /* cospml.c:250-258 */
int Transpose_RowSize(const int row, const int m, const int bw);   /* cospml.c:270-288 */

/* Table producers.  Host output buffers; the values are computed on the GPU
 * in FP64 and the device copy is kept for later transforms.  workspace
 * arguments are accepted and unused.  Returned pointer arrays are malloc'd
 * (bw entries; bw+1 for the mixed tables) and freed by the caller. */
void GenerateCosPmlTable(const int bw, const int m, double* tablespace, double* workspace);   /* :161-242 */
void TransposeCosPmlTable(const int bw, const int m, double* cos_pml_table, double* result);  /* :301-362 */
double** Spharmonic_Pml_Table(const int bw, double* resultspace, double* workspace);          /* :387-401 */
double** Transpose_Spharmonic_Pml_Table(double** spharmonic_pml_table, const int bw,
                                        double* resultspace);                                  /* :415-428 */
double** SemiNaive_Naive_Pml_Table(const int bw, const int m, double* resultspace,
                                   double* workspace);                                         /* :451-475 */
double** Transpose_SemiNaive_Naive_Pml_Table(double** seminaive_naive_pml_table, const int bw, const int m,
                                             double* resultspace, double* workspace);          /* :490-518 */

#ifdef __cplusplus
}
#endif
#endif
