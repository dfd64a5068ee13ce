/* Drop-in for S2kit's include/s2kit/pml.h (pml.h:4). */
#ifndef S2KIT_AMD_PML_H
#define S2KIT_AMD_PML_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces GeneratePmlTable (pml.h:4, pml.c:41-79): L2-normalised P_l^m,
 * l = m..bw-1, at the 2bw grid latitudes, (bw-m) rows of 2bw; host output,
 * computed on the GPU in FP64.  workspace is accepted and unused. */
void GeneratePmlTable(const int bw, const int m, double* pml_table, double* workspace);

#ifdef __cplusplus
}
#endif
#endif
