/* Drop-in for S2kit's include/s2kit/weights.h (weights.h:4). */
#ifndef S2KIT_AMD_WEIGHTS_H
#define S2KIT_AMD_WEIGHTS_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces GenerateWeightsForDLT (weights.h:4, weights.c:32-46): Driscoll-
 * Healy weights, even w[0..2bw) then odd w[2bw..4bw); host output buffer,
 * computed on the GPU in FP64. */
void GenerateWeightsForDLT(const int bw, double* weights);

#ifdef __cplusplus
}
#endif
#endif
