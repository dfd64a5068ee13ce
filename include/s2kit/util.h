/* Drop-in for S2kit's include/s2kit/util.h (util.h:10-17).
 * Implemented by libs2kit_amd.so. */
#ifndef S2KIT_AMD_UTIL_H
#define S2KIT_AMD_UTIL_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sample format of the FST functions (util.h:10-13). */
typedef enum {
    COMPLEX = 0, /* data is complex */
    REAL         /* data is real */
} DataFormat;

/* Replaces IndexOfHarmonicCoeff (util.h:15, util.c:42-49): slot of f(m,l)
 * in the bw*bw coefficient array. Host integer arithmetic. */
int IndexOfHarmonicCoeff(const int m, const int l, const int bw);

/* Replaces TransMult (util.h:17, util.c:68-103): filter (length bw, by
 * degree) times signal coefficients (bw*bw), times sqrt(4 pi/(2l+1)), with
 * the reference's imaginary part x*v - y*u (util.c:26).  Host pointers;
 * computed on the GPU. */
void TransMult(double* rdatacoeffs, double* idatacoeffs, double* rfiltercoeffs, double* ifiltercoeffs,
               double* rres, double* ires, const int bw);

#ifdef __cplusplus
}
#endif
#endif
