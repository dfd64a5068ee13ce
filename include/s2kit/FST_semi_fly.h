/* Drop-in for S2kit's include/s2kit/FST_semi_fly.h (FST_semi_fly.h:8-16).
 *
 * The reference's "on the fly" variants compute the associated Legendre
 * tables inside every call instead of taking them from the caller
 * (FST_semi_fly.c:96-187, 259-331, 416); the transforms are otherwise
 * those of FST_semi_memo.h.  Here the tables are generated on the GPU (FP64,
 * the same recurrence and cosine transform as SemiNaive_Naive_Pml_Table) on
 * the first call for a (bw, cutoff) and kept, so a call costs what the
 * memoised drop-in costs; results are the same.  Prototypes, layouts, the
 * ignored FFTW plans and workspace, and the error behaviour are those of
 * FST_semi_memo.h. */
#ifndef S2KIT_AMD_FST_SEMI_FLY_H
#define S2KIT_AMD_FST_SEMI_FLY_H

#include "s2kit/util.h"

#ifndef S2KIT_AMD_FFTW_PLAN
#define S2KIT_AMD_FFTW_PLAN
typedef struct fftw_plan_s* fftw_plan;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces FSTSemiFly (FST_semi_fly.h:8-9, FST_semi_fly.c:67-211).
 * rdata/idata [4bw^2] -> rcoeffs/icoeffs [bw^2]; weights from
 * GenerateWeightsForDLT. */
void FSTSemiFly(double* rdata, double* idata, double* rcoeffs, double* icoeffs, const int bw, double* workspace,
                DataFormat data_format, const int cutoff, fftw_plan* DCT_plan, fftw_plan* FFT_plan, double* weights);

/* Replaces InvFSTSemiFly (FST_semi_fly.h:11-12, FST_semi_fly.c:235-369). */
void InvFSTSemiFly(double* rcoeffs, double* icoeffs, double* rdata, double* idata, const int bw, double* workspace,
                   DataFormat data_format, const int cutoff, fftw_plan* inv_DCT_plan, fftw_plan* inv_FFT_plan);

/* Replaces FZTSemiFly (FST_semi_fly.h:14, FST_semi_fly.c:390-425).  Writes bw
 * values to rres and ires (REAL: ires = 0; the reference zeroes 2bw doubles
 * there, FST_semi_fly.c:424). */
void FZTSemiFly(double* rdata, double* idata, double* rres, double* ires, const int bw, double* workspace,
                DataFormat data_format, fftw_plan* DCT_plan, double* weights);

/* Replaces ConvOn2SphereSemiFly (FST_semi_fly.h:16, FST_semi_fly.c:454-519):
 * REAL data, seminaive for all orders. */
void ConvOn2SphereSemiFly(double* rdata, double* idata, double* rfilter, double* ifilter, double* rres,
                          double* ires, const int bw, double* workspace);

#ifdef __cplusplus
}
#endif

#endif
