/* Drop-in for S2kit's include/s2kit/FST_semi_memo.h (FST_semi_memo.h:8-16).
 *
 * Same prototypes, same grid (2bw x 2bw, theta-major, split re/im) and
 * coefficient (IndexOfHarmonicCoeff) layouts.  All pointers are host memory,
 * as in the reference; the work runs on the current HIP device.  The FFTW
 * plan arguments are accepted and ignored (NULL is fine): the library links
 * no FFTW.  fftw_plan is declared exactly as FFTW 3 declares it, so this
 * header may be included with or without <fftw3.h>.
 *
 * The functions return void, as the reference's do.  Invalid arguments or a
 * device failure print a message and abort the process, unless
 * s2amd_set_abort_on_error(0) (s2kit_amd.h) switches to recording the error
 * for s2amd_last_error(). */
#ifndef S2KIT_AMD_FST_SEMI_MEMO_H
#define S2KIT_AMD_FST_SEMI_MEMO_H

#include "s2kit/util.h"

#ifndef S2KIT_AMD_FFTW_PLAN
#define S2KIT_AMD_FFTW_PLAN
typedef struct fftw_plan_s* fftw_plan;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces FSTSemiMemo (FST_semi_memo.h:8-9, FST_semi_memo.c:68-202).
 * rdata/idata [4bw^2] -> rcoeffs/icoeffs [bw^2].  Tables from
 * SemiNaive_Naive_Pml_Table / Spharmonic_Pml_Table; weights from
 * GenerateWeightsForDLT; workspace unused. */
void FSTSemiMemo(double* rdata, double* idata, double* rcoeffs, double* icoeffs, const int bw,
                 double** seminaive_naive_table, double* workspace, DataFormat data_format, const int cutoff,
                 fftw_plan* DCT_plan, fftw_plan* FFT_plan, double* weights);

/* Replaces InvFSTSemiMemo (FST_semi_memo.h:11-12, FST_semi_memo.c:228-351). */
void InvFSTSemiMemo(double* rcoeffs, double* icoeffs, double* rdata, double* idata, const int bw,
                    double** transpose_seminaive_naive_table, double* workspace, DataFormat data_format,
                    const int cutoff, fftw_plan* inv_DCT_plan, fftw_plan* inv_FFT_plan);

/* Replaces FZTSemiMemo (FST_semi_memo.h:14, FST_semi_memo.c:374-407).
 * Writes bw values to rres and ires (REAL: ires = 0; the reference zeroes
 * 2bw doubles there). */
void FZTSemiMemo(double* rdata, double* idata, double* rres, double* ires, const int bw, double* cos_pml_table,
                 double* workspace, DataFormat data_format, fftw_plan* DCT_plan, double* weights);

/* Replaces ConvOn2SphereSemiMemo (FST_semi_memo.h:16, FST_semi_memo.c:439-518):
 * REAL data, cutoff = bw.  Tables are generated on the GPU once per bw and
 * kept, where the reference rebuilds them on every call; results are the
 * same. */
void ConvOn2SphereSemiMemo(double* rdata, double* idata, double* rfilter, double* ifilter, double* rres,
                           double* ires, const int bw, double* workspace);

#ifdef __cplusplus
}
#endif
#endif
