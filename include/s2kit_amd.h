/* s2kit_amd.h -- batched device API of libs2kit_amd.so.
 *
 * No counterpart in the reference, which transforms one field per call from
 * host memory.  This is the throughput path: a plan per bandwidth holds the
 * FP64 tables (generated on the GPU), and each call transforms many fields
 * whose grids and coefficients already live in device memory.  Per field,
 * layouts are exactly those of FSTSemiMemo / InvFSTSemiMemo
 * (FST_semi_memo.c:28-54):
 *   grid         re[f*4bw^2 + j*2bw + k], im likewise   (theta_j major, phi_k)
 *   coefficients re[f*bw^2 + IndexOfHarmonicCoeff(m,l,bw)], im likewise
 *
 * Every function returns 0 on success or a negative S2AMD_E* code, with a
 * message for s2amd_last_error() (per thread).  `stream` is a hipStream_t
 * (NULL = the default stream); calls are asynchronous on it.  Calls that
 * share a plan must not run concurrently (the plan owns the scratch).
 */
#ifndef S2KIT_AMD_H
#define S2KIT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    S2AMD_OK = 0,
    S2AMD_EINVAL = -1,  /* bad argument */
    S2AMD_EDEVICE = -2, /* HIP runtime error / no device */
    S2AMD_ENOMEM = -3   /* device allocation failed */
};

/* Arithmetic of the transforms (inputs and outputs are FP64 in every form). */
typedef enum {
    /* FP64 throughout (seminaive: DCT + cosine-series tables, FP64 MFMA).
     * Matches the reference to ~1e-13. */
    S2AMD_FP64 = 0,
    /* seminaive in FP32 (FFTs, DCTs, FP32-MFMA projections). Measured error
     * 1e-6 .. 4e-5 relative: the cosine-space sums cancel.  bw a power of two
     * in [32, 1024]. */
    S2AMD_FP32 = 1,
    /* mixed: FP32 ring FFTs and the naive Legendre transform (naive.c) as an
     * FP32-MFMA GEMM over m on tables from the FP64 recurrence; FP64 in/out.
     * Error ~4e-7 relative.  Any bw in [32, 512] (grids that are not powers
     * of two on FP32 mixed-radix FFTs). */
    S2AMD_MIXED = 2,
    /* split BF16: the mixed form's pipeline with the Legendre GEMMs on BF16
     * MFMA; each FP32 operand is split into three BF16 pieces and six piece
     * products are accumulated in FP32, which reproduces the FP32 GEMM.
     * Error ~4e-7 relative.  bw a power of two in [32, 512]. */
    S2AMD_SPLIT_BF16 = 3
} s2amd_precision;

typedef struct s2amd_plan s2amd_plan;

/* Plan for bandwidth bw (2..1024), up to max_fields fields per call,
 * seminaive for orders < cutoff and naive (naive.c) for orders >= cutoff
 * (cutoff = bw: all seminaive).  Generates the forward and transposed tables
 * on the current device.  FP64 plans with bw in [16, 1024] and FP32 plans on
 * power-of-two grids with bw in [32, 1024] run both branches as MFMA GEMMs at
 * any cutoff (FP64 grids that are not powers of two on mixed-radix FFTs);
 * the mixed and split-BF16 forms use the naive transform for every order
 * whatever the cutoff. */
int s2amd_plan_create(s2amd_plan** plan, int bw, int cutoff, int max_fields, int precision);
/* Waits for the device (work queued on any stream may use the plan), then frees. */
int s2amd_plan_destroy(s2amd_plan* plan);
/* Autotuning (no counterpart in the reference): times forward+inverse of
 * nfields fields on stream for each tile shape the plan's GEMMs accept
 * (FP64/FP32: data columns per block and column tiles per table load; mixed:
 * 64-ring groups per inverse block) and keeps the fastest for later calls.
 * Every shape computes the same sums in the same order: results do not
 * change.  Allocates and frees temporary device fields; synchronises the
 * stream.  Plans whose calls of nfields fields take no GEMM (or split-BF16
 * plans) keep their defaults.  s2amd_plan_tuning reports the choice (0: the
 * built-in heuristic). */
int s2amd_plan_autotune(s2amd_plan* plan, int nfields, void* stream);
int s2amd_plan_tuning(const s2amd_plan* plan, int* cols, int* col_tiles, int* ring_groups);
/* Device bytes held by the plan (tables + scratch). */
size_t s2amd_plan_bytes(const s2amd_plan* plan);
/* Device pointers to the plan's tables (layout of SemiNaive_Naive_Pml_Table /
 * Transpose_..., order m at offset off[m]), for inspection. */
int s2amd_plan_tables(const s2amd_plan* plan, const double** fwd, const double** inv, long* ndoubles);

/* nfields grids -> coefficients (FSTSemiMemo per field). data_format: 0 COMPLEX, 1 REAL. */
int s2amd_forward(s2amd_plan* plan, const double* re, const double* im, double* cre, double* cim, int nfields,
                  int data_format, void* stream);
/* nfields coefficient sets -> grids (InvFSTSemiMemo per field). */
int s2amd_inverse(s2amd_plan* plan, const double* cre, const double* cim, double* re, double* im, int nfields,
                  int data_format, void* stream);
/* s2amd_forward for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, kernels and results
 * (bit for bit) as s2amd_forward on the same plan; the transform's scratch
 * (the order-major intermediate and, for FP64/FP32 plans, the cosine-row
 * buffer) lives in one static device array of S2AMD_NOALLOC_SCRATCH_BYTES
 * instead of the plan's buffers.  The library holds that array on every
 * device it runs on until the process ends.  A call needs per field
 *   mixed / split-BF16 plans:  8 * (2bw)^2 bytes
 *   FP64 / FP32 plans:        16 * (2bw)^2 + 16 * 2bw * bw bytes (+256)
 * (bw=256: 2 MiB / 6 MiB per field, so up to 64 / 21 fields) and returns
 * S2AMD_EINVAL, having launched nothing, when nfields fields need more.  It
 * calls no allocating routine (the first call on a device only looks up the
 * array's address).  Every call on a device shares the array, so calls that
 * overlap on one device (different streams, or host threads) race: serialise
 * them.  Nothing in the array needs to start at zero: each call writes every
 * scratch value before reading it.  The plan's stage timing is not recorded
 * for these calls. */
#define S2AMD_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int s2amd_forward_noalloc(s2amd_plan* plan, const double* re, const double* im, double* cre, double* cim,
                          int nfields, int data_format, void* stream);
/* Zonal coefficients zr/zi [nfields][bw] (FZTSemiMemo per field). */
int s2amd_zonal(s2amd_plan* plan, const double* re, const double* im, double* zr, double* zi, int nfields,
                int data_format, void* stream);
/* ConvOn2SphereSemiMemo per field: signal s, filter h (grids), output o. */
int s2amd_conv(s2amd_plan* plan, const double* sre, const double* sim, const double* hre, const double* him,
               double* ore, double* oim, int nfields, void* stream);

/* ---- convolution batches of uneven bandwidth --------------------------------
 * Many ConvOn2SphereSemiMemo requests, each at its own bandwidth, on device
 * grids.  _create groups the requests by bandwidth (one plan per distinct
 * bandwidth, max_fields = the group's size), and records the whole batch --
 * gathering each group's grids into contiguous staging buffers (skipped for
 * an array whose requests already lie back to back in request order), the
 * group's forward / zonal / TransMult / inverse pipeline, scattering the
 * outputs -- as one HIP graph; _run replays it on a stream.  Requests are
 * copied at creation: their buffers must stay allocated while the batch is
 * used; their contents are read at every _run.  precision: S2AMD_FP64,
 * S2AMD_MIXED or S2AMD_SPLIT_BF16; a group whose bandwidth the chosen form does
 * not support (mixed: outside [32, 512]; split-BF16: not a power of two in
 * [32, 512]) runs in FP64.  No counterpart in the reference,
 * which convolves one pair of host grids per call (FST_semi_memo.c:439-518). */
typedef struct {
    int bw;
    const double* sre; /* signal grid [2bw][2bw], re and im */
    const double* sim;
    const double* hre; /* filter grid */
    const double* him;
    double* ore; /* output grid */
    double* oim;
} s2amd_conv_request;

typedef struct s2amd_conv_batch s2amd_conv_batch;
/* Groups the requests by bandwidth, creates and autotunes (s2amd_plan_autotune)
 * one plan per group, and records the whole batch as one HIP graph. */
int s2amd_conv_batch_create(s2amd_conv_batch** batch, const s2amd_conv_request* req, int n, int precision,
                            void* stream);
int s2amd_conv_batch_run(s2amd_conv_batch* batch, void* stream);
/* Groups in ascending bandwidth: their bandwidths, request counts and the
 * precision each runs in (arrays of at least *ngroups entries, or NULL). */
int s2amd_conv_batch_info(const s2amd_conv_batch* batch, int* ngroups, int* bws, int* counts, int* precisions);
int s2amd_conv_batch_destroy(s2amd_conv_batch* batch);

/* Per-stage kernel timing: when enabled, every launch of the plan is
 * bracketed by HIP events recorded on the call's stream.  stage_times waits
 * for the recorded events, returns the summed milliseconds and launch counts
 * per stage (arrays of S2AMD_NSTAGES), and clears them. */
#define S2AMD_NSTAGES 10
int s2amd_plan_timing(s2amd_plan* plan, int enable);
int s2amd_plan_stage_times(s2amd_plan* plan, double* ms, long* counts);
const char* s2amd_stage_name(int stage);

/* Drop-in error policy: 1 (default) print and abort, 0 record only. */
void s2amd_set_abort_on_error(int on);
const char* s2amd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
