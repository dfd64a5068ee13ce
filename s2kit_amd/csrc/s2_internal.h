// Internal declarations shared by the HIP translation units of libs2kit_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "s2_layout.h"

namespace s2 {

// ---- error handling -------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

struct HipError {
    hipError_t code;
    std::string where;
};

#define S2_HIP(call)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) throw ::s2::HipError{e_, std::string(#call) + " @ " + __FILE__}; \
    } while (0)

struct ArgError {
    std::string msg;
};

// ---- device tables of one (bw, cutoff) -------------------------------------
// Order m's forward table starts at fwd + off[m], its transposed table at
// inv + off[m].  Seminaive orders (m < cutoff) hold the zero-stripped
// cosine-series tables; naive orders hold (bw - m) rows of 2*bw samples.
struct DevTables {
    int bw = 0, cutoff = 0;
    double* fwd = nullptr;
    double* inv = nullptr;
    long* off = nullptr;       // device copy of offsets (bw entries)
    std::vector<long> h_off;   // host copy
    long total = 0;            // doubles per direction
    bool owned = true;
};

// Per-(bw) constant device data: weights, twiddles, phases, sines, degrees.
struct DevConsts {
    int bw = 0, N = 0, logN = 0;
    bool pow2 = false;
    double* w_semi = nullptr;    // 2N: even then odd weights, times sqrt(2pi)/N / sqrt(2N)
    double* w_naive = nullptr;   // N: even weights times sqrt(2pi)/N
    double2* tw = nullptr;       // N: exp(-2 pi i t / N)
    double2* ph = nullptr;       // N: exp(-i pi k / (2N))
    double* sinv = nullptr;      // N: sin(theta_j) on the 2B grid
    int* deg = nullptr;          // bw*bw: degree of each coefficient slot
    float2* tw32 = nullptr;
    float2* ph32 = nullptr;
};

enum class Fmt : int { Complex = 0, Real = 1 };

// ---- kernels' host launchers (s2_kernels.hip) -----------------------------
void consts_create(DevConsts& c, int bw, const double* h_weights /* 4*bw or null: compute on GPU */,
                   hipStream_t s);
void consts_destroy(DevConsts& c);

void tables_generate(DevTables& t, int bw, int cutoff, bool want_fwd, bool want_inv, hipStream_t s);
void tables_destroy(DevTables& t);
// GenerateCosPmlTable for one order into a device buffer of table_size(m,bw).
void gen_cospml_one(int bw, int m, double* d_out, hipStream_t s);
// TransposeCosPmlTable for one order, device to device.
void transpose_cospml_one(int bw, int m, const double* d_in, double* d_out, hipStream_t s);
// GeneratePmlTable for one order into a device buffer of 2*bw*(bw-m).
void gen_pml_one(int bw, int m, double* d_out, hipStream_t s);
// GenerateWeightsForDLT into a device buffer of 4*bw.
void gen_weights(int bw, double* d_w, hipStream_t s);

// FP64 pipeline over F fields, device pointers, field-major layouts:
//   grid  re/im [F][N][N], coefficients re/im [F][bw*bw].
// X: scratch [F][N][N] double2, Cb: scratch [F][N][bw] double2.
void forward_fp64(const DevConsts& c, const DevTables& t, const double* re, const double* im, double* cre,
                  double* cim, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s);
void inverse_fp64(const DevConsts& c, const DevTables& t, const double* cre, const double* cim, double* re,
                  double* im, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s);
// FZT: zonal coefficients [F][bw] of grids [F][N][N].
void zonal_fp64(const DevConsts& c, const DevTables& t, const double* re, const double* im, double* zr, double* zi,
                int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s);
// TransMult over F fields; the filter of field f is hr/hi + f*fstride.
void trans_mult(const DevConsts& c, const double* dr, const double* di, const double* hr, const double* hi,
                long fstride, double* rr, double* ri, int F, hipStream_t s);

}  // namespace s2
