// Internal declarations shared by the HIP translation units of libs2kit_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <exception>
#include <string>
#include <utility>
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

// Makes `dev` current for a scope and restores the caller's device on exit.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        S2_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    DeviceGuard(const DeviceGuard&) = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;
};

// Map the exception escaping an API body to an S2AMD_* code, recording the
// message for s2amd_last_error().
int api_status(const char* fn, const std::exception_ptr& ep);
template <typename F>
int run_api(const char* fn, F&& body) {
    try {
        body();
        return 0;
    } catch (...) {
        return api_status(fn, std::current_exception());
    }
}

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
    float* w_semi32 = nullptr;   // w_semi rounded to FP32
    float* sinv32 = nullptr;
    int* trow = nullptr;         // [bw][bw+1]: start of row i of order m's transposed table
    // Bluestein tables for ring lengths with a prime factor above the register
    // passes' largest radix (else blu_M = 0): M >= 2N-1 values exp(-2 pi i t/M),
    // N chirp values w_t = exp(-i pi t^2/N), and M values FFT_M(b)/M,
    // b_t = conj(w_|t|) at t mod M for |t| < N
    int blu_M = 0;
    double2* blu = nullptr;
    float2* blu32 = nullptr;
};

// Zero-padded, parity-split Legendre matrices for the MFMA GEMMs of the fast
// path (s2_fast.hip) and the mixed form (s2_mixed.hip).  Entry m*2+p describes order m, degrees l = m+p+2i.
struct GemmMeta {
    int Lp;     // number of degrees in the class
    int Kf;     // seminaive: cosine terms of its last (longest) row
    long offF;  // seminaive: AF(m,p) [roundup4(Kf)][roundup16(Lp)]; mixed: NF(m), both parities
    long offI;  // seminaive: AI(m,p) [roundup4(Lp)][roundup16(bw/2)]; mixed: NI(m,p)
    int n0;     // mixed: 16-degree tiles per parity, ceil(Lp(m,0)/16)
    int ki;     // mixed: rows of NI(m,p), Lp padded to whole 16-degree chunks
};

struct GemmTables {
    int bw = 0, prec = 0;  // prec 0: FP64, 1: FP32 (seminaive); 2: FP32 naive (mixed form); 3: split BF16 naive
    int cut = 0;           // prec 0, 1: orders m >= cut are naive (sampled P_l^m tables)
    bool gen = false;      // prec 0, bw not a power of two: natural-order Cb rows of the per-field path
    // tile choices from s2amd_plan_autotune (0: the built-in heuristic)
    int t_ncb = 0, t_ct = 0;  // prec 0, 1: data columns per GEMM block, column tiles per table load
    int t_gj = 0;             // prec 2: 64-ring groups per inverse GEMM block
    void* af = nullptr;
    void* ai = nullptr;
    GemmMeta* meta = nullptr;
    std::vector<GemmMeta> h_meta;
    long nf = 0, ni = 0;
};

enum class Fmt : int { Complex = 0, Real = 1 };

// Optional per-stage timing with HIP events recorded on the launch stream.
enum Stage : int {
    ST_PHI_FWD = 0,
    ST_DCT_FWD,
    ST_LEG_FWD,
    ST_NAIVE_FWD,
    ST_LEG_INV,
    ST_DCT_INV,
    ST_NAIVE_INV,
    ST_PHI_INV,
    ST_ZONAL,
    ST_MULT,
    kStages
};
const char* stage_name(int st);

struct StageTimer {
    bool on = false;
    struct Rec {
        int st;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double ms[kStages] = {};
    long count[kStages] = {};
    hipEvent_t take();
    void collect();  // waits for recorded events, accumulates, recycles them
    void reset();
    ~StageTimer();
};

// RAII span around one launch: records a start and a stop event on `s`.
struct Span {
    StageTimer* t;
    int st;
    hipStream_t s;
    hipEvent_t a = nullptr;
    Span(StageTimer* t_, int st_, hipStream_t s_);
    ~Span();
};

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
                  double* cim, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm = nullptr);
void inverse_fp64(const DevConsts& c, const DevTables& t, const double* cre, const double* cim, double* re,
                  double* im, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm = nullptr);
// FZT: zonal coefficients [F][bw] of grids [F][N][N].
void zonal_fp64(const DevConsts& c, const DevTables& t, const double* re, const double* im, double* zr, double* zi,
                int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm = nullptr);
// TransMult over F fields; the filter of field f is hr/hi + f*fstride.
void trans_mult(const DevConsts& c, const double* dr, const double* di, const double* hr, const double* hi,
                long fstride, double* rr, double* ri, int F, hipStream_t s, StageTimer* tm = nullptr);

// ---- one order (s2_kernels.hip): DLTSemi, InvDLTSemi, DLTNaive, InvDLTNaive on device arrays ----
void dlt_semi_one(const double* data, const double* w, const double* tab, double* out, int B, int m, hipStream_t s);
void inv_dlt_semi_one(const double* coeffs, const double* ttab, const double* sinv, double* out, int B, int m,
                      hipStream_t s);
void dlt_naive_one(const double* data, const double* w, const double* pml, double* out, int B, int m, hipStream_t s);
void inv_dlt_naive_one(const double* coeffs, const double* pml, double* out, int B, int m, hipStream_t s);

// FP64 pipeline with the MFMA GEMMs of s2_fast.hip for both Legendre
// branches: the per-field mixed-radix FFTs and DCTs above, for bandwidths
// that are not powers of two (g.gen).
void forward_fp64_gemm(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                       double* cim, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm);
void inverse_fp64_gemm(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                       double* im, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm);

// ---- GEMM path (s2_fast.hip): MFMA Legendre GEMMs at any cutoff --------
// Power-of-two bw in [32, 1024]: register FFTs, FP64 or FP32.  Other bw in
// [16, 1024]: FP64, on the per-field FFT/DCT kernels.
bool gemm_pow2(int bw);
bool fast_supported(int bw, int cutoff, int prec);
void gemm_fwd_launch(const DevConsts& c, const GemmTables& g, const void* Cb, const void* X, double* cre, double* cim,
                     int F, Fmt fmt, hipStream_t s);
void gemm_inv_launch(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, void* Cb, void* X,
                     int F, Fmt fmt, hipStream_t s);
// Tile shapes the GEMMs accept at bw: (columns per block, column tiles per
// table load), the heuristic's choice first.
std::vector<std::pair<int, int>> gemm_tile_choices(int bw, int prec);
// 64-ring groups per mixed inverse GEMM block that bw admits, the default first.
std::vector<int> mix_gj_choices(int bw);
void gemm_tables_create(GemmTables& g, const DevTables& t, int prec, hipStream_t s);
void gemm_tables_destroy(GemmTables& g);
size_t fast_x_bytes(int bw, int prec);   // scratch X per field
size_t fast_cb_bytes(int bw, int prec);  // scratch Cb per field
void forward_fast(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                  double* cim, int F, Fmt fmt, void* X, void* Cb, hipStream_t s, StageTimer* tm);
void inverse_fast(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                  double* im, int F, Fmt fmt, void* X, void* Cb, hipStream_t s, StageTimer* tm);

// ---- mixed form (s2_mixed.hip): FP32 ring FFTs + naive DLT on FP32 MFMA ----
// FP32 ring FFTs on the mixed-radix kernels (ring lengths that are not powers of two)
void ring_fft_fwd_f32(const DevConsts& c, const double* re, const double* im, float2* X, int F, hipStream_t s);
void ring_fft_inv_f32(const DevConsts& c, const float2* X, double* re, double* im, int F, hipStream_t s);
bool mixed_supported(int bw);  // bw in [32, 512]
bool split_supported(int bw);  // power-of-two bw in [32, 512]
void naive_tables_create(GemmTables& g, const DevConsts& c, hipStream_t s);
void forward_mixed(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                   double* cim, int F, Fmt fmt, void* X, hipStream_t s, StageTimer* tm);
void inverse_mixed(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                   double* im, int F, Fmt fmt, void* X, hipStream_t s, StageTimer* tm);

// ---- split-BF16 form (s2_split.hip): the mixed form's pipeline with the
// Legendre GEMMs on BF16 MFMA, three-piece splits and six products ----------
void split_tables_create(GemmTables& g, const DevConsts& c, hipStream_t s);
void split_fwd_gemm(const DevConsts& c, const GemmTables& g, const float2* X, double* cre, double* cim, int F,
                    Fmt fmt, hipStream_t s);
void split_inv_gemm(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, float2* X, int F,
                    Fmt fmt, hipStream_t s);

}  // namespace s2
