// C-ABI of libs2kit_amd.so: S2kit drop-ins (host pointers) and the batched
// device API (s2kit_amd.h).
//
// Drop-in calls run on the current HIP device, one at a time (a process-wide
// mutex), on a stream owned by the library.  Caller tables are copied to the
// device once and cached: the cache key is (bw, cutoff, direction, the
// caller's per-order pointers) plus a hash of every entry of every order,
// recomputed on every call, so a table edited in place (any entry) or rebuilt
// in the same memory is uploaded again, as the reference reads the caller's
// table on every call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "s2_internal.h"

extern "C" {
#include "s2kit/FST_semi_fly.h"
#include "s2kit/FST_semi_memo.h"
#include "s2kit/naive.h"
#include "s2kit/seminaive.h"
#include "s2kit/cospml.h"
#include "s2kit/pml.h"
#include "s2kit/util.h"
#include "s2kit/weights.h"
#include "s2kit_amd.h"
}

namespace s2 {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static int g_abort_on_error = 1;

void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

static void dropin_fail(const char* fn, const std::string& msg) {
    set_error(std::string(fn) + ": " + msg);
    std::fprintf(stderr, "libs2kit_amd: %s: %s\n", fn, msg.c_str());
    if (g_abort_on_error) std::abort();
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    void reserve(size_t bytes) {
        if (bytes <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        S2_HIP(hipMalloc(&p, bytes));
        n = bytes;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// ---------------------------------------------------------------------------
// drop-in context
// ---------------------------------------------------------------------------
struct CachedTables {
    int bw, cutoff, dir;  // dir 0 forward, 1 transposed, 2 order-0 only (FZT)
    int device;
    std::vector<const double*> ptrs;
    std::vector<unsigned long long> fp;
    DevTables t;
};

struct CachedConsts {
    int bw, device;
    std::vector<double> weights;  // caller weights the prescaled copy was made from (empty: GPU weights)
    DevConsts c;
};

struct Ctx {
    std::mutex mu;
    int device = -1;
    hipStream_t stream = nullptr;
    DBuf re, im, cre, cim, X, Cb, aux1, aux2, aux3, aux4;
    std::list<CachedTables> tables;  // most recent first
    std::list<CachedConsts> consts;
    size_t table_budget = (size_t)8 << 30;
};

static Ctx& ctx() {
    static Ctx* c = new Ctx();  // never destroyed: HIP may be torn down first at exit
    return *c;
}

static void ensure_device(Ctx& c) {
    int dev = 0;
    S2_HIP(hipGetDevice(&dev));
    if (c.stream && c.device == dev) return;
    if (c.stream) {
        // the current device changed: drop the previous device's state and
        // stream, with that device current, then come back
        DeviceGuard g(c.device);
        for (auto& e : c.tables) tables_destroy(e.t);
        for (auto& e : c.consts) consts_destroy(e.c);
        c.tables.clear();
        c.consts.clear();
        for (DBuf* b : {&c.re, &c.im, &c.cre, &c.cim, &c.X, &c.Cb, &c.aux1, &c.aux2, &c.aux3, &c.aux4})
            b->release();
        (void)hipStreamDestroy(c.stream);
        c.stream = nullptr;
    }
    c.device = dev;
    S2_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
}

// Hash of all n entries of a table.  Four interleaved lanes, each a chain of
// s <- (s ^ word) * odd constant: every step is a bijection of the lane state,
// so changing any single entry always changes its lane's final state, and the
// lanes are folded the same way, so it always changes the hash.
static unsigned long long table_hash(const double* p, long n) {
    const unsigned long long K = 0x9E3779B97F4A7C15ULL;  // odd: multiplication is a bijection
    unsigned long long h[4] = {0x243F6A8885A308D3ULL, 0x13198A2E03707344ULL, 0xA4093822299F31D0ULL,
                               0x082EFA98EC4E6C89ULL};
    long i = 0;
    for (; i + 4 <= n; i += 4) {
        unsigned long long v[4];
        std::memcpy(v, p + i, 32);
        for (int k = 0; k < 4; ++k) h[k] = (h[k] ^ v[k]) * K;
    }
    for (; i < n; ++i) {
        unsigned long long v;
        std::memcpy(&v, p + i, 8);
        h[0] = (h[0] ^ v) * K;
    }
    unsigned long long r = (unsigned long long)n * K;
    for (int k = 0; k < 4; ++k) r = (r ^ h[k]) * K;
    return r ^ (r >> 29);
}

static long order_len(int m, int bw, int cutoff) {
    return (m < cutoff) ? table_size(m, bw) : 2L * bw * (bw - m);
}

// Hashes of orders 0..norders-1 of a caller table set; large sets are hashed
// on several host threads (orders split into contiguous ranges).
static std::vector<unsigned long long> table_hashes(const double* const* ptrs, int norders, int bw, int cutoff) {
    std::vector<unsigned long long> h(norders);
    long total = 0;
    for (int m = 0; m < norders; ++m) total += order_len(m, bw, cutoff);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<long>({(long)std::min(hw, 8u), (long)norders, std::max(1L, total >> 21)});
    auto work = [&](int t) {
        for (int m = t; m < norders; m += nt) h[m] = table_hash(ptrs[m], order_len(m, bw, cutoff));
    };
    if (nt <= 1) {
        work(0);
        return h;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    return h;
}

// Device bytes an entry holds (direction 3 entries hold both directions).
static size_t entry_bytes(const CachedTables& e) {
    return sizeof(double) * (size_t)e.t.total * (e.dir == 3 ? 2 : 1);
}

// Every insertion into the table cache: drops an entry of the same key (bw,
// cutoff, direction, device, caller pointers), then evicts least recently
// used entries until the new one fits the budget.  Callers insert only with
// the library stream idle (every drop-in ends with a synchronisation), so no
// queued work reads an evicted table.
static const DevTables& cache_insert(Ctx& c, CachedTables&& e) {
    for (auto it = c.tables.begin(); it != c.tables.end();) {
        if (it->bw == e.bw && it->cutoff == e.cutoff && it->dir == e.dir && it->device == e.device &&
            it->ptrs == e.ptrs) {
            tables_destroy(it->t);
            it = c.tables.erase(it);
        } else {
            ++it;
        }
    }
    size_t used = entry_bytes(e);
    for (const auto& x : c.tables) used += entry_bytes(x);
    while (!c.tables.empty() && used > c.table_budget) {
        used -= entry_bytes(c.tables.back());
        tables_destroy(c.tables.back().t);
        c.tables.pop_back();
    }
    c.tables.push_front(std::move(e));
    return c.tables.front().t;
}

// Device copy of a caller's table set (one direction).  norders = bw, or 1
// for the order-0 table of FZTSemiMemo.
static const DevTables& device_tables(Ctx& c, int bw, int cutoff, int dir, const double* const* ptrs,
                                      int norders) {
    for (int m = 0; m < norders; ++m)
        if (!ptrs[m]) throw ArgError{"table pointer for order " + std::to_string(m) + " is NULL"};
    const std::vector<unsigned long long> fp = table_hashes(ptrs, norders, bw, cutoff);
    for (auto it = c.tables.begin(); it != c.tables.end(); ++it) {
        if (it->bw != bw || it->cutoff != cutoff || it->dir != dir || it->device != c.device) continue;
        if ((int)it->ptrs.size() != norders || !std::equal(it->ptrs.begin(), it->ptrs.end(), ptrs)) continue;
        if (it->fp != fp) continue;
        c.tables.splice(c.tables.begin(), c.tables, it);
        return c.tables.front().t;
    }
    CachedTables e;
    e.bw = bw;
    e.cutoff = cutoff;
    e.dir = dir;
    e.device = c.device;
    e.ptrs.assign(ptrs, ptrs + norders);
    e.fp = fp;
    e.t.bw = bw;
    e.t.cutoff = cutoff;
    e.t.h_off.assign(bw, 0);
    long pos = 0;
    for (int m = 0; m < bw; ++m) {
        e.t.h_off[m] = pos;
        if (m < norders) pos += order_len(m, bw, cutoff);
    }
    e.t.total = pos;
    double* d = nullptr;
    S2_HIP(hipMalloc(&d, sizeof(double) * std::max(1L, pos)));
    for (int m = 0; m < norders; ++m)
        S2_HIP(hipMemcpyAsync(d + e.t.h_off[m], ptrs[m], sizeof(double) * order_len(m, bw, cutoff),
                              hipMemcpyHostToDevice, c.stream));
    S2_HIP(hipMalloc(&e.t.off, sizeof(long) * bw));
    S2_HIP(hipMemcpyAsync(e.t.off, e.t.h_off.data(), sizeof(long) * bw, hipMemcpyHostToDevice, c.stream));
    S2_HIP(hipStreamSynchronize(c.stream));
    if (dir == 1)
        e.t.inv = d;
    else
        e.t.fwd = d;
    e.t.owned = true;
    return cache_insert(c, std::move(e));
}

// Register a device table set that this library produced for the caller's
// host copy (so the first transform finds it without an upload).
static void register_tables(Ctx& c, DevTables&& t, int dir, double** ptrs, int norders) {
    CachedTables e;
    e.bw = t.bw;
    e.cutoff = t.cutoff;
    e.dir = dir;
    e.device = c.device;
    e.ptrs.assign(ptrs, ptrs + norders);
    e.fp = table_hashes(ptrs, norders, t.bw, t.cutoff);
    if (dir == 1 && t.fwd) {
        (void)hipFree(t.fwd);
        t.fwd = nullptr;
    }
    if (dir == 0 && t.inv) {
        (void)hipFree(t.inv);
        t.inv = nullptr;
    }
    e.t = std::move(t);
    (void)cache_insert(c, std::move(e));
}

// Prescaled weights and twiddles for bw; `weights` (4bw, host) if the caller
// supplied them, else the GPU-generated ones.
static const DevConsts& device_consts(Ctx& c, int bw, const double* weights) {
    for (auto it = c.consts.begin(); it != c.consts.end(); ++it) {
        if (it->bw != bw || it->device != c.device) continue;
        bool same = weights ? (it->weights.size() == (size_t)4 * bw &&
                               std::memcmp(it->weights.data(), weights, sizeof(double) * 4 * bw) == 0)
                            : it->weights.empty();
        if (!same) continue;
        c.consts.splice(c.consts.begin(), c.consts, it);
        return c.consts.front().c;
    }
    CachedConsts e;
    e.bw = bw;
    e.device = c.device;
    if (weights) e.weights.assign(weights, weights + 4 * bw);
    consts_create(e.c, bw, weights, c.stream);
    c.consts.push_front(std::move(e));
    while (c.consts.size() > 8) {
        consts_destroy(c.consts.back().c);
        c.consts.pop_back();
    }
    return c.consts.front().c;
}

// The caller's DataFormat as an int, read through its object representation:
// C callers may pass any int, and reading a C++ enum parameter outside the
// enum's values is undefined, so it is checked before it is used as one.
static int fmt_of(const DataFormat& d) {
    static_assert(sizeof(DataFormat) == sizeof(int), "DataFormat is an int-sized enum");
    int v;
    std::memcpy(&v, &d, sizeof v);
    return v;
}

static void check_common(int bw, int fmt) {
    if (bw < 2 || bw > 1024) throw ArgError{"bw must be in [2, 1024], got " + std::to_string(bw)};
    if (fmt != 0 && fmt != 1) throw ArgError{"data_format must be COMPLEX (0) or REAL (1)"};
}

static void need(const void* p, const char* what) {
    if (!p) throw ArgError{std::string(what) + " is NULL"};
}

template <typename F>
static void run_dropin(const char* fn, F&& body) {
    Ctx& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    try {
        body(c);  // validates its arguments, then calls ensure_device()
        if (c.stream) S2_HIP(hipStreamSynchronize(c.stream));
    } catch (const ArgError& e) {
        dropin_fail(fn, e.msg);
    } catch (const HipError& e) {
        dropin_fail(fn, std::string(hipGetErrorString(e.code)) + " (" + e.where + ")");
    } catch (const std::bad_alloc&) {
        dropin_fail(fn, "host allocation failed");
    }
}

// ---------------------------------------------------------------------------
// batched plans
// ---------------------------------------------------------------------------
struct Plan {
    int bw, cutoff, max_fields, precision, device;
    bool fast = false;  // MFMA GEMM path (s2_fast.hip): FP64 bw >= 16, FP32 power-of-two bw >= 32
    DevConsts c;
    DevTables t;
    GemmTables g;
    DBuf X, Cb, cr, ci, zr, zi;
    StageTimer timer;
};

// FP64 plans: the GEMM path from two fields per call, and for one field below
// bw = 512; single fields at bw >= 512 go to the per-field kernels, whose row
// GEMVs stream the tables faster than one field's GEMM tiles use them
// (measured on MI355X, us per fwd+inv call, GEMM vs per-field: bw=256 F=1
// 75 vs 88; bw=512 F=1 210 vs 205, F=2 266 vs 336; bw=1024 F=1 1389 vs 1239).
static bool use_fast(const Plan* p, int F) {
    if (!p->fast) return false;
    if (p->precision != S2AMD_FP64) return true;
    return F >= 2 || p->bw < 512;
}

// The forward pipeline of a plan on scratch X (order-major intermediate) and
// Cb (cosine rows; unused by the mixed and split-BF16 forms).
static void plan_forward_on(Plan* p, const double* re, const double* im, double* cre, double* cim, int F, Fmt fmt,
                            void* X, void* Cb, hipStream_t s, StageTimer* tm) {
    if (p->precision == S2AMD_MIXED || p->precision == S2AMD_SPLIT_BF16)
        forward_mixed(p->c, p->g, re, im, cre, cim, F, fmt, X, s, tm);
    else if (use_fast(p, F))
        forward_fast(p->c, p->g, re, im, cre, cim, F, fmt, X, Cb, s, tm);
    else
        forward_fp64(p->c, p->t, re, im, cre, cim, F, fmt, (double2*)X, (double2*)Cb, s, tm);
}

static void plan_forward(Plan* p, const double* re, const double* im, double* cre, double* cim, int F, Fmt fmt,
                         hipStream_t s) {
    plan_forward_on(p, re, im, cre, cim, F, fmt, p->X.p, p->Cb.p, s, &p->timer);
}

// ---- scratch of s2amd_forward_noalloc --------------------------------------
// One static device array, held by every device the library runs on until the
// process ends; no call allocates it.  In host code `scratch` names a
// host-side stand-in, not device memory: kernels get the array's device
// address, looked up once per device by hipGetSymbolAddress and kept in
// g_scratch_ptr.  Calls that overlap on one device share (and race on) it.
static __device__ __align__(256) unsigned char scratch[S2AMD_NOALLOC_SCRATCH_BYTES];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on the current device `dev`.
static void* scratch_on(int dev) {
    if (dev < 0 || dev >= kScratchDevices)
        throw ArgError{"device index " + std::to_string(dev) + " beyond the " + std::to_string(kScratchDevices) +
                       " devices the no-alloc scratch serves"};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        S2_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw HipError{hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch) returned NULL"};
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// Scratch bytes of a forward over F fields: the plan's own X (and Cb) sizes
// per field, Cb placed at a 256-byte boundary after X.
static size_t noalloc_need(const Plan* p, int F, size_t* cb_off) {
    const size_t N = 2 * (size_t)p->bw;
    if (p->precision == S2AMD_MIXED || p->precision == S2AMD_SPLIT_BF16) {
        *cb_off = 0;
        return sizeof(float2) * N * N * F;
    }
    const size_t x = (sizeof(double2) * N * N * F + 255) / 256 * 256;
    *cb_off = x;
    return x + sizeof(double2) * N * p->bw * F;
}

static void plan_inverse(Plan* p, const double* cre, const double* cim, double* re, double* im, int F, Fmt fmt,
                         hipStream_t s) {
    if (p->precision == S2AMD_MIXED || p->precision == S2AMD_SPLIT_BF16)
        inverse_mixed(p->c, p->g, cre, cim, re, im, F, fmt, p->X.p, s, &p->timer);
    else if (use_fast(p, F))
        inverse_fast(p->c, p->g, cre, cim, re, im, F, fmt, p->X.p, p->Cb.p, s, &p->timer);
    else
        inverse_fp64(p->c, p->t, cre, cim, re, im, F, fmt, p->X.as<double2>(), p->Cb.as<double2>(), s, &p->timer);
}

int api_status(const char* fn, const std::exception_ptr& ep) {
    try {
        std::rethrow_exception(ep);
    } catch (const ArgError& e) {
        set_error(std::string(fn) + ": " + e.msg);
        return S2AMD_EINVAL;
    } catch (const HipError& e) {
        set_error(std::string(fn) + ": " + hipGetErrorString(e.code) + " (" + e.where + ")");
        return e.code == hipErrorOutOfMemory ? S2AMD_ENOMEM : S2AMD_EDEVICE;
    } catch (const std::bad_alloc&) {
        set_error(std::string(fn) + ": host allocation failed");
        return S2AMD_ENOMEM;
    } catch (...) {
        set_error(std::string(fn) + ": unexpected error");
        return S2AMD_EDEVICE;
    }
}

template <typename F>
static int run_plan(const char* fn, F&& body) {
    try {
        body();
        return S2AMD_OK;
    } catch (...) {
        return api_status(fn, std::current_exception());
    }
}

static void plan_check_call(const Plan* p, int nfields, int fmt) {
    if (!p) throw ArgError{"plan is NULL"};
    if (nfields < 0 || nfields > p->max_fields)
        throw ArgError{"nfields must be in [0, " + std::to_string(p->max_fields) + "]"};
    if (fmt != 0 && fmt != 1) throw ArgError{"data_format must be 0 or 1"};
    int dev = -1;
    S2_HIP(hipGetDevice(&dev));
    if (dev != p->device) throw ArgError{"plan was created on another device"};
}

}  // namespace s2

using namespace s2;

// ===========================================================================
// drop-in: layout arithmetic (host)
// ===========================================================================
extern "C" int TableSize(const int m, const int bw) { return s2::table_size(m, bw); }

extern "C" int Spharmonic_TableSize(const int bw) {
    if (bw <= 512) return ((4 * bw * bw * bw) + (6 * bw * bw) - (8 * bw)) / 24 + bw;
    return Reduced_SpharmonicTableSize(bw, bw);
}

extern "C" int Reduced_SpharmonicTableSize(const int bw, const int m) {
    int s = 0;
    for (int i = 0; i < m; ++i) s += s2::table_size(i, bw);
    return s;
}

extern "C" int Reduced_Naive_TableSize(const int bw, const int m) {
    int s = 0;
    for (int i = m; i < bw; ++i) s += bw - i;
    return 2 * bw * s;
}

extern "C" int TableOffset(int m, int l) { return s2::table_offset(m, l); }
extern "C" int RowSize(const int m, const int l) { return s2::row_size(m, l); }
extern "C" int Transpose_RowSize(const int row, const int m, const int bw) {
    return s2::transpose_row_size(row, m, bw);
}
extern "C" int IndexOfHarmonicCoeff(const int m, const int l, const int bw) { return s2::index_of_coeff(m, l, bw); }

// ===========================================================================
// drop-in: tables and weights (GPU, copied back to the caller's buffers)
// ===========================================================================
extern "C" void GenerateWeightsForDLT(const int bw, double* weights) {
    run_dropin("GenerateWeightsForDLT", [&](Ctx& c) {
        check_common(bw, 0);
        need(weights, "weights");
        ensure_device(c);
        c.aux1.reserve(sizeof(double) * 4 * bw);
        gen_weights(bw, c.aux1.as<double>(), c.stream);
        S2_HIP(hipMemcpyAsync(weights, c.aux1.p, sizeof(double) * 4 * bw, hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void GenerateCosPmlTable(const int bw, const int m, double* tablespace, double* workspace) {
    (void)workspace;
    run_dropin("GenerateCosPmlTable", [&](Ctx& c) {
        check_common(bw, 0);
        if (m < 0 || m >= bw) throw ArgError{"m must be in [0, bw)"};
        need(tablespace, "tablespace");
        const long n = table_size(m, bw);
        ensure_device(c);
        c.aux1.reserve(sizeof(double) * n);
        gen_cospml_one(bw, m, c.aux1.as<double>(), c.stream);
        S2_HIP(hipMemcpyAsync(tablespace, c.aux1.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void TransposeCosPmlTable(const int bw, const int m, double* cos_pml_table, double* result) {
    run_dropin("TransposeCosPmlTable", [&](Ctx& c) {
        check_common(bw, 0);
        if (m < 0 || m >= bw) throw ArgError{"m must be in [0, bw)"};
        need(cos_pml_table, "cos_pml_table");
        need(result, "result");
        const long n = table_size(m, bw);
        ensure_device(c);
        c.aux1.reserve(sizeof(double) * n);
        c.aux2.reserve(sizeof(double) * n);
        S2_HIP(hipMemcpyAsync(c.aux1.p, cos_pml_table, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        transpose_cospml_one(bw, m, c.aux1.as<double>(), c.aux2.as<double>(), c.stream);
        S2_HIP(hipMemcpyAsync(result, c.aux2.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void GeneratePmlTable(const int bw, const int m, double* pml_table, double* workspace) {
    (void)workspace;
    run_dropin("GeneratePmlTable", [&](Ctx& c) {
        check_common(bw, 0);
        if (m < 0 || m >= bw) throw ArgError{"m must be in [0, bw)"};
        need(pml_table, "pml_table");
        const long n = 2L * bw * (bw - m);
        ensure_device(c);
        c.aux1.reserve(sizeof(double) * n);
        gen_pml_one(bw, m, c.aux1.as<double>(), c.stream);
        S2_HIP(hipMemcpyAsync(pml_table, c.aux1.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    });
}

// Generate (bw, cutoff) tables on the GPU, copy direction `dir` into the
// caller's resultspace, return its malloc'd pointer array (nptr entries),
// and keep the device copy registered.
static double** produce_tables(const char* fn, int bw, int cutoff, int dir, double* resultspace, int nptr) {
    double** out = nullptr;
    run_dropin(fn, [&](Ctx& c) {
        check_common(bw, 0);
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        need(resultspace, "resultspace");
        ensure_device(c);
        DevTables t;
        tables_generate(t, bw, cutoff, dir == 0, dir == 1, c.stream);
        const double* src = dir == 0 ? t.fwd : t.inv;
        S2_HIP(hipMemcpyAsync(resultspace, src, sizeof(double) * t.total, hipMemcpyDeviceToHost, c.stream));
        S2_HIP(hipStreamSynchronize(c.stream));
        out = static_cast<double**>(std::malloc(sizeof(double*) * nptr));
        if (!out) throw std::bad_alloc();
        for (int m = 0; m < nptr; ++m) out[m] = (m < bw) ? resultspace + t.h_off[m] : nullptr;
        register_tables(c, std::move(t), dir, out, bw);
    });
    return out;
}

extern "C" double** Spharmonic_Pml_Table(const int bw, double* resultspace, double* workspace) {
    (void)workspace;
    return produce_tables("Spharmonic_Pml_Table", bw, bw, 0, resultspace, bw);
}

extern "C" double** SemiNaive_Naive_Pml_Table(const int bw, const int m, double* resultspace, double* workspace) {
    (void)workspace;
    return produce_tables("SemiNaive_Naive_Pml_Table", bw, m, 0, resultspace, bw + 1);
}

// Transposed tables from the caller's forward tables: seminaive orders are
// transposed on the GPU from the caller's values; naive orders are
// regenerated (cospml.c:506-515 regenerates them too).
static double** transpose_tables(const char* fn, double** table, int bw, int cutoff, double* resultspace, int nptr) {
    double** out = nullptr;
    run_dropin(fn, [&](Ctx& c) {
        check_common(bw, 0);
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        need(table, "table");
        need(resultspace, "resultspace");
        ensure_device(c);
        const DevTables& src = device_tables(c, bw, cutoff, 0, table, bw);
        DevTables t;
        t.bw = bw;
        t.cutoff = cutoff;
        t.h_off = src.h_off;
        t.total = src.total;
        S2_HIP(hipMalloc(&t.off, sizeof(long) * bw));
        S2_HIP(hipMemcpyAsync(t.off, t.h_off.data(), sizeof(long) * bw, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMalloc(&t.inv, sizeof(double) * std::max(1L, t.total)));
        for (int m = 0; m < bw; ++m) {
            if (m < cutoff)
                transpose_cospml_one(bw, m, src.fwd + src.h_off[m], t.inv + t.h_off[m], c.stream);
            else
                gen_pml_one(bw, m, t.inv + t.h_off[m], c.stream);
        }
        S2_HIP(hipMemcpyAsync(resultspace, t.inv, sizeof(double) * t.total, hipMemcpyDeviceToHost, c.stream));
        S2_HIP(hipStreamSynchronize(c.stream));
        out = static_cast<double**>(std::malloc(sizeof(double*) * nptr));
        if (!out) throw std::bad_alloc();
        for (int m = 0; m < nptr; ++m) out[m] = (m < bw) ? resultspace + t.h_off[m] : nullptr;
        register_tables(c, std::move(t), 1, out, bw);
    });
    return out;
}

extern "C" double** Transpose_Spharmonic_Pml_Table(double** spharmonic_pml_table, const int bw, double* resultspace) {
    return transpose_tables("Transpose_Spharmonic_Pml_Table", spharmonic_pml_table, bw, bw, resultspace, bw);
}

extern "C" double** Transpose_SemiNaive_Naive_Pml_Table(double** seminaive_naive_pml_table, const int bw,
                                                        const int m, double* resultspace, double* workspace) {
    (void)workspace;
    return transpose_tables("Transpose_SemiNaive_Naive_Pml_Table", seminaive_naive_pml_table, bw, m, resultspace,
                            bw + 1);
}

// ===========================================================================
// drop-in: transforms
// ===========================================================================
// FSTSemiMemo / FSTSemiFly on device tables t (forward layout).
static void dropin_forward(Ctx& c, const DevTables& t, const double* rdata, const double* idata, double* rcoeffs,
                           double* icoeffs, int bw, DataFormat data_format, const double* weights) {
    const int N = 2 * bw;
    const size_t g = sizeof(double) * N * N, k = sizeof(double) * bw * bw;
    const DevConsts& k0 = device_consts(c, bw, weights);
    c.re.reserve(g);
    c.im.reserve(g);
    c.cre.reserve(k);
    c.cim.reserve(k);
    c.X.reserve(2 * g);
    c.Cb.reserve(sizeof(double2) * N * bw);
    S2_HIP(hipMemcpyAsync(c.re.p, rdata, g, hipMemcpyHostToDevice, c.stream));
    S2_HIP(hipMemcpyAsync(c.im.p, idata, g, hipMemcpyHostToDevice, c.stream));
    forward_fp64(k0, t, c.re.as<double>(), c.im.as<double>(), c.cre.as<double>(), c.cim.as<double>(), 1,
                 (Fmt)data_format, c.X.as<double2>(), c.Cb.as<double2>(), c.stream);
    S2_HIP(hipMemcpyAsync(rcoeffs, c.cre.p, k, hipMemcpyDeviceToHost, c.stream));
    S2_HIP(hipMemcpyAsync(icoeffs, c.cim.p, k, hipMemcpyDeviceToHost, c.stream));
}

// InvFSTSemiMemo / InvFSTSemiFly on device tables t (transposed layout).
static void dropin_inverse(Ctx& c, const DevTables& t, const double* rcoeffs, const double* icoeffs, double* rdata,
                           double* idata, int bw, DataFormat data_format) {
    const int N = 2 * bw;
    const size_t g = sizeof(double) * N * N, k = sizeof(double) * bw * bw;
    const DevConsts& k0 = device_consts(c, bw, nullptr);
    c.re.reserve(g);
    c.im.reserve(g);
    c.cre.reserve(k);
    c.cim.reserve(k);
    c.X.reserve(2 * g);
    c.Cb.reserve(sizeof(double2) * N * bw);
    S2_HIP(hipMemcpyAsync(c.cre.p, rcoeffs, k, hipMemcpyHostToDevice, c.stream));
    S2_HIP(hipMemcpyAsync(c.cim.p, icoeffs, k, hipMemcpyHostToDevice, c.stream));
    inverse_fp64(k0, t, c.cre.as<double>(), c.cim.as<double>(), c.re.as<double>(), c.im.as<double>(), 1,
                 (Fmt)data_format, c.X.as<double2>(), c.Cb.as<double2>(), c.stream);
    S2_HIP(hipMemcpyAsync(rdata, c.re.p, g, hipMemcpyDeviceToHost, c.stream));
    S2_HIP(hipMemcpyAsync(idata, c.im.p, g, hipMemcpyDeviceToHost, c.stream));
}

// FZTSemiMemo / FZTSemiFly on device tables t whose order 0 is the cosine table.
static void dropin_zonal(Ctx& c, const DevTables& t, const double* rdata, const double* idata, double* rres,
                         double* ires, int bw, DataFormat data_format, const double* weights) {
    const int N = 2 * bw;
    const size_t g = sizeof(double) * N * N;
    const DevConsts& k0 = device_consts(c, bw, weights);
    c.re.reserve(g);
    c.im.reserve(g);
    c.cre.reserve(sizeof(double) * bw);
    c.cim.reserve(sizeof(double) * bw);
    c.X.reserve(2 * g);
    c.Cb.reserve(sizeof(double2) * N * bw);
    S2_HIP(hipMemcpyAsync(c.re.p, rdata, g, hipMemcpyHostToDevice, c.stream));
    S2_HIP(hipMemcpyAsync(c.im.p, idata, g, hipMemcpyHostToDevice, c.stream));
    zonal_fp64(k0, t, c.re.as<double>(), c.im.as<double>(), c.cre.as<double>(), c.cim.as<double>(), 1,
               (Fmt)data_format, c.X.as<double2>(), c.Cb.as<double2>(), c.stream);
    S2_HIP(hipMemcpyAsync(rres, c.cre.p, sizeof(double) * bw, hipMemcpyDeviceToHost, c.stream));
    S2_HIP(hipMemcpyAsync(ires, c.cim.p, sizeof(double) * bw, hipMemcpyDeviceToHost, c.stream));
}

extern "C" void FSTSemiMemo(double* rdata, double* idata, double* rcoeffs, double* icoeffs, const int bw,
                            double** seminaive_naive_table, double* workspace, DataFormat data_format,
                            const int cutoff, fftw_plan* DCT_plan, fftw_plan* FFT_plan, double* weights) {
    (void)workspace;
    (void)DCT_plan;
    (void)FFT_plan;
    run_dropin("FSTSemiMemo", [&](Ctx& c) {
        check_common(bw, fmt_of(data_format));
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        need(rdata, "rdata");
        need(idata, "idata");
        need(rcoeffs, "rcoeffs");
        need(icoeffs, "icoeffs");
        need(seminaive_naive_table, "seminaive_naive_table");
        need(weights, "weights");
        ensure_device(c);
        const DevTables& t = device_tables(c, bw, cutoff, 0, seminaive_naive_table, bw);
        dropin_forward(c, t, rdata, idata, rcoeffs, icoeffs, bw, data_format, weights);
    });
}


extern "C" void InvFSTSemiMemo(double* rcoeffs, double* icoeffs, double* rdata, double* idata, const int bw,
                               double** transpose_seminaive_naive_table, double* workspace,
                               DataFormat data_format, const int cutoff, fftw_plan* inv_DCT_plan,
                               fftw_plan* inv_FFT_plan) {
    (void)workspace;
    (void)inv_DCT_plan;
    (void)inv_FFT_plan;
    run_dropin("InvFSTSemiMemo", [&](Ctx& c) {
        check_common(bw, fmt_of(data_format));
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        need(rdata, "rdata");
        need(idata, "idata");
        need(rcoeffs, "rcoeffs");
        need(icoeffs, "icoeffs");
        need(transpose_seminaive_naive_table, "transpose_seminaive_naive_table");
        ensure_device(c);
        const DevTables& t = device_tables(c, bw, cutoff, 1, transpose_seminaive_naive_table, bw);
        dropin_inverse(c, t, rcoeffs, icoeffs, rdata, idata, bw, data_format);
    });
}


extern "C" void FZTSemiMemo(double* rdata, double* idata, double* rres, double* ires, const int bw,
                            double* cos_pml_table, double* workspace, DataFormat data_format, fftw_plan* DCT_plan,
                            double* weights) {
    (void)workspace;
    (void)DCT_plan;
    run_dropin("FZTSemiMemo", [&](Ctx& c) {
        check_common(bw, fmt_of(data_format));
        need(rdata, "rdata");
        need(idata, "idata");
        need(rres, "rres");
        need(ires, "ires");
        need(cos_pml_table, "cos_pml_table");
        need(weights, "weights");
        const double* p0 = cos_pml_table;
        ensure_device(c);
        const DevTables& t = device_tables(c, bw, bw, 2, &p0, 1);
        dropin_zonal(c, t, rdata, idata, rres, ires, bw, data_format, weights);
    });
}


extern "C" void TransMult(double* rdatacoeffs, double* idatacoeffs, double* rfiltercoeffs, double* ifiltercoeffs,
                          double* rres, double* ires, const int bw) {
    run_dropin("TransMult", [&](Ctx& c) {
        check_common(bw, 0);
        for (const void* p : {(const void*)rdatacoeffs, (const void*)idatacoeffs, (const void*)rfiltercoeffs,
                              (const void*)ifiltercoeffs, (const void*)rres, (const void*)ires})
            need(p, "coefficient array");
        const size_t k = sizeof(double) * bw * bw, h = sizeof(double) * bw;
        ensure_device(c);
        const DevConsts& k0 = device_consts(c, bw, nullptr);
        c.cre.reserve(k);
        c.cim.reserve(k);
        c.aux1.reserve(h);
        c.aux2.reserve(h);
        c.aux3.reserve(k);
        c.aux4.reserve(k);
        S2_HIP(hipMemcpyAsync(c.cre.p, rdatacoeffs, k, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.cim.p, idatacoeffs, k, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux1.p, rfiltercoeffs, h, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux2.p, ifiltercoeffs, h, hipMemcpyHostToDevice, c.stream));
        trans_mult(k0, c.cre.as<double>(), c.cim.as<double>(), c.aux1.as<double>(), c.aux2.as<double>(), 0,
                   c.aux3.as<double>(), c.aux4.as<double>(), 1, c.stream);
        S2_HIP(hipMemcpyAsync(rres, c.aux3.p, k, hipMemcpyDeviceToHost, c.stream));
        S2_HIP(hipMemcpyAsync(ires, c.aux4.p, k, hipMemcpyDeviceToHost, c.stream));
    });
}

// Library-generated tables (both directions) for the drop-ins that take no
// tables: ConvOn2SphereSemiMemo and the FST_semi_fly functions, which the
// reference builds on every call (FST_semi_memo.c:498-500, FST_semi_fly.c:
// 96-187, 259-331); the same values, generated on the GPU once per (bw, cutoff).
static const DevTables& gen_tables(Ctx& c, int bw, int cutoff) {
    for (auto it = c.tables.begin(); it != c.tables.end(); ++it)
        if (it->bw == bw && it->cutoff == cutoff && it->dir == 3 && it->device == c.device) {
            c.tables.splice(c.tables.begin(), c.tables, it);
            return c.tables.front().t;
        }
    CachedTables e;
    e.bw = bw;
    e.cutoff = cutoff;
    e.dir = 3;
    e.device = c.device;
    tables_generate(e.t, bw, cutoff, true, true, c.stream);
    S2_HIP(hipStreamSynchronize(c.stream));
    return cache_insert(c, std::move(e));
}

// ConvOn2SphereSemiMemo / ConvOn2SphereSemiFly: REAL data, cutoff = bw
// (FST_semi_memo.c:502-507, FST_semi_fly.c:505-511).
static void dropin_conv(const char* fn, double* rdata, double* idata, double* rfilter, double* ifilter, double* rres,
                        double* ires, int bw) {
    run_dropin(fn, [&](Ctx& c) {
        check_common(bw, 0);
        for (const void* p : {(const void*)rdata, (const void*)idata, (const void*)rfilter, (const void*)ifilter,
                              (const void*)rres, (const void*)ires})
            need(p, "grid array");
        const int N = 2 * bw;
        const size_t g = sizeof(double) * N * N, k = sizeof(double) * bw * bw;
        ensure_device(c);
        const DevTables& t = gen_tables(c, bw, bw);
        const DevConsts& k0 = device_consts(c, bw, nullptr);
        c.re.reserve(g);
        c.im.reserve(g);
        c.cre.reserve(k);
        c.cim.reserve(k);
        c.X.reserve(2 * g);
        c.Cb.reserve(sizeof(double2) * N * bw);
        c.aux1.reserve(sizeof(double) * bw);
        c.aux2.reserve(sizeof(double) * bw);
        c.aux3.reserve(k);
        c.aux4.reserve(k);
        // signal: FSTSemiMemo(REAL, cutoff = bw)
        S2_HIP(hipMemcpyAsync(c.re.p, rdata, g, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.im.p, idata, g, hipMemcpyHostToDevice, c.stream));
        forward_fp64(k0, t, c.re.as<double>(), c.im.as<double>(), c.cre.as<double>(), c.cim.as<double>(), 1,
                     Fmt::Real, c.X.as<double2>(), c.Cb.as<double2>(), c.stream);
        // filter: FZTSemiMemo(REAL)
        S2_HIP(hipMemcpyAsync(c.re.p, rfilter, g, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.im.p, ifilter, g, hipMemcpyHostToDevice, c.stream));
        zonal_fp64(k0, t, c.re.as<double>(), c.im.as<double>(), c.aux1.as<double>(), c.aux2.as<double>(), 1,
                   Fmt::Real, c.X.as<double2>(), c.Cb.as<double2>(), c.stream);
        trans_mult(k0, c.cre.as<double>(), c.cim.as<double>(), c.aux1.as<double>(), c.aux2.as<double>(), 0,
                   c.aux3.as<double>(), c.aux4.as<double>(), 1, c.stream);
        inverse_fp64(k0, t, c.aux3.as<double>(), c.aux4.as<double>(), c.re.as<double>(), c.im.as<double>(), 1,
                     Fmt::Real, c.X.as<double2>(), c.Cb.as<double2>(), c.stream);
        S2_HIP(hipMemcpyAsync(rres, c.re.p, g, hipMemcpyDeviceToHost, c.stream));
        S2_HIP(hipMemcpyAsync(ires, c.im.p, g, hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void ConvOn2SphereSemiMemo(double* rdata, double* idata, double* rfilter, double* ifilter, double* rres,
                                      double* ires, const int bw, double* workspace) {
    (void)workspace;
    dropin_conv("ConvOn2SphereSemiMemo", rdata, idata, rfilter, ifilter, rres, ires, bw);
}

// ===========================================================================
// drop-in: seminaive.h, naive.h -- one order at a time
// ===========================================================================
static void check_order(int bw, int m) {
    check_common(bw, 0);
    if (m < 0 || m >= bw) throw ArgError{"m must be in [0, bw)"};
}

extern "C" void DLTSemi(double* data, const int bw, const int m, double* result, double* workspace,
                        double* cos_pml_table, double* weights, fftw_plan* plan) {
    (void)workspace;
    (void)plan;
    run_dropin("DLTSemi", [&](Ctx& c) {
        check_order(bw, m);
        need(data, "data");
        need(result, "result");
        need(cos_pml_table, "cos_pml_table");
        need(weights, "weights");
        ensure_device(c);
        const int N = 2 * bw;
        const size_t nt = table_size(m, bw);
        c.aux1.reserve(sizeof(double) * N);
        c.aux2.reserve(sizeof(double) * nt);
        c.aux3.reserve(sizeof(double) * 2 * N);
        c.aux4.reserve(sizeof(double) * (bw - m));
        S2_HIP(hipMemcpyAsync(c.aux1.p, data, sizeof(double) * N, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux2.p, cos_pml_table, sizeof(double) * nt, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux3.p, weights, sizeof(double) * 2 * N, hipMemcpyHostToDevice, c.stream));
        dlt_semi_one(c.aux1.as<double>(), c.aux3.as<double>(), c.aux2.as<double>(), c.aux4.as<double>(), bw, m,
                     c.stream);
        S2_HIP(hipMemcpyAsync(result, c.aux4.p, sizeof(double) * (bw - m), hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void InvDLTSemi(double* coeffs, const int bw, const int m, double* result, double* trans_cos_pml_table,
                           double* sin_values, double* workspace, fftw_plan* plan) {
    (void)workspace;
    (void)plan;
    run_dropin("InvDLTSemi", [&](Ctx& c) {
        check_order(bw, m);
        need(coeffs, "coeffs");
        need(result, "result");
        need(trans_cos_pml_table, "trans_cos_pml_table");
        need(sin_values, "sin_values");
        ensure_device(c);
        const int N = 2 * bw;
        const size_t nt = table_size(m, bw);
        c.aux1.reserve(sizeof(double) * (bw - m));
        c.aux2.reserve(sizeof(double) * nt);
        c.aux3.reserve(sizeof(double) * N);
        c.aux4.reserve(sizeof(double) * N);
        S2_HIP(hipMemcpyAsync(c.aux1.p, coeffs, sizeof(double) * (bw - m), hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux2.p, trans_cos_pml_table, sizeof(double) * nt, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux3.p, sin_values, sizeof(double) * N, hipMemcpyHostToDevice, c.stream));
        inv_dlt_semi_one(c.aux1.as<double>(), c.aux2.as<double>(), c.aux3.as<double>(), c.aux4.as<double>(), bw,
                         m, c.stream);
        S2_HIP(hipMemcpyAsync(result, c.aux4.p, sizeof(double) * N, hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void DLTNaive(double* data, const int bw, const int m, double* weights, double* result,
                         double* pml_table, double* workspace) {
    (void)workspace;
    run_dropin("DLTNaive", [&](Ctx& c) {
        check_order(bw, m);
        need(data, "data");
        need(weights, "weights");
        need(result, "result");
        need(pml_table, "pml_table");
        ensure_device(c);
        const int N = 2 * bw;
        const size_t nt = (size_t)(bw - m) * N;
        c.aux1.reserve(sizeof(double) * N);
        c.aux2.reserve(sizeof(double) * nt);
        c.aux3.reserve(sizeof(double) * N);
        c.aux4.reserve(sizeof(double) * (bw - m));
        S2_HIP(hipMemcpyAsync(c.aux1.p, data, sizeof(double) * N, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux2.p, pml_table, sizeof(double) * nt, hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux3.p, weights, sizeof(double) * N, hipMemcpyHostToDevice, c.stream));
        dlt_naive_one(c.aux1.as<double>(), c.aux3.as<double>(), c.aux2.as<double>(), c.aux4.as<double>(), bw, m,
                      c.stream);
        S2_HIP(hipMemcpyAsync(result, c.aux4.p, sizeof(double) * (bw - m), hipMemcpyDeviceToHost, c.stream));
    });
}

extern "C" void InvDLTNaive(double* coeffs, const int bw, const int m, double* result, double* pml_table) {
    run_dropin("InvDLTNaive", [&](Ctx& c) {
        check_order(bw, m);
        need(coeffs, "coeffs");
        need(result, "result");
        need(pml_table, "pml_table");
        ensure_device(c);
        const int N = 2 * bw;
        const size_t nt = (size_t)(bw - m) * N;
        c.aux1.reserve(sizeof(double) * (bw - m));
        c.aux2.reserve(sizeof(double) * nt);
        c.aux4.reserve(sizeof(double) * N);
        S2_HIP(hipMemcpyAsync(c.aux1.p, coeffs, sizeof(double) * (bw - m), hipMemcpyHostToDevice, c.stream));
        S2_HIP(hipMemcpyAsync(c.aux2.p, pml_table, sizeof(double) * nt, hipMemcpyHostToDevice, c.stream));
        inv_dlt_naive_one(c.aux1.as<double>(), c.aux2.as<double>(), c.aux4.as<double>(), bw, m, c.stream);
        S2_HIP(hipMemcpyAsync(result, c.aux4.p, sizeof(double) * N, hipMemcpyDeviceToHost, c.stream));
    });
}

// ===========================================================================
// drop-in: FST_semi_fly.h -- the same transforms with tables the library makes
// ===========================================================================
extern "C" void FSTSemiFly(double* rdata, double* idata, double* rcoeffs, double* icoeffs, const int bw,
                           double* workspace, DataFormat data_format, const int cutoff, fftw_plan* DCT_plan,
                           fftw_plan* FFT_plan, double* weights) {
    (void)workspace;
    (void)DCT_plan;
    (void)FFT_plan;
    run_dropin("FSTSemiFly", [&](Ctx& c) {
        check_common(bw, fmt_of(data_format));
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        need(rdata, "rdata");
        need(idata, "idata");
        need(rcoeffs, "rcoeffs");
        need(icoeffs, "icoeffs");
        need(weights, "weights");
        ensure_device(c);
        dropin_forward(c, gen_tables(c, bw, cutoff), rdata, idata, rcoeffs, icoeffs, bw, data_format, weights);
    });
}

extern "C" void InvFSTSemiFly(double* rcoeffs, double* icoeffs, double* rdata, double* idata, const int bw,
                              double* workspace, DataFormat data_format, const int cutoff, fftw_plan* inv_DCT_plan,
                              fftw_plan* inv_FFT_plan) {
    (void)workspace;
    (void)inv_DCT_plan;
    (void)inv_FFT_plan;
    run_dropin("InvFSTSemiFly", [&](Ctx& c) {
        check_common(bw, fmt_of(data_format));
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        need(rdata, "rdata");
        need(idata, "idata");
        need(rcoeffs, "rcoeffs");
        need(icoeffs, "icoeffs");
        ensure_device(c);
        dropin_inverse(c, gen_tables(c, bw, cutoff), rcoeffs, icoeffs, rdata, idata, bw, data_format);
    });
}

extern "C" void FZTSemiFly(double* rdata, double* idata, double* rres, double* ires, const int bw,
                           double* workspace, DataFormat data_format, fftw_plan* DCT_plan, double* weights) {
    (void)workspace;
    (void)DCT_plan;
    run_dropin("FZTSemiFly", [&](Ctx& c) {
        check_common(bw, fmt_of(data_format));
        need(rdata, "rdata");
        need(idata, "idata");
        need(rres, "rres");
        need(ires, "ires");
        need(weights, "weights");
        ensure_device(c);
        dropin_zonal(c, gen_tables(c, bw, bw), rdata, idata, rres, ires, bw, data_format, weights);
    });
}

extern "C" void ConvOn2SphereSemiFly(double* rdata, double* idata, double* rfilter, double* ifilter, double* rres,
                                     double* ires, const int bw, double* workspace) {
    (void)workspace;
    dropin_conv("ConvOn2SphereSemiFly", rdata, idata, rfilter, ifilter, rres, ires, bw);
}

// ===========================================================================
// batched device API
// ===========================================================================
struct s2amd_plan : s2::Plan {};

extern "C" void s2amd_set_abort_on_error(int on) { g_abort_on_error = on ? 1 : 0; }
extern "C" const char* s2amd_last_error(void) { return s2::last_error(); }

extern "C" int s2amd_plan_create(s2amd_plan** out, int bw, int cutoff, int max_fields, int precision) {
    return run_plan("s2amd_plan_create", [&] {
        if (!out) throw ArgError{"plan output pointer is NULL"};
        *out = nullptr;
        check_common(bw, 0);
        if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
        if (max_fields < 1) throw ArgError{"max_fields must be >= 1"};
        if (precision != S2AMD_FP64 && precision != S2AMD_FP32 && precision != S2AMD_MIXED &&
            precision != S2AMD_SPLIT_BF16)
            throw ArgError{"unknown precision"};
        if (precision == S2AMD_MIXED && !mixed_supported(bw))
            throw ArgError{"the mixed form needs bw in [32, 512]"};
        if (precision == S2AMD_SPLIT_BF16 && !split_supported(bw))
            throw ArgError{"the split-BF16 form needs a power-of-two grid with bw in [32, 512]"};
        if (precision == S2AMD_FP32 && !fast_supported(bw, cutoff, 1))
            throw ArgError{"the FP32 form needs a power-of-two grid with bw in [32, 1024]"};
        std::unique_ptr<s2amd_plan> p(new s2amd_plan());
        p->bw = bw;
        p->cutoff = cutoff;
        p->max_fields = max_fields;
        p->precision = precision;
        S2_HIP(hipGetDevice(&p->device));
        hipStream_t s = nullptr;
        S2_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            consts_create(p->c, bw, nullptr, s);
            tables_generate(p->t, bw, cutoff, true, true, s);
            if (precision == S2AMD_MIXED) {
                naive_tables_create(p->g, p->c, s);
            } else if (precision == S2AMD_SPLIT_BF16) {
                split_tables_create(p->g, p->c, s);
            } else {
                p->fast = fast_supported(bw, cutoff, precision);
                if (p->fast) gemm_tables_create(p->g, p->t, precision, s);
            }
            const size_t N = 2 * (size_t)bw, F = max_fields;
            // large enough for the fast layouts and for the FP64 zonal transform
            p->X.reserve(sizeof(double2) * N * N * F);  // largest layout (FP64 generic / zonal)
            p->Cb.reserve(sizeof(double2) * N * bw * F);
            p->cr.reserve(sizeof(double) * bw * bw * F);
            p->ci.reserve(sizeof(double) * bw * bw * F);
            p->zr.reserve(sizeof(double) * bw * F);
            p->zi.reserve(sizeof(double) * bw * F);
            S2_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamDestroy(s);
            consts_destroy(p->c);
            tables_destroy(p->t);
            gemm_tables_destroy(p->g);
            for (DBuf* b : {&p->X, &p->Cb, &p->cr, &p->ci, &p->zr, &p->zi}) b->release();
            throw;
        }
        S2_HIP(hipStreamDestroy(s));
        *out = p.release();
    });
}

// Time plan_forward + plan_inverse of nfields zero fields on stream s:
// best of two runs after a warm-up, in ms.
static float time_pair(Plan* p, int F, hipStream_t s, double* re, double* im, double* cr, double* ci) {
    hipEvent_t e0, e1;
    S2_HIP(hipEventCreate(&e0));
    S2_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
        S2_HIP(hipEventRecord(e0, s));
        plan_forward(p, re, im, cr, ci, F, Fmt::Complex, s);
        plan_inverse(p, cr, ci, re, im, F, Fmt::Complex, s);
        S2_HIP(hipEventRecord(e1, s));
        S2_HIP(hipEventSynchronize(e1));
        float ms = 0;
        S2_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

extern "C" int s2amd_plan_autotune(s2amd_plan* p, int nfields, void* stream) {
    return run_plan("s2amd_plan_autotune", [&] {
        if (!p) throw ArgError{"plan is NULL"};
        if (nfields < 1 || nfields > p->max_fields) throw ArgError{"nfields must be in [1, max_fields]"};
        DeviceGuard dg(p->device);
        const hipStream_t s = (hipStream_t)stream;
        const bool gemm = (p->precision == S2AMD_FP64 || p->precision == S2AMD_FP32) && use_fast(p, nfields);
        const bool mixed = p->precision == S2AMD_MIXED;
        if (!gemm && !mixed) return;  // nothing to choose for this plan and field count
        const size_t N = 2 * (size_t)p->bw, F = nfields;
        struct Scratch {  // temporary fields, freed on every exit
            DBuf re, im, cr, ci;
            ~Scratch() {
                for (DBuf* b : {&re, &im, &cr, &ci}) b->release();
            }
        } t;
        struct Restore {  // the previous choice stays if a trial fails
            GemmTables& g;
            int ncb, ct, gj;
            bool done = false;
            ~Restore() {
                if (!done) g.t_ncb = ncb, g.t_ct = ct, g.t_gj = gj;
            }
        } keep{p->g, p->g.t_ncb, p->g.t_ct, p->g.t_gj};
        t.re.reserve(sizeof(double) * N * N * F);
        t.im.reserve(sizeof(double) * N * N * F);
        t.cr.reserve(sizeof(double) * p->bw * p->bw * F);
        t.ci.reserve(sizeof(double) * p->bw * p->bw * F);
        for (DBuf* b : {&t.re, &t.im, &t.cr, &t.ci}) S2_HIP(hipMemsetAsync(b->p, 0, b->n, s));
        auto run = [&] {
            return time_pair(p, nfields, s, t.re.as<double>(), t.im.as<double>(), t.cr.as<double>(), t.ci.as<double>());
        };
        float best = 1e30f;
        if (gemm) {
            int bn = 0, bc = 0;
            for (const auto& tc : gemm_tile_choices(p->bw, p->precision)) {
                p->g.t_ncb = tc.first;
                p->g.t_ct = tc.second;
                const float ms = run();
                if (ms < best) best = ms, bn = tc.first, bc = tc.second;
            }
            p->g.t_ncb = bn;
            p->g.t_ct = bc;
        } else {
            int bg = 0;
            for (int gj : mix_gj_choices(p->bw)) {
                p->g.t_gj = gj;
                const float ms = run();
                if (ms < best) best = ms, bg = gj;
            }
            p->g.t_gj = bg;
        }
        S2_HIP(hipStreamSynchronize(s));
        keep.done = true;
    });
}

extern "C" int s2amd_plan_tuning(const s2amd_plan* p, int* cols, int* col_tiles, int* ring_groups) {
    return run_plan("s2amd_plan_tuning", [&] {
        if (!p) throw ArgError{"plan is NULL"};
        if (cols) *cols = p->g.t_ncb;
        if (col_tiles) *col_tiles = p->g.t_ct;
        if (ring_groups) *ring_groups = p->g.t_gj;
    });
}

extern "C" int s2amd_plan_destroy(s2amd_plan* p) {
    return run_plan("s2amd_plan_destroy", [&] {
        if (!p) return;
        // work queued on the caller's streams may still read the plan's buffers
        DeviceGuard dg(p->device);
        S2_HIP(hipDeviceSynchronize());
        consts_destroy(p->c);
        tables_destroy(p->t);
        gemm_tables_destroy(p->g);
        for (DBuf* b : {&p->X, &p->Cb, &p->cr, &p->ci, &p->zr, &p->zi}) b->release();
        delete p;
    });
}

extern "C" size_t s2amd_plan_bytes(const s2amd_plan* p) {
    if (!p) return 0;
    const size_t el = p->precision == S2AMD_FP64        ? sizeof(double)
                      : p->precision == S2AMD_SPLIT_BF16 ? sizeof(uint16_t)
                                                         : sizeof(float);  // GEMM table elements
    return 2 * sizeof(double) * p->t.total + el * (p->g.nf + p->g.ni) + p->X.n + p->Cb.n + p->cr.n + p->ci.n +
           p->zr.n + p->zi.n;
}

extern "C" int s2amd_plan_tables(const s2amd_plan* p, const double** fwd, const double** inv, long* n) {
    return run_plan("s2amd_plan_tables", [&] {
        if (!p) throw ArgError{"plan is NULL"};
        if (fwd) *fwd = p->t.fwd;
        if (inv) *inv = p->t.inv;
        if (n) *n = p->t.total;
    });
}

extern "C" int s2amd_forward(s2amd_plan* p, const double* re, const double* im, double* cre, double* cim,
                             int nfields, int data_format, void* stream) {
    return run_plan("s2amd_forward", [&] {
        plan_check_call(p, nfields, data_format);
        if (nfields == 0) return;
        need(re, "re");
        need(im, "im");
        need(cre, "cre");
        need(cim, "cim");
        plan_forward(p, re, im, cre, cim, nfields, (Fmt)data_format, (hipStream_t)stream);
    });
}

extern "C" int s2amd_forward_noalloc(s2amd_plan* p, const double* re, const double* im, double* cre, double* cim,
                                     int nfields, int data_format, void* stream) {
    return run_plan("s2amd_forward_noalloc", [&] {
        plan_check_call(p, nfields, data_format);
        if (nfields == 0) return;
        need(re, "re");
        need(im, "im");
        need(cre, "cre");
        need(cim, "cim");
        size_t cb_off = 0;
        const size_t bytes = noalloc_need(p, nfields, &cb_off);
        if (bytes > S2AMD_NOALLOC_SCRATCH_BYTES) {
            size_t per = 0, off = 0;
            per = noalloc_need(p, 1, &off);
            throw ArgError{std::to_string(nfields) + " fields at bw=" + std::to_string(p->bw) + " need " +
                           std::to_string(bytes) + " bytes of scratch, beyond the " +
                           std::to_string((size_t)S2AMD_NOALLOC_SCRATCH_BYTES) + " of the static array (" +
                           std::to_string(per) + " per field)"};
        }
        unsigned char* base = static_cast<unsigned char*>(scratch_on(p->device));
        plan_forward_on(p, re, im, cre, cim, nfields, (Fmt)data_format, base, base + cb_off, (hipStream_t)stream,
                        nullptr);
    });
}

extern "C" int s2amd_inverse(s2amd_plan* p, const double* cre, const double* cim, double* re, double* im,
                             int nfields, int data_format, void* stream) {
    return run_plan("s2amd_inverse", [&] {
        plan_check_call(p, nfields, data_format);
        if (nfields == 0) return;
        need(re, "re");
        need(im, "im");
        need(cre, "cre");
        need(cim, "cim");
        plan_inverse(p, cre, cim, re, im, nfields, (Fmt)data_format, (hipStream_t)stream);
    });
}

extern "C" int s2amd_zonal(s2amd_plan* p, const double* re, const double* im, double* zr, double* zi, int nfields,
                           int data_format, void* stream) {
    return run_plan("s2amd_zonal", [&] {
        plan_check_call(p, nfields, data_format);
        if (nfields == 0) return;
        need(re, "re");
        need(im, "im");
        need(zr, "zr");
        need(zi, "zi");
        zonal_fp64(p->c, p->t, re, im, zr, zi, nfields, (Fmt)data_format, p->X.as<double2>(), p->Cb.as<double2>(),
                   (hipStream_t)stream, &p->timer);
    });
}

extern "C" int s2amd_conv(s2amd_plan* p, const double* sre, const double* sim, const double* hre,
                          const double* him, double* ore, double* oim, int nfields, void* stream) {
    return run_plan("s2amd_conv", [&] {
        plan_check_call(p, nfields, 1);
        if (nfields == 0) return;
        if (p->cutoff != p->bw) throw ArgError{"convolution needs a plan with cutoff = bw"};
        for (const void* q : {(const void*)sre, (const void*)sim, (const void*)hre, (const void*)him,
                              (const void*)ore, (const void*)oim})
            need(q, "grid array");
        hipStream_t s = (hipStream_t)stream;
        double* cr = p->cr.as<double>();
        double* ci = p->ci.as<double>();
        StageTimer* tm = &p->timer;
        plan_forward(p, sre, sim, cr, ci, nfields, Fmt::Real, s);
        zonal_fp64(p->c, p->t, hre, him, p->zr.as<double>(), p->zi.as<double>(), nfields, Fmt::Real,
                   p->X.as<double2>(), p->Cb.as<double2>(), s, tm);
        trans_mult(p->c, cr, ci, p->zr.as<double>(), p->zi.as<double>(), p->bw, cr, ci, nfields, s, tm);
        plan_inverse(p, cr, ci, ore, oim, nfields, Fmt::Real, s);
    });
}

extern "C" int s2amd_plan_timing(s2amd_plan* p, int enable) {
    return run_plan("s2amd_plan_timing", [&] {
        if (!p) throw ArgError{"plan is NULL"};
        p->timer.collect();
        p->timer.reset();
        p->timer.on = enable != 0;
    });
}

extern "C" int s2amd_plan_stage_times(s2amd_plan* p, double* ms, long* counts) {
    return run_plan("s2amd_plan_stage_times", [&] {
        if (!p) throw ArgError{"plan is NULL"};
        p->timer.collect();
        for (int i = 0; i < S2AMD_NSTAGES; ++i) {
            if (ms) ms[i] = p->timer.ms[i];
            if (counts) counts[i] = p->timer.count[i];
        }
        p->timer.reset();
    });
}

extern "C" const char* s2amd_stage_name(int stage) { return s2::stage_name(stage); }
