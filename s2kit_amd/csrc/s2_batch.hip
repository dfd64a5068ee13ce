// Convolution batches of uneven bandwidth (s2amd_conv_batch_*, s2kit_amd.h).
//
// Requests are grouped by bandwidth; each group gets a plan with max_fields =
// its size and runs the plan's convolution (forward, zonal filter
// coefficients, TransMult, inverse: ConvOn2SphereSemiMemo, FST_semi_memo.c:
// 439-518, per field) over all its requests at once.  Grids that are not
// already back to back in request order are gathered into contiguous staging
// buffers first, and outputs scattered back, by the copy kernels below.  The
// whole batch is recorded once as a HIP graph on a library stream and
// replayed by s2amd_conv_batch_run, so a repeated batch costs one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "s2_internal.h"

extern "C" {
#include "s2kit_amd.h"
}

namespace s2 {

// Copy n request grids of nn doubles each: array a (blockIdx.z) of request f
// (blockIdx.y) between its own buffer and slot f of a contiguous buffer.
struct CopyPlan {
    double* base[4];  // contiguous buffers, one per array
};

__global__ __launch_bounds__(256) void k_gather(const double* const* __restrict__ src, CopyPlan cp, long nn, int n) {
    const int a = blockIdx.z, f = blockIdx.y;
    const double4* s = reinterpret_cast<const double4*>(src[a * n + f]);
    double4* d = reinterpret_cast<double4*>(cp.base[a] + f * nn);
    for (long i = blockIdx.x * 256L + threadIdx.x; i < nn / 4; i += 256L * gridDim.x) d[i] = s[i];
}

__global__ __launch_bounds__(256) void k_scatter(double* const* __restrict__ dst, CopyPlan cp, long nn, int n) {
    const int a = blockIdx.z, f = blockIdx.y;
    const double4* s = reinterpret_cast<const double4*>(cp.base[a] + f * nn);
    double4* d = reinterpret_cast<double4*>(dst[a * n + f]);
    for (long i = blockIdx.x * 256L + threadIdx.x; i < nn / 4; i += 256L * gridDim.x) d[i] = s[i];
}

struct ConvGroup {
    int bw = 0, precision = 0, n = 0;
    s2amd_plan* plan = nullptr;
    const double* in[4] = {};  // sre, sim, hre, him as the pipeline reads them
    double* out[2] = {};       // ore, oim
    int ngather = 0, nscatter = 0;
    CopyPlan gather{}, scatter{};
    void* gather_src = nullptr;   // device [ngather][n] pointers
    void* scatter_dst = nullptr;  // device [nscatter][n] pointers
    std::vector<void*> staging;
};

}  // namespace s2

using namespace s2;

struct s2amd_conv_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<ConvGroup> groups;
};

static void conv_batch_free(s2amd_conv_batch* b) {
    if (!b) return;
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (ConvGroup& g : b->groups) {
        if (g.plan) (void)s2amd_plan_destroy(g.plan);
        for (void* p : g.staging) (void)hipFree(p);
        if (g.gather_src) (void)hipFree(g.gather_src);
        if (g.scatter_dst) (void)hipFree(g.scatter_dst);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    if (prev >= 0) (void)hipSetDevice(prev);
}

static void check_plan(int rc) {
    if (rc == S2AMD_OK) return;
    if (rc == S2AMD_EINVAL) throw ArgError{last_error()};
    throw HipError{rc == S2AMD_ENOMEM ? hipErrorOutOfMemory : hipErrorUnknown, last_error()};
}

static bool form_ok(int precision, int bw) {
    return precision == S2AMD_MIXED ? mixed_supported(bw) : split_supported(bw);
}

// Launch grid of the copy kernels: enough blocks to stream nn doubles per request.
static dim3 copy_grid(long nn, int n, int arrays) {
    const long blocks = std::min<long>(1024, std::max<long>(1, nn / 4 / 256 / 4));
    return dim3((unsigned)blocks, (unsigned)n, (unsigned)arrays);
}

extern "C" int s2amd_conv_batch_create(s2amd_conv_batch** out, const s2amd_conv_request* req, int n, int precision,
                                       void* stream) {
    return run_api("s2amd_conv_batch_create", [&] {
        if (!out) throw ArgError{"batch output pointer is NULL"};
        *out = nullptr;
        if (n < 1) throw ArgError{"need at least one request"};
        if (!req) throw ArgError{"request array is NULL"};
        if (precision != S2AMD_FP64 && precision != S2AMD_MIXED && precision != S2AMD_SPLIT_BF16)
            throw ArgError{"precision must be S2AMD_FP64, S2AMD_MIXED or S2AMD_SPLIT_BF16"};
        std::map<int, std::vector<int>> by_bw;
        for (int i = 0; i < n; ++i) {
            const s2amd_conv_request& r = req[i];
            if (r.bw < 2 || r.bw > 1024) throw ArgError{"request " + std::to_string(i) + ": bw must be in [2, 1024]"};
            for (const void* q : {(const void*)r.sre, (const void*)r.sim, (const void*)r.hre, (const void*)r.him,
                                  (const void*)r.ore, (const void*)r.oim}) {
                if (!q) throw ArgError{"request " + std::to_string(i) + ": a grid pointer is NULL"};
                if ((uintptr_t)q % 32) throw ArgError{"request " + std::to_string(i) + ": grids must be 32-byte aligned"};
            }
            by_bw[r.bw].push_back(i);
        }
        for (auto& kv : by_bw)
            if (kv.second.size() > 65535) throw ArgError{"at most 65535 requests per bandwidth"};
        std::unique_ptr<s2amd_conv_batch, void (*)(s2amd_conv_batch*)> b(new s2amd_conv_batch(), conv_batch_free);
        S2_HIP(hipGetDevice(&b->device));
        S2_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        for (auto& kv : by_bw) {
            b->groups.emplace_back();
            ConvGroup& g = b->groups.back();
            const std::vector<int>& idx = kv.second;
            g.bw = kv.first;
            g.n = (int)idx.size();
            g.precision = (precision != S2AMD_FP64 && form_ok(precision, g.bw)) ? precision : S2AMD_FP64;
            check_plan(s2amd_plan_create(&g.plan, g.bw, g.bw, g.n, g.precision));
            // tile shapes timed at this group's size before the graph fixes them
            check_plan(s2amd_plan_autotune(g.plan, g.n, b->stream));
            const long nn = 4L * g.bw * g.bw;
            const size_t bytes = sizeof(double) * nn * g.n;
            auto ptr_in = [&](int a, int i) -> const double* {
                const s2amd_conv_request& r = req[idx[i]];
                return a == 0 ? r.sre : a == 1 ? r.sim : a == 2 ? r.hre : r.him;
            };
            auto ptr_out = [&](int a, int i) -> double* { return a == 0 ? req[idx[i]].ore : req[idx[i]].oim; };
            // inputs: use a request array in place when its grids lie back to back
            std::vector<const double*> gsrc;
            for (int a = 0; a < 4; ++a) {
                bool contiguous = true;
                for (int i = 1; i < g.n && contiguous; ++i) contiguous = ptr_in(a, i) == ptr_in(a, 0) + i * nn;
                if (contiguous) {
                    g.in[a] = ptr_in(a, 0);
                    continue;
                }
                void* st = nullptr;
                S2_HIP(hipMalloc(&st, bytes));
                g.staging.push_back(st);
                g.gather.base[g.ngather++] = (double*)st;
                g.in[a] = (const double*)st;
                for (int i = 0; i < g.n; ++i) gsrc.push_back(ptr_in(a, i));
            }
            std::vector<double*> sdst;
            for (int a = 0; a < 2; ++a) {
                bool contiguous = true;
                for (int i = 1; i < g.n && contiguous; ++i) contiguous = ptr_out(a, i) == ptr_out(a, 0) + i * nn;
                if (contiguous) {
                    g.out[a] = ptr_out(a, 0);
                    continue;
                }
                void* st = nullptr;
                S2_HIP(hipMalloc(&st, bytes));
                g.staging.push_back(st);
                g.scatter.base[g.nscatter++] = (double*)st;
                g.out[a] = (double*)st;
                for (int i = 0; i < g.n; ++i) sdst.push_back(ptr_out(a, i));
            }
            if (!gsrc.empty()) {
                S2_HIP(hipMalloc(&g.gather_src, sizeof(void*) * gsrc.size()));
                S2_HIP(hipMemcpy(g.gather_src, gsrc.data(), sizeof(void*) * gsrc.size(), hipMemcpyHostToDevice));
            }
            if (!sdst.empty()) {
                S2_HIP(hipMalloc(&g.scatter_dst, sizeof(void*) * sdst.size()));
                S2_HIP(hipMemcpy(g.scatter_dst, sdst.data(), sizeof(void*) * sdst.size(), hipMemcpyHostToDevice));
            }
        }
        // record the batch
        S2_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        try {
            for (ConvGroup& g : b->groups) {
                const long nn = 4L * g.bw * g.bw;
                if (g.ngather) {
                    k_gather<<<copy_grid(nn, g.n, g.ngather), 256, 0, b->stream>>>(
                        (const double* const*)g.gather_src, g.gather, nn, g.n);
                    S2_HIP(hipGetLastError());
                }
                check_plan(s2amd_conv(g.plan, g.in[0], g.in[1], g.in[2], g.in[3], g.out[0], g.out[1], g.n, b->stream));
                if (g.nscatter) {
                    k_scatter<<<copy_grid(nn, g.n, g.nscatter), 256, 0, b->stream>>>(
                        (double* const*)g.scatter_dst, g.scatter, nn, g.n);
                    S2_HIP(hipGetLastError());
                }
            }
        } catch (...) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(b->stream, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
            throw;
        }
        S2_HIP(hipStreamEndCapture(b->stream, &b->graph));
        S2_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        (void)stream;  // creation records on the library stream; nothing is launched
        *out = b.release();
    });
}

extern "C" int s2amd_conv_batch_run(s2amd_conv_batch* b, void* stream) {
    return run_api("s2amd_conv_batch_run", [&] {
        if (!b) throw ArgError{"batch is NULL"};
        int dev = -1;
        S2_HIP(hipGetDevice(&dev));
        if (dev != b->device) throw ArgError{"batch was created on another device"};
        S2_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    });
}

extern "C" int s2amd_conv_batch_info(const s2amd_conv_batch* b, int* ngroups, int* bws, int* counts,
                                     int* precisions) {
    return run_api("s2amd_conv_batch_info", [&] {
        if (!b) throw ArgError{"batch is NULL"};
        if (ngroups) *ngroups = (int)b->groups.size();
        for (size_t i = 0; i < b->groups.size(); ++i) {
            if (bws) bws[i] = b->groups[i].bw;
            if (counts) counts[i] = b->groups[i].n;
            if (precisions) precisions[i] = b->groups[i].precision;
        }
    });
}

extern "C" int s2amd_conv_batch_destroy(s2amd_conv_batch* b) {
    return run_api("s2amd_conv_batch_destroy", [&] { conv_batch_free(b); });
}
