// Device building blocks shared by the batched FP64/FP32 path (s2_fast.hip)
// and the mixed form (s2_mixed.hip): register DFTs, the LDS Stockham FFT,
// XCD-aware block numbering, the four ring-FFT / DCT kernels, MFMA wrappers
// and a launcher for kernels with more than 64 KiB of dynamic LDS.
//
//   k_fphi_fwd   ring FFT along phi, order-major store          (FST_semi_memo.c:81-90)
//   k_fdct_fwd   weights + DCT-II per order row                 (seminaive.c:161-176)
//   k_fdct_inv   DCT-III, x sin, row signs, 1/sqrt(2pi)         (seminaive.c:107-114; FST_semi_memo.c:301-348)
//   k_fphi_inv   inverse ring FFT                               (FST_semi_memo.c:350)
//
// FFTs: in-place Stockham passes of radix 8 (then 4 or 2) with the values of
// each butterfly in registers and one padded LDS buffer (row stride N+1);
// the N twiddles are staged in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include "s2_complex.h"
#include "s2_internal.h"

namespace s2 {

constexpr int kThreads = 256;

// ---------------------------------------------------------------------------
// small DFTs in registers (SIGN -1 forward, +1 inverse)
// ---------------------------------------------------------------------------
template <int SIGN, typename T>
__device__ __forceinline__ T mul_si(T v) {  // v * (SIGN * i)
    using R = typename Real<T>::type;
    return mk<T>(-(R)SIGN * v.y, (R)SIGN * v.x);
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft2(T* x) {
    const T a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft4v(T& x0, T& x1, T& x2, T& x3) {
    const T t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = csub(x1, x3);
    const T u3 = mul_si<SIGN>(t3);
    x0 = cadd(t0, t2);
    x2 = csub(t0, t2);
    x1 = cadd(t1, u3);
    x3 = csub(t1, u3);
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft4(T* x) {
    dft4v<SIGN>(x[0], x[1], x[2], x[3]);
}

template <int SIGN, typename T>
__device__ __forceinline__ void dft8(T* x) {
    using R = typename Real<T>::type;
    const R c = (R)0.70710678118654752440;
    const R s = (R)SIGN;
    T e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    T o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4v<SIGN>(e0, e1, e2, e3);
    dft4v<SIGN>(o0, o1, o2, o3);
    o1 = mk<T>(c * (o1.x - s * o1.y), c * (o1.y + s * o1.x));   // * exp(SIGN i pi/4)
    o2 = mul_si<SIGN>(o2);                                      // * exp(SIGN i pi/2)
    o3 = mk<T>(-c * (o3.x + s * o3.y), c * (s * o3.x - o3.y));  // * exp(SIGN 3 i pi/4)
    x[0] = cadd(e0, o0);
    x[4] = csub(e0, o0);
    x[1] = cadd(e1, o1);
    x[5] = csub(e1, o1);
    x[2] = cadd(e2, o2);
    x[6] = csub(e2, o2);
    x[3] = cadd(e3, o3);
    x[7] = csub(e3, o3);
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dftR(T* x) {
    if constexpr (R == 8)
        dft8<SIGN>(x);
    else if constexpr (R == 4)
        dft4<SIGN>(x);
    else
        dft2<SIGN>(x);
}

// Twiddle exp(-2 pi i t / N) from tw: tw[t], or with TW2 a two-level table
// of 2^(lgN - L) + 2^L values, hi[u] = exp(-2 pi i u 2^L / N) then
// lo[v] = exp(-2 pi i v / N), L = tw2_bits(lgN): hi[t >> L] * lo[t mod 2^L]
// (one rounding more, about 2 sqrt(N) values of LDS instead of N).
__host__ __device__ constexpr int tw2_bits(int lgN) { return (lgN + 1) / 2; }
template <bool TW2, typename T>
__device__ __forceinline__ T twiddle(const T* tw, int t, int lgN) {
    if constexpr (TW2) {
        const int L = tw2_bits(lgN);
        return cmul(tw[t >> L], tw[(1 << (lgN - L)) + (t & ((1 << L) - 1))]);
    } else {
        return tw[t];
    }
}

// One Stockham pass of radix R over the E = V*NTH values of the block of NTH
// threads (E/N transforms of length N, row stride S).  Sub-transform size
// Ns = 2^lgNs.
template <typename T, int SIGN, int V, int R, int LGR, int NTH, bool TW2 = false>
__device__ __forceinline__ void fft_pass(T* a, int lgN, int S, int lgNs, const T* tw) {
    constexpr int PER = V / R;
    const int lgNR = lgN - LGR;
    const int NR = 1 << lgNR;
    const int Ns = 1 << lgNs;
    T x[V];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int b = threadIdx.x + u * NTH;
        const int f = b >> lgNR, j = b & (NR - 1);
        const T* src = a + f * S + j;
#pragma unroll
        for (int r = 0; r < R; ++r) x[u * R + r] = src[r * NR];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int b = threadIdx.x + u * NTH;
        const int f = b >> lgNR, j = b & (NR - 1);
        const int k = j & (Ns - 1);
        const int step = k << (lgNR - lgNs);  // k * N / (R * Ns)
        if (lgNs > 0) {  // the first pass has k = 0: unit twiddles
#pragma unroll
            for (int r = 1; r < R; ++r) {
                T w = twiddle<TW2>(tw, r * step, lgN);
                if (SIGN > 0) w.y = -w.y;
                x[u * R + r] = cmul(x[u * R + r], w);
            }
        }
        dftR<R, SIGN>(x + u * R);
        T* dst = a + f * S + ((j - k) << LGR) + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns] = x[u * R + r];
    }
    __syncthreads();
}

// In-place FFTs of all E/N rows of the block's LDS tile (natural order in and out).
template <typename T, int SIGN, int V, int NTH = kThreads, bool TW2 = false>
__device__ __forceinline__ void fft_block(T* a, int lgN, int S, const T* tw) {
    int lgNs = 0, rem = lgN;
    while (rem >= 3) {
        fft_pass<T, SIGN, V, 8, 3, NTH, TW2>(a, lgN, S, lgNs, tw);
        lgNs += 3;
        rem -= 3;
    }
    if (rem == 2)
        fft_pass<T, SIGN, V, 4, 2, NTH, TW2>(a, lgN, S, lgNs, tw);
    else if (rem == 1)
        fft_pass<T, SIGN, V, 2, 1, NTH, TW2>(a, lgN, S, lgNs, tw);
}

__device__ __forceinline__ int frow_of(int q, int N, int cut) { return q < cut ? q : N - 2 * cut + 1 + q; }

// Streaming accesses of the ring FFTs, non-temporal with NT (no cache
// retention for data read or written once per kernel).  Measured at bw=256
// x1024 (mixed form): in the forward FFT the grid loads and the transposed
// stores both gain (k_fphi_fwd 1.19 -> ~1.08 ms) when its row segments are
// whole 128-byte lines; with 64-byte segments (bw=512 FP64, 4 rings x 16 B)
// it lost 5-13%, and in the inverse FFT either kind lost (1.19 -> 1.23 ms).
// k_fphi_fwd_nt() says which forward instance to launch.
inline bool k_fphi_fwd_nt(int rings_per_block, int value_bytes) { return rings_per_block * value_bytes >= 128; }

template <bool NT, typename E>
__device__ __forceinline__ E ld_stream(const E* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_stream_c(T* p, T v) {  // complex
    if constexpr (NT) {
        using E = typename Real<T>::type;
        typedef E v2 __attribute__((ext_vector_type(2)));
        const v2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
    } else {
        *p = v;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (linear id % 8), and
// each XCD has its own L2.  Renumber blocks so that every XCD works through a
// contiguous range of logical blocks (x fastest): neighbours that share cache
// lines (transposed 64-byte row segments) or a table tile (same order m) then
// meet in the same L2.  A bijection on [0, gridDim.x * gridDim.y).
__device__ __forceinline__ void xcd_block(int& bx, int& by) {
    const int nx = gridDim.x, total = nx * gridDim.y;
    const int id = blockIdx.y * nx + blockIdx.x;
    const int q = total >> 3, r = total & 7, k = id & 7, local = id >> 3;
    const int lid = (k < r) ? k * (q + 1) + local : r * (q + 1) + (k - r) * q + local;
    by = lid / nx;
    bx = lid - by * nx;
}

// GEMM grids (column blocks x orders, orders a multiple of 8): XCD k takes
// the orders m = k, k+8, k+16, ... (the work per order falls with m, so this
// balances the XCDs) and runs all column blocks of one order back to back,
// so that order's table tile stays in its L2.
__device__ __forceinline__ void xcd_block_m(int& bx, int& m) {
    const int nx = gridDim.x;
    const int id = blockIdx.y * nx + blockIdx.x;
    const int k = id & 7, local = id >> 3;
    const int slot = local / nx;
    bx = local - slot * nx;
    m = k + 8 * slot;
}

// position of cosine index k in the parity-split planes of the Cb buffer
__device__ __forceinline__ int cpos(int k, int half) { return (k & 1) ? half + (k >> 1) : (k >> 1); }

// Copy the N twiddles exp(-2 pi i t/N) into LDS (they are read by every
// butterfly of every pass).  Called before a kernel's data loads, its wait
// for the twiddles holds every data load back by one memory round trip, so
// most instances call it after issuing them (bw=256 x 1024: FP64 FFTs and
// DCTs 3-7% faster, the FP32 forward FFT 1.08 -> 1.05 ms); the FP32
// instances noted at their call sites measured faster with it first.
template <typename T, int NTH = kThreads>
__device__ __forceinline__ T* stage_twiddles(unsigned char* smem, const T* __restrict__ tw, int N) {
    T* t = reinterpret_cast<T*>(smem);
    for (int i = threadIdx.x; i < N; i += NTH) t[i] = tw[i];
    return t;
}

// ---------------------------------------------------------------------------
// FFT-type kernels.  grid.y = field, block = NTH (default kThreads) threads,
// LDS = N twiddles + rows*(N+1) values
// ---------------------------------------------------------------------------
template <typename R, int V, int NTH = kThreads, bool NTF = false>
__global__ __launch_bounds__(NTH) void k_fphi_fwd(const double* __restrict__ re, const double* __restrict__ im,
                                                       typename Cplx<R>::type* __restrict__ X, int lgN,
                                                       const typename Cplx<R>::type* __restrict__ tw) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int E = V * NTH;
    const int N = 1 << lgN, S = N + 1;
    const T* twl = reinterpret_cast<const T*>(smem);
    T* a = reinterpret_cast<T*>(smem) + N;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int j0 = bx * rpb;
    const size_t base = ((size_t)f * N + j0) * N;
    {
        double vr[V], vi[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            vr[u] = ld_stream<NTF>(re + base + threadIdx.x + u * NTH);
            vi[u] = ld_stream<NTF>(im + base + threadIdx.x + u * NTH);
        }
        stage_twiddles<T, NTH>(smem, tw, N);
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            a[(idx >> lgN) * S + (idx & (N - 1))] = mk<T>((R)vr[u], (R)vi[u]);
        }
    }
    __syncthreads();
    fft_block<T, -1, V, NTH>(a, lgN, S, twl);
    {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            v[u] = a[(idx & (rpb - 1)) * S + (idx >> lgrpb)];
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            st_stream_c<NTF>(X + ((size_t)f * N + (idx >> lgrpb)) * N + j0 + (idx & (rpb - 1)), v[u]);
        }
    }
}

// Weighted DCT-II of each order row (the data of DLTSemi, seminaive.c:161-176),
// by one N-point complex FFT of the even/odd reordered samples; one block
// transforms rpb = V*NTH/N rows.  Loads and stores are batched per thread.
template <typename R, int V, int NTH = kThreads>
__global__ __launch_bounds__(NTH) void k_fdct_fwd(const typename Cplx<R>::type* __restrict__ X, R* __restrict__ Cb,
                                                  const R* __restrict__ w_semi, int lgN, int cut, int nrows,
                                                  const typename Cplx<R>::type* __restrict__ tw,
                                                  const typename Cplx<R>::type* __restrict__ ph) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const T* twl = reinterpret_cast<const T*>(smem);
    T* a = reinterpret_cast<T*>(smem) + (1 << lgN);
    constexpr int E = V * NTH;
    const int N = 1 << lgN, S = N + 1, B = N >> 1, half = B >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int q0 = bx * rpb;
    // FP32 stages the twiddles before its loads (behind them: bw=512 x 64
    // 0.24 -> 0.26 ms, bw=256 x 1024 0.80 -> 0.77 ms)
    if constexpr (sizeof(R) == 4) stage_twiddles<T, NTH>(smem, tw, N);
    {
        R xr[V], xi[V], ww[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {  // rows past nrows read the last row and weigh it by 0
            const int idx = threadIdx.x + u * NTH;
            const int q = q0 + (idx >> lgN), j = idx & (N - 1);
            const int r = frow_of(min(q, nrows - 1), N, cut);
            const int am = r < B ? r : N - r;
            const T xv = X[((size_t)f * N + r) * N + j];
            xr[u] = xv.x;
            xi[u] = xv.y;
            ww[u] = q < nrows ? w_semi[(am & 1) * N + j] : (R)0;
        }
        if constexpr (sizeof(R) == 8) stage_twiddles<T, NTH>(smem, tw, N);
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            const int rl = idx >> lgN, j = idx & (N - 1);
            const int p = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
            a[rl * S + p] = mk<T>(xr[u] * ww[u], xi[u] * ww[u]);
        }
    }
    __syncthreads();
    fft_block<T, -1, V, NTH>(a, lgN, S, twl);
#pragma unroll
    for (int u = 0; u < V / 2; ++u) {  // rpb*B outputs: c[k] = e^{-i pi k/2N} v[k] + e^{+i pi k/2N} v[N-k]
        const int idx = threadIdx.x + u * NTH;
        const int rl = idx >> (lgN - 1), k = idx & (B - 1);
        const int q = q0 + rl;
        const T* v = a + rl * S;
        const T e = ph[k];
        T c = cadd(cmul(e, v[k]), cmul(cconj(e), v[(N - k) & (N - 1)]));
        if (k == 0) c = cscale(c, (R)0.70710678118654752440);
        if (q < nrows) {
            R* row = Cb + ((size_t)f * N + frow_of(q, N, cut)) * 2 * B;
            const int pos = cpos(k, half);
            row[pos] = c.x;
            row[B + pos] = c.y;
        }
    }
}

// DCT-III of each order row, x sin(theta) for odd orders, row signs and
// 1/sqrt(2pi) (InvDLTSemi's tail, seminaive.c:107-114, and
// FST_semi_memo.c:301-348): V[k] = e^{+i pi k/2N} (X[k] - i X[N-k]), one
// N-point inverse FFT, even/odd unshuffle.  REAL data also fills row N-r.
template <typename R, int V, int NTH = kThreads>
__global__ __launch_bounds__(NTH) void k_fdct_inv(const R* __restrict__ Cb, typename Cplx<R>::type* __restrict__ X,
                                                  const R* __restrict__ sinv, int lgN, int cut, int nrows, int fmt,
                                                  const typename Cplx<R>::type* __restrict__ tw,
                                                  const typename Cplx<R>::type* __restrict__ ph) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const T* twl = reinterpret_cast<const T*>(smem);
    T* a = reinterpret_cast<T*>(smem) + (1 << lgN);
    constexpr int E = V * NTH;
    const int N = 1 << lgN, S = N + 1, B = N >> 1, half = B >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int q0 = bx * rpb;
    {
        // V[k] = conj(ph[k]) * (X[k] - i X[N-k]) with X zero from index B on:
        // X[k] for k < B, -i X[N-k] for k > B, 0 at k = B.
        if constexpr (sizeof(R) == 8) {
            // A thread loads the coefficient of each of its k < B once and forms
            // V at k and at its partner position N - k (B for k = 0, where V is
            // zero).  Loading per output position reads every coefficient twice:
            // FP64 bw=256 x 1024 1.72 -> 1.59 ms with the pairing; FP32 is 12%
            // slower with it.  The phase factors of both positions are loaded
            // with the coefficients: loaded in the second loop, each iteration
            // waited for its own (a branch there kept them from being batched).
            R cr[V / 2], ci[V / 2];
            T pa[V / 2], pb[V / 2];
#pragma unroll
            for (int u = 0; u < V / 2; ++u) {
                const int idx = threadIdx.x + u * NTH;
                const int q = q0 + (idx >> (lgN - 1)), k = idx & (B - 1);
                const R* row = Cb + ((size_t)f * N + frow_of(min(q, nrows - 1), N, cut)) * 2 * B;
                cr[u] = row[cpos(k, half)];
                ci[u] = row[B + cpos(k, half)];
                pa[u] = ph[k];
                pb[u] = ph[k == 0 ? B : N - k];
            }
            stage_twiddles<T, NTH>(smem, tw, N);
#pragma unroll
            for (int u = 0; u < V / 2; ++u) {
                const int idx = threadIdx.x + u * NTH;
                const int rl = idx >> (lgN - 1), k = idx & (B - 1);
                const bool live = q0 + rl < nrows, lb = live && k != 0;
                const R ar = live ? cr[u] : (R)0, ai = live ? ci[u] : (R)0;  // position k
                const R br = lb ? ci[u] : (R)0, bi = lb ? -cr[u] : (R)0;     // partner position
                a[rl * S + k] = cmul(cconj(pa[u]), mk<T>(ar, ai));
                a[rl * S + (k == 0 ? B : N - k)] = cmul(cconj(pb[u]), mk<T>(br, bi));
            }
        } else {
            // one value per output position; FP32 stages the twiddles first
            // (behind its loads this instance ran 1.10 -> 1.39 ms at bw=256 x 1024)
            stage_twiddles<T, NTH>(smem, tw, N);
            R cr[V], ci[V];
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const int idx = threadIdx.x + u * NTH;
                const int q = q0 + (idx >> lgN), k = idx & (N - 1);
                const R* row = Cb + ((size_t)f * N + frow_of(min(q, nrows - 1), N, cut)) * 2 * B;
                const int ks = k < B ? k : (k > B ? N - k : 0);
                cr[u] = row[cpos(ks, half)];
                ci[u] = row[B + cpos(ks, half)];
            }
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const int idx = threadIdx.x + u * NTH;
                const int q = q0 + (idx >> lgN), rl = idx >> lgN, k = idx & (N - 1);
                const bool live = q < nrows && k != B;
                const R tr = live ? (k < B ? cr[u] : ci[u]) : (R)0;
                const R ti = live ? (k < B ? ci[u] : -cr[u]) : (R)0;
                a[rl * S + k] = cmul(cconj(ph[k]), mk<T>(tr, ti));
            }
        }
    }
    // sin(theta_j) of the output phase, loaded while the FFT runs (FP64 B=256
    // x 1024: 1.85 -> 1.75 ms; the same for the forward's phase factors lost 3-6%)
    R sv[V];
#pragma unroll
    for (int u = 0; u < V; ++u) sv[u] = sinv[(threadIdx.x + u * NTH) & (N - 1)];
    __syncthreads();
    fft_block<T, +1, V, NTH>(a, lgN, S, twl);
    const R inv_s2pi = (R)0.39894228040143267794;
    // FP32 (V = 16) unrolled by 4 only: a full unroll keeps every element's
    // value and store addresses live (162 VGPRs, one block per CU instead of two)
    constexpr int UNR = sizeof(R) == 4 ? 4 : V;
#pragma unroll UNR
    for (int u = 0; u < V; ++u) {
        const int idx = threadIdx.x + u * NTH;
        const int rl = idx >> lgN, j = idx & (N - 1);
        const int q = q0 + rl;
        const int n = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
        const T yv = a[rl * S + n];
        if (q >= nrows) continue;
        const int r = frow_of(q, N, cut);
        const int am = r < B ? r : N - r;
        R sc = inv_s2pi;
        if (am & 1) sc *= sv[u];
        if (r > B && (r & 1)) sc = -sc;
        const T y = cscale(yv, sc);
        X[((size_t)f * N + r) * N + j] = y;
        if (fmt == 1 && r > 0) X[((size_t)f * N + (N - r)) * N + j] = cconj(y);
    }
}

// W16: each thread stores two adjacent samples as one 16-byte store per
// plane (re and im must be 16-byte aligned).
template <typename R, int V, int NTH = kThreads, bool W16 = false>
__global__ __launch_bounds__(NTH) void k_fphi_inv(const typename Cplx<R>::type* __restrict__ X,
                                                       double* __restrict__ re, double* __restrict__ im, int lgN,
                                                       const typename Cplx<R>::type* __restrict__ tw) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int E = V * NTH;
    const int N = 1 << lgN, S = N + 1, B = N >> 1;
    const T* twl = reinterpret_cast<const T*>(smem);
    T* a = reinterpret_cast<T*>(smem) + N;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int j0 = bx * rpb;
    // twiddles behind the loads for FP64 and the 16-byte-store instance; the
    // FP32 8-byte-store instance stages them first (behind: bw=512 x 64
    // 0.35 -> 0.39 ms)
    constexpr bool kLate = sizeof(R) == 8 || W16;
    if constexpr (!kLate) stage_twiddles<T, NTH>(smem, tw, N);
    {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            const int r = idx >> lgrpb;
            v[u] = (r == B) ? mk<T>(0, 0) : X[((size_t)f * N + r) * N + j0 + (idx & (rpb - 1))];
        }
        if constexpr (kLate) stage_twiddles<T, NTH>(smem, tw, N);
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            a[(idx & (rpb - 1)) * S + (idx >> lgrpb)] = v[u];
        }
    }
    __syncthreads();
    fft_block<T, +1, V, NTH>(a, lgN, S, twl);
    const size_t base = ((size_t)f * N + j0) * N;
    if constexpr (W16) {  // sample pairs (2p, 2p+1) of one ring (N even)
        T v[V];
#pragma unroll
        for (int u = 0; u < V / 2; ++u) {
            const int idx = 2 * (threadIdx.x + u * NTH);
            const T* row = a + (idx >> lgN) * S + (idx & (N - 1));
            v[2 * u] = row[0];
            v[2 * u + 1] = row[1];
        }
#pragma unroll
        for (int u = 0; u < V / 2; ++u) {
            const size_t o = base + 2 * (threadIdx.x + u * NTH);
            *reinterpret_cast<double2*>(re + o) = make_double2((double)v[2 * u].x, (double)v[2 * u + 1].x);
            *reinterpret_cast<double2*>(im + o) = make_double2((double)v[2 * u].y, (double)v[2 * u + 1].y);
        }
    } else {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * NTH;
            v[u] = a[(idx >> lgN) * S + (idx & (N - 1))];
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            re[base + threadIdx.x + u * NTH] = (double)v[u].x;
            im[base + threadIdx.x + u * NTH] = (double)v[u].y;
        }
    }
}

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

template <typename R>
struct Mma;
template <>
struct Mma<double> {
    using acc = dvec4;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the r-th accumulator of `lane` in the 16x16 result
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mma<float> {
    using acc = fvec4;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

inline int lg2(int n) { return __builtin_ctz(n); }

// Calls of few fields: the default ring-FFT/DCT blocks (512 threads, 8 FP64 or
// 16 FP32 rings) would give fewer than 256 blocks, one per CU (one field at
// bw=256: 64), so one-wave blocks (64 threads; 128 for the longest rings)
// take one or two rings each: eight times the blocks, each block's passes
// synchronised within one or two waves.  Measured per fwd+inv pair of one
// field: bw=64 46.9 -> 31.7 us, 128 51.7 -> 37.8, 256 72.1 -> 64.2; at four
// fields (256 default blocks) the one-wave blocks were slower (92.9 -> 100.3
// us at bw=256).  Returns the block's threads, 0 for the default blocks.
template <int V>
inline int few_field_threads(int N, int F) {
    if ((long)F * N * N >= 512L * V * 256) return 0;
    if (N <= 64 * V) return 64;
    if (N <= 128 * V) return 128;
    return 0;
}

// Launch with `lds` bytes of dynamic LDS; above 64 KiB the kernel must opt
// in (gfx950 has 160 KiB per workgroup).
template <typename K, typename... A>
inline void launch_lds(K kernel, dim3 grid, int threads, size_t lds, hipStream_t s, A... args) {
    if (lds > 65536) S2_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    kernel<<<grid, threads, lds, s>>>(args...);
    S2_HIP(hipGetLastError());
}

}  // namespace s2
