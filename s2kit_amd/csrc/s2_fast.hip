// Batched fast path for power-of-two grids (N = 2*bw = 16 .. 2048), in FP64
// or FP32 arithmetic (FP64 inputs and outputs either way).
//
//   k_fphi_fwd   ring FFT along phi, order-major store          (FST_semi_memo.c:81-90)
//   k_fdct_fwd   weights + DCT-II per order row                 (seminaive.c:161-176)
//   k_gemm_fwd   seminaive projections as MFMA GEMMs            (seminaive.c:183-197)
//   k_gemm_inv   transposed-table cosine sums as MFMA GEMMs     (seminaive.c:74-98)
//   k_fdct_inv   DCT-III, x sin, row signs, 1/sqrt(2pi)         (seminaive.c:107-114; FST_semi_memo.c:301-348)
//   k_fphi_inv   inverse ring FFT                               (FST_semi_memo.c:350)
//
// FFTs: in-place Stockham passes of radix 8 (then 4 or 2) with the values of
// each butterfly in registers and one padded LDS buffer (row stride N+1).
//
// Legendre GEMMs.  For order m split the degrees by parity p = (l-m) mod 2:
// degree l = m+p+2i reads cosine coefficients c[2t+p], t < RowSize(m,l), so
// each (m,p) is a dense matrix A[i][t] (zero above the staircase).  The
// stripped tables are repacked once per plan (k_pack_gemm) into
//   AF(m,p)  [t][i]  (t < roundup4(Kf), i < roundup16(Lp))   forward
//   AI(m,p)  [i][t]  (i < roundup4(Lp), t < roundup16(bw/2)) inverse, built
//            from the transposed table with its scaling folded in
// and the data columns are (field, +-m, re/im).  Per 16x16 tile the K loop
// stops at the staircase, so the padded zeros cost at most one tile row.
#include <algorithm>
#include <cmath>
#include <vector>

#include "s2_complex.h"
#include "s2_internal.h"
#include "s2_recurrence.h"

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

// One Stockham pass of radix R over the E = V*kThreads values of the block
// (E/N transforms of length N, row stride S).  Sub-transform size Ns = 2^lgNs.
template <typename T, int SIGN, int V, int R, int LGR>
__device__ __forceinline__ void fft_pass(T* a, int lgN, int S, int lgNs, const T* tw) {
    constexpr int PER = V / R;
    const int lgNR = lgN - LGR;
    const int NR = 1 << lgNR;
    const int Ns = 1 << lgNs;
    T x[V];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int b = threadIdx.x + u * kThreads;
        const int f = b >> lgNR, j = b & (NR - 1);
        const T* src = a + f * S + j;
#pragma unroll
        for (int r = 0; r < R; ++r) x[u * R + r] = src[r * NR];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int b = threadIdx.x + u * kThreads;
        const int f = b >> lgNR, j = b & (NR - 1);
        const int k = j & (Ns - 1);
        const int step = k << (lgNR - lgNs);  // k * N / (R * Ns)
        if (lgNs > 0) {  // the first pass has k = 0: unit twiddles
#pragma unroll
            for (int r = 1; r < R; ++r) {
                T w = tw[r * step];
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
template <typename T, int SIGN, int V>
__device__ __forceinline__ void fft_block(T* a, int lgN, int S, const T* tw) {
    int lgNs = 0, rem = lgN;
    while (rem >= 3) {
        fft_pass<T, SIGN, V, 8, 3>(a, lgN, S, lgNs, tw);
        lgNs += 3;
        rem -= 3;
    }
    if (rem == 2)
        fft_pass<T, SIGN, V, 4, 2>(a, lgN, S, lgNs, tw);
    else if (rem == 1)
        fft_pass<T, SIGN, V, 2, 1>(a, lgN, S, lgNs, tw);
}

__device__ __forceinline__ int frow_of(int q, int N, int cut) { return q < cut ? q : N - 2 * cut + 1 + q; }

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
// butterfly of every pass).
template <typename T>
__device__ __forceinline__ T* stage_twiddles(unsigned char* smem, const T* __restrict__ tw, int N) {
    T* t = reinterpret_cast<T*>(smem);
    for (int i = threadIdx.x; i < N; i += kThreads) t[i] = tw[i];
    return t;
}

// ---------------------------------------------------------------------------
// FFT-type kernels.  grid.y = field, block = kThreads, LDS = rows*(N+1) values
// ---------------------------------------------------------------------------
template <typename R, int V>
__global__ __launch_bounds__(kThreads) void k_fphi_fwd(const double* __restrict__ re, const double* __restrict__ im,
                                                       typename Cplx<R>::type* __restrict__ X, int lgN,
                                                       const typename Cplx<R>::type* __restrict__ tw) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1;
    const T* twl = stage_twiddles(smem, tw, N);
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
            vr[u] = re[base + threadIdx.x + u * kThreads];
            vi[u] = im[base + threadIdx.x + u * kThreads];
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * kThreads;
            a[(idx >> lgN) * S + (idx & (N - 1))] = mk<T>((R)vr[u], (R)vi[u]);
        }
    }
    __syncthreads();
    fft_block<T, -1, V>(a, lgN, S, twl);
    {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * kThreads;
            v[u] = a[(idx & (rpb - 1)) * S + (idx >> lgrpb)];
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * kThreads;
            X[((size_t)f * N + (idx >> lgrpb)) * N + j0 + (idx & (rpb - 1))] = v[u];
        }
    }
}

template <typename R, int V>
__global__ __launch_bounds__(kThreads) void k_fdct_fwd(const typename Cplx<R>::type* __restrict__ X, R* __restrict__ Cb,
                                                       const R* __restrict__ w_semi, int lgN, int cut, int nrows,
                                                       const typename Cplx<R>::type* __restrict__ tw,
                                                       const typename Cplx<R>::type* __restrict__ ph) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const T* twl = stage_twiddles(smem, tw, 1 << lgN);
    T* a = reinterpret_cast<T*>(smem) + (1 << lgN);
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1, B = N >> 1, half = B >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int q0 = bx * rpb;
    for (int idx = threadIdx.x; idx < E; idx += kThreads) {
        const int rl = idx >> lgN, j = idx & (N - 1);
        const int q = q0 + rl;
        const int p = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
        T x = mk<T>(0, 0);
        if (q < nrows) {
            const int r = frow_of(q, N, cut);
            const int am = r < B ? r : N - r;
            x = cscale(X[((size_t)f * N + r) * N + j], w_semi[(am & 1) * N + j]);
        }
        a[rl * S + p] = x;
    }
    __syncthreads();
    fft_block<T, -1, V>(a, lgN, S, twl);
    for (int idx = threadIdx.x; idx < rpb * B; idx += kThreads) {
        const int rl = idx >> (lgN - 1), k = idx & (B - 1);
        const int q = q0 + rl;
        if (q >= nrows) continue;
        const int r = frow_of(q, N, cut);
        const T* v = a + rl * S;
        const T e = ph[k];
        T c = cadd(cmul(e, v[k]), cmul(cconj(e), v[(N - k) & (N - 1)]));
        if (k == 0) c = cscale(c, (R)0.70710678118654752440);
        R* row = Cb + ((size_t)f * N + r) * 2 * B;
        const int pos = cpos(k, half);
        row[pos] = c.x;
        row[B + pos] = c.y;
    }
}

template <typename R, int V>
__global__ __launch_bounds__(kThreads) void k_fdct_inv(const R* __restrict__ Cb, typename Cplx<R>::type* __restrict__ X,
                                                       const R* __restrict__ sinv, int lgN, int cut, int nrows,
                                                       int fmt, const typename Cplx<R>::type* __restrict__ tw,
                                                       const typename Cplx<R>::type* __restrict__ ph) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const T* twl = stage_twiddles(smem, tw, 1 << lgN);
    T* a = reinterpret_cast<T*>(smem) + (1 << lgN);
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1, B = N >> 1, half = B >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int q0 = bx * rpb;
    // V[k] = e^{+i pi k/2N} (X[k] - i X[N-k]),  X[N] = 0, X[k >= B] = 0
    for (int idx = threadIdx.x; idx < E; idx += kThreads) {
        const int rl = idx >> lgN, k = idx & (N - 1);
        const int q = q0 + rl;
        T v = mk<T>(0, 0);
        if (q < nrows) {
            const int r = frow_of(q, N, cut);
            const R* row = Cb + ((size_t)f * N + r) * 2 * B;
            T xk = mk<T>(0, 0), xn = mk<T>(0, 0);
            if (k < B) xk = mk<T>(row[cpos(k, half)], row[B + cpos(k, half)]);
            const int kn = N - k;
            if (k > 0 && kn < B) xn = mk<T>(row[cpos(kn, half)], row[B + cpos(kn, half)]);
            v = cmul(cconj(ph[k]), mk<T>(xk.x + xn.y, xk.y - xn.x));
        }
        a[rl * S + k] = v;
    }
    __syncthreads();
    fft_block<T, +1, V>(a, lgN, S, twl);
    const R inv_s2pi = (R)0.39894228040143267794;
    for (int idx = threadIdx.x; idx < E; idx += kThreads) {
        const int rl = idx >> lgN, j = idx & (N - 1);
        const int q = q0 + rl;
        if (q >= nrows) continue;
        const int r = frow_of(q, N, cut);
        const int am = r < B ? r : N - r;
        const int n = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
        R sc = inv_s2pi;
        if (am & 1) sc *= sinv[j];
        if (r > B && (r & 1)) sc = -sc;
        const T y = cscale(a[rl * S + n], sc);
        X[((size_t)f * N + r) * N + j] = y;
        if (fmt == 1 && r > 0) X[((size_t)f * N + (N - r)) * N + j] = cconj(y);
    }
}

template <typename R, int V>
__global__ __launch_bounds__(kThreads) void k_fphi_inv(const typename Cplx<R>::type* __restrict__ X,
                                                       double* __restrict__ re, double* __restrict__ im, int lgN,
                                                       const typename Cplx<R>::type* __restrict__ tw) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1, B = N >> 1;
    const T* twl = stage_twiddles(smem, tw, N);
    T* a = reinterpret_cast<T*>(smem) + N;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    int bx, f;
    xcd_block(bx, f);
    const int j0 = bx * rpb;
    {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * kThreads;
            const int r = idx >> lgrpb;
            v[u] = (r == B) ? mk<T>(0, 0) : X[((size_t)f * N + r) * N + j0 + (idx & (rpb - 1))];
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * kThreads;
            a[(idx & (rpb - 1)) * S + (idx >> lgrpb)] = v[u];
        }
    }
    __syncthreads();
    fft_block<T, +1, V>(a, lgN, S, twl);
    const size_t base = ((size_t)f * N + j0) * N;
    {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int idx = threadIdx.x + u * kThreads;
            v[u] = a[(idx >> lgN) * S + (idx & (N - 1))];
        }
#pragma unroll
        for (int u = 0; u < V; ++u) {
            re[base + threadIdx.x + u * kThreads] = (double)v[u].x;
            im[base + threadIdx.x + u * kThreads] = (double)v[u].y;
        }
    }
}

// ---------------------------------------------------------------------------
// GEMM tables
// ---------------------------------------------------------------------------
__host__ __device__ inline int gm_lp(int m, int p, int B) { return (B - m - p > 0) ? (B - m - p + 1) / 2 : 0; }

template <typename R>
__global__ void k_pack_gemm(const double* __restrict__ T, const double* __restrict__ TT, const long* __restrict__ toff,
                            const GemmMeta* __restrict__ meta, R* __restrict__ AF, R* __restrict__ AI, int B) {
    extern __shared__ __align__(16) unsigned char smem[];
    int* rowpos = reinterpret_cast<int*>(smem);
    const int m = blockIdx.x, p = blockIdx.y;
    const int N = 2 * B, half = B / 2;
    const GemmMeta g = meta[m * 2 + p];
    for (int i = threadIdx.x; i < B; i += blockDim.x) rowpos[i + 1] = transpose_row_size(i, m, B);
    __syncthreads();
    if (threadIdx.x == 0) {
        rowpos[0] = 0;
        for (int i = 1; i <= B; ++i) rowpos[i] += rowpos[i - 1];
    }
    __syncthreads();
    const int Li16 = (g.Lp + 15) & ~15, Kf4 = (g.Kf + 3) & ~3;
    const int Li4 = (g.Lp + 3) & ~3, Tp16 = (half + 15) & ~15;
    const double* tm = T + toff[m];
    const double* ttm = TT + toff[m];
    for (int idx = threadIdx.x; idx < Kf4 * Li16; idx += blockDim.x) {
        const int t = idx / Li16, i = idx - t * Li16;
        const int l = m + p + 2 * i;
        double v = 0.0;
        if (i < g.Lp && t < row_size(m, l)) v = tm[table_offset(m, l) + t];
        AF[g.offF + idx] = (R)v;
    }
    const double s0 = 1.0 / sqrt((double)N), s1 = 0.5 / sqrt((double)B);
    for (int idx = threadIdx.x; idx < Li4 * Tp16; idx += blockDim.x) {
        const int i = idx / Tp16, t = idx - i * Tp16;
        double v = 0.0;
        if (i < g.Lp && t < half) {
            const int ci = 2 * t + p;
            const int l = m + p + 2 * i;
            const int l0 = m + transpose_first(ci, m);
            const int rs = rowpos[ci + 1] - rowpos[ci];
            const int d = l - l0;
            if (d >= 0 && !(d & 1) && (d >> 1) < rs) v = ttm[rowpos[ci] + (d >> 1)] * (ci == 0 ? s0 : s1);
        }
        AI[g.offI + idx] = (R)v;
    }
}

// ---------------------------------------------------------------------------
// MFMA GEMMs
// ---------------------------------------------------------------------------
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

// LDS row padding: a row stride of bw + 4 words (FP32) / bw + 2 doubles (FP64)
// keeps both the coalesced fill (consecutive k) and the MFMA operand reads
// (16 columns x 4 consecutive k per wave) free of bank conflicts.
template <typename R>
__host__ __device__ constexpr int lds_pad() {
    return sizeof(R) == 4 ? 4 : 2;
}

// Forward: D'[c][i] = sum_t data[c][t] * A(m,p)[i][t].  grid (ceil(cols/ncb), bw).
// LDS: Bd[c][k2], k2 = parity-split cosine position (p*bw/2 + t).
template <typename R>
__global__ __launch_bounds__(kThreads) void k_gemm_fwd(const R* __restrict__ Cb, const R* __restrict__ AF,
                                                       const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                       double* __restrict__ cim, int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* Bd = reinterpret_cast<R*>(smem);
    const int B = 1 << lgB, N = 2 * B, half = B >> 1;
    int bx, m;
    xcd_block_m(bx, m);
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = bx * ncb;
    if (cbase >= ncols) return;
    const int rs = B + lds_pad<R>();
    for (int idx = threadIdx.x; idx < ncb * B; idx += kThreads) {
        const int c = idx >> lgB, k2 = idx & (B - 1);
        const int col = cbase + c;
        R v = 0;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            const int r = (rem >> 1) ? N - m : m;
            v = Cb[(((size_t)f * N + r) * 2 + (rem & 1)) * B + k2];
        }
        Bd[c * rs + k2] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = ncb >> 4;
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = meta[m * 2 + p];
        if (g.Lp == 0) continue;
        const int Li16 = (g.Lp + 15) & ~15;
        const int ntiles = (Li16 >> 4) * nct;
        for (int item = wave; item < ntiles; item += kThreads / 64) {
            const int it = item / nct, ct = item - it * nct;
            const int i0 = it * 16, c0 = ct * 16;
            const int ilast = min(i0 + 15, g.Lp - 1);
            const int kend = (row_size(m, m + p + 2 * ilast) + 3) & ~3;
            typename Mma<R>::acc acc = {0, 0, 0, 0};
            const R* bcol = Bd + (c0 + (lane & 15)) * rs + p * half;
            const R* arow = AF + g.offF + i0 + (lane & 15);
            for (int t0 = 0; t0 < kend; t0 += 4) {
                const int tk = t0 + (lane >> 4);
                acc = Mma<R>::mma(bcol[tk], arow[(size_t)tk * Li16], acc);
            }
            const int i = i0 + (lane & 15);
            if (i >= g.Lp) continue;
            const int l = m + p + 2 * i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = cbase + c0 + Mma<R>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf;
                const int comp = rem & 1;
                double* out = comp ? cim : cre;
                const double v = (double)acc[r];
                if ((rem >> 1) == 0) {
                    out[f * B2 + index_of_coeff(m, l, B)] = v;
                    if (fmt == 1 && m > 0) out[f * B2 + index_of_coeff(-m, l, B)] = comp ? -sg * v : sg * v;
                } else {
                    out[f * B2 + index_of_coeff(-m, l, B)] = sg * v;
                }
            }
        }
    }
}

// Inverse: D'[c][t] = sum_i coef[c][i] * AI(m,p)[i][t].  grid (ceil(cols/ncb), bw).
// LDS: Cd[c][q], q = i (p = 0) or roundup4(Lp0) + i (p = 1); q < bw always.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_gemm_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                       const R* __restrict__ AI, const GemmMeta* __restrict__ meta,
                                                       R* __restrict__ Cb, int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* Cd = reinterpret_cast<R*>(smem);
    const int B = 1 << lgB, N = 2 * B, half = B >> 1;
    int bx, m;
    xcd_block_m(bx, m);
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = bx * ncb;
    if (cbase >= ncols) return;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int Li4_0 = (g0.Lp + 3) & ~3, Li4_1 = (g1.Lp + 3) & ~3;
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    const int rs = B + lds_pad<R>();
    for (int idx = threadIdx.x; idx < ncb * L; idx += kThreads) {
        const int c = idx / L, li = idx - c * L;
        const int col = cbase + c;
        const int q = (li & 1) ? Li4_0 + (li >> 1) : (li >> 1);
        R v = 0;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            const int mm = (rem >> 1) ? -m : m;
            const double* src = (rem & 1) ? cim : cre;
            v = (R)src[f * B2 + index_of_coeff(mm, m + li, B)];
        }
        Cd[c * rs + q] = v;
    }
    // zero the padding rows of both parities (at most 3 + 3 per column)
    for (int idx = threadIdx.x; idx < ncb * 8; idx += kThreads) {
        const int c = idx >> 3, k = idx & 7;
        const int p = k >> 2, i = (p ? g1.Lp : g0.Lp) + (k & 3);
        if (i < (p ? Li4_1 : Li4_0)) Cd[c * rs + (p ? Li4_0 : 0) + i] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = ncb >> 4;
    const int Tp16 = (half + 15) & ~15;
    const int ntt = Tp16 >> 4;
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = p ? g1 : g0;
        const int Li4 = p ? Li4_1 : Li4_0;
        const int r0 = g.Lp > 0 ? row_size(m, m + p) : 0;
        for (int item = wave; item < ntt * nct; item += kThreads / 64) {
            const int tt = item / nct, ct = item - tt * nct;
            const int t0 = tt * 16, c0 = ct * 16;
            const int istart = max(0, t0 - r0 + 1) & ~3;
            typename Mma<R>::acc acc = {0, 0, 0, 0};
            const R* acol = Cd + (c0 + (lane & 15)) * rs + (p ? Li4_0 : 0);
            const R* brow = AI + g.offI + t0 + (lane & 15);
            for (int i0 = istart; i0 < Li4; i0 += 4) {
                const int ik = i0 + (lane >> 4);
                acc = Mma<R>::mma(acol[ik], brow[(size_t)ik * Tp16], acc);
            }
            const int t = t0 + (lane & 15);
            if (t >= half) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = cbase + c0 + Mma<R>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf;
                const int row = (rem >> 1) ? N - m : m;
                Cb[(((size_t)f * N + row) * 2 + (rem & 1)) * B + p * half + t] = acc[r];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed form: naive DLT (naive.c:34-95) as FP32 MFMA GEMMs on the phi-FFT rows
// ---------------------------------------------------------------------------
// P_l^m(theta_{N-1-j}) = (-1)^(l-m) P_l^m(theta_j) and the quadrature weights
// are symmetric, so each order folds the ring pair (j, N-1-j): degrees with
// l-m even see x_j + x_{N-1-j}, odd ones x_j - x_{N-1-j}, over j < bw.
//   NF(m,p) [j][i]  P_{m+p+2i}^m(theta_j) w_j sqrt(2pi)/N  (j < bw, i < roundup64(Lp))
//   NI(m,p) [i][j]  P_{m+p+2i}^m(theta_j)                  (i < roundup4(Lp), j < bw)
// Both come from the reference's FP64 recurrence (pml.c:41-79), rounded once.
// Columns of both are interleaved in groups of 64: entry 16q + x of a group
// sits at 4x + q, so one float4 per lane feeds four 16-wide MFMA tiles.
__host__ __device__ inline int ilv64(int i) { return (i & ~63) + 4 * (i & 15) + ((i >> 4) & 3); }

// XOR swizzle of an LDS row index: 16 consecutive columns reading the same
// k land in 16 different 4-bank groups (bw is a multiple of 64 here).
__device__ __forceinline__ int swz(int c) { return (c & 15) << 2; }

__global__ void k_pack_naive(const double* __restrict__ w, const GemmMeta* __restrict__ meta, float* __restrict__ NF,
                             float* __restrict__ NI, int B) {
    const int m = blockIdx.x;
    const int N = 2 * B;
    const double cm = (m == 0) ? 0.0 : pmm_coeff(m);
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int L64_0 = (g0.Lp + 63) & ~63, L64_1 = (g1.Lp + 63) & ~63;
    for (int j = threadIdx.x; j < B; j += blockDim.x) {
        const double th = (2.0 * j + 1.0) * M_PI / (2.0 * N);
        const double x = cos((2.0 * j + 1.0) * M_PI / (2.0 * N));
        const double wj = w[j];  // even weight times sqrt(2pi)/N
        double cur = (m == 0) ? 0.70710678118654752440 : cm * pow(sin(th), (double)m);
        double prev = 0.0;
        for (int l = m; l < B; ++l) {
            const int p = (l - m) & 1, i = (l - m) >> 1;
            const GemmMeta& g = p ? g1 : g0;
            NF[g.offF + (long)j * (p ? L64_1 : L64_0) + ilv64(i)] = (float)(cur * wj);
            NI[g.offI + (long)i * B + ilv64(j)] = (float)cur;
            if (l + 1 < B) {
                const double nx = l2_an(m, l) * (cur * x) + l2_cn(m, l) * prev;
                prev = cur;
                cur = nx;
            }
        }
    }
}

// Forward: coefficient[c][i] = sum_j fold_p(x)[c][j] * NF(m,p)[j][i].
// grid (bw, ceil(cols/ncb)); LDS Sd[c][p][j] (swizzled j), 2*bw floats per
// column.  A wave owns 16 columns x 64 degrees (four MFMA tiles).
__global__ __launch_bounds__(kThreads) void k_ngemm_fwd(const float2* __restrict__ X, const float* __restrict__ NF,
                                                        const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                        double* __restrict__ cim, int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* Sd = reinterpret_cast<float*>(smem);
    const int B = 1 << lgB, N = 2 * B;
    const int m = blockIdx.x;  // orders fastest: concurrent blocks stream the same fields' rows
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * ncb;
    if (cbase >= ncols) return;
    // one (field, sign) row feeds two columns (re, im)
    for (int idx = threadIdx.x; idx < (ncb >> 1) * B; idx += kThreads) {
        const int pr = idx >> lgB, j = idx & (B - 1);
        const int c = 2 * pr, col = cbase + c;
        float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            const int r = (rem >> 1) ? N - m : m;
            const float2* row = X + ((size_t)f * N + r) * N;
            a = row[j];
            b = row[N - 1 - j];
        }
        const int j0 = j ^ swz(c), j1 = j ^ swz(c + 1);
        float* d0 = Sd + (size_t)c * 2 * B;
        float* d1 = d0 + 2 * B;
        d0[j0] = a.x + b.x;
        d1[j1] = a.y + b.y;
        d0[B + j0] = a.x - b.x;
        d1[B + j1] = a.y - b.y;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = ncb >> 4;
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int ng0 = (g0.Lp + 63) >> 6, ng1 = (g1.Lp + 63) >> 6;
    for (int item = wave; item < (ng0 + ng1) * nct; item += kThreads / 64) {
        const int gi = item / nct, ct = item - gi * nct;
        const int p = gi >= ng0;
        const GemmMeta g = p ? g1 : g0;
        const int ig = p ? gi - ng0 : gi;
        const int L64 = ((p ? g1.Lp : g0.Lp) + 63) & ~63;
        const int c0 = ct * 16, c = c0 + (lane & 15);
        const float* arow = Sd + (size_t)c * 2 * B + p * B;
        const int sc = swz(c);
        const float* brow = NF + g.offF + ig * 64 + 4 * (lane & 15);
        fvec4 a0[4] = {}, a1[4] = {};
        // table fragments are fetched one step ahead (the last fetch repeats)
        const int kq = lane >> 4;
        float4 ta = *reinterpret_cast<const float4*>(brow + (size_t)kq * L64);
        float4 tb = *reinterpret_cast<const float4*>(brow + (size_t)(kq + 4) * L64);
        for (int j0 = 0; j0 < B; j0 += 8) {
            const int ja = j0 + kq, jb = ja + 4;
            const int jn = min(j0 + 8, B - 8) + kq;
            const float4 na = *reinterpret_cast<const float4*>(brow + (size_t)jn * L64);
            const float4 nb = *reinterpret_cast<const float4*>(brow + (size_t)(jn + 4) * L64);
            const float xa = arow[ja ^ sc], xb = arow[jb ^ sc];
            a0[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.x, a0[0], 0, 0, 0);
            a0[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.y, a0[1], 0, 0, 0);
            a0[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.z, a0[2], 0, 0, 0);
            a0[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.w, a0[3], 0, 0, 0);
            a1[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.x, a1[0], 0, 0, 0);
            a1[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.y, a1[1], 0, 0, 0);
            a1[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.z, a1[2], 0, 0, 0);
            a1[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.w, a1[3], 0, 0, 0);
            ta = na;
            tb = nb;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = ig * 64 + 16 * q + (lane & 15);
            if (i >= g.Lp) continue;
            const int l = m + p + 2 * i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = cbase + c0 + Mma<float>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf;
                const int comp = rem & 1;
                double* out = comp ? cim : cre;
                const double v = (double)a0[q][r] + (double)a1[q][r];
                if ((rem >> 1) == 0) {
                    out[f * B2 + index_of_coeff(m, l, B)] = v;
                    if (fmt == 1 && m > 0) out[f * B2 + index_of_coeff(-m, l, B)] = comp ? -sg * v : sg * v;
                } else {
                    out[f * B2 + index_of_coeff(-m, l, B)] = sg * v;
                }
            }
        }
    }
}

// Forward, streamed form: a block owns TM = 128 columns (32 fields x +-m x
// re/im) of one order and all its degrees, so each order's table is read
// 4F/128 times instead of once per 32 columns.  The folded data goes through
// LDS in double-buffered chunks of KC = 32 rings: the next chunk's global
// loads are in flight while the current chunk feeds the MFMAs.  8 waves, each
// owning 16 columns x every 64-degree group of both parities (G groups per
// parity at most).  LDS chunk layout [p][c][k], k swizzled by ((c>>1)&7)<<2 so
// that a wave's 16 columns x 4 rings hit 64 distinct banks.
constexpr int kNgThreads = 512;
constexpr int kNgTM = 128;
constexpr int kNgKC = 32;

__device__ __forceinline__ int swz32(int c) { return ((c >> 1) & 7) << 2; }

template <int G>
__global__ __launch_bounds__(kNgThreads) void k_ngemm_fwd2(const float2* __restrict__ X, const float* __restrict__ NF,
                                                           const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                           double* __restrict__ cim, int lgB, int F, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* sm = reinterpret_cast<float*>(smem);  // [2 buffers][2 parities][TM][KC]
    constexpr int TM = kNgTM, KC = kNgKC, BUF = 2 * TM * KC;
    const int B = 1 << lgB, N = 2 * B;
    const int m = blockIdx.x;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * TM;
    if (cbase >= ncols) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int ng[2] = {(g0.Lp + 63) >> 6, (g1.Lp + 63) >> 6};
    const int L64[2] = {ng[0] * 64, ng[1] * 64};

    // fill role of this thread: (field, sign) pairs fp and fp+32 (columns
    // 2fp, 2fp+1 and 2fp+64, 2fp+65), rings fk and fk+16 of each chunk
    const int fk = tid & 15, fp = tid >> 4;  // 32 x 16 threads
    const float2* frow[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int col = cbase + 2 * (fp + 32 * e);
        frow[e] = nullptr;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            frow[e] = X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N;
        }
    }
    float2 ra[2][2], rb[2][2];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int j = kc + fk + 16 * h;
                ra[e][h] = frow[e] ? frow[e][j] : make_float2(0.f, 0.f);
                rb[e][h] = frow[e] ? frow[e][N - 1 - j] : make_float2(0.f, 0.f);
            }
    };
    auto stash = [&](float* buf) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int c = 2 * (fp + 32 * e);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = fk + 16 * h;
                const int k0 = k ^ swz32(c), k1 = k ^ swz32(c + 1);
                buf[(0 * TM + c) * KC + k0] = ra[e][h].x + rb[e][h].x;
                buf[(0 * TM + c + 1) * KC + k1] = ra[e][h].y + rb[e][h].y;
                buf[(1 * TM + c) * KC + k0] = ra[e][h].x - rb[e][h].x;
                buf[(1 * TM + c + 1) * KC + k1] = ra[e][h].y - rb[e][h].y;
            }
        }
    };

    const int c = wave * 16 + (lane & 15);  // this lane's column in MFMA operand A
    const int sc = swz32(c), kq = lane >> 4;
    fvec4 acc[2][G][4] = {};
    fetch(0);
    stash(sm);
    __syncthreads();
    for (int kc = 0; kc < B; kc += KC) {
        const float* cur = sm + ((kc / KC) & 1) * BUF;
        const bool more = kc + KC < B;
        if (more) fetch(kc + KC);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const GemmMeta g = p ? g1 : g0;
            const float* arow = cur + (p * TM + c) * KC;
            const float* brow = NF + g.offF + 4 * (lane & 15);
#pragma unroll
            for (int kk = 0; kk < KC; kk += 4) {
                const float a = arow[(kk + kq) ^ sc];
                const size_t jrow = (size_t)(kc + kk + kq) * L64[p];
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    if (q < ng[p]) {
                        const float4 t = *reinterpret_cast<const float4*>(brow + jrow + q * 64);
                        acc[p][q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.x, acc[p][q][0], 0, 0, 0);
                        acc[p][q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.y, acc[p][q][1], 0, 0, 0);
                        acc[p][q][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.z, acc[p][q][2], 0, 0, 0);
                        acc[p][q][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.w, acc[p][q][3], 0, 0, 0);
                    }
                }
            }
        }
        if (more) stash(sm + (((kc / KC) + 1) & 1) * BUF);
        __syncthreads();
    }
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int Lp = p ? g1.Lp : g0.Lp;
#pragma unroll
        for (int q = 0; q < G; ++q) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = q * 64 + 16 * u + (lane & 15);
                if (q >= ng[p] || i >= Lp) continue;
                const int l = m + p + 2 * i;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + wave * 16 + Mma<float>::row(lane, r);
                    if (col >= ncols) continue;
                    const int f = col / cpf, rem = col - f * cpf;
                    const int comp = rem & 1;
                    double* out = comp ? cim : cre;
                    const double v = (double)acc[p][q][u][r];
                    if ((rem >> 1) == 0) {
                        out[f * B2 + index_of_coeff(m, l, B)] = v;
                        if (fmt == 1 && m > 0) out[f * B2 + index_of_coeff(-m, l, B)] = comp ? -sg * v : sg * v;
                    } else {
                        out[f * B2 + index_of_coeff(-m, l, B)] = sg * v;
                    }
                }
            }
        }
    }
}

// Forward, pipelined form.  A block owns TM = 128 columns of one order m and
// all its degrees.  Both operands go through LDS in double-buffered chunks of
// KC = 16 rings: the folded data A[p][c][k] and the table rows T[k][p][G*64].
// The next chunk's global loads are issued before the current chunk's MFMAs
// and stored after them, so the MFMA loop only waits on LDS.  G (64-degree
// groups per parity) is a template constant: orders are launched in buckets
// of equal G, so the loop has no branches (a parity with one group fewer
// reads a zero-filled group).  8 waves, each 16 columns x 2 x G x 4 tiles.
constexpr int kF3Threads = 512;
constexpr int kF3TM = 128;
constexpr int kF3KC = 16;

__device__ __forceinline__ int swz16(int c) { return ((c >> 2) & 3) << 2; }

template <int G>
__global__ __launch_bounds__(kF3Threads) void k_ngemm_fwd3(const float2* __restrict__ X, const float* __restrict__ NF,
                                                           const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                           double* __restrict__ cim, int lgB, int F, int fmt, int m0) {
    constexpr int TM = kF3TM, KC = kF3KC, GW = G * 64;
    constexpr int ABUF = 2 * TM * KC;  // floats of folded data per buffer
    constexpr int TBUF = KC * 2 * GW;  // floats of table per buffer
    constexpr int RW = 2 * GW / 4;     // float4 per table chunk row (both parities)
    extern __shared__ __align__(16) unsigned char smem[];
    float* sm = reinterpret_cast<float*>(smem);
    const int B = 1 << lgB, N = 2 * B;
    const int m = m0 + blockIdx.x;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * TM;
    if (cbase >= ncols) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GemmMeta gm[2] = {meta[m * 2], meta[m * 2 + 1]};
    const int L64[2] = {(gm[0].Lp + 63) & ~63, (gm[1].Lp + 63) & ~63};

    // data fill: pair ap (columns 2ap, 2ap+1), rings k0 and k0+8 of a chunk
    const int ap = tid >> 3, k0 = tid & 7;
    const float2* xrow = nullptr;
    {
        const int col = cbase + 2 * ap;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            xrow = X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N;
        }
    }
    float2 xa[2], xb[2];
    float4 tv[G];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int j = kc + k0 + 8 * h;
            xa[h] = xrow ? xrow[j] : make_float2(0.f, 0.f);
            xb[h] = xrow ? xrow[N - 1 - j] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int e = tid + kF3Threads * u;
            const int k = e / RW, rest = e - k * RW;
            const int p = rest / (GW / 4), i4 = rest - p * (GW / 4);
            tv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (4 * i4 < L64[p])
                tv[u] = *reinterpret_cast<const float4*>(NF + gm[p].offF + (size_t)(kc + k) * L64[p] + 4 * i4);
        }
    };
    auto stash = [&](float* buf) {
        const int c = 2 * ap;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = k0 + 8 * h;
            const int ka = k ^ swz16(c), kb = k ^ swz16(c + 1);
            buf[(0 * TM + c) * KC + ka] = xa[h].x + xb[h].x;
            buf[(0 * TM + c + 1) * KC + kb] = xa[h].y + xb[h].y;
            buf[(1 * TM + c) * KC + ka] = xa[h].x - xb[h].x;
            buf[(1 * TM + c + 1) * KC + kb] = xa[h].y - xb[h].y;
        }
        float* tb = buf + ABUF;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const int e = tid + kF3Threads * u;
            const int k = e / RW, rest = e - k * RW;
            const int p = rest / (GW / 4), i4 = rest - p * (GW / 4);
            *reinterpret_cast<float4*>(tb + (k * 2 + p) * GW + 4 * i4) = tv[u];
        }
    };

    const int c = wave * 16 + (lane & 15);
    const int sc = swz16(c), kq = lane >> 4;
    fvec4 acc[2][G][4] = {};
    fetch(0);
    stash(sm);
    __syncthreads();
    for (int kc = 0, bi = 0; kc < B; kc += KC, bi ^= 1) {
        const float* cur = sm + bi * (ABUF + TBUF);
        const bool more = kc + KC < B;
        if (more) fetch(kc + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const float a = cur[(p * TM + c) * KC + ((kk + kq) ^ sc)];
                const float* trow = cur + ABUF + ((kk + kq) * 2 + p) * GW + 4 * (lane & 15);
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    const float4 t = *reinterpret_cast<const float4*>(trow + q * 64);
                    acc[p][q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.x, acc[p][q][0], 0, 0, 0);
                    acc[p][q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.y, acc[p][q][1], 0, 0, 0);
                    acc[p][q][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.z, acc[p][q][2], 0, 0, 0);
                    acc[p][q][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.w, acc[p][q][3], 0, 0, 0);
                }
            }
        }
        if (more) stash(sm + (bi ^ 1) * (ABUF + TBUF));
        __syncthreads();
    }
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = q * 64 + 16 * u + (lane & 15);
                if (i >= gm[p].Lp) continue;
                const int l = m + p + 2 * i;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + wave * 16 + Mma<float>::row(lane, r);
                    if (col >= ncols) continue;
                    const int f = col / cpf, rem = col - f * cpf;
                    const int comp = rem & 1;
                    double* out = comp ? cim : cre;
                    const double v = (double)acc[p][q][u][r];
                    if ((rem >> 1) == 0) {
                        out[f * B2 + index_of_coeff(m, l, B)] = v;
                        if (fmt == 1 && m > 0) out[f * B2 + index_of_coeff(-m, l, B)] = comp ? -sg * v : sg * v;
                    } else {
                        out[f * B2 + index_of_coeff(-m, l, B)] = sg * v;
                    }
                }
            }
        }
    }
}

// Inverse, pipelined form.  A block owns TM = 128 columns of one order and a
// range of GJ*64 rings (blockIdx.z selects the range).  Chunks of 32
// consecutive degrees (16 per parity) stream through double-buffered LDS:
// coefficients A[c][p][i] and table rows T[p][i][GJ*64].  8 waves, each 16
// columns x GJ x 4 ring tiles x (even, odd) accumulators.
constexpr int kI3KC = 16;  // degrees per parity per chunk

template <int GJ>
__global__ __launch_bounds__(kF3Threads) void k_ngemm_inv3(const double* __restrict__ cre,
                                                           const double* __restrict__ cim,
                                                           const float* __restrict__ NI,
                                                           const GemmMeta* __restrict__ meta,
                                                           float2* __restrict__ X, int lgB, int F, int fmt) {
    constexpr int TM = kF3TM, KC = kI3KC, JW = GJ * 64;
    constexpr int ABUF = TM * 2 * KC;  // floats of coefficients per buffer
    constexpr int TBUF = 2 * KC * JW;  // floats of table per buffer
    extern __shared__ __align__(16) unsigned char smem[];
    float* sm = reinterpret_cast<float*>(smem);
    const int B = 1 << lgB, N = 2 * B;
    const int m = blockIdx.x;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * TM;
    if (cbase >= ncols) return;
    const int jbase = blockIdx.z * JW;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GemmMeta gm[2] = {meta[m * 2], meta[m * 2 + 1]};
    const int Li4[2] = {(gm[0].Lp + 3) & ~3, (gm[1].Lp + 3) & ~3};
    const int L = B - m;
    const size_t B2 = (size_t)B * B;

    // coefficient fill: pair ap (columns 2ap re, 2ap+1 im), degrees l0+d, d = dk + 8h
    const int ap = tid >> 3, dk = tid & 7;
    const double* cr = nullptr;
    const double* ci = nullptr;
    {
        const int col = cbase + 2 * ap;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            const size_t base = f * B2 + index_of_coeff((rem >> 1) ? -m : m, m, B);
            cr = cre + base;
            ci = cim + base;
        }
    }
    double vr[4], vi[4];
    float4 tv[GJ];
    constexpr int RW = JW / 4;  // float4 per table row
    auto fetch = [&](int d0) {  // degrees m + d0 .. m + d0 + 31
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int d = d0 + dk + 8 * h;
            const bool ok = cr && d < L;
            vr[h] = ok ? cr[d] : 0.0;
            vi[h] = ok ? ci[d] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < GJ; ++u) {
            const int e = tid + kF3Threads * u;  // < 2 * KC * RW
            const int row = e / RW, j4 = e - row * RW;
            const int p = row / KC, i = d0 / 2 + (row - p * KC);
            tv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < Li4[p])
                tv[u] = *reinterpret_cast<const float4*>(NI + gm[p].offI + (size_t)i * B + jbase + 4 * j4);
        }
    };
    auto stash = [&](float* buf) {
        const int c = 2 * ap;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int d = dk + 8 * h;  // within the chunk: parity d&1, index d>>1
            const int p = d & 1, i = d >> 1;
            buf[(c * 2 + p) * KC + (i ^ swz16(c))] = (float)vr[h];
            buf[((c + 1) * 2 + p) * KC + (i ^ swz16(c + 1))] = (float)vi[h];
        }
        float* tb = buf + ABUF;
#pragma unroll
        for (int u = 0; u < GJ; ++u) {
            const int e = tid + kF3Threads * u;
            const int row = e / RW, j4 = e - row * RW;
            *reinterpret_cast<float4*>(tb + row * JW + 4 * j4) = tv[u];
        }
    };

    const int c = wave * 16 + (lane & 15);
    const int sc = swz16(c), kq = lane >> 4;
    fvec4 acc[2][GJ][4] = {};
    fetch(0);
    stash(sm);
    __syncthreads();
    for (int d0 = 0, bi = 0; d0 < L; d0 += 2 * KC, bi ^= 1) {
        const float* cur = sm + bi * (ABUF + TBUF);
        const bool more = d0 + 2 * KC < L;
        if (more) fetch(d0 + 2 * KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const float a = cur[(c * 2 + p) * KC + ((kk + kq) ^ sc)];
                const float* trow = cur + ABUF + (p * KC + kk + kq) * JW + 4 * (lane & 15);
#pragma unroll
                for (int q = 0; q < GJ; ++q) {
                    const float4 t = *reinterpret_cast<const float4*>(trow + q * 64);
                    acc[p][q][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.x, acc[p][q][0], 0, 0, 0);
                    acc[p][q][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.y, acc[p][q][1], 0, 0, 0);
                    acc[p][q][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.z, acc[p][q][2], 0, 0, 0);
                    acc[p][q][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, t.w, acc[p][q][3], 0, 0, 0);
                }
            }
        }
        if (more) stash(sm + (bi ^ 1) * (ABUF + TBUF));
        __syncthreads();
    }
    const float inv_s2pi = 0.39894228040143267794f;
#pragma unroll
    for (int q = 0; q < GJ; ++q) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = jbase + q * 64 + 16 * u + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r += 2) {  // rows r, r+1: re, im of one (field, sign)
                const int col = cbase + wave * 16 + Mma<float>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf;
                const int row = (rem >> 1) ? N - m : m;
                const float s = (row > B && (row & 1)) ? -inv_s2pi : inv_s2pi;
                const fvec4& e = acc[0][q][u];
                const fvec4& o = acc[1][q][u];
                const float2 yn = make_float2((e[r] + o[r]) * s, (e[r + 1] + o[r + 1]) * s);
                const float2 ys = make_float2((e[r] - o[r]) * s, (e[r + 1] - o[r + 1]) * s);
                float2* xr = X + ((size_t)f * N + row) * N;
                xr[j] = yn;
                xr[N - 1 - j] = ys;
                if (fmt == 1 && m > 0) {
                    float2* xm = X + ((size_t)f * N + (N - m)) * N;
                    xm[j] = make_float2(yn.x, -yn.y);
                    xm[N - 1 - j] = make_float2(ys.x, -ys.y);
                }
            }
        }
    }
}

// Inverse: E/O[c][j] = sum_i coef_p[c][i] * NI(m,p)[i][j]; ring j gets E+O,
// ring N-1-j gets E-O; x 1/sqrt(2pi), (-1) on odd negative rows, REAL mirror.
// grid (bw, ceil(cols/ncb)); LDS Cd[c][q] (swizzled), q = i or roundup4(Lp0)+i.
// A wave owns 16 columns x 64 rings (four MFMA tiles per parity).
// acol: this lane's LDS column; rows qoff + i (swizzled with sc) for i < K4.
__device__ __forceinline__ void ngemm_inv_k(const float* __restrict__ acol, int qoff, int sc,
                                            const float* __restrict__ brow, int B, int K4, fvec4* acc) {
    fvec4 a1[4] = {};
    int i0 = 0;
    const int kq = (threadIdx.x & 63) >> 4;
    // table fragments one step ahead (clamped: the last fetch repeats)
    // rows are clamped to the last step (K4 - 4 + kq): the tail step, when
    // there is one, finds its fragment in ta; other clamped fetches go unused
    float4 ta = *reinterpret_cast<const float4*>(brow + (size_t)kq * B);
    float4 tb = *reinterpret_cast<const float4*>(brow + (size_t)(min(4, K4 - 4) + kq) * B);
    for (; i0 + 8 <= K4; i0 += 8) {
        const int ia = i0 + kq, ib = ia + 4;
        const float4 na = *reinterpret_cast<const float4*>(brow + (size_t)(min(i0 + 8, K4 - 4) + kq) * B);
        const float4 nb = *reinterpret_cast<const float4*>(brow + (size_t)(min(i0 + 12, K4 - 4) + kq) * B);
        const float xa = acol[(qoff + ia) ^ sc], xb = acol[(qoff + ib) ^ sc];
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.x, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.y, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.z, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.w, acc[3], 0, 0, 0);
        a1[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.x, a1[0], 0, 0, 0);
        a1[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.y, a1[1], 0, 0, 0);
        a1[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.z, a1[2], 0, 0, 0);
        a1[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(xb, tb.w, a1[3], 0, 0, 0);
        ta = na;
        tb = nb;
    }
    if (i0 < K4) {  // one step of 4 left; ta holds it
        const int ia = i0 + kq;
        const float xa = acol[(qoff + ia) ^ sc];
        acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.x, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.y, acc[1], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.z, acc[2], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa, ta.w, acc[3], 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += a1[q];
}

__global__ __launch_bounds__(kThreads) void k_ngemm_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                        const float* __restrict__ NI,
                                                        const GemmMeta* __restrict__ meta, float2* __restrict__ X,
                                                        int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* Cd = reinterpret_cast<float*>(smem);
    const int B = 1 << lgB, N = 2 * B;
    const int m = blockIdx.x;  // orders fastest: concurrent blocks stream the same fields' rows
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * ncb;
    if (cbase >= ncols) return;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int Li4_0 = (g0.Lp + 3) & ~3, Li4_1 = (g1.Lp + 3) & ~3;
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    for (int idx = threadIdx.x; idx < ncb * L; idx += kThreads) {
        const int c = idx / L, li = idx - c * L;
        const int col = cbase + c;
        const int q = (li & 1) ? Li4_0 + (li >> 1) : (li >> 1);
        float v = 0.f;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            const double* src = (rem & 1) ? cim : cre;
            v = (float)src[f * B2 + index_of_coeff((rem >> 1) ? -m : m, m + li, B)];
        }
        Cd[(size_t)c * B + (q ^ swz(c))] = v;
    }
    for (int idx = threadIdx.x; idx < ncb * 8; idx += kThreads) {
        const int c = idx >> 3, k = idx & 7;
        const int p = k >> 2, i = (p ? g1.Lp : g0.Lp) + (k & 3);
        if (i < (p ? Li4_1 : Li4_0)) Cd[(size_t)c * B + (((p ? Li4_0 : 0) + i) ^ swz(c))] = 0.f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = ncb >> 4, ngj = B >> 6;
    const float inv_s2pi = 0.39894228040143267794f;
    for (int item = wave; item < ngj * nct; item += kThreads / 64) {
        const int jg = item / nct, ct = item - jg * nct;
        const int c0 = ct * 16, c = c0 + (lane & 15);
        const int sc = swz(c);
        const float* acol = Cd + (size_t)c * B;
        fvec4 e[4] = {}, o[4] = {};
        ngemm_inv_k(acol, 0, sc, NI + g0.offI + jg * 64 + 4 * (lane & 15), B, Li4_0, e);
        if (g1.Lp > 0) ngemm_inv_k(acol, Li4_0, sc, NI + g1.offI + jg * 64 + 4 * (lane & 15), B, Li4_1, o);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = jg * 64 + 16 * q + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r += 2) {  // rows r, r+1: the re, im columns of one (field, sign)
                const int col = cbase + c0 + Mma<float>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf;
                const int row = (rem >> 1) ? N - m : m;
                const float s = (row > B && (row & 1)) ? -inv_s2pi : inv_s2pi;
                const float2 yn = make_float2((e[q][r] + o[q][r]) * s, (e[q][r + 1] + o[q][r + 1]) * s);
                const float2 ys = make_float2((e[q][r] - o[q][r]) * s, (e[q][r + 1] - o[q][r + 1]) * s);
                float2* xr = X + ((size_t)f * N + row) * N;
                xr[j] = yn;
                xr[N - 1 - j] = ys;
                if (fmt == 1 && m > 0) {
                    float2* xm = X + ((size_t)f * N + (N - m)) * N;
                    xm[j] = make_float2(yn.x, -yn.y);
                    xm[N - 1 - j] = make_float2(ys.x, -ys.y);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int lg2(int n) { return __builtin_ctz(n); }

void gemm_tables_create(GemmTables& g, const DevTables& t, int prec, hipStream_t s) {
    const int B = t.bw, half = B / 2;
    g.bw = B;
    g.prec = prec;
    g.h_meta.assign(2 * B, GemmMeta{});
    long of = 0, oi = 0;
    for (int m = 0; m < B; ++m)
        for (int p = 0; p < 2; ++p) {
            GemmMeta& e = g.h_meta[m * 2 + p];
            e.Lp = gm_lp(m, p, B);
            const int llast = m + p + 2 * (e.Lp - 1);
            e.Kf = e.Lp > 0 ? row_size(m, llast) : 0;
            e.offF = of;
            e.offI = oi;
            of += (long)((e.Kf + 3) & ~3) * ((e.Lp + 15) & ~15);
            oi += (long)((e.Lp + 3) & ~3) * ((half + 15) & ~15);
        }
    g.nf = of;
    g.ni = oi;
    const size_t el = prec == 0 ? sizeof(double) : sizeof(float);
    S2_HIP(hipMalloc(&g.meta, sizeof(GemmMeta) * 2 * B));
    S2_HIP(hipMalloc(&g.af, el * std::max(1L, of)));
    S2_HIP(hipMalloc(&g.ai, el * std::max(1L, oi)));
    S2_HIP(hipMemcpyAsync(g.meta, g.h_meta.data(), sizeof(GemmMeta) * 2 * B, hipMemcpyHostToDevice, s));
    const size_t lds = sizeof(int) * (B + 1);
    if (prec == 0)
        k_pack_gemm<double><<<dim3(B, 2), 256, lds, s>>>(t.fwd, t.inv, t.off, g.meta, (double*)g.af, (double*)g.ai, B);
    else
        k_pack_gemm<float><<<dim3(B, 2), 256, lds, s>>>(t.fwd, t.inv, t.off, g.meta, (float*)g.af, (float*)g.ai, B);
    S2_HIP(hipGetLastError());
    S2_HIP(hipStreamSynchronize(s));
}

void gemm_tables_destroy(GemmTables& g) {
    if (g.meta) (void)hipFree(g.meta);
    if (g.af) (void)hipFree(g.af);
    if (g.ai) (void)hipFree(g.ai);
    g = GemmTables{};
}

size_t fast_x_bytes(int bw, int prec) {
    const size_t N = 2 * (size_t)bw;
    return N * N * (prec == 0 ? sizeof(double2) : sizeof(float2));
}

size_t fast_cb_bytes(int bw, int prec) {
    const size_t N = 2 * (size_t)bw;
    return N * 2 * bw * (prec == 0 ? sizeof(double) : sizeof(float));
}

// Columns per GEMM block: the largest multiple of 16 (at most 64) whose
// padded LDS tile fits in 64 KiB.
static int gemm_ncb(int B, int prec) {
    const int row = prec == 0 ? 8 * (B + lds_pad<double>()) : 4 * (B + lds_pad<float>());
    return std::min(64, 65536 / row) & ~15;
}

static size_t gemm_lds(int B, int prec) {
    const int row = prec == 0 ? 8 * (B + lds_pad<double>()) : 4 * (B + lds_pad<float>());
    return (size_t)gemm_ncb(B, prec) * row;
}

bool fast_supported(int bw, int cutoff, int prec) {
    const int N = 2 * bw;
    // N*N must hold at least one block tile of rows (N >= 64)
    if (bw < 32 || (N & (N - 1)) || N > 2048 || cutoff != bw) return false;
    return gemm_ncb(bw, prec) >= 16;  // GEMM LDS tile of at least 16 columns
}

// Launch with `lds` bytes of dynamic LDS; above 64 KiB the kernel must opt
// in (gfx950 has 160 KiB per workgroup).
template <typename K, typename... A>
static void launch_lds(K kernel, dim3 grid, int threads, size_t lds, hipStream_t s, A... args) {
    if (lds > 65536) S2_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    kernel<<<grid, threads, lds, s>>>(args...);
    S2_HIP(hipGetLastError());
}

template <typename R>
static void fwd_impl(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                     double* cim, int F, Fmt fmt, void* Xv, void* Cbv, hipStream_t s, StageTimer* tm) {
    using T = typename Cplx<R>::type;
    constexpr int V = sizeof(R) == 8 ? 8 : 16;
    const int B = c.bw, N = c.N, lgN = lg2(N);
    const int E = V * kThreads, rpb = E / N;
    const size_t lds = sizeof(T) * ((size_t)rpb * (N + 1) + N);
    T* X = (T*)Xv;
    R* Cb = (R*)Cbv;
    const T* tw = (const T*)(sizeof(R) == 8 ? (const void*)c.tw : (const void*)c.tw32);
    const T* ph = (const T*)(sizeof(R) == 8 ? (const void*)c.ph : (const void*)c.ph32);
    const R* ws = (const R*)(sizeof(R) == 8 ? (const void*)c.w_semi : (const void*)c.w_semi32);
    {
        Span sp(tm, ST_PHI_FWD, s);
        launch_lds(k_fphi_fwd<R, V>, dim3(N / rpb, F), kThreads, lds, s, re, im, X, lgN, tw);
    }
    const int nrows = (fmt == Fmt::Complex) ? 2 * B - 1 : B;
    {
        Span sp(tm, ST_DCT_FWD, s);
        launch_lds(k_fdct_fwd<R, V>, dim3((nrows + rpb - 1) / rpb, F), kThreads, lds, s, (const T*)X, Cb, ws, lgN, B,
                   nrows, tw, ph);
    }
    {
        const int ncb = gemm_ncb(B, g.prec);
        const int cols = 4 * F;
        Span sp(tm, ST_LEG_FWD, s);
        k_gemm_fwd<R><<<dim3((cols + ncb - 1) / ncb, B), kThreads, gemm_lds(B, g.prec), s>>>(
            Cb, (const R*)g.af, g.meta, cre, cim, lg2(B), F, ncb, (int)fmt);
        S2_HIP(hipGetLastError());
    }
}

template <typename R>
static void inv_impl(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                     double* im, int F, Fmt fmt, void* Xv, void* Cbv, hipStream_t s, StageTimer* tm) {
    using T = typename Cplx<R>::type;
    constexpr int V = sizeof(R) == 8 ? 8 : 16;
    const int B = c.bw, N = c.N, lgN = lg2(N);
    const int E = V * kThreads, rpb = E / N;
    const size_t lds = sizeof(T) * ((size_t)rpb * (N + 1) + N);
    T* X = (T*)Xv;
    R* Cb = (R*)Cbv;
    const T* tw = (const T*)(sizeof(R) == 8 ? (const void*)c.tw : (const void*)c.tw32);
    const T* ph = (const T*)(sizeof(R) == 8 ? (const void*)c.ph : (const void*)c.ph32);
    const R* sv = (const R*)(sizeof(R) == 8 ? (const void*)c.sinv : (const void*)c.sinv32);
    {
        const int ncb = gemm_ncb(B, g.prec);
        const int cols = 4 * F;
        Span sp(tm, ST_LEG_INV, s);
        k_gemm_inv<R><<<dim3((cols + ncb - 1) / ncb, B), kThreads, gemm_lds(B, g.prec), s>>>(
            cre, cim, (const R*)g.ai, g.meta, Cb, lg2(B), F, ncb, (int)fmt);
        S2_HIP(hipGetLastError());
    }
    const int nrows = (fmt == Fmt::Complex) ? 2 * B - 1 : B;
    {
        Span sp(tm, ST_DCT_INV, s);
        launch_lds(k_fdct_inv<R, V>, dim3((nrows + rpb - 1) / rpb, F), kThreads, lds, s, (const R*)Cb, X, sv, lgN, B,
                   nrows, (int)fmt, tw, ph);
    }
    {
        Span sp(tm, ST_PHI_INV, s);
        launch_lds(k_fphi_inv<R, V>, dim3(N / rpb, F), kThreads, lds, s, (const T*)X, re, im, lgN, tw);
    }
}

bool mixed_supported(int bw) {
    const int N = 2 * bw;
    return bw >= 64 && bw <= 512 && !(N & (N - 1));
}

void naive_tables_create(GemmTables& g, const DevConsts& c, hipStream_t s) {
    const int B = c.bw;
    g.bw = B;
    g.prec = 2;
    g.h_meta.assign(2 * B, GemmMeta{});
    long of = 0, oi = 0;
    for (int m = 0; m < B; ++m)
        for (int p = 0; p < 2; ++p) {
            GemmMeta& e = g.h_meta[m * 2 + p];
            e.Lp = gm_lp(m, p, B);
            e.Kf = B;
            e.offF = of;
            e.offI = oi;
            of += (long)B * ((e.Lp + 63) & ~63);
            oi += (long)((e.Lp + 3) & ~3) * B;
        }
    g.nf = of;
    g.ni = oi;
    S2_HIP(hipMalloc(&g.meta, sizeof(GemmMeta) * 2 * B));
    S2_HIP(hipMalloc(&g.af, sizeof(float) * std::max(1L, of)));
    S2_HIP(hipMalloc(&g.ai, sizeof(float) * std::max(1L, oi)));
    S2_HIP(hipMemsetAsync(g.af, 0, sizeof(float) * std::max(1L, of), s));
    S2_HIP(hipMemsetAsync(g.ai, 0, sizeof(float) * std::max(1L, oi), s));
    S2_HIP(hipMemcpyAsync(g.meta, g.h_meta.data(), sizeof(GemmMeta) * 2 * B, hipMemcpyHostToDevice, s));
    k_pack_naive<<<B, 256, 0, s>>>(c.w_naive, g.meta, (float*)g.af, (float*)g.ai, B);
    S2_HIP(hipGetLastError());
    S2_HIP(hipStreamSynchronize(s));
}

static int ngemm_ncb(int B, bool fwd) {
    const int row = (fwd ? 2 : 1) * B * (int)sizeof(float);
    return std::min(64, 65536 / row) & ~15;
}

void forward_mixed(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                   double* cim, int F, Fmt fmt, void* Xv, hipStream_t s, StageTimer* tm) {
    constexpr int V = 16;
    const int B = c.bw, N = c.N, lgN = lg2(N);
    const int rpb = V * kThreads / N;
    float2* X = (float2*)Xv;
    {
        Span sp(tm, ST_PHI_FWD, s);
        k_fphi_fwd<float, V><<<dim3(N / rpb, F), kThreads, sizeof(float2) * ((size_t)rpb * (N + 1) + N), s>>>(
            re, im, X, lgN, c.tw32);
        S2_HIP(hipGetLastError());
    }
    {
        Span sp(tm, ST_LEG_FWD, s);
        // orders in buckets of equal 64-degree group count g: m in [B - 128g, B - 128(g-1))
        const int gmax = (B / 2 + 63) / 64;
        for (int gq = 1; gq <= gmax; ++gq) {
            const int lo = std::max(0, B - 128 * gq), hi = std::min(B, B - 128 * (gq - 1));
            if (lo >= hi) continue;
            const dim3 grid(hi - lo, (4 * F + kF3TM - 1) / kF3TM);
            const size_t lds = sizeof(float) * 2 * (2 * kF3TM * kF3KC + kF3KC * 2 * 64 * gq);
            const float* nf = (const float*)g.af;
            switch (gq) {
                case 1: launch_lds(k_ngemm_fwd3<1>, grid, kF3Threads, lds, s, X, nf, g.meta, cre, cim, lg2(B), F, (int)fmt, lo); break;
                case 2: launch_lds(k_ngemm_fwd3<2>, grid, kF3Threads, lds, s, X, nf, g.meta, cre, cim, lg2(B), F, (int)fmt, lo); break;
                case 3: launch_lds(k_ngemm_fwd3<3>, grid, kF3Threads, lds, s, X, nf, g.meta, cre, cim, lg2(B), F, (int)fmt, lo); break;
                default: launch_lds(k_ngemm_fwd3<4>, grid, kF3Threads, lds, s, X, nf, g.meta, cre, cim, lg2(B), F, (int)fmt, lo); break;
            }
        }
    }
}

void inverse_mixed(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                   double* im, int F, Fmt fmt, void* Xv, hipStream_t s, StageTimer* tm) {
    constexpr int V = 16;
    const int B = c.bw, N = c.N, lgN = lg2(N);
    const int rpb = V * kThreads / N;
    float2* X = (float2*)Xv;
    {
        Span sp(tm, ST_LEG_INV, s);
        const int gj = std::min(4, B / 64);
        const dim3 grid(B, (4 * F + kF3TM - 1) / kF3TM, B / (64 * gj));
        const size_t lds = sizeof(float) * 2 * (kF3TM * 2 * kI3KC + 2 * kI3KC * 64 * gj);
        const float* ni = (const float*)g.ai;
        switch (gj) {
            case 1: launch_lds(k_ngemm_inv3<1>, grid, kF3Threads, lds, s, cre, cim, ni, g.meta, X, lg2(B), F, (int)fmt); break;
            case 2: launch_lds(k_ngemm_inv3<2>, grid, kF3Threads, lds, s, cre, cim, ni, g.meta, X, lg2(B), F, (int)fmt); break;
            default: launch_lds(k_ngemm_inv3<4>, grid, kF3Threads, lds, s, cre, cim, ni, g.meta, X, lg2(B), F, (int)fmt); break;
        }
    }
    {
        Span sp(tm, ST_PHI_INV, s);
        k_fphi_inv<float, V><<<dim3(N / rpb, F), kThreads, sizeof(float2) * ((size_t)rpb * (N + 1) + N), s>>>(
            X, re, im, lgN, c.tw32);
        S2_HIP(hipGetLastError());
    }
}

void forward_fast(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                  double* cim, int F, Fmt fmt, void* X, void* Cb, hipStream_t s, StageTimer* tm) {
    if (g.prec == 0)
        fwd_impl<double>(c, g, re, im, cre, cim, F, fmt, X, Cb, s, tm);
    else
        fwd_impl<float>(c, g, re, im, cre, cim, F, fmt, X, Cb, s, tm);
}

void inverse_fast(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                  double* im, int F, Fmt fmt, void* X, void* Cb, hipStream_t s, StageTimer* tm) {
    if (g.prec == 0)
        inv_impl<double>(c, g, cre, cim, re, im, F, fmt, X, Cb, s, tm);
    else
        inv_impl<float>(c, g, cre, cim, re, im, F, fmt, X, Cb, s, tm);
}

}  // namespace s2
