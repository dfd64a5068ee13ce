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
__device__ __forceinline__ void fft_pass(T* a, int lgN, int S, int lgNs, const T* __restrict__ tw) {
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
#pragma unroll
        for (int r = 1; r < R; ++r) {
            T w = tw[r * step];
            if (SIGN > 0) w.y = -w.y;
            x[u * R + r] = cmul(x[u * R + r], w);
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
__device__ void fft_block(T* a, int lgN, int S, const T* __restrict__ tw) {
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

// position of cosine index k in the parity-split planes of the Cb buffer
__device__ __forceinline__ int cpos(int k, int half) { return (k & 1) ? half + (k >> 1) : (k >> 1); }

// ---------------------------------------------------------------------------
// FFT-type kernels.  grid.y = field, block = kThreads, LDS = rows*(N+1) values
// ---------------------------------------------------------------------------
template <typename R, int V>
__global__ __launch_bounds__(kThreads) void k_fphi_fwd(const double* __restrict__ re, const double* __restrict__ im,
                                                       typename Cplx<R>::type* __restrict__ X, int lgN,
                                                       const typename Cplx<R>::type* __restrict__ tw) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    const int f = blockIdx.y, j0 = blockIdx.x * rpb;
    const size_t base = ((size_t)f * N + j0) * N;
    for (int idx = threadIdx.x; idx < E; idx += kThreads)
        a[(idx >> lgN) * S + (idx & (N - 1))] = mk<T>((R)re[base + idx], (R)im[base + idx]);
    __syncthreads();
    fft_block<T, -1, V>(a, lgN, S, tw);
    for (int idx = threadIdx.x; idx < E; idx += kThreads) {
        const int r = idx >> lgrpb, jl = idx & (rpb - 1);
        X[((size_t)f * N + r) * N + j0 + jl] = a[jl * S + r];
    }
}

template <typename R, int V>
__global__ __launch_bounds__(kThreads) void k_fdct_fwd(const typename Cplx<R>::type* __restrict__ X, R* __restrict__ Cb,
                                                       const R* __restrict__ w_semi, int lgN, int cut, int nrows,
                                                       const typename Cplx<R>::type* __restrict__ tw,
                                                       const typename Cplx<R>::type* __restrict__ ph) {
    using T = typename Cplx<R>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1, B = N >> 1, half = B >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    const int f = blockIdx.y, q0 = blockIdx.x * rpb;
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
    fft_block<T, -1, V>(a, lgN, S, tw);
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
    T* a = reinterpret_cast<T*>(smem);
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1, B = N >> 1, half = B >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    const int f = blockIdx.y, q0 = blockIdx.x * rpb;
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
    fft_block<T, +1, V>(a, lgN, S, tw);
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
    T* a = reinterpret_cast<T*>(smem);
    constexpr int E = V * kThreads;
    const int N = 1 << lgN, S = N + 1, B = N >> 1;
    const int lgrpb = __builtin_ctz(E) - lgN, rpb = 1 << lgrpb;
    const int f = blockIdx.y, j0 = blockIdx.x * rpb;
    for (int idx = threadIdx.x; idx < E; idx += kThreads) {
        const int r = idx >> lgrpb, jl = idx & (rpb - 1);
        a[jl * S + r] = (r == B) ? mk<T>(0, 0) : X[((size_t)f * N + r) * N + j0 + jl];
    }
    __syncthreads();
    fft_block<T, +1, V>(a, lgN, S, tw);
    const size_t base = ((size_t)f * N + j0) * N;
    for (int idx = threadIdx.x; idx < E; idx += kThreads) {
        const T v = a[(idx >> lgN) * S + (idx & (N - 1))];
        re[base + idx] = (double)v.x;
        im[base + idx] = (double)v.y;
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

// Forward: D'[c][i] = sum_t data[c][t] * A(m,p)[i][t].  grid (cut, ceil(cols/ncb)).
// LDS: Bd[c][k2], k2 = parity-split cosine position (p*bw/2 + t).
template <typename R>
__global__ __launch_bounds__(kThreads) void k_gemm_fwd(const R* __restrict__ Cb, const R* __restrict__ AF,
                                                       const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                       double* __restrict__ cim, int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* Bd = reinterpret_cast<R*>(smem);
    const int B = 1 << lgB, N = 2 * B, half = B >> 1;
    const int m = blockIdx.x;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * ncb;
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

// Inverse: D'[c][t] = sum_i coef[c][i] * AI(m,p)[i][t].  grid (cut, ceil(cols/ncb)).
// LDS: Cd[c][q], q = i (p = 0) or roundup4(Lp0) + i (p = 1); q < bw always.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_gemm_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                       const R* __restrict__ AI, const GemmMeta* __restrict__ meta,
                                                       R* __restrict__ Cb, int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* Cd = reinterpret_cast<R*>(smem);
    const int B = 1 << lgB, N = 2 * B, half = B >> 1;
    const int m = blockIdx.x;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int cbase = blockIdx.y * ncb;
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
//   NF(m,p) [j][i]  P_{m+p+2i}^m(theta_j) w_j sqrt(2pi)/N  (j < bw, i < roundup16(Lp))
//   NI(m,p) [i][j]  P_{m+p+2i}^m(theta_j)                  (i < roundup4(Lp), j < bw)
// Both come from the reference's FP64 recurrence (pml.c:41-79), rounded once.

// XOR swizzle of an LDS row index: 16 consecutive columns reading the same
// k land in 16 different 4-bank groups (bw is a multiple of 32 here).
__device__ __forceinline__ int swz(int c, int B) { return ((c & 15) << 2) & (B - 1); }

__global__ void k_pack_naive(const double* __restrict__ w, const GemmMeta* __restrict__ meta, float* __restrict__ NF,
                             float* __restrict__ NI, int B) {
    const int m = blockIdx.x;
    const int N = 2 * B;
    const double cm = (m == 0) ? 0.0 : pmm_coeff(m);
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int L16_0 = (g0.Lp + 15) & ~15, L16_1 = (g1.Lp + 15) & ~15;
    for (int j = threadIdx.x; j < B; j += blockDim.x) {
        const double th = (2.0 * j + 1.0) * M_PI / (2.0 * N);
        const double x = cos((2.0 * j + 1.0) * M_PI / (2.0 * N));
        const double wj = w[j];  // even weight times sqrt(2pi)/N
        double cur = (m == 0) ? 0.70710678118654752440 : cm * pow(sin(th), (double)m);
        double prev = 0.0;
        for (int l = m; l < B; ++l) {
            const int p = (l - m) & 1, i = (l - m) >> 1;
            const GemmMeta& g = p ? g1 : g0;
            NF[g.offF + (long)j * (p ? L16_1 : L16_0) + i] = (float)(cur * wj);
            NI[g.offI + (long)i * B + j] = (float)cur;
            if (l + 1 < B) {
                const double nx = l2_an(m, l) * (cur * x) + l2_cn(m, l) * prev;
                prev = cur;
                cur = nx;
            }
        }
    }
}

// Forward: coefficient[c][i] = sum_j fold_p(x)[c][j] * NF(m,p)[j][i].
// grid (bw, ceil(cols/ncb)); LDS Sd[c][p][j] (swizzled j), 2*bw floats per column.
__global__ __launch_bounds__(kThreads) void k_ngemm_fwd(const float2* __restrict__ X, const float* __restrict__ NF,
                                                        const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                        double* __restrict__ cim, int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* Sd = reinterpret_cast<float*>(smem);
    const int B = 1 << lgB, N = 2 * B;
    const int m = blockIdx.x;
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
        const int s0 = swz(c, B), s1 = swz(c + 1, B);
        float* d0 = Sd + (size_t)c * 2 * B;
        float* d1 = d0 + 2 * B;
        d0[j ^ s0] = a.x + b.x;
        d1[j ^ s1] = a.y + b.y;
        d0[B + (j ^ s0)] = a.x - b.x;
        d1[B + (j ^ s1)] = a.y - b.y;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = ncb >> 4;
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = meta[m * 2 + p];
        if (g.Lp == 0) continue;
        const int L16 = (g.Lp + 15) & ~15;
        const int ntiles = (L16 >> 4) * nct;
        for (int item = wave; item < ntiles; item += kThreads / 64) {
            const int it = item / nct, ct = item - it * nct;
            const int i0 = it * 16, c0 = ct * 16;
            const int c = c0 + (lane & 15);
            const float* arow = Sd + (size_t)c * 2 * B + p * B;
            const int sc = swz(c, B);
            const float* brow = NF + g.offF + i0 + (lane & 15);
            fvec4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
            for (int j0 = 0; j0 < B; j0 += 8) {
                const int ja = j0 + (lane >> 4), jb = ja + 4;
                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[ja ^ sc], brow[(size_t)ja * L16], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(arow[jb ^ sc], brow[(size_t)jb * L16], acc1, 0, 0, 0);
            }
            const int i = i0 + (lane & 15);
            if (i >= g.Lp) continue;
            const int l = m + p + 2 * i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = cbase + c0 + Mma<float>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf;
                const int comp = rem & 1;
                double* out = comp ? cim : cre;
                const double v = (double)acc0[r] + (double)acc1[r];
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

// Inverse: E/O[c][j] = sum_i coef_p[c][i] * NI(m,p)[i][j]; ring j gets E+O,
// ring N-1-j gets E-O; x 1/sqrt(2pi), (-1) on odd negative rows, REAL mirror.
// grid (bw, ceil(cols/ncb)); LDS Cd[c][q] (swizzled), q = i or roundup4(Lp0)+i.
__global__ __launch_bounds__(kThreads) void k_ngemm_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                        const float* __restrict__ NI,
                                                        const GemmMeta* __restrict__ meta, float2* __restrict__ X,
                                                        int lgB, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* Cd = reinterpret_cast<float*>(smem);
    const int B = 1 << lgB, N = 2 * B;
    const int m = blockIdx.x;
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
        Cd[(size_t)c * B + (q ^ swz(c, B))] = v;
    }
    for (int idx = threadIdx.x; idx < ncb * 8; idx += kThreads) {
        const int c = idx >> 3, k = idx & 7;
        const int p = k >> 2, i = (p ? g1.Lp : g0.Lp) + (k & 3);
        if (i < (p ? Li4_1 : Li4_0)) Cd[(size_t)c * B + (((p ? Li4_0 : 0) + i) ^ swz(c, B))] = 0.f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nct = ncb >> 4, ntj = B >> 4;
    const float inv_s2pi = 0.39894228040143267794f;
    for (int item = wave; item < ntj * nct; item += kThreads / 64) {
        const int jt = item / nct, ct = item - jt * nct;
        const int j0 = jt * 16, c0 = ct * 16;
        const int c = c0 + (lane & 15);
        const int sc = swz(c, B);
        const float* acol = Cd + (size_t)c * B;
        fvec4 e = {0, 0, 0, 0}, o = {0, 0, 0, 0};
        {
            const float* brow = NI + g0.offI + j0 + (lane & 15);
            fvec4 e1 = {0, 0, 0, 0};
            int i0 = 0;
            for (; i0 + 8 <= Li4_0; i0 += 8) {
                const int ia = i0 + (lane >> 4), ib = ia + 4;
                e = __builtin_amdgcn_mfma_f32_16x16x4f32(acol[ia ^ sc], brow[(size_t)ia * B], e, 0, 0, 0);
                e1 = __builtin_amdgcn_mfma_f32_16x16x4f32(acol[ib ^ sc], brow[(size_t)ib * B], e1, 0, 0, 0);
            }
            if (i0 < Li4_0) {
                const int ia = i0 + (lane >> 4);
                e = __builtin_amdgcn_mfma_f32_16x16x4f32(acol[ia ^ sc], brow[(size_t)ia * B], e, 0, 0, 0);
            }
            e += e1;
        }
        if (g1.Lp > 0) {
            const float* brow = NI + g1.offI + j0 + (lane & 15);
            fvec4 o1 = {0, 0, 0, 0};
            int i0 = 0;
            for (; i0 + 8 <= Li4_1; i0 += 8) {
                const int ia = i0 + (lane >> 4), ib = ia + 4;
                o = __builtin_amdgcn_mfma_f32_16x16x4f32(acol[(Li4_0 + ia) ^ sc], brow[(size_t)ia * B], o, 0, 0, 0);
                o1 = __builtin_amdgcn_mfma_f32_16x16x4f32(acol[(Li4_0 + ib) ^ sc], brow[(size_t)ib * B], o1, 0, 0, 0);
            }
            if (i0 < Li4_1) {
                const int ia = i0 + (lane >> 4);
                o = __builtin_amdgcn_mfma_f32_16x16x4f32(acol[(Li4_0 + ia) ^ sc], brow[(size_t)ia * B], o, 0, 0, 0);
            }
            o += o1;
        }
        const int j = j0 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = cbase + c0 + Mma<float>::row(lane, r);
            if (col >= ncols) continue;
            const int f = col / cpf, rem = col - f * cpf;
            const int row = (rem >> 1) ? N - m : m;
            const float s = (row > B && (row & 1)) ? -inv_s2pi : inv_s2pi;
            const float yn = (e[r] + o[r]) * s, ys = (e[r] - o[r]) * s;
            float* xr = reinterpret_cast<float*>(X + ((size_t)f * N + row) * N) + (rem & 1);
            xr[2 * j] = yn;
            xr[2 * (N - 1 - j)] = ys;
            if (fmt == 1 && m > 0) {
                float* xm = reinterpret_cast<float*>(X + ((size_t)f * N + (N - m)) * N) + (rem & 1);
                const float cs = (rem & 1) ? -1.f : 1.f;
                xm[2 * j] = cs * yn;
                xm[2 * (N - 1 - j)] = cs * ys;
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

template <typename R>
static void fwd_impl(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                     double* cim, int F, Fmt fmt, void* Xv, void* Cbv, hipStream_t s, StageTimer* tm) {
    using T = typename Cplx<R>::type;
    constexpr int V = sizeof(R) == 8 ? 8 : 16;
    const int B = c.bw, N = c.N, lgN = lg2(N);
    const int E = V * kThreads, rpb = E / N;
    const size_t lds = sizeof(T) * (size_t)rpb * (N + 1);
    T* X = (T*)Xv;
    R* Cb = (R*)Cbv;
    const T* tw = (const T*)(sizeof(R) == 8 ? (const void*)c.tw : (const void*)c.tw32);
    const T* ph = (const T*)(sizeof(R) == 8 ? (const void*)c.ph : (const void*)c.ph32);
    const R* ws = (const R*)(sizeof(R) == 8 ? (const void*)c.w_semi : (const void*)c.w_semi32);
    {
        Span sp(tm, ST_PHI_FWD, s);
        k_fphi_fwd<R, V><<<dim3(N / rpb, F), kThreads, lds, s>>>(re, im, X, lgN, tw);
        S2_HIP(hipGetLastError());
    }
    const int nrows = (fmt == Fmt::Complex) ? 2 * B - 1 : B;
    {
        Span sp(tm, ST_DCT_FWD, s);
        k_fdct_fwd<R, V><<<dim3((nrows + rpb - 1) / rpb, F), kThreads, lds, s>>>(X, Cb, ws, lgN, B, nrows, tw, ph);
        S2_HIP(hipGetLastError());
    }
    {
        const int ncb = gemm_ncb(B, g.prec);
        const int cols = 4 * F;
        Span sp(tm, ST_LEG_FWD, s);
        k_gemm_fwd<R><<<dim3(B, (cols + ncb - 1) / ncb), kThreads, gemm_lds(B, g.prec), s>>>(
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
    const size_t lds = sizeof(T) * (size_t)rpb * (N + 1);
    T* X = (T*)Xv;
    R* Cb = (R*)Cbv;
    const T* tw = (const T*)(sizeof(R) == 8 ? (const void*)c.tw : (const void*)c.tw32);
    const T* ph = (const T*)(sizeof(R) == 8 ? (const void*)c.ph : (const void*)c.ph32);
    const R* sv = (const R*)(sizeof(R) == 8 ? (const void*)c.sinv : (const void*)c.sinv32);
    {
        const int ncb = gemm_ncb(B, g.prec);
        const int cols = 4 * F;
        Span sp(tm, ST_LEG_INV, s);
        k_gemm_inv<R><<<dim3(B, (cols + ncb - 1) / ncb), kThreads, gemm_lds(B, g.prec), s>>>(
            cre, cim, (const R*)g.ai, g.meta, Cb, lg2(B), F, ncb, (int)fmt);
        S2_HIP(hipGetLastError());
    }
    const int nrows = (fmt == Fmt::Complex) ? 2 * B - 1 : B;
    {
        Span sp(tm, ST_DCT_INV, s);
        k_fdct_inv<R, V><<<dim3((nrows + rpb - 1) / rpb, F), kThreads, lds, s>>>(Cb, X, sv, lgN, B, nrows, (int)fmt,
                                                                                 tw, ph);
        S2_HIP(hipGetLastError());
    }
    {
        Span sp(tm, ST_PHI_INV, s);
        k_fphi_inv<R, V><<<dim3(N / rpb, F), kThreads, lds, s>>>(X, re, im, lgN, tw);
        S2_HIP(hipGetLastError());
    }
}

bool mixed_supported(int bw) {
    const int N = 2 * bw;
    return bw >= 32 && bw <= 512 && !(N & (N - 1));
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
            of += (long)B * ((e.Lp + 15) & ~15);
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
        k_fphi_fwd<float, V><<<dim3(N / rpb, F), kThreads, sizeof(float2) * (size_t)rpb * (N + 1), s>>>(
            re, im, X, lgN, c.tw32);
        S2_HIP(hipGetLastError());
    }
    {
        const int ncb = ngemm_ncb(B, true);
        Span sp(tm, ST_LEG_FWD, s);
        k_ngemm_fwd<<<dim3(B, (4 * F + ncb - 1) / ncb), kThreads, sizeof(float) * 2 * B * ncb, s>>>(
            X, (const float*)g.af, g.meta, cre, cim, lg2(B), F, ncb, (int)fmt);
        S2_HIP(hipGetLastError());
    }
}

void inverse_mixed(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                   double* im, int F, Fmt fmt, void* Xv, hipStream_t s, StageTimer* tm) {
    constexpr int V = 16;
    const int B = c.bw, N = c.N, lgN = lg2(N);
    const int rpb = V * kThreads / N;
    float2* X = (float2*)Xv;
    {
        const int ncb = ngemm_ncb(B, false);
        Span sp(tm, ST_LEG_INV, s);
        k_ngemm_inv<<<dim3(B, (4 * F + ncb - 1) / ncb), kThreads, sizeof(float) * B * ncb, s>>>(
            cre, cim, (const float*)g.ai, g.meta, X, lg2(B), F, ncb, (int)fmt);
        S2_HIP(hipGetLastError());
    }
    {
        Span sp(tm, ST_PHI_INV, s);
        k_fphi_inv<float, V><<<dim3(N / rpb, F), kThreads, sizeof(float2) * (size_t)rpb * (N + 1), s>>>(
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
