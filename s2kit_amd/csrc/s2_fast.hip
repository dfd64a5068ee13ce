// Batched GEMM path, FP64 or FP32 arithmetic (FP64 inputs and outputs either
// way).  Power-of-two grids (N = 2*bw = 64 .. 2048):
//
//   k_fphi_fwd -> k_fdct_fwd -> k_gemm_fwd     forward  (ring FFT, DCT: s2_blocks.h)
//   k_gemm_inv -> k_fdct_inv -> k_fphi_inv     inverse
//
// Other bw in [16, 1024] (FP64, GEN): the per-field kernels' mixed-radix or
// Bluestein FFTs and DCTs (s2_kernels.hip) around the same GEMMs; those DCTs
// then read and write the same parity-split re/im planes (their planar mode),
// whose rows the GEMMs stream in whole 128-byte lines (with natural-order
// complex rows the inverse GEMM's stores moved 1.66x their bytes at bw=255).
//
//   k_gemm_fwd   seminaive projections as MFMA GEMMs            (seminaive.c:183-197)
//                naive projections (orders >= cutoff) likewise  (naive.c:34-59)
//   k_gemm_inv   transposed-table cosine sums as MFMA GEMMs     (seminaive.c:74-98)
//                naive synthesis (orders >= cutoff) into X      (naive.c:78-95)
//
// With cutoff < bw the DCT stages run over the seminaive order rows only;
// the naive orders' GEMMs read (forward) or write (inverse) the ring-FFT
// rows X directly.  One launch per direction covers all orders, the branch
// taken per workgroup.
//
// Legendre GEMMs.  For order m split the degrees by parity p = (l-m) mod 2:
// degree l = m+p+2i reads cosine coefficients c[2t+p], t < RowSize(m,l), so
// each (m,p) is a dense matrix A[i][t] (zero above the staircase).  The
// stripped tables are repacked once per plan (k_pack_gemm) into
//   AF(m,p)  [t][i]  (t < roundup4(Kf), i < roundup16(Lp))   forward
//   AI(m,p)  [i][t]  (i < roundup4(Lp), t < roundup16(ceil(bw/2))) inverse,
//            built from the transposed table with its scaling folded in
// and the data columns are (field, +-m, re/im).  Per 16x16 tile the K loop
// stops at the staircase, so the padded zeros cost at most one tile row.
#include <algorithm>
#include <cmath>
#include <vector>

#include "s2_blocks.h"
#include "s2_recurrence.h"

namespace s2 {

// ---------------------------------------------------------------------------
// GEMM tables
// ---------------------------------------------------------------------------
// Extents shared by the packing and the GEMMs (bw = B): cosine indices of
// parity p number hE = ceil(B/2) (p = 0) and hO = floor(B/2) (p = 1);
// inverse seminaive rows are roundup16(hE) long; naive forward tables have
// roundup4(B) ring rows and naive inverse rows roundup16(B) rings, zero past B.
__host__ __device__ inline int pk_tp16(int B) { return ((B + 1) / 2 + 15) & ~15; }
__host__ __device__ inline int pk_b4(int B) { return (B + 3) & ~3; }
__host__ __device__ inline int pk_b16(int B) { return (B + 15) & ~15; }

template <typename R>
__global__ void k_pack_gemm(const double* __restrict__ T, const double* __restrict__ TT, const long* __restrict__ toff,
                            const GemmMeta* __restrict__ meta, R* __restrict__ AF, R* __restrict__ AI, int B,
                            int cut) {
    extern __shared__ __align__(16) unsigned char smem[];
    int* rowpos = reinterpret_cast<int*>(smem);
    const int m = blockIdx.x, p = blockIdx.y;
    const int N = 2 * B;
    const GemmMeta g = meta[m * 2 + p];
    if (m >= cut) {  // naive order: sampled P_l^m rows (pml.c:41-79), both directions the same table
        const int Li16 = (g.Lp + 15) & ~15, Li4 = (g.Lp + 3) & ~3;
        const int B4 = pk_b4(B), B16 = pk_b16(B);
        const double* pm = T + toff[m];
        const double* pt = TT + toff[m];
        for (int idx = threadIdx.x; idx < B4 * Li16; idx += blockDim.x) {
            const int j = idx / Li16, i = idx - j * Li16;
            AF[g.offF + idx] = (R)(i < g.Lp && j < B ? pm[(size_t)(p + 2 * i) * N + j] : 0.0);
        }
        const double inv_s2pi = 0.39894228040143267794;
        for (int idx = threadIdx.x; idx < Li4 * B16; idx += blockDim.x) {
            const int i = idx / B16, j = idx - i * B16;
            AI[g.offI + idx] = (R)(i < g.Lp && j < B ? pt[(size_t)(p + 2 * i) * N + j] * inv_s2pi : 0.0);
        }
        return;
    }
    for (int i = threadIdx.x; i < B; i += blockDim.x) rowpos[i + 1] = transpose_row_size(i, m, B);
    __syncthreads();
    if (threadIdx.x == 0) {
        rowpos[0] = 0;
        for (int i = 1; i <= B; ++i) rowpos[i] += rowpos[i - 1];
    }
    __syncthreads();
    const int Li16 = (g.Lp + 15) & ~15, Kf4 = (g.Kf + 3) & ~3;
    const int Li4 = (g.Lp + 3) & ~3, Tp16 = pk_tp16(B), cnt = p ? B / 2 : (B + 1) / 2;
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
        if (i < g.Lp && t < cnt) {
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
// LDS row padding: a row stride of bw + 4 words (FP32) / bw + 2 doubles (FP64)
// keeps both the coalesced fill (consecutive k) and the MFMA operand reads
// (16 columns x 4 consecutive k per wave) free of bank conflicts.
template <typename R>
__host__ __device__ constexpr int lds_pad() {
    return sizeof(R) == 4 ? 4 : 2;
}

// Extents of the GEMMs at bandwidth B.  Cosine index k = 2t + p: hE = ceil(B/2)
// even and hO = floor(B/2) odd ones, at positions t and hE + t of the Cb
// planes.  Any B that is not a power of two (gen) takes an LDS row long
// enough for the K loops' round-up (B + 3) and the inverse's two padded
// parities (B + 6), zero past B.
struct GGeo {
    int B, N, hE, hO, rs, B4, B16, Tp16;
};
template <typename R>
__host__ __device__ inline GGeo ggeo(int B, bool gen) {
    GGeo g;
    g.B = B;
    g.N = 2 * B;
    g.hE = (B + 1) / 2;
    g.hO = B / 2;
    g.rs = gen ? B + 8 : B + lds_pad<R>();
    g.B4 = (B + 3) & ~3;
    g.B16 = (B + 15) & ~15;
    g.Tp16 = (g.hE + 15) & ~15;
    return g;
}

#ifndef S2_GEMM_RUN
#define S2_GEMM_RUN 8
#endif
// S K steps of MFMAs for CT column tiles on table values g already loaded.
template <typename R, int CT, int S>
__device__ __forceinline__ void mma_batch(typename Mma<R>::acc (&acc)[CT], const R* lds, int cs, const R (&g)[S],
                                          int k, int kq) {
    R l[CT][S];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int u = 0; u < S; ++u) l[c][u] = lds[c * cs + k + 4 * u + kq];
#pragma unroll
    for (int u = 0; u < S; ++u)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = Mma<R>::mma(l[c][u], g[u], acc[c]);
}

// The table values of S K steps (one per lane and step).
template <typename R, int S>
__device__ __forceinline__ void table_batch(R (&g)[S], const R* __restrict__ glb, size_t stride, int k, int kq) {
#pragma unroll
    for (int u = 0; u < S; ++u) g[u] = glb[(size_t)(k + 4 * u + kq) * stride];
}

// acc[c] += sum over k in [k0, k1), step 4, of the 16x16x4 MFMA of the LDS
// operand lds[c*cs + k + kq] (column tile c < CT) and the table operand
// glb[(k + kq) * stride], kq = lane/16.  Each table value feeds CT MFMAs, and
// the table values of S steps are loaded together, one batch ahead of their
// MFMAs (two register buffers), so the L2 round trips overlap the MFMAs
// instead of each MFMA waiting for its own load.  The K order of every
// accumulator is ascending whatever S, CT or the column blocking.
template <typename R, int CT, int S = S2_GEMM_RUN>
__device__ __forceinline__ void mma_run(typename Mma<R>::acc (&acc)[CT], const R* lds, int cs,
                                        const R* __restrict__ glb, size_t stride, int k0, int k1) {
    const int kq = (threadIdx.x & 63) >> 4;
    int k = k0;
    // two table buffers: batch k + 4S is in flight while batch k's MFMAs run
    // (one buffer, each batch waited before its MFMAs: FP32 bw=256 forward
    // GEMM 1.84 ms against 1.49)
    R ga[S], gb[S];
    if (k + 4 * S <= k1) table_batch<R, S>(ga, glb, stride, k, kq);
    while (k + 4 * S <= k1) {
        if (k + 8 * S <= k1) table_batch<R, S>(gb, glb, stride, k + 4 * S, kq);
        mma_batch<R, CT, S>(acc, lds, cs, ga, k, kq);
        k += 4 * S;
        if (k + 4 * S > k1) break;
        if (k + 8 * S <= k1) table_batch<R, S>(ga, glb, stride, k + 4 * S, kq);
        mma_batch<R, CT, S>(acc, lds, cs, gb, k, kq);
        k += 4 * S;
    }
    for (; k < k1; k += 4) {
        const R g = glb[(size_t)(k + kq) * stride];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = Mma<R>::mma(lds[c * cs + k + kq], g, acc[c]);
    }
}

// Output tiles of a block: its waves take every (kThreads/64 * gridDim.z)-th
// tile, slice blockIdx.z of the tiles when grid.z splits an order's column
// block over several blocks (calls of few columns, whose blocks are each one
// order's table stream).  A tile's sums do not depend on the split.
__device__ __forceinline__ int item0(int wave) { return wave + (kThreads / 64) * blockIdx.z; }
__device__ __forceinline__ int item_step() { return (kThreads / 64) * gridDim.z; }

// Orders m of a GEMM launch: grid (column blocks, roundup8(bw), tile slices);
// blocks past the last order return at once.
__device__ __forceinline__ bool gemm_block(int B, int& bx, int& m) {
    xcd_block_m(bx, m);
    return m < B;
}

// Coefficient (m, l) of data column col (field, +-m row, re/im) from the
// projection v; REAL data also fills the -m run (f(l,-m) = (-1)^m conj f(l,m)).
__device__ __forceinline__ void fwd_store(double v, int col, int m, int l, int B, int cpf, int fmt,
                                          double* __restrict__ cre, double* __restrict__ cim) {
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    const int f = col / cpf, rem = col - f * cpf;
    const int comp = rem & 1;
    double* out = comp ? cim : cre;
    if ((rem >> 1) == 0) {
        out[f * B2 + index_of_coeff(m, l, B)] = v;
        if (fmt == 1 && m > 0) out[f * B2 + index_of_coeff(-m, l, B)] = comp ? -sg * v : sg * v;
    } else {
        out[f * B2 + index_of_coeff(-m, l, B)] = sg * v;
    }
}

// Seminaive order m < cutoff.  LDS: Bd[c][k2], k2 = parity-split cosine
// position (p*hE + t).
template <typename R, bool GEN, int CT>
__device__ __forceinline__ void gemm_fwd_semi(R* __restrict__ Bd, const R* __restrict__ Cb, const R* __restrict__ AF,
                                              const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                              double* __restrict__ cim, int m, int cbase, const GGeo& G, int F,
                                              int ncb, int fmt) {
    const int B = G.B, N = G.N;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int rs = G.rs;
    if constexpr (GEN) {  // the same planes at any B; zero past B (the K loops' round-up)
        for (int idx = threadIdx.x; idx < ncb * rs; idx += kThreads) {
            const int c = idx / rs, k2 = idx - c * rs;
            const int col = cbase + c;
            R v = 0;
            if (col < ncols && k2 < B) {
                const int f = col / cpf, rem = col - f * cpf;
                const int r = (rem >> 1) ? N - m : m;
                v = Cb[(((size_t)f * N + r) * 2 + (rem & 1)) * B + k2];
            }
            Bd[c * rs + k2] = v;
        }
    } else {
        const int lgB = __builtin_ctz(B);
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
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nctg = (ncb >> 4) / CT;  // groups of CT column tiles
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = meta[m * 2 + p];
        if (g.Lp == 0) continue;
        const int Li16 = (g.Lp + 15) & ~15;
        const int ntiles = (Li16 >> 4) * nctg;
        for (int item = item0(wave); item < ntiles; item += item_step()) {
            const int it = item / nctg, ct = item - it * nctg;
            const int i0 = it * 16, c0 = ct * 16 * CT;
            const int ilast = min(i0 + 15, g.Lp - 1);
            const int kend = (row_size(m, m + p + 2 * ilast) + 3) & ~3;
            typename Mma<R>::acc acc[CT] = {};
            const R* bcol = Bd + (c0 + (lane & 15)) * rs + p * G.hE;
            const R* arow = AF + g.offF + i0 + (lane & 15);
            mma_run<R, CT>(acc, bcol, 16 * rs, arow, Li16, 0, kend);
            const int i = i0 + (lane & 15);
            if (i >= g.Lp) continue;
#pragma unroll
            for (int c = 0; c < CT; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + c0 + 16 * c + Mma<R>::row(lane, r);
                    if (col < ncols) fwd_store((double)acc[c][r], col, m, m + p + 2 * i, B, cpf, fmt, cre, cim);
                }
        }
    }
}

// Naive order m >= cutoff (naive.c:34-59), from the ring-FFT rows X.  The
// grid is symmetric, x_{N-1-j} = -x_j, so P_l^m(x_{N-1-j}) = (-1)^{l-m}
// P_l^m(x_j): degree l = m+p+2i is the sum over the bw rings j < bw of
// (w_j X_j + (-1)^p w_{N-1-j} X_{N-1-j}) P_l^m(x_j).  One parity at a time,
// LDS: Bd[c][j], zero for j >= bw.
template <typename R, int CT>
__device__ __forceinline__ void gemm_fwd_naive(R* __restrict__ Bd, const typename Cplx<R>::type* __restrict__ X,
                                               const double* __restrict__ wn, const R* __restrict__ AF,
                                               const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                               double* __restrict__ cim, int m, int cbase, const GGeo& G, int F,
                                               int ncb, int fmt) {
    using T = typename Cplx<R>::type;
    const int B = G.B, N = G.N, B4 = G.B4;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const int rs = G.rs;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nctg = (ncb >> 4) / CT;
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = meta[m * 2 + p];
        if (g.Lp == 0) break;
        if (p) __syncthreads();  // parity 0's reads of Bd are done
        for (int idx = threadIdx.x; idx < ncb * B4; idx += kThreads) {
            const int c = idx / B4, j = idx - c * B4;
            const int col = cbase + c;
            R v = 0;
            if (col < ncols && j < B) {
                const int f = col / cpf, rem = col - f * cpf;
                const T* xr = X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N;
                const T a = xr[j], b = xr[N - 1 - j];
                const R wa = (R)wn[j] * ((rem & 1) ? a.y : a.x);
                const R wb = (R)wn[N - 1 - j] * ((rem & 1) ? b.y : b.x);
                v = p ? wa - wb : wa + wb;
            }
            Bd[c * rs + j] = v;
        }
        __syncthreads();
        const int Li16 = (g.Lp + 15) & ~15;
        const int ntiles = (Li16 >> 4) * nctg;
        for (int item = item0(wave); item < ntiles; item += item_step()) {
            const int it = item / nctg, ct = item - it * nctg;
            const int i0 = it * 16, c0 = ct * 16 * CT;
            typename Mma<R>::acc acc[CT] = {};
            const R* bcol = Bd + (c0 + (lane & 15)) * rs;
            const R* arow = AF + g.offF + i0 + (lane & 15);
            mma_run<R, CT>(acc, bcol, 16 * rs, arow, Li16, 0, B4);
            const int i = i0 + (lane & 15);
            if (i >= g.Lp) continue;
#pragma unroll
            for (int c = 0; c < CT; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + c0 + 16 * c + Mma<R>::row(lane, r);
                    if (col < ncols) fwd_store((double)acc[c][r], col, m, m + p + 2 * i, B, cpf, fmt, cre, cim);
                }
        }
    }
}

// Forward: D'[c][i] = sum_k data[c][k] * A(m,p)[i][k].  grid (ceil(cols/ncb), roundup8(bw)).
template <typename R, bool GEN, int CT>
__global__ __launch_bounds__(kThreads) void k_gemm_fwd(const R* __restrict__ Cb,
                                                       const typename Cplx<R>::type* __restrict__ X,
                                                       const double* __restrict__ wn, const R* __restrict__ AF,
                                                       const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                       double* __restrict__ cim, int B, int cut, int mlo, int F,
                                                       int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* Bd = reinterpret_cast<R*>(smem);
    int bx, m;
    if (!gemm_block(B, bx, m) || m < mlo) return;  // orders below mlo: the K-chunked launch
    const int cbase = bx * ncb;
    if (cbase >= F * ((m > 0 && fmt == 0) ? 4 : 2)) return;
    const GGeo G = ggeo<R>(B, GEN);
    if (m < cut)
        gemm_fwd_semi<R, GEN, CT>(Bd, Cb, AF, meta, cre, cim, m, cbase, G, F, ncb, fmt);
    else
        gemm_fwd_naive<R, CT>(Bd, X, wn, AF, meta, cre, cim, m, cbase, G, F, ncb, fmt);
}

// Coefficient columns of order m into LDS: Cd[c][q], q = i (p = 0) or
// roundup4(Lp0) + i (p = 1), with the padding rows of both parities zeroed;
// q < bw + 6.  `negodd`: negate the -m columns of odd orders.
template <typename R>
__device__ __forceinline__ void inv_fill(R* __restrict__ Cd, const double* __restrict__ cre,
                                         const double* __restrict__ cim, const GemmMeta& g0, const GemmMeta& g1,
                                         int m, int cbase, const GGeo& G, int ncb, int cpf, int ncols, bool negodd) {
    const int B = G.B;
    const int Li4_0 = (g0.Lp + 3) & ~3, Li4_1 = (g1.Lp + 3) & ~3;
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    const int rs = G.rs;
    const bool neg = negodd && (m & 1);
    for (int idx = threadIdx.x; idx < ncb * L; idx += kThreads) {
        const int c = idx / L, li = idx - c * L;
        const int col = cbase + c;
        const int q = (li & 1) ? Li4_0 + (li >> 1) : (li >> 1);
        R v = 0;
        if (col < ncols) {
            const int f = col / cpf, rem = col - f * cpf;
            const int mm = (rem >> 1) ? -m : m;
            const double* src = (rem & 1) ? cim : cre;
            const double x = src[f * B2 + index_of_coeff(mm, m + li, B)];
            v = (R)((neg && (rem >> 1)) ? -x : x);
        }
        Cd[c * rs + q] = v;
    }
    for (int idx = threadIdx.x; idx < ncb * 8; idx += kThreads) {
        const int c = idx >> 3, k = idx & 7;
        const int p = k >> 2, i = (p ? g1.Lp : g0.Lp) + (k & 3);
        if (i < (p ? Li4_1 : Li4_0)) Cd[c * rs + (p ? Li4_0 : 0) + i] = 0;
    }
}

// Seminaive order m < cutoff: D'[c][t] = sum_i coef[c][i] * AI(m,p)[i][t],
// cosine coefficients k = 2t + p into the Cb row of (field, +-m).
template <typename R, bool GEN, int CT>
__device__ __forceinline__ void gemm_inv_semi(R* __restrict__ Cd, const double* __restrict__ cre,
                                              const double* __restrict__ cim, const R* __restrict__ AI,
                                              const GemmMeta* __restrict__ meta, R* __restrict__ Cb, int m, int cbase,
                                              const GGeo& G, int F, int ncb, int fmt) {
    const int B = G.B, N = G.N;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int Li4_0 = (g0.Lp + 3) & ~3, Li4_1 = (g1.Lp + 3) & ~3;
    const int rs = G.rs;
    inv_fill<R>(Cd, cre, cim, g0, g1, m, cbase, G, ncb, cpf, ncols, false);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nctg = (ncb >> 4) / CT;
    const int Tp16 = G.Tp16;
    const int ntt = Tp16 >> 4;
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = p ? g1 : g0;
        const int Li4 = p ? Li4_1 : Li4_0;
        const int cnt = p ? G.hO : G.hE;
        const int r0 = g.Lp > 0 ? row_size(m, m + p) : 0;
        for (int item = item0(wave); item < ntt * nctg; item += item_step()) {
            const int tt = item / nctg, ct = item - tt * nctg;
            const int t0 = tt * 16, c0 = ct * 16 * CT;
            const int istart = max(0, t0 - r0 + 1) & ~3;
            typename Mma<R>::acc acc[CT] = {};
            const R* acol = Cd + (c0 + (lane & 15)) * rs + (p ? Li4_0 : 0);
            const R* brow = AI + g.offI + t0 + (lane & 15);
            mma_run<R, CT>(acc, acol, 16 * rs, brow, Tp16, istart, Li4);
            const int t = t0 + (lane & 15);
            if (t >= cnt) continue;
#pragma unroll
            for (int c = 0; c < CT; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + c0 + 16 * c + Mma<R>::row(lane, r);
                    if (col >= ncols) continue;
                    const int f = col / cpf, rem = col - f * cpf;
                    const int row = (rem >> 1) ? N - m : m;
                    Cb[(((size_t)f * N + row) * 2 + (rem & 1)) * B + p * G.hE + t] = acc[c][r];
                }
        }
    }
}

// Naive order m >= cutoff (naive.c:78-95), straight into the ring-FFT rows
// X: per ring j < bw the parity sums E (p = 0) and O (p = 1) over the table
// rows (which carry 1/sqrt(2pi)) give ring j <- E+O and ring N-1-j <- E-O.
// Row N-m of an odd order carries a minus sign (FST_semi_memo.c:301-348),
// put on its coefficients; REAL data also fills the conjugate row N-m.
template <typename R, int CT>
__device__ __forceinline__ void gemm_inv_naive(R* __restrict__ Cd, const double* __restrict__ cre,
                                               const double* __restrict__ cim, const R* __restrict__ AI,
                                               const GemmMeta* __restrict__ meta, typename Cplx<R>::type* __restrict__ X,
                                               int m, int cbase, const GGeo& G, int F, int ncb, int fmt) {
    const int B = G.B, N = G.N, B16 = G.B16;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int Li4_0 = (g0.Lp + 3) & ~3, Li4_1 = (g1.Lp + 3) & ~3;
    const int rs = G.rs;
    inv_fill<R>(Cd, cre, cim, g0, g1, m, cbase, G, ncb, cpf, ncols, true);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nctg = (ncb >> 4) / CT;
    const int ntt = B16 >> 4;
    const bool mirror = fmt == 1 && m > 0;
    for (int item = item0(wave); item < ntt * nctg; item += item_step()) {
        const int tt = item / nctg, ct = item - tt * nctg;
        const int t0 = tt * 16, c0 = ct * 16 * CT;
        typename Mma<R>::acc e[CT] = {}, o[CT] = {};
        const R* acol = Cd + (c0 + (lane & 15)) * rs;
        const R* b0 = AI + g0.offI + t0 + (lane & 15);
        mma_run<R, CT>(e, acol, 16 * rs, b0, B16, 0, Li4_0);
        const R* b1 = AI + g1.offI + t0 + (lane & 15);
        mma_run<R, CT>(o, acol + Li4_0, 16 * rs, b1, B16, 0, Li4_1);
        const int j = t0 + (lane & 15);
        if (j >= B) continue;
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = cbase + c0 + 16 * c + Mma<R>::row(lane, r);
                if (col >= ncols) continue;
                const int f = col / cpf, rem = col - f * cpf, comp = rem & 1;
                R* xr = reinterpret_cast<R*>(X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N) + comp;
                const R lo = e[c][r] + o[c][r], hi = e[c][r] - o[c][r];
                xr[2 * j] = lo;
                xr[2 * (N - 1 - j)] = hi;
                if (mirror) {  // conj into row N-m
                    R* xm = reinterpret_cast<R*>(X + ((size_t)f * N + (N - m)) * N) + comp;
                    xm[2 * j] = comp ? -lo : lo;
                    xm[2 * (N - 1 - j)] = comp ? -hi : hi;
                }
            }
    }
}

// Inverse.  grid (ceil(cols/ncb), roundup8(bw)).
template <typename R, bool GEN, int CT>
__global__ __launch_bounds__(kThreads) void k_gemm_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                       const R* __restrict__ AI, const GemmMeta* __restrict__ meta,
                                                       R* __restrict__ Cb, typename Cplx<R>::type* __restrict__ X,
                                                       int B, int cut, int mlo, int F, int ncb, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* Cd = reinterpret_cast<R*>(smem);
    int bx, m;
    if (!gemm_block(B, bx, m) || m < mlo) return;  // orders below mlo: the K-chunked launch
    const int cbase = bx * ncb;
    if (cbase >= F * ((m > 0 && fmt == 0) ? 4 : 2)) return;
    const GGeo G = ggeo<R>(B, GEN);
    if (m < cut)
        gemm_inv_semi<R, GEN, CT>(Cd, cre, cim, AI, meta, Cb, m, cbase, G, F, ncb, fmt);
    else
        gemm_inv_naive<R, CT>(Cd, cre, cim, AI, meta, X, m, cbase, G, F, ncb, fmt);
}

// ---------------------------------------------------------------------------
// K-chunked seminaive GEMMs (every order seminaive, many columns)
// ---------------------------------------------------------------------------
// The kernels above hold whole data rows in LDS: at bw=256 FP64 a column takes
// 2 KiB, so four blocks per CU leave 16 columns per block and every table value
// loaded feeds one MFMA.  Here a block owns kKcCols = 64 columns of one order
// and the data stream through two LDS chunk buffers of kKc = 16 K positions
// (staged in registers one chunk ahead); wave w owns the 16-wide tile w of
// each parity (eight waves up to bw = 256, sixteen at bw = 512: one wave per
// tile) with all four column tiles, so a table value feeds four
// MFMAs and a block reads its order's table once per 64 columns.  The table
// values of the next chunk are loaded before the current chunk's MFMAs.  Each
// accumulator takes its K steps in the same ascending order as the kernels
// above: the results are the same.
constexpr int kKcCols = 64, kKc = 16, kKcNS = 1, kKcRS = kKc + 2;
constexpr int kKcSteps = kKc / 4;                   // K steps of 4 per chunk

// Data column col of order m: the start of its Cb row (field, +-m, re/im).
template <typename R>
__device__ __forceinline__ const R* kc_row(const R* __restrict__ Cb, int col, int m, int N, int B, int cpf) {
    const int f = col / cpf, rem = col - f * cpf;
    const int r = (rem >> 1) ? N - m : m;
    return Cb + (((size_t)f * N + r) * 2 + (rem & 1)) * B;
}

// Forward (seminaive.c:183-197): D'[c][i] = sum_t data[c][p*hE + t] AF(m,p)[t][i].
template <typename R, int NTH>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(4))) void k_gemm_fwd_kc(const R* __restrict__ Cb, const R* __restrict__ AF,
                                                          const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                          double* __restrict__ cim, int B, int cut, int F, int fmt) {
    __shared__ R Bd[2][kKcCols * kKcRS];
    int bx, m;
    if (!gemm_block(B, bx, m) || m >= cut) return;  // seminaive orders only
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf, cbase = bx * kKcCols;
    if (cbase >= ncols) return;
    const int N = 2 * B, hE = (B + 1) / 2, hO = B / 2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, kq = lane >> 4, ln = lane & 15;
    constexpr int kKcWaves = NTH / 64, kKcPer = kKcCols * kKc / NTH;  // waves; chunk values staged per thread
    const int fc = tid / (kKc / kKcPer), fk = kKcPer * (tid % (kKc / kKcPer));  // staging: column fc, positions fk ..
    const bool fcol = cbase + fc < ncols;
    const R* frow = fcol ? kc_row(Cb, cbase + fc, m, N, B, cpf) : Cb;
    for (int p = 0; p < 2; ++p) {
        const GemmMeta g = meta[m * 2 + p];
        if (g.Lp == 0) continue;
        const int Li16 = (g.Lp + 15) & ~15, ntile = Li16 >> 4;
        const int cnt = p ? hO : hE;
        const R* src = frow + p * hE;
        const R* ga = AF + g.offF;  // this parity's table (32-bit offsets from here)
        // K extent of each of this wave's tiles (multiples of 4), and of the block
        int kend[kKcNS];
#pragma unroll
        for (int s = 0; s < kKcNS; ++s) {
            const int it = wave + kKcWaves * s;
            kend[s] = it < ntile ? (row_size(m, m + p + 2 * min(16 * it + 15, g.Lp - 1)) + 3) & ~3 : 0;
        }
        const int nq = ((row_size(m, m + p + 2 * (g.Lp - 1)) + 3) / 4 * 4 + kKc - 1) / kKc;
        typename Mma<R>::acc acc[kKcNS][4] = {};
        R v[kKcPer], ta[kKcNS][kKcSteps], tb[kKcNS][kKcSteps];
        bool vok[kKcPer];
        auto stage = [&](int q) __attribute__((always_inline)) {
#pragma unroll
            for (int e = 0; e < kKcPer; ++e) {
                const int t = q * kKc + fk + e;
                vok[e] = fcol && t < cnt;
                v[e] = *(vok[e] ? src + t : Cb);  // unconditional load (see table); zeroed in put
            }
        };
        auto table = [&](int q, R (&tv)[kKcNS][kKcSteps]) __attribute__((always_inline)) {
#pragma unroll
            for (int s = 0; s < kKcNS; ++s)
#pragma unroll
                for (int u = 0; u < kKcSteps; ++u) {
                    // Every lane loads, from the array's first entry when the
                    // entry is past its tile (a load under a branch left the
                    // compiler unable to count loads in flight, and it waited
                    // for all of them, this prefetch included, before each
                    // MFMA).  Such values are never used: chunk() runs an
                    // MFMA only under the same condition.
                    const int k = q * kKc + 4 * u;
                    tv[s][u] = *(k < kend[s] ? ga + ((k + kq) * Li16 + 16 * (wave + kKcWaves * s) + ln) : AF);
                }
        };
        auto put = [&](int b) __attribute__((always_inline)) {
#pragma unroll
            for (int e = 0; e < kKcPer; ++e) Bd[b][fc * kKcRS + fk + e] = vok[e] ? v[e] : (R)0;
        };
        auto chunk = [&](int q, const R (&tv)[kKcNS][kKcSteps]) __attribute__((always_inline)) {
            const R* bd = Bd[q & 1] + ln * kKcRS + kq;
#pragma unroll
            for (int s = 0; s < kKcNS; ++s)
#pragma unroll
                for (int u = 0; u < kKcSteps; ++u)
                    if (q * kKc + 4 * u < kend[s]) {
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            acc[s][c] = Mma<R>::mma(bd[16 * c * kKcRS + 4 * u], tv[s][u], acc[s][c]);
                    }
        };
        stage(0);
        table(0, ta);
        put(0);
        __syncthreads();
        for (int q = 0; q < nq; q += 2) {
            // The next chunk's loads, issued unconditionally (past the last
            // chunk they read the arrays' first entries, unused): with the
            // count of loads in flight fixed, each MFMA waits only for the
            // table value it reads, loaded one chunk earlier, and the empty
            // asm keeps the loads ahead of this chunk's MFMAs.
            stage(q + 1);
            table(q + 1, tb);
            asm volatile("" ::: "memory");
            chunk(q, ta);
            if (q + 1 < nq) put(1);
            __syncthreads();
            if (q + 1 >= nq) break;
            stage(q + 2);
            table(q + 2, ta);
            asm volatile("" ::: "memory");
            chunk(q + 1, tb);
            if (q + 2 < nq) put(0);
            __syncthreads();
        }
        // the thread index through an empty asm: the store arithmetic stays
        // after the sums instead of holding VGPRs across the K loop (FP64: 128
        // VGPRs + 6 spilled -> 112; bw=256 x 1024 forward GEMM 1.61 -> 1.48 ms,
        // FP32 1.15 -> 0.97 ms)
        int tix = threadIdx.x;
        asm volatile("" : "+v"(tix));
        const int elane = tix & 63, ewave = tix >> 6;
#pragma unroll
        for (int s = 0; s < kKcNS; ++s) {
            const int i = 16 * (ewave + kKcWaves * s) + (elane & 15);
            if (i >= g.Lp) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + 16 * c + Mma<R>::row(elane, r);
                    if (col < ncols) fwd_store((double)acc[s][c][r], col, m, m + p + 2 * i, B, cpf, fmt, cre, cim);
                }
        }
    }
}

// Inverse (seminaive.c:74-98): D'[c][t] = sum_i coef[c][m+p+2i] AI(m,p)[i][t],
// into Cb row (field, +-m, re/im) at p*hE + t.
template <typename R, int NTH>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(4))) void k_gemm_inv_kc(const double* __restrict__ cre,
                                                          const double* __restrict__ cim, const R* __restrict__ AI,
                                                          const GemmMeta* __restrict__ meta, R* __restrict__ Cb, int B,
                                                          int cut, int F, int fmt) {
    __shared__ R Cd[2][kKcCols * kKcRS];
    int bx, m;
    if (!gemm_block(B, bx, m) || m >= cut) return;  // seminaive orders only
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf, cbase = bx * kKcCols;
    if (cbase >= ncols) return;
    const int N = 2 * B, hE = (B + 1) / 2, hO = B / 2, Tp16 = (hE + 15) & ~15, ntt = Tp16 >> 4;
    const size_t B2 = (size_t)B * B;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, kq = lane >> 4, ln = lane & 15;
    constexpr int kKcWaves = NTH / 64, kKcPer = kKcCols * kKc / NTH;  // waves; chunk values staged per thread
    const int fc = tid / (kKc / kKcPer), fk = kKcPer * (tid % (kKc / kKcPer));  // staging: column fc, positions fk ..
    const bool fcol = cbase + fc < ncols;
    int fmm = m;
    const double* fsrc = cre;
    if (fcol) {
        const int col = cbase + fc;
        const int f = col / cpf, rem = col - f * cpf;
        fmm = (rem >> 1) ? -m : m;
        fsrc = ((rem & 1) ? cim : cre) + f * B2;
    }
    for (int p = 0; p < 2; ++p) {
        // a parity without degrees (Lp = 0) still stores its zero cosine
        // coefficients: the DCT reads every position below its count
        const GemmMeta g = meta[m * 2 + p];
        const int Li4 = (g.Lp + 3) & ~3;
        const int cnt = p ? hO : hE;
        const int r0 = row_size(m, m + p);
        const R* ga = AI + g.offI;  // this parity's table (32-bit offsets from here)
        // K start of each of this wave's output tiles (multiples of 4); K ends at Li4
        int kst[kKcNS];
#pragma unroll
        for (int s = 0; s < kKcNS; ++s) {
            const int tt = wave + kKcWaves * s;
            kst[s] = tt < ntt ? max(0, 16 * tt - r0 + 1) & ~3 : Li4;
        }
        const int nq = (Li4 + kKc - 1) / kKc;
        typename Mma<R>::acc acc[kKcNS][4] = {};
        R v[kKcPer], ta[kKcNS][kKcSteps], tb[kKcNS][kKcSteps];
        bool vok[kKcPer];
        auto stage = [&](int q) __attribute__((always_inline)) {
#pragma unroll
            for (int e = 0; e < kKcPer; ++e) {
                const int i = q * kKc + fk + e;
                vok[e] = fcol && i < g.Lp;
                v[e] = *(vok[e] ? fsrc + index_of_coeff(fmm, m + p + 2 * i, B) : cre);  // unconditional; zeroed in put
            }
        };
        auto active = [&](int s, int k) __attribute__((always_inline)) { return k >= kst[s] && k < Li4; };
        auto table = [&](int q, R (&tv)[kKcNS][kKcSteps]) __attribute__((always_inline)) {
#pragma unroll
            for (int s = 0; s < kKcNS; ++s)
#pragma unroll
                for (int u = 0; u < kKcSteps; ++u) {
                    const int k = q * kKc + 4 * u;  // unconditional load, used only where active, as in the forward
                    tv[s][u] = *(active(s, k) ? ga + ((k + kq) * Tp16 + 16 * (wave + kKcWaves * s) + ln) : AI);
                }
        };
        auto put = [&](int b) __attribute__((always_inline)) {
#pragma unroll
            for (int e = 0; e < kKcPer; ++e) Cd[b][fc * kKcRS + fk + e] = vok[e] ? v[e] : (R)0;
        };
        auto chunk = [&](int q, const R (&tv)[kKcNS][kKcSteps]) __attribute__((always_inline)) {
            const R* cd = Cd[q & 1] + ln * kKcRS + kq;
#pragma unroll
            for (int s = 0; s < kKcNS; ++s)
#pragma unroll
                for (int u = 0; u < kKcSteps; ++u)
                    if (active(s, q * kKc + 4 * u)) {
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            acc[s][c] = Mma<R>::mma(cd[16 * c * kKcRS + 4 * u], tv[s][u], acc[s][c]);
                    }
        };
        stage(0);
        table(0, ta);
        put(0);
        __syncthreads();
        for (int q = 0; q < nq; q += 2) {
            // The next chunk's loads, issued unconditionally (past the last
            // chunk they read the arrays' first entries, unused): with the
            // count of loads in flight fixed, each MFMA waits only for the
            // table value it reads, loaded one chunk earlier, and the empty
            // asm keeps the loads ahead of this chunk's MFMAs.
            stage(q + 1);
            table(q + 1, tb);
            asm volatile("" ::: "memory");
            chunk(q, ta);
            if (q + 1 < nq) put(1);
            __syncthreads();
            if (q + 1 >= nq) break;
            stage(q + 2);
            table(q + 2, ta);
            asm volatile("" ::: "memory");
            chunk(q + 1, tb);
            if (q + 2 < nq) put(0);
            __syncthreads();
        }
        int tix = threadIdx.x;  // as in the forward epilogue
        asm volatile("" : "+v"(tix));
        const int elane = tix & 63, ewave = tix >> 6;
#pragma unroll
        for (int s = 0; s < kKcNS; ++s) {
            const int t = 16 * (ewave + kKcWaves * s) + (elane & 15);
            if (t >= cnt) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = cbase + 16 * c + Mma<R>::row(elane, r);
                    if (col >= ncols) continue;
                    const int f = col / cpf, rem = col - f * cpf;
                    const int row = (rem >> 1) ? N - m : m;
                    Cb[(((size_t)f * N + row) * 2 + (rem & 1)) * B + p * hE + t] = acc[s][c][r];
                }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

void gemm_tables_create(GemmTables& g, const DevTables& t, int prec, hipStream_t s) {
    const int B = t.bw, cut = t.cutoff;
    g.bw = B;
    g.prec = prec;
    g.cut = cut;
    g.gen = !gemm_pow2(B);
    g.h_meta.assign(2 * B, GemmMeta{});
    long of = 0, oi = 0;
    for (int m = 0; m < B; ++m)
        for (int p = 0; p < 2; ++p) {
            GemmMeta& e = g.h_meta[m * 2 + p];
            e.Lp = gm_lp(m, p, B);
            e.offF = of;
            e.offI = oi;
            if (m >= cut) {  // naive: bw folded rings in both directions
                e.Kf = B;
                of += (long)pk_b4(B) * ((e.Lp + 15) & ~15);
                oi += (long)((e.Lp + 3) & ~3) * pk_b16(B);
                continue;
            }
            const int llast = m + p + 2 * (e.Lp - 1);
            e.Kf = e.Lp > 0 ? row_size(m, llast) : 0;
            of += (long)((e.Kf + 3) & ~3) * ((e.Lp + 15) & ~15);
            oi += (long)((e.Lp + 3) & ~3) * pk_tp16(B);
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
        k_pack_gemm<double><<<dim3(B, 2), 256, lds, s>>>(t.fwd, t.inv, t.off, g.meta, (double*)g.af, (double*)g.ai, B,
                                                          cut);
    else
        k_pack_gemm<float><<<dim3(B, 2), 256, lds, s>>>(t.fwd, t.inv, t.off, g.meta, (float*)g.af, (float*)g.ai, B,
                                                         cut);
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

// LDS row bytes of a GEMM data column.
static int gemm_row(int B, int prec) {
    const bool gen = !gemm_pow2(B);
    return prec == 0 ? 8 * ggeo<double>(B, gen).rs : 4 * ggeo<float>(B, gen).rs;
}

// Columns per GEMM block: the widest of 64, 32, 16 whose LDS tile lets four
// blocks share a CU (40 KiB each of gfx950's 160 KiB; waves then run two
// column tiles per table load when there are 32 or more), else 16 in up to
// 160 KiB.  Occupancy outweighs table reuse: at bw=256 FP64, 32 columns in
// 66 KiB (two blocks per CU) ran the GEMMs 1.2-1.3x slower than 16 in 33 KiB.
static int gemm_ncb(int B, int prec) {
    const int row = gemm_row(B, prec);
    for (int n = 64; n >= 16; n >>= 1)
        if (n * row <= 40960) return n;
    return 16 * row <= 163840 ? 16 : 0;
}

// Column tiles per wave item: two from 32 columns, when an order has enough
// 16-degree tiles (bw >= 64) that pairing still gives all four waves work.
static int gemm_ct(int ncb, int B) { return ncb >= 32 && B >= 64 ? 2 : 1; }


bool gemm_pow2(int bw) { return bw >= 32 && bw <= 1024 && !(bw & (bw - 1)); }

bool fast_supported(int bw, int cutoff, int prec) {
    if (cutoff < 0 || cutoff > bw || bw > 1024) return false;
    // power-of-two grids with N >= 64: the register FFTs (s2_blocks.h) and the
    // GEMMs in FP64 or FP32; other bw >= 16: the per-field path's mixed-radix
    // FFTs and DCTs with the GEMMs, FP64 only
    if (!gemm_pow2(bw) && (prec != 0 || bw < 16)) return false;
    return gemm_ncb(bw, prec) >= 16;  // GEMM LDS tile of at least 16 columns
}

// Tile slices per order and column block: enough blocks for the chip
// (1024) when the call has few columns, at most 8.
static int gemm_slices(int nblocks) { return std::max(1, std::min(8, 1024 / std::max(1, nblocks))); }

template <typename R, bool GEN, int CT>
static void fwd_launch(const DevConsts& c, const GemmTables& g, const void* Cb, const void* X, double* cre,
                       double* cim, int F, Fmt fmt, int ncb, int mlo, hipStream_t s) {
    using T = typename Cplx<R>::type;
    const int B = c.bw, cols = 4 * F;
    const int nx = (cols + ncb - 1) / ncb, ny = (B + 7) & ~7;
    const dim3 grid(nx, ny, gemm_slices(nx * ny));
    launch_lds(k_gemm_fwd<R, GEN, CT>, grid, kThreads, (size_t)ncb * gemm_row(B, g.prec), s, (const R*)Cb, (const T*)X,
               (const double*)c.w_naive, (const R*)g.af, (const GemmMeta*)g.meta, cre, cim, B, g.cut, mlo, F, ncb,
               (int)fmt);
}

template <typename R, bool GEN, int CT>
static void inv_launch(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, void* Cb,
                       void* X, int F, Fmt fmt, int ncb, int mlo, hipStream_t s) {
    using T = typename Cplx<R>::type;
    const int B = c.bw, cols = 4 * F;
    const int nx = (cols + ncb - 1) / ncb, ny = (B + 7) & ~7;
    const dim3 grid(nx, ny, gemm_slices(nx * ny));
    launch_lds(k_gemm_inv<R, GEN, CT>, grid, kThreads, (size_t)ncb * gemm_row(B, g.prec), s, cre, cim, (const R*)g.ai,
               (const GemmMeta*)g.meta, (R*)Cb, (T*)X, B, g.cut, mlo, F, ncb, (int)fmt);
}

std::vector<std::pair<int, int>> gemm_tile_choices(int bw, int prec) {
    const int row = gemm_row(bw, prec), n0 = gemm_ncb(bw, prec);
    std::vector<std::pair<int, int>> v{{n0, gemm_ct(n0, bw)}};
    for (int n = 16; n <= 64; n <<= 1) {
        if (n * row > 163840) continue;
        for (int ct = 1; ct <= (n >= 32 ? 2 : 1); ++ct)
            if (n != v[0].first || ct != v[0].second) v.push_back({n, ct});
    }
    return v;
}

// The plan's tuned tile shape, or the heuristic's.
static void gemm_tiles(const DevConsts& c, const GemmTables& g, int& ncb, bool& two) {
    ncb = g.t_ncb ? g.t_ncb : gemm_ncb(c.bw, g.prec);
    two = (g.t_ncb ? g.t_ct : gemm_ct(ncb, c.bw)) == 2;
}

// The K-chunked kernels take the seminaive orders (m < cutoff) of calls of
// many fields, bw in [128, 512] (at most sixteen 16-wide tiles per parity:
// one per wave; at bw = 64 they were slower, 31.5 -> 34.0 us forward at 256
// fields); the naive orders of a plan with cutoff < bw follow in a second
// launch of the whole-row kernels.
static bool kc_ok(const DevConsts& c, const GemmTables& g, int F) {
    return g.cut > 0 && c.bw >= 128 && c.bw <= 512 && F >= 64;
}
// threads of a K-chunked block: one wave per 16-wide tile of a parity
static int kc_threads(int B) { return B <= 256 ? 512 : 1024; }

static dim3 kc_grid(int B, int F) { return dim3((4 * F + kKcCols - 1) / kKcCols, (B + 7) & ~7); }

void gemm_fwd_launch(const DevConsts& c, const GemmTables& g, const void* Cb, const void* X, double* cre, double* cim,
                     int F, Fmt fmt, hipStream_t s) {
    if (kc_ok(c, g, F)) {
        if (g.prec != 0)
            (kc_threads(c.bw) == 512 ? k_gemm_fwd_kc<float, 512> : k_gemm_fwd_kc<float, 1024>)<<<kc_grid(c.bw, F), kc_threads(c.bw), 0, s>>>((const float*)Cb, (const float*)g.af,
                                                                     (const GemmMeta*)g.meta, cre, cim, c.bw, g.cut, F,
                                                                     (int)fmt);
        else
            (kc_threads(c.bw) == 512 ? k_gemm_fwd_kc<double, 512> : k_gemm_fwd_kc<double, 1024>)<<<kc_grid(c.bw, F), kc_threads(c.bw), 0, s>>>((const double*)Cb, (const double*)g.af,
                                                                      (const GemmMeta*)g.meta, cre, cim, c.bw, g.cut, F,
                                                                      (int)fmt);
        S2_HIP(hipGetLastError());
        if (g.cut == c.bw) return;
    }
    const int mlo = kc_ok(c, g, F) ? g.cut : 0;
    int ncb;
    bool two;
    gemm_tiles(c, g, ncb, two);
    if (g.prec != 0)
        (two ? fwd_launch<float, false, 2> : fwd_launch<float, false, 1>)(c, g, Cb, X, cre, cim, F, fmt, ncb, mlo, s);
    else if (g.gen)
        (two ? fwd_launch<double, true, 2> : fwd_launch<double, true, 1>)(c, g, Cb, X, cre, cim, F, fmt, ncb, mlo, s);
    else
        (two ? fwd_launch<double, false, 2> : fwd_launch<double, false, 1>)(c, g, Cb, X, cre, cim, F, fmt, ncb, mlo, s);
}

void gemm_inv_launch(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, void* Cb, void* X,
                     int F, Fmt fmt, hipStream_t s) {
    if (kc_ok(c, g, F)) {
        if (g.prec != 0)
            (kc_threads(c.bw) == 512 ? k_gemm_inv_kc<float, 512> : k_gemm_inv_kc<float, 1024>)<<<kc_grid(c.bw, F), kc_threads(c.bw), 0, s>>>(cre, cim, (const float*)g.ai,
                                                                     (const GemmMeta*)g.meta, (float*)Cb, c.bw, g.cut, F,
                                                                     (int)fmt);
        else
            (kc_threads(c.bw) == 512 ? k_gemm_inv_kc<double, 512> : k_gemm_inv_kc<double, 1024>)<<<kc_grid(c.bw, F), kc_threads(c.bw), 0, s>>>(cre, cim, (const double*)g.ai,
                                                                      (const GemmMeta*)g.meta, (double*)Cb, c.bw, g.cut, F,
                                                                      (int)fmt);
        S2_HIP(hipGetLastError());
        if (g.cut == c.bw) return;
    }
    const int mlo = kc_ok(c, g, F) ? g.cut : 0;
    int ncb;
    bool two;
    gemm_tiles(c, g, ncb, two);
    if (g.prec != 0)
        (two ? inv_launch<float, false, 2> : inv_launch<float, false, 1>)(c, g, cre, cim, Cb, X, F, fmt, ncb, mlo, s);
    else if (g.gen)
        (two ? inv_launch<double, true, 2> : inv_launch<double, true, 1>)(c, g, cre, cim, Cb, X, F, fmt, ncb, mlo, s);
    else
        (two ? inv_launch<double, false, 2> : inv_launch<double, false, 1>)(c, g, cre, cim, Cb, X, F, fmt, ncb, mlo, s);
}

template <typename R>
static void fwd_impl(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                     double* cim, int F, Fmt fmt, void* Xv, void* Cbv, hipStream_t s, StageTimer* tm) {
    using T = typename Cplx<R>::type;
    constexpr int V = sizeof(R) == 8 ? 8 : 16;
    const int N = c.N, lgN = lg2(N);
    const int E = V * kThreads, rpb = E / N;
    const size_t lds = sizeof(T) * ((size_t)rpb * (N + 1) + N);
    T* X = (T*)Xv;
    R* Cb = (R*)Cbv;
    const T* tw = (const T*)(sizeof(R) == 8 ? (const void*)c.tw : (const void*)c.tw32);
    const T* ph = (const T*)(sizeof(R) == 8 ? (const void*)c.ph : (const void*)c.ph32);
    const R* ws = (const R*)(sizeof(R) == 8 ? (const void*)c.w_semi : (const void*)c.w_semi32);
    const int few = few_field_threads<V>(N, F);
    const int rpf = few ? V * few / N : 0;  // rings per one-wave block
    const size_t ldf = few ? sizeof(T) * ((size_t)rpf * (N + 1) + N) : 0;
    {
        Span sp(tm, ST_PHI_FWD, s);
        // 512-thread blocks (twice the rings, full 128-byte row segments) when a block still holds whole rows
        if (few == 64) {
            launch_lds(k_fphi_fwd<R, V, 64>, dim3(N / rpf, F), 64, ldf, s, re, im, X, lgN, tw);
        } else if (few == 128) {
            launch_lds(k_fphi_fwd<R, V, 128>, dim3(N / rpf, F), 128, ldf, s, re, im, X, lgN, tw);
        } else if ((long)N * N >= 512L * V) {
            const int rp2 = V * 512 / N;
            const size_t l2 = sizeof(T) * ((size_t)rp2 * (N + 1) + N);
            if (k_fphi_fwd_nt(rp2, sizeof(T)))
                launch_lds(k_fphi_fwd<R, V, 512, true>, dim3(N / rp2, F), 512, l2, s, re, im, X, lgN, tw);
            else
                launch_lds(k_fphi_fwd<R, V, 512>, dim3(N / rp2, F), 512, l2, s, re, im, X, lgN, tw);
        } else {
            launch_lds(k_fphi_fwd<R, V>, dim3(N / rpb, F), kThreads, lds, s, re, im, X, lgN, tw);
        }
    }
    // weighted DCT of the seminaive order rows only
    const int cut = g.cut, nrows = (fmt == Fmt::Complex) ? 2 * cut - 1 : cut;
    if (cut > 0) {
        Span sp(tm, ST_DCT_FWD, s);
        if (few == 64) {
            launch_lds(k_fdct_fwd<R, V, 64>, dim3((nrows + rpf - 1) / rpf, F), 64, ldf, s, (const T*)X, Cb, ws, lgN,
                       cut, nrows, tw, ph);
        } else if (few == 128) {
            launch_lds(k_fdct_fwd<R, V, 128>, dim3((nrows + rpf - 1) / rpf, F), 128, ldf, s, (const T*)X, Cb, ws, lgN,
                       cut, nrows, tw, ph);
        } else if ((long)N * N >= 512L * V) {
            const int rp2 = V * 512 / N;
            launch_lds(k_fdct_fwd<R, V, 512>, dim3((nrows + rp2 - 1) / rp2, F), 512,
                       sizeof(T) * ((size_t)rp2 * (N + 1) + N), s, (const T*)X, Cb, ws, lgN, cut, nrows, tw, ph);
        } else {
            launch_lds(k_fdct_fwd<R, V>, dim3((nrows + rpb - 1) / rpb, F), kThreads, lds, s, (const T*)X, Cb, ws, lgN,
                       cut, nrows, tw, ph);
        }
    }
    {
        Span sp(tm, ST_LEG_FWD, s);
        gemm_fwd_launch(c, g, Cb, X, cre, cim, F, fmt, s);
    }
}

template <typename R>
static void inv_impl(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                     double* im, int F, Fmt fmt, void* Xv, void* Cbv, hipStream_t s, StageTimer* tm) {
    using T = typename Cplx<R>::type;
    constexpr int V = sizeof(R) == 8 ? 8 : 16;
    const int N = c.N, lgN = lg2(N);
    const int E = V * kThreads, rpb = E / N;
    const size_t lds = sizeof(T) * ((size_t)rpb * (N + 1) + N);
    T* X = (T*)Xv;
    R* Cb = (R*)Cbv;
    const T* tw = (const T*)(sizeof(R) == 8 ? (const void*)c.tw : (const void*)c.tw32);
    const T* ph = (const T*)(sizeof(R) == 8 ? (const void*)c.ph : (const void*)c.ph32);
    const R* sv = (const R*)(sizeof(R) == 8 ? (const void*)c.sinv : (const void*)c.sinv32);
    const int few = few_field_threads<V>(N, F);
    const int rpf = few ? V * few / N : 0;  // rings per one-wave block
    const size_t ldf = few ? sizeof(T) * ((size_t)rpf * (N + 1) + N) : 0;
    {
        Span sp(tm, ST_LEG_INV, s);
        gemm_inv_launch(c, g, cre, cim, Cb, X, F, fmt, s);
    }
    // DCT-III of the seminaive order rows (the naive rows are already in X)
    const int cut = g.cut, nrows = (fmt == Fmt::Complex) ? 2 * cut - 1 : cut;
    if (cut > 0) {
        Span sp(tm, ST_DCT_INV, s);
        if (few == 64) {
            launch_lds(k_fdct_inv<R, V, 64>, dim3((nrows + rpf - 1) / rpf, F), 64, ldf, s, (const R*)Cb, X, sv, lgN,
                       cut, nrows, (int)fmt, tw, ph);
        } else if (few == 128) {
            launch_lds(k_fdct_inv<R, V, 128>, dim3((nrows + rpf - 1) / rpf, F), 128, ldf, s, (const R*)Cb, X, sv, lgN,
                       cut, nrows, (int)fmt, tw, ph);
        } else if ((long)N * N >= 512L * V) {
            const int rp2 = V * 512 / N;
            launch_lds(k_fdct_inv<R, V, 512>, dim3((nrows + rp2 - 1) / rp2, F), 512,
                       sizeof(T) * ((size_t)rp2 * (N + 1) + N), s, (const R*)Cb, X, sv, lgN, cut, nrows, (int)fmt, tw,
                       ph);
        } else {
            launch_lds(k_fdct_inv<R, V>, dim3((nrows + rpb - 1) / rpb, F), kThreads, lds, s, (const R*)Cb, X, sv, lgN,
                       cut, nrows, (int)fmt, tw, ph);
        }
    }
    {
        Span sp(tm, ST_PHI_INV, s);
        if (few == 64) {
            launch_lds(k_fphi_inv<R, V, 64>, dim3(N / rpf, F), 64, ldf, s, (const T*)X, re, im, lgN, tw);
        } else if (few == 128) {
            launch_lds(k_fphi_inv<R, V, 128>, dim3(N / rpf, F), 128, ldf, s, (const T*)X, re, im, lgN, tw);
        } else if ((long)N * N >= 512L * V) {
            const int rp2 = V * 512 / N;
            launch_lds(k_fphi_inv<R, V, 512>, dim3(N / rp2, F), 512, sizeof(T) * ((size_t)rp2 * (N + 1) + N), s,
                       (const T*)X, re, im, lgN, tw);
        } else {
            launch_lds(k_fphi_inv<R, V>, dim3(N / rpb, F), kThreads, lds, s, (const T*)X, re, im, lgN, tw);
        }
    }
}

void forward_fast(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                  double* cim, int F, Fmt fmt, void* X, void* Cb, hipStream_t s, StageTimer* tm) {
    if (g.gen)
        forward_fp64_gemm(c, g, re, im, cre, cim, F, fmt, (double2*)X, (double2*)Cb, s, tm);
    else if (g.prec == 0)
        fwd_impl<double>(c, g, re, im, cre, cim, F, fmt, X, Cb, s, tm);
    else
        fwd_impl<float>(c, g, re, im, cre, cim, F, fmt, X, Cb, s, tm);
}

void inverse_fast(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                  double* im, int F, Fmt fmt, void* X, void* Cb, hipStream_t s, StageTimer* tm) {
    if (g.gen)
        inverse_fp64_gemm(c, g, cre, cim, re, im, F, fmt, (double2*)X, (double2*)Cb, s, tm);
    else if (g.prec == 0)
        inv_impl<double>(c, g, cre, cim, re, im, F, fmt, X, Cb, s, tm);
    else
        inv_impl<float>(c, g, cre, cim, re, im, F, fmt, X, Cb, s, tm);
}

}  // namespace s2
