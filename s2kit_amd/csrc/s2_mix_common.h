// Pieces shared by the two naive-Legendre GEMM forms on FP32 ring FFTs:
// FP32 MFMA (s2_mixed.hip) and split-BF16 MFMA (s2_split.hip).  Both read and
// write the same FP32 intermediate X [F][N rows][N rings] (complex), run one
// launch per direction over all orders with a workgroup per order pair
// (m, bw-1-m) and column block, and store their results the same way: the
// inverse tables carry 1/sqrt(2pi) and the inverse operand the row sign.
#pragma once

#include "s2_blocks.h"

namespace s2 {

// Waves per GEMM block, 16 data columns each.
#ifndef S2_MIX_WAVES
#define S2_MIX_WAVES 4
#endif
#ifndef S2_MIX_FFT_THREADS
#define S2_MIX_FFT_THREADS 512
#endif
// Ring-FFT blocks of the FP32 forms: 16 values per thread, so a block holds
// 16 rings at bw=256 and stores full 128-byte row segments.
constexpr int kMixFftThreads = S2_MIX_FFT_THREADS;
constexpr int kMixThreads = 64 * S2_MIX_WAVES;
constexpr int kMixCols = 16 * S2_MIX_WAVES;

// 16-degree tiles per parity of order m (the even parity has the most degrees).
__host__ __device__ inline int mix_n0(int m, int B) { return (gm_lp(m, 0, B) + 15) / 16; }

// Workgroup id -> an order pair (m, bw-1-m) and a column block cb of nb per
// order.  A workgroup runs both orders of its pair, so every workgroup has
// about the same work (bw-m + m+1 degrees): blocks of very different lengths
// would leave the dispatcher's round-robin placement unbalanced.  Ids go
// round-robin over the 8 XCDs (id mod 8); XCD k takes the pairs m = 8q+k, all
// column blocks of one pair back to back, so the pair's table tiles stay in
// that XCD's L2.  bw % 16 == 0.
__device__ __forceinline__ void mix_pair(int nb, int& m, int& cb) {  // grid: mix_npairs8(bw) * nb
    const int id = blockIdx.x, k = id & 7, local = id >> 3;
    const int q = local / nb;
    cb = local - q * nb;
    m = 8 * q + k;  // < bw/2; its partner is bw-1-m
}

// Folded rings the mixed form's GEMMs run over: bw rounded up to whole
// 64-ring groups (32 at bw = 32); the tables are zero past bw, and the
// inverse stores no ring past bw.  Equal to bw for powers of two.
__host__ __device__ inline int mix_bp(int B) { return B <= 32 ? 32 : (B + 63) & ~63; }
// Order pairs (m, bw-1-m) of a GEMM launch, ceil(bw/2) (the middle order of
// an odd bw pairs with itself), rounded up to the 8 XCDs of mix_pair.
__host__ __device__ inline int mix_npairs(int B) { return (B + 1) / 2; }
__host__ __device__ inline int mix_npairs8(int B) { return (mix_npairs(B) + 7) & ~7; }

// Folded rings (JW) and 16-ring tiles (JT) of an inverse block: 64 GJ rings,
// or the 32 rings of bw = 32 for GJ = 0.
template <int GJ>
__host__ __device__ constexpr int mix_jw() {
    return GJ ? 64 * GJ : 32;
}
template <int GJ>
__host__ __device__ constexpr int mix_jt() {
    return GJ ? 4 * GJ : 2;
}

// Data columns of order m: (field, +m or -m row, re or im), cpf per field.
__host__ __device__ inline int mix_cpf(int m, int fmt) { return (m > 0 && fmt == 0) ? 4 : 2; }

typedef double dpair8 __attribute__((ext_vector_type(2))) __attribute__((aligned(8)));

// Forward epilogue from registers: lane (x16, kq) holds, for column 4kq + r,
// degree offsets 32t + 2x16 (acc[t], even parity) and 32t + 2x16 + 1
// (acc[N0+t], odd): one 16-byte store per (t, r), 16 lanes covering 256
// contiguous bytes of the column's run of L = bw-m coefficients
// (IndexOfHarmonicCoeff keeps an order's degrees contiguous).
template <int N0>
__device__ __forceinline__ void mix_fwd_store(const fvec4 (&acc)[2 * N0], double* __restrict__ cre,
                                              double* __restrict__ cim, int m, int cb, int B, int F, int fmt) {
    // the thread index through an empty asm: the column and address arithmetic
    // below then stays after the sums instead of being hoisted ahead of the K
    // loop, where it held VGPRs for the whole loop (k_mix_fwd<16>: 256 VGPRs
    // + 48 spilled -> 246, bw=512 x 128 forward GEMM 0.763 -> 0.734 ms;
    // k_mix_fwd<8>: 168 + 6 -> 155, bw=256 unchanged; results identical)
    int tix = threadIdx.x;
    asm volatile("" : "+v"(tix));
    const int lane = tix & 63, wave = tix >> 6;
    const int x16 = lane & 15, kq = lane >> 4;
    const int cpf = mix_cpf(m, fmt), ncols = F * cpf;
    const size_t B2 = (size_t)B * B;
    const bool odd = m & 1;
    const int L = B - m;
    const int bp = index_of_coeff(m, m, B), bn = index_of_coeff(-m, m, B);
    auto flip = [](float v, unsigned s) __attribute__((always_inline)) {
        return (double)__uint_as_float(__float_as_uint(v) ^ s);
    };
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int c = cb * kMixCols + wave * 16 + 4 * kq + r;
        if (c >= ncols) continue;
        const int f = c / cpf, rem = c - f * cpf, comp = rem & 1;
        double* base = (comp ? cim : cre) + f * B2;
        const bool neg = (rem >> 1) != 0;
        double* o1 = base + (neg ? bn : bp);
        const unsigned s1 = (neg && odd) ? 0x80000000u : 0u;  // rows -m carry (-1)^m
        const bool mir = !neg && fmt == 1 && m > 0;  // REAL data: the -m run is the conjugate mirror
        double* o2 = base + bn;
        const unsigned s2 = ((comp != 0) != odd) ? 0x80000000u : 0u;
#pragma unroll
        for (int t = 0; t < N0; ++t) {
            const int d = 32 * t + 2 * x16;
            const float v0 = acc[t][r], v1 = acc[N0 + t][r];
            if (d + 1 < L) {
                *reinterpret_cast<dpair8*>(o1 + d) = dpair8{flip(v0, s1), flip(v1, s1)};
                if (mir) *reinterpret_cast<dpair8*>(o2 + d) = dpair8{flip(v0, s2), flip(v1, s2)};
            } else if (d < L) {
                o1[d] = flip(v0, s1);
                if (mir) o2[d] = flip(v0, s2);
            }
        }
    }
}

// Inverse epilogue: acc[p][t] holds the parity-p sums E (p = 0) and O
// (p = 1) at rings jbase + 16t + x16 of columns 4(lane/16) + r (the table
// carries 1/sqrt(2pi), the operand the row sign); rows r, r+1 are the re, im
// parts of one (field, sign) row.  Ring j <- E+O, ring N-1-j <- E-O (j < bw:
// padded rings are not stored), and for REAL data the conjugate mirror row
// N-m.
template <int JT>
__device__ __forceinline__ void mix_inv_store(const fvec4 (&acc)[2][JT], float2* __restrict__ X, int m, int cb,
                                              int jbase, int B, int F, int fmt) {
    // the thread index through an empty asm, as in the forward epilogue
    int tix = threadIdx.x;
    asm volatile("" : "+v"(tix));
    const int lane = tix & 63, wave = tix >> 6;
    const int x16 = lane & 15, kq = lane >> 4;
    const int N = 2 * B;
    const int cpf = mix_cpf(m, fmt), ncols = F * cpf;
    const bool mirror = fmt == 1 && m > 0;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int c = cb * kMixCols + wave * 16 + 4 * kq + 2 * rr;
        if (c >= ncols) continue;
        const int f = c / cpf, rem = c - f * cpf;
        float2* xr = X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N;
        float2* xm = X + ((size_t)f * N + (N - m)) * N;
        float2* plo = xr + jbase + x16;
        float2* phi = xr + (N - 1 - jbase - x16);
        const int r = 2 * rr;
#pragma unroll
        for (int t = 0; t < JT; ++t) {
            if (jbase + 16 * t + x16 >= B) continue;
            const fvec4 e = acc[0][t], o = acc[1][t];
            plo[16 * t] = make_float2(e[r] + o[r], e[r + 1] + o[r + 1]);
            phi[-16 * t] = make_float2(e[r] - o[r], e[r + 1] - o[r + 1]);
        }
        if (mirror) {
            float2* qlo = xm + jbase + x16;
            float2* qhi = xm + (N - 1 - jbase - x16);
#pragma unroll
            for (int t = 0; t < JT; ++t) {
                if (jbase + 16 * t + x16 >= B) continue;
                const fvec4 e = acc[0][t], o = acc[1][t];
                qlo[16 * t] = make_float2(e[r] + o[r], -(e[r + 1] + o[r + 1]));
                qhi[-16 * t] = make_float2(e[r] - o[r], -(e[r + 1] - o[r + 1]));
            }
        }
    }
}

}  // namespace s2
