// Split-BF16 form: FP32 ring FFTs (as in s2_mixed.hip) and the naive Legendre
// transform (naive.c:34-95) as BF16 MFMA GEMMs with compensated products.
//
// Every FP32 operand x is split into three BF16 pieces, x = h + m + l with
// h = bf16(x), m = bf16(x - h), l = bf16(x - h - m) (24 significant bits in
// all), and each product a*b is formed from the six piece products whose
// weight is at least 2^-16 of h_a*h_b:
//     l_a h_b + h_a l_b + m_a m_b + m_a h_b + h_a m_b + h_a h_b
// (smallest first), each an exact BF16 x BF16 product accumulated in FP32 by
// v_mfma_f32_16x16x32_bf16.  The dropped terms are below 2^-24 |a b|, so the
// result matches an FP32 GEMM; the tables' pieces come from the FP64
// recurrence (pml.c:41-79), split once when the plan is made.
//
//   forward   k_fphi_fwd<float> -> k_split_fwd
//   inverse   k_split_inv -> k_fphi_inv<float>
//
// The GEMMs, folding and column layout are those of the FP32-MFMA form
// (s2_mixed.hip): per order m, data columns (field, +m or -m row, re or im),
//   forward  C[col][l]  = sum_{j<bw} fold_p(X[col])[j] * w_j P_l^m(theta_j)
//   inverse  E/O[col][j] = sum_i coef[col][m+p+2i] * P_{m+p+2i}^m(theta_j).
// K runs 32 at a time (one MFMA): lane group kg = lane/16 holds k = 8kg..8kg+7.
//
// Table stages stream global -> LDS with global_load_lds_dwordx4 (no staging
// registers), two 24 KiB buffers per block: stage s+1 is issued before stage
// s's MFMAs, and each wave waits for its own transfers (vmcnt 0) before the
// barrier that hands the buffer to the block.  The data operand is loaded
// straight into registers one chunk (32 rings or degrees) ahead.
#include <algorithm>
#include <cmath>

#include "s2_mix_common.h"
#include "s2_recurrence.h"

namespace s2 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

#define S2_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

#ifndef S2_SPLIT_GJMAX
#define S2_SPLIT_GJMAX 4
#endif
// Column blocks (kMixCols columns each) that one workgroup streams per order.
#ifndef S2_SPLIT_BLOCKS
#define S2_SPLIT_BLOCKS 1
#endif
constexpr int kSplitBlocks = S2_SPLIT_BLOCKS;
// Table tiles whose six products are interleaved (a group's pieces are read
// while the previous group's MFMAs run).
#ifndef S2_SPLIT_G
#define S2_SPLIT_G 1
#endif
constexpr int kSplitG = S2_SPLIT_G;
// Waves per SIMD the register allocation of each GEMM kernel aims for.
#ifndef S2_SPLIT_WPE_FWD
#define S2_SPLIT_WPE_FWD 2
#endif
#ifndef S2_SPLIT_WPE_INV
#define S2_SPLIT_WPE_INV 2
#endif

constexpr int kSplitUnit = 1024;            // bytes of one piece of one 16-wide tile: 64 lanes x 8 bf16
constexpr int kSplitTile = 3 * kSplitUnit;  // the three pieces of a tile
constexpr int kSplitStageTiles = 8;
constexpr int kSplitStage = kSplitStageTiles * kSplitTile;  // 24 KiB

// Tables (bf16 elements; a "tile" is 3 pieces x 64 lanes x 8 values, lane
// kg*16 + x holding k = 8kg..8kg+7 of column x):
//   SF(m) [ch < bw/32][t < 2 n0][piece][lane][8]: ring j = 32ch + 8kg + e,
//         degree l = m + 2(16t + x) (t < n0) or m + 1 + 2(16(t-n0) + x),
//         value w_j P_l^m(theta_j) (w_j: even DH weight times sqrt(2pi)/N);
//   SI(m) [c < ki/32][p][g < bw/16][piece][lane][8]: degree l = m+p+2i,
//         i = 32c + 8kg + e, ring j = 16g + x, value P_l^m(theta_j)/sqrt(2pi).
// Entries past the last degree are zero.
__host__ __device__ inline int split_ki(int m, int B) { return 32 * ((gm_lp(m, 0, B) + 31) / 32); }

__device__ __forceinline__ void split3(double v, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)(float)v;
    const double r = v - (double)(float)h;
    m = (__bf16)(float)r;
    l = (__bf16)(float)(r - (double)(float)m);
}

__global__ void k_split_pack(const double* __restrict__ w, const GemmMeta* __restrict__ meta, __bf16* __restrict__ SF,
                             __bf16* __restrict__ SI, int B) {
    const int m = blockIdx.x;
    const int N = 2 * B, G = B / 16;
    const GemmMeta g0 = meta[m * 2];
    const int nt = 2 * g0.n0;
    const double cm = (m == 0) ? 0.0 : pmm_coeff(m);
    for (int j = threadIdx.x; j < B; j += blockDim.x) {
        const double th = (2.0 * j + 1.0) * M_PI / (2.0 * N);
        const double x = cos(th);
        const double wj = w[j];
        double cur = (m == 0) ? 0.70710678118654752440 : cm * pow(sin(th), (double)m);
        double prev = 0.0;
        for (int l = m; l < B; ++l) {
            const int p = (l - m) & 1, i = (l - m) >> 1;
            __bf16 h, md, lo;
            {
                const int t = p * g0.n0 + (i >> 4), lane = ((j >> 3) & 3) * 16 + (i & 15);
                const size_t at = (size_t)g0.offF + (((size_t)(j >> 5) * nt + t) * 3 * 64 + lane) * 8 + (j & 7);
                split3(cur * wj, h, md, lo);
                SF[at] = h;
                SF[at + 512] = md;
                SF[at + 1024] = lo;
            }
            {
                const int lane = ((i >> 3) & 3) * 16 + (j & 15);
                const size_t at =
                    (size_t)g0.offI + (((((size_t)(i >> 5) * 2 + p) * G + (j >> 4)) * 3) * 64 + lane) * 8 + (i & 7);
                split3(cur * 0.39894228040143267794, h, md, lo);  // times 1/sqrt(2pi)
                SI[at] = h;
                SI[at + 512] = md;
                SI[at + 1024] = lo;
            }
            if (l + 1 < B) {
                const double nx = l2_an(m, l) * (cur * x) + l2_cn(m, l) * prev;
                prev = cur;
                cur = nx;
            }
        }
    }
}

// Three BF16 pieces of eight FP32 values.
__device__ __forceinline__ void split8(const float (&x)[8], bf16x8 (&a)[3]) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)x[e];
        const float r = x[e] - (float)h;
        const __bf16 md = (__bf16)r;
        a[0][e] = h;
        a[1][e] = md;
        a[2][e] = (__bf16)(r - (float)md);
    }
}

// Product k of the six, smallest first: l*h, h*l, m*m, m*h, h*m, h*h.
__device__ __forceinline__ fvec4 mfma_piece(int k, const bf16x8 (&a)[3], const bf16x8 (&b)[3], fvec4 c) {
    constexpr int ia[6] = {2, 0, 1, 1, 0, 0}, ib[6] = {0, 2, 1, 0, 1, 0};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ia[k]], b[ib[k]], c, 0, 0, 0);
}

__device__ __forceinline__ fvec4 mfma6(const bf16x8 (&a)[3], const bf16x8 (&b)[3], fvec4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
    return c;
}

// Pieces of tile tt of a stage buffer, as this lane's B operand.
__device__ __forceinline__ void stage_tile(const unsigned char* buf, int tt, int lane, bf16x8 (&b)[3]) {
#pragma unroll
    for (int pc = 0; pc < 3; ++pc)
        b[pc] = *reinterpret_cast<const bf16x8*>(buf + (tt * 3 + pc) * kSplitUnit + lane * 16);
}

// Copy n tiles starting at src (global, contiguous) into an LDS stage buffer;
// the block's waves take the 1 KiB units round robin.
__device__ __forceinline__ void stage_copy(const unsigned char* src, unsigned char* dst, int n, int wave, int lane) {
    for (int u = wave; u < 3 * n; u += kMixThreads / 64)
        __builtin_amdgcn_global_load_lds((const void*)(src + (size_t)u * kSplitUnit + lane * 16),
                                         S2_LDS_PTR(dst + u * kSplitUnit), 16, 0, 0);
}

// This wave's LDS transfers (and every other load) have landed; then the block syncs.
__device__ __forceinline__ void stage_sync() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// Order m over column blocks cb0, cb0+1, ... (nblk of them): one stream of
// chunks (32 folded rings each; the order's NT tiles go through the stage
// buffers SPC at a time), so the next block's first data and table stage are
// in flight while the current block finishes.
template <int N0>
__device__ __forceinline__ void split_fwd_body(const float2* __restrict__ X, const __bf16* __restrict__ SF,
                                               const GemmMeta& g0, double* __restrict__ cre, double* __restrict__ cim,
                                               int m, int cb0, int nblk, int B, int F, int fmt, unsigned char* sm) {
    constexpr int NT = 2 * N0;
    constexpr int SPC = (NT + kSplitStageTiles - 1) / kSplitStageTiles;
    const int N = 2 * B;
    const int cpf = mix_cpf(m, fmt), ncols = F * cpf;
    const int nact = min(nblk, (ncols - cb0 * kMixCols + kMixCols - 1) / kMixCols);
    if (nact <= 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, kg = lane >> 4;
    const bool imcol = x16 & 1;  // columns alternate re, im

    // operand A: this lane's column of block cb, rings 32ch + 8kg + e and their mirrors
    auto column = [&](int cb) __attribute__((always_inline)) {
        const int col = cb * kMixCols + wave * 16 + x16;
        const int cc = col < ncols ? col : 0;  // past the end: read column 0, store nothing
        const int f = cc / cpf, rem = cc - f * cpf;
        return X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N;
    };
    const unsigned char* tsrc = reinterpret_cast<const unsigned char*>(SF + g0.offF);
    const int nch = B / 32, spb = nch * SPC, nst = nact * spb;

    auto copy_stage = [&](int s) __attribute__((always_inline)) {
        const int sb = s % spb, ch = sb / SPC, q = sb - ch * SPC;
        const int t0 = q * kSplitStageTiles;
        stage_copy(tsrc + ((size_t)ch * NT + t0) * kSplitTile, sm + (s & 1) * kSplitStage,
                   min(kSplitStageTiles, NT - t0), wave, lane);
    };
    fvec4 xa[4], xb[4];  // rings j0..j0+7 and N-8-j0..N-1-j0, complex pairs
    auto load = [&](const float2* xr, int ch) __attribute__((always_inline)) {
        const int j0 = 32 * ch + 8 * kg;
        const fvec4* pa = reinterpret_cast<const fvec4*>(xr + j0);
        const fvec4* pb = reinterpret_cast<const fvec4*>(xr + N - 8 - j0);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            xa[h] = pa[h];
            xb[h] = pb[h];
        }
    };
    bf16x8 ap[2][3];  // folded operands x_j + x_{N-1-j} (even degrees), x_j - x_{N-1-j} (odd)
    auto prep = [&]() __attribute__((always_inline)) {
        float a0[8], a1[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const fvec4 va = xa[e >> 1];
            const int eb = 7 - e;  // ring N-1-j0-e is entry 7-e of the mirror run
            const fvec4 vb = xb[eb >> 1];
            const float u = (e & 1) ? (imcol ? va.w : va.z) : (imcol ? va.y : va.x);
            const float v = (eb & 1) ? (imcol ? vb.w : vb.z) : (imcol ? vb.y : vb.x);
            a0[e] = u + v;
            a1[e] = u - v;
        }
        split8(a0, ap[0]);
        split8(a1, ap[1]);
    };

    const float2* xr = column(cb0);
    copy_stage(0);
    load(xr, 0);
    stage_sync();
    for (int it = 0; it < nact; ++it) {
        const float2* xn = column(cb0 + min(it + 1, nact - 1));
        fvec4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = fvec4{0.f, 0.f, 0.f, 0.f};
        for (int ch = 0; ch < nch; ++ch) {
            prep();
            if (ch + 1 < nch)
                load(xr, ch + 1);
            else if (it + 1 < nact)
                load(xn, 0);
#pragma unroll
            for (int q = 0; q < SPC; ++q) {
                const int s = it * spb + ch * SPC + q;
                if (s + 1 < nst) copy_stage(s + 1);
                const unsigned char* buf = sm + (s & 1) * kSplitStage;
                // tiles in groups of kSplitG (stage tile counts are even): the six
                // products run across the group, so consecutive MFMAs are independent;
                // the next group's table pieces are read while this group's MFMAs run
                constexpr int GN = kSplitStageTiles / kSplitG;
                bf16x8 bq[2][kSplitG][3];
#pragma unroll
                for (int g = 0; g < kSplitG; ++g)
                    if (q * kSplitStageTiles + g < NT) stage_tile(buf, g, lane, bq[0][g]);
#pragma unroll
                for (int gi = 0; gi < GN; ++gi) {
                    const int t0 = q * kSplitStageTiles + gi * kSplitG;
                    if (t0 < NT) {
#pragma unroll
                        for (int g = 0; g < kSplitG; ++g) {
                            const int tn = (gi + 1) * kSplitG + g;
                            if (gi + 1 < GN && q * kSplitStageTiles + tn < NT)
                                stage_tile(buf, tn, lane, bq[(gi + 1) & 1][g]);
                        }
                        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of these MFMAs
#pragma unroll
                        for (int k = 0; k < 6; ++k)
#pragma unroll
                            for (int g = 0; g < kSplitG; ++g) {
                                const int t = t0 + g;
                                if (t < NT) acc[t] = mfma_piece(k, ap[t < N0 ? 0 : 1], bq[gi & 1][g], acc[t]);
                            }
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                stage_sync();
            }
        }
        mix_fwd_store<N0>(acc, cre, cim, m, cb0 + it, B, F, fmt);
        xr = xn;
    }
}

template <int MAXN0>
__global__ __launch_bounds__(kMixThreads) __attribute__((amdgpu_waves_per_eu(S2_SPLIT_WPE_FWD))) void k_split_fwd(const float2* __restrict__ X, const __bf16* __restrict__ SF,
                                                           const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                           double* __restrict__ cim, int lgB, int F, int fmt, int nbw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int B = 1 << lgB;
    int mp, cw;
    mix_pair(nbw, mp, cw);
    const int cb0 = cw * kSplitBlocks;
    for (int h = 0; h < 2; ++h) {
        const int m = h ? B - 1 - mp : mp;
        const GemmMeta g0 = meta[m * 2];
#define S2_SPLIT_FWD(n)                                                                                        \
    case n:                                                                                                    \
        if (n <= MAXN0)                                                                                        \
            split_fwd_body<(n <= MAXN0 ? n : 1)>(X, SF, g0, cre, cim, m, cb0, kSplitBlocks, B, F, fmt, smem); \
        break;
        switch (g0.n0) {
            S2_SPLIT_FWD(1)
            S2_SPLIT_FWD(2)
            S2_SPLIT_FWD(3)
            S2_SPLIT_FWD(4)
            S2_SPLIT_FWD(5)
            S2_SPLIT_FWD(6)
            S2_SPLIT_FWD(7)
            S2_SPLIT_FWD(8)
            S2_SPLIT_FWD(9)
            S2_SPLIT_FWD(10)
            S2_SPLIT_FWD(11)
            S2_SPLIT_FWD(12)
            S2_SPLIT_FWD(13)
            S2_SPLIT_FWD(14)
            S2_SPLIT_FWD(15)
            S2_SPLIT_FWD(16)
            default: break;
        }
#undef S2_SPLIT_FWD
        __syncthreads();  // the partner order's first stage reuses the buffers
    }
}

// ---------------------------------------------------------------------------
// inverse
// ---------------------------------------------------------------------------
// Order m, ring range jr (JT = 4*GJ ring tiles, 64*GJ of the bw folded
// rings; 2 tiles, the 32 rings of bw = 32, for GJ = 0), column blocks cb0,
// cb0+1, ... (nblk): one stream of chunks of 32
// degree indices of both parities; a stage is one parity's tiles (8 at most).
template <int GJ>
__device__ __forceinline__ void split_inv_body(const double* __restrict__ cre, const double* __restrict__ cim,
                                               const __bf16* __restrict__ SI, const GemmMeta* __restrict__ meta,
                                               float2* __restrict__ X, int B, int F, int fmt, int m, int cb0,
                                               int nblk, int jr, unsigned char* sm) {
    constexpr int JT = mix_jt<GJ>();
    constexpr int TPS = JT < kSplitStageTiles ? JT : kSplitStageTiles;
    constexpr int SPP = JT / TPS;  // stages per parity
    constexpr int SPC = 2 * SPP;   // stages per chunk
    constexpr int SB = TPS * kSplitTile;
    const int jbase = jr * 16 * JT;
    const int cpf = mix_cpf(m, fmt), ncols = F * cpf;
    const int nact = min(nblk, (ncols - cb0 * kMixCols + kMixCols - 1) / kMixCols);
    if (nact <= 0) return;
    const GemmMeta g0 = meta[m * 2];
    const int L = B - m, G = B / 16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, kg = lane >> 4;
    const size_t B2 = (size_t)B * B;

    // operand A: coefficients of this lane's column of block cb, degree
    // offsets d = 64c + 16kg + e, e < 16 (parity e & 1, index 32c + 8kg + (e >> 1))
    auto column = [&](int cb) __attribute__((always_inline)) {
        const int col = cb * kMixCols + wave * 16 + x16;
        const int cc = col < ncols ? col : 0;
        const int f = cc / cpf, rem = cc - f * cpf;
        return ((rem & 1) ? cim : cre) + f * B2 + index_of_coeff((rem >> 1) ? -m : m, m, B);
    };
    // rows -m of odd m carry the sign (-1)^m: applied to the operand
    const unsigned nsign = ((((wave * 16 + x16) % cpf) >> 1) && (m & 1)) ? 0x80000000u : 0u;
    const unsigned char* tsrc = reinterpret_cast<const unsigned char*>(SI + g0.offI);
    const int nch = g0.ki / 32, spb = nch * SPC, nst = nact * spb;
    const int gb = jbase / 16;

    auto copy_stage = [&](int s) __attribute__((always_inline)) {
        const int sb = s % spb, c = sb / SPC, r = sb - c * SPC;
        const int p = r / SPP, hh = r - p * SPP;
        stage_copy(tsrc + (((size_t)c * 2 + p) * G + gb + hh * TPS) * kSplitTile, sm + (s & 1) * SB, TPS, wave,
                   lane);
    };
    dpair8 cv[8];  // degree offsets d0 + 2h, d0 + 2h + 1
    auto load = [&](const double* cp, int c) __attribute__((always_inline)) {
        const int d0 = 64 * c + 16 * kg;
        if (d0 + 16 <= L) {
#pragma unroll
            for (int h = 0; h < 8; ++h) cv[h] = *reinterpret_cast<const dpair8*>(cp + d0 + 2 * h);
        } else {  // the run ends inside this lane's 16 degrees
#pragma unroll
            for (int h = 0; h < 8; ++h)
                cv[h] = dpair8{d0 + 2 * h < L ? cp[d0 + 2 * h] : 0.0, d0 + 2 * h + 1 < L ? cp[d0 + 2 * h + 1] : 0.0};
        }
    };
    bf16x8 ap[2][3];
    auto prep = [&]() __attribute__((always_inline)) {
        float a[2][8];
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            a[0][h] = __uint_as_float(__float_as_uint((float)cv[h].x) ^ nsign);
            a[1][h] = __uint_as_float(__float_as_uint((float)cv[h].y) ^ nsign);
        }
        split8(a[0], ap[0]);
        split8(a[1], ap[1]);
    };

    const double* cp = column(cb0);
    copy_stage(0);
    load(cp, 0);
    stage_sync();
    for (int it = 0; it < nact; ++it) {
        const double* cn = column(cb0 + min(it + 1, nact - 1));
        fvec4 acc[2][JT];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int t = 0; t < JT; ++t) acc[p][t] = fvec4{0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < nch; ++c) {
            prep();
            if (c + 1 < nch)
                load(cp, c + 1);
            else if (it + 1 < nact)
                load(cn, 0);
#pragma unroll
            for (int r = 0; r < SPC; ++r) {
                const int s = it * spb + c * SPC + r;
                const int p = r / SPP, hh = r - p * SPP;
                if (s + 1 < nst) copy_stage(s + 1);
                const unsigned char* buf = sm + (s & 1) * SB;
                constexpr int GN = TPS / kSplitG;
                bf16x8 bq[2][kSplitG][3];
#pragma unroll
                for (int g = 0; g < kSplitG; ++g) stage_tile(buf, g, lane, bq[0][g]);
#pragma unroll
                for (int gi = 0; gi < GN; ++gi) {
#pragma unroll
                    for (int g = 0; g < kSplitG; ++g)
                        if (gi + 1 < GN) stage_tile(buf, (gi + 1) * kSplitG + g, lane, bq[(gi + 1) & 1][g]);
                    __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of these MFMAs
#pragma unroll
                    for (int k = 0; k < 6; ++k)
#pragma unroll
                        for (int g = 0; g < kSplitG; ++g) {
                            const int t = hh * TPS + gi * kSplitG + g;
                            acc[p][t] = mfma_piece(k, ap[p], bq[gi & 1][g], acc[p][t]);
                        }
                    __builtin_amdgcn_sched_barrier(0);
                }
                stage_sync();
            }
        }
        mix_inv_store<JT>(acc, X, m, cb0 + it, jbase, B, F, fmt);
        cp = cn;
    }
}

template <int GJ>
__global__ __launch_bounds__(kMixThreads) __attribute__((amdgpu_waves_per_eu(S2_SPLIT_WPE_INV))) void k_split_inv(const double* __restrict__ cre,
                                                           const double* __restrict__ cim,
                                                           const __bf16* __restrict__ SI,
                                                           const GemmMeta* __restrict__ meta, float2* __restrict__ X,
                                                           int lgB, int F, int fmt, int nbw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int B = 1 << lgB;
    const int nz = B / mix_jw<GJ>();
    int mp, cz;
    mix_pair(nbw * nz, mp, cz);
    const int cw = cz / nz, jr = cz - cw * nz;
    split_inv_body<GJ>(cre, cim, SI, meta, X, B, F, fmt, mp, cw * kSplitBlocks, kSplitBlocks, jr, smem);
    __syncthreads();  // the partner order's first stage reuses the buffers
    split_inv_body<GJ>(cre, cim, SI, meta, X, B, F, fmt, B - 1 - mp, cw * kSplitBlocks, kSplitBlocks, jr, smem);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool split_supported(int bw) {
    const int N = 2 * bw;
    return bw >= 32 && bw <= 512 && !(N & (N - 1));
}

void split_tables_create(GemmTables& g, const DevConsts& c, hipStream_t s) {
    const int B = c.bw;
    g.bw = B;
    g.prec = 3;
    g.h_meta.assign(2 * B, GemmMeta{});
    long of = 0, oi = 0;  // bf16 elements
    for (int m = 0; m < B; ++m) {
        const int n0 = mix_n0(m, B), ki = split_ki(m, B);
        for (int p = 0; p < 2; ++p) {
            GemmMeta& e = g.h_meta[m * 2 + p];
            e.Lp = gm_lp(m, p, B);
            e.Kf = B;
            e.n0 = n0;
            e.ki = ki;
            e.offF = of;
            e.offI = oi;
        }
        of += (long)(B / 32) * 2 * n0 * (kSplitTile / 2);
        oi += (long)(ki / 32) * 2 * (B / 16) * (kSplitTile / 2);
    }
    g.nf = of;
    g.ni = oi;
    S2_HIP(hipMalloc(&g.meta, sizeof(GemmMeta) * 2 * B));
    S2_HIP(hipMalloc(&g.af, sizeof(__bf16) * of));
    S2_HIP(hipMalloc(&g.ai, sizeof(__bf16) * oi));
    S2_HIP(hipMemsetAsync(g.af, 0, sizeof(__bf16) * of, s));
    S2_HIP(hipMemsetAsync(g.ai, 0, sizeof(__bf16) * oi, s));
    S2_HIP(hipMemcpyAsync(g.meta, g.h_meta.data(), sizeof(GemmMeta) * 2 * B, hipMemcpyHostToDevice, s));
    k_split_pack<<<B, 256, 0, s>>>(c.w_naive, g.meta, (__bf16*)g.af, (__bf16*)g.ai, B);
    S2_HIP(hipGetLastError());
    S2_HIP(hipStreamSynchronize(s));
}

void split_fwd_gemm(const DevConsts& c, const GemmTables& g, const float2* X, double* cre, double* cim, int F,
                    Fmt fmt, hipStream_t s) {
    const int B = c.bw;
    const int nb = (4 * F + kMixCols - 1) / kMixCols;
    const int nbw = (nb + kSplitBlocks - 1) / kSplitBlocks;
    const dim3 grid(B / 2 * nbw);  // order pairs x groups of column blocks
    const int n0max = mix_n0(0, B);
    const size_t lds = 2 * kSplitStage;
    const __bf16* sf = (const __bf16*)g.af;
    const GemmMeta* meta = g.meta;
    if (n0max <= 2)
        launch_lds(k_split_fwd<2>, grid, kMixThreads, lds, s, X, sf, meta, cre, cim, lg2(B), F, (int)fmt, nbw);
    else if (n0max <= 4)
        launch_lds(k_split_fwd<4>, grid, kMixThreads, lds, s, X, sf, meta, cre, cim, lg2(B), F, (int)fmt, nbw);
    else if (n0max <= 8)
        launch_lds(k_split_fwd<8>, grid, kMixThreads, lds, s, X, sf, meta, cre, cim, lg2(B), F, (int)fmt, nbw);
    else
        launch_lds(k_split_fwd<16>, grid, kMixThreads, lds, s, X, sf, meta, cre, cim, lg2(B), F, (int)fmt, nbw);
}

void split_inv_gemm(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, float2* X, int F,
                    Fmt fmt, hipStream_t s) {
    const int B = c.bw;
    const int gj = B < 64 ? 0 : std::min(S2_SPLIT_GJMAX, B / 64);
    const int jw = gj ? 64 * gj : 32, jt = jw / 16;
    const int nb = (4 * F + kMixCols - 1) / kMixCols;
    const int nbw = (nb + kSplitBlocks - 1) / kSplitBlocks;
    const dim3 grid(B / 2 * nbw * (B / jw));  // order pairs x groups of column blocks x ring ranges
    const size_t lds = 2 * (size_t)std::min(jt, kSplitStageTiles) * kSplitTile;
    const __bf16* si = (const __bf16*)g.ai;
    const GemmMeta* meta = g.meta;
    switch (gj) {
        case 0: launch_lds(k_split_inv<0>, grid, kMixThreads, lds, s, cre, cim, si, meta, X, lg2(B), F, (int)fmt, nbw); break;
        case 1: launch_lds(k_split_inv<1>, grid, kMixThreads, lds, s, cre, cim, si, meta, X, lg2(B), F, (int)fmt, nbw); break;
        case 2: launch_lds(k_split_inv<2>, grid, kMixThreads, lds, s, cre, cim, si, meta, X, lg2(B), F, (int)fmt, nbw); break;
        default: launch_lds(k_split_inv<4>, grid, kMixThreads, lds, s, cre, cim, si, meta, X, lg2(B), F, (int)fmt, nbw); break;
    }
}

}  // namespace s2
