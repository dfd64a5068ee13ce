// Mixed form: FP32 ring FFTs and the naive Legendre transform (naive.c:34-95)
// as FP32 MFMA GEMMs over the orders; FP64 inputs and outputs, tables from
// the FP64 three-term recurrence (pml.c:41-79) rounded once to FP32.
//
//   forward   k_fphi_fwd<float> (s2_blocks.h) -> k_mix_fwd
//   inverse   k_mix_inv -> k_fphi_inv<float>
//
// Folding.  P_l^m(theta_{N-1-j}) = (-1)^(l-m) P_l^m(theta_j) and the
// quadrature weights are symmetric, so degrees with l-m even see
// x_j + x_{N-1-j} and odd ones x_j - x_{N-1-j}, over the bw rings j < bw.
//
// GEMMs.  The data columns are (field, +m or -m row, re or im); per order m
//   forward  C[col][l]  = sum_{j<bw} fold_p(X[col])[j] * w_j P_l^m(theta_j)
//   inverse  E/O[col][j] = sum_i coef[col][m+p+2i] * P_{m+p+2i}^m(theta_j)
// with ring j <- E+O and ring N-1-j <- E-O.  One launch covers every order:
// a block owns kMixCols columns of one order; workgroups are numbered so that
// each XCD takes orders m = 8q + k and bw-1-(8q+k) alternately, a cheap
// order (few degrees, bound by its data stream) next to an expensive one.
//
// Operands.  A (data or coefficients) goes from global memory straight into
// registers in the MFMA operand layout; the K index is permuted so that each
// lane reads S consecutive rings (degrees) per chunk.  B (the table) streams
// through two LDS buffers, one barrier per chunk: the next chunk is copied
// global -> LDS directly (global_load_lds, no staging registers) before the
// current chunk's MFMAs, and the forward reads its table fragments one tile
// group ahead of their MFMAs.  The registers this frees give three waves per
// SIMD (forward up to n0 = 8, inverse).  Measured against register staging,
// fragments read up front and two waves per SIMD, results bit-identical:
// bw=256 x 1024 forward GEMM 0.944 -> 0.873 ms, inverse 0.961 -> 0.905 ms
// (step 4.12 -> 4.00 ms); bw=512 x 128 forward 0.798 -> 0.764 ms (its n0 = 16
// instance's spills 118 -> 48 VGPRs; at three waves it spilled 155 and
// slowed to 0.91 ms, so it keeps two), inverse 0.797 -> 0.738 ms.
#include <algorithm>
#include <cmath>

#include "s2_mix_common.h"
#include "s2_recurrence.h"

namespace s2 {

typedef float fvec2 __attribute__((ext_vector_type(2)));

#ifndef S2_MIX_WPE_FWD
#define S2_MIX_WPE_FWD 3  // waves per SIMD; the n0 = 16 instance (bw > 256) keeps 2
#endif
#ifndef S2_MIX_WPE_INV
#define S2_MIX_WPE_INV 3
#endif
#ifndef S2_MIX_GJMAX
#define S2_MIX_GJMAX 4
#endif

// Tables.
//   NF(m) [j < bw][W], W = 32 n0(m): tiles t < n0 hold degrees l = m+2(16t+x),
//         tiles n0 + t hold l = m+1+2(16t+x), as w_j P_l^m(theta_j)
//         (w_j: even DH weight times sqrt(2pi)/N).  Tiles are interleaved in
//         fours -- entry x of tile 4g+v sits at 64g + 4x + v, a last pair at
//         64G + 2x + v -- so one float4 (float2) per lane feeds four (two)
//         16x16 MFMA tiles.  Entries past the last degree are zero.
//   NI(m,p) [i < ki(m)][j < bw]: P_{m+p+2i}^m(theta_j)/sqrt(2pi), rings interleaved in
//         groups of 64 the same way; rows past the last degree are zero.
#ifndef S2_MIX_INV_KC
#define S2_MIX_INV_KC 8
#endif
constexpr int kMixInvKC = S2_MIX_INV_KC;  // degrees per parity per inverse chunk
__host__ __device__ inline int mix_ki(int m, int B) {
    return kMixInvKC * ((B - m + 2 * kMixInvKC - 1) / (2 * kMixInvKC));
}
__host__ __device__ inline int mix_pos(int t, int x, int nt) {
    const int g4 = nt >> 2;
    return (t < 4 * g4) ? 64 * (t >> 2) + 4 * x + (t & 3) : 64 * g4 + 2 * x + (t & 1);
}
__host__ __device__ inline int ilv64(int j) { return (j & ~63) + 4 * (j & 15) + ((j >> 4) & 3); }
// bw = 32 has one group of 32 folded rings: position 2x + v holds ring 16v + x
// (one float2 per lane feeds two 16x16 MFMA tiles).
__host__ __device__ inline int ilv_ring(int j, int B) { return B >= 64 ? ilv64(j) : 2 * (j & 15) + ((j >> 4) & 1); }


// Rings per forward chunk: at least 16 MFMAs per wave between barriers.
#ifndef S2_MIX_KC_SMALL
#define S2_MIX_KC_SMALL 32
#endif
#ifndef S2_MIX_KC_BIG
#define S2_MIX_KC_BIG 16
#endif
__host__ __device__ constexpr int mix_kc(int n0) { return n0 <= 2 ? S2_MIX_KC_SMALL : S2_MIX_KC_BIG; }

__global__ void k_mix_pack(const double* __restrict__ w, const GemmMeta* __restrict__ meta, float* __restrict__ NF,
                           float* __restrict__ NI, int B) {
    const int Bp = mix_bp(B);
    const int m = blockIdx.x;
    const int N = 2 * B;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int nt = 2 * g0.n0, W = 16 * nt;
    const double cm = (m == 0) ? 0.0 : pmm_coeff(m);
    for (int j = threadIdx.x; j < B; j += blockDim.x) {
        const double th = (2.0 * j + 1.0) * M_PI / (2.0 * N);
        const double x = cos(th);
        const double wj = w[j];
        double cur = (m == 0) ? 0.70710678118654752440 : cm * pow(sin(th), (double)m);
        double prev = 0.0;
        for (int l = m; l < B; ++l) {
            const int p = (l - m) & 1, i = (l - m) >> 1;
            NF[g0.offF + (long)j * W + mix_pos(p * g0.n0 + (i >> 4), i & 15, nt)] = (float)(cur * wj);
            NI[(p ? g1.offI : g0.offI) + (long)i * Bp + ilv_ring(j, Bp)] = (float)(cur * 0.39894228040143267794);
            if (l + 1 < B) {
                const double nx = l2_an(m, l) * (cur * x) + l2_cn(m, l) * prev;
                prev = cur;
                cur = nx;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <int N0>
__device__ __forceinline__ void mix_fwd_body(const float2* __restrict__ X, const float* __restrict__ NF,
                                             const GemmMeta& g0, const GemmMeta& g1, double* __restrict__ cre,
                                             double* __restrict__ cim, int m, int cb, int B, int F, int fmt,
                                             float* sm) {
    constexpr int NT = 2 * N0, G4 = NT / 4, R2 = NT % 4;  // R2: 0, or 2 tiles in a last pair
    constexpr int W = 16 * NT;
    constexpr int KC = mix_kc(N0), S = KC / 4;
    constexpr int NV = KC * W / 4;  // float4 per table chunk (a whole number of KiB)
    static_assert(NV % 64 == 0, "table chunks go to LDS in 1 KiB wave transfers");
    const int N = 2 * B;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    if (cb * kMixCols >= ncols) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, kq = lane >> 4;

    // operand A: this lane's column, rings kc + kq*S + s (s < S) of a chunk
    const int col = cb * kMixCols + wave * 16 + x16;
    const bool imcol = col & 1;
    const float2* xr;
    {
        const int cc = col < ncols ? col : 0;  // past the end: read column 0, store nothing
        const int f = cc / cpf, rem = cc - f * cpf;
        xr = X + ((size_t)f * N + ((rem >> 1) ? N - m : m)) * N;
    }
    const fvec4* tsrc = reinterpret_cast<const fvec4*>(NF + g0.offF);

    fvec4 xa[S / 2], xb[S / 2];  // rings j0 .. j0+S-1 and N-j0-S .. N-j0-1 (pairs of complex values)
    auto fetch = [&](int ch) __attribute__((always_inline)) {
        // the table chunk straight into its LDS buffer (global_load_lds, no
        // staging registers), 1 KiB per wave instruction
        const unsigned char* src = reinterpret_cast<const unsigned char*>(tsrc + (size_t)ch * NV);
        unsigned char* dst = reinterpret_cast<unsigned char*>(sm + (ch & 1) * (KC * W));
        for (int u = wave; u < NV / 64; u += kMixThreads / 64)
            __builtin_amdgcn_global_load_lds((const void*)(src + u * 1024 + lane * 16),
                                             (__attribute__((address_space(3))) void*)(dst + u * 1024), 16, 0, 0);
        const int j0 = ch * KC + kq * S;
        const fvec4* pa = reinterpret_cast<const fvec4*>(xr + j0);
        const fvec4* pb = reinterpret_cast<const fvec4*>(xr + N - j0 - S);
#pragma unroll
        for (int h = 0; h < S / 2; ++h) {
            xa[h] = pa[h];
            xb[h] = pb[h];
        }
    };
    float a0[S], a1[S];  // folded operands: x_j + x_{N-1-j}, x_j - x_{N-1-j}
    auto fold = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const fvec4 va = xa[s >> 1];
            const fvec4 vb = xb[(S - 1 - s) >> 1];  // ring N-1-j0-s is entry S-1-s of the mirror run
            const float u = (s & 1) ? (imcol ? va.w : va.z) : (imcol ? va.y : va.x);
            const int sb = S - 1 - s;
            const float v = (sb & 1) ? (imcol ? vb.w : vb.z) : (imcol ? vb.y : vb.x);
            a0[s] = u + v;
            a1[s] = u - v;
        }
    };

    fvec4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = fvec4{0.f, 0.f, 0.f, 0.f};
    constexpr int TB = KC * W;  // floats per table buffer
    fetch(0);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS transfers have landed
    __syncthreads();
    const int nch = mix_bp(B) / KC;  // rings past bw meet zero table rows
    for (int ch = 0; ch < nch; ++ch) {
        const float* tb = sm + (ch & 1) * TB + kq * S * W;  // this lane's table rows: ring kq*S + s
        fold();
        const bool more = ch + 1 < nch;
        if (more) fetch(ch + 1);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const float* r = tb + s * W;
            // table fragments streamed one tile group ahead of their MFMAs
            // (two float4 live instead of G4)
            fvec2 tp;
            if (R2) tp = *reinterpret_cast<const fvec2*>(r + 64 * G4 + 2 * x16);
            fvec4 tcur = *reinterpret_cast<const fvec4*>(r + 4 * x16);
#pragma unroll
            for (int g = 0; g < G4; ++g) {
                const fvec4 tnext = *reinterpret_cast<const fvec4*>(r + 64 * (g + 1 < G4 ? g + 1 : g) + 4 * x16);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int t = 4 * g + v;
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(t < N0 ? a0[s] : a1[s], tcur[v], acc[t], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
                tcur = tnext;
            }
            if (R2) {
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const int t = 4 * G4 + v;
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(t < N0 ? a0[s] : a1[s], tp[v], acc[t], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS transfers (chunk ch + 1) have landed
        __syncthreads();
    }

    mix_fwd_store<N0>(acc, cre, cim, m, cb, B, F, fmt);
}

template <int MAXN0>
__global__ __launch_bounds__(kMixThreads) __attribute__((amdgpu_waves_per_eu(MAXN0 > 8 ? 2 : S2_MIX_WPE_FWD))) void k_mix_fwd(const float2* __restrict__ X, const float* __restrict__ NF,
                                                        const GemmMeta* __restrict__ meta, double* __restrict__ cre,
                                                        double* __restrict__ cim, int B, int F, int fmt, int nb) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* sm = reinterpret_cast<float*>(smem);
    int mp, cb;
    mix_pair(nb, mp, cb);
    if (mp >= mix_npairs(B)) return;
    for (int h = 0; h < 2; ++h) {
        if (h && B - 1 - mp == mp) break;  // odd bw: the middle order pairs with itself
        const int m = h ? B - 1 - mp : mp;
        const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
#define S2_MIX_FWD(n) \
    case n:           \
        if (n <= MAXN0) mix_fwd_body<(n <= MAXN0 ? n : 1)>(X, NF, g0, g1, cre, cim, m, cb, B, F, fmt, sm); \
        break;
    switch (g0.n0) {
        S2_MIX_FWD(1)
        S2_MIX_FWD(2)
        S2_MIX_FWD(3)
        S2_MIX_FWD(4)
        S2_MIX_FWD(5)
        S2_MIX_FWD(6)
        S2_MIX_FWD(7)
        S2_MIX_FWD(8)
        S2_MIX_FWD(9)
        S2_MIX_FWD(10)
        S2_MIX_FWD(11)
        S2_MIX_FWD(12)
        S2_MIX_FWD(13)
        S2_MIX_FWD(14)
        S2_MIX_FWD(15)
        S2_MIX_FWD(16)
        default: break;
    }
#undef S2_MIX_FWD
    }
}

// ---------------------------------------------------------------------------
// inverse
// ---------------------------------------------------------------------------
// A block owns kMixCols columns of one order and JW = 64*GJ of the bw folded
// rings (32 for GJ = 0, bw = 32; jb selects which); chunks of KC degrees per
// parity.
template <int GJ>
__device__ __forceinline__ void mix_inv_body(const double* __restrict__ cre, const double* __restrict__ cim,
                                             const float* __restrict__ NI, const GemmMeta* __restrict__ meta,
                                             float2* __restrict__ X, int B, int F, int fmt, int m, int cz, float* sm) {
    constexpr int JW = mix_jw<GJ>(), JT = mix_jt<GJ>();
    constexpr int KC = kMixInvKC, S = KC / 4;
    constexpr int NV = 2 * KC * JW / 4;  // float4 per table chunk (both parities)
    constexpr int PT = (NV + kMixThreads - 1) / kMixThreads;
    static_assert(NV % kMixThreads == 0 || NV < kMixThreads, "table chunk must split evenly over the block");
    constexpr bool FULL = NV % kMixThreads == 0;  // else threads tid >= NV carry no table float4
    const int Bp = mix_bp(B);  // rings of the table rows (zero past bw)
    const int nz = Bp / JW;
    const int cb = cz / nz, jbase = (cz - cb * nz) * JW;
    const int cpf = (m > 0 && fmt == 0) ? 4 : 2;
    const int ncols = F * cpf;
    if (cb * kMixCols >= ncols) return;
    const GemmMeta g0 = meta[m * 2], g1 = meta[m * 2 + 1];
    const int L = B - m;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x16 = lane & 15, kq = lane >> 4;
    const size_t B2 = (size_t)B * B;

    // operand A: coefficients of this lane's column, degree offsets
    // d = 2(ch*KC + kq*S) + e, e < 2S (parity e & 1, index s = e >> 1)
    const int col = cb * kMixCols + wave * 16 + x16;
    const double* cp;
    {
        const int cc = col < ncols ? col : 0;
        const int f = cc / cpf, rem = cc - f * cpf;
        cp = ((rem & 1) ? cim : cre) + f * B2 + index_of_coeff((rem >> 1) ? -m : m, m, B);
    }
    // table: float4 e = tid + u*kMixThreads of a chunk is parity p, row rw, quad c4
    const float* tsrc[PT];
#pragma unroll
    for (int u = 0; u < PT; ++u) {
        const int e = FULL ? tid + u * kMixThreads : min(tid, NV - 1);
        const int p = e / (KC * JW / 4), rest = e - p * (KC * JW / 4);
        const int rw = rest / (JW / 4), c4 = rest - rw * (JW / 4);
        tsrc[u] = NI + (p ? g1.offI : g0.offI) + (size_t)rw * Bp + jbase + 4 * c4;
    }
    // rows -m of odd m carry the sign (-1)^m: applied to the operand
    const unsigned nsign = ((((wave * 16 + x16) % cpf) >> 1) && (m & 1)) ? 0x80000000u : 0u;
    dpair8 cv[S];  // degree offsets d0 + 2h (even parity), d0 + 2h + 1 (odd)
    constexpr int TB = 2 * KC * JW;  // floats per table buffer
    auto fetch = [&](int ch) __attribute__((always_inline)) {
        // the table chunk gathered straight into its LDS buffer (global_load_lds:
        // a global address per lane, 64 lanes' float4 contiguous in LDS): float4
        // e of the chunk at position e
        unsigned char* dst = reinterpret_cast<unsigned char*>(sm + (ch & 1) * TB);
#pragma unroll
        for (int u = 0; u < PT; ++u)
            if (FULL || u * kMixThreads + wave * 64 < NV)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(tsrc[u] + (size_t)ch * KC * Bp),
                    (__attribute__((address_space(3))) void*)(dst + (u * kMixThreads + wave * 64) * 16), 16, 0, 0);
        const int d0 = 2 * (ch * KC + kq * S);
        if (d0 + 2 * S <= L) {
#pragma unroll
            for (int h = 0; h < S; ++h) cv[h] = *reinterpret_cast<const dpair8*>(cp + d0 + 2 * h);
        } else {  // the run ends inside this lane's degrees
#pragma unroll
            for (int h = 0; h < S; ++h)
                cv[h] = dpair8{d0 + 2 * h < L ? cp[d0 + 2 * h] : 0.0, d0 + 2 * h + 1 < L ? cp[d0 + 2 * h + 1] : 0.0};
        }
    };
    float a[2][S];
    auto convert = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int h = 0; h < S; ++h) {
            a[0][h] = __uint_as_float(__float_as_uint((float)cv[h].x) ^ nsign);
            a[1][h] = __uint_as_float(__float_as_uint((float)cv[h].y) ^ nsign);
        }
    };

    fvec4 acc[2][JT];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int t = 0; t < JT; ++t) acc[p][t] = fvec4{0.f, 0.f, 0.f, 0.f};
    fetch(0);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS transfers have landed
    __syncthreads();
    const int nch = g0.ki / KC;
    for (int ch = 0; ch < nch; ++ch) {
        const float* tb = sm + (ch & 1) * TB;
        convert();
        const bool more = ch + 1 < nch;
        if (more) fetch(ch + 1);
        // table operands of one K step: a float4 feeds 4 ring tiles (a float2
        // the 2 tiles of a 32-ring block)
        constexpr int GQ = GJ ? GJ : 1, TV = GJ ? 4 : 2;
        fvec4 tq[2][2][GQ];
        auto lds = [&](int s, int b) __attribute__((always_inline)) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const float* r = tb + (p * KC + kq * S + s) * JW;
                if constexpr (GJ == 0) {
                    const fvec2 h = *reinterpret_cast<const fvec2*>(r + 2 * x16);
                    tq[b][p][0] = fvec4{h.x, h.y, 0.f, 0.f};
                } else {
#pragma unroll
                    for (int g = 0; g < GJ; ++g)
                        tq[b][p][g] = *reinterpret_cast<const fvec4*>(r + 64 * g + 4 * x16);
                }
            }
        };
        lds(0, 0);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int b = s & 1;
            if (s + 1 < S) lds(s + 1, b ^ 1);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int g = 0; g < GQ; ++g) {
#pragma unroll
                    for (int v = 0; v < TV; ++v)
                        acc[p][TV * g + v] =
                            __builtin_amdgcn_mfma_f32_16x16x4f32(a[p][s], tq[b][p][g][v], acc[p][TV * g + v], 0, 0, 0);
                }
        }
        __builtin_amdgcn_s_waitcnt(0);  // this wave's LDS transfers (chunk ch + 1) have landed
        __syncthreads();
    }

    mix_inv_store<JT>(acc, X, m, cb, jbase, B, F, fmt);
}

template <int GJ>
__global__ __launch_bounds__(kMixThreads) __attribute__((amdgpu_waves_per_eu(S2_MIX_WPE_INV))) void k_mix_inv(
    const double* __restrict__ cre, const double* __restrict__ cim, const float* __restrict__ NI,
    const GemmMeta* __restrict__ meta, float2* __restrict__ X, int B, int F, int fmt, int nb) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* sm = reinterpret_cast<float*>(smem);
    int mp, cz;
    mix_pair(nb * (mix_bp(B) / mix_jw<GJ>()), mp, cz);
    if (mp >= mix_npairs(B)) return;
    mix_inv_body<GJ>(cre, cim, NI, meta, X, B, F, fmt, mp, cz, sm);
    if (B - 1 - mp == mp) return;  // odd bw: the middle order pairs with itself
    __syncthreads();  // the partner order's prologue reuses the table buffers
    mix_inv_body<GJ>(cre, cim, NI, meta, X, B, F, fmt, B - 1 - mp, cz, sm);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// 64-ring groups per inverse block that divide the padded ring count,
// the default (the largest) first.
static int mix_gj_default(int bw) {
    const int g = mix_bp(bw) / 64;
    if (g == 0) return 0;  // one 32-ring block
    int gj = 1;
    while (2 * gj <= S2_MIX_GJMAX && g % (2 * gj) == 0) gj *= 2;
    return gj;
}

std::vector<int> mix_gj_choices(int bw) {
    std::vector<int> v{mix_gj_default(bw)};
    if (v[0] == 0) return v;
    for (int gj = 1; gj < v[0]; gj <<= 1) v.push_back(gj);
    return v;
}

bool mixed_supported(int bw) { return bw >= 32 && bw <= 512; }

void naive_tables_create(GemmTables& g, const DevConsts& c, hipStream_t s) {
    const int B = c.bw;
    g.bw = B;
    g.prec = 2;
    g.h_meta.assign(2 * B, GemmMeta{});
    long of = 0, oi = 0;
    for (int m = 0; m < B; ++m) {
        const int n0 = mix_n0(m, B), ki = mix_ki(m, B);
        for (int p = 0; p < 2; ++p) {
            GemmMeta& e = g.h_meta[m * 2 + p];
            e.Lp = gm_lp(m, p, B);
            e.Kf = B;
            e.n0 = n0;
            e.ki = ki;
            e.offF = of;
            e.offI = oi;
            oi += (long)ki * mix_bp(B);
        }
        of += (long)mix_bp(B) * 32 * n0;
    }
    g.nf = of;
    g.ni = oi;
    S2_HIP(hipMalloc(&g.meta, sizeof(GemmMeta) * 2 * B));
    S2_HIP(hipMalloc(&g.af, sizeof(float) * std::max(1L, of)));
    S2_HIP(hipMalloc(&g.ai, sizeof(float) * std::max(1L, oi)));
    S2_HIP(hipMemsetAsync(g.af, 0, sizeof(float) * std::max(1L, of), s));
    S2_HIP(hipMemsetAsync(g.ai, 0, sizeof(float) * std::max(1L, oi), s));
    S2_HIP(hipMemcpyAsync(g.meta, g.h_meta.data(), sizeof(GemmMeta) * 2 * B, hipMemcpyHostToDevice, s));
    k_mix_pack<<<B, 256, 0, s>>>(c.w_naive, g.meta, (float*)g.af, (float*)g.ai, B);
    S2_HIP(hipGetLastError());
    S2_HIP(hipStreamSynchronize(s));
}

// LDS of k_mix_fwd: two table chunks of the widest order.
static size_t mix_fwd_lds(int B) {
    size_t mx = 0;
    for (int n0 = 1; n0 <= mix_n0(0, B); ++n0) mx = std::max(mx, sizeof(float) * 2 * mix_kc(n0) * 32 * n0);
    return mx;
}

static void mix_fwd_gemm(const DevConsts& c, const GemmTables& g, const float2* X, double* cre, double* cim, int F,
                         Fmt fmt, hipStream_t s) {
    const int B = c.bw;
    const int nb = (4 * F + kMixCols - 1) / kMixCols;
    const dim3 grid(mix_npairs8(B) * nb);  // order pairs x column blocks
    const size_t lds = mix_fwd_lds(B);
    const float* nf = (const float*)g.af;
    const GemmMeta* meta = g.meta;
    const int n0max = mix_n0(0, B);
    if (n0max <= 2)
        launch_lds(k_mix_fwd<2>, grid, kMixThreads, lds, s, X, nf, meta, cre, cim, B, F, (int)fmt, nb);
    else if (n0max <= 4)
        launch_lds(k_mix_fwd<4>, grid, kMixThreads, lds, s, X, nf, meta, cre, cim, B, F, (int)fmt, nb);
    else if (n0max <= 8)
        launch_lds(k_mix_fwd<8>, grid, kMixThreads, lds, s, X, nf, meta, cre, cim, B, F, (int)fmt, nb);
    else
        launch_lds(k_mix_fwd<16>, grid, kMixThreads, lds, s, X, nf, meta, cre, cim, B, F, (int)fmt, nb);
}

static void mix_inv_gemm(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, float2* X,
                         int F, Fmt fmt, hipStream_t s) {
    const int B = c.bw;
    const int gj = B <= 32 ? 0 : g.t_gj ? g.t_gj : mix_gj_default(B);
    const int jw = gj ? 64 * gj : 32;
    const int nb = (4 * F + kMixCols - 1) / kMixCols;
    const dim3 grid(mix_npairs8(B) * nb * (mix_bp(B) / jw));  // order pairs x column blocks x ring ranges
    const size_t lds = sizeof(float) * 2 * (2 * kMixInvKC * jw);
    const float* ni = (const float*)g.ai;
    const GemmMeta* meta = g.meta;
    switch (gj) {
        case 0: launch_lds(k_mix_inv<0>, grid, kMixThreads, lds, s, cre, cim, ni, meta, X, B, F, (int)fmt, nb); break;
        case 1: launch_lds(k_mix_inv<1>, grid, kMixThreads, lds, s, cre, cim, ni, meta, X, B, F, (int)fmt, nb); break;
        case 2: launch_lds(k_mix_inv<2>, grid, kMixThreads, lds, s, cre, cim, ni, meta, X, B, F, (int)fmt, nb); break;
        default: launch_lds(k_mix_inv<4>, grid, kMixThreads, lds, s, cre, cim, ni, meta, X, B, F, (int)fmt, nb); break;
    }
}

void forward_mixed(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                   double* cim, int F, Fmt fmt, void* Xv, hipStream_t s, StageTimer* tm) {
    constexpr int V = 16;
    const int N = c.N, lgN = lg2(N);
    const int rpb = V * kMixFftThreads / N;
    float2* X = (float2*)Xv;
    {
        Span sp(tm, ST_PHI_FWD, s);
        const size_t lds = sizeof(float2) * ((size_t)rpb * (N + 1) + N);
        if (!c.pow2) {  // mixed-radix FP32 ring FFTs
            ring_fft_fwd_f32(c, re, im, X, F, s);
        } else if (rpb > N) {  // bw = 32: the 64 rings of a field in one 256-thread block
            constexpr int T2 = 256;
            const size_t l2 = sizeof(float2) * ((size_t)N * (N + 1) + N);
            launch_lds(k_fphi_fwd<float, V, T2, true>, dim3(1, F), T2, l2, s, re, im, X, lgN,
                       (const float2*)c.tw32);
        } else if (k_fphi_fwd_nt(rpb, sizeof(float2)))
            launch_lds(k_fphi_fwd<float, V, kMixFftThreads, true>, dim3(N / rpb, F), kMixFftThreads, lds, s, re, im,
                       X, lgN, (const float2*)c.tw32);
        else
            launch_lds(k_fphi_fwd<float, V, kMixFftThreads>, dim3(N / rpb, F), kMixFftThreads, lds, s, re, im, X,
                       lgN, (const float2*)c.tw32);
    }
    {
        Span sp(tm, ST_LEG_FWD, s);
        if (g.prec == 3)
            split_fwd_gemm(c, g, X, cre, cim, F, fmt, s);
        else
            mix_fwd_gemm(c, g, X, cre, cim, F, fmt, s);
    }
}

void inverse_mixed(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                   double* im, int F, Fmt fmt, void* Xv, hipStream_t s, StageTimer* tm) {
    constexpr int V = 16;
    const int N = c.N, lgN = lg2(N);
    const int rpb = V * kMixFftThreads / N;
    float2* X = (float2*)Xv;
    {
        Span sp(tm, ST_LEG_INV, s);
        if (g.prec == 3)
            split_inv_gemm(c, g, cre, cim, X, F, fmt, s);
        else
            mix_inv_gemm(c, g, cre, cim, X, F, fmt, s);
    }
    {
        Span sp(tm, ST_PHI_INV, s);
        const size_t lds = sizeof(float2) * ((size_t)rpb * (N + 1) + N);
        if (!c.pow2) {  // mixed-radix FP32 ring FFTs
            ring_fft_inv_f32(c, X, re, im, F, s);
        } else if (rpb > N) {  // bw = 32: the 64 rings of a field in one 256-thread block
            constexpr int T2 = 256;
            const size_t l2 = sizeof(float2) * ((size_t)N * (N + 1) + N);
            launch_lds(k_fphi_inv<float, V, T2>, dim3(1, F), T2, l2, s, (const float2*)X, re, im, lgN,
                       (const float2*)c.tw32);
        } else if ((((uintptr_t)re | (uintptr_t)im) & 15) == 0)
            launch_lds(k_fphi_inv<float, V, kMixFftThreads, true>, dim3(N / rpb, F), kMixFftThreads, lds, s,
                       (const float2*)X, re, im, lgN, (const float2*)c.tw32);
        else
            launch_lds(k_fphi_inv<float, V, kMixFftThreads>, dim3(N / rpb, F), kMixFftThreads, lds, s,
                       (const float2*)X, re, im, lgN, (const float2*)c.tw32);
    }
}

}  // namespace s2
