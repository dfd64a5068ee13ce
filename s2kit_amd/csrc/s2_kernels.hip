// HIP kernels of the memoised seminaive SHT for gfx950 (MI355X).
//
// Forward, per field (FSTSemiMemo, FST_semi_memo.c:68-202):
//   k_phi_fwd   ring FFT along phi, order-major (transposed) store     (:81)
//   k_dct_fwd   quadrature weights + DCT-II along theta per order row  (seminaive.c:161-176)
//   k_leg_fwd   projections on the zero-stripped cosine tables         (seminaive.c:183-197)
//   k_naive_fwd dense projection on sampled P_l^m for orders >= cutoff (naive.c:34-59)
// Inverse (InvFSTSemiMemo, :228-351):
//   k_leg_inv   cosine sums against the transposed tables              (seminaive.c:74-98)
//   k_dct_inv   DCT-III, x sin for odd orders, signs, 1/sqrt(2pi)      (seminaive.c:107-114, :301-348)
//   k_naive_inv dense synthesis for orders >= cutoff                   (naive.c:78-95)
//   k_phi_inv   inverse ring FFT along phi                             (:350)
// Tables (cospml.c:161-362, pml.c:41-79, weights.c:32-46) are generated on
// the GPU in FP64 by k_gen_cospml / k_transpose / k_gen_pml / k_weights.
//
// The ring FFTs of these kernels come in four instances: radix-2 Stockham
// passes in LDS for power-of-two lengths, in-place register passes of radix
// 2/4/8 and odd primes up to 19 for lengths with no larger prime factor, and
// Bluestein's convolution (power-of-two or 7-smooth length) otherwise;
// power-of-two grids of 64 .. 2048 points take the register radix-8 kernels
// of s2_blocks.h instead.  Both DCTs use one complex FFT of length N = 2B on
// a reordered sequence (Makhoul); the DCTs are real-linear with real kernels,
// so one complex row carries the real and imaginary transforms together.
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <vector>

#include "s2_complex.h"
#include "s2_internal.h"
#include "s2_recurrence.h"
#include "s2_blocks.h"

namespace s2 {

// ---------------------------------------------------------------------------
// block-wide FFTs in LDS
// ---------------------------------------------------------------------------

// Division by a divisor 1 <= d < 2^16 of a value x < 2^16: floor(x/d) =
// (x * ceil(2^32/d)) >> 32 (the rounding error x*(ceil(2^32/d)*d - 2^32)/2^32
// stays below 1/d).
__device__ __forceinline__ unsigned long long magic16(unsigned d) {
    // ceil(2^32/d) = floor((2^32-1)/d) + 1 for d >= 2 (a 32-bit division)
    return d == 1 ? 0x100000000ull : (unsigned long long)(0xFFFFFFFFu / d) + 1;
}
__device__ __forceinline__ int div16(int x, unsigned long long mg) { return (int)(((unsigned long long)x * mg) >> 32); }

// Threads of the mixed-radix and Bluestein instances (two blocks of 8 waves
// per CU keep their LDS passes' latency covered).
#ifndef S2_MR_THREADS
#define S2_MR_THREADS 512
#endif
constexpr int kMrThreads = S2_MR_THREADS;
// The in-place instances (MODE >= 1) hold two 512-thread blocks per CU when
// they stay within 128 VGPRs (4 waves per SIMD).
#define S2_FFT_OCCUPANCY(MODE) __attribute__((amdgpu_waves_per_eu((MODE) >= 1 ? 4 : 1)))

// ---------------------------------------------------------------------------
// In-place mixed-radix FFTs with register butterflies (lengths whose prime
// factors are at most kRegMaxPrime).  Each pass of radix R is one Stockham
// step: every thread loads the R inputs of each of its butterflies (up to
// PER of them) into registers, the block synchronises, and each thread
// twiddles, transforms and stores its butterflies over the same LDS buffer.
// Against two-buffer Stockham passes this halves the LDS of a block (twice
// the rings per block); each input is read once per pass.
// ---------------------------------------------------------------------------
constexpr int kRegMaxPrime = 19;
// Bluestein's convolution lengths use radices up to 7 only: the radix-11..19
// butterflies' registers would not leave room for two blocks per CU there.
constexpr int kBluMaxPrime = 7;
#ifndef S2_RM_ELEMS64
#define S2_RM_ELEMS64 4096
#endif
#ifndef S2_RM_ELEMS32
#define S2_RM_ELEMS32 8192
#endif
// values per block of the register instances (FP64: 64 KiB, FP32: 64 KiB)
template <typename T>
constexpr int rm_elems() {
    return sizeof(T) == 16 ? S2_RM_ELEMS64 : S2_RM_ELEMS32;
}

// a block-uniform value in scalar registers
__device__ __forceinline__ double uniform(double x) {
    const long long b = __builtin_bit_cast(long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float uniform(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

// R-point DFT of v stored to out[p*Ns], y_p = sum_q v_q w^{pq},
// w = exp(SIGN 2 pi i/R).  Odd R (prime or not) by conjugate pairs: with s_q = v_q + v_{R-q},
// d_q = v_q - v_{R-q} (q <= H = (R-1)/2),
//   y_p, y_{R-p} = v_0 + sum_q s_q cos(2 pi pq/R) +- i SIGN sum_q d_q sin(2 pi pq/R),
// cs[t], sn[t] = cos, sin(2 pi t/R) for t <= H.  v is overwritten.
template <typename T, int SIGN, int R>
__device__ __forceinline__ void dft_store(T* v, const typename Real<T>::type* cs, const typename Real<T>::type* sn,
                                          T* out, int Ns) {
    using F = typename Real<T>::type;
    if constexpr (R == 2) {
        out[0] = cadd(v[0], v[1]);
        out[Ns] = csub(v[0], v[1]);
    } else if constexpr (R == 8) {
        dft8<SIGN>(v);
#pragma unroll
        for (int p = 0; p < 8; ++p) out[p * Ns] = v[p];
    } else if constexpr (R == 4) {
        const T s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const T s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        const T r13 = SIGN < 0 ? mk<T>(d13.y, -d13.x) : mk<T>(-d13.y, d13.x);  // SIGN i d13
        out[0] = cadd(s02, s13);
        out[2 * Ns] = csub(s02, s13);
        out[Ns] = cadd(d02, r13);
        out[3 * Ns] = csub(d02, r13);
    } else {
        constexpr int H = (R - 1) / 2;
        T y0 = v[0];
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            const T s = cadd(v[q], v[R - q]), d = csub(v[q], v[R - q]);
            v[q] = s;
            v[R - q] = d;
            y0 = cadd(y0, s);
        }
        out[0] = y0;
#pragma unroll
        for (int p = 1; p <= H; ++p) {
            F ar = v[0].x, ai = v[0].y, br = 0, bi = 0;
#pragma unroll
            for (int q = 1; q <= H; ++q) {
                const int r = (p * q) % R;
                const int t = r <= H ? r : R - r;
                const F sv = r <= H ? sn[t] : -sn[t];
                ar += v[q].x * cs[t];
                ai += v[q].y * cs[t];
                br += v[R - q].x * sv;
                bi += v[R - q].y * sv;
            }
            // +- i SIGN (br + i bi) = +- SIGN (-bi + i br)
            out[p * Ns] = mk<T>(ar - (F)SIGN * bi, ai + (F)SIGN * br);
            out[(R - p) * Ns] = mk<T>(ar + (F)SIGN * bi, ai - (F)SIGN * br);
        }
    }
}

// One in-place pass of radix R over nfft length-N transforms (row stride N):
// butterfly j = jq*Ns + k (k < Ns, j < N/R) reads a[j + q*N/R], applies the
// twiddles exp(SIGN 2 pi i q k / (Ns*R)), runs an R-point DFT and writes
// output p to a[jq*Ns*R + p*Ns + k].  tw[t] = exp(-2 pi i t / N) (in LDS).
template <typename T, int SIGN, int R, int PER>
__device__ __forceinline__ void mr_reg_pass(T* a, int N, int Ns, int nfft, const T* __restrict__ tw) {
    using F = typename Real<T>::type;
    constexpr int H = (R - 1) / 2;
    const int NR = N / R, step = N / (Ns * R);
    const unsigned long long mNR = magic16(NR), mNs = magic16(Ns);
    const int total = nfft * NR;
    T v[PER][R];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int bf = threadIdx.x + u * kMrThreads;
        if (bf < total) {
            const int f = div16(bf, mNR), j = bf - f * NR;
            const T* in = a + f * N + j;
#pragma unroll
            for (int q = 0; q < R; ++q) v[u][q] = in[q * NR];
        }
    }
    // cos and sin of 2 pi t/R from the table: tw[t N/R] = exp(-2 pi i t/R)
    // (t = 0 is met by composite R: pq = 0 mod R)
    F cs[H + 1], sn[H + 1];
    cs[0] = (F)1;
    sn[0] = (F)0;
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        const T w = tw[t * NR];
        cs[t] = uniform(w.x);
        sn[t] = uniform(-w.y);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int bf = threadIdx.x + u * kMrThreads;
        if (bf < total) {
            const int f = div16(bf, mNR), j = bf - f * NR;
            const int jq = div16(j, mNs), k = j - jq * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    T w = tw[q * k * step];  // q*k*step < R*Ns*step = N
                    if (SIGN > 0) w.y = -w.y;
                    v[u][q] = cmul(v[u][q], w);
                }
            }
            dft_store<T, SIGN, R>(v[u], cs, sn, a + f * N + jq * Ns * R + k, Ns);
        }
    }
    __syncthreads();
}

// Butterflies per thread of a radix-R pass over at most E values.
template <int E, int R>
constexpr int rm_per() {
    return (E + R * kMrThreads - 1) / (R * kMrThreads);
}

// Radix of the next pass over the remaining length rem: 8, 4 or 2 while
// rem is even, then the smallest odd prime factor.  (Odd composites 15 and 9
// whole in registers, fewer passes, measured on MI355X: FP64 bw=263 -7%,
// bw=37 +9% per step as the smooth-length instance went from 86 to 124
// VGPRs, bw=255 and 100 unchanged; in FP32 their O(R) conjugate-pair sums
// took the mixed form at bw = 255 from below 1e-6 to 1.02e-6.  Not kept.)
__device__ __forceinline__ int mr_reg_radix(int rem) {
    if (rem % 8 == 0) return 8;
    if (rem % 4 == 0) return 4;
    if (rem % 2 == 0) return 2;
    int R = 3;
    while (rem % R) R += 2;
    return R;
}

// In-place FFTs (natural order in and out) of nfft rows of length N, row
// stride N, nfft*N <= rm_elems<T>(); every prime factor of N at most
// PMAX (rm_supported).  One pass per radix of mr_reg_radix.
template <typename T, int SIGN, int PMAX = kRegMaxPrime>
__device__ __forceinline__ void mr_reg_fft(T* a, int N, int nfft, const T* __restrict__ tw) {
    constexpr int E = rm_elems<T>();
    int rem = N, Ns = 1;
    while (rem > 1) {
        const int R = mr_reg_radix(rem);
        switch (R) {
            case 2: mr_reg_pass<T, SIGN, 2, rm_per<E, 2>()>(a, N, Ns, nfft, tw); break;
            case 3: mr_reg_pass<T, SIGN, 3, rm_per<E, 3>()>(a, N, Ns, nfft, tw); break;
            case 4: mr_reg_pass<T, SIGN, 4, rm_per<E, 4>()>(a, N, Ns, nfft, tw); break;
            case 8: mr_reg_pass<T, SIGN, 8, rm_per<E, 8>()>(a, N, Ns, nfft, tw); break;
            case 5: mr_reg_pass<T, SIGN, 5, rm_per<E, 5>()>(a, N, Ns, nfft, tw); break;
            case 7: mr_reg_pass<T, SIGN, 7, rm_per<E, 7>()>(a, N, Ns, nfft, tw); break;
            case 11: if constexpr (PMAX >= 11) mr_reg_pass<T, SIGN, 11, rm_per<E, 11>()>(a, N, Ns, nfft, tw); break;
            case 13: if constexpr (PMAX >= 13) mr_reg_pass<T, SIGN, 13, rm_per<E, 13>()>(a, N, Ns, nfft, tw); break;
            case 17: if constexpr (PMAX >= 17) mr_reg_pass<T, SIGN, 17, rm_per<E, 17>()>(a, N, Ns, nfft, tw); break;
            case 19: if constexpr (PMAX >= 19) mr_reg_pass<T, SIGN, 19, rm_per<E, 19>()>(a, N, Ns, nfft, tw); break;
            default: break;  // excluded on the host (rm_supported, bluestein_length)
        }
        rem /= R;
        Ns *= R;
    }
}
// Loads per thread that the load phases of the per-field FFT/DCT kernels keep
// in flight before storing to LDS (4 covers a 2048-value block at 512
// threads in one batch).
constexpr int kFillBatch = 4;


// nfft independent power-of-two transforms of length N stored back to back
// in a, radix-2 Stockham passes between a and b (scratch of the same size).
// SIGN = -1: forward, +1: inverse (unnormalised).  tw[t] = exp(-2 pi i t / N).
// Returns the buffer holding the result.
template <typename T, int SIGN>
__device__ T* radix2_fft(T* a, T* b, int N, int logN, int nfft, const T* __restrict__ tw) {
    const int half = N >> 1;
    const int total = nfft * half;
    for (int s = 0, Ns = 1; s < logN; ++s, Ns <<= 1) {
        const int tstride = half >> s;  // N / (2 Ns)
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int f = idx >> (logN - 1);
            const int j = idx & (half - 1);
            const T* in = a + f * N;
            T* out = b + f * N;
            const int k = j & (Ns - 1);
            T w = tw[k * tstride];
            if (SIGN > 0) w.y = -w.y;
            const T x0 = in[j];
            const T x1 = cmul(in[j + half], w);
            const int o = ((j - k) << 1) + k;
            out[o] = cadd(x0, x1);
            out[o + Ns] = csub(x0, x1);
        }
        __syncthreads();
        T* t = a;
        a = b;
        b = t;
    }
    return a;
}

// ---------------------------------------------------------------------------
// Bluestein's algorithm for ring lengths with a prime factor above
// kRegMaxPrime: with w_t = exp(SIGN i pi t^2 / N),
//   y_k = sum_n x_n exp(SIGN 2 pi i nk/N) = w_k sum_n (x_n w_n) conj(w_{k-n}),
// a circular convolution of length M >= 2N-1 run as two FFTs of length M:
// radix-8 register passes (fft_block) for a power of two, the register
// mixed-radix passes (mr_reg_fft) for M whose primes are at most
// kRegMaxPrime.  Tables (consts_create), for the forward sign:
// g[0..M) = exp(-2 pi i t/M), g[M..M+N) = w_t, g[M+N..2M+N) = FFT_M(b)/M with
// b_t = conj(w_|t|) at t mod M (|t| < N); the inverse sign uses conj(w) and
// conj(Bh[-k mod M]).  blu_rows rows of blu_stride values per block, all of
// them transformed (rows >= nfft are zeroed first).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr bool is_pow2(int M) { return (M & (M - 1)) == 0; }
__host__ __device__ constexpr int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ constexpr int imax(int a, int b) { return a > b ? a : b; }
// row stride: power-of-two rows padded against bank conflicts
__host__ __device__ constexpr int blu_stride(int M) { return is_pow2(M) ? M + 1 : M; }
// rows per block: fft_block transforms exactly rm_elems/M rows; mixed-radix
// rows: at most rm_elems values, and with the M twiddles within 80 KiB of
// LDS (two blocks per CU)
template <typename T>
__host__ __device__ constexpr int blu_rows(int M) {
    return is_pow2(M) ? rm_elems<T>() / M
                      : imax(1, imin(rm_elems<T>() / M, (81920 / (int)sizeof(T) - M) / M));
}

// A power-of-two M keeps its M twiddles in LDS while the block stays within
// 80 KiB (two blocks per CU); beyond that it stages the two-level table of
// twiddle() instead (bw=509 x 64 fields FP64: 2.33 -> 1.53 ms per FFT/DCT
// launch; where the full table fits, the extra complex product per twiddle
// cost up to 5%).
template <typename T>
__host__ __device__ constexpr bool blu_tw2(int M) {
    return is_pow2(M) && (size_t)sizeof(T) * ((size_t)blu_rows<T>(M) * blu_stride(M) + M) > 81920;
}
__host__ __device__ constexpr int tw2_values(int M) {
    return (M >> tw2_bits(__builtin_ctz(M))) + (1 << tw2_bits(__builtin_ctz(M)));
}

template <typename T, int SIGN, bool P2>
__device__ __forceinline__ void bluestein_rows(T* a, int N, int M, int nfft, const T* twl, const T* __restrict__ g) {
    constexpr int V = rm_elems<T>() / kMrThreads;
    constexpr bool p2 = P2;  // is_pow2(M): separate instances keep each path's registers
    const int S = blu_stride(M), rows = blu_rows<T>(M), tot = rows * M;
    const unsigned long long mM = magic16(M);
    const T* __restrict__ w = g + M;
    const T* __restrict__ bh = g + M + N;
    for (int i = threadIdx.x; i < tot; i += kMrThreads) {
        const int r = div16(i, mM), n = i - r * M;
        T v = mk<T>(0, 0);
        if (n < N && r < nfft) {
            T c = w[n];
            if (SIGN > 0) c.y = -c.y;
            v = cmul(a[r * S + n], c);
        }
        a[r * S + n] = v;
    }
    __syncthreads();
    // Both FFTs forward (one instance of the passes per kernel): the inverse
    // FFT of z is conj(FFT(conj z)), the conjugations folded into the products.
    if constexpr (p2) {
        if (blu_tw2<T>(M)) fft_block<T, -1, V, kMrThreads, true>(a, __builtin_ctz(M), S, twl);
        else fft_block<T, -1, V, kMrThreads, false>(a, __builtin_ctz(M), S, twl);
    } else {
        mr_reg_fft<T, -1, kBluMaxPrime>(a, M, rows, twl);
    }
    for (int i = threadIdx.x; i < tot; i += kMrThreads) {
        const int r = div16(i, mM), k = i - r * M;
        T c = bh[SIGN < 0 ? k : (k == 0 ? 0 : M - k)];
        if (SIGN > 0) c.y = -c.y;
        a[r * S + k] = cconj(cmul(a[r * S + k], c));
    }
    __syncthreads();
    if constexpr (p2) {
        if (blu_tw2<T>(M)) fft_block<T, -1, V, kMrThreads, true>(a, __builtin_ctz(M), S, twl);
        else fft_block<T, -1, V, kMrThreads, false>(a, __builtin_ctz(M), S, twl);
    } else {
        mr_reg_fft<T, -1, kBluMaxPrime>(a, M, rows, twl);
    }
    const unsigned long long mN = magic16(N);
    for (int i = threadIdx.x; i < nfft * N; i += kMrThreads) {
        const int r = div16(i, mN), k = i - r * N;
        T c = w[k];
        if (SIGN > 0) c.y = -c.y;
        a[r * S + k] = cmul(cconj(a[r * S + k]), c);
    }
    __syncthreads();
}

// Row stride of the per-field kernels' LDS rows: N, or blu_stride(M) for the
// Bluestein instance (whose logN argument holds M).
template <int MODE>
__device__ __forceinline__ int row_stride(int N, int logN) {
    return MODE >= 2 ? blu_stride(logN) : N;
}
// Rows of the block's LDS buffer: rpb, or blu_rows(M) for Bluestein (which
// transforms all of them).
template <typename T, int MODE>
__device__ __forceinline__ int lds_rows(int rpb, int logN) {
    return MODE >= 2 ? blu_rows<T>(logN) : rpb;
}
// LDS position of the flat index i = row * N + column.
template <int MODE>
__device__ __forceinline__ int lds_row(int i, int N, int RS, unsigned long long mN) {
    if constexpr (MODE >= 2) {
        const int r = div16(i, mN);
        return r * RS + (i - r * N);
    } else {
        return i;
    }
}
// Stage the FFT twiddles in LDS at b, past the rows the in-place instances
// (MODE >= 1) transform: N values, M for Bluestein, or the two-level table of
// twiddle() (hi then lo) where blu_tw2.
template <typename T, int MODE>
__device__ __forceinline__ const T* stage_fft_twiddles(T* b, int N, int logN, const T* __restrict__ tw) {
    T* twl = b;
    if (MODE == 2 && blu_tw2<T>(logN)) {
        const int M = logN, lgM = __builtin_ctz(M), L = tw2_bits(lgM), nhi = M >> L;
        for (int i = threadIdx.x; i < nhi + (1 << L); i += blockDim.x) twl[i] = i < nhi ? tw[i << L] : tw[i - nhi];
    } else {
        const int n = MODE >= 2 ? logN : N;
        for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = tw[i];
    }
    return twl;
}

// The FFT stage of the per-field kernels: MODE 0 radix-2 passes between two
// buffers (power-of-two N), 1 the in-place register mixed-radix passes
// (mr_reg_fft), 2 and 3 Bluestein with M a power of two and M of small
// primes (bluestein_rows; logN holds M, tw the LDS copy of the M twiddles,
// gtab the global tables).  Returns the buffer holding the result.
template <typename T, int SIGN, int MODE>
__device__ __forceinline__ T* fft_rows(T* a, T* b, int N, int logN, int pow2, int nfft, const T* __restrict__ tw,
                                       const T* __restrict__ gtab) {
    if constexpr (MODE >= 2) {
        bluestein_rows<T, SIGN, MODE == 2>(a, N, logN, nfft, tw, gtab);
        return a;
    } else if constexpr (MODE == 1) {
        mr_reg_fft<T, SIGN>(a, N, nfft, tw);
        return a;
    } else {
        return radix2_fft<T, SIGN>(a, b, N, logN, nfft, tw);
    }
}

// Row of the order-major buffers that holds semi/naive row number q:
// q < cut -> order q; then the negative orders -(cut-1) .. -1 in rows
// N-cut+1 .. N-1 (rows bw+1 .. N-1 when cut = bw).
__device__ __forceinline__ int row_of(int q, int N, int cut) { return q < cut ? q : N - 2 * cut + 1 + q; }

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

// grid (ceil(N/rpb), F): rpb rings of one field per block.
template <typename T, int MODE>
__global__ __launch_bounds__(MODE ? kMrThreads : 256) S2_FFT_OCCUPANCY(MODE) void k_phi_fwd(const double* __restrict__ re, const double* __restrict__ im,
                                                 T* __restrict__ X, int N, int logN, int pow2, int rpb,
                                                 const T* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    const int RS = row_stride<MODE>(N, logN);
    T* b = a + lds_rows<T, MODE>(rpb, logN) * RS;
    constexpr int FB = kFillBatch;
    int bx = blockIdx.x, f = blockIdx.y;
    if constexpr (MODE != 0) xcd_block(bx, f);  // neighbouring ring blocks share 128-byte lines: same L2
    const int j0 = bx * rpb;
    const int nr = min(rpb, N - j0);
    const size_t base = ((size_t)f * N + j0) * N;
    const int tot = nr * N;  // < 2^16 (div16)
    const unsigned long long mN = magic16(N);
    using F = typename Real<T>::type;
    for (int i0 = threadIdx.x; i0 < tot; i0 += FB * blockDim.x) {
        double vr[FB], vi[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                vr[u] = re[base + i];
                vi[u] = im[base + i];
            }
        }
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) a[lds_row<MODE>(i, N, RS, mN)] = mk<T>((F)vr[u], (F)vi[u]);
        }
    }
    const T* twp = tw;
    if constexpr (MODE != 0) twp = stage_fft_twiddles<T, MODE>(b, N, logN, tw);
    __syncthreads();
    T* r = fft_rows<T, -1, MODE>(a, b, N, logN, pow2, nr, twp, tw);
    const unsigned long long mnr = magic16(nr);
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int row = div16(i, mnr), jl = i - row * nr;
        X[((size_t)f * N + row) * N + j0 + jl] = r[jl * RS + row];
    }
}

// Weighted DCT-II of every seminaive order row, first bw outputs kept.
// grid (ceil(nrows/rpb), F).  w_semi holds the even then the odd weights,
// premultiplied by sqrt(2 pi)/N (FFT scaling) and 1/sqrt(2N) (DCT scaling).
template <typename T, int MODE>
__global__ __launch_bounds__(MODE ? kMrThreads : 256) S2_FFT_OCCUPANCY(MODE) void k_dct_fwd(const T* __restrict__ X, T* __restrict__ Cb,
                                                 const double* __restrict__ w_semi, const T* __restrict__ tw,
                                                 const T* __restrict__ ph, int N, int logN, int pow2, int B,
                                                 int cut, int nrows, int rpb, int planar) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    const int RS = row_stride<MODE>(N, logN);
    T* b = a + lds_rows<T, MODE>(rpb, logN) * RS;
    constexpr int FB = kFillBatch;
    int bx = blockIdx.x, f = blockIdx.y;
    if constexpr (MODE != 0) xcd_block(bx, f);
    const int q0 = bx * rpb;
    const int nr = min(rpb, nrows - q0);
    const int tot = nr * N;  // < 2^16 (div16)
    const unsigned long long mN = magic16(N);
    for (int i0 = threadIdx.x; i0 < tot; i0 += FB * blockDim.x) {
        T xv[FB];
        double wv[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                const int rl = div16(i, mN), j = i - rl * N;
                const int r = row_of(q0 + rl, N, cut);
                const int am = r < B ? r : N - r;
                wv[u] = w_semi[(am & 1) * N + j];
                xv[u] = X[((size_t)f * N + r) * N + j];
            }
        }
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                const int rl = div16(i, mN), j = i - rl * N;
                const int p = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
                a[rl * RS + p] = cscale(xv[u], (typename Real<T>::type)wv[u]);
            }
        }
    }
    const T* twp = tw;
    if constexpr (MODE != 0) twp = stage_fft_twiddles<T, MODE>(b, N, logN, tw);
    __syncthreads();
    T* V = fft_rows<T, -1, MODE>(a, b, N, logN, pow2, nr, twp, tw);
    const unsigned long long mB = magic16(B);
    for (int i = threadIdx.x; i < nr * B; i += blockDim.x) {
        const int rl = div16(i, mB), k = i - rl * B;
        const int r = row_of(q0 + rl, N, cut);
        const T* v = V + rl * RS;
        const T e = ph[k];
        T c = cadd(cmul(e, v[k]), cmul(cconj(e), v[k ? N - k : 0]));
        if (k == 0) c = cscale(c, (typename Real<T>::type)0.70710678118654752440);
        if (planar) {  // re and im planes, even cosine indices first (the GEMMs' layout)
            auto* row = reinterpret_cast<typename Real<T>::type*>(Cb) + ((size_t)f * N + r) * 2 * B;
            const int pos = (k & 1) ? ((B + 1) >> 1) + (k >> 1) : (k >> 1);
            row[pos] = c.x;
            row[B + pos] = c.y;
        } else {
            Cb[((size_t)f * N + r) * B + k] = c;
        }
    }
}

// Seminaive projections of orders m < cut.  grid (cut, ceil(F/fc)); block m
// handles rows m and N-m of fc fields.  zonal: write only degree l of order 0
// into zr/zi [F][B] (FZTSemiMemo).
__global__ __launch_bounds__(256) void k_leg_fwd(const double2* __restrict__ Cb, const double* __restrict__ T,
                                                 const long* __restrict__ toff, double* __restrict__ cre,
                                                 double* __restrict__ cim, int N, int B, int F, int fc, int fmt,
                                                 int zonal) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* cs = reinterpret_cast<double2*>(smem);
    const int m = blockIdx.x, f0 = blockIdx.y * fc;
    const int nf = min(fc, F - f0);
    const int nsign = (m > 0 && fmt == 0) ? 2 : 1;
    for (int i = threadIdx.x; i < nf * nsign * B; i += blockDim.x) {
        const int k = i % B, rest = i / B, s = rest % nsign, fl = rest / nsign;
        const int r = s ? N - m : m;
        cs[i] = Cb[((size_t)(f0 + fl) * N + r) * B + k];
    }
    __syncthreads();
    const double* tm = T + toff[m];
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    for (int it = threadIdx.x; it < L * nf * nsign; it += blockDim.x) {
        const int li = it % L, rest = it / L, s = rest % nsign, fl = rest / nsign;
        const int l = m + li;
        const int rs = row_size(m, l);
        const double* tp = tm + table_offset(m, l);
        const double2* c = cs + (fl * nsign + s) * B + (li & 1);
        double ar = 0.0, ai = 0.0;
        for (int t = 0; t < rs; ++t) {
            const double tv = tp[t];
            const double2 cv = c[2 * t];
            ar += cv.x * tv;
            ai += cv.y * tv;
        }
        const int f = f0 + fl;
        if (zonal) {
            cre[(size_t)f * B + l] = ar;
            cim[(size_t)f * B + l] = (fmt == 0) ? ai : 0.0;
        } else if (s == 0) {
            const int p = index_of_coeff(m, l, B);
            cre[f * B2 + p] = ar;
            cim[f * B2 + p] = ai;
            if (fmt == 1 && m > 0) {  // REAL: f(l,-m) = (-1)^m conj f(l,m)
                const int q = index_of_coeff(-m, l, B);
                cre[f * B2 + q] = sg * ar;
                cim[f * B2 + q] = -sg * ai;
            }
        } else {
            const int q = index_of_coeff(-m, l, B);
            cre[f * B2 + q] = sg * ar;
            cim[f * B2 + q] = sg * ai;
        }
    }
}

// Naive projections of orders m >= cut (DLTNaive: even weights for every
// order).  grid (B-cut, ceil(F/fc)).  w_naive = even weights * sqrt(2pi)/N.
__global__ __launch_bounds__(256) void k_naive_fwd(const double2* __restrict__ X, const double* __restrict__ P,
                                                   const long* __restrict__ toff,
                                                   const double* __restrict__ w_naive, double* __restrict__ cre,
                                                   double* __restrict__ cim, int N, int B, int cut, int F, int fc,
                                                   int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* xs = reinterpret_cast<double2*>(smem);
    const int m = cut + blockIdx.x, f0 = blockIdx.y * fc;
    const int nf = min(fc, F - f0);
    const int nsign = (m > 0 && fmt == 0) ? 2 : 1;
    for (int i = threadIdx.x; i < nf * nsign * N; i += blockDim.x) {
        const int j = i % N, rest = i / N, s = rest % nsign, fl = rest / nsign;
        const int r = s ? N - m : m;
        xs[i] = cscale(X[((size_t)(f0 + fl) * N + r) * N + j], w_naive[j]);
    }
    __syncthreads();
    const double* pm = P + toff[m];
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    for (int it = threadIdx.x; it < L * nf * nsign; it += blockDim.x) {
        const int li = it % L, rest = it / L, s = rest % nsign, fl = rest / nsign;
        const int l = m + li;
        const double* pr = pm + (size_t)li * N;
        const double2* x = xs + (fl * nsign + s) * N;
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < N; ++j) {
            ar += x[j].x * pr[j];
            ai += x[j].y * pr[j];
        }
        const int f = f0 + fl;
        if (s == 0) {
            const int p = index_of_coeff(m, l, B);
            cre[f * B2 + p] = ar;
            cim[f * B2 + p] = ai;
            if (fmt == 1) {
                const int q = index_of_coeff(-m, l, B);
                cre[f * B2 + q] = sg * ar;
                cim[f * B2 + q] = -sg * ai;
            }
        } else {
            const int q = index_of_coeff(-m, l, B);
            cre[f * B2 + q] = sg * ar;
            cim[f * B2 + q] = sg * ai;
        }
    }
}

// ---------------------------------------------------------------------------
// inverse
// ---------------------------------------------------------------------------

// Cosine coefficients of every seminaive order from the transposed tables.
// grid (cut, ceil(F/fc)).  Output rows of Cb as the forward uses them.
// ---- one order at a time: DLTSemi / InvDLTSemi / DLTNaive / InvDLTNaive ----
// The reference's per-order Legendre transforms (seminaive.c, naive.c) as
// single-block FP64 kernels.  The cosine transforms are the REDFT10 / REDFT01
// definitions summed directly over the bw cosine indices that are nonzero,
// with cos(pi q / 2N) from an LDS table of the first quadrant.
__device__ __forceinline__ double cos_q(const double* c1, long q, int N) {  // cos(pi q / (2N))
    long r = q % (4L * N);
    if (r > 2L * N) r = 4L * N - r;
    return (r > N) ? -c1[2 * N - r] : c1[r];
}

// DLTSemi (seminaive.c:153-198): weights, DCT-II, projections on the
// zero-stripped cosine table of order m.  LDS: (N+1) + N + (bw+2) doubles.
__global__ __launch_bounds__(256) void k_dlt_semi_one(const double* __restrict__ data, const double* __restrict__ w,
                                                      const double* __restrict__ tab, double* __restrict__ out,
                                                      int B, int m) {
    extern __shared__ double sh[];
    const int N = 2 * B;
    double* c1 = sh;
    double* wd = c1 + N + 1;
    double* cd = wd + N;
    for (int r = threadIdx.x; r <= N; r += blockDim.x) c1[r] = cospi((double)r / (2.0 * N));
    for (int i = threadIdx.x; i < N; i += blockDim.x) wd[i] = data[i] * w[(m & 1) ? N + i : i];
    __syncthreads();
    const int nk = min(N, B + 2);
    const double nrm = 1.0 / sqrt(2.0 * N);
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
        double v = 0.0;
        for (int j = 0; j < N; ++j) v += wd[j] * cos_q(c1, (long)k * (2 * j + 1), N);
        v *= 2.0 * nrm;
        cd[k] = (k == 0) ? v * 0.70710678118654752440 : v;
    }
    __syncthreads();
    for (int i = m + threadIdx.x; i < B; i += blockDim.x) {
        const int toggle = (i - m) & 1;
        const double* p = tab + table_offset(m, i);
        double v = 0.0;
        for (int j = 0; j < i / 2; ++j) v += cd[2 * j + toggle] * p[j];
        if (!((i - m) & 1) || !(m & 1)) v += cd[2 * (i / 2) + toggle] * p[i / 2];
        out[i - m] = v;
    }
}

// InvDLTSemi (seminaive.c:56-115): transposed projections to the cosine
// series, DCT-III, times sin(theta) for odd m.  LDS: (N+1) + bw doubles.
__global__ __launch_bounds__(256) void k_inv_dlt_semi_one(const double* __restrict__ coeffs,
                                                          const double* __restrict__ ttab,
                                                          const double* __restrict__ sinv, double* __restrict__ out,
                                                          int B, int m) {
    extern __shared__ double sh[];
    const int N = 2 * B;
    double* c1 = sh;
    double* fc = c1 + N + 1;
    for (int r = threadIdx.x; r <= N; r += blockDim.x) c1[r] = cospi((double)r / (2.0 * N));
    const double coeff = 0.5 / sqrt((double)B);
    for (int i = threadIdx.x; i < B; i += blockDim.x) {
        double v = 0.0;
        if (!(i == B - 1 && (m & 1))) {
            long off = 0;
            for (int r = 0; r < i; ++r) off += transpose_row_size(r, m, B);
            const double* a = coeffs + ((i > m) ? (i - m) + (m & 1) : (i & 1));
            const int rs = transpose_row_size(i, m, B);
            for (int j = 0; j < rs; ++j) v += a[2 * j] * ttab[off + j];
            v = (i == 0) ? v / sqrt((double)N) : v * coeff;
        }
        fc[i] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        double v = 0.0;
        for (int j = 1; j < B; ++j) v += fc[j] * cos_q(c1, (long)j * (2 * k + 1), N);
        v = fc[0] + 2.0 * v;
        out[k] = (m & 1) ? v * sinv[k] : v;
    }
}

// DLTNaive (naive.c:34-59): weighted samples against the (bw-m) x N sampled
// table.  InvDLTNaive (naive.c:78-95): the table's rows summed with the
// coefficients.
__global__ __launch_bounds__(256) void k_dlt_naive_one(const double* __restrict__ data, const double* __restrict__ w,
                                                       const double* __restrict__ pml, double* __restrict__ out,
                                                       int B, int m) {
    const int N = 2 * B;
    for (int i = threadIdx.x; i < B - m; i += blockDim.x) {
        const double* row = pml + (size_t)i * N;
        double v = 0.0;
        for (int j = 0; j < N; ++j) v += data[j] * w[j] * row[j];
        out[i] = v;
    }
}

__global__ __launch_bounds__(256) void k_inv_dlt_naive_one(const double* __restrict__ coeffs,
                                                           const double* __restrict__ pml, double* __restrict__ out,
                                                           int B, int m) {
    const int N = 2 * B;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        double v = 0.0;
        for (int i = 0; i < B - m; ++i) v += coeffs[i] * pml[(size_t)i * N + j];
        out[j] = v;
    }
}

void dlt_semi_one(const double* data, const double* w, const double* tab, double* out, int B, int m, hipStream_t s) {
    const int N = 2 * B;
    k_dlt_semi_one<<<1, 256, sizeof(double) * ((N + 1) + N + (B + 2)), s>>>(data, w, tab, out, B, m);  // <= 41 KiB
    S2_HIP(hipGetLastError());
}
void inv_dlt_semi_one(const double* coeffs, const double* ttab, const double* sinv, double* out, int B, int m,
                      hipStream_t s) {
    const int N = 2 * B;
    k_inv_dlt_semi_one<<<1, 256, sizeof(double) * ((N + 1) + B), s>>>(coeffs, ttab, sinv, out, B, m);
    S2_HIP(hipGetLastError());
}
void dlt_naive_one(const double* data, const double* w, const double* pml, double* out, int B, int m,
                   hipStream_t s) {
    k_dlt_naive_one<<<1, 256, 0, s>>>(data, w, pml, out, B, m);
    S2_HIP(hipGetLastError());
}
void inv_dlt_naive_one(const double* coeffs, const double* pml, double* out, int B, int m, hipStream_t s) {
    k_inv_dlt_naive_one<<<1, 256, 0, s>>>(coeffs, pml, out, B, m);
    S2_HIP(hipGetLastError());
}

// ---- projections for a few fields: the tables streamed as coalesced GEMVs ----
// With one to four fields a call reads every table entry once and uses it
// for at most 16 products, so the projections are bound by the table stream.
// LPR = 16 or 32 lanes take one table row (16 or 32 contiguous 16-byte
// loads instead of one row per lane; every load of a row issued at once), a
// block takes rpb rows of one order, and the order's data (DCT output, or
// coefficients) sits in LDS.

typedef double dpair8 __attribute__((ext_vector_type(2))) __attribute__((aligned(8)));

// A table row of rs entries over LPR lanes: lane j holds entries
// 2*LPR*it + 2j and 2*LPR*it + 2j + 1 (zero past the row), all MAXI loads
// issued at once.
template <int MAXI, int LPR>
__device__ __forceinline__ void row_load(const double* __restrict__ tp, int rs, int j, dpair8 (&tv)[MAXI]) {
#pragma unroll
    for (int it = 0; it < MAXI; ++it) {
        const int t = 2 * LPR * it + 2 * j;
        if (t + 1 < rs)
            tv[it] = *reinterpret_cast<const dpair8*>(tp + t);
        else if (t < rs)
            tv[it] = dpair8{tp[t], 0.0};
        else
            tv[it] = dpair8{0.0, 0.0};
    }
}

// Sum over the LPR lanes (16 or 32) that share a table row.
template <int LPR>
__device__ __forceinline__ double sum_row(double v) {
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPR);
    return v;
}

// Forward (DLTSemi's projections, seminaive.c:183-197) for orders m < cutoff:
// row l of order m against the order's cosine coefficients Cb[f][+-m][k].
// NTH threads: 512 from bw = 512 on (twice the row groups per block for the
// same data fill: bw=512, one field, 59 -> 53 us; the inverse lost, 65 -> 72).
template <int NF, int MAXI, int LPR, int NTH>
__global__ __launch_bounds__(NTH) void k_leg_fwd_rows(const double2* __restrict__ Cb, const double* __restrict__ T,
                                                      const long* __restrict__ toff, double* __restrict__ cre,
                                                      double* __restrict__ cim, int N, int B, int F, int fmt, int rpb,
                                                      int planar) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* cs = reinterpret_cast<double2*>(smem);
    const int m = blockIdx.y, f0 = blockIdx.z * NF;
    const int L = B - m;
    const int l0 = blockIdx.x * rpb;  // first row (degree offset l - m) of this block
    if (l0 >= L) return;
    const int nf = min(NF, F - f0);
    const int nsign = (m > 0 && fmt == 0) ? 2 : 1;
    for (int i = threadIdx.x; i < nf * nsign * B; i += blockDim.x) {
        const int k = i % B, rest = i / B, s = rest % nsign, fl = rest / nsign;
        const size_t row = ((size_t)(f0 + fl) * N + (s ? N - m : m)) * B;
        if (planar) {  // re then im plane, even then odd k (the register DCT's rows)
            const double* pr = reinterpret_cast<const double*>(Cb + row);
            const int pos = cpos(k, B >> 1);
            cs[i] = make_double2(pr[pos], pr[B + pos]);
        } else {
            cs[i] = Cb[row + k];
        }
    }
    __syncthreads();
    const double* tm = T + toff[m];
    const int g = threadIdx.x / LPR, j = threadIdx.x % LPR;
    const size_t B2 = (size_t)B * B;
    const double sg = (m & 1) ? -1.0 : 1.0;
    const int lend = min(L, l0 + rpb);
    for (int li = l0 + g; li < lend; li += NTH / LPR) {
        const int l = m + li;
        const int rs = row_size(m, l);
        const double* tp = tm + table_offset(m, l);
        const double2* c = cs + (li & 1);
        double ar[NF][2], ai[NF][2];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int s = 0; s < 2; ++s) ar[f][s] = ai[f][s] = 0.0;
        dpair8 tv[MAXI];
        row_load<MAXI, LPR>(tp, rs, j, tv);
#pragma unroll
        for (int it = 0; it < MAXI; ++it)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int t = 2 * LPR * it + 2 * j + e;
                if (t < rs) {
#pragma unroll
                    for (int f = 0; f < NF; ++f)
#pragma unroll
                        for (int s = 0; s < 2; ++s)
                            if (f < nf && s < nsign) {
                                const double2 cv = c[(f * nsign + s) * B + 2 * t];
                                ar[f][s] += cv.x * tv[it][e];
                                ai[f][s] += cv.y * tv[it][e];
                            }
                }
            }
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                ar[f][s] = sum_row<LPR>(ar[f][s]);
                ai[f][s] = sum_row<LPR>(ai[f][s]);
            }
        if (j != 0) continue;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            if (f >= nf) break;
            const size_t fb = (size_t)(f0 + f) * B2;
            const int p = index_of_coeff(m, l, B);
            cre[fb + p] = ar[f][0];
            cim[fb + p] = ai[f][0];
            if (nsign == 2) {
                const int q = index_of_coeff(-m, l, B);
                cre[fb + q] = sg * ar[f][1];
                cim[fb + q] = sg * ai[f][1];
            } else if (fmt == 1 && m > 0) {  // REAL: f(l,-m) = (-1)^m conj f(l,m)
                const int q = index_of_coeff(-m, l, B);
                cre[fb + q] = sg * ar[f][0];
                cim[fb + q] = -sg * ai[f][0];
            }
        }
    }
}

// Inverse (InvDLTSemi's transposed projections, seminaive.c:74-95) for orders
// m < cutoff: row i of order m's transposed table against the coefficients
// m+first, m+first+2, ...; trow[m][i] is where row i starts in the order.
template <int NF, int MAXI, int LPR>
__global__ __launch_bounds__(256) void k_leg_inv_rows(const double* __restrict__ cre, const double* __restrict__ cim,
                                                      const double* __restrict__ TT, const long* __restrict__ toff,
                                                      const int* __restrict__ trow, double2* __restrict__ Cb, int N,
                                                      int B, int F, int fmt, int rpb, int planar) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* co = reinterpret_cast<double2*>(smem);
    const int m = blockIdx.y, f0 = blockIdx.z * NF;
    const int i0 = blockIdx.x * rpb;
    if (i0 >= B) return;
    const int nf = min(NF, F - f0);
    const int nsign = (m > 0 && fmt == 0) ? 2 : 1;
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    for (int i = threadIdx.x; i < nf * nsign * L; i += blockDim.x) {
        const int li = i % L, rest = i / L, s = rest % nsign, fl = rest / nsign;
        const int p = index_of_coeff(s ? -m : m, m + li, B);
        co[i] = make_double2(cre[(f0 + fl) * B2 + p], cim[(f0 + fl) * B2 + p]);
    }
    // The block's row offsets trow[m][i0 .. iend] go to LDS with the data, so
    // that a row's table loads wait on no global load of their own (one
    // field: B=512 67 -> 62 us, B=1024 466 -> 435 us; bit-identical).
    const int iend = min(B, i0 + rpb);
    int* rps = reinterpret_cast<int*>(co + 2 * B * NF);
    {
        const int* rp = trow + (size_t)m * (B + 1);
        for (int t = threadIdx.x; t <= iend - i0; t += 256) rps[t] = rp[i0 + t];
    }
    __syncthreads();
    const double* tt = TT + toff[m];
    const double scale = 0.5 / sqrt((double)B);
    const double scale0 = 1.0 / sqrt((double)N);
    const int g = threadIdx.x / LPR, j = threadIdx.x % LPR;
    for (int i = i0 + g; i < iend; i += 256 / LPR) {
        double vr[NF][2], vi[NF][2];
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int s = 0; s < 2; ++s) vr[f][s] = vi[f][s] = 0.0;
        if (!(i == B - 1 && (m & 1))) {
            const int first = transpose_first(i, m);
            const int r0 = rps[i - i0], rs = rps[i - i0 + 1] - r0;
            const double* tp = tt + r0;
            const double2* a = co + first;
            dpair8 tv[MAXI];
            row_load<MAXI, LPR>(tp, rs, j, tv);
#pragma unroll
            for (int it = 0; it < MAXI; ++it)
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int k = 2 * LPR * it + 2 * j + e;
                    if (k < rs) {
#pragma unroll
                        for (int f = 0; f < NF; ++f)
#pragma unroll
                            for (int s = 0; s < 2; ++s)
                                if (f < nf && s < nsign) {
                                    const double2 av = a[(f * nsign + s) * L + 2 * k];
                                    vr[f][s] += av.x * tv[it][e];
                                    vi[f][s] += av.y * tv[it][e];
                                }
                    }
                }
            const double sc = (i == 0) ? scale0 : scale;
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    vr[f][s] = sum_row<LPR>(vr[f][s]) * sc;
                    vi[f][s] = sum_row<LPR>(vi[f][s]) * sc;
                }
        }
        if (j != 0) continue;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (f < nf && s < nsign) {
                    const size_t row = ((size_t)(f0 + f) * N + (s ? N - m : m)) * B;
                    if (planar) {  // re then im plane, even then odd index (the register DCT's rows)
                        double* pr = reinterpret_cast<double*>(Cb + row);
                        const int pos = cpos(i, B >> 1);
                        pr[pos] = vr[f][s];
                        pr[B + pos] = vi[f][s];
                    } else {
                        Cb[row + i] = make_double2(vr[f][s], vi[f][s]);
                    }
                }
    }
}

__global__ __launch_bounds__(256) void k_leg_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                 const double* __restrict__ TT, const long* __restrict__ toff,
                                                 double2* __restrict__ Cb, int N, int B, int F, int fc, int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int m = blockIdx.x, f0 = blockIdx.y * fc;
    const int nf = min(fc, F - f0);
    const int nsign = (m > 0 && fmt == 0) ? 2 : 1;
    const int L = B - m;
    int* rowpos = reinterpret_cast<int*>(smem);                       // B+1 ints
    double2* co = reinterpret_cast<double2*>(smem + ((B + 2) * sizeof(int) + 15) / 16 * 16);
    const size_t B2 = (size_t)B * B;
    for (int i = threadIdx.x; i < nf * nsign * L; i += blockDim.x) {
        const int li = i % L, rest = i / L, s = rest % nsign, fl = rest / nsign;
        const int p = index_of_coeff(s ? -m : m, m + li, B);
        co[i] = make_double2(cre[(f0 + fl) * B2 + p], cim[(f0 + fl) * B2 + p]);
    }
    for (int i = threadIdx.x; i < B; i += blockDim.x) rowpos[i + 1] = transpose_row_size(i, m, B);
    __syncthreads();
    if (threadIdx.x == 0) {
        rowpos[0] = 0;
        for (int i = 1; i <= B; ++i) rowpos[i] += rowpos[i - 1];
    }
    __syncthreads();
    const double* tt = TT + toff[m];
    const double scale = 0.5 / sqrt((double)B);
    const double scale0 = 1.0 / sqrt((double)N);
    for (int it = threadIdx.x; it < B * nf * nsign; it += blockDim.x) {
        const int i = it % B, rest = it / B, s = rest % nsign, fl = rest / nsign;
        double vr = 0.0, vi = 0.0;
        if (!(i == B - 1 && (m & 1))) {
            const int first = transpose_first(i, m);
            const int rs = rowpos[i + 1] - rowpos[i];
            const double* tp = tt + rowpos[i];
            const double2* a = co + (fl * nsign + s) * L + first;
            for (int k = 0; k < rs; ++k) {
                const double tv = tp[k];
                const double2 av = a[2 * k];
                vr += av.x * tv;
                vi += av.y * tv;
            }
            const double sc = (i == 0) ? scale0 : scale;
            vr *= sc;
            vi *= sc;
        }
        const int r = s ? N - m : m;
        Cb[((size_t)(f0 + fl) * N + r) * B + i] = make_double2(vr, vi);
    }
}

// DCT-III of every seminaive order row (inputs 0..B-1, zero above), then
// x sin(theta) for odd orders, x (-1) on odd negative rows, x 1/sqrt(2pi).
// REAL mirrors row N-m = conj(row m).  grid (ceil(nrows/rpb), F).
template <typename T, int MODE>
__global__ __launch_bounds__(MODE ? kMrThreads : 256) S2_FFT_OCCUPANCY(MODE) void k_dct_inv(const T* __restrict__ Cb, T* __restrict__ X,
                                                 const T* __restrict__ tw, const T* __restrict__ ph,
                                                 const double* __restrict__ sinv, int N, int logN, int pow2, int B,
                                                 int cut, int nrows, int rpb, int fmt, int planar) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    const int RS = row_stride<MODE>(N, logN);
    T* b = a + lds_rows<T, MODE>(rpb, logN) * RS;
    constexpr int FB = kFillBatch;
    int bx = blockIdx.x, f = blockIdx.y;
    if constexpr (MODE != 0) xcd_block(bx, f);
    const int q0 = bx * rpb;
    const int nr = min(rpb, nrows - q0);
    using R = typename Real<T>::type;
    // V[k] = e^{+i pi k/2N} (X[k] - i X[N-k]), X[N] = 0, X[k >= B] = 0: one
    // of the two terms is nonzero, X[k] for k < B, X[N-k] for k > B.
    const int tot = nr * N;  // < 2^16 (div16)
    const unsigned long long mN = magic16(N);
    for (int i0 = threadIdx.x; i0 < tot; i0 += FB * blockDim.x) {
        T xv[FB], e[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                const int rl = div16(i, mN), k = i - rl * N;
                const int r = row_of(q0 + rl, N, cut);
                const int ks = k < B ? k : N - k;
                if (k == B) {
                    xv[u] = mk<T>(0, 0);
                } else if (planar) {  // re and im planes, even cosine indices first
                    const R* row = reinterpret_cast<const R*>(Cb) + ((size_t)f * N + r) * 2 * B;
                    const int pos = (ks & 1) ? ((B + 1) >> 1) + (ks >> 1) : (ks >> 1);
                    xv[u] = mk<T>(row[pos], row[B + pos]);
                } else {
                    xv[u] = Cb[((size_t)f * N + r) * B + ks];
                }
                e[u] = ph[k];
            }
        }
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                const int rl = div16(i, mN), k = i - rl * N;
                const T x = xv[u];
                a[rl * RS + k] = cmul(cconj(e[u]), (k <= B) ? x : mk<T>(x.y, -x.x));  // k = B: x is zero
            }
        }
    }
    const T* twp = tw;
    if constexpr (MODE != 0) twp = stage_fft_twiddles<T, MODE>(b, N, logN, tw);
    __syncthreads();
    T* u = fft_rows<T, +1, MODE>(a, b, N, logN, pow2, nr, twp, tw);
    const R inv_s2pi = (R)0.39894228040143267794;  // 1/sqrt(2 pi)
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int rl = div16(i, mN), j = i - rl * N;
        const int r = row_of(q0 + rl, N, cut);
        const int am = r < B ? r : N - r;
        const int n = (j & 1) ? N - 1 - (j >> 1) : (j >> 1);
        R sc = inv_s2pi;
        if (am & 1) sc *= (R)sinv[j];
        if (r > B && (r & 1)) sc = -sc;
        const T y = cscale(u[rl * RS + n], sc);
        X[((size_t)f * N + r) * N + j] = y;
        if (fmt == 1 && r > 0) X[((size_t)f * N + (N - r)) * N + j] = cconj(y);
    }
}

// Dense synthesis of orders m >= cut (InvDLTNaive), with the same row
// signs / scaling / REAL mirror as k_dct_inv.  grid (B-cut, ceil(F/fc)).
__global__ __launch_bounds__(256) void k_naive_inv(const double* __restrict__ cre, const double* __restrict__ cim,
                                                   const double* __restrict__ P, const long* __restrict__ toff,
                                                   double2* __restrict__ X, int N, int B, int cut, int F, int fc,
                                                   int fmt) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* co = reinterpret_cast<double2*>(smem);
    const int m = cut + blockIdx.x, f0 = blockIdx.y * fc;
    const int nf = min(fc, F - f0);
    const int nsign = (m > 0 && fmt == 0) ? 2 : 1;
    const int L = B - m;
    const size_t B2 = (size_t)B * B;
    for (int i = threadIdx.x; i < nf * nsign * L; i += blockDim.x) {
        const int li = i % L, rest = i / L, s = rest % nsign, fl = rest / nsign;
        const int p = index_of_coeff(s ? -m : m, m + li, B);
        co[i] = make_double2(cre[(f0 + fl) * B2 + p], cim[(f0 + fl) * B2 + p]);
    }
    __syncthreads();
    const double* pm = P + toff[m];
    const double inv_s2pi = 0.39894228040143267794;
    for (int it = threadIdx.x; it < N * nf * nsign; it += blockDim.x) {
        const int j = it % N, rest = it / N, s = rest % nsign, fl = rest / nsign;
        const double2* a = co + (fl * nsign + s) * L;
        double yr = 0.0, yi = 0.0;
        for (int li = 0; li < L; ++li) {
            const double pv = pm[(size_t)li * N + j];
            yr += a[li].x * pv;
            yi += a[li].y * pv;
        }
        const int r = s ? N - m : m;
        double sc = inv_s2pi;
        if (r > B && (r & 1)) sc = -sc;
        const double2 y = make_double2(yr * sc, yi * sc);
        const int f = f0 + fl;
        X[((size_t)f * N + r) * N + j] = y;
        if (fmt == 1 && r > 0) X[((size_t)f * N + (N - r)) * N + j] = make_double2(y.x, -y.y);
    }
}

// Inverse ring FFT (+i, unnormalised) per ring; row B (Nyquist) is zero.
// grid (ceil(N/rpb), F).
template <typename T, int MODE>
__global__ __launch_bounds__(MODE ? kMrThreads : 256) S2_FFT_OCCUPANCY(MODE) void k_phi_inv(const T* __restrict__ X, double* __restrict__ re,
                                                 double* __restrict__ im, int N, int logN, int pow2, int B, int rpb,
                                                 const T* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* a = reinterpret_cast<T*>(smem);
    const int RS = row_stride<MODE>(N, logN);
    T* b = a + lds_rows<T, MODE>(rpb, logN) * RS;
    constexpr int FB = kFillBatch;
    int bx = blockIdx.x, f = blockIdx.y;
    if constexpr (MODE != 0) xcd_block(bx, f);  // neighbouring ring blocks share 128-byte lines: same L2
    const int j0 = bx * rpb;
    const int nr = min(rpb, N - j0);
    const int tot = nr * N;  // < 2^16 (div16)
    const unsigned long long mnr = magic16(nr);
    for (int i0 = threadIdx.x; i0 < tot; i0 += FB * blockDim.x) {
        T v[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                const int r = div16(i, mnr), jl = i - r * nr;
                v[u] = (r == B) ? mk<T>(0, 0) : X[((size_t)f * N + r) * N + j0 + jl];
            }
        }
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * blockDim.x;
            if (i < tot) {
                const int r = div16(i, mnr), jl = i - r * nr;
                a[jl * RS + r] = v[u];
            }
        }
    }
    const T* twp = tw;
    if constexpr (MODE != 0) twp = stage_fft_twiddles<T, MODE>(b, N, logN, tw);
    __syncthreads();
    T* y = fft_rows<T, +1, MODE>(a, b, N, logN, pow2, nr, twp, tw);
    const size_t base = ((size_t)f * N + j0) * N;
    const unsigned long long mN = magic16(N);
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const T v = y[lds_row<MODE>(i, N, RS, mN)];
        re[base + i] = (double)v.x;
        im[base + i] = (double)v.y;
    }
}

// m = 0 bin of each ring as a direct sum (FST_semi_memo.c:386-398), into
// row 0 of X.  One wavefront per ring; grid (ceil(N/4), F), block 256.
__global__ __launch_bounds__(256) void k_ring_sum(const double* __restrict__ re, const double* __restrict__ im,
                                                  double2* __restrict__ X, int N) {
    const int f = blockIdx.y;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= N) return;
    const size_t base = ((size_t)f * N + j) * N;
    double sr = 0.0, si = 0.0;
    for (int k = lane; k < N; k += 64) {
        sr += re[base + k];
        si += im[base + k];
    }
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
    }
    if (lane == 0) X[(size_t)f * N * N + j] = make_double2(sr, si);
}

// TransMult (util.c:68-103): filter degree l times signal coefficient, with
// the reference's imaginary part x*v - y*u, times sqrt(4 pi/(2l+1)).
__global__ void k_trans_mult(const double* dr, const double* di,
                             const double* __restrict__ hr, const double* __restrict__ hi, long fstride,
                             double* rr, double* ri, const int* __restrict__ deg, int B,
                             int F) {
    const size_t B2 = (size_t)B * B;
    const size_t total = B2 * F;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t f = i / B2, p = i - f * B2;
        const int l = deg[p];
        const double x = hr[f * fstride + l], y = hi[f * fstride + l];
        const double u = dr[i], v = di[i];
        const double s = sqrt(4.0 * M_PI / (2.0 * l + 1.0));
        rr[i] = (x * u - y * v) * s;
        ri[i] = (x * v - y * u) * s;
    }
}

// ---------------------------------------------------------------------------
// tables and weights (FP64)
// ---------------------------------------------------------------------------

// cosT[k*B + i] = cos(pi (2i+1) k / (2B)), argument reduced mod 2 pi exactly.
__global__ void k_cos_table(double* __restrict__ cosT, int B) {
    const int total = B * B;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const int k = idx / B, i = idx - k * B;
        const long e = ((long)(2 * i + 1) * k) % (4L * B);
        cosT[idx] = cos(M_PI * (double)e / (2.0 * B));
    }
}

// GenerateCosPmlTable (cospml.c:161-242) for orders m0 + blockIdx.x.  One
// thread per Chebyshev node runs the recurrence; each degree's samples go
// through LDS and the parity-selected DCT-II outputs are formed as direct
// sums against cosT.  out + (toff ? toff[m] : 0) receives order m.
__global__ void k_gen_cospml(double* __restrict__ out, const long* __restrict__ toff, const double* __restrict__ cosT,
                             int B, int m0) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* P = reinterpret_cast<double*>(smem);
    const int m = m0 + blockIdx.x;
    double* tab = out + (toff ? toff[m] : 0);
    const double cm = (m == 0) ? 0.0 : pmm_coeff(m);
    const double fudge = 1.0 / sqrt((double)B);
    for (int i = threadIdx.x; i < B; i += blockDim.x) {
        const double th = (2.0 * i + 1.0) * M_PI / (2.0 * B);
        double p = (m == 0) ? 0.70710678118654752440 : cm * pow(sin(th), (double)m);
        if (m & 1) p /= sin(th);
        P[i] = p;
    }
    // each thread keeps the recurrence state of its nodes in registers
    // (B <= 1024 = 4 nodes per thread with 256 threads)
    double cur[4], prev[4], xs[4];
    for (int u = 0; u < 4; ++u) {
        const int i = threadIdx.x + u * blockDim.x;
        cur[u] = 0.0;
        prev[u] = 0.0;
        xs[u] = 0.0;
        if (i < B) xs[u] = cos((2.0 * i + 1.0) * M_PI / (2.0 * B));
    }
    __syncthreads();
    for (int u = 0; u < 4; ++u) {
        const int i = threadIdx.x + u * blockDim.x;
        if (i < B) cur[u] = P[i];
    }
    for (int l = m; l < B; ++l) {
        const int tau = (l - m) & 1, rs = row_size(m, l);
        double* dst = tab + table_offset(m, l);
        for (int t = threadIdx.x; t < rs; t += blockDim.x) {
            const int k = 2 * t + tau;
            const double* ck = cosT + (size_t)k * B;
            double s = 0.0;
            for (int i = 0; i < B; ++i) s += P[i] * ck[i];
            double c = 2.0 * s;
            if (k == 0) c *= 0.70710678118654752440;
            dst[t] = c * fudge;
        }
        __syncthreads();
        if (l + 1 < B) {
            const double an = l2_an(m, l), cn = l2_cn(m, l);
            for (int u = 0; u < 4; ++u) {
                const int i = threadIdx.x + u * blockDim.x;
                if (i < B) {
                    const double nx = an * (cur[u] * xs[u]) + cn * prev[u];
                    prev[u] = cur[u];
                    cur[u] = nx;
                    P[i] = nx;
                }
            }
        }
        __syncthreads();
    }
}

// TransposeCosPmlTable (cospml.c:301-362), restated by what it holds: row i
// lists cosine coefficient i of degrees l0(i), l0(i)+2, ... (the degrees
// InvDLTSemi pairs with it), Transpose_RowSize(i) of them.
__global__ void k_transpose(const double* __restrict__ in, double* __restrict__ outp,
                            const long* __restrict__ toff, int B, int m0) {
    extern __shared__ __align__(16) unsigned char smem[];
    int* rowpos = reinterpret_cast<int*>(smem);
    const int m = m0 + blockIdx.x;
    const long o = toff ? toff[m] : 0;
    const double* t = in + o;
    double* tt = outp + o;
    if (m == B - 1) {  // one degree: the two layouts coincide
        const int n = table_size(m, B);
        for (int i = threadIdx.x; i < n; i += blockDim.x) tt[i] = t[i];
        return;
    }
    for (int i = threadIdx.x; i < B; i += blockDim.x)
        rowpos[i + 1] = (i == B - 1 && (m & 1)) ? 0 : transpose_row_size(i, m, B);
    __syncthreads();
    if (threadIdx.x == 0) {
        rowpos[0] = 0;
        for (int i = 1; i <= B; ++i) rowpos[i] += rowpos[i - 1];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < B; i += blockDim.x) {
        const int rs = rowpos[i + 1] - rowpos[i];
        const int l0 = m + transpose_first(i, m);
        for (int k = 0; k < rs; ++k) tt[rowpos[i] + k] = t[table_offset(m, l0 + 2 * k) + i / 2];
    }
}

// GeneratePmlTable (pml.c:41-79): P_l^m at the 2B grid latitudes, orders
// m0 + blockIdx.x; each thread runs the recurrence at its own nodes.
__global__ void k_gen_pml(double* __restrict__ out, const long* __restrict__ toff, int B, int m0) {
    const int m = m0 + blockIdx.x;
    const int N = 2 * B;
    double* tab = out + (toff ? toff[m] : 0);
    const double cm = (m == 0) ? 0.0 : pmm_coeff(m);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double th = (2.0 * j + 1.0) * M_PI / (2.0 * N);
        const double x = cos((2.0 * j + 1.0) * M_PI / (2.0 * N));
        double cur = (m == 0) ? 0.70710678118654752440 : cm * pow(sin(th), (double)m);
        double prev = 0.0;
        tab[j] = cur;
        for (int l = m; l + 1 < B; ++l) {
            const double nx = l2_an(m, l) * (cur * x) + l2_cn(m, l) * prev;
            prev = cur;
            cur = nx;
            tab[(size_t)(l - m + 1) * N + j] = nx;
        }
    }
}

// GenerateWeightsForDLT (weights.c:32-46).
__global__ void k_weights(double* __restrict__ w, int B) {
    const int N = 2 * B;
    const double c = M_PI / (4.0 * B);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double k = 2.0 * i + 1.0;
        double s = 0.0;
        for (int j = 0; j < B; ++j) s += 1.0 / (2.0 * j + 1.0) * sin(k * (2.0 * j + 1.0) * c);
        s *= 2.0 * sin(k * c) / B;
        w[i] = s;
        w[i + N] = s * sin(k * c);
    }
}

// Prescaled weights: w_semi = w * sqrt(2pi)/N / sqrt(2N) (2N entries),
// w_naive = w[0:N] * sqrt(2pi)/N.
__global__ void k_prescale_weights(const double* __restrict__ w, double* __restrict__ ws, double* __restrict__ wn,
                                   int N) {
    const double nc = sqrt(2.0 * M_PI) / N;
    const double dc = 1.0 / sqrt(2.0 * N);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * N; i += gridDim.x * blockDim.x) {
        ws[i] = w[i] * nc * dc;
        if (i < N) wn[i] = w[i] * nc;
    }
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = (float)in[i];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
const char* stage_name(int st) {
    static const char* names[kStages] = {"phi_fwd", "dct_fwd", "leg_fwd", "naive_fwd", "leg_inv",
                                         "dct_inv", "naive_inv", "phi_inv", "zonal", "mult"};
    return (st >= 0 && st < kStages) ? names[st] : "?";
}

hipEvent_t StageTimer::take() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    S2_HIP(hipEventCreate(&e));
    return e;
}

void StageTimer::collect() {
    for (auto& r : recs) {
        float t = 0.f;
        S2_HIP(hipEventSynchronize(r.b));
        S2_HIP(hipEventElapsedTime(&t, r.a, r.b));
        ms[r.st] += t;
        count[r.st] += 1;
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    recs.clear();
}

void StageTimer::reset() {
    for (int i = 0; i < kStages; ++i) {
        ms[i] = 0.0;
        count[i] = 0;
    }
}

StageTimer::~StageTimer() {
    for (auto& r : recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
}

Span::Span(StageTimer* t_, int st_, hipStream_t s_) : t(t_ && t_->on ? t_ : nullptr), st(st_), s(s_) {
    if (!t) return;
    a = t->take();
    S2_HIP(hipEventRecord(a, s));
}

Span::~Span() {
    if (!t) return;
    try {
        hipEvent_t b = t->take();
        if (hipEventRecord(b, s) == hipSuccess) {
            t->recs.push_back({st, a, b});
            return;
        }
        t->pool.push_back(b);
    } catch (...) {
    }
    t->pool.push_back(a);
}
static int ilog2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return ((1 << l) == n) ? l : -1;
}

// The per-field FFT/DCT kernels come in four instances (MODE of fft_rows):
// 0 for power-of-two N (radix-2 passes, 256 threads, 2048 values per block
// in two buffers), 1 for N whose prime factors are at most kRegMaxPrime
// (in-place register passes, 512 threads, rm_elems<T>() values per block), 2
// and 3 for the other N (Bluestein with a power-of-two or a 7-smooth
// convolution length M, 512 threads, blu_rows(M) rows per block).  The
// in-place instances also hold their twiddles in LDS.  Separate instances
// keep each path's registers out of the others.
template <typename T>
struct FftCfg {
    int mode;    // instance
    int rpb;     // rows (rings or order rows) per block
    int lg;      // the kernels' logN argument: log2 N, or M for Bluestein
    size_t lds;  // dynamic LDS bytes
    const T* tw; // the kernels' tw argument: twiddles, or the Bluestein tables
};

// N has no prime factor above pmax
static bool rm_supported(int N, int pmax = kRegMaxPrime) {
    int rem = N;
    while (rem % 2 == 0) rem /= 2;
    for (int p = 3; p <= pmax; p += 2)
        while (rem % p == 0) rem /= p;
    return rem == 1;
}

template <typename T>
static FftCfg<T> fft_cfg(const DevConsts& c) {
    const int N = c.N;
    if (N > 2048) throw ArgError{"ring length above 2048"};  // block loops index below 2^16
    constexpr bool f64 = sizeof(T) == 16;
    const T* tw = reinterpret_cast<const T*>(f64 ? (const void*)c.tw : (const void*)c.tw32);
    constexpr int kLdsValues = 163840 / (int)sizeof(T);  // gfx950: 160 KiB per workgroup
    if (c.pow2) {
        const int rpb = N >= 2048 ? 1 : 2048 / N;
        return {0, rpb, c.logN, 2 * sizeof(T) * rpb * N, tw};
    }
    if (rm_supported(N)) {
        const int rpb = std::max(1, std::min(rm_elems<T>() / N, (kLdsValues - N) / N));
        return {1, rpb, c.logN, sizeof(T) * ((size_t)rpb * N + N), tw};
    }
    if (c.blu_M == 0) throw ArgError{"no FFT tables for this ring length"};  // consts_create builds them
    const int M = c.blu_M, rpb = blu_rows<T>(M);
    const T* g = reinterpret_cast<const T*>(f64 ? (const void*)c.blu : (const void*)c.blu32);
    const int ntw = blu_tw2<T>(M) ? tw2_values(M) : M;  // LDS twiddles
    return {is_pow2(M) ? 2 : 3, rpb, M, sizeof(T) * ((size_t)rpb * blu_stride(M) + ntw), g};
}

// FFT_INSTANCES(k, T): the four instances of a per-field FFT/DCT kernel.
#define FFT_INSTANCES(k, T) k<T, 0>, k<T, 1>, k<T, 2>, k<T, 3>
template <typename T, typename K0, typename K1, typename K2, typename K3, typename... A>
static void fft_launch(const FftCfg<T>& g, K0 k0, K1 k1, K2 k2, K3 k3, dim3 grid, hipStream_t s, A... args) {
    switch (g.mode) {
        case 0: launch_lds(k0, grid, 256, g.lds, s, args...); break;
        case 1: launch_lds(k1, grid, kMrThreads, g.lds, s, args...); break;
        case 2: launch_lds(k2, grid, kMrThreads, g.lds, s, args...); break;
        default: launch_lds(k3, grid, kMrThreads, g.lds, s, args...); break;
    }
}

// Bluestein tables of ring length N (layout in DevConsts), in long double.
// M: the power of two >= 2N-1 when it is at most 5/4 of the smallest
// length >= 2N-1 whose primes are at most kBluMaxPrime (radix-8 passes beat
// the mixed-radix ones on nearly equal lengths), else that smooth length.
// The chirp's angle is reduced exactly (t^2 mod 2N); FFT_M(b) is a
// recursive mixed-radix transform whose twiddles are evaluated directly.
typedef long double LD;
static const LD kPiL = 3.141592653589793238462643383279502884L;

static void host_fft(std::vector<LD>& re, std::vector<LD>& im) {  // exp(-2 pi i / n), any n
    const int n = (int)re.size();
    if (n == 1) return;
    int p = 2;
    while (n % p) ++p;
    const int m = n / p;
    std::vector<std::vector<LD>> sr(p, std::vector<LD>(m)), si(p, std::vector<LD>(m));
    for (int r = 0; r < p; ++r)
        for (int j = 0; j < m; ++j) {
            sr[r][j] = re[(size_t)j * p + r];
            si[r][j] = im[(size_t)j * p + r];
        }
    for (int r = 0; r < p; ++r) host_fft(sr[r], si[r]);
    for (int k = 0; k < n; ++k) {
        LD ar = 0, ai = 0;
        for (int r = 0; r < p; ++r) {
            const LD a = -2 * kPiL * (LD)(((long long)r * k) % n) / n, cr = cosl(a), ci = sinl(a);
            const LD xr = sr[r][k % m], xi = si[r][k % m];
            ar += xr * cr - xi * ci;
            ai += xr * ci + xi * cr;
        }
        re[k] = ar;
        im[k] = ai;
    }
}

static int bluestein_length(int N) {
    int p2 = 1;
    while (p2 < 2 * N - 1) p2 <<= 1;
    int ms = 2 * N - 1;
    while (!rm_supported(ms, kBluMaxPrime)) ++ms;
    return 4 * p2 <= 5 * ms ? p2 : ms;
}

static void bluestein_tables(int N, int& M, std::vector<double2>& out) {
    M = bluestein_length(N);
    out.assign((size_t)2 * M + N, make_double2(0.0, 0.0));
    for (int t = 0; t < M; ++t) {
        const LD a = 2 * kPiL * t / M;
        out[t] = make_double2((double)cosl(a), (double)-sinl(a));
    }
    std::vector<LD> br(M, 0), bi(M, 0);
    for (int t = 0; t < N; ++t) {
        const long long q = ((long long)t * t) % (2LL * N);
        const LD a = kPiL * (LD)q / N, wr = cosl(a), wi = -sinl(a);
        out[M + t] = make_double2((double)wr, (double)wi);
        br[t] = wr;  // b_t = conj(w_t)
        bi[t] = -wi;
        if (t > 0) {
            br[M - t] = wr;
            bi[M - t] = -wi;
        }
    }
    host_fft(br, bi);
    for (int k = 0; k < M; ++k) out[M + N + k] = make_double2((double)(br[k] / M), (double)(bi[k] / M));
}

void consts_create(DevConsts& c, int bw, const double* h_weights, hipStream_t s) {
    c.bw = bw;
    c.N = 2 * bw;
    const int N = c.N;
    const int lg = ilog2_exact(N);
    c.pow2 = lg >= 0;
    c.logN = lg < 0 ? 0 : lg;
    double* d_w = nullptr;
    S2_HIP(hipMalloc(&d_w, sizeof(double) * 4 * bw));
    if (h_weights)
        S2_HIP(hipMemcpyAsync(d_w, h_weights, sizeof(double) * 4 * bw, hipMemcpyHostToDevice, s));
    else
        k_weights<<<(N + 255) / 256, 256, 0, s>>>(d_w, bw);
    S2_HIP(hipMalloc(&c.w_semi, sizeof(double) * 2 * N));
    S2_HIP(hipMalloc(&c.w_naive, sizeof(double) * N));
    k_prescale_weights<<<(2 * N + 255) / 256, 256, 0, s>>>(d_w, c.w_semi, c.w_naive, N);
    S2_HIP(hipGetLastError());
    std::vector<double2> tw(N), ph(N);
    std::vector<float2> tw32(N), ph32(N);
    std::vector<double> sv(N);
    std::vector<float> sv32(N);
    for (int t = 0; t < N; ++t) {
        const double a = 2.0 * M_PI * (double)t / N;
        tw[t] = make_double2(cos(a), -sin(a));
        const double b = M_PI * (double)t / (2.0 * N);
        ph[t] = make_double2(cos(b), -sin(b));
        tw32[t] = make_float2((float)tw[t].x, (float)tw[t].y);
        ph32[t] = make_float2((float)ph[t].x, (float)ph[t].y);
        sv[t] = sin((2.0 * t + 1.0) * M_PI / (2.0 * N));
        sv32[t] = (float)sv[t];
    }
    std::vector<int> trow((size_t)bw * (bw + 1));
    for (int m = 0; m < bw; ++m) {
        int* r = trow.data() + (size_t)m * (bw + 1);
        r[0] = 0;
        for (int i = 0; i < bw; ++i) r[i + 1] = r[i] + transpose_row_size(i, m, bw);
    }
    S2_HIP(hipMalloc(&c.trow, sizeof(int) * trow.size()));
    S2_HIP(hipMemcpy(c.trow, trow.data(), sizeof(int) * trow.size(), hipMemcpyHostToDevice));
    std::vector<int> deg((size_t)bw * bw);
    for (int m = -(bw - 1); m < bw; ++m)
        for (int l = m < 0 ? -m : m; l < bw; ++l) deg[index_of_coeff(m, l, bw)] = l;
    S2_HIP(hipMalloc(&c.tw, sizeof(double2) * N));
    S2_HIP(hipMalloc(&c.ph, sizeof(double2) * N));
    S2_HIP(hipMalloc(&c.tw32, sizeof(float2) * N));
    S2_HIP(hipMalloc(&c.ph32, sizeof(float2) * N));
    S2_HIP(hipMalloc(&c.sinv, sizeof(double) * N));
    S2_HIP(hipMalloc(&c.sinv32, sizeof(float) * N));
    S2_HIP(hipMalloc(&c.w_semi32, sizeof(float) * 2 * N));
    S2_HIP(hipMalloc(&c.deg, sizeof(int) * deg.size()));
    S2_HIP(hipMemcpyAsync(c.tw, tw.data(), sizeof(double2) * N, hipMemcpyHostToDevice, s));
    S2_HIP(hipMemcpyAsync(c.ph, ph.data(), sizeof(double2) * N, hipMemcpyHostToDevice, s));
    S2_HIP(hipMemcpyAsync(c.tw32, tw32.data(), sizeof(float2) * N, hipMemcpyHostToDevice, s));
    S2_HIP(hipMemcpyAsync(c.ph32, ph32.data(), sizeof(float2) * N, hipMemcpyHostToDevice, s));
    S2_HIP(hipMemcpyAsync(c.sinv, sv.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    S2_HIP(hipMemcpyAsync(c.sinv32, sv32.data(), sizeof(float) * N, hipMemcpyHostToDevice, s));
    k_to_float<<<(2 * N + 255) / 256, 256, 0, s>>>(c.w_semi, c.w_semi32, 2 * N);
    S2_HIP(hipGetLastError());
    S2_HIP(hipMemcpyAsync(c.deg, deg.data(), sizeof(int) * deg.size(), hipMemcpyHostToDevice, s));
    std::vector<double2> blu;
    std::vector<float2> blu32;
    if (!c.pow2 && !rm_supported(N)) {
        bluestein_tables(N, c.blu_M, blu);
        blu32.resize(blu.size());
        for (size_t i = 0; i < blu.size(); ++i) blu32[i] = make_float2((float)blu[i].x, (float)blu[i].y);
        S2_HIP(hipMalloc(&c.blu, sizeof(double2) * blu.size()));
        S2_HIP(hipMalloc(&c.blu32, sizeof(float2) * blu.size()));
        S2_HIP(hipMemcpyAsync(c.blu, blu.data(), sizeof(double2) * blu.size(), hipMemcpyHostToDevice, s));
        S2_HIP(hipMemcpyAsync(c.blu32, blu32.data(), sizeof(float2) * blu.size(), hipMemcpyHostToDevice, s));
    }
    S2_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    S2_HIP(hipFree(d_w));
}

void consts_destroy(DevConsts& c) {
    for (void* p : {(void*)c.w_semi, (void*)c.w_naive, (void*)c.tw, (void*)c.ph, (void*)c.sinv, (void*)c.deg,
                    (void*)c.tw32, (void*)c.ph32, (void*)c.w_semi32, (void*)c.sinv32, (void*)c.trow, (void*)c.blu,
                    (void*)c.blu32})
        if (p) (void)hipFree(p);
    c = DevConsts{};
}

static void check_bw(int bw) {
    if (bw < 2 || bw > 1024) throw ArgError{"bandwidth must be in [2, 1024]"};
}

void gen_weights(int bw, double* d_w, hipStream_t s) {
    check_bw(bw);
    k_weights<<<(2 * bw + 255) / 256, 256, 0, s>>>(d_w, bw);
    S2_HIP(hipGetLastError());
}

static double* make_cos_table(int bw, hipStream_t s) {
    double* cosT = nullptr;
    S2_HIP(hipMalloc(&cosT, sizeof(double) * bw * bw));
    k_cos_table<<<std::min(4096, (bw * bw + 255) / 256), 256, 0, s>>>(cosT, bw);
    S2_HIP(hipGetLastError());
    return cosT;
}

void gen_cospml_one(int bw, int m, double* d_out, hipStream_t s) {
    check_bw(bw);
    double* cosT = make_cos_table(bw, s);
    k_gen_cospml<<<1, 256, sizeof(double) * bw, s>>>(d_out, nullptr, cosT, bw, m);
    S2_HIP(hipGetLastError());
    S2_HIP(hipStreamSynchronize(s));
    S2_HIP(hipFree(cosT));
}

void transpose_cospml_one(int bw, int m, const double* d_in, double* d_out, hipStream_t s) {
    check_bw(bw);
    S2_HIP(hipMemsetAsync(d_out, 0, sizeof(double) * table_size(m, bw), s));
    k_transpose<<<1, 256, sizeof(int) * (bw + 1), s>>>(d_in, d_out, nullptr, bw, m);
    S2_HIP(hipGetLastError());
}

void gen_pml_one(int bw, int m, double* d_out, hipStream_t s) {
    check_bw(bw);
    k_gen_pml<<<1, 256, 0, s>>>(d_out, nullptr, bw, m);
    S2_HIP(hipGetLastError());
}

void tables_generate(DevTables& t, int bw, int cutoff, bool want_fwd, bool want_inv, hipStream_t s) {
    check_bw(bw);
    if (cutoff < 0 || cutoff > bw) throw ArgError{"cutoff must be in [0, bw]"};
    t.bw = bw;
    t.cutoff = cutoff;
    t.h_off.assign(bw, 0);
    long pos = 0;
    for (int m = 0; m < bw; ++m) {
        t.h_off[m] = pos;
        pos += (m < cutoff) ? table_size(m, bw) : 2L * bw * (bw - m);
    }
    t.total = pos;
    S2_HIP(hipMalloc(&t.off, sizeof(long) * bw));
    S2_HIP(hipMemcpyAsync(t.off, t.h_off.data(), sizeof(long) * bw, hipMemcpyHostToDevice, s));
    double* fwd = nullptr;
    S2_HIP(hipMalloc(&fwd, sizeof(double) * pos));
    if (cutoff > 0) {
        double* cosT = make_cos_table(bw, s);
        k_gen_cospml<<<cutoff, 256, sizeof(double) * bw, s>>>(fwd, t.off, cosT, bw, 0);
        S2_HIP(hipGetLastError());
        S2_HIP(hipStreamSynchronize(s));
        S2_HIP(hipFree(cosT));
    }
    if (cutoff < bw) {
        k_gen_pml<<<bw - cutoff, 256, 0, s>>>(fwd, t.off, bw, cutoff);
        S2_HIP(hipGetLastError());
    }
    if (want_inv) {
        S2_HIP(hipMalloc(&t.inv, sizeof(double) * pos));
        // for odd bw the transposed rows do not fill an order's span; keep the tail zero
        S2_HIP(hipMemsetAsync(t.inv, 0, sizeof(double) * pos, s));
        if (cutoff > 0) {
            k_transpose<<<cutoff, 256, sizeof(int) * (bw + 1), s>>>(fwd, t.inv, t.off, bw, 0);
            S2_HIP(hipGetLastError());
        }
        if (cutoff < bw) {
            const long o = t.h_off[cutoff];
            S2_HIP(hipMemcpyAsync(t.inv + o, fwd + o, sizeof(double) * (pos - o), hipMemcpyDeviceToDevice, s));
        }
    }
    if (want_fwd) {
        t.fwd = fwd;
    } else {
        S2_HIP(hipStreamSynchronize(s));
        S2_HIP(hipFree(fwd));
    }
    t.owned = true;
}

void tables_destroy(DevTables& t) {
    if (t.owned) {
        if (t.fwd) (void)hipFree(t.fwd);
        if (t.inv) (void)hipFree(t.inv);
    }
    if (t.off) (void)hipFree(t.off);
    t = DevTables{};
}

static void check_launch_shape(const DevConsts& c) {
    if (c.N > 2048) throw ArgError{"bandwidth above 1024 is not supported"};
}

// Calls of at most kRowsMaxFields fields use the row-GEMV projections, nf
// fields per block (their data must fit 64 KiB of LDS: 2 signs x bw x 16 B each).
constexpr int kRowsMaxFields = 4;
static int rows_fields(int F, int B) {
    const int cap = std::max(1, 65536 / (2 * B * (int)sizeof(double2)));
    const int nf = F <= 1 ? 1 : F <= 2 ? 2 : 4;
    return nf <= cap ? nf : (cap >= 2 ? 2 : 1);
}

// Table rows per block: long enough that a block's table stream dwarfs its
// data fill (measured on MI355X: 64 rows at bw <= 128 up to 512 at bw = 1024).
static int rows_per_block(int B) { return std::max(64, B / 2); }

// Row loads per lane: a row holds at most bw/2 + 1 entries (32 per load round).
// Row-GEMV instance for bw and nf fields: 16 lanes per row, a lane holding up
// to 5, 9 or 17 pair loads of the longest row (bw/2 + 1 entries).  With
// `wide`, bandwidths above 542 take 32 lanes per row instead (9 pairs per
// lane: 112 VGPRs and 4 waves per SIMD against 190 and 2).  Measured at
// bw=1024, one field: forward 455 -> 407 us with 32 lanes, inverse 444 ->
// 497 us (and both slower with 32 lanes at bw=512), so only the forward
// passes `wide`.  pick(NF, MAXI, LPR) returns the kernel.
template <typename L, typename P>
static void rows_launch(int B, int nf, bool wide, L&& launch, P&& pick) {
    using std::integral_constant;
    auto fields = [&](auto mx, auto lp) {
        if (nf == 1) launch(pick(integral_constant<int, 1>{}, mx, lp));
        else if (nf == 2) launch(pick(integral_constant<int, 2>{}, mx, lp));
        else launch(pick(integral_constant<int, 4>{}, mx, lp));
    };
    const int row = B / 2 + 1;
    if (B <= 542 || !wide) {
        if ((row + 31) / 32 <= 5) fields(integral_constant<int, 5>{}, integral_constant<int, 16>{});
        else if ((row + 31) / 32 <= 9) fields(integral_constant<int, 9>{}, integral_constant<int, 16>{});
        else fields(integral_constant<int, 17>{}, integral_constant<int, 16>{});
    } else {
        if ((row + 63) / 64 <= 5) fields(integral_constant<int, 5>{}, integral_constant<int, 32>{});
        else fields(integral_constant<int, 9>{}, integral_constant<int, 32>{});
    }
}

// Power-of-two N in [64, 2048]: the per-field pipeline's ring FFTs run on
// the register kernels of the batched path (radix-8 passes in registers,
// 256-thread blocks of 2048 values, the same order-major X), which beat the
// radix-2 LDS passes at the same block count.
static bool reg_rings(const DevConsts& c) { return c.pow2 && c.N >= 64 && c.N <= 2048; }
// Blocks of those kernels for F fields: one-wave blocks for calls of few
// fields (few_field_threads), else 256 threads; rings per block and LDS.
// Results are bit-identical either way.  Measured on one field, the four
// ring stages (FFTs and DCTs) per fwd+inv pair: bw=64 34.1 -> 27.9 us,
// bw=128 34.4 -> 29.2 us, bw=256 unchanged (37.7 / 38.0 us).
struct RegBlocks {
    int threads, rpb;
    size_t lds;
};
static RegBlocks reg_blocks(int N, int F) {
    const int few = few_field_threads<8>(N, F);
    const int th = few ? few : kThreads;
    const int rpb = 8 * th / N;
    return {th, rpb, sizeof(double2) * ((size_t)rpb * (N + 1) + N)};
}
// kern<NTH>: the instance for the block's threads
template <typename L>
static void reg_launch(const RegBlocks& rb, L&& launch) {
    if (rb.threads == 64) launch(std::integral_constant<int, 64>{});
    else if (rb.threads == 128) launch(std::integral_constant<int, 128>{});
    else launch(std::integral_constant<int, kThreads>{});
}

// FP32 ring FFTs of grids whose ring length is not a power of two (the mixed
// form's first and last stages at such bandwidths), on the register
// mixed-radix or Bluestein instances; same order-major FP32 intermediate X
// as the power-of-two register kernels.
void ring_fft_fwd_f32(const DevConsts& c, const double* re, const double* im, float2* X, int F, hipStream_t s) {
    const int N = c.N;
    const FftCfg<float2> fg = fft_cfg<float2>(c);
    const int rpb = fg.rpb;
    fft_launch(fg, FFT_INSTANCES(k_phi_fwd, float2), dim3((N + rpb - 1) / rpb, F), s,
               re, im, X, N, fg.lg, 0, rpb, fg.tw);
    S2_HIP(hipGetLastError());
}

void ring_fft_inv_f32(const DevConsts& c, const float2* X, double* re, double* im, int F, hipStream_t s) {
    const int N = c.N;
    const FftCfg<float2> fg = fft_cfg<float2>(c);
    const int rpb = fg.rpb;
    fft_launch(fg, FFT_INSTANCES(k_phi_inv, float2), dim3((N + rpb - 1) / rpb, F), s,
               X, re, im, N, fg.lg, 0, c.bw, rpb, fg.tw);
    S2_HIP(hipGetLastError());
}

void forward_fp64(const DevConsts& c, const DevTables& t, const double* re, const double* im, double* cre,
                  double* cim, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm) {
    check_launch_shape(c);
    const int N = c.N, B = c.bw, cut = t.cutoff, f = (int)fmt;
    const FftCfg<double2> fg = fft_cfg<double2>(c);
    const int rpb = fg.rpb;
    const RegBlocks rb = reg_blocks(N, F);
    {
    Span sp(tm, ST_PHI_FWD, s);
    if (reg_rings(c))
        reg_launch(rb, [&](auto nth) {
            launch_lds(k_fphi_fwd<double, 8, decltype(nth)::value>, dim3(N / rb.rpb, F), rb.threads, rb.lds, s, re, im,
                       X, c.logN, c.tw);
        });
    else
        fft_launch(fg, FFT_INSTANCES(k_phi_fwd, double2), dim3((N + rpb - 1) / rpb, F),
                   s, re, im, X, N, fg.lg, c.pow2, rpb, fg.tw);
    S2_HIP(hipGetLastError());
    }
    if (cut > 0) {
        const int nrows = (fmt == Fmt::Complex) ? 2 * cut - 1 : cut;
        // the row GEMVs also read the register DCT's planar rows
        const int planar = (reg_rings(c) && F <= kRowsMaxFields) ? 1 : 0;
        {
        Span sp(tm, ST_DCT_FWD, s);
        if (planar)
            reg_launch(rb, [&](auto nth) {
                launch_lds(k_fdct_fwd<double, 8, decltype(nth)::value>, dim3((nrows + rb.rpb - 1) / rb.rpb, F),
                           rb.threads, rb.lds, s, (const double2*)X, (double*)Cb, c.w_semi, c.logN, cut, nrows, c.tw,
                           c.ph);
            });
        else
            fft_launch(fg, FFT_INSTANCES(k_dct_fwd, double2),
                       dim3((nrows + rpb - 1) / rpb, F), s, X, Cb, c.w_semi, fg.tw, c.ph, N, fg.lg, c.pow2, B, cut, nrows, rpb, 0);
        S2_HIP(hipGetLastError());
        }
        Span sp(tm, ST_LEG_FWD, s);
        if (F <= kRowsMaxFields) {
            const int nf = rows_fields(F, B);
            const int rpb = rows_per_block(B);
            const dim3 grid((B + rpb - 1) / rpb, cut, (F + nf - 1) / nf);
            const size_t lds = sizeof(double2) * 2 * B * nf;
            if (B >= 512)
                rows_launch(B, nf, true, [&](auto kern) { kern<<<grid, 512, lds, s>>>(Cb, t.fwd, t.off, cre, cim, N, B, F, f, rpb, planar); },
                            [](auto nf_, auto mx, auto lp) {
                                return k_leg_fwd_rows<decltype(nf_)::value, decltype(mx)::value, decltype(lp)::value, 512>;
                            });
            else
                rows_launch(B, nf, true, [&](auto kern) { kern<<<grid, 256, lds, s>>>(Cb, t.fwd, t.off, cre, cim, N, B, F, f, rpb, planar); },
                            [](auto nf_, auto mx, auto lp) {
                                return k_leg_fwd_rows<decltype(nf_)::value, decltype(mx)::value, decltype(lp)::value, 256>;
                            });
        } else {
            const int fc = std::max(1, 65536 / (2 * B * (int)sizeof(double2)));
            k_leg_fwd<<<dim3(cut, (F + fc - 1) / fc), 256, sizeof(double2) * 2 * B * std::min(fc, F), s>>>(
                Cb, t.fwd, t.off, cre, cim, N, B, F, fc, f, 0);
        }
        S2_HIP(hipGetLastError());
    }
    if (cut < B) {
        const int fc = std::max(1, 65536 / (2 * N * (int)sizeof(double2)));
        Span sp(tm, ST_NAIVE_FWD, s);
        k_naive_fwd<<<dim3(B - cut, (F + fc - 1) / fc), 256, sizeof(double2) * 2 * N * std::min(fc, F), s>>>(
            X, t.fwd, t.off, c.w_naive, cre, cim, N, B, cut, F, fc, f);
        S2_HIP(hipGetLastError());
    }
}

void inverse_fp64(const DevConsts& c, const DevTables& t, const double* cre, const double* cim, double* re,
                  double* im, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm) {
    check_launch_shape(c);
    const int N = c.N, B = c.bw, cut = t.cutoff, f = (int)fmt;
    const FftCfg<double2> fg = fft_cfg<double2>(c);
    const int rpb = fg.rpb;
    // the register DCT reads the row GEMVs' planar rows
    const int planar = (reg_rings(c) && F <= kRowsMaxFields) ? 1 : 0;
    const RegBlocks rb = reg_blocks(N, F);
    if (cut > 0) {
        const int head = ((B + 2) * (int)sizeof(int) + 15) / 16 * 16;  // row offsets
        const int fc = std::max(1, (65536 - head) / (2 * B * (int)sizeof(double2)));
        const size_t lds = head + sizeof(double2) * 2 * B * std::min(fc, F);
        {
        Span sp(tm, ST_LEG_INV, s);
        if (F <= kRowsMaxFields) {
            const int nf = rows_fields(F, B);
            const int rpb = rows_per_block(B);
            const dim3 grid((B + rpb - 1) / rpb, cut, (F + nf - 1) / nf);
            const size_t lr = sizeof(double2) * 2 * B * nf + sizeof(int) * (rpb + 1);  // data, row offsets
            rows_launch(B, nf, false, [&](auto kern) { launch_lds(kern, grid, 256, lr, s, cre, cim, t.inv, t.off, c.trow, Cb, N, B, F, f, rpb, planar); },
                        [](auto nf_, auto mx, auto lp) {
                            return k_leg_inv_rows<decltype(nf_)::value, decltype(mx)::value, decltype(lp)::value>;
                        });
        } else {
            k_leg_inv<<<dim3(cut, (F + fc - 1) / fc), 256, lds, s>>>(cre, cim, t.inv, t.off, Cb, N, B, F, fc, f);
        }
        S2_HIP(hipGetLastError());
        }
        const int nrows = (fmt == Fmt::Complex) ? 2 * cut - 1 : cut;
        Span sp(tm, ST_DCT_INV, s);
        if (planar)
            reg_launch(rb, [&](auto nth) {
                launch_lds(k_fdct_inv<double, 8, decltype(nth)::value>, dim3((nrows + rb.rpb - 1) / rb.rpb, F),
                           rb.threads, rb.lds, s, (const double*)Cb, X, c.sinv, c.logN, cut, nrows, f, c.tw, c.ph);
            });
        else
            fft_launch(fg, FFT_INSTANCES(k_dct_inv, double2),
                       dim3((nrows + rpb - 1) / rpb, F), s, Cb, X, fg.tw, c.ph, c.sinv, N, fg.lg, c.pow2, B, cut, nrows, rpb, f, 0);
        S2_HIP(hipGetLastError());
    }
    if (cut < B) {
        const int fc = std::max(1, 65536 / (2 * B * (int)sizeof(double2)));
        Span sp(tm, ST_NAIVE_INV, s);
        k_naive_inv<<<dim3(B - cut, (F + fc - 1) / fc), 256, sizeof(double2) * 2 * B * std::min(fc, F), s>>>(
            cre, cim, t.inv, t.off, X, N, B, cut, F, fc, f);
        S2_HIP(hipGetLastError());
    }
    Span sp(tm, ST_PHI_INV, s);
    if (reg_rings(c))
        reg_launch(rb, [&](auto nth) {
            launch_lds(k_fphi_inv<double, 8, decltype(nth)::value>, dim3(N / rb.rpb, F), rb.threads, rb.lds, s,
                       (const double2*)X, re, im, c.logN, c.tw);
        });
    else
        fft_launch(fg, FFT_INSTANCES(k_phi_inv, double2), dim3((N + rpb - 1) / rpb, F),
                   s, X, re, im, N, fg.lg, c.pow2, B, rpb, fg.tw);
    S2_HIP(hipGetLastError());
}

void forward_fp64_gemm(const DevConsts& c, const GemmTables& g, const double* re, const double* im, double* cre,
                       double* cim, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm) {
    check_launch_shape(c);
    const int N = c.N, B = c.bw, cut = g.cut;
    const FftCfg<double2> fg = fft_cfg<double2>(c);
    const int rpb = fg.rpb;
    {
        Span sp(tm, ST_PHI_FWD, s);
        fft_launch(fg, FFT_INSTANCES(k_phi_fwd, double2), dim3((N + rpb - 1) / rpb, F),
                   s, re, im, X, N, fg.lg, c.pow2, rpb, fg.tw);
        S2_HIP(hipGetLastError());
    }
    if (cut > 0) {
        const int nrows = (fmt == Fmt::Complex) ? 2 * cut - 1 : cut;
        Span sp(tm, ST_DCT_FWD, s);
        fft_launch(fg, FFT_INSTANCES(k_dct_fwd, double2),
                       dim3((nrows + rpb - 1) / rpb, F), s, X, Cb, c.w_semi, fg.tw, c.ph, N, fg.lg, c.pow2, B, cut, nrows, rpb, 1);
        S2_HIP(hipGetLastError());
    }
    Span sp(tm, ST_LEG_FWD, s);
    gemm_fwd_launch(c, g, Cb, X, cre, cim, F, fmt, s);
}

void inverse_fp64_gemm(const DevConsts& c, const GemmTables& g, const double* cre, const double* cim, double* re,
                       double* im, int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm) {
    check_launch_shape(c);
    const int N = c.N, B = c.bw, cut = g.cut, f = (int)fmt;
    const FftCfg<double2> fg = fft_cfg<double2>(c);
    const int rpb = fg.rpb;
    {
        Span sp(tm, ST_LEG_INV, s);
        gemm_inv_launch(c, g, cre, cim, Cb, X, F, fmt, s);
    }
    if (cut > 0) {
        const int nrows = (fmt == Fmt::Complex) ? 2 * cut - 1 : cut;
        Span sp(tm, ST_DCT_INV, s);
        fft_launch(fg, FFT_INSTANCES(k_dct_inv, double2),
                       dim3((nrows + rpb - 1) / rpb, F), s, Cb, X, fg.tw, c.ph, c.sinv, N, fg.lg, c.pow2, B, cut, nrows, rpb, f, 1);
        S2_HIP(hipGetLastError());
    }
    Span sp(tm, ST_PHI_INV, s);
    fft_launch(fg, FFT_INSTANCES(k_phi_inv, double2), dim3((N + rpb - 1) / rpb, F), s,
               X, re, im, N, fg.lg, c.pow2, B, rpb, fg.tw);
    S2_HIP(hipGetLastError());
}

void zonal_fp64(const DevConsts& c, const DevTables& t, const double* re, const double* im, double* zr, double* zi,
                int F, Fmt fmt, double2* X, double2* Cb, hipStream_t s, StageTimer* tm) {
    check_launch_shape(c);
    Span sp(tm, ST_ZONAL, s);
    if (t.cutoff < 1) throw ArgError{"the zonal transform needs the order-0 cosine table"};
    const int N = c.N, B = c.bw;
    k_ring_sum<<<dim3((N + 3) / 4, F), 256, 0, s>>>(re, im, X, N);
    S2_HIP(hipGetLastError());
    const FftCfg<double2> fg = fft_cfg<double2>(c);  // one order row (m = 0) per field
    fft_launch(fg, FFT_INSTANCES(k_dct_fwd, double2), dim3(1, F), s, X, Cb, c.w_semi,
               fg.tw, c.ph, N, fg.lg, c.pow2, B, t.cutoff, 1, 1, 0);
    S2_HIP(hipGetLastError());
    const int fc = std::max(1, 65536 / (2 * B * (int)sizeof(double2)));
    k_leg_fwd<<<dim3(1, (F + fc - 1) / fc), 256, sizeof(double2) * 2 * B * std::min(fc, F), s>>>(
        Cb, t.fwd, t.off, zr, zi, N, B, F, fc, (int)fmt, 1);
    S2_HIP(hipGetLastError());
}

void trans_mult(const DevConsts& c, const double* dr, const double* di, const double* hr, const double* hi,
                long fstride, double* rr, double* ri, int F, hipStream_t s, StageTimer* tm) {
    Span sp(tm, ST_MULT, s);
    const size_t total = (size_t)c.bw * c.bw * F;
    const int blocks = (int)std::min<size_t>(8192, (total + 255) / 256);
    k_trans_mult<<<blocks, 256, 0, s>>>(dr, di, hr, hi, fstride, rr, ri, c.deg, c.bw, F);
    S2_HIP(hipGetLastError());
}

}  // namespace s2
