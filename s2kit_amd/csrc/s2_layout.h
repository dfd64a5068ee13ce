// Integer layout arithmetic of the memoised seminaive SHT, shared by host
// code and kernels.  These functions fix where callers' tables and
// coefficients live, so they reproduce the reference's layout exactly:
//   RowSize           cospml.c:250-258
//   TableOffset       cospml.c:123-134
//   TableSize         cospml.c:39-59
//   Transpose_RowSize cospml.c:270-288 (including its short count for odd bw)
//   IndexOfHarmonicCoeff util.c:42-49
#pragma once

#include <hip/hip_runtime.h>

namespace s2 {

// Cosine coefficients stored for degree l of order m: the series of P_l^m
// (even m) or P_l^m/sin (odd m) has degree l (resp. l-1) and every other
// index vanishes.
__host__ __device__ inline int row_size(int m, int l) {
    if (l < m) return 0;
    int deg = (m & 1) ? l - 1 : l;
    return deg / 2 + 1;
}

// Start of degree l inside order m's zero-stripped table: the prefix sum of
// row sizes, in closed form; an odd order is the even order m-1 one degree
// lower.
__host__ __device__ inline int table_offset(int m, int l) {
    if (m & 1) {
        m -= 1;
        l -= 1;
    }
    int hl = l / 2, hm = m / 2;
    int off = hl * (hl + 1) - hm * (hm + 1);
    if (l & 1) off += hl + 1;
    return off;
}

// Entries of order m's table: table_offset one degree past the last.
__host__ __device__ inline int table_size(int m, int bw) { return table_offset(m, bw); }

// Entries in row `row` (a cosine index) of order m's transposed table.
__host__ __device__ inline int transpose_row_size(int row, int m, int bw) {
    if (row >= bw) return 0;
    int even_m = m & ~1, r = row;
    if (m & 1) {
        if (row == bw - 1) return 0;
        r = row + 1;  // an odd order reads like the even order m-1, one index up
    }
    int n = (r <= even_m) ? (bw - even_m) / 2 : (bw - r - 1) / 2 + 1;
    if ((m & 1) && row < m) n -= (row & 1);
    return n;
}

// First degree read by row i of the transposed table, as an offset from m.
__host__ __device__ inline int transpose_first(int i, int m) { return (i > m) ? (i - m) + (m & 1) : (i & 1); }

// Degrees l = m+p+2i (i = 0, 1, ...) of order m below bw with l-m of parity p.
__host__ __device__ inline int gm_lp(int m, int p, int bw) { return (bw - m - p > 0) ? (bw - m - p + 1) / 2 : 0; }

__host__ __device__ inline int index_of_coeff(int m, int l, int bw) {
    int L = bw - 1;
    if (m >= 0) return m * (L + 1) - (m * (m - 1)) / 2 + (l - m);
    int am = -m;
    return (L * (L + 3)) / 2 + 1 + ((L + m) * (L + m + 1)) / 2 + (l - am);
}

}  // namespace s2
