// FP64 seed and three-term-recurrence coefficients of the L2-normalised
// associated Legendre functions, in the reference's operation order:
//   P_m^m = c_m sin^m(theta), c_m from pmm.c:21-34
//   P_{l+1}^m = an(m,l) x P_l^m + cn(m,l) P_{l-1}^m, l2_norms.c:16-38
#pragma once

#include <hip/hip_runtime.h>

namespace s2 {

// L2-normalised seed P_m^m(cos th) (pmm.c:21-34) and recurrence coefficients
// (l2_norms.c:16-38), in the reference's operation order.
__device__ inline double pmm_coeff(int m) {
    double c = sqrt(m + 0.5);
    for (int i = 0; i < m; ++i) c *= sqrt((m - (i / 2.0)) / ((double)m - i));
    if (m) c *= pow(2.0, -m / 2.0);
    if (m & 1) c = -c;
    return c;
}
__device__ __forceinline__ double l2_an(int m, int l) {
    return sqrt(((2.0 * l + 3.0) / (2.0 * l + 1.0)) * ((l - m + 1.0) / (l + m + 1.0))) *
           ((2.0 * l + 1.0) / (l - m + 1.0));
}
__device__ __forceinline__ double l2_cn(int m, int l) {
    if (l == 0) return 0.0;
    return -1.0 *
           sqrt(((2.0 * l + 3.0) / (2.0 * l - 1.0)) * ((l - m + 1.0) / (l + m + 1.0)) *
                (((double)l - m) / ((double)l + m))) *
           ((l + m) / (l - m + 1.0));
}

}  // namespace s2
