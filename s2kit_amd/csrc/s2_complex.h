// Complex helpers on HIP's double2 / float2 vector types.
#pragma once

#include <hip/hip_runtime.h>

namespace s2 {

template <typename T>
struct Real;
template <>
struct Real<double2> {
    using type = double;
};
template <>
struct Real<float2> {
    using type = float;
};

template <typename R>
struct Cplx;
template <>
struct Cplx<double> {
    using type = double2;
};
template <>
struct Cplx<float> {
    using type = float2;
};

template <typename T>
__device__ __forceinline__ T mk(typename Real<T>::type x, typename Real<T>::type y) {
    T r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename T>
__device__ __forceinline__ T cadd(T a, T b) {
    return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ T csub(T a, T b) {
    return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ T cmul(T a, T b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ T cconj(T a) {
    return mk<T>(a.x, -a.y);
}
template <typename T>
__device__ __forceinline__ T cscale(T a, typename Real<T>::type s) {
    return mk<T>(a.x * s, a.y * s);
}

}  // namespace s2
