// Shared definitions of the trans_bl_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/trans_bl_amd.h"

#define TBL_HIP(call)                                                   \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) {                                         \
            tbl_set_last_error(hipGetErrorString(e_), __FILE__, __LINE__); \
            return TBL_ERR_HIP;                                         \
        }                                                               \
    } while (0)

void tbl_set_last_error(const char *msg, const char *file, int line);

namespace tbl {

// complex value with real component type T (interleaved re, im)
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a) { return {-a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> operator/(cx<T> a, T s) { return {a.re / s, a.im / s}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// warp-level shuffles of complex values within groups of G lanes
template <typename T>
__device__ __forceinline__ cx<T> shfl_up(cx<T> v, int d, int G)
{
    return {__shfl_up(v.re, d, G), __shfl_up(v.im, d, G)};
}
template <typename T>
__device__ __forceinline__ cx<T> shfl_down(cx<T> v, int d, int G)
{
    return {__shfl_down(v.re, d, G), __shfl_down(v.im, d, G)};
}

// Per-level geometry of the internal layout.  Fields of one level are
// stored mode-major, then x column, then y contiguous with a padded column
// stride ld: index (k*nx + i)*ld + j, i and j 0-based.
struct Level {
    int nx, ny, ld;
    double dx2;       // (dx * 2^(l-1))^2
    size_t plane;     // nx * ld (elements per mode)
};

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

}  // namespace tbl
