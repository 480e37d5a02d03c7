// Raw buffer-resource loads (gfx950).
#pragma once

#include "tbl_common.h"

namespace tbl {

// ---------------------------------------------------------------------------
// raw buffer-resource loads (stride 0, range in bytes, 32-bit offsets from a
// uniform base): an offset past the range reads zero, so a masked load needs
// no branch (and no register copy at a branch join that would wait for it),
// and a constant part of the offset goes into the instruction
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr unsigned kBufOff = 0x80000000u;    // past every range used here
constexpr int kBufFlags = 0x00020000;        // gfx9 buffer resource word 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, size_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), static_cast<short>(0),
                                             static_cast<int>(bytes), kBufFlags);
}
// Same for a base and range the compiler cannot prove wavefront-uniform
// (they must be): read from the first lane, so the resource lives in
// scalar registers instead of being re-formed per load.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc_uniform(const void *p, size_t bytes)
{
    const unsigned long long a = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(a));
    const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(a >> 32));
    const int n = __builtin_amdgcn_readfirstlane(static_cast<int>(bytes));
    void *q = reinterpret_cast<void *>((static_cast<unsigned long long>(hi) << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(q, static_cast<short>(0), n, kBufFlags);
}
__device__ __forceinline__ cx<double> buf_load(__amdgpu_buffer_rsrc_t r, unsigned off, double)
{
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    cx<double> x;
    __builtin_memcpy(&x, &v, sizeof(x));
    return x;
}
__device__ __forceinline__ cx<float> buf_load(__amdgpu_buffer_rsrc_t r, unsigned off, float)
{
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    cx<float> x;
    __builtin_memcpy(&x, &v, sizeof(x));
    return x;
}
// scalar loads; soff is a wavefront-uniform byte offset (a scalar register)
__device__ __forceinline__ double buf_scalar(__amdgpu_buffer_rsrc_t r, unsigned off, double,
                                             int soff = 0)
{
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
    double x;
    __builtin_memcpy(&x, &v, sizeof(x));
    return x;
}
__device__ __forceinline__ float buf_scalar(__amdgpu_buffer_rsrc_t r, unsigned off, float,
                                            int soff = 0)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}

}  // namespace tbl
