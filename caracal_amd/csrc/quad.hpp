// Cross-lane moves inside a quad (4 consecutive lanes) with DPP quad_perm.
// All patterns are compile-time constants; a 64-bit value moves as two
// 32-bit halves.  Every lane of the quad must be active.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cam {
namespace quad {

template <int CTRL>
__device__ __forceinline__ int dpp32(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = dpp32<CTRL>((int)(uint32_t)u);
    const int hi = dpp32<CTRL>((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __builtin_bit_cast(float, dpp32<CTRL>(__builtin_bit_cast(int, v)));
}

// quad_perm control word: lane i of the quad reads lane p_i.
constexpr int perm(int p0, int p1, int p2, int p3) { return p0 | (p1 << 2) | (p2 << 4) | (p3 << 6); }

// Lane i receives the value of lane (i + N) mod 4.
template <int N, typename T>
__device__ __forceinline__ T rot(T v)
{
    if constexpr ((N & 3) == 0) return v;
    else return dpp<perm(N & 3, (1 + N) & 3, (2 + N) & 3, (3 + N) & 3)>(v);
}

// Every lane receives the value of lane S.
template <int S, typename T>
__device__ __forceinline__ T bcast(T v)
{
    return dpp<perm(S, S, S, S)>(v);
}

// Sum over the quad, identical bits on all four lanes.
template <typename T>
__device__ __forceinline__ T sum(T v)
{
    const T a = v + dpp<perm(1, 0, 3, 2)>(v);
    return a + dpp<perm(2, 3, 0, 1)>(a);
}

}  // namespace quad
}  // namespace cam
