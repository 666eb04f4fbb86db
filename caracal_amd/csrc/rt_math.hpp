// Overloaded device math for the two arithmetic types of the engine.  The
// float overloads call the accurate single-precision library routines (not
// the __expf-style approximations) because the mixed path is held to a 1e-5
// relative force error against FP64.
#pragma once
#include <hip/hip_runtime.h>

namespace cam {
namespace rt {
__device__ __forceinline__ double exp(double x) { return ::exp(x); }
__device__ __forceinline__ float exp(float x) { return ::expf(x); }
__device__ __forceinline__ double log(double x) { return ::log(x); }
__device__ __forceinline__ float log(float x) { return ::logf(x); }
__device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ float sqrt(float x) { return ::sqrtf(x); }
__device__ __forceinline__ double tanh(double x) { return ::tanh(x); }
__device__ __forceinline__ float tanh(float x) { return ::tanhf(x); }
__device__ __forceinline__ double cosh(double x) { return ::cosh(x); }
__device__ __forceinline__ float cosh(float x) { return ::coshf(x); }
__device__ __forceinline__ double acos(double x) { return ::acos(x); }
__device__ __forceinline__ float acos(float x) { return ::acosf(x); }
__device__ __forceinline__ double cos(double x) { return ::cos(x); }
__device__ __forceinline__ double sin(double x) { return ::sin(x); }
}  // namespace rt
}  // namespace cam
