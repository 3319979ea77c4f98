// Internal helpers shared by the engine's HIP translation units (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/cirq_amd.h"

namespace cqs {

constexpr int WG = 256;  // threads per workgroup of every tile kernel (4 wave64)

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail_hip(hipError_t e, const char* what);

#define CQS_HIP(call)                                        \
  do {                                                       \
    hipError_t _e = (call);                                  \
    if (_e != hipSuccess) return ::cqs::fail_hip(_e, #call); \
  } while (0)

#define CQS_REQUIRE(cond, ...)          \
  do {                                  \
    if (!(cond)) {                      \
      ::cqs::set_error(__VA_ARGS__);    \
      return CQS_ERR_ARG;               \
    }                                   \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- complex64 helpers
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a*b + c
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 c) {
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}
// conj(a)*b + c
__device__ __forceinline__ float2 cfma_conj(float2 a, float2 b, float2 c) {
  return make_float2(fmaf(a.x, b.x, fmaf(a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(-a.y, b.x, c.y)));
}

// Deposit the low bits of x at the given ascending-or-not positions.
__device__ __forceinline__ uint64_t deposit(uint32_t x, const int32_t* pos, int nb) {
  uint64_t r = 0;
  for (int b = 0; b < nb; ++b) r |= (uint64_t)((x >> b) & 1u) << pos[b];
  return r;
}
__device__ __forceinline__ uint64_t deposit64(uint64_t x, const int32_t* pos, int nb) {
  uint64_t r = 0;
  for (int b = 0; b < nb; ++b) r |= ((x >> b) & 1ull) << pos[b];
  return r;
}
// deposit64 for x < 2^MAXB (a workgroup's tile index): unrolled to a compile-time bound so that
// the position loads are issued together instead of one dependent scalar load per bit.
template <int MAXB>
__device__ __forceinline__ uint64_t deposit_bounded(uint64_t x, const int32_t* pos, int nb) {
  uint64_t r = 0;
#pragma unroll
  for (int b = 0; b < MAXB; ++b)
    if (b < nb) r |= ((x >> b) & 1ull) << pos[b];
  return r;
}

// Insert a zero bit at position b of x.
__device__ __forceinline__ int insert0(int x, int b) {
  return ((x >> b) << (b + 1)) | (x & ((1 << b) - 1));
}

// FP64 block reduction of one value per thread (WG threads); result valid in thread 0.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace cqs
