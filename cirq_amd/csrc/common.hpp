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

// v + (v moved by one DPP pattern), FP64 as two 32-bit DPP moves; lanes whose source is outside
// the row, or whose row is masked off, add +0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_add(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, ROW_MASK, 0xf, true);
  return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum of v over the 64 lanes of a wave, returned to every lane.  Row prefix sums by DPP row shifts
// (1, 2, 4, 8: lane 15 of each 16-lane row holds its row), then row_bcast:15 and row_bcast:31 fold
// the rows into lane 63, which is read back as a uniform value: register moves only, where a
// butterfly of shuffles would make six dependent round trips through the LDS crossbar.  The
// summation order is fixed, so results are deterministic.
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0x111, 0xf>(v);  // row_shr:1
  v = dpp_add<0x112, 0xf>(v);  // row_shr:2
  v = dpp_add<0x114, 0xf>(v);  // row_shr:4
  v = dpp_add<0x118, 0xf>(v);  // row_shr:8
  v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)x, 63);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

}  // namespace cqs
