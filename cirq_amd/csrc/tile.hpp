// Tile program interpreter: the device side of one pass (see cqs_pass in cirq_amd.h).
//
// A workgroup owns one tile: the 2^L amplitudes whose flat indices differ only in the L
// "local" bit positions lpos[] (ascending).  Tile index j (0..2^L) maps to flat offset
// deposit(j, lpos); the tile number maps to the remaining "outer" positions opos[].
// Thread tid owns elements j = tid + e*WG for the load/store and all elementwise ops, so
// elementwise ops need no barrier between them; dense ops are group-parallel and are
// fenced by barriers.
#pragma once

#include "common.hpp"

namespace cqs {

struct TileProg {
  int32_t n;         // total flat-index bits
  int32_t L;         // local bits
  int32_t nouter;    // n - L
  int32_t op_count;  // ops in this pass
  int32_t max_k;     // largest dense k the LDS matrix stage was sized for
  float mscale;      // amplitude scale of the split-FP16 MFMA path (2^15 for normalised states)
  int32_t lpos[CQS_MAX_LOCAL];
  int32_t opos[CQS_MAX_BITS];
  const cqs_op* ops;
  const float2* mats;
  const int32_t* aux;
};

// LDS layout of a tile kernel: [2^L tile amplitudes][4^max_k staged matrix entries if max_k >= 3]

__device__ __forceinline__ bool outer_controls_ok(const cqs_op& op, uint64_t base) {
  const uint64_t m = ((uint64_t)op.co_mask_hi << 32) | op.co_mask_lo;
  const uint64_t v = ((uint64_t)op.co_val_hi << 32) | op.co_val_lo;
  return (base & m) == v;
}

// Dense 2^K x 2^K matrix on tile bits op.t[0..K); M row-major, bit m of the matrix index is
// tile bit op.t[m].  Each thread owns whole groups, so no barrier is needed inside.
template <int K>
__device__ __forceinline__ void tile_dense(float2* tile, int L, const cqs_op& op, const float2* M) {
  constexpr int D = 1 << K;
  int ts[K];
#pragma unroll
  for (int m = 0; m < K; ++m) ts[m] = op.t[m];
#pragma unroll
  for (int a = 1; a < K; ++a)
#pragma unroll
    for (int b = a; b > 0; --b)
      if (ts[b - 1] > ts[b]) {
        int x = ts[b - 1];
        ts[b - 1] = ts[b];
        ts[b] = x;
      }
  int pat[D];
#pragma unroll
  for (int m = 0; m < D; ++m) {
    int p = 0;
#pragma unroll
    for (int b = 0; b < K; ++b) p |= ((m >> b) & 1) << op.t[b];
    pat[m] = p;
  }
  const int G = (1 << L) >> K;
  const int cm = op.cl_mask, cv = op.cl_val;
  for (int g = threadIdx.x; g < G; g += WG) {
    int j0 = g;
#pragma unroll
    for (int s = 0; s < K; ++s) j0 = insert0(j0, ts[s]);
    if ((j0 & cm) != cv) continue;
    float2 v[D];
#pragma unroll
    for (int m = 0; m < D; ++m) v[m] = tile[j0 + pat[m]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) acc = cfma(M[r * D + c], v[c], acc);
      tile[j0 + pat[r]] = acc;
    }
  }
}

// ---------------------------------------------------------------- fused blocks on MFMA
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

// Y = U X on tile bits op.t[0..K) with v_mfma_f32_32x32x16_f16.  The tile is viewed as a
// (2^K rows) x (2^(L-K) columns) complex matrix; in real-interleaved form (row 2r = Re, 2r+1 = Im)
// it is R (2^(K+1) square) times X.  Each wave owns whole 32-column blocks, so reads and writes of
// different waves never overlap and the block needs no barrier inside.
template <int K>
__device__ __forceinline__ void tile_block_mfma(float2* tile, int L, const cqs_op& op, const half8_t* A,
                                                float scale, float inv_scale) {
  constexpr int D = 1 << K;
  constexpr int R = 2 * D;
  constexpr int KS = R / 16;
  constexpr int RB = R / 32;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, c32 = lane & 31;
  int tmask = 0;
#pragma unroll
  for (int m = 0; m < K; ++m) tmask |= 1 << op.t[m];
  // tile offset of complex row r
  auto rowoff = [&](int r) {
    int j = 0;
#pragma unroll
    for (int m = 0; m < K; ++m) j |= ((r >> m) & 1) << op.t[m];
    return j;
  };
  const int ncol = 1 << (L - K);
  const int ncb = (ncol + 31) >> 5;
  for (int cb = wave; cb < ncb; cb += WG / 64) {
    const int col = cb * 32 + c32;
    const bool valid = col < ncol;
    int jc = 0;
    for (int b = 0, i = 0; b < L; ++b) {
      if (!((tmask >> b) & 1)) {
        jc |= ((col >> i) & 1) << b;
        ++i;
      }
    }
    floatx16_t acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int a = 0; a < 16; ++a) acc[rb][a] = 0.f;
#pragma unroll 1
    for (int ks = 0; ks < KS; ++ks) {
      half8_t bhi, blo;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = ks * 8 + 4 * h + e;
        const float2 v = valid ? tile[jc | rowoff(r)] : make_float2(0.f, 0.f);
        const float xr = v.x * scale, xi = v.y * scale;
        const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
        bhi[2 * e] = hr;
        bhi[2 * e + 1] = hi;
        blo[2 * e] = (_Float16)(xr - (float)hr);
        blo[2 * e + 1] = (_Float16)(xi - (float)hi);
      }
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        // A operands stream from the L1/L2 caches (every workgroup reads the same matrix)
        const half8_t ahi = A[(ks * RB + rb) * 64 + lane];
        const half8_t alo = A[(KS * RB + ks * RB + rb) * 64 + lane];
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bhi, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bhi, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, blo, acc[rb], 0, 0, 0);
      }
    }
    if (valid) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int a = 0; a < 16; a += 2) {
          const int i = rb * 32 + (a & 3) + 8 * (a >> 2) + 4 * h;  // even: real part of row i/2
          tile[jc | rowoff(i >> 1)] = make_float2(acc[rb][a] * inv_scale, acc[rb][a + 1] * inv_scale);
        }
    }
  }
}

// Global flat index of this thread's e-th element.
__device__ __forceinline__ uint64_t elem_index(const TileProg& P, uint64_t base, uint64_t off_lo,
                                               int nlo, int e) {
  return base + off_lo + deposit((uint32_t)e, P.lpos + nlo, P.L - nlo);
}

__device__ __forceinline__ void tile_diag(float2* tile, const TileProg& P, const cqs_op& op,
                                          uint64_t base, uint64_t off_lo, int nlo, int per) {
  const float2* Dg = P.mats + op.mat;
  const int k = op.k;
  for (int e = 0; e < per; ++e) {
    const uint64_t gi = elem_index(P, base, off_lo, nlo, e);
    int idx = 0;
    for (int m = 0; m < k; ++m) idx |= (int)((gi >> op.t[m]) & 1ull) << m;
    const int j = threadIdx.x + e * WG;
    tile[j] = cmul(tile[j], Dg[idx]);
  }
}

__device__ __forceinline__ void tile_pairphase(float2* tile, const TileProg& P, const cqs_op& op,
                                               uint64_t base, uint64_t off_lo, int nlo, int per) {
  const float2* ph = P.mats + op.mat;
  const int32_t* pr = P.aux + op.t[0];
  const int np = op.k;
  for (int e = 0; e < per; ++e) {
    const uint64_t gi = elem_index(P, base, off_lo, nlo, e);
    const int j = threadIdx.x + e * WG;
    float2 v = tile[j];
    for (int p = 0; p < np; ++p) {
      if ((gi >> pr[2 * p]) & (gi >> pr[2 * p + 1]) & 1ull) v = cmul(v, ph[p]);
    }
    tile[j] = v;
  }
}

__device__ __forceinline__ void tile_scale(float2* tile, const TileProg& P, const cqs_op& op, int per) {
  const float2 s = P.mats[op.mat];
  for (int e = 0; e < per; ++e) {
    const int j = threadIdx.x + e * WG;
    tile[j] = cmul(tile[j], s);
  }
}

// Stage a matrix into LDS (for K >= 3 dense ops).  Caller fences before and after.
__device__ __forceinline__ void stage_matrix(float2* dst, const float2* src, int count) {
  for (int i = threadIdx.x; i < count; i += WG) dst[i] = src[i];
}

__device__ __forceinline__ void run_dense(float2* tile, float2* mstage, int L, int max_k, const cqs_op& op,
                                          const float2* mats) {
  const float2* M = mats + op.mat;
  if (op.k > max_k && op.k >= 3) return;  // the stage was not sized for it (host bug)
  // k = 3 stages its 8x8 matrix in LDS; k <= 2 reads it through the scalar cache
  switch (op.k) {
    case 1: tile_dense<1>(tile, L, op, M); break;
    case 2: tile_dense<2>(tile, L, op, M); break;
    case 3:
      stage_matrix(mstage, M, 64);
      __syncthreads();
      tile_dense<3>(tile, L, op, mstage);
      break;
    default: break;  // wider matrices run as CQS_OP_BLOCK on MFMA
  }
}

// Runs the program of P on one tile that is already in LDS.  The caller has loaded the
// tile with the element mapping (no barrier yet).  Returns with the tile complete in LDS in
// a state where the caller must fence before re-reading it with the element mapping iff
// the function returns false.
__device__ __forceinline__ bool run_program(float2* tile, float2* mstage, const TileProg& P,
                                            const cqs_op* ops, int op_count, uint64_t base,
                                            uint64_t off_lo, int nlo, int per) {
  bool elem_mode = true;
  for (int i = 0; i < op_count; ++i) {
    const cqs_op& op = ops[i];
    if (!outer_controls_ok(op, base)) continue;
    if (op.kind == CQS_OP_DENSE) {
      __syncthreads();
      run_dense(tile, mstage, P.L, P.max_k, op, P.mats);
      elem_mode = false;
    } else if (op.kind == CQS_OP_BLOCK) {
      __syncthreads();
      const half8_t* A = reinterpret_cast<const half8_t*>(P.mats + op.mat);
      if (op.k == 4)
        tile_block_mfma<4>(tile, P.L, op, A, P.mscale, 1.f / P.mscale);
      else if (op.k == 5)
        tile_block_mfma<5>(tile, P.L, op, A, P.mscale, 1.f / P.mscale);
      elem_mode = false;
    } else {
      if (!elem_mode) {
        __syncthreads();
        elem_mode = true;
      }
      if (op.kind == CQS_OP_DIAG) {
        tile_diag(tile, P, op, base, off_lo, nlo, per);
      } else if (op.kind == CQS_OP_PAIRPHASE) {
        tile_pairphase(tile, P, op, base, off_lo, nlo, per);
      } else if (op.kind == CQS_OP_SCALE) {
        tile_scale(tile, P, op, per);
      }
    }
  }
  return elem_mode;
}

}  // namespace cqs
