// Device side of adjoint differentiation, shared by the single-state and batched kernels.
//
// The host supplies the reversed program: every gate U_k becomes U_k^dagger, applied to both
// tiles when it has no parameter, and split into [U_k^dagger on psi][GRAD dU_k][U_k^dagger on phi]
// when it has one, so that GRAD sees psi_{k-1} and phi_k and accumulates
//   g_k = 2 Re <phi_k | dU_k | psi_{k-1}>        (Jones & Gacon 2020, arXiv:2009.02823)
// in FP64, one partial per tile.
#pragma once

#include "tile.hpp"

namespace cqs {

constexpr int kSelBoth = 0, kSelPsi = 1, kSelPhi = 2;

__device__ __forceinline__ int op_kind(sop_t& op) { return op.kind & 0xff; }
__device__ __forceinline__ int op_sel(sop_t& op) { return (op.kind >> 8) & 3; }

// Elementwise op on one LDS tile (this thread's elements).
template <int PER, int NT, class Prog>
__device__ __forceinline__ void ew_on_tile(float2* tile, const Prog& P, sop_t& op, uint64_t base,
                                           bool act) {
  if (!act) return;
  float2 v[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) v[e] = tile[pad_nat<NT>(e)];
  const int kind = op_kind(op);
  if (kind == CQS_OP_DIAG)
    ew_diag<PER, NT>(v, P, op, base);
  else if (kind == CQS_OP_PAIRPHASE)
    ew_pairs<PER, NT>(v, P, op, base);
  else if (kind == CQS_OP_SCALE)
    ew_scale<PER>(v, P, op);
#pragma unroll
  for (int e = 0; e < PER; ++e) tile[pad_nat<NT>(e)] = v[e];
}

// sum over the tile of 2 Re(conj(phi) dU psi) for a dense dU (k <= 3) on tile bits op.t.
template <int K, int NT>
__device__ __forceinline__ double grad_dense(const float2* psi, const float2* phi, int L, sop_t& op, smat_t* M) {
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
  double acc = 0.0;
  const int G = (1 << L) >> K;
  for (int g = threadIdx.x; g < G; g += NT) {
    int j0 = g;
#pragma unroll
    for (int s = 0; s < K; ++s) j0 = insert0(j0, ts[s]);
    float2 x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = psi[pad(j0 + pat[c])];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double yr = 0.0, yi = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const float2 m = ldm(M, r * D + c);
        yr += (double)m.x * x[c].x - (double)m.y * x[c].y;
        yi += (double)m.x * x[c].y + (double)m.y * x[c].x;
      }
      const float2 f = phi[pad(j0 + pat[r])];
      acc += (double)f.x * yr + (double)f.y * yi;  // Re(conj(f) y)
    }
  }
  return 2.0 * acc;
}

// sum over the tile of 2 Re(conj(phi_j) d(j) psi_j) for a diagonal derivative d.
template <int PER, int NT, class Prog>
__device__ __forceinline__ double grad_diag(const float2* psi, const float2* phi, const Prog& P, sop_t& op,
                                            uint64_t base, bool act) {
  if (!act) return 0.0;
  constexpr int TB = __builtin_ctz(NT);
  const int tid = threadIdx.x;
  int fixed = 0, emask = 0;
#pragma unroll
  for (int m = 0; m < 6; ++m) {  // k <= 6
    if (m >= op.k) break;
    const int t = op.t[m];
    if (t >= kOuter)
      fixed |= (int)((base >> (t - kOuter)) & 1ull) << m;
    else if (t < TB)
      fixed |= ((tid >> t) & 1) << m;
    else
      emask |= 1 << m;
  }
  const float2* Dg = P.mats + op.mat;
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    int idx = fixed;
#pragma unroll
    for (int m = 0; m < 6; ++m)
      if (m < op.k && ((emask >> m) & 1)) idx |= ((e >> (op.t[m] - TB)) & 1) << m;
    const float2 d = Dg[idx], x = psi[pad_nat<NT>(e)], f = phi[pad_nat<NT>(e)];
    const double yr = (double)d.x * x.x - (double)d.y * x.y;
    const double yi = (double)d.x * x.y + (double)d.y * x.x;
    acc += (double)f.x * yr + (double)f.y * yi;
  }
  return 2.0 * acc;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// One workgroup's share of a reverse-sweep pass: both tiles (psi, phi) in LDS; GRAD ops write the
// tile's partial to partial[slot * stride + tile].
template <int L, class Prog>
__device__ __forceinline__ void adjoint_tile(float2* __restrict__ psi, float2* __restrict__ phi, const Prog& P,
                                             const int32_t* lpos, uint64_t base, double* __restrict__ partial,
                                             int64_t stride, int64_t tile, float2* lds, double* red) {
  constexpr int NT = TileShape<L>::NT, PER = TileShape<L>::PER;
  float2* tA = lds;
  float2* tB = lds + tile_words(L);
  float2* mstage = lds + 2 * tile_words(L);
  const bool act = L >= 8 || (int)threadIdx.x < (1 << L);
  uint64_t off_lo, off_hi[PER];
  tile_offsets<L>(lpos, off_lo, off_hi);
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      tA[pad_nat<NT>(e)] = psi[base + off_lo + off_hi[e]];
      tB[pad_nat<NT>(e)] = phi[base + off_lo + off_hi[e]];
    }
  }
  for (int i = 0; i < P.op_count; ++i) {
    sop_t& op = scalar_ptr(P.ops)[i];
    __syncthreads();
    const bool ok = outer_controls_ok(op, base);
    const int kind = op_kind(op), sel = op_sel(op);
    if (kind == CQS_OP_GRAD_DENSE || kind == CQS_OP_GRAD_DIAG) {
      double g = 0.0;
      if (!ok) {
        // the gate does not act on this tile: its derivative contributes nothing here
      } else if (kind == CQS_OP_GRAD_DENSE) {
        smat_t* M = scalar_ptr(P.mats) + op.mat;
        if (op.k == 1)
          g = grad_dense<1, NT>(tA, tB, L, op, M);
        else if (op.k == 2)
          g = grad_dense<2, NT>(tA, tB, L, op, M);
        else if (op.k == 3)
          g = grad_dense<3, NT>(tA, tB, L, op, M);
      } else {
        g = grad_diag<PER, NT>(tA, tB, P, op, base, act);
      }
      g = block_sum<NT>(g, red);
      if (threadIdx.x == 0) partial[(int64_t)op.slot * stride + tile] = g;
      continue;
    }
    if (!ok) continue;
    if (kind == CQS_OP_DENSE || kind == CQS_OP_BLOCK) {
      if (sel != kSelPhi) run_mixing_op<NT>(tA, mstage, P, op, base);
      if (sel == kSelBoth) __syncthreads();  // the matrix stage is shared by both tiles
      if (sel != kSelPsi) run_mixing_op<NT>(tB, mstage, P, op, base);
    } else {
      if (sel != kSelPhi) ew_on_tile<PER, NT>(tA, P, op, base, act);
      if (sel != kSelPsi) ew_on_tile<PER, NT>(tB, P, op, base, act);
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      psi[base + off_lo + off_hi[e]] = tA[pad_nat<NT>(e)];
      phi[base + off_lo + off_hi[e]] = tB[pad_nat<NT>(e)];
    }
  }
}

}  // namespace cqs
