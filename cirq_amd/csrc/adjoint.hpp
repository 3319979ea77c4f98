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

// This thread's share (its groups) of 2 Re sum conj(phi) dU psi over the LDS tiles for a dense dU
// (k <= 3) on tile bits op.t.
template <int K, int NT, class Op>
__device__ __forceinline__ double grad_dense(const float2* psi, const float2* phi, int L, Op& op, smat_t* M) {
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
  // one group at a time: with the op's bits as constants the trip count is known, and unrolling it
  // would keep several groups' operands live at once (spills at 128 VGPRs)
#pragma unroll 1
  for (int g = threadIdx.x; g < G; g += NT) {
    int j0 = g;
#pragma unroll
    for (int s = 0; s < K; ++s) j0 = insert0(j0, ts[s]);
    float2 x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = psi[pad(j0 + pat[c])];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 y = make_float2(0.f, 0.f);  // (dU x)_r in FP32, as the state itself
#pragma unroll
      for (int c = 0; c < D; ++c) y = cfma(ldm(M, r * D + c), x[c], y);
      const float2 f = phi[pad(j0 + pat[r])];
      acc += (double)f.x * y.x + (double)f.y * y.y;  // Re(conj(f) y), summed in FP64
    }
  }
  return 2.0 * acc;
}

// GRAD ops run on register-resident tiles when their targets are register bits; otherwise the
// tile goes through LDS.  A DENSE or GRAD_DENSE op whose targets are register bits (ascending)
// stays in registers.
template <int PER, int NT>
__device__ __forceinline__ bool adj_register_dense(sop_t& op) {
  constexpr int TB = __builtin_ctz(NT);
  constexpr int NR = __builtin_ctz(PER);
  const int kind = op_kind(op);
  if ((kind != CQS_OP_DENSE && kind != CQS_OP_GRAD_DENSE) || op.k < 1 || op.k > 3 || op.k > NR) return false;
  if (op.t[0] < TB) return false;
#pragma unroll
  for (int m = 1; m < 3; ++m)
    if (m < op.k && op.t[m] <= op.t[m - 1]) return false;
  return true;
}

template <int PER, int NT>
__device__ __forceinline__ bool adj_mixing(sop_t& op) {
  const int kind = op_kind(op);
  if (kind == CQS_OP_BLOCK) return true;
  return (kind == CQS_OP_DENSE || kind == CQS_OP_GRAD_DENSE) && !adj_register_dense<PER, NT>(op);
}

// 2 Re sum conj(f) dU x over this thread's groups of a register-bit dense dU (targets R0 < R1 <
// R2, matrix index bit m <-> target m): dU x in FP32, the products with conj(f) summed in FP64.
template <int PER, int K, int R0, int R1, int R2>
__device__ __forceinline__ double grad_reg_dense_k(const float2 (&x)[PER], const float2 (&f)[PER], smat_t* M) {
  constexpr int D = 1 << K;
  constexpr int RS[3] = {R0, R1, R2};
  constexpr int TMASK = (1 << R0) | (K > 1 ? 1 << R1 : 0) | (K > 2 ? 1 << R2 : 0);
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    if (e & TMASK) continue;
    int idx[D];
#pragma unroll
    for (int m = 0; m < D; ++m) {
      int v = e;
#pragma unroll
      for (int b = 0; b < K; ++b) v |= ((m >> b) & 1) << RS[b];
      idx[m] = v;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 y = make_float2(0.f, 0.f);  // (dU x)_r in FP32, as the state itself
#pragma unroll
      for (int c = 0; c < D; ++c) y = cfma(ldm(M, r * D + c), x[idx[c]], y);
      acc += (double)f[idx[r]].x * y.x + (double)f[idx[r]].y * y.y;  // summed in FP64
    }
  }
  return 2.0 * acc;
}

template <int PER, int NT, class Prog>
__device__ __forceinline__ double grad_reg_dense(const float2 (&x)[PER], const float2 (&f)[PER], const Prog& P,
                                                 sop_t& op) {
  constexpr int TB = __builtin_ctz(NT);
  constexpr int NR = __builtin_ctz(PER);
  smat_t* M = scalar_ptr(P.mats) + op.mat;
  const int r0 = op.t[0] - TB, r1 = op.t[1] - TB, r2 = op.t[2] - TB;
  const int code = op.k == 1 ? r0 : (op.k == 2 ? 16 + r0 * 4 + r1 : 32 + r0 * 16 + r1 * 4 + r2);
#define CQS_G1(A) \
  case A: if constexpr (NR > A) return grad_reg_dense_k<PER, 1, A, 0, 0>(x, f, M); return 0.0;
#define CQS_G2(A, B) \
  case 16 + A * 4 + B: if constexpr (NR > B) return grad_reg_dense_k<PER, 2, A, B, 0>(x, f, M); return 0.0;
#define CQS_G3(A, B, C) \
  case 32 + A * 16 + B * 4 + C: if constexpr (NR > C) return grad_reg_dense_k<PER, 3, A, B, C>(x, f, M); return 0.0;
  switch (code) {
    CQS_G1(0) CQS_G1(1) CQS_G1(2) CQS_G1(3)
    CQS_G2(0, 1) CQS_G2(0, 2) CQS_G2(0, 3) CQS_G2(1, 2) CQS_G2(1, 3) CQS_G2(2, 3)
    CQS_G3(0, 1, 2) CQS_G3(0, 1, 3) CQS_G3(0, 2, 3) CQS_G3(1, 2, 3)
    default: return 0.0;
  }
#undef CQS_G1
#undef CQS_G2
#undef CQS_G3
}

// 2 Re sum conj(f_j) d(j) x_j over this thread's register elements for a diagonal derivative d.
template <int PER, int NT, class Prog, class Op>
__device__ __forceinline__ double grad_reg_diag(const float2 (&x)[PER], const float2 (&f)[PER], const Prog& P,
                                                Op& op, uint64_t base) {
  const DiagTable<PER, NT, Prog> D(P, op, base);
  double acc = 0.0;
  if (D.small) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const float2 y = cmul(D.pick(D.index(e)), x[e]);
      acc += (double)f[e].x * y.x + (double)f[e].y * y.y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const float2 y = cmul(D.Dg[D.index(e)], x[e]);
      acc += (double)f[e].x * y.x + (double)f[e].y * y.y;
    }
  }
  return 2.0 * acc;
}

// Every wave writes its own partial of a GRAD op: partial[slot * stride + tile * NW + wave] (a
// shuffle reduction inside the wave, no workgroup barrier; the host sums the NW entries per tile).
template <int NT>
__device__ __forceinline__ void write_wave_partial(double g, double* __restrict__ partial, int64_t stride,
                                                   int64_t tile, int slot) {
  g = wave_sum(g);
  if ((threadIdx.x & 63) == 0)
    partial[(int64_t)slot * stride + tile * (NT / 64) + (threadIdx.x >> 6)] = g;
}

// Waves per workgroup of the adjoint kernels at tile width L (the factor between tiles and
// per-slot partial entries).
__host__ __device__ constexpr int adjoint_waves(int L) { return (L <= 12 ? 256 : (1 << L) / 16) / 64; }

// One workgroup's share of a reverse-sweep pass.  psi and phi stay in registers (this thread's
// elements, natural layout) through runs of elementwise ops, register-bit dense ops and GRAD ops on
// register bits; ops that mix elements of different threads (other dense ops, MFMA blocks, GRAD
// ops on thread bits) run on the two LDS tiles between barriers.  GRAD ops write one FP64 partial
// per wave.  `stride` is the per-slot partial stride (tiles x waves).
template <int L, class Prog>
__device__ __forceinline__ void adjoint_tile(float2* __restrict__ psi, float2* __restrict__ phi, const Prog& P,
                                             const int32_t* lpos, uint64_t base, double* __restrict__ partial,
                                             int64_t stride, int64_t tile, float2* lds) {
  constexpr int NT = TileShape<L>::NT, PER = TileShape<L>::PER;
  float2* tA = lds;
  float2* tB = lds + tile_words(L);
  float2* mstage = lds + 2 * tile_words(L);
  const bool act = L >= 8 || (int)threadIdx.x < (1 << L);
  warm_program<NT>(P);
  float2 va[PER], vb[PER];
  {  // offsets are recomputed for the store, so they are not live across the program
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(lpos, off_lo, off_hi);
    if (act) {
#pragma unroll
      for (int e = 0; e < PER; ++e) {
        va[e] = psi[base + off_lo + off_hi[e]];
        vb[e] = phi[base + off_lo + off_hi[e]];
      }
    } else {
#pragma unroll
      for (int e = 0; e < PER; ++e) va[e] = vb[e] = make_float2(0.f, 0.f);
    }
  }
  int i = 0;
  for (;;) {
    // register phase
    for (; i < P.op_count && !adj_mixing<PER, NT>(scalar_ptr(P.ops)[i]); ++i) {
      sop_t& op = scalar_ptr(P.ops)[i];
      const int kind = op_kind(op), sel = op_sel(op);
      const bool ok = outer_controls_ok(op, base);
      if (kind == CQS_OP_GRAD_DIAG || kind == CQS_OP_GRAD_DENSE) {
        double g = 0.0;
        if (ok && act) g = kind == CQS_OP_GRAD_DIAG ? grad_reg_diag<PER, NT>(va, vb, P, op, base)
                                                    : grad_reg_dense<PER, NT>(va, vb, P, op);
        write_wave_partial<NT>(g, partial, stride, tile, op.slot);
        continue;
      }
      if (!ok) continue;
      if (kind == CQS_OP_DENSE) {
        if (sel != kSelPhi) reg_dense<PER, NT>(va, P, op, base);
        if (sel != kSelPsi) reg_dense<PER, NT>(vb, P, op, base);
      } else if (kind == CQS_OP_DIAG) {
        if (sel != kSelPhi) ew_diag<PER, NT>(va, P, op, base);
        if (sel != kSelPsi) ew_diag<PER, NT>(vb, P, op, base);
      } else if (kind == CQS_OP_PAIRPHASE) {
        if (sel != kSelPhi) ew_pairs<PER, NT>(va, P, op, base);
        if (sel != kSelPsi) ew_pairs<PER, NT>(vb, P, op, base);
      } else if (kind == CQS_OP_SCALE) {
        if (sel != kSelPhi) ew_scale<PER>(va, P, op);
        if (sel != kSelPsi) ew_scale<PER>(vb, P, op);
      }
    }
    if (i >= P.op_count) break;
    // LDS phase: both tiles out of registers, then every consecutive op that mixes threads
    regs_to_tile<PER, NT>(va, tA, act);
    regs_to_tile<PER, NT>(vb, tB, act);
    for (; i < P.op_count && adj_mixing<PER, NT>(scalar_ptr(P.ops)[i]); ++i) {
      sop_t& op = scalar_ptr(P.ops)[i];
      __syncthreads();
      const int kind = op_kind(op), sel = op_sel(op);
      const bool ok = outer_controls_ok(op, base);
      if (kind == CQS_OP_GRAD_DENSE) {
        double g = 0.0;
        if (ok) {
          smat_t* M = scalar_ptr(P.mats) + op.mat;
          if (op.k == 1)
            g = grad_dense<1, NT>(tA, tB, L, op, M);
          else if (op.k == 2)
            g = grad_dense<2, NT>(tA, tB, L, op, M);
          else if (op.k == 3)
            g = grad_dense<3, NT>(tA, tB, L, op, M);
        }
        write_wave_partial<NT>(g, partial, stride, tile, op.slot);
        continue;
      }
      if (!ok) continue;
      if (sel != kSelPhi) run_mixing_op<NT>(tA, mstage, P, op, base);
      if (sel == kSelBoth) __syncthreads();  // the matrix stage is shared by both tiles
      if (sel != kSelPsi) run_mixing_op<NT>(tB, mstage, P, op, base);
    }
    __syncthreads();
    tile_to_regs<PER, NT>(va, tA, act);
    tile_to_regs<PER, NT>(vb, tB, act);  // each thread reads back only its own slots: no barrier until the next mix
  }
  uint64_t off_lo, off_hi[PER];
  tile_offsets<L>(lpos, off_lo, off_hi);
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      psi[base + off_lo + off_hi[e]] = va[e];
      phi[base + off_lo + off_hi[e]] = vb[e];
    }
  }
}

}  // namespace cqs
