// Tile program interpreter: the device side of one pass (see cqs_pass in cirq_amd.h).
//
// A workgroup owns one tile: the 2^L amplitudes whose flat indices differ only in the L
// "local" bit positions lpos[] (ascending).  Tile index j (0..2^L) maps to flat offset
// deposit(j, lpos); the workgroup number maps to the remaining "outer" positions opos[].
//
// Thread tid of NT owns elements j = tid + e*NT (e < PER): it loads them from HBM into
// registers and keeps them there through elementwise ops (diagonals, pair phases, scales) and
// dense ops on register bits (tile bits >= log2(NT)).  The tile goes through LDS only for ops
// that mix elements of different threads (other dense ops, MFMA blocks) and for PERMUTE, which
// re-lays the tile so that the next run of gates finds its targets among the register bits.
#pragma once

#include "common.hpp"

namespace cqs {

struct TileProg {
  int32_t n;         // total flat-index bits
  int32_t L;         // local bits
  int32_t nouter;    // n - L
  int32_t op_count;  // ops in this pass
  int32_t max_k;     // largest VALU dense k the LDS matrix stage was sized for
  float mscale;      // amplitude scale of the split-FP16 MFMA path (2^15 for normalised states)
  int32_t warm;      // nonzero: touch the pass's op descriptors and matrices first (warm_program)
  int32_t lpos[CQS_MAX_LOCAL];      // tile bit -> flat bit when the pass loads
  int32_t lpos_out[CQS_MAX_LOCAL];  // tile bit -> flat bit when it stores (after PERMUTE ops)
  int32_t opos[CQS_MAX_BITS];
  const cqs_op* ops;
  const float2* mats;
  const int32_t* aux;
  int64_t basis;  // >= 0: the pass starts from basis state |basis> instead of reading psi
  // The state is zero outside {x : x & zmask == zval} (a chain of passes from a basis state: the
  // bits no earlier pass had local still hold the basis's values), restricted here to the pass's
  // outer bits: a tile whose outer bits disagree holds only zeros.  zmask = 0: no such knowledge.
  uint64_t zmask, zval;
  // Tiles whose outer bits disagree with (wmask, wval) are not stored at all: the next pass of the
  // same run reads none of their elements (it stores them itself, as zeros).  wmask = 0: store all.
  uint64_t wmask, wval;
};

// The part of a pass program the tile interpreter reads (TileProg carries the same fields).
struct ProgView {
  int32_t L;
  int32_t op_count;
  int32_t max_k;
  float mscale;
  int32_t warm;
  const cqs_op* ops;
  const float2* mats;
  const int32_t* aux;
};

// LDS tiles hold one empty slot after every 32 amplitudes: lanes that walk a tile with a stride
// of 32 amplitudes or more (MFMA block columns when the block's targets are low tile bits, dense
// groups, PERMUTE writes) then fall into different banks instead of one.
constexpr int kPadShift = 5;
__device__ __forceinline__ int pad(int j) { return j + (j >> kPadShift); }
__host__ __device__ constexpr int tile_words(int L) { return (1 << L) + ((1 << L) >> kPadShift); }
// pad(threadIdx.x + e * NT) for NT a multiple of 2^kPadShift: one per-thread base plus a constant
// per element (which the LDS instructions take as an immediate offset).
template <int NT>
__device__ __forceinline__ int pad_nat(int e) {
  static_assert(NT % (1 << kPadShift) == 0, "natural layout needs NT a multiple of the pad period");
  return pad((int)threadIdx.x) + e * (NT + (NT >> kPadShift));
}

// Ops, matrices and aux entries are the same for every lane of a workgroup: they are read through
// the constant address space, so that uniform addresses become scalar loads (the scalar data
// cache) instead of one vector-memory load per lane with a full memory latency per op.  These
// pointers are only ever read.
#define CQS_SCALAR __attribute__((address_space(4)))
typedef const CQS_SCALAR cqs_op sop_t;
typedef const CQS_SCALAR float2 smat_t;
typedef const CQS_SCALAR int32_t saux_t;
__device__ __forceinline__ sop_t* scalar_ptr(const cqs_op* p) { return (sop_t*)p; }
__device__ __forceinline__ smat_t* scalar_ptr(const float2* p) { return (smat_t*)p; }
__device__ __forceinline__ saux_t* scalar_ptr(const int32_t* p) { return (saux_t*)p; }
// element i of a matrix, from the constant address space or (staged) from LDS
__device__ __forceinline__ float2 ldm(smat_t* p, int i) {
  const CQS_SCALAR float* f = (const CQS_SCALAR float*)p;
  return make_float2(f[2 * i], f[2 * i + 1]);
}
__device__ __forceinline__ float2 ldm(const float2* p, int i) { return p[i]; }

// Bit encoding of DIAG / PAIRPHASE positions: values < kOuter are tile-local indices, values
// >= kOuter are kOuter + a global position outside the tile.
constexpr int kOuter = 64;

// Fused DENSE / BLOCK ops may hold several matrices, one per value of up to 3 outer bits: t[5] is
// count << 24 | the bits' flat positions (7 bits each); the tile's outer bits pick the variant.
__device__ __forceinline__ int variant_of(sop_t& op, uint64_t base) {
  const int code = op.t[5];
  const int cnt = (code >> 24) & 3;
  int v = 0;
  for (int i = 0; i < cnt; ++i) v |= (int)((base >> ((code >> (7 * i)) & 127)) & 1ull) << i;
  return v;
}

__device__ __forceinline__ bool outer_controls_ok(sop_t& op, uint64_t base) {
  const uint64_t m = ((uint64_t)op.co_mask_hi << 32) | op.co_mask_lo;
  const uint64_t v = ((uint64_t)op.co_val_hi << 32) | op.co_val_lo;
  return (base & m) == v;
}

// ---------------------------------------------------------------- dense ops on the LDS tile
// 2^K x 2^K matrix on tile bits op.t[0..K); M row-major, bit m of the matrix index is tile bit
// op.t[m].  Each thread owns whole groups, so no barrier is needed inside.
template <int K, int NT, class Op, class MPtr>
__device__ __forceinline__ void tile_dense(float2* tile, int L, Op& op, MPtr M) {
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
  // Group g = tid + g0 (g0 a multiple of NT: disjoint bits).  Inserting zeros at the targets and
  // pad() are additive over disjoint bits, so the lane's part is placed once and the uniform
  // part g0 and the row offsets pat[m] on the scalar unit.
  int jt = threadIdx.x;
#pragma unroll
  for (int s = 0; s < K; ++s) jt = insert0(jt, ts[s]);
  const int pt = pad(jt);
  int ppat[D];
#pragma unroll
  for (int m = 0; m < D; ++m) ppat[m] = pad(pat[m]);
  if ((int)threadIdx.x >= G) return;  // tiles with fewer groups than threads
  // one group at a time: with the op's bits as constants the trip count is known, and unrolling it
  // would keep several groups' operands live at once (spills at 128 VGPRs)
#pragma unroll 1
  for (int g0 = 0; g0 < G; g0 += NT) {
    int jh = g0;
#pragma unroll
    for (int s = 0; s < K; ++s) jh = insert0(jh, ts[s]);
    if (((jt | jh) & cm) != cv) continue;
    const int b = pt + pad(jh);
    float2 v[D];
#pragma unroll
    for (int m = 0; m < D; ++m) v[m] = tile[b + ppat[m]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) acc = cfma(ldm(M, r * D + c), v[c], acc);
      tile[b + ppat[r]] = acc;
    }
  }
}

// ---------------------------------------------------------------- fused blocks on MFMA
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Split x * s into FP16 hi + lo (x * s = hi + lo + O(2^-22 |x s|)) for a complex value, packed as
// (re, im) pairs.  Each half comes from one mixed-precision FMA that rounds to nearest straight
// into the low or high half of a register: hi = f16(x * s), lo = f16(x * s - hi), where the FMA
// computes x * s - hi exactly before its one rounding.
__device__ __forceinline__ void split_f16(float xr, float xi, float s, half2_t& hi, half2_t& lo) {
  // the low-half writes are plain outputs: the high-half write that follows fills the rest
  asm("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(hi) : "v"(xr), "v"(s));
  asm("v_fma_mixhi_f16 %0, %1, %2, 0" : "+v"(hi) : "v"(xi), "v"(s));
  asm("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]" : "=v"(lo) : "v"(xr), "v"(s), "v"(hi));
  asm("v_fma_mixhi_f16 %0, %1, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(lo) : "v"(xi), "v"(s), "v"(hi));
}

// Y = U X on tile bits op.t[0..K) with v_mfma_f32_32x32x16_f16.  The tile is viewed as a
// (2^K rows) x (2^(L-K) columns) complex matrix; in real-interleaved form (row 2r = Re, 2r+1 = Im)
// it is R (2^(K+1) square) times X.  Each wave owns whole 32-column blocks, so reads and writes of
// different waves never overlap and the block needs no barrier inside.  X is split as
// hi + lo FP16 after scaling by `scale` (split_f16); U's split comes from the host in A-operand
// order.
//
// Addressing: the tile index of (row r, column c) is c with zeros inserted at the targets, OR the
// bits of r deposited at the targets.  pad() is additive over disjoint bit sets, so the padded
// index is a per-lane column part + a per-lane part from the lane's half (h) + a part that depends
// only on compile-time row bits and the op's targets (scalar).
// kFull: every 32-column block is whole (L - K >= 5), so no column needs a bounds test.
template <int K, int NT, bool kFull, class Op>
__device__ __forceinline__ void tile_block_mfma(float2* tile, int L, Op& op, const half8_t* A,
                                                float scale, float inv_scale) {
  constexpr int D = 1 << K;
  constexpr int R = 2 * D;
  constexpr int KS = R / 16;
  constexpr int RB = R / 32;
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, h = lane >> 5, c32 = lane & 31;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);  // uniform: column blocks in SGPRs
  int pbit[K];  // pad(1 << t[m])
  int ts[K];    // targets in ascending order
#pragma unroll
  for (int m = 0; m < K; ++m) {
    const int b = 1 << op.t[m];
    pbit[m] = b + (b >> kPadShift);
    ts[m] = op.t[m];
  }
#pragma unroll
  for (int a = 1; a < K; ++a)
#pragma unroll
    for (int b = K - 1; b >= a; --b)
      if (ts[b - 1] > ts[b]) {
        const int x = ts[b - 1];
        ts[b - 1] = ts[b];
        ts[b] = x;
      }
  auto prow = [&](int r) {  // padded offset of row r (r is a compile-time constant at every use)
    int o = 0;
#pragma unroll
    for (int m = 0; m < K; ++m)
      if ((r >> m) & 1) o += pbit[m];
    return o;
  };
  const int ncol = 1 << (L - K);
  const int ncb = (ncol + 31) >> 5;
  const int h_in = h ? pbit[2] : 0;   // B operand row r = 8 ks + 4 h + e: bit 2 is h
  const int h_out = h ? pbit[1] : 0;  // output row 16 rb + 4 (a >> 2) + 2 h + ((a >> 1) & 1)
  // A fragments stay in registers across column blocks when they are few (K = 4)
  constexpr bool kHoistA = KS * RB * 2 <= 4;
  half8_t ahold[kHoistA ? KS * RB * 2 : 1];
  if constexpr (kHoistA) {
#pragma unroll
    for (int f = 0; f < KS * RB * 2; ++f) ahold[f] = A[f * 64 + lane];
  }
  // Which non-target tile bits carry the lane's column bits (c32) and which the column block's.
  // Any split of the non-target bits between the two visits every column once, and pad() is
  // additive over disjoint bits, so the lane part is placed once and the block part per block on
  // the scalar unit.  LDS serves an 8-byte access 16 lanes per cycle over 32 four-byte banks: the
  // 16 lanes of one cycle (c32 bits 0-3) need padded word offsets that differ modulo 16.
  // pad(1 << b) is 2^b modulo 16 for b < 4 and 2^(b-5) for 5 <= b < 9, so (whole column blocks)
  // c32 bit j takes tile bit j, or bit j + 5 when j is a target; c32 bit 4, and a bit j whose two
  // choices are both targets, take the lowest free bits; the block index takes the rest in order.
  // Partial column blocks (L - K < 5) keep the five lowest non-target bits for the lane.
  int jl = 0;
  uint32_t gmask = 0;  // whole column blocks: the tile bits of the column block index
  if constexpr (kFull) {
    uint32_t used = 0;
#pragma unroll
    for (int m = 0; m < K; ++m) used |= 1u << ts[m];
    int lp[5];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      lp[j] = !((used >> j) & 1) ? j : !((used >> (j + 5)) & 1) ? j + 5 : -1;
      if (lp[j] >= 0) used |= 1u << lp[j];
    }
    lp[4] = -1;
#pragma unroll
    for (int j = 0; j < 5; ++j)
      if (lp[j] < 0) {
        lp[j] = __builtin_ctz(~used);
        used |= 1u << lp[j];
      }
#pragma unroll
    for (int j = 0; j < 5; ++j) jl |= ((c32 >> j) & 1) << lp[j];
    gmask = ~used & ((1u << L) - 1);
  } else {
    jl = c32;
#pragma unroll
    for (int m = 0; m < K; ++m) jl = insert0(jl, ts[m]);
  }
  const int pl = jl + (jl >> kPadShift);
  const int b_in0 = pl + h_in, b_out0 = pl + h_out;
  // A column block's padded input / output offsets and its LDS writes.
  auto offsets = [&](int cb, int& b_in, int& b_out) {
    int jh = 0;
    if constexpr (kFull) {
      uint32_t g = gmask;
      for (int i = 0; g != 0; ++i) {  // deposit cb's bits at the block bits, lowest first
        const int b = __builtin_ctz(g);
        g &= g - 1;
        jh |= ((cb >> i) & 1) << b;
      }
    } else {
      jh = cb * 32;
#pragma unroll
      for (int m = 0; m < K; ++m) jh = insert0(jh, ts[m]);
    }
    const int ph = jh + (jh >> kPadShift);
    b_in = b_in0 + ph;
    b_out = b_out0 + ph;
  };
  auto store = [&](int b_out, bool valid, const floatx16_t (&acc)[RB]) {
    if (valid) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int a = 0; a < 16; a += 2)  // a even: real part, a + 1: imaginary part of one row
          tile[b_out + prow(rb * 16 + 4 * (a >> 2) + ((a >> 1) & 1))] =
              make_float2(acc[rb][a] * inv_scale, acc[rb][a + 1] * inv_scale);
    }
  };
#pragma unroll 1  // one column block at a time, each K-slice split as it is read
  for (int cb = wave; cb < ncb; cb += NW) {
    int b_in, b_out;
    offsets(cb, b_in, b_out);
    const bool valid = kFull || cb * 32 + c32 < ncol;
    floatx16_t acc[RB];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      half2_t hh[4], ll[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float2 v = valid ? tile[b_in + prow(ks * 8 + e)] : make_float2(0.f, 0.f);
        split_f16(v.x, v.y, scale, hh[e], ll[e]);
      }
      const half8_t bhi = __builtin_shufflevector(__builtin_shufflevector(hh[0], hh[1], 0, 1, 2, 3),
                                                  __builtin_shufflevector(hh[2], hh[3], 0, 1, 2, 3), 0, 1, 2, 3,
                                                  4, 5, 6, 7);
      const half8_t blo = __builtin_shufflevector(__builtin_shufflevector(ll[0], ll[1], 0, 1, 2, 3),
                                                  __builtin_shufflevector(ll[2], ll[3], 0, 1, 2, 3), 0, 1, 2, 3,
                                                  4, 5, 6, 7);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        half8_t ahi, alo;
        if constexpr (kHoistA) {
          ahi = ahold[ks * RB + rb];
          alo = ahold[KS * RB + ks * RB + rb];
        } else {
          // A operands stream from the L1/L2 caches (every workgroup reads the same matrix)
          ahi = A[(ks * RB + rb) * 64 + lane];
          alo = A[(KS * RB + ks * RB + rb) * 64 + lane];
        }
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, bhi, ks == 0 ? floatx16_t{} : acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(alo, bhi, acc[rb], 0, 0, 0);
        acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ahi, blo, acc[rb], 0, 0, 0);
      }
    }
    store(b_out, valid, acc);
  }
}

template <int NT>
__device__ __forceinline__ void stage_matrix(float2* dst, smat_t* src, int count) {
  for (int i = threadIdx.x; i < count; i += NT) dst[i] = ldm(src, i);
}

// Dense (VALU, k <= 3) or block (MFMA, k = 4, 5) op on the LDS tile.  Caller fences before.
template <int NT, class Prog>
__device__ __forceinline__ void run_mixing_op(float2* tile, float2* mstage, const Prog& P, sop_t& op,
                                              uint64_t base) {
  const int var = variant_of(op, base);
  if ((op.kind & 0xff) == CQS_OP_BLOCK) {  // bits 8-9 may carry an adjoint tile selector
    const half8_t* A = reinterpret_cast<const half8_t*>(P.mats + op.mat + var * (2 << (2 * op.k)));
    const bool full = P.L - op.k >= 5;
    if (op.k == 4) {
      if (full)
        tile_block_mfma<4, NT, true>(tile, P.L, op, A, P.mscale, 1.f / P.mscale);
      else
        tile_block_mfma<4, NT, false>(tile, P.L, op, A, P.mscale, 1.f / P.mscale);
    } else if (op.k == 5) {
      if (full)
        tile_block_mfma<5, NT, true>(tile, P.L, op, A, P.mscale, 1.f / P.mscale);
      else
        tile_block_mfma<5, NT, false>(tile, P.L, op, A, P.mscale, 1.f / P.mscale);
    }
    return;
  }
  smat_t* M = scalar_ptr(P.mats) + op.mat + var * (1 << (2 * op.k));
  switch (op.k) {
    case 1: tile_dense<1, NT>(tile, P.L, op, M); break;
    case 2: tile_dense<2, NT>(tile, P.L, op, M); break;
    case 3:
      if (P.max_k < 3) return;  // the stage was not sized for it (host bug)
      stage_matrix<NT>(mstage, M, 64);
      __syncthreads();
      tile_dense<3, NT>(tile, P.L, op, mstage);
      break;
    default: break;
  }
}

// ---------------------------------------------------------------- elementwise ops in registers
// v[e] holds tile element j = threadIdx.x + e*NT: tile bits below log2(NT) are thread bits (the
// same for all of a thread's elements), the rest are register bits (bits of e, the same for every
// thread).  Both ops below split their bit conditions that way, so the per-element work is one
// complex multiply with everything else decided once per op.

// Diagonal over k bits (tile-local or outer): index bits from thread / outer bits form `fixed`
// (per lane); index bits from register bit r contribute w[r] (uniform), so element e reads entry
// fixed | sub(e).  Tables of at most 4 entries (one- and two-qubit diagonals: Rz, ZZ, CZPow) are
// read through scalar loads and picked per lane with selects, so the op waits on no vector load
// (a per-lane gather costs a memory latency per op, which few workgroups per CU cannot hide);
// larger tables are gathered per lane.
template <int PER, int NT, class Prog>
struct DiagTable {
  static constexpr int TB = __builtin_ctz(NT);
  static constexpr int NR = __builtin_ctz(PER);
  int fixed = 0;
  int w[NR > 0 ? NR : 1];
  int any_reg = 0;
  bool small;
  float2 d0, d1, d2, d3;
  const float2* Dg;

  template <class Op>
  __device__ __forceinline__ DiagTable(const Prog& P, Op& op, uint64_t base) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int r = 0; r < (NR > 0 ? NR : 1); ++r) w[r] = 0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {  // k <= 6; unrolled so the target loads issue together
      if (m >= op.k) break;
      const int t = op.t[m];
      if (t >= kOuter) {
        fixed |= (int)((base >> (t - kOuter)) & 1ull) << m;
      } else if (t < TB) {
        fixed |= ((tid >> t) & 1) << m;
      } else {
        any_reg = 1;
#pragma unroll
        for (int r = 0; r < NR; ++r)
          if (t - TB == r) w[r] |= 1 << m;
      }
    }
    small = op.k <= 2;
    Dg = P.mats + op.mat;
    if (small) {
      smat_t* Ds = scalar_ptr(P.mats) + op.mat;
      d0 = ldm(Ds, 0);
      d1 = ldm(Ds, 1);
      d2 = d0;
      d3 = d1;
      if (op.k == 2) {
        d2 = ldm(Ds, 2);
        d3 = ldm(Ds, 3);
      }
    }
  }
  __device__ __forceinline__ int index(int e) const {
    int idx = fixed;
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if ((e >> r) & 1) idx |= w[r];
    return idx;
  }
  __device__ __forceinline__ float2 pick(int idx) const {
    const float2 lo = (idx & 1) ? d1 : d0;
    const float2 hi = (idx & 1) ? d3 : d2;
    return (idx & 2) ? hi : lo;
  }
  // entry of element e (uniform branch on `small`)
  __device__ __forceinline__ float2 at(int e) const { return small ? pick(index(e)) : Dg[index(e)]; }
};

template <int PER, int NT, class Prog, class Op>
__device__ __forceinline__ void ew_diag(float2 (&v)[PER], const Prog& P, Op& op, uint64_t base) {
  const DiagTable<PER, NT, Prog> D(P, op, base);
  if (!D.any_reg) {
    const float2 d = D.small ? D.pick(D.fixed) : D.Dg[D.fixed];
#pragma unroll
    for (int e = 0; e < PER; ++e) v[e] = cmul(v[e], d);
    return;
  }
  if (D.small) {
#pragma unroll
    for (int e = 0; e < PER; ++e) v[e] = cmul(v[e], D.pick(D.index(e)));
    return;
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) v[e] = cmul(v[e], D.Dg[D.index(e)]);
}

// Products of phases on bit pairs (a CZ layer).  Pairs on thread / outer bits only multiply one
// per-lane phase; pairs touching register bits multiply the elements whose e has those bits.
template <int PER, int NT, class Prog, class Op>
__device__ __forceinline__ void ew_pairs(float2 (&v)[PER], const Prog& P, Op& op, uint64_t base) {
  constexpr int TB = __builtin_ctz(NT);
  saux_t* pr = scalar_ptr(P.aux) + op.t[0];
  smat_t* ph = scalar_ptr(P.mats) + op.mat;
  const int tid = threadIdx.x;
  float2 cph = make_float2(1.f, 0.f);
  for (int p0 = 0; p0 < op.k; p0 += 8)  // chunks of 8 pairs, unrolled: their aux loads issue together
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int p = p0 + q;
    if (p >= op.k) break;
    const int ab[2] = {pr[2 * p], pr[2 * p + 1]};
    bool en = true;
    int tneed = 0, rneed = 0;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int x = ab[q];
      if (x >= kOuter)
        en = en && ((base >> (x - kOuter)) & 1ull);
      else if (x < TB)
        tneed |= 1 << x;
      else
        rneed |= 1 << (x - TB);
    }
    if (!en) continue;
    const bool tc = (tid & tneed) == tneed;
    const float2 f = ldm(ph, p);
    if (rneed == 0) {
      const float2 g = cmul(cph, f);
      cph = tc ? g : cph;
      continue;
    }
#pragma unroll
    for (int e = 0; e < PER; ++e)
      if ((e & rneed) == rneed) {  // uniform test; the lane condition is a select
        const float2 g = cmul(v[e], f);
        v[e] = tc ? g : v[e];
      }
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) v[e] = cmul(v[e], cph);
}

template <int PER, class Prog, class Op>
__device__ __forceinline__ void ew_scale(float2 (&v)[PER], const Prog& P, Op& op) {
  const float2 s = ldm(scalar_ptr(P.mats), op.mat);
#pragma unroll
  for (int e = 0; e < PER; ++e) v[e] = cmul(v[e], s);
}

// Dense ops on register bits stay inside one thread.  Register bit r is tile bit log2(NT) + r, i.e.
// bit r of the element index e; targets come sorted (R0 < R1 < R2) with matrix index bit m <->
// target m.  Local controls are tested on the group's base element (controls are never targets).
// Without local controls (kCtl false) the groups are straight-line code: no per-group branch, so
// results land in their own registers instead of being copied back at every branch join.
template <int PER, int NT, int K, int R0, int R1, int R2, bool kCtl>
__device__ __forceinline__ void reg_dense_k(float2 (&v)[PER], smat_t* M, int cm, int cv) {
  constexpr int D = 1 << K;
  constexpr int RS[3] = {R0, R1, R2};
  constexpr int TMASK = (1 << R0) | (K > 1 ? 1 << R1 : 0) | (K > 2 ? 1 << R2 : 0);
  const int tm = (int)threadIdx.x & cm;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    if (e & TMASK) continue;
    if (kCtl && (tm | ((e * NT) & cm)) != cv) continue;
    int idx[D];
#pragma unroll
    for (int m = 0; m < D; ++m) {
      int x = e;
#pragma unroll
      for (int b = 0; b < K; ++b) x |= ((m >> b) & 1) << RS[b];
      idx[m] = x;
    }
    float2 x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = v[idx[c]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = cmul(ldm(M, r * D), x[0]);
#pragma unroll
      for (int c = 1; c < D; ++c) acc = cfma(ldm(M, r * D + c), x[c], acc);
      v[idx[r]] = acc;
    }
  }
}

template <int PER, int NT>
__device__ __forceinline__ bool is_register_dense(sop_t& op) {
  constexpr int TB = __builtin_ctz(NT);
  constexpr int NR = __builtin_ctz(PER);
  if (op.kind != CQS_OP_DENSE || op.k < 1 || op.k > 3 || op.k > NR) return false;
  if (op.t[0] < TB) return false;
#pragma unroll
  for (int m = 1; m < 3; ++m)
    if (m < op.k && op.t[m] <= op.t[m - 1]) return false;
  return true;
}

template <int PER, int NT, class Prog>
__device__ __forceinline__ void reg_dense(float2 (&v)[PER], const Prog& P, sop_t& op, uint64_t base) {
  constexpr int TB = __builtin_ctz(NT);
  constexpr int NR = __builtin_ctz(PER);
  smat_t* M = scalar_ptr(P.mats) + op.mat + variant_of(op, base) * (1 << (2 * op.k));
  const int cm = op.cl_mask, cv = op.cl_val;
  const int r0 = op.t[0] - TB, r1 = op.t[1] - TB, r2 = op.t[2] - TB;
  const int code = op.k == 1 ? r0 : (op.k == 2 ? 16 + r0 * 4 + r1 : 32 + r0 * 16 + r1 * 4 + r2);
#define CQS_RK(K, A, B, C)                                  \
  if (cm == 0)                                              \
    reg_dense_k<PER, NT, K, A, B, C, false>(v, M, 0, 0);    \
  else                                                      \
    reg_dense_k<PER, NT, K, A, B, C, true>(v, M, cm, cv);
#define CQS_R1(A) \
  case A: if constexpr (NR > A) { CQS_RK(1, A, 0, 0) } return;
#define CQS_R2(A, B) \
  case 16 + A * 4 + B: if constexpr (NR > B) { CQS_RK(2, A, B, 0) } return;
#define CQS_R3(A, B, C) \
  case 32 + A * 16 + B * 4 + C: if constexpr (NR > C) { CQS_RK(3, A, B, C) } return;
  switch (code) {
    CQS_R1(0) CQS_R1(1) CQS_R1(2) CQS_R1(3)
    CQS_R2(0, 1) CQS_R2(0, 2) CQS_R2(0, 3) CQS_R2(1, 2) CQS_R2(1, 3) CQS_R2(2, 3)
    CQS_R3(0, 1, 2) CQS_R3(0, 1, 3) CQS_R3(0, 2, 3) CQS_R3(1, 2, 3)
    default: return;
  }
#undef CQS_RK
#undef CQS_R1
#undef CQS_R2
#undef CQS_R3
}

template <int PER, int NT>
__device__ __forceinline__ bool is_mixing(sop_t& op) {
  return (op.kind == CQS_OP_DENSE && !is_register_dense<PER, NT>(op)) || op.kind == CQS_OP_BLOCK;
}

// This thread's elements between registers and the LDS tile, natural layout (j = tid + e*NT).
template <int PER, int NT>
__device__ __forceinline__ void tile_to_regs(float2 (&v)[PER], const float2* tile, bool act) {
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) v[e] = tile[pad_nat<NT>(e)];
  }
}
template <int PER, int NT>
__device__ __forceinline__ void regs_to_tile(const float2 (&v)[PER], float2* tile, bool act) {
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) tile[pad_nat<NT>(e)] = v[e];
  }
}

// Element j = tid + e*NT goes to tile index sum_b bit_b(j) << dst[b].
template <int PER, int NT>
__device__ __forceinline__ void regs_to_tile_permuted(const float2 (&v)[PER], float2* tile, int L, saux_t* dst,
                                                      bool act) {
  constexpr int TB = __builtin_ctz(NT);
  constexpr int NR = __builtin_ctz(PER);
  const int tb = L < TB ? L : TB;
  int jt = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b)
    if (b < tb) jt |= ((threadIdx.x >> b) & 1) << dst[b];
  int eb[NR > 0 ? NR : 1];
#pragma unroll
  for (int r = 0; r < NR; ++r) eb[r] = 1 << dst[TB + r];
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      int j = jt;
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if ((e >> r) & 1) j |= eb[r];
      tile[pad(j)] = v[e];
    }
  }
}

// Runs the pass program.  On entry v holds this thread's elements (natural layout); on return it
// holds them after the last op, in the layout of the last PERMUTE.  The program alternates
// register phases (elementwise ops, register-bit dense ops, PERMUTE) and LDS phases (ops that mix
// elements of different threads, LDS to LDS between barriers); v is dead during an LDS phase.
template <int PER, int NT, class Prog>
__device__ __forceinline__ void run_program(float2 (&v)[PER], float2* tile, float2* mstage, const Prog& P,
                                            uint64_t base, bool act) {
  int i = 0;
  for (;;) {
    for (; i < P.op_count && !is_mixing<PER, NT>(scalar_ptr(P.ops)[i]); ++i) {
      sop_t& op = scalar_ptr(P.ops)[i];
      if (op.kind == CQS_OP_PERMUTE) {
        __syncthreads();  // every thread has read its elements before any is overwritten
        regs_to_tile_permuted<PER, NT>(v, tile, P.L, scalar_ptr(P.aux) + op.t[0], act);
        __syncthreads();
        tile_to_regs<PER, NT>(v, tile, act);
        continue;
      }
      if (!outer_controls_ok(op, base)) continue;
      if (op.kind == CQS_OP_DENSE)
        reg_dense<PER, NT>(v, P, op, base);
      else if (op.kind == CQS_OP_DIAG)
        ew_diag<PER, NT>(v, P, op, base);
      else if (op.kind == CQS_OP_PAIRPHASE)
        ew_pairs<PER, NT>(v, P, op, base);
      else if (op.kind == CQS_OP_SCALE)
        ew_scale<PER>(v, P, op);
    }
    if (i >= P.op_count) return;
    regs_to_tile<PER, NT>(v, tile, act);
    for (; i < P.op_count && is_mixing<PER, NT>(scalar_ptr(P.ops)[i]); ++i) {
      __syncthreads();
      sop_t& op = scalar_ptr(P.ops)[i];
      if (outer_controls_ok(op, base)) run_mixing_op<NT>(tile, mstage, P, op, base);
    }
    __syncthreads();
    tile_to_regs<PER, NT>(v, tile, act);
  }
}

// Threads per workgroup for a tile of 2^L amplitudes (L <= 14): 256 threads up to L = 12 (threads
// beyond 2^L idle when L < 8), then 16 elements per thread.  NT is always a multiple of 64, so the
// wave loops of the MFMA path always advance.
template <int L>
struct TileShape {
  static constexpr int NT = L <= 12 ? 256 : (1 << L) / 16;
  static constexpr int PER = L <= 8 ? 1 : (1 << L) / NT;
  static constexpr int TB = __builtin_ctz(NT);
  static constexpr int TBL = L < TB ? L : TB;  // thread bits that are tile bits
};

// Flat offsets of this thread's elements: off_lo (thread bits) + off_hi[e] (element bits).
template <int L>
__device__ __forceinline__ void tile_offsets(const int32_t* pos, uint64_t& off_lo,
                                             uint64_t (&off_hi)[TileShape<L>::PER]) {
  constexpr int TBL = TileShape<L>::TBL, PER = TileShape<L>::PER;
  off_lo = 0;
#pragma unroll
  for (int b = 0; b < TBL; ++b) off_lo |= (uint64_t)((threadIdx.x >> b) & 1) << pos[b];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < L - TBL; ++b) o |= (uint64_t)((e >> b) & 1) << pos[TBL + b];
    off_hi[e] = o;
  }
}

// With few workgroups per CU (small states) the interpreter's op reads miss every cache: each op
// waits for its descriptor and then its matrix through dependent scalar loads from memory.  With
// P.warm set, the workgroup first touches every op's descriptor and the first line of its matrix
// with vector loads, all in flight together, so that those scalar loads find them in L2.
template <int NT, class Prog>
__device__ __forceinline__ void warm_program(const Prog& P) {
  if (!P.warm) return;
  for (int i = threadIdx.x; i < P.op_count; i += NT) {
    const int32_t* o = (const int32_t*)(P.ops + i);
    const int kind = *(volatile const int32_t*)(o + 0) & 0xff;
    const int mat = *(volatile const int32_t*)(o + 2);
    if (kind != CQS_OP_PERMUTE) (void)*(volatile const float*)(P.mats + mat);
  }
}

// Zeros at the tile's amplitudes (layout lpos_out).
template <int L, int PER, int NT>
__device__ __forceinline__ void store_zero_tile(float2* __restrict__ psi, const int32_t* lpos_out, uint64_t base,
                                                bool act) {
  uint64_t off_lo, off_hi[PER];
  tile_offsets<L>(lpos_out, off_lo, off_hi);
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) psi[base + off_lo + off_hi[e]] = make_float2(0.f, 0.f);
  }
}

// One workgroup's whole pass: load the tile from HBM (layout lpos) into registers, run the
// program, store it (layout lpos_out).  `base` is the flat offset of the tile's outer bits.  With
// basis >= 0 nothing is read: the tile starts as its part of the basis state |basis>.  Tiles whose
// outer bits disagree with (zmask, zval) hold only zeros and store zeros; tiles whose outer bits
// disagree with (wmask, wval) store nothing.
template <int L, class Prog>
__device__ __forceinline__ void pass_tile(float2* __restrict__ psi, const Prog& P, const int32_t* lpos,
                                          const int32_t* lpos_out, uint64_t base, float2* lds,
                                          int64_t basis = -1, uint64_t zmask = 0, uint64_t zval = 0,
                                          uint64_t wmask = 0, uint64_t wval = 0) {
  constexpr int NT = TileShape<L>::NT, PER = TileShape<L>::PER;
  float2* tile = lds;
  float2* mstage = lds + tile_words(L);
  const bool act = L >= 8 || (int)threadIdx.x < (1 << L);
  if ((base & wmask) != wval) return;  // nobody reads this tile before the next pass stores it
  if ((base & zmask) != zval) {
    // every op of a pass acts within its tile and maps zero to zero: a tile that holds only zeros
    // stays zero, unread (uniform across the workgroup, so no barrier is skipped by part of it)
    store_zero_tile<L, PER, NT>(psi, lpos_out, base, act);
    return;
  }
  warm_program<NT>(P);
  float2 v[PER];
  {
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(lpos, off_lo, off_hi);
    if (act) {
      if (basis >= 0) {
#pragma unroll
        for (int e = 0; e < PER; ++e)
          v[e] = make_float2(base + off_lo + off_hi[e] == (uint64_t)basis ? 1.f : 0.f, 0.f);
      } else {
#pragma unroll
        for (int e = 0; e < PER; ++e) v[e] = psi[base + off_lo + off_hi[e]];
      }
    }
  }
  run_program<PER, NT>(v, tile, mstage, P, base, act);
  {
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(lpos_out, off_lo, off_hi);
    if (act) {
#pragma unroll
      for (int e = 0; e < PER; ++e) psi[base + off_lo + off_hi[e]] = v[e];
    }
  }
}

}  // namespace cqs
