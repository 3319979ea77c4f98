// Adjoint differentiation: one reverse sweep over psi (the forward state) and phi = H psi.
//
// The host supplies the reversed program: every gate U_k becomes U_k^dagger, applied to both
// tiles when it has no parameter, and split into [U_k^dagger on psi][GRAD dU_k][U_k^dagger on phi]
// when it has one, so that GRAD sees psi_{k-1} and phi_k and accumulates
//   g_k = 2 Re <phi_k | dU_k | psi_{k-1}>        (Jones & Gacon 2020, arXiv:2009.02823)
// in FP64, one partial per tile, reduced by cqs_grad_reduce.
#include "tile.hpp"

namespace cqs {

constexpr int kSelBoth = 0, kSelPsi = 1, kSelPhi = 2;

__device__ __forceinline__ int op_kind(const cqs_op& op) { return op.kind & 0xff; }
__device__ __forceinline__ int op_sel(const cqs_op& op) { return (op.kind >> 8) & 3; }

// Elementwise op on one LDS tile (this thread's elements).
template <int PER, int NT>
__device__ __forceinline__ void ew_on_tile(float2* tile, const TileProg& P, const cqs_op& op, uint64_t base,
                                           bool act) {
  if (!act) return;
  float2 v[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) v[e] = tile[threadIdx.x + e * NT];
  cqs_op o = op;
  o.kind = op_kind(op);
  if (o.kind == CQS_OP_DIAG)
    ew_diag<PER, NT>(v, P, o, base);
  else if (o.kind == CQS_OP_PAIRPHASE)
    ew_pairs<PER, NT>(v, P, o, base);
  else if (o.kind == CQS_OP_SCALE)
    ew_scale<PER>(v, P, o);
#pragma unroll
  for (int e = 0; e < PER; ++e) tile[threadIdx.x + e * NT] = v[e];
}

// sum over the tile of 2 Re(conj(phi) dU psi) for a dense dU (k <= 3) on tile bits op.t.
template <int K, int NT>
__device__ __forceinline__ double grad_dense(const float2* psi, const float2* phi, int L, const cqs_op& op,
                                             const float2* M) {
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
    for (int c = 0; c < D; ++c) x[c] = psi[j0 + pat[c]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double yr = 0.0, yi = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const float2 m = M[r * D + c];
        yr += (double)m.x * x[c].x - (double)m.y * x[c].y;
        yi += (double)m.x * x[c].y + (double)m.y * x[c].x;
      }
      const float2 f = phi[j0 + pat[r]];
      acc += (double)f.x * yr + (double)f.y * yi;  // Re(conj(f) y)
    }
  }
  return 2.0 * acc;
}

// sum over the tile of 2 Re(conj(phi_j) d(j) psi_j) for a diagonal derivative d.
template <int PER, int NT>
__device__ __forceinline__ double grad_diag(const float2* psi, const float2* phi, const TileProg& P, const cqs_op& op,
                                            uint64_t base, bool act) {
  if (!act) return 0.0;
  constexpr int TB = __builtin_ctz(NT);
  const int tid = threadIdx.x;
  int fixed = 0, emask = 0;
  for (int m = 0; m < op.k; ++m) {
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
    for (int m = 0; m < op.k; ++m)
      if ((emask >> m) & 1) idx |= ((e >> (op.t[m] - TB)) & 1) << m;
    const float2 d = Dg[idx], x = psi[tid + e * NT], f = phi[tid + e * NT];
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

template <int L>
__global__ __launch_bounds__(TileShape<L>::NT) void k_adjoint_pass(float2* __restrict__ psi, float2* __restrict__ phi,
                                                                 TileProg P, double* __restrict__ partial,
                                                                 int64_t ntiles) {
  constexpr int NT = TileShape<L>::NT, PER = TileShape<L>::PER;
  extern __shared__ float2 lds[];
  __shared__ double red[NT / 64];
  float2* tA = lds;
  float2* tB = lds + (1 << L);
  float2* mstage = lds + 2 * (1 << L);
  const bool act = L >= 8 || (int)threadIdx.x < (1 << L);
  const uint64_t base = deposit64((uint64_t)blockIdx.x, P.opos, P.nouter);
  uint64_t off_lo, off_hi[PER];
  tile_offsets<L>(P.lpos, off_lo, off_hi);
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      tA[threadIdx.x + e * NT] = psi[base + off_lo + off_hi[e]];
      tB[threadIdx.x + e * NT] = phi[base + off_lo + off_hi[e]];
    }
  }
  for (int i = 0; i < P.op_count; ++i) {
    const cqs_op& op = P.ops[i];
    __syncthreads();
    const bool ok = outer_controls_ok(op, base);
    const int kind = op_kind(op), sel = op_sel(op);
    if (kind == CQS_OP_GRAD_DENSE || kind == CQS_OP_GRAD_DIAG) {
      double g = 0.0;
      if (!ok) {
        // the gate does not act on this tile: its derivative contributes nothing here
      } else if (kind == CQS_OP_GRAD_DENSE) {
        const float2* M = P.mats + op.mat;
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
      if (threadIdx.x == 0) partial[(int64_t)op.slot * ntiles + blockIdx.x] = g;
      continue;
    }
    if (!ok) continue;
    cqs_op o = op;
    o.kind = kind;
    if (kind == CQS_OP_DENSE || kind == CQS_OP_BLOCK) {
      if (sel != kSelPhi) run_mixing_op<NT>(tA, mstage, P, o);
      if (sel == kSelBoth) __syncthreads();  // the matrix stage is shared by both tiles
      if (sel != kSelPsi) run_mixing_op<NT>(tB, mstage, P, o);
    } else {
      if (sel != kSelPhi) ew_on_tile<PER, NT>(tA, P, o, base, act);
      if (sel != kSelPsi) ew_on_tile<PER, NT>(tB, P, o, base, act);
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      psi[base + off_lo + off_hi[e]] = tA[threadIdx.x + e * NT];
      phi[base + off_lo + off_hi[e]] = tB[threadIdx.x + e * NT];
    }
  }
}

typedef void (*AdjKernel)(float2*, float2*, TileProg, double*, int64_t);
static const AdjKernel kAdjKernels[15] = {
    k_adjoint_pass<0>, k_adjoint_pass<1>, k_adjoint_pass<2>,  k_adjoint_pass<3>,  k_adjoint_pass<4>,
    k_adjoint_pass<5>, k_adjoint_pass<6>, k_adjoint_pass<7>,  k_adjoint_pass<8>,  k_adjoint_pass<9>,
    k_adjoint_pass<10>, k_adjoint_pass<11>, k_adjoint_pass<12>, k_adjoint_pass<13>, k_adjoint_pass<14>};

// out[g] = sum_t partial[g * ntiles + t]  (fixed order)
__global__ __launch_bounds__(64) void k_grad_reduce(const double* __restrict__ partial, int64_t ntiles,
                                                    double* __restrict__ out) {
  const int g = blockIdx.x;
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < ntiles; t += 64) s += partial[(int64_t)g * ntiles + t];
  s = wave_sum(s);
  if (threadIdx.x == 0) out[g] = s;
}

bool make_prog(TileProg& P, int n, const cqs_pass& ps, const cqs_op* ops, const float2* mats, const int32_t* aux);

}  // namespace cqs

using namespace cqs;

extern "C" {

int cqs_adjoint_passes(void* psi, void* phi, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                       const cqs_op* ops, const void* mats, const int32_t* aux, double* grad_partial,
                       int64_t partial_stride, void* stream) {
  CQS_REQUIRE(psi != nullptr && phi != nullptr && psi != phi, "psi/phi must be distinct non-NULL");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_passes >= 0 && (n_passes == 0 || passes != nullptr), "bad pass list");
  hipStream_t s = as_stream(stream);
  for (int i = 0; i < n_passes; ++i) {
    const cqs_pass& ps = passes[i];
    if (ps.op_count <= 0) continue;
    CQS_REQUIRE(ops != nullptr && mats != nullptr, "ops/mats are NULL");
    TileProg P;
    if (!make_prog(P, n_bits, ps, ops, (const float2*)mats, aux)) return CQS_ERR_ARG;
    for (int b = 0; b < ps.L; ++b)
      CQS_REQUIRE(ps.lpos_out[b] == ps.lpos[b], "adjoint passes cannot re-lay their tiles (PERMUTE)");
    const uint64_t tiles = 1ull << P.nouter;
    CQS_REQUIRE((int64_t)tiles <= partial_stride, "partial_stride %lld < tiles %llu", (long long)partial_stride,
                (unsigned long long)tiles);
    const int mk = ps.max_k;
    const size_t stage = mk >= 3 ? (size_t(1) << (2 * mk)) : 0;
    const size_t shmem = ((size_t(2) << ps.L) + stage) * sizeof(float2);
    CQS_REQUIRE(shmem <= 160 * 1024 - 1024, "adjoint pass needs %zu bytes of LDS", shmem);
    if (shmem > 64 * 1024)
      CQS_HIP(hipFuncSetAttribute((const void*)kAdjKernels[ps.L], hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shmem));
    const int nt = ps.L <= 12 ? 256 : (1 << ps.L) / 16;
    hipLaunchKernelGGL(kAdjKernels[ps.L], dim3((unsigned)tiles), dim3(nt), shmem, s, (float2*)psi, (float2*)phi, P,
                       grad_partial, partial_stride);
    CQS_HIP(hipGetLastError());
  }
  return CQS_OK;
}

int cqs_grad_reduce(const double* grad_partial, int32_t n_slots, int64_t partial_stride, double* out, void* stream) {
  CQS_REQUIRE(n_slots >= 0 && partial_stride >= 1, "bad sizes");
  if (n_slots == 0) return CQS_OK;
  CQS_REQUIRE(grad_partial != nullptr && out != nullptr, "NULL pointer");
  hipLaunchKernelGGL(k_grad_reduce, dim3(n_slots), dim3(64), 0, as_stream(stream), grad_partial, partial_stride, out);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

}  // extern "C"
