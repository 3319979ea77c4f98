// Batched trajectories: R copies of a state, one per repetition, stored one after the other
// (copy r at amps + r * 2^n_bits).  Deterministic stretches of a circuit run on all copies at
// once as one pass program over n_bits + log2(R) bits (the copy index is a block of outer bits no
// op touches); what differs per copy is here: the projection after a measurement, the matrix a
// mixture or Kraus channel picked, and the Kraus weights ||K_i psi_r||^2 that the pick depends on.
// Grids put the copy in blockIdx.y, so per-copy values are uniform within a workgroup.
#include "common.hpp"

namespace cqs {

__device__ __forceinline__ uint64_t insert0_64(uint64_t x, int b) {
  return ((x >> b) << (b + 1)) | (x & ((1ull << b) - 1ull));
}

struct SelTargets {
  int t[3];   // flat bit of matrix index bit m
  int ts[3];  // the same bits in ascending order
};

// Flat index of group g's first element (zeros at the targets) and its 2^K members.
template <int K>
__device__ __forceinline__ void group_indices(uint64_t g, const SelTargets& T, uint64_t (&idx)[1 << K]) {
  uint64_t j0 = g;
#pragma unroll
  for (int s = 0; s < K; ++s) j0 = insert0_64(j0, T.ts[s]);
#pragma unroll
  for (int m = 0; m < (1 << K); ++m) {
    uint64_t j = j0;
#pragma unroll
    for (int b = 0; b < K; ++b) j |= (uint64_t)((m >> b) & 1) << T.t[b];
    idx[m] = j;
  }
}

// amps_r[i] = ((i & mask) == values[r]) ? amps_r[i] * scales[r] : 0
__global__ __launch_bounds__(256) void k_collapse_batch(float2* __restrict__ amps, uint64_t per, uint64_t mask,
                                                        const uint64_t* __restrict__ values,
                                                        const float* __restrict__ scales) {
  float2* s = amps + (uint64_t)blockIdx.y * per;
  const uint64_t v = values[blockIdx.y];
  const float sc = scales[blockIdx.y];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < per; i += stride) {
    const float2 x = s[i];
    s[i] = ((i & mask) == v) ? make_float2(x.x * sc, x.y * sc) : make_float2(0.f, 0.f);
  }
}

// Copy r: every group of 2^K amplitudes on the targets <- scales[r] * (M_{select[r]} x), the
// product accumulated from zero in column order.  A copy whose select is out of range is left
// as it is (the host checks the selection before it uploads it).
template <int K>
__global__ __launch_bounds__(256) void k_select_batch(float2* __restrict__ amps, uint64_t per, SelTargets T,
                                                      const float2* __restrict__ mats, int n_mats,
                                                      const int32_t* __restrict__ select,
                                                      const float* __restrict__ scales) {
  constexpr int D = 1 << K;
  const int sel = select[blockIdx.y];
  if (sel < 0 || sel >= n_mats) return;
  const float2* M = mats + (uint64_t)sel * D * D;
  const float sc = scales[blockIdx.y];
  float2* s = amps + (uint64_t)blockIdx.y * per;
  const uint64_t groups = per >> K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    uint64_t idx[D];
    group_indices<K>(g, T, idx);
    float2 x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = s[idx[c]];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < D; ++c) acc = cfma(M[r * D + c], x[c], acc);
      s[idx[r]] = make_float2(acc.x * sc, acc.y * sc);
    }
  }
}

// Several per-copy selections in one sweep of the copies: a run of mixtures whose targets lie in
// a union of at most 4 bits.  Each thread loads the 2^u amplitudes of one group of the union into
// registers, applies the run's ops in order (op i: copy r's matrix select[row_i][r] of op i's set,
// on union positions p0 (and p1) of the group), and stores them: one read and one write of the
// copies for the whole run instead of one per op.
constexpr int kRunOps = 16;
struct SelRun {
  int u;                 // bits in the union (1..4)
  int bits[4];           // their flat positions, ascending
  int nops;
  int k[kRunOps];        // 1 or 2 targets
  int p0[kRunOps];       // union position of matrix index bit 0
  int p1[kRunOps];       // union position of matrix index bit 1 (k = 2)
  int mat[kRunOps];      // offset (complex entries) of the op's matrices in mats
  int n_mats[kRunOps];
  int row[kRunOps];      // row of the op in the selection table (one int32 per copy)
};

// Thread t's 16 amplitudes live in LDS at xs[e * 256 + t] (consecutive threads, consecutive
// addresses), so the ops below index them by run-time union positions.
__device__ __forceinline__ void run_op1(float2* xs, int D, int p0, const float2* M) {
  const int t = threadIdx.x, b0 = 1 << p0;
  for (int e = 0; e < D; ++e) {
    if (e & b0) continue;
    const float2 a = xs[e * 256 + t], b = xs[(e | b0) * 256 + t];
    xs[e * 256 + t] = cfma(M[1], b, cmul(M[0], a));
    xs[(e | b0) * 256 + t] = cfma(M[3], b, cmul(M[2], a));
  }
}

__device__ __forceinline__ void run_op2(float2* xs, int D, int p0, int p1, const float2* M) {
  const int t = threadIdx.x, b0 = 1 << p0, b1 = 1 << p1;
  for (int e = 0; e < D; ++e) {
    if (e & (b0 | b1)) continue;
    const int id[4] = {e, e | b0, e | b1, e | b0 | b1};  // matrix index bit 0 <-> p0, bit 1 <-> p1
    float2 v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = xs[id[c] * 256 + t];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc = cfma(M[r * 4 + c], v[c], acc);
      xs[id[r] * 256 + t] = acc;
    }
  }
}

__global__ __launch_bounds__(256) void k_select_run(float2* __restrict__ amps, uint64_t per, SelRun R,
                                                    const float2* __restrict__ mats,
                                                    const int32_t* __restrict__ table, int n_states) {
  // the op fields go to LDS (compile-time indices here), so the op loop below reads them with a
  // run-time index without copying the argument block to scratch memory
  __shared__ int sk[kRunOps], sp0[kRunOps], sp1[kRunOps], smat[kRunOps], snm[kRunOps], srow[kRunOps];
  __shared__ float2 xs[16 * 256];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kRunOps; ++i) {
      sk[i] = R.k[i];
      sp0[i] = R.p0[i];
      sp1[i] = R.p1[i];
      smat[i] = R.mat[i];
      snm[i] = R.n_mats[i];
      srow[i] = R.row[i];
    }
  }
  __syncthreads();
  const int copy = blockIdx.y;
  float2* s = amps + (uint64_t)copy * per;
  const int D = 1 << R.u;
  const int t = threadIdx.x;
  uint64_t pb[4];  // flat offsets of the union bits
#pragma unroll
  for (int b = 0; b < 4; ++b) pb[b] = b < R.u ? 1ull << R.bits[b] : 0;
  const uint64_t groups = per >> R.u;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + t; g < groups; g += stride) {
    uint64_t j0 = g;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (b < R.u) j0 = insert0_64(j0, R.bits[b]);
    for (int e = 0; e < D; ++e) {
      const uint64_t j = j0 | ((e & 1) ? pb[0] : 0) | ((e & 2) ? pb[1] : 0) | ((e & 4) ? pb[2] : 0) |
                         ((e & 8) ? pb[3] : 0);
      xs[e * 256 + t] = s[j];
    }
    for (int i = 0; i < R.nops; ++i) {
      const int sel = table[(size_t)srow[i] * n_states + copy];
      if (sel < 0 || sel >= snm[i]) continue;
      const float2* M = mats + smat[i] + (size_t)sel * (sk[i] == 1 ? 4 : 16);
      if (sk[i] == 1)
        run_op1(xs, D, sp0[i], M);
      else
        run_op2(xs, D, sp0[i], sp1[i], M);
    }
    for (int e = 0; e < D; ++e) {
      const uint64_t j = j0 | ((e & 1) ? pb[0] : 0) | ((e & 2) ? pb[1] : 0) | ((e & 4) ? pb[2] : 0) |
                         ((e & 8) ? pb[3] : 0);
      s[j] = xs[e * 256 + t];
    }
  }
}

constexpr int kMaxKraus = 16;

// partial[(r * gridDim.x + blockIdx.x) * n_mats + i] = sum over this workgroup's groups of copy r
// of ||K_i x||^2, in FP64 (fixed order: per-thread sums, then a wave tree, then the 4 waves).
template <int K>
__global__ __launch_bounds__(256) void k_kraus_weights(const float2* __restrict__ amps, uint64_t per, SelTargets T,
                                                       const float2* __restrict__ mats, int n_mats,
                                                       double* __restrict__ partial) {
  constexpr int D = 1 << K;
  __shared__ float2 sm[kMaxKraus * D * D];
  __shared__ double red[4][kMaxKraus];
  for (int i = threadIdx.x; i < n_mats * D * D; i += blockDim.x) sm[i] = mats[i];
  __syncthreads();
  const float2* s = amps + (uint64_t)blockIdx.y * per;
  const uint64_t groups = per >> K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double acc[kMaxKraus];
#pragma unroll
  for (int i = 0; i < kMaxKraus; ++i) acc[i] = 0.0;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    uint64_t idx[D];
    group_indices<K>(g, T, idx);
    float2 x[D];
#pragma unroll
    for (int c = 0; c < D; ++c) x[c] = s[idx[c]];
#pragma unroll
    for (int i = 0; i < kMaxKraus; ++i) {
      if (i >= n_mats) break;
      const float2* M = sm + i * D * D;
      double w = 0.0;
#pragma unroll
      for (int r = 0; r < D; ++r) {
        float2 y = make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < D; ++c) y = cfma(M[r * D + c], x[c], y);
        w += (double)y.x * (double)y.x + (double)y.y * (double)y.y;
      }
      acc[i] += w;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxKraus; ++i) {
    if (i >= n_mats) break;
    const double v = wave_sum(acc[i]);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < n_mats) {
    const int i = threadIdx.x;
    partial[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * n_mats + i] =
        ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
  }
}

// Two single-qubit Kraus channels in a row, A then B, in one sweep of the copies: copy r applies
// A's chosen operator (matsA[selA[r]] times scaleA[r], as k_select_batch<1> does) on bit a, stores
// the result, and accumulates ||K_i x||^2 of B's operators on bit b for the weights of B's draw
// (partial sums as k_kraus_weights writes them).  Groups span the union of a and b.
__global__ __launch_bounds__(256) void k_select_then_weights(float2* __restrict__ amps, uint64_t per, int bit_a,
                                                             const float2* __restrict__ matsA, int n_matsA,
                                                             const int32_t* __restrict__ selA,
                                                             const float* __restrict__ scaleA, int bit_b,
                                                             const float2* __restrict__ matsB, int n_matsB,
                                                             double* __restrict__ partial) {
  __shared__ float2 sm[kMaxKraus * 4];
  __shared__ double red[4][kMaxKraus];
  for (int i = threadIdx.x; i < n_matsB * 4; i += blockDim.x) sm[i] = matsB[i];
  __syncthreads();
  const int sel = selA[blockIdx.y];
  const bool apply = sel >= 0 && sel < n_matsA;
  const float2* MA = matsA + (apply ? sel : 0) * 4;
  const float sc = scaleA[blockIdx.y];
  float2* s = amps + (uint64_t)blockIdx.y * per;
  const bool same = bit_a == bit_b;
  const int lo = bit_a < bit_b ? bit_a : bit_b, hi = bit_a < bit_b ? bit_b : bit_a;
  const uint64_t A = 1ull << bit_a, B = 1ull << bit_b;
  const uint64_t groups = same ? per >> 1 : per >> 2;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double acc[kMaxKraus];
#pragma unroll
  for (int i = 0; i < kMaxKraus; ++i) acc[i] = 0.0;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    const uint64_t j0 = same ? insert0_64(g, lo) : insert0_64(insert0_64(g, lo), hi);
    // x[e]: e bit 0 <-> bit a, e bit 1 <-> bit b (unused when a == b)
    const int ne = same ? 2 : 4;
    uint64_t id[4];
    float2 x[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      id[e] = j0 | ((e & 1) ? A : 0) | ((e & 2) ? B : 0);
      x[e] = e < ne ? s[id[e]] : make_float2(0.f, 0.f);
    }
    if (apply) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // the pairs along a: (x[2h], x[2h + 1])
        if (h == 1 && same) break;
        const float2 a0 = x[2 * h], a1 = x[2 * h + 1];
        float2 y0 = cfma(MA[0], a0, make_float2(0.f, 0.f)), y1 = cfma(MA[2], a0, make_float2(0.f, 0.f));
        y0 = cfma(MA[1], a1, y0);
        y1 = cfma(MA[3], a1, y1);
        x[2 * h] = make_float2(y0.x * sc, y0.y * sc);
        x[2 * h + 1] = make_float2(y1.x * sc, y1.y * sc);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (e < ne) s[id[e]] = x[e];
    }
    // B's weights: pairs along b (along a again when a == b)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && same) break;
      const float2 b0 = same ? x[0] : x[h], b1 = same ? x[1] : x[h + 2];
#pragma unroll
      for (int i = 0; i < kMaxKraus; ++i) {
        if (i >= n_matsB) break;
        const float2* M = sm + i * 4;
        const float2 z0 = cfma(M[1], b1, cfma(M[0], b0, make_float2(0.f, 0.f)));
        const float2 z1 = cfma(M[3], b1, cfma(M[2], b0, make_float2(0.f, 0.f)));
        acc[i] += ((double)z0.x * (double)z0.x + (double)z0.y * (double)z0.y) +
                  ((double)z1.x * (double)z1.x + (double)z1.y * (double)z1.y);
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxKraus; ++i) {
    if (i >= n_matsB) break;
    const double v = wave_sum(acc[i]);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < n_matsB) {
    const int i = threadIdx.x;
    partial[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * n_matsB + i] =
        ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
  }
}

// out[r * n_mats + i] = sum over b of partial[(r * nblk + b) * n_mats + i], b ascending.
__global__ void k_kraus_finish(const double* __restrict__ partial, int nblk, int n_mats, int n_states,
                               double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_states * n_mats) return;
  const int r = e / n_mats, i = e % n_mats;
  double v = 0.0;
  for (int b = 0; b < nblk; ++b) v += partial[((uint64_t)r * nblk + b) * n_mats + i];
  out[e] = v;
}

// Copy r's Kraus operator, as the walk of cirq-core/cirq/sim/state_vector_simulation_state.py:224-250
// picks it from p = u[r]: p -= w_i in order until p < 0; the heaviest operator seen so far (the
// first of equals) when the walk ends with p >= 0 or on a zero weight.  select[r] and
// scale[r] = 1 / sqrt(w) feed k_select_batch.
__global__ void k_kraus_choose(const double* __restrict__ weights, const double* __restrict__ u, int n_states,
                               int n_mats, int32_t* __restrict__ select, float* __restrict__ scale) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_states) return;
  const double* w = weights + (size_t)r * n_mats;
  double p = u[r], fb_w = 0.0, last = 0.0;
  int fb = 0, sel = -1;
  for (int i = 0; i < n_mats; ++i) {
    last = w[i];
    if (last > fb_w) {
      fb = i;
      fb_w = last;
    }
    p -= last;
    if (p < 0) {
      sel = i;
      break;
    }
  }
  double wsel = last;
  if (sel < 0 || last == 0.0) {
    sel = fb;
    wsel = fb_w;
  }
  select[r] = sel;
  scale[r] = (float)(1.0 / sqrt(wsel));
}

// Workgroups per copy: enough for about 4096 workgroups in all, at most one per 256 groups.
static int blocks_per_copy(uint64_t groups, int n_states) {
  uint64_t want = (groups + 255) / 256;
  uint64_t cap = (uint64_t)(4096 / (n_states > 0 ? n_states : 1));
  if (cap < 1) cap = 1;
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  return (int)want;
}

static int make_targets(SelTargets& T, int n_bits, int k, const int32_t* targets) {
  CQS_REQUIRE(k >= 1 && k <= 3, "batched channel ops act on 1-3 qubits (got %d)", k);
  CQS_REQUIRE(targets != nullptr, "targets is NULL");
  uint64_t seen = 0;
  for (int m = 0; m < k; ++m) {
    const int b = targets[m];
    CQS_REQUIRE(b >= 0 && b < n_bits && !((seen >> b) & 1ull), "bad or repeated target bit %d", b);
    seen |= 1ull << b;
    T.t[m] = b;
    T.ts[m] = b;
  }
  for (int a = 1; a < k; ++a)
    for (int b = a; b > 0 && T.ts[b - 1] > T.ts[b]; --b) {
      const int x = T.ts[b - 1];
      T.ts[b - 1] = T.ts[b];
      T.ts[b] = x;
    }
  for (int m = k; m < 3; ++m) T.t[m] = T.ts[m] = 0;
  return CQS_OK;
}

}  // namespace cqs

using namespace cqs;

extern "C" {

int cqs_collapse_batch(void* amps, int32_t n_bits, int32_t n_states, uint64_t mask, const uint64_t* values,
                       const float* scales, void* stream) {
  CQS_REQUIRE(amps != nullptr && values != nullptr && scales != nullptr, "NULL pointer");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_states >= 1 && n_states <= 65535, "n_states=%d out of range (1..65535)", n_states);
  const uint64_t per = 1ull << n_bits;
  const int gx = blocks_per_copy(per, n_states);
  hipLaunchKernelGGL(k_collapse_batch, dim3(gx, n_states), dim3(256), 0, as_stream(stream), (float2*)amps, per, mask,
                     values, scales);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_apply_select_batch(void* amps, int32_t n_bits, int32_t n_states, int32_t k, const int32_t* targets,
                           const void* mats, int32_t n_mats, const int32_t* select, const float* scales,
                           void* stream) {
  CQS_REQUIRE(amps != nullptr && mats != nullptr && select != nullptr && scales != nullptr, "NULL pointer");
  CQS_REQUIRE(n_bits >= 1 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_states >= 1 && n_states <= 65535, "n_states=%d out of range (1..65535)", n_states);
  CQS_REQUIRE(n_mats >= 1, "n_mats=%d", n_mats);
  SelTargets T;
  if (int rc = make_targets(T, n_bits, k, targets)) return rc;
  const uint64_t per = 1ull << n_bits;
  const int gx = blocks_per_copy(per >> k, n_states);
  const dim3 grid(gx, n_states);
  hipStream_t s = as_stream(stream);
  const float2* M = (const float2*)mats;
  switch (k) {
    case 1: hipLaunchKernelGGL(k_select_batch<1>, grid, dim3(256), 0, s, (float2*)amps, per, T, M, n_mats, select, scales); break;
    case 2: hipLaunchKernelGGL(k_select_batch<2>, grid, dim3(256), 0, s, (float2*)amps, per, T, M, n_mats, select, scales); break;
    default: hipLaunchKernelGGL(k_select_batch<3>, grid, dim3(256), 0, s, (float2*)amps, per, T, M, n_mats, select, scales); break;
  }
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

size_t cqs_kraus_workspace_bytes(int32_t n_bits, int32_t n_states, int32_t k, int32_t n_mats) {
  if (n_bits < 1 || n_states < 1 || k < 1 || k > 3 || n_mats < 1) return 0;
  const int gx = blocks_per_copy((1ull << n_bits) >> k, n_states);
  return (size_t)n_states * gx * n_mats * sizeof(double);
}

int cqs_kraus_weights_batch(const void* amps, int32_t n_bits, int32_t n_states, int32_t k, const int32_t* targets,
                            const void* mats, int32_t n_mats, double* out, void* workspace, void* stream) {
  CQS_REQUIRE(amps != nullptr && mats != nullptr && out != nullptr && workspace != nullptr, "NULL pointer");
  CQS_REQUIRE(n_bits >= 1 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_states >= 1 && n_states <= 65535, "n_states=%d out of range (1..65535)", n_states);
  CQS_REQUIRE(n_mats >= 1 && n_mats <= kMaxKraus, "n_mats=%d out of range (1..%d)", n_mats, kMaxKraus);
  SelTargets T;
  if (int rc = make_targets(T, n_bits, k, targets)) return rc;
  const uint64_t per = 1ull << n_bits;
  const int gx = blocks_per_copy(per >> k, n_states);
  const dim3 grid(gx, n_states);
  hipStream_t s = as_stream(stream);
  const float2* M = (const float2*)mats;
  double* partial = (double*)workspace;
  switch (k) {
    case 1: hipLaunchKernelGGL(k_kraus_weights<1>, grid, dim3(256), 0, s, (const float2*)amps, per, T, M, n_mats, partial); break;
    case 2: hipLaunchKernelGGL(k_kraus_weights<2>, grid, dim3(256), 0, s, (const float2*)amps, per, T, M, n_mats, partial); break;
    default: hipLaunchKernelGGL(k_kraus_weights<3>, grid, dim3(256), 0, s, (const float2*)amps, per, T, M, n_mats, partial); break;
  }
  CQS_HIP(hipGetLastError());
  const int total = n_states * n_mats;
  hipLaunchKernelGGL(k_kraus_finish, dim3((total + 255) / 256), dim3(256), 0, s, (const double*)partial, gx, n_mats,
                     (int)n_states, out);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}


int cqs_kraus_choose_batch(const double* weights, const double* uniforms, int32_t n_states, int32_t n_mats,
                           int32_t* select, float* scales, void* stream) {
  CQS_REQUIRE(weights != nullptr && uniforms != nullptr && select != nullptr && scales != nullptr, "NULL pointer");
  CQS_REQUIRE(n_states >= 1 && n_mats >= 1 && n_mats <= kMaxKraus, "bad sizes");
  hipLaunchKernelGGL(k_kraus_choose, dim3((n_states + 255) / 256), dim3(256), 0, as_stream(stream), weights, uniforms,
                     (int)n_states, (int)n_mats, select, scales);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}


int cqs_select_run_batch(void* amps, int32_t n_bits, int32_t n_states, int32_t n_ops, const int32_t* op_k,
                         const int32_t* op_targets, const int32_t* op_mat, const int32_t* op_n_mats,
                         const int32_t* op_row, const void* mats, int64_t mat_entries, const int32_t* table,
                         int32_t table_rows, void* stream) {
  CQS_REQUIRE(amps != nullptr && mats != nullptr && table != nullptr, "NULL pointer");
  CQS_REQUIRE(op_k && op_targets && op_mat && op_n_mats && op_row, "NULL op arrays");
  CQS_REQUIRE(n_bits >= 1 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_states >= 1 && n_states <= 65535, "n_states=%d out of range (1..65535)", n_states);
  CQS_REQUIRE(n_ops >= 1 && n_ops <= kRunOps, "n_ops=%d out of range (1..%d)", n_ops, kRunOps);
  SelRun R;
  R.u = 0;
  R.nops = n_ops;
  uint64_t un = 0;
  for (int i = 0; i < n_ops; ++i) {
    CQS_REQUIRE(op_k[i] == 1 || op_k[i] == 2, "op %d acts on %d bits (1 or 2)", i, op_k[i]);
    CQS_REQUIRE(op_n_mats[i] >= 1 && op_mat[i] >= 0 && op_row[i] >= 0 && op_row[i] < table_rows,
                "op %d: bad matrix set or table row", i);
    const int64_t d = op_k[i] == 1 ? 2 : 4;
    CQS_REQUIRE((int64_t)op_mat[i] + (int64_t)op_n_mats[i] * d * d <= mat_entries, "op %d: matrices past the end", i);
    for (int m = 0; m < op_k[i]; ++m) {
      const int b = op_targets[2 * i + m];
      CQS_REQUIRE(b >= 0 && b < n_bits, "op %d: target bit %d out of range", i, b);
      un |= 1ull << b;
    }
    CQS_REQUIRE(op_k[i] == 1 || op_targets[2 * i] != op_targets[2 * i + 1], "op %d: repeated target", i);
  }
  for (int b = 0; b < n_bits; ++b)
    if ((un >> b) & 1ull) {
      CQS_REQUIRE(R.u < 4, "the run's targets span more than 4 bits");
      R.bits[R.u++] = b;
    }
  for (int b = R.u; b < 4; ++b) R.bits[b] = 0;
  auto pos = [&](int bit) {
    for (int q = 0; q < R.u; ++q)
      if (R.bits[q] == bit) return q;
    return 0;
  };
  for (int i = 0; i < kRunOps; ++i) {
    const bool on = i < n_ops;
    R.k[i] = on ? op_k[i] : 1;
    R.p0[i] = on ? pos(op_targets[2 * i]) : 0;
    R.p1[i] = on && op_k[i] == 2 ? pos(op_targets[2 * i + 1]) : 0;
    R.mat[i] = on ? op_mat[i] : 0;
    R.n_mats[i] = on ? op_n_mats[i] : 1;
    R.row[i] = on ? op_row[i] : 0;
  }
  const uint64_t per = 1ull << n_bits;
  const int gx = blocks_per_copy(per >> R.u, n_states);
  hipLaunchKernelGGL(k_select_run, dim3(gx, n_states), dim3(256), 0, as_stream(stream), (float2*)amps, per, R,
                     (const float2*)mats, table, (int)n_states);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}


int cqs_select_then_kraus_weights(void* amps, int32_t n_bits, int32_t n_states, int32_t bit_a, const void* mats_a,
                                  int32_t n_mats_a, const int32_t* select_a, const float* scales_a, int32_t bit_b,
                                  const void* mats_b, int32_t n_mats_b, double* out, void* workspace, void* stream) {
  CQS_REQUIRE(amps && mats_a && select_a && scales_a && mats_b && out && workspace, "NULL pointer");
  CQS_REQUIRE(n_bits >= 2 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_states >= 1 && n_states <= 65535, "n_states=%d out of range (1..65535)", n_states);
  CQS_REQUIRE(bit_a >= 0 && bit_a < n_bits && bit_b >= 0 && bit_b < n_bits, "bad target bits");
  CQS_REQUIRE(n_mats_a >= 1 && n_mats_b >= 1 && n_mats_b <= kMaxKraus, "bad operator counts");
  const uint64_t per = 1ull << n_bits;
  // the same workgroup count as cqs_kraus_weights_batch for a one-qubit channel, so the
  // workspace size and cqs_kraus_workspace_bytes agree
  const int gx = blocks_per_copy(per >> 1, n_states);
  hipStream_t s = as_stream(stream);
  double* partial = (double*)workspace;
  hipLaunchKernelGGL(k_select_then_weights, dim3(gx, n_states), dim3(256), 0, s, (float2*)amps, per, (int)bit_a,
                     (const float2*)mats_a, (int)n_mats_a, select_a, scales_a, (int)bit_b, (const float2*)mats_b,
                     (int)n_mats_b, partial);
  CQS_HIP(hipGetLastError());
  const int total = n_states * n_mats_b;
  hipLaunchKernelGGL(k_kraus_finish, dim3((total + 255) / 256), dim3(256), 0, s, (const double*)partial, gx, n_mats_b,
                     (int)n_states, out);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

}  // extern "C"
