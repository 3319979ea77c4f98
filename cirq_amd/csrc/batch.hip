// Batched sweeps: many states (different sizes, programs and parameter values) per launch.
//
// Every launch covers a list of items sorted by their first workgroup; a workgroup finds its item
// by binary search, then runs the same per-tile code as the single-state kernels on that item's
// state, pass descriptor and matrix buffer.
#include <algorithm>

#include "adjoint.hpp"

namespace cqs {

template <class Item>
__device__ __forceinline__ int find_item(const Item* __restrict__ items, int n, int64_t wg) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)items[mid].tile_begin <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int find_terms(const cqs_batch_terms* __restrict__ items, int n, int64_t wg) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)items[mid].block_begin <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ ProgView view_of(const cqs_batch_pass& B, const cqs_batch_item& I, const cqs_op* ops,
                                            const float2* mats, const int32_t* aux) {
  ProgView P;
  P.L = B.L;
  P.op_count = B.op_count;
  P.max_k = B.max_k;
  P.mscale = 32768.f;
  P.warm = 0;
  P.ops = ops + B.op_begin;
  P.mats = mats + I.mats;
  P.aux = aux;
  return P;
}

__global__ __launch_bounds__(256) void k_batch_init(float2* __restrict__ arena, const cqs_batch_item* __restrict__ items,
                                                    int n_items, const cqs_batch_pass* __restrict__ passes,
                                                    int chunk_bits) {
  const cqs_batch_item& I = items[find_item(items, n_items, blockIdx.x)];
  const int n = passes[I.pass].n_bits;
  const int cb = n < chunk_bits ? n : chunk_bits;
  const int64_t chunk = (int64_t)blockIdx.x - I.tile_begin;
  float2* s = arena + I.state + (chunk << cb);
  for (int64_t i = threadIdx.x; i < (int64_t(1) << cb); i += 256)
    s[i] = make_float2(chunk == 0 && i == 0 ? 1.f : 0.f, 0.f);
}

template <int L>
__global__ __launch_bounds__(TileShape<L>::NT, L <= 12 ? 4 : 1) void k_batch_pass(
    float2* __restrict__ arena, const cqs_batch_item* __restrict__ items, int n_items,
    const cqs_batch_pass* __restrict__ passes, const cqs_op* __restrict__ ops, const float2* __restrict__ mats,
    const int32_t* __restrict__ aux) {
  extern __shared__ float2 lds[];
  const cqs_batch_item& I = items[find_item(items, n_items, blockIdx.x)];
  const cqs_batch_pass& B = passes[I.pass];
  const ProgView P = view_of(B, I, ops, mats, aux);
  const uint64_t base = deposit_bounded<31>((uint64_t)blockIdx.x - (uint64_t)I.tile_begin, B.opos, B.nouter);
  pass_tile<L>(arena + I.state, P, B.lpos, B.lpos_out, base, lds);
}

template <int L>
__global__ __launch_bounds__(TileShape<L>::NT) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_batch_adjoint(
    float2* __restrict__ arena, const cqs_batch_item* __restrict__ items, int n_items,
    const cqs_batch_pass* __restrict__ passes, const cqs_op* __restrict__ ops, const float2* __restrict__ mats,
    const int32_t* __restrict__ aux, double* __restrict__ partial) {
  extern __shared__ float2 lds[];
  const cqs_batch_item& I = items[find_item(items, n_items, blockIdx.x)];
  const cqs_batch_pass& B = passes[I.pass];
  const ProgView P = view_of(B, I, ops, mats, aux);
  const int64_t tile = (int64_t)blockIdx.x - I.tile_begin;
  const uint64_t base = deposit_bounded<31>((uint64_t)tile, B.opos, B.nouter);
  adjoint_tile<L>(arena + I.state, arena + I.state2, P, B.lpos, base, partial + I.partial,
                  (int64_t(1) << B.nouter) * adjoint_waves(L), tile, lds);
}

// phi[j] = sum_t c_t (-1)^popcount((j ^ x_t) & z_t) psi[j ^ x_t] over the block's indices j.
__global__ __launch_bounds__(256) void k_batch_apply(float2* __restrict__ arena,
                                                     const cqs_batch_terms* __restrict__ items, int n_items,
                                                     const uint64_t* __restrict__ xm, const uint64_t* __restrict__ zm,
                                                     const double* __restrict__ coef, int chunk_bits) {
  const cqs_batch_terms& I = items[find_terms(items, n_items, blockIdx.x)];
  const int cb = I.n_bits < chunk_bits ? I.n_bits : chunk_bits;
  const int64_t blk = (int64_t)blockIdx.x - I.block_begin;
  const float2* psi = arena + I.state;
  float2* phi = arena + I.state2;
  for (int64_t jj = threadIdx.x; jj < (int64_t(1) << cb); jj += 256) {
    const uint64_t j = ((uint64_t)blk << cb) + jj;
    double sr = 0.0, si = 0.0;
    for (int t = I.term_begin; t < I.term_begin + I.n_terms; ++t) {
      const uint64_t src = j ^ xm[t];
      const float2 v = psi[src];
      const double sg = (__popcll(src & zm[t]) & 1) ? -1.0 : 1.0;
      const double vr = sg * v.x, vi = sg * v.y, cr = coef[2 * t], ci = coef[2 * t + 1];
      sr += cr * vr - ci * vi;
      si += cr * vi + ci * vr;
    }
    phi[j] = make_float2((float)sr, (float)si);
  }
}

// One block per (item, chunk): for every term of the item, the block's share of
// sum_i (-1)^popcount(i & z) conj(psi[i ^ x]) psi[i].
__global__ __launch_bounds__(256) void k_batch_expect(const float2* __restrict__ arena,
                                                      const cqs_batch_terms* __restrict__ items, int n_items,
                                                      const uint64_t* __restrict__ xm,
                                                      const uint64_t* __restrict__ zm, double* __restrict__ partial,
                                                      int chunk_bits) {
  __shared__ double red[4][2];
  const cqs_batch_terms& I = items[find_terms(items, n_items, blockIdx.x)];
  const int cb = I.n_bits < chunk_bits ? I.n_bits : chunk_bits;
  const int64_t blk = (int64_t)blockIdx.x - I.block_begin;
  const int64_t nblk = int64_t(1) << (I.n_bits - cb);
  const float2* psi = arena + I.state;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = 0; t < I.n_terms; ++t) {
    const uint64_t x = xm[I.term_begin + t], z = zm[I.term_begin + t];
    double ar = 0.0, ai = 0.0;
    for (int64_t ii = threadIdx.x; ii < (int64_t(1) << cb); ii += 256) {
      const uint64_t i = ((uint64_t)blk << cb) + ii;
      const float2 a = psi[i], b = psi[i ^ x];
      double pr = (double)b.x * a.x + (double)b.y * a.y;  // conj(b) * a
      double pi = (double)b.x * a.y - (double)b.y * a.x;
      if (__popcll(i & z) & 1) {
        pr = -pr;
        pi = -pi;
      }
      ar += pr;
      ai += pi;
    }
    ar = wave_sum(ar);
    ai = wave_sum(ai);
    if (lane == 0) {
      red[w][0] = ar;
      red[w][1] = ai;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      const double s = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
      partial[2 * (I.out + (int64_t)t * nblk + blk) + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_segment_sum(const double* __restrict__ in, const int64_t* __restrict__ seg,
                                                    int width, double* __restrict__ out) {
  const int s = blockIdx.x;
  const int64_t b = seg[s], e = seg[s + 1];
  double a0 = 0.0, a1 = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += 64) {
    if (width == 2) {
      a0 += in[2 * i];
      a1 += in[2 * i + 1];
    } else {
      a0 += in[i];
    }
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  if (threadIdx.x == 0) {
    if (width == 2) {
      out[2 * s] = a0;
      out[2 * s + 1] = a1;
    } else {
      out[s] = a0;
    }
  }
}

typedef void (*BatchPassKernel)(float2*, const cqs_batch_item*, int, const cqs_batch_pass*, const cqs_op*,
                                const float2*, const int32_t*);
static const BatchPassKernel kBatchPass[15] = {
    k_batch_pass<0>,  k_batch_pass<1>,  k_batch_pass<2>,  k_batch_pass<3>,  k_batch_pass<4>,
    k_batch_pass<5>,  k_batch_pass<6>,  k_batch_pass<7>,  k_batch_pass<8>,  k_batch_pass<9>,
    k_batch_pass<10>, k_batch_pass<11>, k_batch_pass<12>, k_batch_pass<13>, k_batch_pass<14>};

typedef void (*BatchAdjKernel)(float2*, const cqs_batch_item*, int, const cqs_batch_pass*, const cqs_op*,
                               const float2*, const int32_t*, double*);
static const BatchAdjKernel kBatchAdj[15] = {
    k_batch_adjoint<0>,  k_batch_adjoint<1>,  k_batch_adjoint<2>,  k_batch_adjoint<3>,  k_batch_adjoint<4>,
    k_batch_adjoint<5>,  k_batch_adjoint<6>,  k_batch_adjoint<7>,  k_batch_adjoint<8>,  k_batch_adjoint<9>,
    k_batch_adjoint<10>, k_batch_adjoint<11>, k_batch_adjoint<12>, k_batch_adjoint<13>, k_batch_adjoint<14>};

// LDS of a launch: every item's pass may need the 4^3 matrix stage, so size for it.
static size_t batch_lds(int L, bool adjoint) {
  return ((adjoint ? 2 : 1) * (size_t)tile_words(L) + 64) * sizeof(float2);
}

template <class K>
static int ensure_lds(K kernel, size_t shmem) {
  if (shmem > 64 * 1024) {
    CQS_REQUIRE(shmem <= 160 * 1024 - 1024, "batched launch needs %zu bytes of LDS", shmem);
    CQS_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  }
  return CQS_OK;
}

}  // namespace cqs

using namespace cqs;

extern "C" {

int cqs_batch_validate(const cqs_batch_pass* passes, int32_t n_passes, int32_t n_ops, const cqs_batch_item* items,
                       int32_t n_items, int32_t L, int64_t arena_elems, int64_t mats_elems, int32_t adjoint) {
  CQS_REQUIRE(passes != nullptr && items != nullptr && n_passes > 0 && n_items > 0, "empty batch");
  CQS_REQUIRE(L >= 0 && L <= 14, "L=%d out of range", L);
  for (int p = 0; p < n_passes; ++p) {
    const cqs_batch_pass& B = passes[p];
    CQS_REQUIRE(B.n_bits >= B.L && B.n_bits <= CQS_MAX_BITS && B.L >= 0 && B.L <= 14, "pass %d: bad sizes", p);
    CQS_REQUIRE(B.nouter == B.n_bits - B.L, "pass %d: nouter != n_bits - L", p);
    CQS_REQUIRE(B.op_begin >= 0 && B.op_count >= 0 && B.op_begin + B.op_count <= n_ops, "pass %d: ops out of range",
                p);
    CQS_REQUIRE(B.max_k >= 0 && B.max_k <= 3, "pass %d: max_k=%d (batched passes stage at most 3-bit matrices)", p,
                B.max_k);
    uint64_t in = 0, out = 0;
    for (int b = 0; b < B.L; ++b) {
      const int a = B.lpos[b], c = B.lpos_out[b];
      CQS_REQUIRE(a >= 0 && a < B.n_bits && !((in >> a) & 1), "pass %d: bad lpos", p);
      CQS_REQUIRE(c >= 0 && c < B.n_bits && !((out >> c) & 1), "pass %d: bad lpos_out", p);
      in |= 1ull << a;
      out |= 1ull << c;
    }
    CQS_REQUIRE(in == out, "pass %d: lpos_out is not a permutation of lpos", p);
    int k = 0;
    for (int q = 0; q < B.n_bits; ++q)
      if (!((in >> q) & 1)) {
        CQS_REQUIRE(B.opos[k] == q, "pass %d: opos must list the outer positions in ascending order", p);
        ++k;
      }
  }
  int64_t next = 0;
  for (int i = 0; i < n_items; ++i) {
    const cqs_batch_item& I = items[i];
    CQS_REQUIRE(I.pass >= 0 && I.pass < n_passes, "item %d: pass index out of range", i);
    const cqs_batch_pass& B = passes[I.pass];
    CQS_REQUIRE(B.L == L, "item %d: its pass has L=%d, launch has L=%d", i, B.L, L);
    CQS_REQUIRE(I.tile_begin == next, "item %d: tile_begin %d, expected %lld", i, I.tile_begin, (long long)next);
    next += int64_t(1) << B.nouter;
    CQS_REQUIRE(next < (int64_t(1) << 31), "too many workgroups in one launch");
    const int64_t size = int64_t(1) << B.n_bits;
    CQS_REQUIRE(I.state >= 0 && I.state + size <= arena_elems, "item %d: state outside the arena", i);
    if (adjoint) {
      CQS_REQUIRE(I.state2 >= 0 && I.state2 + size <= arena_elems, "item %d: phi outside the arena", i);
      CQS_REQUIRE(I.state2 >= I.state + size || I.state >= I.state2 + size, "item %d: psi and phi overlap", i);
      CQS_REQUIRE(I.partial >= 0, "item %d: negative partial offset", i);
    }
    CQS_REQUIRE(I.mats >= 0 && I.mats < mats_elems && (I.mats % 2) == 0, "item %d: bad mats offset", i);
  }
  return CQS_OK;
}

int cqs_batch_init(void* arena, const cqs_batch_item* items, int32_t n_items, int64_t n_chunks,
                   const cqs_batch_pass* passes, int32_t chunk_bits, void* stream) {
  CQS_REQUIRE(arena && items && passes, "NULL pointer");
  CQS_REQUIRE(n_items > 0 && n_chunks > 0 && n_chunks < (int64_t(1) << 31), "bad sizes");
  CQS_REQUIRE(chunk_bits >= 0 && chunk_bits <= 20, "chunk_bits out of range");
  hipLaunchKernelGGL(k_batch_init, dim3((unsigned)n_chunks), dim3(256), 0, as_stream(stream), (float2*)arena, items,
                     n_items, passes, chunk_bits);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_batch_passes(void* arena, const cqs_batch_item* items, int32_t n_items, int64_t n_tiles, int32_t L,
                     const cqs_batch_pass* passes, const cqs_op* ops, const void* mats, const int32_t* aux,
                     void* stream) {
  CQS_REQUIRE(arena && items && passes && ops && mats, "NULL pointer");
  CQS_REQUIRE(L >= 0 && L <= 14, "L=%d out of range", L);
  CQS_REQUIRE(n_items > 0 && n_tiles > 0 && n_tiles < (int64_t(1) << 31), "bad sizes");
  const size_t shmem = batch_lds(L, false);
  if (int rc = ensure_lds(kBatchPass[L], shmem)) return rc;
  const int nt = L <= 12 ? 256 : (1 << L) / 16;
  hipLaunchKernelGGL(kBatchPass[L], dim3((unsigned)n_tiles), dim3(nt), shmem, as_stream(stream), (float2*)arena,
                     items, n_items, passes, ops, (const float2*)mats, aux);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_batch_adjoint(void* arena, const cqs_batch_item* items, int32_t n_items, int64_t n_tiles, int32_t L,
                      const cqs_batch_pass* passes, const cqs_op* ops, const void* mats, const int32_t* aux,
                      double* partial, void* stream) {
  CQS_REQUIRE(arena && items && passes && ops && mats && partial, "NULL pointer");
  CQS_REQUIRE(L >= 0 && L <= 13, "L=%d out of range", L);
  CQS_REQUIRE(n_items > 0 && n_tiles > 0 && n_tiles < (int64_t(1) << 31), "bad sizes");
  const size_t shmem = batch_lds(L, true);
  if (int rc = ensure_lds(kBatchAdj[L], shmem)) return rc;
  const int nt = L <= 12 ? 256 : (1 << L) / 16;
  hipLaunchKernelGGL(kBatchAdj[L], dim3((unsigned)n_tiles), dim3(nt), shmem, as_stream(stream), (float2*)arena,
                     items, n_items, passes, ops, (const float2*)mats, aux, partial);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_batch_apply_paulisum(void* arena, const cqs_batch_terms* items, int32_t n_items, int64_t n_blocks,
                             const uint64_t* xmask, const uint64_t* zmask, const double* coef, int32_t chunk_bits,
                             void* stream) {
  CQS_REQUIRE(arena && items && xmask && zmask && coef, "NULL pointer");
  CQS_REQUIRE(n_items > 0 && n_blocks > 0 && n_blocks < (int64_t(1) << 31), "bad sizes");
  CQS_REQUIRE(chunk_bits >= 0 && chunk_bits <= 20, "chunk_bits out of range");
  hipLaunchKernelGGL(k_batch_apply, dim3((unsigned)n_blocks), dim3(256), 0, as_stream(stream), (float2*)arena, items,
                     n_items, xmask, zmask, coef, chunk_bits);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_batch_expect(const void* arena, const cqs_batch_terms* items, int32_t n_items, int64_t n_blocks,
                     const uint64_t* xmask, const uint64_t* zmask, double* partial, int32_t chunk_bits,
                     void* stream) {
  CQS_REQUIRE(arena && items && xmask && zmask && partial, "NULL pointer");
  CQS_REQUIRE(n_items > 0 && n_blocks > 0 && n_blocks < (int64_t(1) << 31), "bad sizes");
  CQS_REQUIRE(chunk_bits >= 0 && chunk_bits <= 20, "chunk_bits out of range");
  hipLaunchKernelGGL(k_batch_expect, dim3((unsigned)n_blocks), dim3(256), 0, as_stream(stream),
                     (const float2*)arena, items, n_items, xmask, zmask, partial, chunk_bits);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_segment_sum(const double* in, const int64_t* seg, int32_t n_seg, int32_t width, double* out, void* stream) {
  CQS_REQUIRE(in && seg && out, "NULL pointer");
  CQS_REQUIRE(width == 1 || width == 2, "width must be 1 or 2");
  if (n_seg <= 0) return CQS_OK;
  hipLaunchKernelGGL(k_segment_sum, dim3((unsigned)n_seg), dim3(64), 0, as_stream(stream), in, seg, width, out);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

}  // extern "C"
