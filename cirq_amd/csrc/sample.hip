// Measurement sampling, bit-exact with NumPy's RandomState.choice as Cirq calls it.
//
// Cirq (cirq-core/cirq/sim/state_vector.py:219-226, cirq-core/cirq/sim/simulation_utils.py:62-64):
//   q = (psi * conj(psi)).real                  float32; NumPy evaluates fma(a, a, b*b)
//   q = clip(q, 0, None); p = q / q.sum()       float32; q.sum() = sequential float32 sum over
//                                               8192-element buffers of pairwise sums
//   choice: cdf = cumsum(float64(p)); cdf /= cdf[-1]; idx = searchsorted(cdf, u, 'right')
//
// The float64 cumsum is sequential.  It is reproduced exactly in parallel: inside one binade
// [2^E, 2^(E+1)) every partial sum is a multiple of u = 2^(E-52), and rounding s + x depends on s
// only through the parity of s/u (ties to even).  So a chunk of additions that stays inside one
// binade maps s to s + d[parity(s/u)]*u; such maps compose associatively
// (compose(f, g)[p] = f[p] + g[p ^ (f[p] & 1)]).  Chunks that cross a binade edge (a few dozen as
// the sum climbs from 0 to 1) are summed element by element.  Every map is verified against the
// exact incoming value before use, so a misclassified chunk only costs time.
#include <string.h>

#include "common.hpp"

namespace cqs {

constexpr int kChunk = 256;         // elements per chunk
constexpr int kSuper = 1024;        // chunks per superchunk
constexpr int kNpBuf = 8192;        // NumPy's reduction buffer (elements)
constexpr int kChunksPerWG = 32;    // chunks staged per workgroup in the chunk kernels

struct ChunkFn {
  long long d0, d1;  // increments in ulps for incoming parity 0 / 1
  int E;             // binade exponent
  int ok;            // 1 = usable map
  long long pad;
};

struct SampleHdr {
  float T32;        // NumPy's float32 normalisation sum
  float pad0;
  double total;     // cdf[-1] before normalisation (exact NumPy value)
  int n_seq_chunks; // chunks summed element by element (diagnostic)
  int pad1[59];
};

struct Layout {
  uint64_t N, NC, Cl, NSC, SCl, NB;
  size_t o_hdr, o_csum32, o_capx, o_cin, o_cfn, o_scsum, o_scpre, o_scfn, o_scstart, o_scmode,
      o_cstart, o_t32sum, o_t32map, bytes;
};

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

static Layout layout_for(int n) {
  Layout l;
  l.N = 1ull << n;
  l.NC = l.N >= (uint64_t)kChunk ? l.N / kChunk : 1;
  l.Cl = l.N / l.NC;
  l.NSC = l.NC >= (uint64_t)kSuper ? l.NC / kSuper : 1;
  l.SCl = l.NC / l.NSC;
  l.NB = l.N >= (uint64_t)kNpBuf ? l.N / kNpBuf : 1;
  size_t o = 0;
  l.o_hdr = o; o += al(sizeof(SampleHdr));
  l.o_csum32 = o; o += al(l.NB * sizeof(float));
  l.o_capx = o; o += al(l.NC * sizeof(double));
  l.o_cin = o; o += al(l.NC * sizeof(double));
  l.o_cfn = o; o += al(l.NC * sizeof(ChunkFn));
  l.o_scsum = o; o += al(l.NSC * sizeof(double));
  l.o_scpre = o; o += al(l.NSC * sizeof(double));
  l.o_scfn = o; o += al(l.NSC * sizeof(ChunkFn));
  l.o_scstart = o; o += al(l.NSC * sizeof(double));
  l.o_scmode = o; o += al(l.NSC * sizeof(int));
  l.o_cstart = o; o += al(l.NC * sizeof(double));
  const uint64_t nch32 = (l.NB + 63) / 64;  // chunks of the float32 total's exact walk
  l.o_t32sum = o; o += al(nch32 * sizeof(double));
  l.o_t32map = o; o += al(nch32 * 16);
  l.bytes = o;
  return l;
}

// Probability source: |psi_i|^2, Re rho_ii, or a caller buffer; clipped at 0 like np.clip.
// With permute != 0, sample index i is first mapped to the state index whose bit perm[b] is
// bit b of i (all qubits measured in a non-natural order).
struct Loader {
  const float2* amps;
  const float* probs;
  uint64_t N;
  int mode;  // 0: state vector, 1: density-matrix diagonal, 2: float buffer
  int nbits;
  int permute;
  int8_t perm[CQS_MAX_BITS];
  __device__ __forceinline__ uint64_t map(uint64_t i) const {
    if (!permute) return i;
    uint64_t r = 0;
    for (int b = 0; b < nbits; ++b) r |= ((i >> b) & 1ull) << perm[b];
    return r;
  }
  __device__ __forceinline__ float q(uint64_t i) const {
    float v;
    if (mode == 0) {
      const float2 a = amps[map(i)];
      v = fmaf(a.x, a.x, a.y * a.y);
    } else if (mode == 1) {
      const uint64_t k = map(i);
      v = amps[k * (N + 1)].x;
    } else {
      v = probs[i];
    }
    return v < 0.f ? 0.f : v;
  }
  // q(i), ..., q(i + 3) for i a multiple of 4: one 16-byte load of probabilities or two of
  // amplitudes when the layout allows it, else four single reads.
  __device__ __forceinline__ float4 q4(uint64_t i) const {
    if (!permute && mode == 2 && (reinterpret_cast<uintptr_t>(probs + i) & 15) == 0) {
      const float4 v = *reinterpret_cast<const float4*>(probs + i);
      return make_float4(v.x < 0.f ? 0.f : v.x, v.y < 0.f ? 0.f : v.y, v.z < 0.f ? 0.f : v.z, v.w < 0.f ? 0.f : v.w);
    }
    if (!permute && mode == 0 && (reinterpret_cast<uintptr_t>(amps + i) & 15) == 0) {
      const float4 a = *reinterpret_cast<const float4*>(amps + i);
      const float4 b = *reinterpret_cast<const float4*>(amps + i + 2);
      const float p0 = fmaf(a.x, a.x, a.y * a.y), p1 = fmaf(a.z, a.z, a.w * a.w);
      const float p2 = fmaf(b.x, b.x, b.y * b.y), p3 = fmaf(b.z, b.z, b.w * b.w);
      return make_float4(p0 < 0.f ? 0.f : p0, p1 < 0.f ? 0.f : p1, p2 < 0.f ? 0.f : p2, p3 < 0.f ? 0.f : p3);
    }
    return make_float4(q(i), q(i + 1), q(i + 2), q(i + 3));
  }
  // Uniform: q4 at every multiple of 4 is plain 16-byte loads (no permutation, aligned base).
  __device__ __forceinline__ bool vec4() const {
    if (permute) return false;
    if (mode == 2) return (reinterpret_cast<uintptr_t>(probs) & 15) == 0;
    if (mode == 0) return (reinterpret_cast<uintptr_t>(amps) & 15) == 0;
    return false;
  }
  // R consecutive groups of 4 at i0, i0 + stride, ... (vec4() holds): every load issued before
  // any is used, so their latencies overlap.
  template <int R>
  __device__ __forceinline__ void q4_batch(uint64_t i0, uint64_t stride, float4 (&out)[R]) const {
    if (mode == 2) {
#pragma unroll
      for (int r = 0; r < R; ++r) out[r] = *reinterpret_cast<const float4*>(probs + i0 + r * stride);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float4 v = out[r];
        out[r] = make_float4(v.x < 0.f ? 0.f : v.x, v.y < 0.f ? 0.f : v.y, v.z < 0.f ? 0.f : v.z,
                             v.w < 0.f ? 0.f : v.w);
      }
    } else {
      float4 a[R], b[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r] = *reinterpret_cast<const float4*>(amps + i0 + r * stride);
        b[r] = *reinterpret_cast<const float4*>(amps + i0 + r * stride + 2);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float p0 = fmaf(a[r].x, a[r].x, a[r].y * a[r].y), p1 = fmaf(a[r].z, a[r].z, a[r].w * a[r].w);
        const float p2 = fmaf(b[r].x, b[r].x, b[r].y * b[r].y), p3 = fmaf(b[r].z, b[r].z, b[r].w * b[r].w);
        out[r] = make_float4(p0 < 0.f ? 0.f : p0, p1 < 0.f ? 0.f : p1, p2 < 0.f ? 0.f : p2, p3 < 0.f ? 0.f : p3);
      }
    }
  }
};

__device__ __forceinline__ double ldexp_d(double x, int e) { return ldexp(x, e); }

// ---------------------------------------------------------------- NumPy float32 sum
// Pairwise sum of one buffer of len (power of two <= 8192) exactly as NumPy's
// FLOAT_pairwise_sum: blocks of 128 with 8 strided accumulators, halves combined recursively.
// craw (optional, chunks of Cl elements): each chunk's sum of q in FP64, before normalisation.
__global__ __launch_bounds__(256) void k_sum_buffers(Loader ld, uint64_t len, float* __restrict__ out,
                                                     uint64_t Cl = 0, double* __restrict__ craw = nullptr) {
  __shared__ float sq[kNpBuf + kNpBuf / 128];
  const uint64_t off = (uint64_t)blockIdx.x * len;
  if (len == (uint64_t)kNpBuf && ld.vec4()) {  // full buffers: 8 groups of 4 per thread, loads batched
    constexpr int R = kNpBuf / 1024;
    float4 v[R];
    ld.q4_batch<R>(off + 4 * threadIdx.x, 1024, v);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = 4 * threadIdx.x + (uint64_t)r * 1024;
      float* d = sq + i + (i >> 7);
      d[0] = v[r].x;
      d[1] = v[r].y;
      d[2] = v[r].z;
      d[3] = v[r].w;
    }
  } else if (len % 4 == 0) {
    for (uint64_t i = 4 * threadIdx.x; i < len; i += 1024) {
      const float4 v = ld.q4(off + i);  // i .. i + 3 lie in one 128-element block
      float* d = sq + i + (i >> 7);
      d[0] = v.x;
      d[1] = v.y;
      d[2] = v.z;
      d[3] = v.w;
    }
  } else {
    for (uint64_t i = threadIdx.x; i < len; i += 256) sq[i + (i >> 7)] = ld.q(off + i);
  }
  __syncthreads();
  if (craw != nullptr && len >= Cl) {
    // the buffer's chunks (len / Cl of them, at most 32): 8 threads per chunk, strided, then a
    // shuffle tree; the order does not matter here (the chunk maps verify every use)
    const int nch = (int)(len / Cl);
    for (int c = threadIdx.x >> 3; c < nch; c += 32) {
      double a = 0.0;
      for (uint64_t i = threadIdx.x & 7; i < Cl; i += 8) {
        const uint64_t k = (uint64_t)c * Cl + i;
        a += (double)sq[k + (k >> 7)];
      }
      a += __shfl_xor(a, 1, 8);
      a += __shfl_xor(a, 2, 8);
      a += __shfl_xor(a, 4, 8);
      if ((threadIdx.x & 7) == 0) craw[off / Cl + c] = a;
    }
  }
  if (threadIdx.x >= 64) return;
  const int t = threadIdx.x;
  if (len < 128) {
    if (t == 0) {
      float res;
      if (len < 8) {
        res = 0.f;
        for (uint64_t i = 0; i < len; ++i) res = res + sq[i];
      } else {
        float r[8];
        for (int j = 0; j < 8; ++j) r[j] = sq[j];
        for (uint64_t i = 8; i < len; i += 8)
          for (int j = 0; j < 8; ++j) r[j] = r[j] + sq[i + j];
        res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
      }
      out[blockIdx.x] = res;
    }
    return;
  }
  const int nblk = (int)(len / 128);
  float s = 0.f;
  if (t < nblk) {
    const float* b = sq + t * 129;
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = b[j];
    for (int i = 8; i < 128; i += 8)
#pragma unroll
      for (int j = 0; j < 8; ++j) r[j] = r[j] + b[i + j];
    s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  }
  // adjacent-pair tree == NumPy's recursive halving for power-of-two lengths
  for (int o = 1; o < nblk; o <<= 1) s = s + __shfl_xor(s, o, 64);
  if (t == 0) out[blockIdx.x] = s;
}

// Sequential float32 accumulation of the buffer sums (NumPy's reduce loop).
// The buffer sums are added one after the other in float32 (NumPy's order); the block stages
// 8192 of them at a time in LDS so that the one adding thread never waits on global memory.
__global__ __launch_bounds__(256) void k_total32(const float* __restrict__ bsum, uint64_t nb, SampleHdr* hdr) {
  constexpr int kStage = 8192;
  __shared__ __align__(16) float st[kStage];
  float S = 0.f;
  for (uint64_t c0 = 0; c0 < nb; c0 += kStage) {
    const int cnt = (int)((nb - c0) < (uint64_t)kStage ? (nb - c0) : (uint64_t)kStage);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += 256) st[i] = bsum[c0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      // the same sequential adds; 16-byte LDS reads issued ahead of the dependent chain
      int i = 0;
      const float4* st4 = reinterpret_cast<const float4*>(st);
#pragma unroll 8
      for (; i + 4 <= cnt; i += 4) {
        const float4 v = st4[i >> 2];
        S = S + v.x;
        S = S + v.y;
        S = S + v.z;
        S = S + v.w;
      }
      for (; i < cnt; ++i) S = S + st[i];
    }
  }
  if (threadIdx.x == 0) hdr->T32 = S;
}

__device__ __forceinline__ double px(const Loader& ld, float T32, uint64_t i) {
  return (double)__fdiv_rn(ld.q(i), T32);
}

// Approximate chunk sums: the FP64 chunk sums of q from k_sum_buffers divided by NumPy's float32
// total (normalised here, in place).  Against the sum of the float32 quotients p_i the chunks use,
// the relative error is at most 2^-24 (each quotient is one rounding); the chunk maps take a binade
// from these only when the chunk lies 2^-22 (relative) inside it, and every map is verified against
// the exact running sum before use.  Then the exclusive scan inside each superchunk; superchunk
// totals out.
__global__ __launch_bounds__(256) void k_scan_local(double* __restrict__ capx, const SampleHdr* hdr, uint64_t SCl,
                                                    double* __restrict__ cin, double* __restrict__ scsum) {
  __shared__ double sh[256];
  const uint64_t base = (uint64_t)blockIdx.x * SCl;
  const int per = (int)((SCl + 255) / 256);
  const int t = threadIdx.x;
  const double T32 = (double)hdr->T32;
  double loc = 0.0;
  for (int k = 0; k < per; ++k) {
    const uint64_t i = (uint64_t)t * per + k;
    if (i < SCl) {
      const double v = capx[base + i] / T32;
      capx[base + i] = v;
      loc += v;
    }
  }
  sh[t] = loc;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (int k = 0; k < 256; ++k) {
      const double v = sh[k];
      sh[k] = run;
      run += v;
    }
    scsum[blockIdx.x] = run;
  }
  __syncthreads();
  double run = sh[t];
  for (int k = 0; k < per; ++k) {
    const uint64_t i = (uint64_t)t * per + k;
    if (i < SCl) {
      cin[base + i] = run;
      run += capx[base + i];
    }
  }
}

// Exclusive prefix of the superchunk sums (approximate values: any summation order will do), one
// workgroup, each thread over a contiguous run.
__global__ __launch_bounds__(1024) void k_scan_top(const double* __restrict__ scsum, uint64_t NSC,
                                                   double* __restrict__ scpre) {
  __shared__ double tot[1024];
  const int t = threadIdx.x;
  const uint64_t per = (NSC + 1023) / 1024;
  const uint64_t b = (uint64_t)t * per, e = b + per < NSC ? b + per : NSC;
  double a = 0.0;
  for (uint64_t i = b; i < e; ++i) a += scsum[i];
  tot[t] = a;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (int k = 0; k < 1024; ++k) {
      const double v = tot[k];
      tot[k] = run;
      run += v;
    }
  }
  __syncthreads();
  double run = tot[t];
  for (uint64_t i = b; i < e; ++i) {
    scpre[i] = run;
    run += scsum[i];
  }
}

// ---------------------------------------------------------------- chunk maps
__device__ __forceinline__ int binade_of(double x) {
  int e;
  frexp(x, &e);
  return e - 1;  // x in [2^E, 2^(E+1))
}

__global__ __launch_bounds__(256) void k_chunk_fn(Loader ld, const SampleHdr* hdr, uint64_t NC,
                                                  uint64_t Cl, uint64_t SCl, const double* __restrict__ capx,
                                                  const double* __restrict__ cin, const double* __restrict__ scpre,
                                                  ChunkFn* __restrict__ cfn) {
  __shared__ float sp[kChunksPerWG * (kChunk + 1)];
  __shared__ int ovf[2 * kChunksPerWG];
  const float T32 = hdr->T32;
  const uint64_t c0 = (uint64_t)blockIdx.x * kChunksPerWG;
  const uint64_t nloc = (NC - c0) < (uint64_t)kChunksPerWG ? (NC - c0) : (uint64_t)kChunksPerWG;
  if (Cl == (uint64_t)kChunk && nloc == (uint64_t)kChunksPerWG && ld.vec4()) {  // loads batched
    constexpr int R = kChunksPerWG * kChunk / 1024;
    float4 v[R];
    ld.q4_batch<R>(c0 * Cl + 4 * threadIdx.x, 1024, v);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint64_t i = 4 * threadIdx.x + (uint64_t)r * 1024;
      float* d = sp + (i / kChunk) * (kChunk + 1) + (i % kChunk);
      d[0] = __fdiv_rn(v[r].x, T32);
      d[1] = __fdiv_rn(v[r].y, T32);
      d[2] = __fdiv_rn(v[r].z, T32);
      d[3] = __fdiv_rn(v[r].w, T32);
    }
  } else if (Cl % 4 == 0) {
    for (uint64_t i = 4 * threadIdx.x; i < nloc * Cl; i += 1024) {
      const uint64_t lc = i / Cl, k = i % Cl;  // k .. k + 3 lie in one chunk
      const float4 v = ld.q4(c0 * Cl + i);
      float* d = sp + lc * (kChunk + 1) + k;
      d[0] = __fdiv_rn(v.x, T32);
      d[1] = __fdiv_rn(v.y, T32);
      d[2] = __fdiv_rn(v.z, T32);
      d[3] = __fdiv_rn(v.w, T32);
    }
  } else {
    for (uint64_t i = threadIdx.x; i < nloc * Cl; i += 256) {
      const uint64_t lc = i / Cl, k = i % Cl;
      sp[lc * (kChunk + 1) + k] = __fdiv_rn(ld.q((c0 + lc) * Cl + k), T32);
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  const int lc = t & (kChunksPerWG - 1);
  const int par = t / kChunksPerWG;  // 0 or 1 for t < 64
  const uint64_t c = c0 + lc;
  int E = 0, clean = 0;
  long long d = 0;
  int over = 1;
  if (t < 2 * kChunksPerWG && c < NC) {
    const double a = cin[c] + scpre[c / SCl];
    const double b = a + capx[c];
    const double eps = 0x1p-22;  // 4x the approximate sums' relative error
    if (a > 0.0) {
      E = binade_of(a);
      const double lo = ldexp_d(1.0, E), hi = ldexp_d(1.0, E + 1);
      clean = (a * (1.0 - eps) >= lo) && (b * (1.0 + eps) < hi);
    }
    if (clean) {
      const double u = ldexp_d(1.0, E - 52);
      const double s0 = ldexp_d(1.0, E) + (par ? u : 0.0);
      double s = s0;
      const float* row = sp + lc * (kChunk + 1);
      for (uint64_t k = 0; k < Cl; ++k) s = s + (double)row[k];
      over = !(s < ldexp_d(1.0, E + 1));
      d = (long long)ldexp_d(s - s0, 52 - E);
    }
    ovf[t] = over;
  }
  __syncthreads();
  if (t < 2 * kChunksPerWG && c < NC) {
    if (par == 0) {
      cfn[c].d0 = d;
      cfn[c].E = E;
      cfn[c].ok = clean && !ovf[lc] && !ovf[lc + kChunksPerWG];
    } else {
      cfn[c].d1 = d;
    }
  }
}

__device__ __forceinline__ void compose(long long& a0, long long& a1, long long b0, long long b1) {
  // (apply a, then b)
  const long long r0 = a0 + ((a0 & 1) ? b1 : b0);
  const long long r1 = a1 + ((a1 & 1) ? b0 : b1);
  a0 = r0;
  a1 = r1;
}

// Compose the chunk maps of each superchunk when all share one binade.
__global__ __launch_bounds__(256) void k_super_fn(const ChunkFn* __restrict__ cfn, uint64_t SCl,
                                                  ChunkFn* __restrict__ scfn) {
  __shared__ long long s0[256], s1[256];
  __shared__ int okall, E0;
  const uint64_t base = (uint64_t)blockIdx.x * SCl;
  const int t = threadIdx.x;
  if (t == 0) {
    okall = 1;
    E0 = cfn[base].E;
  }
  __syncthreads();
  const int per = (int)((SCl + 255) / 256);
  long long a0 = 0, a1 = 0;
  int ok = 1;
  for (int k = 0; k < per; ++k) {
    const uint64_t i = (uint64_t)t * per + k;
    if (i < SCl) {
      const ChunkFn f = cfn[base + i];
      if (!f.ok || f.E != E0) ok = 0;
      compose(a0, a1, f.d0, f.d1);
    }
  }
  if (!ok) okall = 0;  // benign race: every writer stores 0
  s0[t] = a0;
  s1[t] = a1;
  __syncthreads();
  if (t == 0) {
    long long r0 = 0, r1 = 0;
    for (int k = 0; k < 256; ++k) compose(r0, r1, s0[k], s1[k]);
    ChunkFn out;
    out.d0 = r0;
    out.d1 = r1;
    out.E = E0;
    out.ok = okall;
    out.pad = 0;
    scfn[blockIdx.x] = out;
  }
}

// Inside one binade a double is its bit pattern: S in [2^E, 2^(E+1)) exactly when its exponent
// field is E + 1023 (S is a positive normal double here), the ulp parity is the pattern's lowest
// bit, and adding d >= 0 ulps (the maps only add probabilities) is adding d to the pattern, exact
// while the exponent field stays the same.
//
// One step of the walk over maps[0 .. avail) (avail <= 64) with one wave.  Lane j composes maps
// 0 .. j (an inclusive scan of compose across the wave, earlier maps first), so a run of maps that
// are usable in S's binade and keep the running value inside it advances S in one step; applying
// the composed map is the same integer arithmetic as applying the maps one after another (the
// value stays in the binade after every map because the sums only grow).  Returns the number of
// maps applied (0 .. avail, uniform); `before` of lane j is the value before map j, for every
// j <= the returned count.
__device__ __forceinline__ int walk_maps(double& S, const ChunkFn* maps, int avail, double& before) {
  constexpr long long kExp = 0x7ff0000000000000ll;
  const int lane = threadIdx.x & 63;
  long long d0 = 0, d1 = 0;
  int E = 0, ok = 0;
  if (lane < avail) {
    const ChunkFn f = maps[lane];
    d0 = f.d0;
    d1 = f.d1;
    E = f.E;
    ok = f.ok;
  }
  const long long b = __double_as_longlong(S);
  const long long ebits = b & kExp;
  const bool usable = S > 0.0 && lane < avail && ok && ((long long)(E + 1023) << 52) == ebits;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    long long p0 = __shfl_up(d0, o, 64), p1 = __shfl_up(d1, o, 64);
    if (lane >= o) {
      compose(p0, p1, d0, d1);  // maps before this lane's window first
      d0 = p0;
      d1 = p1;
    }
  }
  const long long r = b + ((b & 1) ? d1 : d0);
  const unsigned long long stays = __ballot(usable && (r & kExp) == ebits);
  const int n = ~stays ? __builtin_ctzll(~stays) : 64;  // first lane whose map does not apply
  const long long rp = __shfl_up(r, 1, 64);
  before = lane == 0 ? S : __longlong_as_double(rp);
  if (n > 0) S = __longlong_as_double(__shfl(r, n - 1, 64));
  return n;
}

// One workgroup of 64 lanes walks the superchunks in order, up to 64 maps per step (walk_maps).
// All lanes keep identical copies of S.  A superchunk whose composed map does not apply is walked
// by its chunk maps; a chunk whose map does not apply (it crosses a binade edge, or S is still 0)
// is summed element by element from staged values.
__global__ __launch_bounds__(64) void k_scan_serial(Loader ld, SampleHdr* hdr, uint64_t NSC, uint64_t SCl,
                                                    uint64_t Cl, const ChunkFn* __restrict__ cfn,
                                                    const ChunkFn* __restrict__ scfn,
                                                    double* __restrict__ scstart, int* __restrict__ scmode,
                                                    double* __restrict__ cstart) {
  constexpr int kStage = 512;  // superchunk maps staged in LDS, so the walk reads LDS only
  __shared__ double xs[kChunk];
  __shared__ ChunkFn sf[kStage];
  __shared__ ChunkFn cf[kSuper];  // the chunk maps of a superchunk walked chunk by chunk
  const float T32 = hdr->T32;
  const int t = threadIdx.x;
  const int ncs = (int)SCl;  // chunks per superchunk (<= kSuper)
  double S = 0.0;
  int nseq = 0;
  for (uint64_t s0 = 0; s0 < NSC; s0 += kStage) {
    const int ns = (int)(NSC - s0 < (uint64_t)kStage ? NSC - s0 : (uint64_t)kStage);
    __syncthreads();
#pragma unroll  // every lane's loads issued before any LDS store (one latency, not eight)
    for (int r = 0; r < kStage / 64; ++r) {
      const int i = t + 64 * r;
      if (i < ns) sf[i] = scfn[s0 + i];
    }
    __syncthreads();
    for (int k = 0; k < ns;) {
      double before;
      const int adv = walk_maps(S, sf + k, ns - k < 64 ? ns - k : 64, before);
      if (t <= adv && k + t < ns) scstart[s0 + k + t] = before;  // lane adv: the next superchunk's start
      if (t < adv) scmode[s0 + k + t] = 1;
      k += adv;
      if (adv == 64 || k >= ns) continue;
      const uint64_t sc = s0 + k;  // walked by its chunk maps
      if (t == 0) scmode[sc] = 0;
      __syncthreads();
#pragma unroll
      for (int r = 0; r < kSuper / 64; ++r) {
        const int i = t + 64 * r;
        if (i < ncs) cf[i] = cfn[sc * SCl + i];
      }
      __syncthreads();
      for (int c = 0; c < ncs;) {
        double cb;
        const int a = walk_maps(S, cf + c, ncs - c < 64 ? ncs - c : 64, cb);
        if (t <= a && c + t < ncs) cstart[sc * SCl + c + t] = cb;
        c += a;
        if (a == 64 || c >= ncs) continue;
        const uint64_t ch = sc * SCl + c;  // summed element by element
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kChunk / 64; ++r) {
          const uint64_t i = t + 64 * r;
          if (i < Cl) xs[i] = px(ld, T32, ch * Cl + i);
        }
        __syncthreads();
        for (uint64_t i = 0; i < Cl; ++i) S = S + xs[i];
        ++nseq;
        ++c;
      }
      ++k;
    }
  }
  if (t == 0) {
    hdr->total = S;
    hdr->n_seq_chunks = nseq;
  }
}

// ---------------------------------------------------------------- NumPy's float32 total, many buffers
// The same exact-walk idea in float32 for the sequential float32 additions of the buffer sums
// (k_total32 adds them one at a time in one thread): a float32 in [2^E, 2^(E+1)) (E >= -126) is a
// multiple of u = 2^(E-23), and rounding s + x depends on s only through the parity of s/u.  The
// buffer sums are cut into chunks of kT32Chunk; a chunk whose approximate start and end (FP64 prefix
// of FP64 chunk sums) lie well inside one binade gets a map (ulps added for each incoming parity),
// computed by adding its elements to 2^E + parity*u in float32; one wave walks the maps up to 64 at
// a time (the float32 twin of walk_maps) and adds the other chunks element by element.
constexpr int kT32Chunk = 64;
constexpr int kT32MinBuffers = 1024;  // below this the one-thread loop is as fast
constexpr int kT32Stage = 1024;       // maps staged in LDS per round of the walk
struct Map32 {
  int d0, d1;  // ulps for incoming parity 0 / 1
  int E;       // binade exponent
  int ok;      // 1 = usable map
};
static_assert(sizeof(Map32) == 16, "the workspace layout reserves 16 bytes per map");

__global__ __launch_bounds__(256) void k_t32_chunk_sums(const float* __restrict__ bsum, uint64_t nb,
                                                        double* __restrict__ csum) {
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t nch = (nb + kT32Chunk - 1) / kT32Chunk;
  if (c >= nch) return;
  const uint64_t e = (c + 1) * kT32Chunk < nb ? (c + 1) * kT32Chunk : nb;
  double a = 0.0;
  for (uint64_t i = c * kT32Chunk; i < e; ++i) a += (double)bsum[i];
  csum[c] = a;
}

// Exclusive prefix of the chunk sums in place (one workgroup; each thread scans a contiguous run).
__global__ __launch_bounds__(1024) void k_t32_prefix(double* __restrict__ csum, uint64_t nch) {
  __shared__ double tot[1024];
  const int t = threadIdx.x;
  const uint64_t per = (nch + 1023) / 1024;
  const uint64_t b = (uint64_t)t * per, e = b + per < nch ? b + per : nch;
  double a = 0.0;
  for (uint64_t i = b; i < e; ++i) a += csum[i];
  tot[t] = a;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (int k = 0; k < 1024; ++k) {
      const double v = tot[k];
      tot[k] = run;
      run += v;
    }
  }
  __syncthreads();
  double run = tot[t];
  for (uint64_t i = b; i < e; ++i) {
    const double v = csum[i];
    csum[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(256) void k_t32_maps(const float* __restrict__ bsum, uint64_t nb,
                                                  const double* __restrict__ cpre, Map32* __restrict__ maps) {
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t nch = (nb + kT32Chunk - 1) / kT32Chunk;
  if (c >= nch) return;
  const uint64_t b0 = c * kT32Chunk, e = b0 + kT32Chunk < nb ? b0 + kT32Chunk : nb;
  const double a = cpre[c];
  const double b = c + 1 < nch ? cpre[c + 1] : a;  // the next chunk's start (the last chunk: unused)
  Map32 m = {0, 0, 0, 0};
  if (a > 0.0 && c + 1 < nch) {
    int ex;
    frexp(a, &ex);
    const int E = ex - 1;
    // The FP64 prefix tracks the exact sum, not the float32 running sum (which drifts from it by
    // float32 roundings): a wrong binade only makes the map unusable at its verification.
    const double eps = 0x1p-30;
    if (E >= -126 && E <= 126 && a * (1.0 - eps) >= ldexp(1.0, E) && b * (1.0 + eps) < ldexp(1.0, E + 1)) {
      const float lo = ldexpf(1.f, E), hi = ldexpf(1.f, E + 1), u = ldexpf(1.f, E - 23);
      int d[2], over = 0;
      for (int p = 0; p < 2; ++p) {
        const float s0 = lo + (p ? u : 0.f);
        float s = s0;
        for (uint64_t i = b0; i < e; ++i) s = s + bsum[i];
        over |= !(s < hi);
        d[p] = __float_as_int(s) - __float_as_int(s0);
      }
      m.d0 = d[0];
      m.d1 = d[1];
      m.E = E;
      m.ok = !over;
    }
  }
  maps[c] = m;
}

// walk_maps in float32: applies up to `avail` (<= 64) maps to S with one wave; returns how many.
__device__ __forceinline__ int walk_maps32(float& S, const Map32* maps, int avail) {
  constexpr int kExp = 0x7f800000;
  const int lane = threadIdx.x & 63;
  long long d0 = 0, d1 = 0;
  int E = 0, ok = 0;
  if (lane < avail) {
    const Map32 f = maps[lane];
    d0 = f.d0;
    d1 = f.d1;
    E = f.E;
    ok = f.ok;
  }
  const int bits = __float_as_int(S);
  const int ebits = bits & kExp;
  const bool usable = S > 0.f && lane < avail && ok && ((E + 127) << 23) == ebits;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    long long p0 = __shfl_up(d0, o, 64), p1 = __shfl_up(d1, o, 64);
    if (lane >= o) {
      compose(p0, p1, d0, d1);
      d0 = p0;
      d1 = p1;
    }
  }
  const long long r = (long long)bits + ((bits & 1) ? d1 : d0);
  const unsigned long long stays = __ballot(usable && r < 0x7f800000ll && ((int)r & kExp) == ebits);
  const int n = ~stays ? __builtin_ctzll(~stays) : 64;
  if (n > 0) S = __int_as_float((int)__shfl(r, n - 1, 64));
  return n;
}

__global__ __launch_bounds__(64) void k_t32_walk(const float* __restrict__ bsum, uint64_t nb,
                                                 const Map32* __restrict__ maps, SampleHdr* hdr) {
  __shared__ Map32 sm[kT32Stage];
  const int t = threadIdx.x;
  const uint64_t nch = (nb + kT32Chunk - 1) / kT32Chunk;
  float S = 0.f;
  for (uint64_t s0 = 0; s0 < nch; s0 += kT32Stage) {
    const int ns = (int)(nch - s0 < (uint64_t)kT32Stage ? nch - s0 : (uint64_t)kT32Stage);
    __syncthreads();
    for (int i = t; i < ns; i += 64) sm[i] = maps[s0 + i];
    __syncthreads();
    for (int k = 0; k < ns;) {
      const int adv = walk_maps32(S, sm + k, ns - k < 64 ? ns - k : 64);
      k += adv;
      if (adv == 64 || k >= ns) continue;
      // chunk s0 + k added element by element: one load per lane, then the lanes' values in order
      // (every lane adds the same values, so S stays uniform)
      const uint64_t i0 = (s0 + k) * kT32Chunk;
      const int cnt = (int)(nb - i0 < (uint64_t)kT32Chunk ? nb - i0 : (uint64_t)kT32Chunk);
      const int v = t < cnt ? __float_as_int(bsum[i0 + t]) : 0;
      for (int i = 0; i < cnt; ++i) S = S + __int_as_float(__builtin_amdgcn_readlane(v, i));
      ++k;
    }
  }
  if (t == 0) hdr->T32 = S;
}

// Chunk start values inside superchunks that were advanced by one composed map.
__global__ __launch_bounds__(256) void k_chunk_starts(const ChunkFn* __restrict__ cfn, uint64_t SCl,
                                                      const double* __restrict__ scstart,
                                                      const int* __restrict__ scmode,
                                                      const ChunkFn* __restrict__ scfn,
                                                      double* __restrict__ cstart) {
  if (!scmode[blockIdx.x]) return;
  __shared__ long long g0[256], g1[256];
  const uint64_t base = (uint64_t)blockIdx.x * SCl;
  const int t = threadIdx.x;
  const int per = (int)((SCl + 255) / 256);
  long long a0 = 0, a1 = 0;
  for (int k = 0; k < per; ++k) {
    const uint64_t i = (uint64_t)t * per + k;
    if (i < SCl) compose(a0, a1, cfn[base + i].d0, cfn[base + i].d1);
  }
  g0[t] = a0;
  g1[t] = a1;
  __syncthreads();
  // inclusive Hillis-Steele scan; element t becomes (x_0 ... x_t) applied in order
  for (int o = 1; o < 256; o <<= 1) {
    long long p0 = 0, p1 = 0, c0 = g0[t], c1 = g1[t];
    const bool has = t >= o;
    if (has) {
      p0 = g0[t - o];
      p1 = g1[t - o];
    }
    __syncthreads();
    if (has) {
      compose(p0, p1, c0, c1);  // earlier prefix first, then this element
      g0[t] = p0;
      g1[t] = p1;
    }
    __syncthreads();
  }
  const int E = scfn[blockIdx.x].E;
  const double S0 = scstart[blockIdx.x];
  const long long k = (long long)ldexp_d(S0, 52 - E);
  const int p = (int)(k & 1);
  long long acc = 0;
  if (t > 0) acc = p ? g1[t - 1] : g0[t - 1];
  for (int kk = 0; kk < per; ++kk) {
    const uint64_t i = (uint64_t)t * per + kk;
    if (i < SCl) {
      cstart[base + i] = S0 + ldexp_d((double)acc, E - 52);
      const ChunkFn f = cfn[base + i];
      const int pp = p ^ (int)(acc & 1);
      acc += pp ? f.d1 : f.d0;
    }
  }
}

// ---------------------------------------------------------------- draws
__global__ __launch_bounds__(256) void k_draw(Loader ld, const SampleHdr* hdr, uint64_t NC, uint64_t Cl,
                                              const double* __restrict__ cstart,
                                              const double* __restrict__ uni, int64_t reps,
                                              int64_t* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= reps) return;
  const double T = hdr->total;
  const float T32 = hdr->T32;
  const double u = uni[r];
  // first chunk whose end value, normalised, exceeds u
  uint64_t lo = 0, hi = NC - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const double end = (mid + 1 < NC) ? cstart[mid + 1] : T;
    if (end / T > u)
      hi = mid;
    else
      lo = mid + 1;
  }
  const uint64_t c = lo;
  double s = cstart[c];
  int64_t idx = (int64_t)(c * Cl + Cl - 1);
  if (Cl % 4 == 0) {  // four probabilities per load, added one at a time as NumPy's cumsum does
    for (uint64_t i = 0; i < Cl; i += 4) {
      const float4 q = ld.q4(c * Cl + i);
      const float pq[4] = {q.x, q.y, q.z, q.w};
      int hit = -1;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s = s + (double)__fdiv_rn(pq[k], T32);
        if (hit < 0 && s / T > u) hit = k;
      }
      if (hit >= 0) {
        idx = (int64_t)(c * Cl + i + hit);
        break;
      }
    }
  } else {
    for (uint64_t i = 0; i < Cl; ++i) {
      s = s + px(ld, T32, c * Cl + i);
      if (s / T > u) {
        idx = (int64_t)(c * Cl + i);
        break;
      }
    }
  }
  out[r] = idx;
}

__global__ void k_bits(const int64_t* __restrict__ idx, int64_t reps, int m, uint8_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t total = reps * m;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t r = e / m;
    const int k = (int)(e % m);
    out[e] = (uint8_t)((idx[r] >> (m - 1 - k)) & 1);
  }
}

// Marginal probabilities: measured axes first (in the given order), then the rest; sum over the
// rest with NumPy's pairwise order per row (rows are contiguous after the transpose).
struct AxisBits {
  int b[64];
};

// Pairwise sum (NumPy order) of a row of length len <= 128 held by one thread.
template <typename F>
__device__ __forceinline__ float pairwise_short(uint64_t len, F at) {
  if (len < 8) {
    float res = 0.f;
    for (uint64_t i = 0; i < len; ++i) res = res + at(i);
    return res;
  }
  float r[8];
  for (int j = 0; j < 8; ++j) r[j] = at(j);
  for (uint64_t i = 8; i < len; i += 8)
    for (int j = 0; j < 8; ++j) r[j] = r[j] + at(i + j);
  return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
}

// Short rows (<= 128 unmeasured combinations): one thread per output row.  blockIdx.y picks one
// of several states stored `stride` entries apart (and its own 2^m outputs).
__global__ __launch_bounds__(256) void k_marginal_short(Loader ld, int n, AxisBits ab, int m, uint64_t rows,
                                                        float* __restrict__ out, uint64_t stride) {
  ld.amps += (uint64_t)blockIdx.y * stride;
  out += (uint64_t)blockIdx.y * rows;
  const uint64_t row = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= rows) return;
  const int rlen_bits = n - m;
  uint64_t rbase = 0;
  for (int b = 0; b < m; ++b) rbase |= ((row >> (m - 1 - b)) & 1ull) << ab.b[b];
  auto at = [&](uint64_t k) {
    uint64_t fi = rbase;
    for (int b = 0; b < rlen_bits; ++b) fi |= ((k >> (rlen_bits - 1 - b)) & 1ull) << ab.b[m + b];
    return ld.q(fi);
  };
  out[row] = pairwise_short(1ull << rlen_bits, at);
}

__global__ __launch_bounds__(256) void k_marginal(Loader ld, int n, AxisBits ab, int m, float* __restrict__ out,
                                                  uint64_t stride) {
  ld.amps += (uint64_t)blockIdx.y * stride;
  out += (uint64_t)blockIdx.y * gridDim.x;
  const int* axes_bits = ab.b;
  // axes_bits[0..m) : flat bit of measured qubit k (k-th most significant output bit)
  // axes_bits[m..n) : flat bits of the unmeasured qubits in ascending axis order
  __shared__ float sq[kNpBuf + kNpBuf / 128];
  const int rlen_bits = n - m;
  const uint64_t rlen = 1ull << rlen_bits;
  const uint64_t row = blockIdx.x;
  // flat index of (row, k): row bits -> measured positions, k bits -> unmeasured positions
  uint64_t rbase = 0;
  for (int b = 0; b < m; ++b) rbase |= ((row >> (m - 1 - b)) & 1ull) << axes_bits[b];
  float S = 0.f;
  for (uint64_t bufo = 0; bufo < rlen; bufo += kNpBuf) {
    const uint64_t len = (rlen - bufo) < (uint64_t)kNpBuf ? (rlen - bufo) : (uint64_t)kNpBuf;
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < len; i += 256) {
      const uint64_t k = bufo + i;
      uint64_t fi = rbase;
      for (int b = 0; b < rlen_bits; ++b) fi |= ((k >> (rlen_bits - 1 - b)) & 1ull) << axes_bits[m + b];
      sq[i + (i >> 7)] = ld.q(fi);
    }
    __syncthreads();
    float res = 0.f;
    if (threadIdx.x < 64) {
      const int t = threadIdx.x;
      if (len < 128) {
        if (len < 8) {
          for (uint64_t i = 0; i < len; ++i) res = res + sq[i];
        } else {
          float r[8];
          for (int j = 0; j < 8; ++j) r[j] = sq[j];
          for (uint64_t i = 8; i < len; i += 8)
            for (int j = 0; j < 8; ++j) r[j] = r[j] + sq[i + j];
          res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        }
      } else {
        const int nblk = (int)(len / 128);
        float s = 0.f;
        if (t < nblk) {
          const float* bb = sq + t * 129;
          float r[8];
          for (int j = 0; j < 8; ++j) r[j] = bb[j];
          for (int i = 8; i < 128; i += 8)
            for (int j = 0; j < 8; ++j) r[j] = r[j] + bb[i + j];
          s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        }
        for (int o = 1; o < nblk; o <<= 1) s = s + __shfl_xor(s, o, 64);
        res = s;
      }
    }
    if (threadIdx.x == 0) S = S + res;
  }
  if (threadIdx.x == 0) out[row] = S;
}


// out[j] = p(map(j)) where output bit b is input flat bit perm[b] and p = |amp|^2 as NumPy rounds
// it (state vector) or Re rho[k, k] (density matrix).  One workgroup per tile of 2^T output
// indices over the output bits S[0..T) (ascending); loads walk the tile with the bits that feed
// input bits 0..5 lowest (contiguous reads), stores with output bits 0..5 lowest (contiguous
// writes), and LDS holds the tile in between.
struct PermTile {
  int n, T, nouter, is_dm;
  int S[16];
  int A[16];   // A[k]: tile bit (index into S) of load-order bit k
  int Ain[16]; // perm[S[A[k]]]: the input bit of load-order bit k (resolved on the host)
  int O[CQS_MAX_BITS];
  int Oin[CQS_MAX_BITS];  // perm[O[b]]: the input bit of outer bit b
};

constexpr int kPermT = 12;
constexpr int kPermRun = 6;  // contiguous run bits on each side (2 x 6 <= kPermT)

__global__ __launch_bounds__(256) void k_perm_probs(const float2* __restrict__ amps, float* __restrict__ out,
                                                    PermTile pt) {
  __shared__ float tile[(1 << kPermT) + ((1 << kPermT) >> 5)];  // one pad slot per 32: fewer bank conflicts
  const uint64_t N = 1ull << pt.n;
  uint64_t base_out = 0, base_in = 0;
  for (int b = 0; b < pt.nouter; ++b)
    if ((blockIdx.x >> b) & 1) {
      base_out |= 1ull << pt.O[b];
      base_in |= 1ull << pt.Oin[b];
    }
  const int T = pt.T, cnt = 1 << T;
  // thread part of the load-order index: bits 0..7 (or fewer when T < 8)
  const int tb = T < 8 ? T : 8;
  int lt = 0;
  uint64_t it = 0;
  for (int k = 0; k < tb; ++k)
    if ((threadIdx.x >> k) & 1) {
      lt |= 1 << pt.A[k];
      it |= 1ull << pt.Ain[k];
    }
  const bool act = (int)threadIdx.x < cnt;
  // element loops unrolled to their compile-time bound (2^(kPermT - 8)): every element's offsets
  // (uniform) are formed and its load issued before any value is used
  constexpr int kE = 1 << (kPermT - 8);
  const int ne = cnt >> tb;
  int le[kE];
  uint64_t ie[kE];
#pragma unroll
  for (int e = 0; e < kE; ++e) {
    le[e] = 0;
    ie[e] = 0;
#pragma unroll
    for (int kk = 0; kk < kPermT - 8; ++kk) {
      const int k = tb + kk;
      if (k < T && ((e >> kk) & 1)) {
        le[e] |= 1 << pt.A[k];
        ie[e] |= 1ull << pt.Ain[k];
      }
    }
  }
  float v[kE];
  if (pt.is_dm) {
#pragma unroll
    for (int e = 0; e < kE; ++e) v[e] = act && e < ne ? amps[(base_in | it | ie[e]) * (N + 1)].x : 0.f;
  } else {
    float2 a[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) a[e] = act && e < ne ? amps[base_in | it | ie[e]] : make_float2(0.f, 0.f);
#pragma unroll
    for (int e = 0; e < kE; ++e) v[e] = fmaf(a[e].x, a[e].x, a[e].y * a[e].y);
  }
#pragma unroll
  for (int e = 0; e < kE; ++e) {
    if (act && e < ne) {
      const int w = lt | le[e];
      tile[w + (w >> 5)] = v[e];
    }
  }
  __syncthreads();
  // store order: tile index j = tid + e*256; the thread part of the output offset once per thread
  uint64_t ot = base_out;
  for (int b = 0; b < tb; ++b)
    if ((threadIdx.x >> b) & 1) ot |= 1ull << pt.S[b];
#pragma unroll
  for (int e = 0; e < kE; ++e) {
    if (e >= ne) break;
    uint64_t oe = 0;
#pragma unroll
    for (int kk = 0; kk < kPermT - 8; ++kk) {
      const int b = tb + kk;
      if (b < T && ((e >> kk) & 1)) oe |= 1ull << pt.S[b];
    }
    if (act) {
      const int j = (int)threadIdx.x + (e << tb);
      out[ot | oe] = tile[j + (j >> 5)];
    }
  }
}

// ---------------------------------------------------------------- one draw per row
// Row r of probs (n_rows rows of 2^m float32 marginals, one per copy of a batched trajectory):
// p = clip(row, 0) / T32_r with T32_r NumPy's float32 sum of the row, then choice(2^m, p) for the
// uniform u[r]: the float64 sequential cumsum, normalised by its last entry, searchsorted right.
// bsum holds the row's NumPy buffer sums (k_sum_buffers over buffers of min(2^m, 8192)); the
// row total adds them one after the other in float32, as NumPy's reduce loop does.
__global__ __launch_bounds__(64) void k_draw_rows(const float* __restrict__ probs, int m, int n_rows,
                                                  const float* __restrict__ bsum, int bufs_per_row,
                                                  const double* __restrict__ uni, int64_t* __restrict__ out_index,
                                                  float* __restrict__ out_p) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= n_rows) return;
  float T32 = 0.f;
  for (int b = 0; b < bufs_per_row; ++b) T32 = T32 + bsum[(size_t)r * bufs_per_row + b];
  const uint64_t len = 1ull << m;
  const float* row = probs + (size_t)r * len;
  auto pr = [&](uint64_t i) {
    const float q = row[i];
    return __fdiv_rn(q < 0.f ? 0.f : q, T32);
  };
  double total = 0.0;
  for (uint64_t i = 0; i < len; ++i) total = total + (double)pr(i);
  const double u = uni[r];
  double S = 0.0;
  uint64_t idx = 0;
  for (; idx < len; ++idx) {  // count of i with cdf_i / total <= u (cdf is non-decreasing)
    S = S + (double)pr(idx);
    if (!(S / total <= u)) break;
  }
  if (idx >= len) idx = len - 1;  // u < 1 and cdf[-1] / total == 1 make this unreachable
  out_index[r] = (int64_t)idx;
  out_p[r] = pr(idx);
}

}  // namespace cqs

using namespace cqs;

extern "C" {

size_t cqs_sample_workspace_bytes(int32_t n_qubits) {
  if (n_qubits < 0 || n_qubits > 40) return 0;
  return layout_for(n_qubits).bytes;
}

static Loader make_loader(const void* amps, const float* probs_in, int n, int is_dm, const int32_t* perm = nullptr) {
  Loader ld;
  ld.amps = (const float2*)amps;
  ld.probs = probs_in;
  ld.N = 1ull << n;
  ld.mode = probs_in ? 2 : (is_dm ? 1 : 0);
  ld.nbits = n;
  ld.permute = 0;
  for (int b = 0; b < CQS_MAX_BITS; ++b) ld.perm[b] = 0;
  if (perm != nullptr && !probs_in) {
    ld.permute = 1;
    for (int b = 0; b < n; ++b) ld.perm[b] = (int8_t)perm[b];
  }
  return ld;
}

// qubit_axes (HOST, n entries or NULL): sampling order of the qubits.  Returns false if it is not
// a permutation; fills perm with the flat bit of each sample-index bit.
static bool axes_to_perm(const int32_t* axes, int n, int32_t* perm) {
  if (axes == nullptr) return true;
  uint64_t seen = 0;
  for (int k = 0; k < n; ++k) {
    const int a = axes[k];
    if (a < 0 || a >= n || ((seen >> a) & 1ull)) return false;
    seen |= 1ull << a;
    perm[n - 1 - k] = n - 1 - a;  // sample bit (n-1-k) <- qubit axes[k] <- flat bit n-1-a
  }
  return true;
}

int cqs_sample_prepare(const void* amps, const float* probs_in, int32_t n_qubits, int32_t is_dm,
                       const int32_t* qubit_axes, void* workspace, void* stream) {
  CQS_REQUIRE(n_qubits >= 0 && n_qubits <= 40, "n_qubits=%d out of range", n_qubits);
  int32_t perm[64];
  CQS_REQUIRE(axes_to_perm(qubit_axes, n_qubits, perm), "qubit_axes is not a permutation of 0..%d", n_qubits - 1);
  CQS_REQUIRE(amps != nullptr || probs_in != nullptr, "no probability source");
  CQS_REQUIRE(workspace != nullptr, "workspace is NULL");
  const Layout l = layout_for(n_qubits);
  char* ws = (char*)workspace;
  SampleHdr* hdr = (SampleHdr*)(ws + l.o_hdr);
  float* csum32 = (float*)(ws + l.o_csum32);
  double* capx = (double*)(ws + l.o_capx);
  double* cin = (double*)(ws + l.o_cin);
  ChunkFn* cfn = (ChunkFn*)(ws + l.o_cfn);
  double* scsum = (double*)(ws + l.o_scsum);
  double* scpre = (double*)(ws + l.o_scpre);
  ChunkFn* scfn = (ChunkFn*)(ws + l.o_scfn);
  double* scstart = (double*)(ws + l.o_scstart);
  int* scmode = (int*)(ws + l.o_scmode);
  double* cstart = (double*)(ws + l.o_cstart);
  const Loader ld = make_loader(amps, probs_in, n_qubits, is_dm, qubit_axes ? perm : nullptr);
  hipStream_t s = as_stream(stream);
  const uint64_t buflen = l.N / l.NB;
  CQS_REQUIRE(buflen >= l.Cl && buflen % l.Cl == 0, "buffer of %llu elements, chunks of %llu",
              (unsigned long long)buflen, (unsigned long long)l.Cl);
  hipLaunchKernelGGL(k_sum_buffers, dim3((unsigned)l.NB), dim3(256), 0, s, ld, buflen, csum32, l.Cl, capx);
  if (l.NB >= (uint64_t)kT32MinBuffers) {
    const uint64_t nch = (l.NB + kT32Chunk - 1) / kT32Chunk;
    double* t32sum = (double*)(ws + l.o_t32sum);
    Map32* t32map = (Map32*)(ws + l.o_t32map);
    const unsigned g = (unsigned)((nch + 255) / 256);
    hipLaunchKernelGGL(k_t32_chunk_sums, dim3(g), dim3(256), 0, s, (const float*)csum32, l.NB, t32sum);
    hipLaunchKernelGGL(k_t32_prefix, dim3(1), dim3(1024), 0, s, t32sum, nch);
    hipLaunchKernelGGL(k_t32_maps, dim3(g), dim3(256), 0, s, (const float*)csum32, l.NB, (const double*)t32sum,
                       t32map);
    hipLaunchKernelGGL(k_t32_walk, dim3(1), dim3(64), 0, s, (const float*)csum32, l.NB, (const Map32*)t32map, hdr);
  } else {
    hipLaunchKernelGGL(k_total32, dim3(1), dim3(256), 0, s, csum32, l.NB, hdr);
  }
  const unsigned gch = (unsigned)((l.NC + kChunksPerWG - 1) / kChunksPerWG);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)l.NSC), dim3(256), 0, s, capx, (const SampleHdr*)hdr, l.SCl, cin,
                     scsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, (const double*)scsum, l.NSC, scpre);
  hipLaunchKernelGGL(k_chunk_fn, dim3(gch), dim3(256), 0, s, ld, (const SampleHdr*)hdr, l.NC, l.Cl, l.SCl,
                     (const double*)capx, (const double*)cin, (const double*)scpre, cfn);
  hipLaunchKernelGGL(k_super_fn, dim3((unsigned)l.NSC), dim3(256), 0, s, (const ChunkFn*)cfn, l.SCl, scfn);
  CQS_REQUIRE(l.SCl <= (uint64_t)kSuper, "superchunk of %llu chunks", (unsigned long long)l.SCl);
  hipLaunchKernelGGL(k_scan_serial, dim3(1), dim3(64), 0, s, ld, hdr, l.NSC, l.SCl, l.Cl,
                     (const ChunkFn*)cfn, (const ChunkFn*)scfn, scstart, scmode, cstart);
  hipLaunchKernelGGL(k_chunk_starts, dim3((unsigned)l.NSC), dim3(256), 0, s, (const ChunkFn*)cfn, l.SCl,
                     (const double*)scstart, (const int*)scmode, (const ChunkFn*)scfn, cstart);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_sample_draw(const void* amps, const float* probs_in, int32_t n_qubits, int32_t is_dm,
                    const int32_t* qubit_axes, const double* uniforms, int64_t reps, int64_t* out_index,
                    void* workspace, void* stream) {
  CQS_REQUIRE(n_qubits >= 0 && n_qubits <= 40, "n_qubits=%d out of range", n_qubits);
  int32_t perm[64];
  CQS_REQUIRE(axes_to_perm(qubit_axes, n_qubits, perm), "qubit_axes is not a permutation of 0..%d", n_qubits - 1);
  CQS_REQUIRE(amps != nullptr || probs_in != nullptr, "no probability source");
  CQS_REQUIRE(workspace != nullptr, "workspace is NULL");
  CQS_REQUIRE(reps >= 0, "reps < 0");
  if (reps == 0) return CQS_OK;
  CQS_REQUIRE(uniforms != nullptr && out_index != nullptr, "uniforms/out are NULL");
  const Layout l = layout_for(n_qubits);
  char* ws = (char*)workspace;
  const Loader ld = make_loader(amps, probs_in, n_qubits, is_dm, qubit_axes ? perm : nullptr);
  const unsigned g = (unsigned)((reps + 255) / 256);
  hipLaunchKernelGGL(k_draw, dim3(g), dim3(256), 0, as_stream(stream), ld, (const SampleHdr*)(ws + l.o_hdr),
                     l.NC, l.Cl, (const double*)(ws + l.o_cstart), uniforms, reps, out_index);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_permuted_probs(const void* amps, int32_t n_qubits, int32_t is_dm, const int32_t* qubit_axes, float* out,
                       void* stream) {
  CQS_REQUIRE(amps != nullptr && out != nullptr && qubit_axes != nullptr, "NULL pointer");
  CQS_REQUIRE(n_qubits >= 1 && n_qubits <= (is_dm ? 20 : 40), "n_qubits=%d out of range", n_qubits);
  int32_t perm[64];
  CQS_REQUIRE(axes_to_perm(qubit_axes, n_qubits, perm), "qubit_axes is not a permutation of 0..%d", n_qubits - 1);
  const int n = n_qubits;
  PermTile pt;
  memset(&pt, 0, sizeof(pt));
  pt.n = n;
  pt.is_dm = is_dm ? 1 : 0;
  const int T = n < kPermT ? n : kPermT;
  // tile bits: output bits 0..5, the output bits feeding input bits 0..5, then the lowest others
  // (loads and stores in runs of at least 64 elements; 12 tile bits hold both sets)
  bool in_s[64] = {false};
  int inv[64];
  for (int b = 0; b < n; ++b) inv[perm[b]] = b;
  int cnt = 0;
  auto take = [&](int b) {
    if (!in_s[b] && cnt < T) {
      in_s[b] = true;
      ++cnt;
    }
  };
  for (int b = 0; b < kPermRun && b < n; ++b) take(b);
  for (int b = 0; b < kPermRun && b < n; ++b) take(inv[b]);
  for (int b = 0; b < n; ++b) take(b);
  int k = 0;
  int tile_of[64];
  for (int b = 0; b < n; ++b)
    if (in_s[b]) {
      tile_of[b] = k;
      pt.S[k++] = b;
    }
  pt.T = T;
  // load order: tile bits feeding input bits 0, 1, 2, ... first
  bool used[16] = {false};
  k = 0;
  for (int q = 0; q < n && k < T; ++q) {
    const int b = inv[q];
    if (in_s[b] && !used[tile_of[b]] && q < kPermRun) {
      used[tile_of[b]] = true;
      pt.A[k++] = tile_of[b];
    }
  }
  for (int t = 0; t < T; ++t)
    if (!used[t]) pt.A[k++] = t;
  for (int t = 0; t < T; ++t) pt.Ain[t] = perm[pt.S[pt.A[t]]];
  pt.nouter = n - T;
  k = 0;
  for (int b = 0; b < n; ++b)
    if (!in_s[b]) {
      pt.Oin[k] = perm[b];
      pt.O[k++] = b;
    }
  const uint64_t grid = 1ull << pt.nouter;
  CQS_REQUIRE(grid < (1ull << 31), "too many tiles");
  hipLaunchKernelGGL(k_perm_probs, dim3((unsigned)grid), dim3(256), 0, as_stream(stream), (const float2*)amps, out,
                     pt);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_indices_to_bits(const int64_t* index, int64_t reps, int32_t m, uint8_t* out, void* stream) {
  CQS_REQUIRE(reps >= 0 && m >= 0 && m <= 63, "bad reps/m");
  if (reps == 0 || m == 0) return CQS_OK;
  CQS_REQUIRE(index != nullptr && out != nullptr, "NULL pointer");
  const int64_t total = reps * m;
  int64_t g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_bits, dim3((unsigned)g), dim3(256), 0, as_stream(stream), index, reps, m, out);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_marginal_probs(const void* amps, int32_t n_qubits, int32_t is_dm, const int32_t* qubit_axes,
                       int32_t m, float* out, void* stream) {
  return cqs_marginal_probs_batch(amps, n_qubits, is_dm, 1, qubit_axes, m, out, stream);
}

int cqs_marginal_probs_batch(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_states,
                             const int32_t* qubit_axes, int32_t m, float* out, void* stream) {
  CQS_REQUIRE(amps != nullptr && out != nullptr && qubit_axes != nullptr, "NULL pointer");
  CQS_REQUIRE(n_states >= 1 && n_states <= 65535, "n_states=%d out of range (1..65535)", n_states);
  CQS_REQUIRE(n_qubits >= 1 && n_qubits <= 40 && m >= 1 && m <= n_qubits && m <= 31, "bad sizes");
  // The device needs the flat bit of every axis: measured first, then the unmeasured ones.
  // qubit_axes is a HOST array of m Cirq axes.
  AxisBits bits;
  int used[64] = {0};
  for (int k = 0; k < m; ++k) {
    const int a = qubit_axes[k];
    CQS_REQUIRE(a >= 0 && a < n_qubits && !used[a], "bad or repeated axis %d", a);
    used[a] = 1;
    bits.b[k] = n_qubits - 1 - a;
  }
  int w = m;
  for (int a = 0; a < n_qubits; ++a)
    if (!used[a]) bits.b[w++] = n_qubits - 1 - a;
  const Loader ld = make_loader(amps, nullptr, n_qubits, is_dm);
  const uint64_t rows = 1ull << m;
  const uint64_t stride = is_dm ? (1ull << (2 * n_qubits)) : (1ull << n_qubits);  // entries per state
  if (n_qubits - m <= 7) {
    CQS_REQUIRE((rows + 255) / 256 <= 0x7fffffffull, "too many outcomes");
    hipLaunchKernelGGL(k_marginal_short, dim3((unsigned)((rows + 255) / 256), n_states), dim3(256), 0,
                       as_stream(stream), ld, (int)n_qubits, bits, (int)m, rows, out, stride);
  } else {
    CQS_REQUIRE(m <= 24, "marginals over more than 2^24 outcomes with long rows are not supported");
    hipLaunchKernelGGL(k_marginal, dim3((unsigned)rows, n_states), dim3(256), 0, as_stream(stream), ld, (int)n_qubits,
                       bits, (int)m, out, stride);
  }
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}


size_t cqs_draw_rows_workspace_bytes(int32_t m, int32_t n_rows) {
  if (m < 0 || m > 30 || n_rows < 1) return 0;
  const uint64_t len = 1ull << m;
  const uint64_t per = len >= (uint64_t)kNpBuf ? len / kNpBuf : 1;
  return (size_t)(per * (uint64_t)n_rows * sizeof(float));
}

int cqs_draw_rows(const float* probs, int32_t m, int32_t n_rows, const double* uniforms, int64_t* out_index,
                  float* out_p, void* workspace, void* stream) {
  CQS_REQUIRE(probs != nullptr && uniforms != nullptr && out_index != nullptr && out_p != nullptr &&
                  workspace != nullptr, "NULL pointer");
  CQS_REQUIRE(m >= 0 && m <= 30, "m=%d out of range", m);
  CQS_REQUIRE(n_rows >= 1, "n_rows=%d", n_rows);
  const uint64_t len = 1ull << m;
  const uint64_t buflen = len >= (uint64_t)kNpBuf ? (uint64_t)kNpBuf : len;
  const uint64_t nbuf = (len / buflen) * (uint64_t)n_rows;
  CQS_REQUIRE(nbuf <= 0x7fffffffull, "too many buffers");
  hipStream_t s = as_stream(stream);
  const Loader ld = make_loader(nullptr, probs, 0, 0);
  float* bsum = (float*)workspace;
  hipLaunchKernelGGL(k_sum_buffers, dim3((unsigned)nbuf), dim3(256), 0, s, ld, buflen, bsum);
  hipLaunchKernelGGL(k_draw_rows, dim3((unsigned)((n_rows + 63) / 64)), dim3(64), 0, s, probs, (int)m, (int)n_rows,
                     (const float*)bsum, (int)(len / buflen), uniforms, out_index, out_p);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

}  // extern "C"
