// Pauli-string expectations, PauliSum application and norms (FP64 accumulation).
#include <algorithm>
#include <atomic>
#include <numeric>
#include <vector>

#include <string.h>

#include "common.hpp"

namespace cqs {

constexpr int kTermsPerLaunch = 16;
constexpr int kRedBlocks = 512;  // fixed grid => deterministic reduction order

struct TermBatch {
  uint64_t x;                    // shared flip mask
  int count;
  int pad;
  uint64_t z[kTermsPerLaunch];
};

// partials[(blockIdx * 16 + k) * 2 + {0,1}] = this block's share of term k.
__global__ __launch_bounds__(256) void k_expect(const float2* __restrict__ amps, int n, int is_dm,
                                                TermBatch tb, double* __restrict__ partials) {
  __shared__ double red[4][kTermsPerLaunch][2];
  const uint64_t N = 1ull << n;
  double ar[kTermsPerLaunch], ai[kTermsPerLaunch];
#pragma unroll
  for (int k = 0; k < kTermsPerLaunch; ++k) ar[k] = ai[k] = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    double pr, pi;
    if (is_dm) {
      const float2 v = amps[i * N + (i ^ tb.x)];
      pr = v.x;
      pi = v.y;
    } else {
      const float2 a = amps[i];
      const float2 b = amps[i ^ tb.x];
      const double a_r = a.x, a_i = a.y, b_r = b.x, b_i = b.y;
      pr = b_r * a_r + b_i * a_i;  // conj(b) * a
      pi = b_r * a_i - b_i * a_r;
    }
#pragma unroll
    for (int k = 0; k < kTermsPerLaunch; ++k) {
      if (k < tb.count) {
        const bool neg = __popcll(i & tb.z[k]) & 1;
        ar[k] += neg ? -pr : pr;
        ai[k] += neg ? -pi : pi;
      }
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kTermsPerLaunch; ++k) {
    if (k < tb.count) {
      const double sr = wave_sum(ar[k]), si = wave_sum(ai[k]);
      if (lane == 0) {
        red[w][k][0] = sr;
        red[w][k][1] = si;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < tb.count * 2) {
    const int k = threadIdx.x >> 1, c = threadIdx.x & 1;
    const double s = (red[0][k][c] + red[1][k][c]) + (red[2][k][c] + red[3][k][c]);
    partials[((size_t)blockIdx.x * kTermsPerLaunch + k) * 2 + c] = s;
  }
}

// partials[2 * blockIdx + {0,1}] = this block's share of sum_i conj(a_i) b_i (FP64, fixed order).
__global__ __launch_bounds__(256) void k_inner(const float2* __restrict__ a, const float2* __restrict__ b,
                                               uint64_t N, double* __restrict__ partials) {
  __shared__ double red[4][2];
  double sr = 0.0, si = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const float2 x = a[i], y = b[i];
    const double xr = x.x, xi = x.y, yr = y.x, yi = y.y;
    sr += xr * yr + xi * yi;
    si += xr * yi - xi * yr;
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[w][0] = sr;
    red[w][1] = si;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int c = threadIdx.x;
    partials[2 * blockIdx.x + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
  }
}

__global__ __launch_bounds__(64) void k_inner_final(const double* __restrict__ partials, int nblocks,
                                                    double* __restrict__ out) {
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 64) {
    sr += partials[2 * b];
    si += partials[2 * b + 1];
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (threadIdx.x == 0) {
    out[0] = sr;
    out[1] = si;
  }
}

struct FinalBatch {
  int count;
  int nblocks;
  int out_idx[kTermsPerLaunch];
  double cr[kTermsPerLaunch], ci[kTermsPerLaunch];
};

__global__ __launch_bounds__(64) void k_expect_final(const double* __restrict__ partials, FinalBatch fb,
                                                     double* __restrict__ out) {
  const int lane = threadIdx.x;
  for (int k = 0; k < fb.count; ++k) {
    double sr = 0.0, si = 0.0;
    for (int b = lane; b < fb.nblocks; b += 64) {
      sr += partials[((size_t)b * kTermsPerLaunch + k) * 2 + 0];
      si += partials[((size_t)b * kTermsPerLaunch + k) * 2 + 1];
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) {
      out[2 * fb.out_idx[k] + 0] = fb.cr[k] * sr - fb.ci[k] * si;
      out[2 * fb.out_idx[k] + 1] = fb.cr[k] * si + fb.ci[k] * sr;
    }
  }
}

// Terms without X or Y (x mask 0): sum_i (-1)^popcount(i & z) |psi_i|^2, up to 64 terms per read
// of the state.  Each thread takes 16 consecutive amplitudes i = 16g + j, so the sign splits into
// (-1)^popcount(g & z>>4), the same for the 16, times (-1)^popcount(j & z&15), the same for every
// thread: one FP64 FMA per amplitude and distinct low mask.  A thread's groups are
// g = lo + k * stride (lo < stride, stride = 2^sbits the grid's threads), so the group sign is
// parity(lo & z>>4), fixed per thread, times parity(k * stride & z>>4), which depends only on the
// z bits at 4 + sbits and above.  Terms that agree in their low 4 bits and in those high bits form
// one class: a thread accumulates one FP64 sum per class (lane j evaluates the iteration sign of
// class j and a ballot shares it) and gives each term its class sum with the term's fixed sign
// (negation is exact, so every term's sum is the same as with the sign applied per group).
// |psi_i|^2 is exact in FP64 up to its final rounding.  Which thread sums which groups does not
// depend on the batch, so every term's reduction order is fixed.
constexpr int kZTerms = 64;       // terms per launch
constexpr int kZBatch = kZTerms;
constexpr int kZClasses = 24;     // classes per launch (FP64 accumulators in registers)
constexpr int kZBlocks = 2048;    // fixed grid => deterministic reduction order
struct ZBatch {
  int count;                      // terms
  int nclass;                     // classes
  uint64_t newrun;                // bit j: class j's low 4 bits differ from class j - 1's (bit 0 set)
  uint64_t z[kZBatch];            // the terms' z masks
  uint64_t zc[kZClasses];         // a z mask of each class (its low 4 bits and high bits)
  uint8_t cls[kZBatch];           // term -> class
};

__global__ __launch_bounds__(256) void k_expect_z(const float2* __restrict__ amps, int n, ZBatch zb,
                                                  double* __restrict__ partials) {
  __shared__ double red[4][kZTerms];
  __shared__ int zls[kZClasses];  // low 4 z bits per class, read inside the loop (not kept in registers)
  const int tid = threadIdx.x;
  const int cnt = zb.count, ncl = zb.nclass;
  if (tid < kZClasses) zls[tid] = tid < ncl ? (int)(zb.zc[tid] & 15) : 0;
  __syncthreads();
  double acc[kZClasses];
#pragma unroll
  for (int j = 0; j < kZClasses; ++j) acc[j] = 0.0;
  const uint64_t G = 1ull << (n - 4);
  const uint64_t stride = (uint64_t)gridDim.x * 256;  // 2^19 whenever G >= 2^19 (full grid)
  const uint64_t lo = (uint64_t)blockIdx.x * 256 + tid;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t zh_lane = lane < ncl ? zb.zc[lane] >> 4 : 0;
  const uint64_t iters = G > stride ? G / stride : 1;  // every thread runs them all: the ballot sees all lanes
  for (uint64_t k = 0; k < iters; ++k) {
    const uint64_t kpar = __ballot(__popcll((k * stride) & zh_lane) & 1);
    const uint64_t g = lo + k * stride;
    if (g >= G) continue;
    const float4* src = reinterpret_cast<const float4*>(amps + 16 * g);
    double p[16];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float4 v = src[q];
      p[2 * q] = fma((double)v.x, (double)v.x, (double)v.y * (double)v.y);
      p[2 * q + 1] = fma((double)v.z, (double)v.z, (double)v.w * (double)v.w);
    }
    // The host orders classes by their low 4 z bits, so the signed 16-term sum is formed once per
    // run of equal low masks (a uniform branch) and shared by the run's classes.  The run's mask is
    // read from LDS there (a barrier to code motion keeps the masks out of the registers).
    asm volatile("" ::: "memory");
    double w = 0.0;
#pragma unroll
    for (int j = 0; j < kZClasses; ++j) {
      if (j < ncl) {
        if ((zb.newrun >> j) & 1) {
          const int zl = __builtin_amdgcn_readfirstlane(zls[j]);
          w = 0.0;
#pragma unroll
          for (int q = 0; q < 16; ++q) w = fma((__popc(q & zl) & 1) ? -1.0 : 1.0, p[q], w);
        }
        acc[j] += ((kpar >> j) & 1) ? -w : w;
      }
    }
  }
  // each term: its class's sum with the term's fixed sign parity(lo & z>>4)
  for (int t = 0; t < cnt; ++t) {
    const int c = zb.cls[t];
    double v = acc[0];
#pragma unroll
    for (int j = 1; j < kZClasses; ++j) v = c == j ? acc[j] : v;
    const double sv = wave_sum((__popcll(lo & (zb.z[t] >> 4)) & 1) ? -v : v);
    if (lane == 0) red[wv][t] = sv;
  }
  __syncthreads();
  if (tid < cnt) partials[(size_t)blockIdx.x * kZBatch + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
}

struct ZFinal {
  int count;
  int nblocks;
  int out_idx[kZBatch];
  double cr[kZBatch], ci[kZBatch];
};

// One wave per term (workgroup k sums term k's block partials in a fixed order).
__global__ __launch_bounds__(64) void k_expect_z_final(const double* __restrict__ partials, ZFinal fb,
                                                       double* __restrict__ out) {
  const int k = blockIdx.x;
  double sv = 0.0;
  for (int b = threadIdx.x; b < fb.nblocks; b += 64) sv += partials[(size_t)b * kZBatch + k];
  sv = wave_sum(sv);
  if (threadIdx.x == 0) {
    out[2 * fb.out_idx[k] + 0] = fb.cr[k] * sv;
    out[2 * fb.out_idx[k] + 1] = fb.ci[k] * sv;
  }
}

constexpr int kApplyTerms = 32;
struct ApplyBatch {
  int count;
  int accumulate;
  uint64_t x[kApplyTerms];
  uint64_t z[kApplyTerms];
  double cr[kApplyTerms], ci[kApplyTerms];
};

// phi[j] (+)= sum_t c_t (-1)^popcount((j ^ x_t) & z_t) psi[j ^ x_t]
__global__ __launch_bounds__(256) void k_apply_paulisum(const float2* __restrict__ psi, float2* __restrict__ phi,
                                                        uint64_t N, ApplyBatch ab) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride) {
    double sr = 0.0, si = 0.0;
    if (ab.accumulate) {
      const float2 o = phi[j];
      sr = o.x;
      si = o.y;
    }
    for (int t = 0; t < ab.count; ++t) {
      const uint64_t src = j ^ ab.x[t];
      const float2 v = psi[src];
      const double s = (__popcll(src & ab.z[t]) & 1) ? -1.0 : 1.0;
      const double vr = s * v.x, vi = s * v.y;
      sr += ab.cr[t] * vr - ab.ci[t] * vi;
      si += ab.cr[t] * vi + ab.ci[t] * vr;
    }
    phi[j] = make_float2((float)sr, (float)si);
  }
}

// Term checks shared by the expectation entry points.
static int check_terms(int32_t n_qubits, int32_t is_dm, int32_t n_terms, const uint64_t* xmask, const uint64_t* zmask,
                       const double* coef) {
  CQS_REQUIRE(n_qubits >= 0 && n_qubits <= (is_dm ? 20 : CQS_MAX_BITS), "n_qubits=%d out of range", n_qubits);
  CQS_REQUIRE(n_terms >= 0, "n_terms < 0");
  if (n_terms == 0) return CQS_OK;
  CQS_REQUIRE(xmask && zmask && coef, "NULL term arrays");
  const uint64_t valid = n_qubits >= 64 ? ~0ull : ((1ull << n_qubits) - 1);
  for (int t = 0; t < n_terms; ++t)
    CQS_REQUIRE(!(xmask[t] & ~valid) && !(zmask[t] & ~valid), "term %d has masks beyond %d qubits", t,
                n_qubits);
  return CQS_OK;
}

// Launches over the terms in `order` (identity when NULL): runs of terms with one X mask share a
// k_expect launch (up to 16 terms), runs of Z-only terms a k_expect_z launch (up to 64).  Every
// term's sum has a fixed reduction order whatever batch it lands in, so any order of the terms
// gives the same bits.  `partials` holds one launch's per-block partial sums.
static int expect_batches(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_terms, const int* order,
                          const uint64_t* xmask, const uint64_t* zmask, const double* coef, double* out,
                          double* partials, hipStream_t s) {
  auto term = [&](int i) { return order ? order[i] : i; };
  const uint64_t N = 1ull << n_qubits;
  int grid = (int)std::min<uint64_t>(kRedBlocks, (N + 255) / 256);
  if (grid < 1) grid = 1;
  const int zgrid = n_qubits >= 4 ? (int)std::min<uint64_t>(kZBlocks, ((N >> 4) + 255) / 256) : 0;
  for (int i = 0; i < n_terms;) {
    if (!is_dm && n_qubits >= 4 && xmask[term(i)] == 0) {
      ZBatch zb;
      ZFinal zf;
      memset(&zb, 0, sizeof(zb));
      // class key: the low 4 z bits and the z bits that the iteration index reaches (from bit
      // 4 + sbits up; none when the grid covers every group in one iteration)
      const uint64_t G = N >> 4, stride = (uint64_t)zgrid * 256;
      const int sbits = __builtin_ctzll(stride);
      const uint64_t hmask = G > stride ? ~0ull << (4 + sbits) : 0;
      auto key = [&](uint64_t z) { return (z & 15) | (z & hmask); };
      while (i < n_terms && zb.count < kZBatch && xmask[term(i)] == 0) {
        const int t = term(i);
        const uint64_t kz = key(zmask[t]);
        int c = 0;
        while (c < zb.nclass && key(zb.zc[c]) != kz) ++c;
        if (c == zb.nclass) {
          if (zb.nclass == kZClasses) break;  // the next launch takes it
          if (c == 0 || (kz & 15) != (zb.zc[c - 1] & 15)) zb.newrun |= 1ull << c;
          zb.zc[c] = zmask[t];
          ++zb.nclass;
        }
        zb.cls[zb.count] = (uint8_t)c;
        zb.z[zb.count] = zmask[t];
        zf.out_idx[zb.count] = t;
        zf.cr[zb.count] = coef[2 * t];
        zf.ci[zb.count] = coef[2 * t + 1];
        ++zb.count;
        ++i;
      }
      zf.count = zb.count;
      zf.nblocks = zgrid;
      hipLaunchKernelGGL(k_expect_z, dim3(zgrid), dim3(256), 0, s, (const float2*)amps, (int)n_qubits, zb,
                         partials);
      hipLaunchKernelGGL(k_expect_z_final, dim3(zf.count), dim3(64), 0, s, (const double*)partials, zf, out);
      continue;
    }
    TermBatch tb;
    FinalBatch fb;
    tb.x = xmask[term(i)];
    tb.count = 0;
    tb.pad = 0;
    while (i < n_terms && tb.count < kTermsPerLaunch && xmask[term(i)] == tb.x) {
      const int t = term(i);
      tb.z[tb.count] = zmask[t];
      fb.out_idx[tb.count] = t;
      fb.cr[tb.count] = coef[2 * t];
      fb.ci[tb.count] = coef[2 * t + 1];
      ++tb.count;
      ++i;
    }
    for (int k = tb.count; k < kTermsPerLaunch; ++k) tb.z[k] = 0;
    fb.count = tb.count;
    fb.nblocks = grid;
    hipLaunchKernelGGL(k_expect, dim3(grid), dim3(256), 0, s, (const float2*)amps, (int)n_qubits, (int)is_dm,
                       tb, partials);
    hipLaunchKernelGGL(k_expect_final, dim3(1), dim3(64), 0, s, (const double*)partials, fb, out);
  }
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

// Static scratch for cqs_expect_paulis_noalloc: one launch's partial sums, the largest any batch
// writes (2048 blocks x 64 Z terms, or 512 blocks x 16 terms x 2 doubles): 1 MiB per GPU, held
// until the process ends.  In host code `scratch` is only a handle for hipGetSymbolAddress; the
// kernels get the device address that call returns, kept per device.
static __device__ __align__(256) double scratch[kZBatch * kZBlocks > 2 * kTermsPerLaunch * kRedBlocks
                                                    ? kZBatch * kZBlocks
                                                    : 2 * kTermsPerLaunch * kRedBlocks];
constexpr int kMaxDevices = 64;
static std::atomic<double*> g_scratch_ptr[kMaxDevices];

// The current device's address of `scratch`, looked up once per device (hipGetSymbolAddress with
// that device current).
static int scratch_for_current_device(double** out) {
  int dev = 0;
  CQS_HIP(hipGetDevice(&dev));
  CQS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the scratch table covers", dev, kMaxDevices);
  double* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (p == nullptr) {
    void* addr = nullptr;
    CQS_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
    CQS_REQUIRE(addr != nullptr, "hipGetSymbolAddress returned NULL for the expectation scratch");
    p = (double*)addr;
    g_scratch_ptr[dev].store(p, std::memory_order_release);
  }
  *out = p;
  return CQS_OK;
}

}  // namespace cqs

using namespace cqs;

extern "C" {

size_t cqs_expect_workspace_bytes(int32_t n_terms) {
  (void)n_terms;
  return sizeof(double) * std::max(2 * kTermsPerLaunch * kRedBlocks, kZBatch * kZBlocks);
}

int cqs_expect_paulis(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_terms,
                      const uint64_t* xmask, const uint64_t* zmask, const double* coef, double* out,
                      void* workspace, void* stream) {
  CQS_REQUIRE(amps != nullptr && out != nullptr && workspace != nullptr, "NULL pointer");
  if (int rc = check_terms(n_qubits, is_dm, n_terms, xmask, zmask, coef)) return rc;
  if (n_terms == 0) return CQS_OK;
  std::vector<int> order(n_terms);
  std::iota(order.begin(), order.end(), 0);
  // by X mask (terms sharing one launch); Z-only terms also by their low 4 z bits (k_expect_z)
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (xmask[a] != xmask[b]) return xmask[a] < xmask[b];
    return xmask[a] == 0 && (zmask[a] & 15) < (zmask[b] & 15);
  });
  return expect_batches(amps, n_qubits, is_dm, n_terms, order.data(), xmask, zmask, coef, out, (double*)workspace,
                        as_stream(stream));
}

int cqs_expect_paulis_noalloc(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_terms,
                              const uint64_t* xmask, const uint64_t* zmask, const double* coef, double* out,
                              void* stream) {
  CQS_REQUIRE(amps != nullptr && out != nullptr, "NULL pointer");
  if (int rc = check_terms(n_qubits, is_dm, n_terms, xmask, zmask, coef)) return rc;
  if (n_terms == 0) return CQS_OK;
  CQS_REQUIRE(cqs_expect_workspace_bytes(n_terms) <= sizeof(scratch),
              "expectation scratch of %zu bytes exceeds the static %zu-byte array", cqs_expect_workspace_bytes(n_terms),
              sizeof(scratch));
  double* partials = nullptr;
  if (int rc = scratch_for_current_device(&partials)) return rc;
  // terms in the caller's order: runs of equal X masks share launches (no host sort, no allocation)
  return expect_batches(amps, n_qubits, is_dm, n_terms, nullptr, xmask, zmask, coef, out, partials,
                        as_stream(stream));
}

int cqs_norm2(const void* amps, int32_t n_bits, double* out, void* workspace, void* stream) {
  // A 1-term "expectation" of the identity over all 2^n_bits entries, real part in out[0].
  CQS_REQUIRE(amps != nullptr && out != nullptr && workspace != nullptr, "NULL pointer");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  const uint64_t x = 0, z = 0;
  const double c[2] = {1.0, 0.0};
  return cqs_expect_paulis(amps, n_bits, 0, 1, &x, &z, c, out, workspace, stream);
}

int cqs_inner(const void* a, const void* b, int32_t n_bits, double* out, void* workspace, void* stream) {
  CQS_REQUIRE(a != nullptr && b != nullptr && out != nullptr && workspace != nullptr, "NULL pointer");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  const uint64_t N = 1ull << n_bits;
  hipStream_t s = as_stream(stream);
  const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(kRedBlocks, (N + 255) / 256));
  hipLaunchKernelGGL(k_inner, dim3(grid), dim3(256), 0, s, (const float2*)a, (const float2*)b, N, (double*)workspace);
  hipLaunchKernelGGL(k_inner_final, dim3(1), dim3(64), 0, s, (const double*)workspace, grid, out);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_apply_paulisum(const void* psi, void* phi, int32_t n_qubits, int32_t n_terms, const uint64_t* xmask,
                       const uint64_t* zmask, const double* coef, void* stream) {
  CQS_REQUIRE(psi != nullptr && phi != nullptr && psi != phi, "psi/phi must be distinct non-NULL");
  CQS_REQUIRE(n_qubits >= 0 && n_qubits <= CQS_MAX_BITS, "n_qubits=%d out of range", n_qubits);
  CQS_REQUIRE(n_terms >= 0, "n_terms < 0");
  CQS_REQUIRE(n_terms == 0 || (xmask && zmask && coef), "NULL term arrays");
  const uint64_t valid = (1ull << n_qubits) - 1;
  for (int t = 0; t < n_terms; ++t)
    CQS_REQUIRE(!(xmask[t] & ~valid) && !(zmask[t] & ~valid), "term %d has masks beyond %d qubits", t, n_qubits);
  const uint64_t N = 1ull << n_qubits;
  hipStream_t s = as_stream(stream);
  int grid = (int)std::min<uint64_t>(65536, (N + 255) / 256);
  if (grid < 1) grid = 1;
  int done = 0;
  do {
    ApplyBatch ab;
    ab.count = std::min(kApplyTerms, n_terms - done);
    ab.accumulate = done > 0;
    for (int k = 0; k < kApplyTerms; ++k) {
      const bool v = k < ab.count;
      ab.x[k] = v ? xmask[done + k] : 0;
      ab.z[k] = v ? zmask[done + k] : 0;
      ab.cr[k] = v ? coef[2 * (done + k)] : 0.0;
      ab.ci[k] = v ? coef[2 * (done + k) + 1] : 0.0;
    }
    hipLaunchKernelGGL(k_apply_paulisum, dim3(grid), dim3(256), 0, s, (const float2*)psi, (float2*)phi, N, ab);
    done += ab.count;
  } while (done < n_terms);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

}  // extern "C"
