// Library plumbing, state set-up and the gate-pass kernel.
#include <stdarg.h>
#include <string.h>

#include <vector>

#include "tile.hpp"

namespace cqs {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail_hip(hipError_t e, const char* what) {
  set_error("%s failed: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? CQS_ERR_OOM : CQS_ERR_HIP;
}

// ------------------------------------------------------------------ kernels

__global__ void k_init_basis(float2* __restrict__ a, uint64_t n, uint64_t index) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = make_float2(i == index ? 1.f : 0.f, 0.f);
}

__global__ void k_scale(float2* __restrict__ a, uint64_t n, float2 s) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = cmul(a[i], s);
}

// One pass over a single state: workgroup b owns tile b (2^L amplitudes).  The tile goes from HBM
// straight into registers and back; LDS only carries the ops that mix elements across threads.
// Every tile width runs 16 waves per CU (4 per SIMD, 128 VGPRs): 4 workgroups of 256 threads up to
// L = 12, 2 of 512 at L = 13, 1 of 1024 at L = 14 (the LDS tile allows no more).
template <int L>
__global__ __launch_bounds__(TileShape<L>::NT) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_pass(
    float2* __restrict__ psi, TileProg P) {
  extern __shared__ float2 lds[];
  const uint64_t base = deposit_bounded<31>((uint64_t)blockIdx.x, P.opos, P.nouter);
  pass_tile<L>(psi, P, P.lpos, P.lpos_out, base, lds, P.basis);
}

// amps[i] = ((i & mask) == value) ? amps[i] * s : 0
__global__ void k_collapse(float2* __restrict__ a, uint64_t n, uint64_t mask, uint64_t value, float2 s) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    a[i] = ((i & mask) == value) ? cmul(a[i], s) : make_float2(0.f, 0.f);
}

typedef void (*PassKernel)(float2*, TileProg);
static const PassKernel kPassKernels[15] = {k_pass<0>,  k_pass<1>,  k_pass<2>,  k_pass<3>, k_pass<4>,
                                            k_pass<5>,  k_pass<6>,  k_pass<7>,  k_pass<8>, k_pass<9>,
                                            k_pass<10>, k_pass<11>, k_pass<12>, k_pass<13>, k_pass<14>};
static int threads_for_L(int L) { return L <= 12 ? 256 : (1 << L) / 16; }

static int grid_for(uint64_t n) {
  uint64_t g = (n + WG - 1) / WG;
  return (int)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

// Fill a TileProg for one pass over a state of n bits.  Returns false on a bad pass.
bool make_prog(TileProg& P, int n, const cqs_pass& ps, const cqs_op* ops, const float2* mats,
               const int32_t* aux) {
  if (ps.L < 0 || ps.L > 14 || ps.L > n) {
    set_error("pass has L=%d (n=%d); need 0 <= L <= min(14, n)", ps.L, n);
    return false;
  }
  P.n = n;
  P.L = ps.L;
  P.nouter = n - ps.L;
  P.op_count = ps.op_count;
  uint64_t used = 0, used_out = 0;
  for (int b = 0; b < ps.L; ++b) {
    const int p = ps.lpos[b], q = ps.lpos_out[b];
    if (p < 0 || p >= n || ((used >> p) & 1ull)) {
      set_error("pass local positions must be distinct and < n (bit %d = %d)", b, p);
      return false;
    }
    if (q < 0 || q >= n || ((used_out >> q) & 1ull)) {
      set_error("pass output positions must be distinct and < n (bit %d = %d)", b, q);
      return false;
    }
    used |= 1ull << p;
    used_out |= 1ull << q;
    P.lpos[b] = p;
    P.lpos_out[b] = q;
  }
  if (used != used_out) {
    set_error("pass output positions are not a permutation of its local positions");
    return false;
  }
  int k = 0;
  for (int p = 0; p < n; ++p)
    if (!((used >> p) & 1ull)) P.opos[k++] = p;
  if (ps.max_k < 0 || ps.max_k > 5) {
    set_error("pass max_k=%d out of range", ps.max_k);
    return false;
  }
  P.max_k = ps.max_k;
  P.mscale = 32768.f;
  P.basis = -1;
  P.ops = ops + ps.op_begin;
  P.mats = mats;
  P.aux = aux;
  return true;
}

}  // namespace cqs

using namespace cqs;

extern "C" {

int cqs_abi_version(void) { return CQS_ABI_VERSION; }

const char* cqs_last_error(void) { return g_last_error.c_str(); }

int cqs_device_info(int32_t* n_cu, uint64_t* total_mem_bytes) {
  int dev = 0;
  CQS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CQS_HIP(hipGetDeviceProperties(&prop, dev));
  if (n_cu) *n_cu = prop.multiProcessorCount;
  if (total_mem_bytes) *total_mem_bytes = prop.totalGlobalMem;
  return CQS_OK;
}

int cqs_init_basis(void* amps, int32_t n_bits, uint64_t index, void* stream) {
  CQS_REQUIRE(amps != nullptr, "amps is NULL");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  const uint64_t n = 1ull << n_bits;
  CQS_REQUIRE(index < n, "basis index %llu >= 2^%d", (unsigned long long)index, n_bits);
  hipLaunchKernelGGL(k_init_basis, dim3(grid_for(n)), dim3(WG), 0, as_stream(stream),
                     (float2*)amps, n, index);
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_scale(void* amps, int32_t n_bits, float re, float im, void* stream) {
  CQS_REQUIRE(amps != nullptr, "amps is NULL");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  const uint64_t n = 1ull << n_bits;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(WG), 0, as_stream(stream), (float2*)amps, n,
                     make_float2(re, im));
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

int cqs_collapse(void* amps, int32_t n_bits, uint64_t mask, uint64_t value, float re, float im,
                 void* stream) {
  CQS_REQUIRE(amps != nullptr, "amps is NULL");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  const uint64_t n = 1ull << n_bits;
  hipLaunchKernelGGL(k_collapse, dim3(grid_for(n)), dim3(WG), 0, as_stream(stream), (float2*)amps, n, mask,
                     value, make_float2(re, im));
  CQS_HIP(hipGetLastError());
  return CQS_OK;
}

// Launches needing more than the default 64 KB of dynamic LDS raise the kernel's limit first.
static int ensure_pass_lds(int L, size_t shmem) {
  static size_t granted[15] = {0};
  if (shmem > 64 * 1024 && shmem > granted[L]) {
    CQS_REQUIRE(shmem <= 160 * 1024, "pass needs %zu bytes of LDS (> 160 KB)", shmem);
    CQS_HIP(hipFuncSetAttribute((const void*)kPassKernels[L], hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shmem));
    granted[L] = shmem;
  }
  return CQS_OK;
}

static int run_passes(void* amps, int32_t n_bits, int64_t basis, const cqs_pass* passes, int32_t n_passes,
                      const cqs_op* ops, const void* mats, const int32_t* aux, void* stream) {
  CQS_REQUIRE(amps != nullptr, "amps is NULL");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_passes >= 0 && (n_passes == 0 || passes != nullptr), "bad pass list");
  CQS_REQUIRE(basis < 0 || (uint64_t)basis < (1ull << n_bits), "basis state %lld out of range",
              (long long)basis);
  hipStream_t s = as_stream(stream);
  int64_t start = basis;  // consumed by the first pass that runs
  for (int i = 0; i < n_passes; ++i) {
    const cqs_pass& ps = passes[i];
    if (ps.op_count <= 0) continue;
    CQS_REQUIRE(ops != nullptr && mats != nullptr, "ops/mats are NULL");
    TileProg P;
    if (!make_prog(P, n_bits, ps, ops, (const float2*)mats, aux)) return CQS_ERR_ARG;
    P.basis = start;
    start = -1;
    const uint64_t tiles = 1ull << P.nouter;
    CQS_REQUIRE(tiles <= 0x7fffffffull, "too many tiles (n=%d, L=%d)", n_bits, ps.L);
    const int mk = ps.max_k;
    const size_t stage = mk >= 3 ? (size_t(1) << (2 * mk)) : 0;
    const size_t shmem = ((size_t)tile_words(ps.L) + stage) * sizeof(float2);
    if (int rc = ensure_pass_lds(ps.L, shmem)) return rc;
    hipLaunchKernelGGL(kPassKernels[ps.L], dim3((unsigned)tiles), dim3(threads_for_L(ps.L)), shmem, s,
                       (float2*)amps, P);
    CQS_HIP(hipGetLastError());
  }
  if (start >= 0) return cqs_init_basis(amps, n_bits, (uint64_t)start, stream);  // no pass ran
  return CQS_OK;
}

int cqs_run_passes(void* amps, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                   const cqs_op* ops, const void* mats, const int32_t* aux, void* stream) {
  return run_passes(amps, n_bits, -1, passes, n_passes, ops, mats, aux, stream);
}

int cqs_run_passes_from_basis(void* amps, int32_t n_bits, uint64_t basis, const cqs_pass* passes,
                              int32_t n_passes, const cqs_op* ops, const void* mats, const int32_t* aux,
                              void* stream) {
  CQS_REQUIRE(basis < (1ull << 62), "basis state out of range");
  return run_passes(amps, n_bits, (int64_t)basis, passes, n_passes, ops, mats, aux, stream);
}

}  // extern "C"
