// Library plumbing, state set-up and the gate-pass kernel.
#include <stdarg.h>
#include <string.h>

#include <mutex>
#include <unordered_map>
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
  // Workgroups are dispatched to the 8 XCDs round-robin (b mod 8); tile b' = (b mod 8) G/8 + b/8
  // gives each XCD a contiguous eighth of the tiles, so the workgroups resident on one XCD touch
  // neighbouring addresses (measured 1.6% faster over the 30-qubit grid circuit's passes).
  uint32_t b = blockIdx.x;
  const uint32_t G = gridDim.x;
  if ((G & 7) == 0) b = (b & 7) * (G >> 3) + (b >> 3);
  const uint64_t base = deposit_bounded<31>((uint64_t)b, P.opos, P.nouter);
  pass_tile<L>(psi, P, P.lpos, P.lpos_out, base, lds, P.basis, P.zmask, P.zval, P.wmask, P.wval);
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
// Passes of at most 2^kWarmOuterBits tiles (a few workgroups per CU) warm their program first.
constexpr int kWarmOuterBits = 11;

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
  P.warm = P.nouter <= kWarmOuterBits;
  P.basis = -1;
  P.zmask = 0;
  P.zval = 0;
  P.wmask = 0;
  P.wval = 0;
  P.ops = ops + ps.op_begin;
  P.mats = mats;
  P.aux = aux;
  return true;
}

// A DIAG / PAIRPHASE position code: a tile bit (< L) or kOuter + a flat position outside the tile.
static bool valid_code(int c, int L, int n, uint64_t local) {
  if (c >= 0 && c < L) return true;
  const int g = c - kOuter;
  return c >= kOuter && g < n && !((local >> g) & 1ull);
}

// Every op of one pass: kinds, target bits, matrix / aux extents, controls and outer-bit variants
// (the checks the kernels' indexing assumes).  mats_c: complex entries in the matrix buffer.
bool validate_pass_ops(int n, const cqs_pass& ps, const cqs_op* ops, int32_t n_ops, int64_t mats_c,
                       const int32_t* aux, int64_t n_aux, bool adjoint) {
  const int L = ps.L;
  if (ps.op_begin < 0 || ps.op_count < 0 || (int64_t)ps.op_begin + ps.op_count > n_ops) {
    set_error("pass ops [%d, %d) outside the %d-op table", ps.op_begin, ps.op_begin + ps.op_count, n_ops);
    return false;
  }
  uint64_t local = 0;
  for (int b = 0; b < L; ++b) local |= 1ull << ps.lpos[b];
  const uint64_t nmask = n >= 64 ? ~0ull : ((1ull << n) - 1);
  for (int i = ps.op_begin; i < ps.op_begin + ps.op_count; ++i) {
    const cqs_op& o = ops[i];
    const int kind = o.kind & 0xff, sel = (o.kind >> 8) & 3;
    if ((o.kind & ~0x3ff) != 0 || sel == 3) {
      set_error("op %d: kind word 0x%x has bits set outside the kind and tile selector", i, o.kind);
      return false;
    }
    if (sel != 0 && !adjoint) {
      set_error("op %d: tile selector bits (kind word 0x%x) belong to adjoint programs only", i, o.kind);
      return false;
    }
    if ((kind == CQS_OP_GRAD_DENSE || kind == CQS_OP_GRAD_DIAG) && !adjoint) {
      set_error("op %d: gradient ops belong to adjoint programs only", i);
      return false;
    }
    const uint64_t com = ((uint64_t)o.co_mask_hi << 32) | o.co_mask_lo;
    const uint64_t cov = ((uint64_t)o.co_val_hi << 32) | o.co_val_lo;
    if ((com & ~nmask) || (com & local) || (cov & ~com)) {
      set_error("op %d: outer controls must name bits outside the tile (mask 0x%llx)", i, (unsigned long long)com);
      return false;
    }
    if (L < 31 && ((o.cl_mask & ~((1 << L) - 1)) || (o.cl_val & ~o.cl_mask))) {
      set_error("op %d: local controls outside the %d-bit tile", i, L);
      return false;
    }
    auto need_mats = [&](int64_t count) {
      if (o.mat < 0 || (int64_t)o.mat + count > mats_c) {
        set_error("op %d: matrix entries [%d, %lld) outside the %lld-entry buffer", i, o.mat,
                  (long long)(o.mat + count), (long long)mats_c);
        return false;
      }
      return true;
    };
    switch (kind) {
      case CQS_OP_DENSE:
      case CQS_OP_BLOCK:
      case CQS_OP_GRAD_DENSE: {
        const bool block = kind == CQS_OP_BLOCK;
        if (block ? (o.k < 4 || o.k > 5) : (o.k < 1 || o.k > 3)) {
          set_error("op %d: k=%d out of range for its kind %d", i, o.k, kind);
          return false;
        }
        if (kind == CQS_OP_DENSE && o.k > ps.max_k) {
          set_error("op %d: dense k=%d exceeds the pass's max_k=%d (the LDS matrix stage)", i, o.k, ps.max_k);
          return false;
        }
        int seen = 0;
        for (int m = 0; m < o.k; ++m) {
          if (o.t[m] < 0 || o.t[m] >= L || ((seen >> o.t[m]) & 1)) {
            set_error("op %d: targets must be distinct tile bits < %d", i, L);
            return false;
          }
          seen |= 1 << o.t[m];
        }
        if (seen & o.cl_mask) {
          set_error("op %d: a local control is also a target", i);
          return false;
        }
        const int code = o.t[5], cnt = (code >> 24) & 3;
        if (code & ~((3 << 24) | ((1 << 21) - 1))) {
          set_error("op %d: bad variant code 0x%x", i, code);
          return false;
        }
        for (int v = 0; v < cnt; ++v) {
          const int pos = (code >> (7 * v)) & 127;
          if (pos >= n || ((local >> pos) & 1ull)) {
            set_error("op %d: variant bit %d is not an outer bit", i, pos);
            return false;
          }
        }
        const int64_t per = block ? (int64_t)2 << (2 * o.k) : (int64_t)1 << (2 * o.k);
        if (block && (o.mat & 1)) {
          set_error("op %d: MFMA block operands must start 16-byte aligned", i);
          return false;
        }
        if (!need_mats(per << cnt)) return false;
        break;
      }
      case CQS_OP_DIAG:
      case CQS_OP_GRAD_DIAG:
        if (o.k < 1 || o.k > 6) {
          set_error("op %d: diagonal on k=%d bits", i, o.k);
          return false;
        }
        for (int m = 0; m < o.k; ++m)
          if (!valid_code(o.t[m], L, n, local)) {
            set_error("op %d: diagonal bit code %d", i, o.t[m]);
            return false;
          }
        if (!need_mats((int64_t)1 << o.k)) return false;
        break;
      case CQS_OP_PAIRPHASE:
        if (o.k < 1 || aux == nullptr || o.t[0] < 0 || (int64_t)o.t[0] + 2 * o.k > n_aux) {
          set_error("op %d: %d pairs at aux[%d] outside the %lld-entry aux table", i, o.k, o.t[0], (long long)n_aux);
          return false;
        }
        for (int m = 0; m < 2 * o.k; ++m)
          if (!valid_code(aux[o.t[0] + m], L, n, local)) {
            set_error("op %d: pair bit code %d", i, aux[o.t[0] + m]);
            return false;
          }
        if (!need_mats(o.k)) return false;
        break;
      case CQS_OP_SCALE:
        if (!need_mats(1)) return false;
        break;
      case CQS_OP_PERMUTE: {
        if (adjoint || o.k != L || aux == nullptr || o.t[0] < 0 || (int64_t)o.t[0] + L > n_aux) {
          set_error("op %d: PERMUTE needs k = L and L aux entries (forward programs only)", i);
          return false;
        }
        int seen = 0;
        for (int b = 0; b < L; ++b) {
          const int d = aux[o.t[0] + b];
          if (d < 0 || d >= L || ((seen >> d) & 1)) {
            set_error("op %d: PERMUTE destinations are not a permutation of the tile bits", i);
            return false;
          }
          seen |= 1 << d;
        }
        break;
      }
      default:
        set_error("op %d: unknown kind %d", i, kind);
        return false;
    }
  }
  return true;
}

}  // namespace cqs

using namespace cqs;

extern "C" {

int cqs_program_validate(int32_t n_bits, const cqs_pass* passes, int32_t n_passes, const cqs_op* ops, int32_t n_ops,
                         int64_t mats_complex, const int32_t* aux, int64_t n_aux, int32_t adjoint) {
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_passes >= 0 && (n_passes == 0 || passes != nullptr), "bad pass list");
  CQS_REQUIRE(n_ops >= 0 && (n_ops == 0 || ops != nullptr), "bad op table");
  for (int i = 0; i < n_passes; ++i) {
    const cqs_pass& ps = passes[i];
    if (ps.op_count <= 0) continue;
    TileProg P;
    if (!make_prog(P, n_bits, ps, ops, nullptr, aux)) return CQS_ERR_ARG;
    if (adjoint)
      for (int b = 0; b < ps.L; ++b)
        CQS_REQUIRE(ps.lpos_out[b] == ps.lpos[b], "pass %d: adjoint passes cannot re-lay their tiles", i);
    if (!validate_pass_ops(n_bits, ps, ops, n_ops, mats_complex, aux, n_aux, adjoint != 0)) {
      std::string msg = "pass " + std::to_string(i) + ": " + g_last_error;
      g_last_error = msg;
      return CQS_ERR_ARG;
    }
  }
  return CQS_OK;
}

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

// Dynamic LDS above 64 KB must be granted per kernel; specialised kernels are granted once each.
static int ensure_kernel_lds(const void* fn, size_t shmem) {
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> granted;
  if (shmem <= 64 * 1024) return CQS_OK;
  CQS_REQUIRE(shmem <= 160 * 1024, "pass needs %zu bytes of LDS (> 160 KB)", shmem);
  std::lock_guard<std::mutex> lock(mu);
  size_t& g = granted[fn];
  if (shmem > g) {
    CQS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    g = shmem;
  }
  return CQS_OK;
}

static int run_passes(void* amps, int32_t n_bits, int64_t basis, const cqs_pass* passes, int32_t n_passes,
                      const cqs_op* ops, const void* mats, const int32_t* aux, void* stream,
                      const void* const* kernels = nullptr) {
  CQS_REQUIRE(amps != nullptr, "amps is NULL");
  CQS_REQUIRE(n_bits >= 0 && n_bits <= CQS_MAX_BITS, "n_bits=%d out of range", n_bits);
  CQS_REQUIRE(n_passes >= 0 && (n_passes == 0 || passes != nullptr), "bad pass list");
  CQS_REQUIRE(basis < 0 || (uint64_t)basis < (1ull << n_bits), "basis state %lld out of range",
              (long long)basis);
  hipStream_t s = as_stream(stream);
  int64_t start = basis;  // consumed by the first pass that runs
  // From a basis state, the bits that no pass so far had local still hold the basis's values:
  // the state is zero elsewhere, so tiles that disagree on them are skipped (written as zeros, and
  // not written at all where the next pass does not read them either).
  const uint64_t all = n_bits >= 64 ? ~0ull : (1ull << n_bits) - 1;
  uint64_t fixed = basis >= 0 ? all : 0;
  for (int i = 0; i < n_passes; ++i) {
    const cqs_pass& ps = passes[i];
    if (ps.op_count <= 0) continue;
    CQS_REQUIRE(ops != nullptr && mats != nullptr, "ops/mats are NULL");
    TileProg P;
    if (!make_prog(P, n_bits, ps, ops, (const float2*)mats, aux)) return CQS_ERR_ARG;
    P.basis = start;
    start = -1;
    uint64_t local = 0;
    for (int b = 0; b < ps.L; ++b) local |= 1ull << ps.lpos[b];
    P.zmask = fixed & ~local;
    P.zval = (uint64_t)basis & P.zmask;
    fixed &= ~local;
    // The next pass that runs reads only the tiles its own (zmask, zval) keeps and stores zeros
    // over the rest: this pass leaves the rest unwritten.  Its mask lies within this pass's outer
    // bits (fixed no longer holds this pass's local bits), so whole tiles are kept or left.
    for (int j = i + 1; fixed != 0 && j < n_passes; ++j) {
      if (passes[j].op_count <= 0) continue;
      CQS_REQUIRE(passes[j].L >= 0 && passes[j].L <= CQS_MAX_LOCAL, "pass %d: L=%d out of range", j, passes[j].L);
      uint64_t next = 0;
      for (int b = 0; b < passes[j].L; ++b) {
        CQS_REQUIRE(passes[j].lpos[b] >= 0 && passes[j].lpos[b] < n_bits, "pass %d: lpos[%d]=%d out of range", j,
                    b, passes[j].lpos[b]);
        next |= 1ull << passes[j].lpos[b];
      }
      P.wmask = fixed & ~next;
      P.wval = (uint64_t)basis & P.wmask;
      break;
    }
    const uint64_t tiles = 1ull << P.nouter;
    CQS_REQUIRE(tiles <= 0x7fffffffull, "too many tiles (n=%d, L=%d)", n_bits, ps.L);
    const int mk = ps.max_k;
    const size_t stage = mk >= 3 ? (size_t(1) << (2 * mk)) : 0;
    const size_t shmem = ((size_t)tile_words(ps.L) + stage) * sizeof(float2);
    if (kernels != nullptr) {
      CQS_REQUIRE(kernels[i] != nullptr, "no specialised kernel for pass %d", i);
      if (int rc = ensure_kernel_lds(kernels[i], shmem)) return rc;
      float2* psi = (float2*)amps;
      void* args[] = {&psi, &P};
      CQS_HIP(hipLaunchKernel(kernels[i], dim3((unsigned)tiles), dim3(threads_for_L(ps.L)), args, shmem, s));
    } else {
      if (int rc = ensure_pass_lds(ps.L, shmem)) return rc;
      hipLaunchKernelGGL(kPassKernels[ps.L], dim3((unsigned)tiles), dim3(threads_for_L(ps.L)), shmem, s,
                         (float2*)amps, P);
    }
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

int cqs_run_passes_specialized(void* amps, int32_t n_bits, int64_t basis, const cqs_pass* passes, int32_t n_passes,
                               const cqs_op* ops, const void* mats, const int32_t* aux,
                               const void* const* kernels, void* stream) {
  CQS_REQUIRE(kernels != nullptr || n_passes == 0, "kernels is NULL");
  return run_passes(amps, n_bits, basis < 0 ? -1 : basis, passes, n_passes, ops, mats, aux, stream, kernels);
}

}  // extern "C"
