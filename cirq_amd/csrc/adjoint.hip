// Adjoint differentiation: one reverse sweep over psi (the forward state) and phi = H psi, FP64
// per-tile partials reduced by cqs_grad_reduce.  The device code is in adjoint.hpp.
#include "adjoint.hpp"

namespace cqs {

template <int L>
__global__ __launch_bounds__(TileShape<L>::NT) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_adjoint_pass(float2* __restrict__ psi, float2* __restrict__ phi,
                                                                 TileProg P, double* __restrict__ partial,
                                                                 int64_t stride) {
  extern __shared__ float2 lds[];
  const uint64_t base = deposit_bounded<31>((uint64_t)blockIdx.x, P.opos, P.nouter);
  adjoint_tile<L>(psi, phi, P, P.lpos, base, partial, stride, blockIdx.x, lds);
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

static int adjoint_passes(void* psi, void* phi, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                          const cqs_op* ops, const void* mats, const int32_t* aux, double* grad_partial,
                          int64_t partial_stride, const void* const* kernels, void* stream) {
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
    const int64_t entries = (int64_t)tiles * adjoint_waves(ps.L);  // one partial per wave of every tile
    CQS_REQUIRE(entries <= partial_stride, "partial_stride %lld < tiles x waves %lld", (long long)partial_stride,
                (long long)entries);
    const int mk = ps.max_k;
    const size_t stage = mk >= 3 ? (size_t(1) << (2 * mk)) : 0;
    const size_t shmem = (2 * (size_t)tile_words(ps.L) + stage) * sizeof(float2);
    CQS_REQUIRE(shmem <= 160 * 1024 - 1024, "adjoint pass needs %zu bytes of LDS", shmem);
    if (shmem > 64 * 1024)
      CQS_HIP(hipFuncSetAttribute((const void*)kAdjKernels[ps.L], hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shmem));
    const int nt = ps.L <= 12 ? 256 : (1 << ps.L) / 16;
    if (kernels != nullptr) {
      CQS_REQUIRE(kernels[i] != nullptr, "no specialised kernel for reverse pass %d", i);
      if (shmem > 64 * 1024)
        CQS_HIP(hipFuncSetAttribute(kernels[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
      float2* a = (float2*)psi;
      float2* b = (float2*)phi;
      void* args[] = {&a, &b, &P, &grad_partial, &partial_stride};
      CQS_HIP(hipLaunchKernel(kernels[i], dim3((unsigned)tiles), dim3(nt), args, shmem, s));
    } else {
      hipLaunchKernelGGL(kAdjKernels[ps.L], dim3((unsigned)tiles), dim3(nt), shmem, s, (float2*)psi, (float2*)phi,
                         P, grad_partial, partial_stride);
    }
    CQS_HIP(hipGetLastError());
  }
  return CQS_OK;
}

int cqs_adjoint_passes(void* psi, void* phi, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                       const cqs_op* ops, const void* mats, const int32_t* aux, double* grad_partial,
                       int64_t partial_stride, void* stream) {
  return adjoint_passes(psi, phi, n_bits, passes, n_passes, ops, mats, aux, grad_partial, partial_stride, nullptr,
                        stream);
}

int cqs_adjoint_passes_specialized(void* psi, void* phi, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                                   const cqs_op* ops, const void* mats, const int32_t* aux, double* grad_partial,
                                   int64_t partial_stride, const void* const* kernels, void* stream) {
  CQS_REQUIRE(kernels != nullptr || n_passes == 0, "kernels is NULL");
  return adjoint_passes(psi, phi, n_bits, passes, n_passes, ops, mats, aux, grad_partial, partial_stride, kernels,
                        stream);
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
