/*
 * cirq_amd.h - C ABI of the MI355X (gfx950) state-vector / density-matrix engine.
 *
 * The engine replaces the per-operation NumPy work that Cirq's simulators do on the host.
 * Every entry point takes raw device pointers (for example a torch tensor's data_ptr()),
 * explicit sizes and a hipStream_t passed as void*.  Calls are stream-ordered; nothing
 * synchronises unless its comment says so.  Every function returns CQS_OK (0) or an error
 * code; cqs_last_error() returns a thread-local message for the last failure.
 * The library never frees caller memory.
 *
 * Layout conventions (identical to Cirq's big-endian flattening):
 *   - A state vector over n qubits is complex64 (float re, float im interleaved), length
 *     2^n.  Qubit axis i of Cirq's qubit_order is flat-index bit (n-1-i)
 *     (cirq-core/cirq/sim/sparse_simulator.py:246-267, reshape(-1) of a (2,)*n tensor).
 *   - A density matrix over n qubits is the same layout over 2n bits: row r, column c at
 *     flat index r*2^n + c, so row qubit i is bit (2n-1-i) and column qubit i is bit (n-1-i)
 *     (cirq-core/cirq/sim/density_matrix_simulation_state.py:177-178: left axes i, right
 *     axes i+n).
 *   - "bit positions" below always mean flat-index bit positions.
 */
#ifndef CIRQ_AMD_H
#define CIRQ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CQS_ABI_VERSION 2

/* status codes */
#define CQS_OK 0
#define CQS_ERR_ARG 1         /* bad argument (maps to ValueError on the Python side)   */
#define CQS_ERR_HIP 2         /* HIP runtime error                                      */
#define CQS_ERR_OOM 3         /* allocation failure (maps to MemoryError)               */
#define CQS_ERR_UNSUPPORTED 4 /* op kind or size the engine does not handle (TypeError) */

/* op kinds inside a pass program */
#define CQS_OP_DENSE 1     /* 2^k x 2^k complex matrix on k tile-local bits (k <= 3)        */
#define CQS_OP_DIAG 2      /* 2^k diagonal on k arbitrary bit positions (k <= 6)            */
#define CQS_OP_PAIRPHASE 3 /* product of phases on bit pairs (a CZ / CZPow layer)           */
#define CQS_OP_SCALE 4     /* multiply every amplitude by one complex number (global phase) */
#define CQS_OP_BLOCK 5     /* fused 2^k x 2^k block (k = 4 or 5) on MFMA, split-FP16 operands  */
#define CQS_OP_GRAD_DENSE 6 /* adjoint only: partial of 2 Re<phi|dU|psi> for a dense dU (k <= 3) */
#define CQS_OP_GRAD_DIAG 7  /* adjoint only: the same for a diagonal dU (DIAG bit encoding)     */
#define CQS_OP_PERMUTE 8    /* re-layout of the tile: tile bit b moves to tile bit aux[t[0] + b] */
/* In adjoint programs bits 8-9 of `kind` select the tile an op acts on: 0 both, 1 psi, 2 phi. */

#define CQS_MAX_LOCAL 16 /* largest tile: 2^16 amplitudes are never used; L <= 14 in practice */
#define CQS_MAX_BITS 48  /* largest flat index width handled by the pass kernels            */

/*
 * One operation of a pass program (64 bytes).
 *   DENSE:     t[0..k-1] are tile-local bit indices (0..L-1); bit m of the matrix row/column
 *              index corresponds to tile bit t[m].  The matrix is row-major at mats[mat].
 *              cl_mask/cl_val: controls on tile-local bits (the op acts only on groups whose
 *              tile index satisfies (j & cl_mask) == cl_val); co_*: controls on bits outside
 *              the tile (global positions, 64-bit masks split in two words).
 *   DIAG:      t[0..k-1] name k bits, each either a tile-local bit index (< 64) or 64 + a flat
 *              bit position outside the tile; bit m of the diagonal index is that bit of the
 *              amplitude's flat index.  Diagonal at mats[mat .. mat+2^k).
 *   PAIRPHASE: k pairs of bits, encoded as for DIAG, at aux[t[0] + 2p], aux[t[0] + 2p + 1]; the
 *              phase of pair p is mats[mat + p] and applies when both bits are 1.
 *   SCALE:     every amplitude times mats[mat].
 *   BLOCK:     like DENSE (targets t[0..k-1], k = 4 or 5) but applied with f16 MFMA: the matrix is
 *              given as its real form R (R[2r+a][2c+b] = Re/Im parts of U[r][c], size 2^(k+1)),
 *              split as R = Rhi + Rlo in FP16 and pre-arranged in MFMA A-operand order:
 *              half8 entries [part][ks][rb][lane] = R_part[rb*32 + lane%32][ks*16 + 8*(lane/32) + e]
 *              starting at mats + mat (16-byte aligned).  Amplitudes are scaled by 2^15 and split
 *              the same way; three products (hi*hi, lo*hi, hi*lo) accumulate in FP32.
 *   PERMUTE:   k = L; tile element j moves to the index whose bit aux[t[0] + b] is bit b of j.
 *              Later ops of the pass name tile bits in the new layout.  Forward passes only.
 *   slot:      parameter slot for gradient bookkeeping (-1 when unused); GRAD ops write their
 *              per-tile partial to grad_partial[slot * partial_stride + tile].
 */
typedef struct cqs_op {
  int32_t kind;
  int32_t k;
  int32_t mat;
  int32_t slot;
  int32_t t[6];
  int32_t cl_mask;
  int32_t cl_val;
  uint32_t co_mask_lo, co_mask_hi;
  uint32_t co_val_lo, co_val_hi;
} cqs_op;

/*
 * One pass: every workgroup gathers the 2^L amplitudes whose flat indices differ only in
 * the L distinct bit positions lpos[0..L-1] (tile index bit b <-> flat bit lpos[b]), runs
 * ops[op_begin .. op_begin+op_count) on them and writes them back through lpos_out (equal to
 * lpos unless the pass holds PERMUTE ops; then it is the layout after the last one).  One pass
 * reads and writes the state once.
 * Thread t of T holds tile elements t + e*T: the lowest log2(T) tile bits are thread bits, the
 * highest log2(2^L / T) are register bits.  DENSE ops with k <= 3 whose targets are all register
 * bits, given in ascending order, run in registers; other DENSE and BLOCK ops exchange the tile
 * through LDS.
 */
typedef struct cqs_pass {
  int32_t L;
  int32_t op_begin;
  int32_t op_count;
  int32_t max_k; /* largest DENSE k in the pass (sizes the LDS matrix stage) */
  int32_t lpos[CQS_MAX_LOCAL];
  int32_t lpos_out[CQS_MAX_LOCAL];
} cqs_pass;

/* ---------------------------------------------------------------- library */

/* ABI version (CQS_ABI_VERSION).  No reference counterpart. */
int cqs_abi_version(void);

/* Thread-local message of the last failing call ("" if none). */
const char* cqs_last_error(void);

/* Number of compute units and total device memory of the current device. */
int cqs_device_info(int32_t* n_cu, uint64_t* total_mem_bytes);

/* ---------------------------------------------------------- state set-up */

/*
 * amps[0 .. 2^n_bits) = 0 except amps[index] = 1.
 * Replaces the initial-state construction of
 * cirq-core/cirq/sim/state_vector_simulation_state.py:49-76 (qis.to_valid_state_vector)
 * and, with n_bits = 2n and index = b*2^n + b, the density-matrix one of
 * cirq-core/cirq/sim/density_matrix_simulation_state.py:41-76.
 */
int cqs_init_basis(void* amps, int32_t n_bits, uint64_t index, void* stream);

/* ------------------------------------------------------------ gate passes */

/*
 * Runs n_passes passes in order on one state (2^n_bits complex64 amplitudes, in place).
 * passes is a HOST array; ops/mats/aux are DEVICE arrays (mats: complex64).
 * Replaces the per-operation loop of cirq-core/cirq/sim/simulator_base.py:198-212 and the
 * NumPy kernels it reaches: linalg.apply_matrix_to_slices
 * (cirq-core/cirq/linalg/transformations.py:310-380), linalg.targeted_left_multiply
 * (cirq-core/cirq/linalg/transformations.py:105-172), the gates' own _apply_unitary_
 * (cirq-core/cirq/ops/common_gates.py:123-133,454-464,658-669,1072-1082) and, on a density
 * matrix, protocols.apply_channel (cirq-core/cirq/protocols/apply_channel_protocol.py:168-356).
 */
int cqs_run_passes(void* amps, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                   const cqs_op* ops, const void* mats, const int32_t* aux, void* stream);

/*
 * Host-side check of a pass program (HOST copies of passes, ops and aux) before it is uploaded
 * for cqs_run_passes / cqs_run_passes_from_basis (adjoint = 0) or cqs_adjoint_passes
 * (adjoint = 1): pass layouts, op kinds, target bits (distinct, inside the tile), local and
 * outer controls, outer-bit matrix variants, DENSE k against the pass's max_k, matrix extents
 * against mats_complex entries, PAIRPHASE / PERMUTE entries against n_aux.  The kernels trust
 * every one of these; a program that fails here would read or write out of bounds.  Returns
 * CQS_ERR_ARG with the first problem in cqs_last_error.  The program takes the place of the
 * op-by-op dispatch of cirq-core/cirq/sim/simulator_base.py:198-212, whose Python objects need
 * no such check.
 */
int cqs_program_validate(int32_t n_bits, const cqs_pass* passes, int32_t n_passes, const cqs_op* ops, int32_t n_ops,
                         int64_t mats_complex, const int32_t* aux, int64_t n_aux, int32_t adjoint);

/*
 * cqs_init_basis(amps, n_bits, basis) followed by cqs_run_passes, with the initial state never
 * written or read: the first pass that runs builds its tiles of |basis> in registers (amps may
 * hold anything on entry).  Tiles the state cannot reach yet are not read, and a pass leaves
 * unwritten the tiles the next pass does not read; every element is written by the time the call's
 * work completes on the stream.  With no pass to run it is cqs_init_basis.  The same initial states as
 * cirq-core/cirq/sim/state_vector_simulation_state.py:49-76 and
 * cirq-core/cirq/sim/density_matrix_simulation_state.py:41-76, one HBM write and read fewer.
 */
int cqs_run_passes_from_basis(void* amps, int32_t n_bits, uint64_t basis, const cqs_pass* passes,
                              int32_t n_passes, const cqs_op* ops, const void* mats, const int32_t* aux,
                              void* stream);

/*
 * cqs_run_passes (basis < 0) or cqs_run_passes_from_basis (basis >= 0) with pass i run by
 * kernels[i], a kernel compiled for that pass's op sequence instead of the op interpreter:
 * kernels[i] is the host launch handle of a `__global__ void (float2*, TileProg)` built by
 * cirq_amd/specialize.py against this library's csrc/tile.hpp (the same device functions,
 * called in the pass's order with its bit positions as constants).  The program arrays are
 * checked as for cqs_run_passes.  Same reference as cqs_run_passes
 * (cirq-core/cirq/sim/simulator_base.py:198-212 and the kernels it reaches).
 */
int cqs_run_passes_specialized(void* amps, int32_t n_bits, int64_t basis, const cqs_pass* passes, int32_t n_passes,
                               const cqs_op* ops, const void* mats, const int32_t* aux,
                               const void* const* kernels, void* stream);

/* ---------------------------------------------------------- adjoint gradient */

/*
 * One reverse sweep of an adjoint program over psi (the final state) and phi (H psi), both
 * updated in place, writing FP64 per-tile partials of every GRAD op.  The program is built by the
 * host from the forward circuit reversed and inverted (cirq_amd/adjoint.py).  Cirq itself has no
 * gradient machinery; callers differentiate expectation values from
 * cirq-core/cirq/sim/simulator.py:349-371 (simulate_expectation_values_sweep) by parameter shifts,
 * which this replaces with one sweep (Jones & Gacon 2020, arXiv:2009.02823).
 */
int cqs_adjoint_passes(void* psi, void* phi, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                       const cqs_op* ops, const void* mats, const int32_t* aux, double* grad_partial,
                       int64_t partial_stride, void* stream);

/* out[g] = sum over t < partial_stride of grad_partial[g * partial_stride + t], g < n_slots. */
/*
 * cqs_adjoint_passes with reverse pass i run by kernels[i], a kernel compiled for that pass's op
 * sequence by cirq_amd/specialize.py (adjoint=True) against this library's csrc/adjoint.hpp:
 * the host launch handle of a `__global__ void (float2* psi, float2* phi, TileProg, double*
 * partial, int64_t stride)`.  Arguments are checked as for cqs_adjoint_passes.  Same reference
 * (the reverse sweep of Jones & Gacon, arXiv:2009.02823; no Cirq counterpart).
 */
int cqs_adjoint_passes_specialized(void* psi, void* phi, int32_t n_bits, const cqs_pass* passes, int32_t n_passes,
                                   const cqs_op* ops, const void* mats, const int32_t* aux, double* grad_partial,
                                   int64_t partial_stride, const void* const* kernels, void* stream);

int cqs_grad_reduce(const double* grad_partial, int32_t n_slots, int64_t partial_stride, double* out, void* stream);

/* ---------------------------------------------------------- batched sweeps */

/*
 * Many states of different sizes, each running its own pass program with its own parameter
 * values, advance together: one launch runs pass p of every state whose program has a pass p with
 * tile size L.  Replaces the per-resolver loop of cirq-core/cirq/sim/simulator_base.py:275-320
 * (simulate_sweep_iter: resolve, copy the initial state, simulate, one point at a time) and its
 * expectation-value form cirq-core/cirq/sim/sparse_simulator.py:193-218.
 * All descriptor arrays below are DEVICE arrays laid out as these structs; cqs_batch_validate
 * checks HOST copies of them before the caller uploads.
 */
typedef struct cqs_batch_pass { /* one pass of one program (352 bytes) */
  int32_t n_bits;
  int32_t L;
  int32_t op_begin; /* into the batch's ops array */
  int32_t op_count;
  int32_t max_k;
  int32_t nouter; /* n_bits - L */
  int32_t pad[2];
  int32_t lpos[CQS_MAX_LOCAL];
  int32_t lpos_out[CQS_MAX_LOCAL];
  int32_t opos[CQS_MAX_BITS]; /* the n_bits - L positions outside the tile, ascending */
} cqs_batch_pass;

typedef struct cqs_batch_item { /* one state's share of one launch (40 bytes) */
  int64_t state;      /* complex64 offset of the state (psi) in the arena */
  int64_t state2;     /* complex64 offset of phi (adjoint launches) */
  int64_t mats;       /* complex64 offset of this state's matrix buffer (16-byte aligned) */
  int64_t partial;    /* adjoint: double offset of this state's [slot][tile] partials */
  int32_t pass;       /* index into the batch passes */
  int32_t tile_begin; /* first workgroup of this item; items are sorted by it */
} cqs_batch_item;

typedef struct cqs_batch_terms { /* one state's Pauli terms (40 bytes) */
  int64_t state;       /* psi */
  int64_t state2;      /* phi (cqs_batch_apply_paulisum) */
  int64_t out;         /* cqs_batch_expect: first partial (complex double) of [term][block] */
  int32_t n_bits;
  int32_t term_begin;  /* into xmask / zmask / coef */
  int32_t n_terms;
  int32_t block_begin; /* first workgroup; each covers 2^min(chunk_bits, n_bits) amplitudes */
} cqs_batch_terms;

/* HOST-side consistency check of a launch's descriptors (positions, offsets within the arenas,
 * contiguous tile numbering, L of every item's pass).  adjoint != 0 also checks state2. */
int cqs_batch_validate(const cqs_batch_pass* passes, int32_t n_passes, int32_t n_ops, const cqs_batch_item* items,
                       int32_t n_items, int32_t L, int64_t arena_elems, int64_t mats_elems, int32_t adjoint);

/* |0...0> in every item's state; item tile_begin counts chunks of 2^min(chunk_bits, n_bits). */
int cqs_batch_init(void* arena, const cqs_batch_item* items, int32_t n_items, int64_t n_chunks,
                   const cqs_batch_pass* passes, int32_t chunk_bits, void* stream);

/* One forward pass level: n_tiles workgroups of tile size L over the items. */
int cqs_batch_passes(void* arena, const cqs_batch_item* items, int32_t n_items, int64_t n_tiles, int32_t L,
                     const cqs_batch_pass* passes, const cqs_op* ops, const void* mats, const int32_t* aux,
                     void* stream);

/* One reverse-sweep level of adjoint programs (cqs_adjoint_passes for many states). */
int cqs_batch_adjoint(void* arena, const cqs_batch_item* items, int32_t n_items, int64_t n_tiles, int32_t L,
                      const cqs_batch_pass* passes, const cqs_op* ops, const void* mats, const int32_t* aux,
                      double* partial, void* stream);

/* phi = sum_t coef_t P_t psi for every item (cqs_apply_paulisum per state); masks and coef
 * (complex double pairs) are DEVICE arrays. */
int cqs_batch_apply_paulisum(void* arena, const cqs_batch_terms* items, int32_t n_items, int64_t n_blocks,
                             const uint64_t* xmask, const uint64_t* zmask, const double* coef, int32_t chunk_bits,
                             void* stream);

/* Per item, term and block: partial[out + t * blocks_of_item + b] = sum over the block's indices i
 * of (-1)^popcount(i & z_t) conj(psi[i ^ x_t]) psi[i] (complex double; coefficients are applied
 * by the caller).  Reduce with cqs_segment_sum. */
int cqs_batch_expect(const void* arena, const cqs_batch_terms* items, int32_t n_items, int64_t n_blocks,
                     const uint64_t* xmask, const uint64_t* zmask, double* partial, int32_t chunk_bits,
                     void* stream);

/* out[s] = sum of in[seg[s] .. seg[s+1]) in a fixed order, for width = 1 (double) or 2 (complex
 * double, offsets counted in complex elements).  seg: DEVICE int64[n_seg + 1]. */
int cqs_segment_sum(const double* in, const int64_t* seg, int32_t n_seg, int32_t width, double* out, void* stream);

/* ---------------------------------------------------------- reductions */

/*
 * Pauli-string expectations, FP64 accumulation.  Term t has flip mask xmask[t] (qubits with
 * X or Y), sign mask zmask[t] (qubits with Z or Y) and complex coefficient coef[2t],coef[2t+1]
 * that already includes i^(number of Y).  xmask/zmask/coef are HOST arrays; out is DEVICE
 * double[2*n_terms] (re, im).  For a state vector (is_dm = 0):
 *   out[t] = coef_t * sum_i (-1)^popcount(i & zmask) conj(psi[i ^ xmask]) psi[i]
 * which is cirq-core/cirq/ops/pauli_string.py:625-656 summed as in
 * cirq-core/cirq/ops/linear_combinations.py:717-719.  For a density matrix over n qubits
 * (is_dm = 1, amps has 2^(2n) entries):
 *   out[t] = coef_t * sum_r (-1)^popcount(r & zmask) rho[r, r ^ xmask]
 * which is cirq-core/cirq/ops/pauli_string.py:734-764 (before its final real part).
 * workspace: DEVICE scratch of cqs_expect_workspace_bytes(n_terms) bytes.
 */
size_t cqs_expect_workspace_bytes(int32_t n_terms);
int cqs_expect_paulis(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_terms,
                      const uint64_t* xmask, const uint64_t* zmask, const double* coef,
                      double* out, void* workspace, void* stream);

/*
 * cqs_expect_paulis without a workspace argument and without any allocation (no hipMalloc,
 * hipFree or allocating library call, host or device): the per-block partial sums live in a
 * static 1 MiB device array per GPU, located once per device with hipGetSymbolAddress and kept
 * until the process ends.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * Terms are batched in the caller's order (runs of one X mask share a launch); the values are
 * bit-identical to cqs_expect_paulis.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (different streams or threads) race and must be serialised by the caller.
 * Sizes beyond the array are refused with CQS_ERR_ARG (the array holds the largest batch any
 * n_terms needs, so this does not happen for valid arguments).
 */
int cqs_expect_paulis_noalloc(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_terms,
                              const uint64_t* xmask, const uint64_t* zmask, const double* coef,
                              double* out, void* stream);

/*
 * phi = sum_t coef_t P_t psi (the PauliSum applied to a state vector), complex64 out; the
 * term arrays are HOST arrays as above.
 * Used by the adjoint gradient.  No direct reference counterpart: Cirq evaluates each term
 * separately (cirq-core/cirq/ops/pauli_string.py:648-656).
 */
int cqs_apply_paulisum(const void* psi, void* phi, int32_t n_qubits, int32_t n_terms,
                       const uint64_t* xmask, const uint64_t* zmask, const double* coef,
                       void* stream);

/* out[0] + i out[1] = sum_i conj(a[i]) b[i] over 2^n_bits complex64 entries, FP64 accumulation in
 * a fixed order (DEVICE doubles); workspace: DEVICE scratch of cqs_expect_workspace_bytes(1) bytes.
 * The adjoint gradient's <H> = s Re<psi|phi> once phi = (H / s) psi is formed, the expectation
 * cirq-core/cirq/ops/linear_combinations.py:659-719 (PauliSum.expectation_from_state_vector)
 * computes term by term. */
int cqs_inner(const void* a, const void* b, int32_t n_bits, double* out, void* workspace, void* stream);

/* out[0] = sum_i |amps[i]|^2 in FP64 (DEVICE double).  Norm check of
 * cirq-core/cirq/qis/states.py:865-897 and cirq-core/cirq/sim/state_vector_simulator.py:131-146. */
int cqs_norm2(const void* amps, int32_t n_bits, double* out, void* workspace, void* stream);

/* amps[i] = ((i & mask) == value) ? amps[i] * (re + i im) : 0: projection onto a measurement
 * outcome and renormalisation (cirq-core/cirq/sim/state_vector.py:235-322,
 * cirq-core/cirq/sim/density_matrix_utils.py:97-182 with the masks of both row and column bits). */
int cqs_collapse(void* amps, int32_t n_bits, uint64_t mask, uint64_t value, float re, float im, void* stream);

/* amps *= scale (complex64) over 2^n_bits entries: the renormalisation of
 * cirq-core/cirq/sim/state_vector_simulator.py:139-146. */
int cqs_scale(void* amps, int32_t n_bits, float re, float im, void* stream);

/* ---------------------------------------------------------- sampling */

/*
 * Sampling all qubits of a state (is_dm = 0: |psi_i|^2; is_dm = 1: Re rho[i,i]), bit-exact
 * with the NumPy path of cirq-core/cirq/sim/state_vector.py:219-226 /
 * cirq-core/cirq/sim/density_matrix_utils.py:84-89:
 *   p = (psi * conj(psi)).real (float32), p /= p.sum() (NumPy's float32 summation order),
 *   cdf = cumsum(float64(p)) (sequential), cdf /= cdf[-1],
 *   idx = searchsorted(cdf, u, 'right') for the caller's uniforms u (RandomState.random_sample).
 * cqs_sample_prepare builds the per-chunk CDF state in workspace; cqs_sample_draw maps
 * uniforms (DEVICE double[reps]) to indices (DEVICE int64[reps]).
 * probs_in (optional, DEVICE float32[2^n_qubits]): use these probabilities instead of the
 * amplitudes (marginal measurements); pass NULL otherwise.
 * qubit_axes (optional, HOST int32[n_qubits]): measure all qubits in this order (Cirq axes; the
 * first is the most significant bit of the drawn index), the transpose of
 * cirq-core/cirq/sim/simulation_utils.py:44-48 with nothing summed; NULL = natural order.
 */
size_t cqs_sample_workspace_bytes(int32_t n_qubits);
int cqs_sample_prepare(const void* amps, const float* probs_in, int32_t n_qubits, int32_t is_dm,
                       const int32_t* qubit_axes, void* workspace, void* stream);
int cqs_sample_draw(const void* amps, const float* probs_in, int32_t n_qubits, int32_t is_dm,
                    const int32_t* qubit_axes, const double* uniforms, int64_t reps, int64_t* out_index,
                    void* workspace, void* stream);

/*
 * Marginal probabilities over measured qubits, in NumPy's order: p (float32, as above),
 * transposed so that measured qubits come first in the order given (qubit_axes[0..m) are Cirq
 * axes), then summed over the rest (cirq-core/cirq/sim/simulation_utils.py:41-60).
 * qubit_axes: HOST int32[m].  out: DEVICE float32[2^m], not normalised (cqs_sample_prepare
 * normalises).
 */
int cqs_marginal_probs(const void* amps, int32_t n_qubits, int32_t is_dm, const int32_t* qubit_axes,
                       int32_t m, float* out, void* stream);

/*
 * Batched trajectories.  n_states copies of a state (one per repetition of a circuit whose
 * measurements or state-vector channels are not all terminal) lie one after the other in amps,
 * copy r at amps + r * 2^n_bits entries (2^(2 n_qubits) for density matrices).  They replace the
 * per-repetition loop of cirq-core/cirq/sim/simulator_base.py:248-273: the deterministic stretches
 * run on all copies as one pass program (cqs_run_passes over n_bits + log2(n_states) bits), and
 * the calls below do what differs per copy.  Per-copy arrays are DEVICE arrays of n_states
 * entries; 1 <= n_states <= 65535.
 *
 * cqs_marginal_probs_batch: cqs_marginal_probs of every copy, out[r * 2^m ..] (the same float32
 * arithmetic per copy), for measure_state_vector / measure_density_matrix
 * (cirq-core/cirq/sim/state_vector.py:235-322, cirq-core/cirq/sim/density_matrix_utils.py:97-182).
 */
int cqs_marginal_probs_batch(const void* amps, int32_t n_qubits, int32_t is_dm, int32_t n_states,
                             const int32_t* qubit_axes, int32_t m, float* out, void* stream);

/* One draw per row of probs (DEVICE float32, n_rows rows of 2^m unnormalised marginals, e.g.
 * from cqs_marginal_probs_batch), each as NumPy's RandomState.choice(2^m, p) makes it for the
 * row's uniform (DEVICE double[n_rows]): p = clip(row, 0) / its NumPy float32 sum, the float64
 * sequential cumsum normalised by its last entry, searchsorted right
 * (cirq-core/cirq/sim/state_vector.py:235-322 draws one measurement this way).  out_index: DEVICE
 * int64[n_rows]; out_p: DEVICE float32[n_rows], the drawn outcome's normalised probability.
 * workspace: DEVICE, cqs_draw_rows_workspace_bytes bytes. */
size_t cqs_draw_rows_workspace_bytes(int32_t m, int32_t n_rows);
int cqs_draw_rows(const float* probs, int32_t m, int32_t n_rows, const double* uniforms, int64_t* out_index,
                  float* out_p, void* workspace, void* stream);

/* Copy r: amps_r[i] = ((i & mask) == values[r]) ? amps_r[i] * scales[r] : 0 (each copy's
 * measurement outcome projected and renormalised, as cqs_collapse does for one state). */
int cqs_collapse_batch(void* amps, int32_t n_bits, int32_t n_states, uint64_t mask, const uint64_t* values,
                       const float* scales, void* stream);

/* Copy r: the k-qubit matrix mats[select[r]] (complex64, row-major 2^k x 2^k, matrix index bit m
 * <-> flat bit targets[m]; targets: HOST int32[k], 1 <= k <= 3) times scales[r], on every group of
 * 2^k amplitudes.  The unitary a mixture drew (cirq-core/cirq/sim/state_vector_simulation_state.py:183-203)
 * or the Kraus operator a channel drew, renormalised (:205-257).  Copies whose select is outside
 * [0, n_mats) are left unchanged. */
int cqs_apply_select_batch(void* amps, int32_t n_bits, int32_t n_states, int32_t k, const int32_t* targets,
                           const void* mats, int32_t n_mats, const int32_t* select, const float* scales,
                           void* stream);

/* out[r * n_mats + i] = ||K_i psi_r||^2 in FP64 (DEVICE double[n_states * n_mats], fixed-order
 * sums; 1 <= n_mats <= 16): the weights cirq-core/cirq/sim/state_vector_simulation_state.py:224-250
 * walks to pick a Kraus operator.  workspace: DEVICE, cqs_kraus_workspace_bytes bytes. */
size_t cqs_kraus_workspace_bytes(int32_t n_bits, int32_t n_states, int32_t k, int32_t n_mats);

/* A run of per-copy channel choices in one sweep of the copies (mixtures whose draws are known:
 * cirq-core/cirq/sim/state_vector_simulation_state.py:183-203 for each in order).  Op i acts on
 * op_k[i] (1 or 2) flat bits op_targets[2i], op_targets[2i + 1] (matrix index bit m <-> bit
 * op_targets[2i + m]); copy r applies matrix table[op_row[i] * n_states + r] of the op's set of
 * op_n_mats[i] row-major 2^k x 2^k complex64 matrices at mats + op_mat[i] (complex entries).  All
 * targets together span at most 4 bits; 1 <= n_ops <= 16.  op_*: HOST int32 arrays; mats
 * (mat_entries complex64 entries) and table (table_rows x n_states int32): DEVICE.  Copies whose
 * entry is outside the op's set skip that op. */
int cqs_select_run_batch(void* amps, int32_t n_bits, int32_t n_states, int32_t n_ops, const int32_t* op_k,
                         const int32_t* op_targets, const int32_t* op_mat, const int32_t* op_n_mats,
                         const int32_t* op_row, const void* mats, int64_t mat_entries, const int32_t* table,
                         int32_t table_rows, void* stream);

/* Two one-qubit Kraus channels in a row, A on flat bit bit_a then B on bit_b, in one sweep of the
 * copies: applies A's chosen operators (as cqs_apply_select_batch with select_a / scales_a, DEVICE)
 * and writes B's weights ||K_i psi_r||^2 (as cqs_kraus_weights_batch) to out (DEVICE
 * double[n_states * n_mats_b]).  workspace: cqs_kraus_workspace_bytes(n_bits, n_states, 1,
 * n_mats_b) bytes.  The walk of cirq-core/cirq/sim/state_vector_simulation_state.py:205-257 for a
 * run of channels, one read and write of the copies per channel. */
int cqs_select_then_kraus_weights(void* amps, int32_t n_bits, int32_t n_states, int32_t bit_a, const void* mats_a,
                                  int32_t n_mats_a, const int32_t* select_a, const float* scales_a, int32_t bit_b,
                                  const void* mats_b, int32_t n_mats_b, double* out, void* workspace, void* stream);

/* Copy r's Kraus operator from its weights (DEVICE double[n_states * n_mats], as
 * cqs_kraus_weights_batch writes them) and uniform (DEVICE double[n_states]), by the walk of
 * cirq-core/cirq/sim/state_vector_simulation_state.py:224-250: select[r] (DEVICE int32) and
 * scales[r] = 1 / sqrt(its weight) (DEVICE float32), ready for cqs_apply_select_batch. */
int cqs_kraus_choose_batch(const double* weights, const double* uniforms, int32_t n_states, int32_t n_mats,
                           int32_t* select, float* scales, void* stream);
int cqs_kraus_weights_batch(const void* amps, int32_t n_bits, int32_t n_states, int32_t k, const int32_t* targets,
                            const void* mats, int32_t n_mats, double* out, void* workspace, void* stream);

/*
 * out[j] (DEVICE float32[2^n_qubits]) = the probability of the basis state whose big-endian digits,
 * qubit qubit_axes[0] first, spell j: p = float32(|amp|^2) rounded as NumPy rounds it (state
 * vector) or Re rho[k, k] (density matrix).  The transpose Cirq's product state applies when it
 * merges factors before sampling (cirq-core/cirq/sim/simulation_product_state.py:68-81, 154-172),
 * done once with contiguous reads and writes; pass out as probs_in to cqs_sample_prepare / _draw
 * (with qubit_axes NULL).  qubit_axes: HOST int32[n_qubits], a permutation.
 */
int cqs_permuted_probs(const void* amps, int32_t n_qubits, int32_t is_dm, const int32_t* qubit_axes, float* out,
                       void* stream);

/*
 * Index -> big-endian digits: out[r*m + k] = bit (m-1-k) of index[r]
 * (cirq-core/cirq/value/digits.py:125-170 as used by cirq-core/cirq/sim/state_vector.py:228-232).
 */
int cqs_indices_to_bits(const int64_t* index, int64_t reps, int32_t m, uint8_t* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CIRQ_AMD_H */
