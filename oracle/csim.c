/* Test infrastructure: complex128 state-vector and density-matrix kernels in plain C (OpenMP).
 *
 * These restate the same contractions as the NumPy oracle (oracle/cirq_np.py), for sizes where the
 * NumPy restatement takes hours (28-qubit states, 12-qubit density matrices over thousands of
 * ops).  They are checkers only: tests and the fixture scripts under tests/golden/ call them.
 *
 *   csim_apply          cirq-core/cirq/linalg/transformations.py:105-172 (targeted_left_multiply:
 *                       a 2^k x 2^k matrix on k target axes, big-endian: the first target axis is
 *                       the most significant bit of the matrix index)
 *   csim_apply_channel  cirq-core/cirq/protocols/apply_channel_protocol.py:297-356
 *                       (sum_k K rho K^dagger, left axes i, right axes i + n,
 *                       cirq-core/cirq/sim/density_matrix_simulation_state.py:177-178)
 *
 * Layout: a state over n qubits is complex128[2^n] with qubit axis a at flat bit n-1-a (Cirq's
 * C-order flattening of the (2,)*n tensor); a density matrix is complex128[4^n] with row axis a at
 * flat bit 2n-1-a and column axis a at flat bit n-1-a.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define KMAX 6

/* Index of group g (over the n_bits - k non-target bits) with all target bits zero. */
static inline uint64_t spread(uint64_t g, const int *sorted_bits, int k) {
    for (int i = 0; i < k; i++) {
        uint64_t b = sorted_bits[i];
        uint64_t lo = g & ((1ULL << b) - 1);
        g = ((g >> b) << (b + 1)) | lo;
    }
    return g;
}

static void sort_bits(int *b, int k) {
    for (int i = 1; i < k; i++)
        for (int j = i; j > 0 && b[j - 1] > b[j]; j--) {
            int t = b[j];
            b[j] = b[j - 1];
            b[j - 1] = t;
        }
}

/* offsets[j] = flat offset of matrix index j, whose bit (k-1-i) is the value of target bit tb[i]. */
static void offsets_for(const int *tb, int k, uint64_t *off) {
    for (int j = 0; j < (1 << k); j++) {
        uint64_t o = 0;
        for (int i = 0; i < k; i++)
            if ((j >> (k - 1 - i)) & 1) o |= 1ULL << tb[i];
        off[j] = o;
    }
}

/* psi <- M on the flat bits tb[0..k) (tb[0] the most significant matrix bit). */
static int apply_bits(cplx *psi, int n_bits, int k, const int *tb, const cplx *M) {
    if (k < 1 || k > KMAX || n_bits < k) return -1;
    int sb[KMAX];
    memcpy(sb, tb, sizeof(int) * k);
    sort_bits(sb, k);
    for (int i = 0; i < k; i++)
        if (sb[i] < 0 || sb[i] >= n_bits || (i && sb[i] == sb[i - 1])) return -2;
    uint64_t off[1 << KMAX];
    offsets_for(tb, k, off);
    const int D = 1 << k;
    const int64_t groups = (int64_t)1 << (n_bits - k);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < groups; g++) {
        cplx v[1 << KMAX];
        uint64_t base = spread((uint64_t)g, sb, k);
        for (int j = 0; j < D; j++) v[j] = psi[base | off[j]];
        for (int r = 0; r < D; r++) {
            cplx acc = 0;
            for (int c = 0; c < D; c++) acc += M[r * D + c] * v[c];
            psi[base | off[r]] = acc;
        }
    }
    return 0;
}

int csim_apply(cplx *psi, int n, int k, const int *axes, const cplx *M) {
    int tb[KMAX];
    if (k < 1 || k > KMAX) return -1;
    for (int i = 0; i < k; i++) tb[i] = n - 1 - axes[i];
    return apply_bits(psi, n, k, tb, M);
}

/* rho <- sum_q K_q rho K_q^dagger on qubit axes `axes` of an n-qubit density matrix. */
int csim_apply_channel(cplx *rho, int n, int k, const int *axes, int nk, const cplx *kraus) {
    if (k < 1 || 2 * k > KMAX || nk < 1) return -1;
    int rb[KMAX], cb[KMAX], all[2 * KMAX];
    for (int i = 0; i < k; i++) {
        rb[i] = 2 * n - 1 - axes[i];
        cb[i] = n - 1 - axes[i];
        all[i] = rb[i];
        all[k + i] = cb[i];
    }
    sort_bits(all, 2 * k);
    for (int i = 0; i < 2 * k; i++)
        if (all[i] < 0 || all[i] >= 2 * n || (i && all[i] == all[i - 1])) return -2;
    const int D = 1 << k;
    uint64_t ro[1 << KMAX], co[1 << KMAX];
    offsets_for(rb, k, ro);
    offsets_for(cb, k, co);
    const int64_t groups = (int64_t)1 << (2 * n - 2 * k);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < groups; g++) {
        cplx R[8 * 8], T[8 * 8], O[8 * 8];
        uint64_t base = spread((uint64_t)g, all, 2 * k);
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++) R[r * D + c] = rho[base | ro[r] | co[c]];
        for (int i = 0; i < D * D; i++) O[i] = 0;
        for (int q = 0; q < nk; q++) {
            const cplx *K = kraus + (size_t)q * D * D;
            for (int r = 0; r < D; r++) /* T = K R */
                for (int c = 0; c < D; c++) {
                    cplx a = 0;
                    for (int m = 0; m < D; m++) a += K[r * D + m] * R[m * D + c];
                    T[r * D + c] = a;
                }
            for (int r = 0; r < D; r++) /* O += T K^dagger */
                for (int c = 0; c < D; c++) {
                    cplx a = 0;
                    for (int m = 0; m < D; m++) a += T[r * D + m] * conj(K[c * D + m]);
                    O[r * D + c] += a;
                }
        }
        for (int r = 0; r < D; r++)
            for (int c = 0; c < D; c++) rho[base | ro[r] | co[c]] = O[r * D + c];
    }
    return 0;
}

/* sum over basis states of sign(popcount(i & zmask)) |psi_i|^2, for each of n_terms masks. */
void csim_expect_z(const cplx *psi, int n_bits, int n_terms, const uint64_t *zmask, double *out) {
    const int64_t N = (int64_t)1 << n_bits;
    for (int t = 0; t < n_terms; t++) {
        double s = 0;
        const uint64_t z = zmask[t];
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (int64_t i = 0; i < N; i++) {
            double p = creal(psi[i]) * creal(psi[i]) + cimag(psi[i]) * cimag(psi[i]);
            s += (__builtin_popcountll((uint64_t)i & z) & 1) ? -p : p;
        }
        out[t] = s;
    }
}
