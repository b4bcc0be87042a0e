/*
 * wqc_tn.h — C ABI of the MI355X (gfx950) MPS tensor-network engine.
 *
 * This is the drop-in boundary for wqc-core's TN backend seam. Every entry point
 * replaces one method of the reference's `MpsState` (src/tn/mps.rs) or one piece of
 * the backend selection (src/tn/backend.rs, src/tn/engine_status.rs); the citation on
 * each declaration names the reference item it stands in for.
 *
 * Conventions
 *  - Plain C types only: opaque handles, pointers, sizes. No exceptions cross the ABI.
 *  - Every function that can fail returns a wqc_tn_status. The thread-local message of
 *    the last failure is available from wqc_tn_last_error().
 *  - Complex numbers are interleaved doubles (re, im).
 *  - Site tensors use the reference layout (src/tn/mps.rs:56): a site of shape
 *    (l, 2, r) is row-major, element (a, s, b) at index a*2r + s*r + b.
 *  - Qubit q of a basis index is bit q (src/tn/mps.rs:742-780).
 *  - The library owns device memory; the caller owns every host buffer it passes.
 *  - A state handle is single-threaded; a device handle may be shared by states
 *    created on the same thread. Calls are synchronous at the ABI: results are ready
 *    when a function returns.
 */
#ifndef WQC_TN_H
#define WQC_TN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference returns Result<_, EngineError> (src/engine.rs:68-78). */
typedef enum wqc_tn_status {
    WQC_TN_OK = 0,
    WQC_TN_INVALID_ARG = 1,     /* InvalidQubitCount, QubitIndexOutOfBounds, bad sizes      */
    WQC_TN_OUT_OF_MEMORY = 2,   /* InsufficientMemory (src/tn/mps.rs:36-48), on device VRAM  */
    WQC_TN_EXEC_FAILED = 3,     /* ExecutionFailed(String): numeric or runtime failure       */
    WQC_TN_CANCELLED = 4,       /* cancel::check (src/cancel.rs:52-63)                       */
    WQC_TN_BOND_EXCEEDED = 5,   /* "mps bond exceeded chi" (src/tn/mps.rs:1137-1143)         */
    WQC_TN_NO_DEVICE = 6        /* no usable gfx950 device                                   */
} wqc_tn_status;

/* Gate kinds. The numeric values are the reference's own gate ids
 * (Gate::to_stark_id, src/engine.rs:284-300); MEASURE/RESET/IF follow. */
typedef enum wqc_tn_gate_kind {
    WQC_TN_GATE_X = 1,
    WQC_TN_GATE_Y = 2,
    WQC_TN_GATE_Z = 3,
    WQC_TN_GATE_H = 4,
    WQC_TN_GATE_S = 5,
    WQC_TN_GATE_T = 6,
    WQC_TN_GATE_CNOT = 7,   /* q0 = control, q1 = target */
    WQC_TN_GATE_CZ = 8,     /* q0 = control, q1 = target */
    WQC_TN_GATE_CCNOT = 9,  /* q0, q1 = controls, q2 = target */
    WQC_TN_GATE_RX = 10,    /* q0 = target, angle */
    WQC_TN_GATE_RY = 11,
    WQC_TN_GATE_RZ = 12,
    WQC_TN_GATE_MEASURE = 13, /* no-op on the state (src/tn/mps.rs:294) */
    WQC_TN_GATE_RESET = 14,   /* no-op on the state (src/tn/mps.rs:294) */
    WQC_TN_GATE_IF = 15       /* no-op on the state (src/tn/mps.rs:294) */
} wqc_tn_gate_kind;

/* One gate of src/engine.rs:257-279 (Gate enum), flattened. */
typedef struct wqc_tn_gate {
    int32_t kind;   /* wqc_tn_gate_kind */
    int32_t q0;
    int32_t q1;
    int32_t q2;
    double angle;   /* RX/RY/RZ only */
} wqc_tn_gate;

/* One independent slice of a batch: a |0...0> register of qubit_count qubits with bond cap
 * max_bond_dim, driven by gates[gate_offset .. gate_offset + gate_count) of a shared array. */
typedef struct wqc_tn_slice {
    size_t qubit_count;
    size_t max_bond_dim;
    size_t gate_offset;
    size_t gate_count;
} wqc_tn_slice;

/* Arithmetic of the two-site split.
 *  FP64       : plain complex128 everywhere (the drop-in form).
 *  MIXED_FP32 : the Jacobi SVD runs its early sweeps on complex64 rows (FP32 MFMA Gram and
 *               rotation, FP32 eigensolve) while the accumulated rotation stays unitary in
 *               FP64; the rows are then recomputed in FP64 and the last sweeps, the
 *               singular values, the truncation decision and the factors are FP64.
 *  MIXED_BF16 : as MIXED_FP32, with the early sweeps' Gram matrices on BF16 MFMA in split-BF16
 *               (each FP32 operand as two bf16, three products, FP32 accumulation); their
 *               rotations stay FP32, the finish FP64. */
typedef enum wqc_tn_precision {
    WQC_TN_PRECISION_FP64 = 0,
    WQC_TN_PRECISION_MIXED_FP32 = 1,
    WQC_TN_PRECISION_MIXED_BF16 = 2
} wqc_tn_precision;

typedef struct wqc_tn_device wqc_tn_device;
typedef struct wqc_tn_state wqc_tn_state;

/* Counters and timings of one state (WorkReport fields and test counters). */
typedef struct wqc_tn_stats {
    uint64_t two_qubit_units;     /* apply_two_qubit_at_positions calls (src/tn/mps.rs:377) */
    uint64_t one_qubit_gates;
    uint64_t exact_splits;        /* svd_split_two_site (src/tn/mps.rs:1182)              */
    uint64_t rsvd_kept;           /* RSVD_KEPT (src/tn/mps.rs:797-801)                    */
    uint64_t rsvd_fallbacks;      /* RSVD_FALLBACKS                                       */
    uint64_t svd_check_failures;  /* SVD_RETRIES analogue (src/tn/svd_guard.rs:141-144)   */
    uint64_t jacobi_sweeps;       /* total outer Jacobi sweeps run                        */
    uint64_t truncated_splits;    /* splits that dropped a singular value above 1e-12     */
    uint64_t peak_vram_bytes;     /* vram_peak_bytes (src/tn/mps.rs:157-166)              */
    double max_discarded_weight;  /* max over splits of sum_{k>=chi} s_k^2 / sum s_k^2     */
    double sum_discarded_weight;
    uint64_t jacobi_inner_sweeps; /* sweeps of the per-block-pair eigensolves               */
    uint64_t jacobi_pair_solves;  /* block-pair eigensolves run                             */
    uint64_t jacobi_sweeps_lo;    /* low-precision sweeps of the mixed form                 */
    uint64_t jacobi_stalled;      /* Jacobi jobs stopped at their rounding floor            */
    uint64_t jacobi_unconverged;  /* Jacobi jobs still rotating at the sweep cap (60); kept,
                                     like the reference's Jacobi at its cap, only when the
                                     split passes the norm and probe-vector checks of
                                     svd_reproduces (src/tn/svd_guard.rs:242-283)             */
} wqc_tn_stats;

/* ---- errors ------------------------------------------------------------------ */

/* Thread-local text of the last failure on this thread ("" when none). */
const char *wqc_tn_last_error(void);

/* ---- backend seam (src/tn/backend.rs:12-29, src/tn/engine_status.rs:17-117) ---- */

/* Parses WQC_TN_BACKEND like tn_backend_from_env: returns "hip" for "hip"/"gpu"/"rocm",
 * otherwise "cpu" (the reference CPU path stays the caller's). Never fails. */
const char *wqc_tn_backend_from_env(void);

/* Probe like probe_tn_engine_status: writes "hip" and the device name into `active`
 * (capacity `cap`), the usable VRAM budget into *vram_budget_bytes (cf. the
 * `vram_budget_bytes` of /sysinfo, src/tn/engine_status.rs:105-117). */
int wqc_tn_engine_status(char *active, size_t cap, uint64_t *vram_budget_bytes);

/* ---- device ------------------------------------------------------------------ */

/* Open device `ordinal` (0 = first visible GPU). Replaces GpuMpsDevice::try_new
 * (src/tn/gpu/device.rs:68-214). */
int wqc_tn_device_open(int ordinal, wqc_tn_device **out);
/* Frees the device's cache of device blocks and pinned staging that freed states gave back
 * (at most 48 GiB of device memory; it saves the next states of the device hipMalloc and
 * hipFree, which waits for the whole device). Every state of the device must be freed first. */
void wqc_tn_device_close(wqc_tn_device *dev);

/* ---- state lifetime ------------------------------------------------------------ */

/* MpsState::try_new_with_bond (src/tn/mps.rs:82-126): |0...0> product state, every
 * site (1,2,1) = [1, 0]. qubit_count == 0 -> INVALID_ARG; max_bond_dim == 0 ->
 * EXEC_FAILED; N*chi^2*32 bytes above the free VRAM -> OUT_OF_MEMORY. */
int wqc_tn_state_new(wqc_tn_device *dev, size_t qubit_count, size_t max_bond_dim,
                     int precision, wqc_tn_state **out);
void wqc_tn_state_free(wqc_tn_state *st);

size_t wqc_tn_qubit_count(const wqc_tn_state *st);
/* pub max_bond_dim field (mutated by src/tn/tree.rs:288). */
int wqc_tn_set_max_bond_dim(wqc_tn_state *st, size_t max_bond_dim);
/* private reject_truncation flag (src/tn/mps.rs:61, set at :1270,1274). */
int wqc_tn_set_reject_truncation(wqc_tn_state *st, int reject);

/* ---- gate application ------------------------------------------------------------ */

/* MpsState::apply_gate (src/tn/mps.rs:278-297). */
int wqc_tn_apply_gate(wqc_tn_state *st, const wqc_tn_gate *gate);

/* The apply loop of trace::execute_with_trace_using (src/tn/trace.rs:71-161) without
 * trace rows: MEASURE/RESET/IF skipped, same-target same-axis rotation runs applied
 * as one net rotation, runs of H applied one by one, and qubit-disjoint gates batched
 * into one device launch per layer (disjoint updates touch disjoint sites, so the
 * result equals gate-by-gate application). This is contract_slice's gate loop
 * (src/tn/contract.rs:10-30). */
int wqc_tn_apply_circuit(wqc_tn_state *st, const wqc_tn_gate *gates, size_t count);

/* MpsState::apply_adjacent_disjoint_gates (src/tn/mps.rs:520-565). */
int wqc_tn_apply_adjacent_disjoint_gates(wqc_tn_state *st, const wqc_tn_gate *gates,
                                         size_t count);

/* Sweep-tree replay hooks (src/tn/mps.rs:621-655). `gate` of apply_unary_at must be a
 * one-qubit gate; apply_pair_at takes CNOT or CZ; site+1 must exist. */
int wqc_tn_apply_unary_at(wqc_tn_state *st, size_t site, const wqc_tn_gate *gate);
int wqc_tn_apply_tdg_at(wqc_tn_state *st, size_t site);
int wqc_tn_apply_swap_at(wqc_tn_state *st, size_t site);
int wqc_tn_apply_pair_at(wqc_tn_state *st, size_t site, const wqc_tn_gate *gate,
                         int ctrl_on_left);
int wqc_tn_apply_cnot_at(wqc_tn_state *st, size_t site, int ctrl_on_left);

/* Cancellation (src/cancel.rs:22-63): sets a flag polled before every two-qubit unit
 * and every layer; the running call then returns WQC_TN_CANCELLED. Thread-safe. */
void wqc_tn_request_cancel(wqc_tn_state *st);
void wqc_tn_clear_cancel(wqc_tn_state *st);

/* ---- readouts ---------------------------------------------------------------------- */

/* <bra|ket> of two states of the same register (overlap, src/tn/mps.rs:1421-1474): the
 * left-to-right environment (ket bond x bra bond) may hold at most 256 entries, else
 * EXEC_FAILED "mps overlap environment exceeds 256"; INVALID_ARG for different registers. */
int wqc_tn_overlap(wqc_tn_state *ket, wqc_tn_state *bra, double out[2]);

/* contract_time_split (src/tn/mps.rs:1241-1249 with time_split_boundaries :1358-1392): the ket
 * is |0...0> through gates[0, split), the bra |0...0> through the adjoints of gates[split, count)
 * in reverse order (adjoint_gate_sequence :1398-1419), both with reject_truncation at
 * max_bond_dim; out = <bra|ket> = <0...0| U |0...0>. EXEC_FAILED when split > count, when the
 * second half holds MEASURE/RESET/IF, or when the overlap environment exceeds 256 entries;
 * BOND_EXCEEDED when either half needs truncation. */
int wqc_tn_contract_time_split(wqc_tn_device *dev, size_t qubit_count, const wqc_tn_gate *gates, size_t count,
                               size_t split, size_t max_bond_dim, int precision, double out[2]);

/* MpsState::amplitude_at_compact_zero (src/tn/mps.rs:262-276): <0...0|psi>, FP64. */
int wqc_tn_amplitude_at_compact_zero(wqc_tn_state *st, double out[2]);

/* <0...0|psi> as wqc_tn_amplitude_at_compact_zero (same kernels, same result), allocating
 * nothing. No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops). Its vectors live in a fixed 2 MiB device
 * array per GPU, enough for every bond up to 1024 (a larger bond -> INVALID_ARG). Every call on
 * one GPU shares that array, so calls that overlap on one GPU race. */
int wqc_tn_amplitude_at_compact_zero_noalloc(wqc_tn_state *st, double out[2]);

/* MpsState::site_amplitudes (src/tn/mps.rs:679-693): out = {a0.re, a0.im, a1.re, a1.im}. */
int wqc_tn_site_amplitudes(wqc_tn_state *st, size_t target, double out[4]);
/* MpsState::control_probability (src/tn/mps.rs:695-698). */
int wqc_tn_control_probability(wqc_tn_state *st, size_t control, double *out);
/* site_amplitudes of many targets of one state (src/tn/mps.rs:679-693), for the trace's
 * per-layer probes (first touch of each target and every control, src/tn/trace.rs:346-420,
 * 640-657): out[4k..4k+3] = (a0.re, a0.im, a1.re, a1.im) of targets[k]; the control
 * probability of a control c is |a1|^2 (mps.rs:695-698). One sweep of environments serves
 * every target. INVALID_ARG for a target outside the register. */
int wqc_tn_site_amplitudes_batch(wqc_tn_state *st, const int32_t *targets, size_t count, double *out);

/* MpsState::contract_to_statevector (src/tn/mps.rs:215-260): n <= 20, else
 * EXEC_FAILED "full statevector projection requires qubit_count <= 20".
 * `out` holds 2^n complex values (2^(n+1) doubles). */
int wqc_tn_contract_to_statevector(wqc_tn_state *st, double *out, size_t len_complex);

/* Replaces pauli_string_expectation over contract_to_statevector (src/expectation.rs:140-167,
 * src/tn/mps.rs:215-260, which refuse qubit_count > 20): <psi|P_k|psi> for n_terms Pauli
 * strings, contracted on the device with transfer matrices for any qubit_count. `labels` holds
 * the strings back to back, each exactly qubit_count characters from {I, X, Y, Z}, in Qiskit
 * order (the last character acts on qubit 0, src/expectation.rs:98-132). Un-normalized, as the
 * reference. out[k] is the real part (the reference drops the rounding-level imaginary part);
 * out_imag (optional, may be NULL) receives the imaginary part. INVALID_ARG on any other
 * character. */
int wqc_tn_pauli_expectations(wqc_tn_state *st, const char *labels, size_t n_terms, double *out,
                              double *out_imag);
/* Adjoint gradient of E(theta) = sum_k coeffs[k] <psi|P_k|psi> (real coefficients, Pauli
 * labels as wqc_tn_pauli_expectations) with respect to every RX/RY/RZ angle of the circuit; no
 * counterpart in the reference (its expectation mode has no gradients). The forward pass is the
 * plain FP64 form with exact splits (precision is validated, the range finder is not used: its
 * sketch would have to be differentiated too), truncating to max_bond_dim by the reference's
 * rule, one gate at a time, with the state saved at the start of every segment of
 * checkpoint_every runs of one-qubit gates and the gates after them (0: every run). The
 * gradient is reverse mode over the site tensors of that forward pass, truncations included:
 * the final site cotangents come from lambda = H psi (an MPO-MPS product), and each segment,
 * replayed from its checkpoint, is walked back op by op (a one-qubit update passes U^H back; a
 * unit passes the cotangents of its two sites back through the adjoint of its truncated split
 * with the sqrt(S) gauge, its gate and its merge). grad[i] (count entries) is dE/dtheta_i, 0 for
 * non-rotation gates; value = {Re E, Im E}. adjoint_bond is unused (no adjoint state is
 * carried). */
int wqc_tn_backward(wqc_tn_device *dev, size_t qubit_count, size_t max_bond_dim, int precision,
                    const wqc_tn_gate *gates, size_t count, const char *labels,
                    const double *coeffs, size_t n_terms, size_t checkpoint_every,
                    size_t adjoint_bond, double *grad, double value[2]);

/* A batch of independent slices, contract_slice (src/tn/contract.rs:10-30) per slice: new
 * register, the slice's gates, <0...0|psi> into out_amps[2i], out_amps[2i+1].
 * workers == 0: one batch. Every slice's gates are scheduled into layers of site-disjoint
 * operations, and step t runs layer t of all slices together: one one-qubit launch and shared
 * two-site batches (merge, Jacobi rounds, factors over units of many slices of different widths
 * and bond caps). Should a batch fail, the slices are rerun one by one for per-slice statuses.
 * workers > 0: that many host threads, each driving one state on its own stream, largest
 * estimated cost first (workers < 0: 8 threads).
 * out_status[i] is the slice's status code (its amplitude is NaN when it is not OK); the call
 * returns EXEC_FAILED naming the first failed slice, OK when all succeeded. */
int wqc_tn_run_slices(wqc_tn_device *dev, const wqc_tn_slice *slices, size_t n_slices,
                      const wqc_tn_gate *gates, int precision, int workers, double *out_amps,
                      int *out_status);

/* Terminal sampling, sample_terminal_measurements (src/sample.rs:192-211).
 *  method 0: the reference's semantics (n <= 20, EXEC_FAILED above it, as the reference's
 *            statevector projection): snapped Born table of the projected statevector keyed by
 *            the classical register, StdRng::seed_from_u64(seed) (ChaCha12), the first
 *            cumulative entry >= r per shot (src/sample.rs:85-189).
 *  method 1: MPS sampler for any n <= 64: shot k draws qubits 0..n-1 from the MPS's
 *            conditional Born probabilities using draw k*n+q of the same seeded stream, so
 *            counts are exact samples of |psi_MPS|^2 (no counterpart in the reference, which
 *            refuses n > 20).
 *  method 2: method 0 when n <= 20, else method 1.
 * measure k maps qubit measure_qubits[k] to cbit measure_cbits[k]; classical_bits in 1..64
 * (EXEC_FAILED "classical_bit_count must be > 0 for sample_counts" at 0, as the reference).
 * Writes up to `cap` (key, count) pairs in key order, key bit c = classical bit c (the
 * reference's Qiskit bitstring has cbit 0 rightmost); *n_out = pairs written; INVALID_ARG if
 * cap is too small (cap >= min(shots, 2^classical_bits) always suffices). */
int wqc_tn_sample_counts(wqc_tn_state *st, const int32_t *measure_qubits,
                         const int32_t *measure_cbits, size_t n_measures, size_t classical_bits,
                         uint64_t shots, uint64_t seed, int method, uint64_t *keys,
                         uint64_t *counts, size_t cap, size_t *n_out);

/* compute_outcome_probabilities (src/sample.rs:85-131) for qubit_count <= 20: the snapped
 * Born table of the projected statevector, one (key, probability) per outcome with p > 0 in
 * key order (bit c of a key = classical bit c); the table build_terminal_distribution_segment
 * (src/distribution_proof.rs:139-185) and the exact sampler draw from. Same argument checks
 * and errors as wqc_tn_sample_counts; INVALID_ARG when cap is too small. */
int wqc_tn_outcome_probabilities(wqc_tn_state *st, const int32_t *measure_qubits, const int32_t *measure_cbits,
                                 size_t n_measures, size_t classical_bits, uint64_t *keys, double *probs,
                                 size_t cap, size_t *n_out);
/* observable_expectation + compute_expectations (src/expectation.rs:170-200): sum_k coeff_k
 * <P_k> with complex coefficients (coeffs = n_terms (re, im) pairs), each component snapped to
 * the 1e-6 grid (snap_complex_result, src/engine.rs:30-40). out = {re, im}. */
int wqc_tn_observable_expectation(wqc_tn_state *st, const char *labels, const double *coeffs,
                                  size_t n_terms, double out[2]);

/* Bond dimensions: out[i] = right bond of site i, i < n (len >= n). */
int wqc_tn_bond_dims(wqc_tn_state *st, size_t *out, size_t len);

/* Site download/upload (the `pub sites` field, src/tn/mps.rs:56). download writes
 * l*2*r complex values when len_complex allows and always reports l and r. upload
 * replaces the site; its l must equal the right bond of site-1 (1 for site 0) and r
 * the left bond of site+1 (1 for the last site), except when the caller uploads a
 * whole chain site by site (then shapes are checked at the next gate). */
int wqc_tn_download_site(wqc_tn_state *st, size_t site, double *out, size_t len_complex,
                         size_t *l, size_t *r);
int wqc_tn_upload_site(wqc_tn_state *st, size_t site, const double *data, size_t l,
                       size_t r);
/* dst.sites = src.sites.clone() (the `pub sites` field, src/tn/mps.rs:56) without a host round
 * trip: every site of src copied device to device into dst. Both states must be on one device
 * and hold the same qubit count (INVALID_ARG otherwise); precision, bond cap and counters of dst
 * are kept. */
int wqc_tn_copy_sites(wqc_tn_state *dst, const wqc_tn_state *src);

/* ---- reporting ---------------------------------------------------------------------- */

/* MpsState::backend_label (src/tn/mps.rs:146-155): "hip". */
const char *wqc_tn_backend_label(const wqc_tn_state *st);
/* MpsState::peak_vram_bytes (src/tn/mps.rs:157-166): largest device footprint of this
 * state (sites plus split workspace), in bytes. */
uint64_t wqc_tn_peak_vram_bytes(const wqc_tn_state *st);
int wqc_tn_get_stats(const wqc_tn_state *st, wqc_tn_stats *out);

/* Device time (ms) accumulated per kernel family since the state was created, for
 * measurement: index 0 = one-qubit, 1 = merge, 2 = range-finder GEMMs, 3 = Jacobi,
 * 4 = factors, 5 = readouts, 6 = LQ preconditioning of the Jacobi, 7 = the mixed forms'
 * change of basis between their FP32 and FP64 sweeps (Newton-Schulz on J, X <- J X).
 * Returns the number of entries written. */
int wqc_tn_kernel_times(const wqc_tn_state *st, double *ms, size_t len);
int wqc_tn_enable_timing(wqc_tn_state *st, int enable);
/* Launch counts per kernel family (same indexing as wqc_tn_kernel_times). */
int wqc_tn_kernel_launches(const wqc_tn_state *st, uint64_t *count, size_t len);

/* Split log for measurement: when enabled, every two-qubit unit appends eight int32
 * values {rows, cols, bond, path, ell, chi_kept, fp64_sweeps, low_sweeps} (path 0 exact,
 * 1 range finder kept, 2 range finder fell back to exact; sweeps of the unit's Jacobi). wqc_tn_split_log copies up to cap entries and
 * returns the number recorded. */
int wqc_tn_enable_split_log(wqc_tn_state *st, int enable);
size_t wqc_tn_split_log(const wqc_tn_state *st, int32_t *out, size_t cap_entries);
/* The discarded weight sum_{k >= chi} s_k^2 / sum_k s_k^2 of each logged unit, in split-log order
 * (the range finder's B spectrum where that path kept its split): the per-bond diagnostic of the
 * truncation, no counterpart in the reference. Copies up to cap entries, returns the number
 * recorded. */
size_t wqc_tn_split_discarded(const wqc_tn_state *st, double *out, size_t cap_entries);
/* The probe residual of each logged unit's kept factorization, in split-log order: the larger of
 * |A x_p - U S V^H x_p| / (|A|_F |x_p|) over the two probe vectors of svd_reproduces
 * (src/tn/svd_guard.rs:261-280; A is the matrix the checked SVD took). The reference computes it
 * inside svd_reproduces and keeps only the verdict; this exposes the value. Copies up to cap
 * entries, returns the number recorded. */
size_t wqc_tn_split_probe_residuals(const wqc_tn_state *st, double *out, size_t cap_entries);

#ifdef __cplusplus
}
#endif

#endif /* WQC_TN_H */
