/* iqs_hip.h -- C ABI of the MI355X state-vector engine (libiqs_hip.so).
 *
 * This is the boundary between host bookkeeping (the iqs::QubitRegister drop-in, the
 * Python bindings) and the hand-written gfx950 kernels.  Plain pointers and sizes only.
 *
 * Conventions
 *   - A state is 2^n complex amplitudes in HBM, in DATA order: bit k of an index is the
 *     qubit stored at data position k (the reference's `position`, include/qureg.hpp:83-95).
 *     Program-qubit -> position maps stay on the host; every entry point takes positions.
 *   - Complex numbers cross the ABI as interleaved double pairs (re, im); matrices are
 *     row-major (TinyMatrix layout, include/tinymatrix.hpp:32-35).  For IQSH_C64 states the
 *     library narrows them to float once per call.
 *   - 4x4 / 2-qubit index convention: row = 2*b(first position) + b(second position)
 *     (src/qureg_apply2qubitgate.cpp:53-62, src/qureg_applydiag.cpp:165-171).
 *   - Every function returns IQSH_OK (0) or an error code, never aborts or throws;
 *     iqsh_last_error() describes the last failure of the calling thread.
 *   - Work is stream-ordered on the library stream (iqsh_set_stream).  Functions that
 *     return host values synchronize that stream; the others are asynchronous.
 *   - Reductions accumulate in FP64 for both precisions, in a fixed two-stage order
 *     (bitwise reproducible run to run).
 *   - One library stream per process.  Host threads may share the library: the shared
 *     workspaces (reduction partial sums, block-matrix ring, layer partial sums) are each
 *     used under a lock held until the call has finished with them, and all work is
 *     ordered on the one stream.  Switching the stream while another thread has work in
 *     flight is the caller's to order.
 */
#ifndef IQS_HIP_H
#define IQS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum iqsh_status {
  IQSH_OK = 0,
  IQSH_ERR_ARG = 1,    /* invalid argument (position, size, handle, dtype mismatch) */
  IQSH_ERR_ALLOC = 2,  /* device allocation failed */
  IQSH_ERR_HIP = 3,    /* HIP runtime error */
  IQSH_ERR_NOTSUP = 4  /* unsupported combination */
};

enum iqsh_dtype {
  IQSH_C64 = 0,  /* std::complex<float>  (ComplexSP, include/utils.hpp:24) */
  IQSH_C128 = 1  /* std::complex<double> (ComplexDP, include/utils.hpp:25) */
};

/* SpecializeV2 gate kinds, numbered as iqs::GateSpec1Q / GateSpec2Q (include/gate_spec.hpp:10-25);
 * IQSH_SPEC_CPHASE is GateSpec2Q::CPhase. */
enum iqsh_spec {
  IQSH_SPEC_H = 0, IQSH_SPEC_RX, IQSH_SPEC_RY, IQSH_SPEC_RZ,
  IQSH_SPEC_X, IQSH_SPEC_Y, IQSH_SPEC_Z, IQSH_SPEC_T_OR_CPHASE
};

typedef struct iqsh_state *iqsh_handle;

/* ---- errors, device, stream ------------------------------------------------------- */
const char *iqsh_last_error(void);
/* Use `stream` (a hipStream_t, NULL = library default stream) for all later work. */
int iqsh_set_stream(void *stream);
int iqsh_sync(void);
/* Compute units of the current device (0 if unknown). */
int iqsh_device_cus(int *cus);
/* Timing on the library stream: record a start event; then record a stop event, wait for
 * it and return the elapsed milliseconds (hipEventElapsedTime). */
int iqsh_timer_start(void);
int iqsh_timer_stop(double *ms);

/* ---- lifetime and data movement -------------------------------------------------- */
/* Allocate a register of n qubits (0 <= n <= 40; n = 0 holds one amplitude).  Replaces the reference's _mm_malloc allocation
 * (src/qureg_init.cpp:149-188).  Contents are undefined until initialized. */
int iqsh_create(unsigned num_qubits, int dtype, iqsh_handle *out);
/* Wrap existing device memory of 2^n amplitudes (not owned, not freed by destroy).
 * Counterpart of the importing constructor QubitRegister(n, Type*) (src/qureg_init.cpp:193-200). */
int iqsh_wrap(unsigned num_qubits, int dtype, void *device_ptr, iqsh_handle *out);
int iqsh_destroy(iqsh_handle h);
int iqsh_info(iqsh_handle h, unsigned *num_qubits, int *dtype, void **device_ptr);
/* Host <-> device copies of `count` amplitudes starting at data index `offset`
 * (operator[], GetGlobalAmplitude/SetGlobalAmplitude, src/qureg_utils.cpp:78-121). */
int iqsh_upload(iqsh_handle h, const void *host, uint64_t offset, uint64_t count);
int iqsh_download(iqsh_handle h, void *host, uint64_t offset, uint64_t count);
/* dst := src (copy constructor, src/qureg_init.cpp:364-379). */
int iqsh_copy(iqsh_handle dst, iqsh_handle src);
/* dst := src converted to dst's precision (same number of qubits); used to compare a
 * complex64 register against a complex128 one.  No counterpart in the reference. */
int iqsh_convert(iqsh_handle dst, iqsh_handle src);

/* ---- initialization (Initialize("base"|"++++"), src/qureg_init.cpp:225-359) ------- */
int iqsh_init_basis(iqsh_handle h, uint64_t data_index);
/* Every amplitude := (re, im) (InitializationWithSameAmplitudeEverywhere, qureg_utils.cpp:173-184). */
int iqsh_init_uniform(iqsh_handle h, double re, double im);

/* ---- gates -------------------------------------------------------------------------- */
/* 2x2 m on pairs (i, i + 2^pos), bit pos of i clear: Loop_DN (src/highperfkernels.cpp:287-380)
 * as called by Apply1QubitGate_helper (src/qureg_apply1qubitgate.cpp:173-230).  Matrix
 * structure is exploited exactly (identity: no pass; diag with m00 = 1: half the state). */
int iqsh_apply_1q(iqsh_handle h, unsigned pos, const double m[8]);
/* SpecializeV2 1q bodies (src/spec_kernels.cpp:36-102, Loop_DN spec :249-332). */
int iqsh_apply_1q_spec(iqsh_handle h, unsigned pos, int spec, double angle);
/* The 2x2 m on the pairs (k, k + c[8]) of the reference's Loop_TN loop nest
 * (src/highperfkernels.cpp:397-486): for (i = c[0]; i < c[1]; i += c[2])
 * for (j = i + c[3]; j < i + c[4]; j += c[5]) for (k = j + c[6]; k < j + c[7]; ++k).
 * Loop_DN(gstart, gend, pos) is {gstart, gend, 2^(pos+1), 0, 1, 1, 0, 2^pos, 2^pos}.  For the
 * sub-range gate helpers (Apply1QubitGate_helper / ApplyControlled1QubitGate_helper with
 * sind, eind: src/qureg_apply1qubitgate.cpp:173-230, src/qureg_applyctrl1qubitgate.cpp:
 * 228-406).  Every index touched must lie in the register (checked); the pairs must be
 * disjoint, as in the reference. */
int iqsh_apply_nest(iqsh_handle h, const uint64_t c[9], const double m[8]);
/* The 2x2 matrix the SpecializeV2 path applies for `spec` (controlled: GateSpec2Q meaning
 * of IQSH_SPEC_T_OR_CPHASE), with the reference's constants in `dtype`'s precision
 * (src/spec_kernels.cpp:172-178, 264-270, 356-361). */
int iqsh_spec_matrix(int dtype, int spec, int controlled, double angle, double m[8]);
/* Controlled 2x2 on target tpos where bit cpos = 1: Loop_TN (src/highperfkernels.cpp:397-486)
 * as called by ApplyControlled1QubitGate_helper (src/qureg_applyctrl1qubitgate.cpp:228-406). */
int iqsh_apply_c1q(iqsh_handle h, unsigned cpos, unsigned tpos, const double m[8]);
/* SpecializeV2 controlled bodies (src/spec_kernels.cpp:104-153, Loop_TN spec :334-417). */
int iqsh_apply_c1q_spec(iqsh_handle h, unsigned cpos, unsigned tpos, int spec, double angle);
/* 2x2 m on a target with several control positions (ctrl_mask bits all 1).  The reference
 * reaches this through ApplyToffoli's five controlled gates (src/qureg_applytoffoli.cpp:24-47). */
int iqsh_apply_mc1q(iqsh_handle h, uint64_t ctrl_mask, unsigned tpos, const double m[8]);
/* m on (|bit p1 = 1, bit p2 = 0>, |bit p1 = 0, bit p2 = 1>), p1 < p2 after ordering:
 * ApplySwap_helper (src/qureg_applyswap.cpp:114-234). */
int iqsh_apply_swaplike(iqsh_handle h, unsigned p1, unsigned p2, const double m[8]);
/* Diagonal d[2*b(p1) + b(p2)] (4 complex): ApplyDiag (src/qureg_applydiag.cpp:55-227). */
int iqsh_apply_diag2(iqsh_handle h, unsigned p1, unsigned p2, const double d[8]);
/* Dense 4x4 with row index 2*b(ph) + b(pl): Apply2QubitGate (src/qureg_apply2qubitgate.cpp:15-73). */
int iqsh_apply_2q(iqsh_handle h, unsigned ph, unsigned pl, const double m[32]);
/* Dense 2^k x 2^k block on positions pos[0..k-1] (row-major, row index bit j = bit pos[j]),
 * 1 <= k <= 5.  Replaces the reference's fused-gate replay (src/qureg_fusion.cpp:56-94):
 * a run of gates is multiplied into one block on the host and applied in one pass. */
int iqsh_apply_kq(iqsh_handle h, unsigned k, const unsigned *pos, const double *U);
/* Select the kernel for 5-qubit blocks (and 3-/4-qubit blocks padded to 5): 1 (default)
 * matrix cores, three real 32x32 products per block (Ur xr, Ui xi, (Ur+Ui)(xr+xi));
 * 2 matrix cores, one real 64x64 product; 0 vector ALUs.  Process-wide. */
int iqsh_set_block_mfma(int mode);
/* state *= f (AmplitudeWiseScalarMultiplication / Normalize, src/qureg_utils.cpp:162-201). */
int iqsh_scale(iqsh_handle h, double re, double im);
/* h += f * x (AmplitudeWiseSum, src/qureg_utils.cpp:208-231). */
int iqsh_axpy(iqsh_handle h, iqsh_handle x, double re, double im);
/* Zero the half with bit pos != value (CollapseQubit, src/qureg_measure.cpp:93-126). */
int iqsh_collapse(iqsh_handle h, unsigned pos, int value);

/* ---- permutation (PermuteLocalQubits, src/qureg_permute.cpp:57-105) ------------------ */
/* New state: bit b of each new data index is taken from bit src_pos[b] of the old index,
 * i.e. new[j] = old[sum_b bit_b(j) << src_pos[b]].  Bit-exact move.  Uses a second buffer
 * of the state's size (allocated on first use, kept until destroy). */
int iqsh_permute(iqsh_handle h, const unsigned *src_pos);

/* ---- reductions (FP64) --------------------------------------------------------------- */
/* sum |psi_i|^2 over bit pos = 1 (GetProbability, src/qureg_measure.cpp:136-178). */
int iqsh_prob1(iqsh_handle h, unsigned pos, double *out);
/* max |psi_i|^2 over the half with bit pos == value (the scans of IsClassicalBit /
 * GetClassicalValue, src/qureg_measure.cpp:19-81, 184-262). */
int iqsh_maxnorm_half(iqsh_handle h, unsigned pos, int value, double *out);
/* sum |psi_i|^2 (ComputeNorm squared, src/qureg_utils.cpp:236-256). */
int iqsh_norm2(iqsh_handle h, double *out);
/* <psi|P|psi> for the Pauli string P with X-or-Y positions xmask, Y-or-Z positions zmask
 * and ny Y factors, in one read-only pass; *norm2 receives sum |psi|^2 from the same pass.
 * Replaces ExpectationValue's basis change + parity reduction + undo
 * (src/qureg_expectval.cpp:102-213). */
int iqsh_expval_pauli(iqsh_handle h, uint64_t xmask, uint64_t zmask, unsigned ny,
                      double *value, double *norm2);
/* iqsh_expval_pauli without any allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (tight serving loops, code that forbids hipMalloc).
 * Same kernels and summation order as iqsh_expval_pauli (identical results); its partial
 * sums live in a fixed 64 KiB device array per GPU (at most 2048 partial sums per value,
 * which covers every register size: the reduction grid is capped there), whose address is
 * taken on the first call on each device.  Every call on a GPU uses that GPU's one array
 * and leaves its partial sums there; calls that overlapped on one GPU would race on it,
 * so calls take turns (a lock held until the result is back on the host). */
int iqsh_expval_pauli_noalloc(iqsh_handle h, uint64_t xmask, uint64_t zmask, unsigned ny,
                              double *value, double *norm2);
/* <x|h> = sum conj(x_i) h_i (ComputeOverlap, src/qureg_utils.cpp:267-300). */
int iqsh_overlap(iqsh_handle h, iqsh_handle x, double *re, double *im);
/* max |h_i - s x_i| (MaxAbsDiff, src/qureg_utils.cpp:40-68); s = (sre, sim). */
int iqsh_maxabsdiff(iqsh_handle h, iqsh_handle x, double sre, double sim, double *out);
/* sum |h_i - x_i|^2 (MaxL2NormDiff, src/qureg_utils.cpp:129-157). */
int iqsh_l2diff2(iqsh_handle h, iqsh_handle x, double *out);

/* ---- QAOA MaxCut (src/qaoa_features.cpp) -------------------------------------------- */
/* The cut value of a data index x is sum_e w_e [bit u_e(x) != bit v_e(x)], with u_e, v_e
 * DATA positions (the caller maps program vertices through its permutation).  The cut is
 * evaluated on the fly from edge-parity tables; no cost register is read. */
/* psi_i *= exp(-i gamma cut_i) (ImplementQaoaLayerBasedOnCostFunction, :248-264). */
int iqsh_qaoa_phase(iqsh_handle h, unsigned ne, const unsigned *u, const unsigned *v,
                    const double *w, double gamma);
/* sum cut_i |psi_i|^2 (GetExpectationValueFromCostFunction, :273-295). */
int iqsh_qaoa_expect(iqsh_handle h, unsigned ne, const unsigned *u, const unsigned *v,
                     const double *w, double *out);
/* Write cut_i into the real part of a cost register (InitializeVectorAs[Weighted]MaxCut-
 * CostFunction, :73-239); returns the maximum cut in *max_cut.  Any number of edges (one
 * pass per 64 edges); the on-the-fly functions above take at most 64. */
int iqsh_maxcut_diag(iqsh_handle diag, unsigned ne, const unsigned *u, const unsigned *v,
                     const double *w, double *max_cut);
/* Reference form with a cost register: psi_i *= (cos(g d_i), -sin(g d_i)), d_i = Re diag_i. */
int iqsh_qaoa_phase_diag(iqsh_handle psi, iqsh_handle diag, double gamma);
/* sum Re(diag_i) |psi_i|^2 (and with squared cost: GetExpectationValueSquaredFromCostFunction). */
int iqsh_qaoa_expect_diag(iqsh_handle psi, iqsh_handle diag, int squared, double *out);

/* ---- adjoint differentiation (no counterpart in the reference) ---------------------- */
/* lam_i := cut_i psi_i and *cost := sum cut_i |psi_i|^2 in one pass (lam = H_C psi). */
int iqsh_qaoa_cost_apply(iqsh_handle psi, iqsh_handle lam, unsigned ne, const unsigned *u,
                         const unsigned *v, const double *w, double *cost);
/* Backward step through psi <- exp(-i gamma C) psi for both psi and lam:
 * *grad := 2 Im <lam| C |psi>, then psi, lam *= exp(+i gamma cut). */
int iqsh_qaoa_phase_adjoint(iqsh_handle psi, iqsh_handle lam, unsigned ne, const unsigned *u,
                            const unsigned *v, const double *w, double gamma, double *grad);
/* Backward step through a 1q gate U on pos for both states: *grad := 2 Im <lam| G |psi>
 * (G a 2x2 generator, e.g. X/2 for RX), then psi := Uinv psi, lam := Uinv lam.
 * G == NULL skips the reduction (non-parameterized gate). */
int iqsh_adjoint_1q(iqsh_handle psi, iqsh_handle lam, unsigned pos, const double Uinv[8],
                    const double *G, double *grad);

/* ---- layers of one gate on every qubit -------------------------------------------------
 * psi := m (x) m (x) ... (x) m psi (the same 2x2 matrix on every position) in
 * 1 + ceil((n - 8) / 5) passes instead of n (a matrix with a real diagonal and an imaginary
 * off-diagonal, such as RX, takes shorter kernels).  Replaces the per-qubit loop that applies a
 * QAOA mixer, RX(2 beta) on every qubit, after ImplementQaoaLayerBasedOnCostFunction
 * (src/qaoa_features.cpp:73-110 users; src/qureg_apply1qubitgate.cpp:277-285).  n >= 10. */
int iqsh_apply_1q_all(iqsh_handle h, const double m[8]);
/* Backward step through such a layer for both states: *grad := sum_q 2 Im <lam| G_q |psi>
 * (G on position q, each term as iqsh_adjoint_1q would take it), then psi := Uinv^(x)n psi,
 * lam := Uinv^(x)n lam, in 1 + ceil((n - 8) / 4) passes (G equal to X exactly, and Uinv of
 * the RX form, take shorter kernels).  n >= 10. */
int iqsh_adjoint_1q_all(iqsh_handle psi, iqsh_handle lam, const double Uinv[8], const double *G, double *grad);

/* One QAOA layer, psi := mixer^(x)n exp(-i gamma C) psi with C the cut of the edge list (as
 * iqsh_qaoa_phase), the phase fused into the mixer's first pass: the reference's
 * ImplementQaoaLayerBasedOnCostFunction followed by the per-qubit mixer loop
 * (src/qaoa_features.cpp:248-264, unit_test/include/qaoa_features_test.hpp:75-76).  n >= 10. */
int iqsh_qaoa_layer(iqsh_handle h, unsigned ne, const unsigned *u, const unsigned *v, const double *w, double gamma,
                    const double mixer[8]);
/* Backward step through one QAOA layer for both states: *dmixer := sum_q 2 Im <lam| G_q |psi>
 * (as iqsh_adjoint_1q_all), the mixer undone with mixer_inv, then *dgamma := 2 Im <lam| C |psi>
 * and the phase undone (as iqsh_qaoa_phase_adjoint) -- fused into the same passes.  n >= 10. */
int iqsh_qaoa_layer_adjoint(iqsh_handle psi, iqsh_handle lam, unsigned ne, const unsigned *u, const unsigned *v,
                            const double *w, double gamma, const double mixer_inv[8], const double *G,
                            double *dgamma, double *dmixer);

/* ---- batches of hardware-efficient-ansatz circuits ------------------------------------
 * No counterpart in the reference; the nearest is its pool of independent states
 * (src/mpi_env.cpp:239-404, one state per MPI rank group).  A batch holds `count`
 * independent registers, instance b with n_b qubits (1 <= n_b <= 26) and d_b >= 1 layers.
 * Every instance starts from |0...0> and layer l applies, on every qubit q,
 * RotationY(theta[l][q]) then RotationZ(phi[l][q]) (src/qureg_apply1qubitgate.cpp:297-325:
 * RY = [[c,-s],[s,c]], RZ = diag(e^{-i phi/2}, e^{+i phi/2}), c,s of theta/2), followed by
 * CPauliZ(i, i+1) for i = 0..n_b-2.  The cost is <sum_q Z_q> in the final state.
 * Parameters and gradients share one layout: instance after instance, then layer, qubit,
 * and the pair (theta, phi) -- 2 * n_b * d_b doubles per instance.
 * A run executes forward passes and, with IQSH_HEA_GRAD, the adjoint gradient
 * d cost / d angle of every angle, for all instances in one persistent kernel launch
 * (plus a fill, a matrix set-up and a reduction kernel), captured in a HIP graph on first
 * use and replayed afterwards.  Amplitudes are complex64 or complex128 (dtype); costs,
 * gradient sums and their reductions are FP64 in a fixed order (bitwise reproducible for
 * a given tuning). */
typedef struct iqsh_hea *iqsh_hea_handle;

enum iqsh_hea_flags {
  IQSH_HEA_GRAD = 1,       /* adjoint gradient in addition to the cost */
  IQSH_HEA_KEEP_STATE = 2, /* keep every final state (forward-only runs) for iqsh_hea_state */
  IQSH_HEA_NO_GRAPH = 4    /* launch the kernels directly instead of replaying the graph */
};

/* Tuning of a batch: `variant` picks the kernel's tile (tile qubits, amplitudes per thread,
 * see iqsh_hea_variant_info); `workers_per_cu` persistent workgroups per compute unit
 * (0 = as many as fit); `order` the task order: 0 = tiled instances first, single-tile
 * instances last; 1 = single-tile first; 2 = single-tile instances interleaved among the
 * steps of the tiled ones; 3 = as 0 with every tiled instance's steps aligned to end on
 * the last round.  `group_mib` > 0 processes the tiled instances in consecutive groups
 * whose states (psi and lam) total at most that many MiB, so that a group's passes can be
 * served from the last-level cache; single-tile instances are then spread over the groups
 * and `order` is not used. */
typedef struct {
  int variant;
  int workers_per_cu;
  int order;
  int group_mib;
} iqsh_hea_tune;

typedef struct {
  iqsh_hea_tune tune;
  double ms; /* mean milliseconds per run */
} iqsh_hea_trial;

/* Number of kernel variants for a dtype, and the shape of one: instances with
 * n <= tile_qubits run whole in one workgroup; larger ones (n <= 2 * tile_qubits) run as
 * tile passes.  `for_tests` variants have small tiles and exist to exercise the tiled path
 * on small registers; the autotuner skips them. */
int iqsh_hea_variants(int dtype, int *count);
int iqsh_hea_variant_info(int dtype, int variant, int *tile_qubits, int *amps_per_thread,
                          int *threads, int *for_tests);
/* Create a batch (allocates every state and work buffer); tune NULL = default variant 0,
 * as many workers as fit, order 0. */
int iqsh_hea_create(int dtype, int count, const int *nqubits, const int *depth,
                    const iqsh_hea_tune *tune, iqsh_hea_handle *out);
int iqsh_hea_destroy(iqsh_hea_handle h);
int iqsh_hea_num_params(iqsh_hea_handle h, uint64_t *nparams);
int iqsh_hea_set_tune(iqsh_hea_handle h, const iqsh_hea_tune *tune);
int iqsh_hea_get_tune(iqsh_hea_handle h, iqsh_hea_tune *tune);
/* Angles from host memory or from device memory (both nparams doubles). */
int iqsh_hea_set_angles(iqsh_hea_handle h, const double *host_angles);
int iqsh_hea_set_angles_device(iqsh_hea_handle h, const double *device_angles);
/* Run the batch (asynchronous on the library stream). */
int iqsh_hea_run(iqsh_hea_handle h, int flags);
/* Costs (count doubles) and, after a gradient run, gradients (nparams doubles); NULL
 * skips one.  The host form synchronizes; the device form copies on the stream. */
int iqsh_hea_results(iqsh_hea_handle h, double *cost, double *grad);
int iqsh_hea_results_device(iqsh_hea_handle h, double *cost_dev, double *grad_dev);
/* Both forms first read the run's control words (the device form therefore synchronizes
 * too) and return IQSH_ERR_HIP, copying nothing, when a dependency wait gave up or a task
 * queue was not drained.  The batch's tile tasks wait for the tiles of the previous step
 * with bounded polling: at most max_polls polls (default 2^20, about one second) before
 * the wait gives up, the run aborts and every workgroup exits.  unmet != 0 makes every
 * such wait unsatisfiable -- a fault injection for testing that path from the host.  The
 * next run recaptures its graph. */
int iqsh_hea_set_wait_limit(iqsh_hea_handle h, uint32_t max_polls, int unmet);
/* Final state of instance `inst` (2^n amplitudes of the batch dtype) after a run with
 * IQSH_HEA_KEEP_STATE. */
int iqsh_hea_state(iqsh_hea_handle h, int inst, void *host);
/* Algorithmic HBM bytes of one run (amplitude reads and writes of the tiled instances;
 * single-tile instances stay on chip). */
int iqsh_hea_traffic(iqsh_hea_handle h, int flags, double *bytes);
/* Try every tuning (non-test variants that fit the batch x workers {1, 2} x (orders
 * {0, 1, 2, 3} ungrouped, and groups of 64 and 160 MiB)), `reps` timed runs each after
 * one warm-up, keep the fastest.  Trials are
 * written to `trials` (up to max_trials; *ntrials = how many were run). */
int iqsh_hea_autotune(iqsh_hea_handle h, int flags, int reps, iqsh_hea_trial *trials,
                      int max_trials, int *ntrials, double *best_ms);

/* ---- statistics ---------------------------------------------------------------------- */
/* With p_i = |psi_i|^2, in FP64: out[0] = -sum p ln p, out[1] = -sum_{p > 0} ln p,
 * out[k - 1] = sum p^k for k = 2..10.  Replaces the loops of QubitRegister::Entropy and
 * ::GoogleStats (src/qureg_utils.cpp:306-450), which scale these sums on the host. */
int iqsh_prob_moments(iqsh_handle h, double out[11]);
/* hist[b] = sum of |psi_i|^2 over i whose cost value c_i = Re diag_i falls in bin b:
 * mode 0 b = (int)c (GetHistogramFromCostFunction, src/qaoa_features.cpp:335-393),
 * mode 1 b = floor(c + 1e-7) (...WithWeightsRounded, :402-462), mode 2
 * b = floor(c / bin_width + 1e-7) (...WithWeightsBinned, :471-529).  A cost value outside
 * [0, max_value] (mode 0) or [-1e-7, max_value + 1e-7] (modes 1, 2), or a bin outside
 * [0, nbins), is an error (the reference asserts).  nbins <= 4096. */
int iqsh_cost_histogram(iqsh_handle psi, iqsh_handle diag, int mode, double max_value, double bin_width,
                        double *hist, int nbins);

/* ---- pool of states ------------------------------------------------------------------
 * A pool is an ordinary register of state_qubits + k positions: the top k positions index
 * up to 2^k states of state_qubits qubits, state s holding amplitudes
 * [s * 2^state_qubits, (s + 1) * 2^state_qubits).  Every gate entry point above, on
 * positions < state_qubits, applies the same gate to every state in one launch.  The
 * entry points below act state by state; states >= nstates (padding up to 2^k) are left
 * untouched by them and reported by none.  Counterpart: the reference's pool of states,
 * one state per MPI rank group (src/mpi_env.cpp:239-404), and
 * Environment::IncoherentSumOverAllStatesOfPool (src/mpi_env.cpp:468-486), which sums one
 * value per state -- the per-state values here. */
enum iqsh_pool_reduce_kind {
  IQSH_POOL_NORM2 = 0,   /* sum_i |psi_i|^2 */
  IQSH_POOL_PROB1 = 1,   /* sum over i with bit `arg` set (GetProbability, unnormalized) */
  IQSH_POOL_ZSTRING = 2  /* sum_i (-1)^popcount(i & arg) |psi_i|^2 (<Z...Z> on the mask) */
};
/* Every state s < nstates set to |index> (style 0) or to the uniform superposition
 * (style 1, amplitude 2^(-state_qubits/2)); padding states zero.  Replaces
 * Initialize("base"/"++++") run in every rank group (src/qureg_init.cpp:225-300). */
int iqsh_pool_init(iqsh_handle h, unsigned state_qubits, int style, uint64_t index, uint64_t nstates);
/* A different 2x2 matrix per state on position pos < state_qubits: mats holds nstates
 * row-major complex 2x2 matrices (8 doubles each).  One launch for the whole pool; what the
 * reference does with one Apply1QubitGate per rank group (noise gates and channel branches
 * differ from state to state). */
int iqsh_pool_apply_1q_each(iqsh_handle h, unsigned state_qubits, unsigned pos, const double *mats,
                            uint64_t nstates);
/* A different dense 4x4 matrix per state on positions (ph, pl) < state_qubits, row index
 * 2*b(ph) + b(pl) as iqsh_apply_2q: mats holds nstates row-major complex 4x4 matrices
 * (32 doubles each).  What the reference does with one Apply2QubitGate per rank group for
 * the branches of a 2-qubit channel (src/qureg_apply_channel.cpp:66-106). */
int iqsh_pool_apply_2q_each(iqsh_handle h, unsigned state_qubits, unsigned ph, unsigned pl, const double *mats,
                            uint64_t nstates);
/* out[s] = the reduction `kind` of state s (FP64, fixed order), for s < nstates.  Replaces
 * per-rank-group GetProbability / ComputeNorm / ExpectationValue of Z strings
 * (src/qureg_measure.cpp:136-178, src/qureg_utils.cpp:237-262, src/qureg_expectval.cpp:104-375). */
int iqsh_pool_reduce(iqsh_handle h, unsigned state_qubits, int kind, uint64_t arg, double *out, uint64_t nstates);
/* out[2s], out[2s+1] = Re, Im of <ref|state s> = sum_i conj(ref_i) psi_{s,i}; ref is a
 * register of state_qubits qubits and the pool's precision (ComputeOverlap per rank group,
 * src/qureg_utils.cpp:268-299). */
int iqsh_pool_overlap(iqsh_handle h, unsigned state_qubits, iqsh_handle ref, double *out, uint64_t nstates);

#ifdef __cplusplus
}
#endif

#endif /* IQS_HIP_H */
