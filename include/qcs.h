#ifndef QCS_H
#define QCS_H

/* ======================================================================== */
/* QCS -- state-vector quantum circuit simulator, MI355X (gfx950) backend    */
/* ======================================================================== */
/*
 * Drop-in for the reference's C89 API (SegusFaultise/qcs include/qcs.h):
 * the declarations in the first part of this header keep the reference's
 * names, argument meanings and error behaviour (messages on stderr, no
 * return codes).  The state vector lives in HBM for the life of a circuit;
 * gates run as hand-written HIP kernels through the C ABI in qcs_hip.h.
 *
 * Build mode: the reference selected a backend with QCS_MULTI_THREAD /
 * QCS_CPU_OPENMP / QCS_GPU_OPENCL / QCS_SIMD_ONLY before including this
 * header.  This library is the QCS_GPU_HIP mode; there is no host fallback.
 *
 * Semantics: controlled gates, QFT and Grover follow the textbook
 * definitions.  qc_create_ex(n, QCS_REFERENCE_COMPAT) reproduces the
 * reference's sequential maps as written instead (controlled-U writes only
 * the control=1,target=0 half; Grover's oracle/diffusion swap buffers).
 */

/* ======================================================================== */
/* QUANTUM CIRCUIT BUILDER */
/* ======================================================================== */
typedef struct t_q_circuit t_q_circuit;

/* Circuit Creation */
t_q_circuit *qc_create(int num_qubits);
void qc_destroy(t_q_circuit *circuit);

/* Basic Gates */
void qc_h(t_q_circuit *circuit, int qubit);
void qc_x(t_q_circuit *circuit, int qubit);
void qc_y(t_q_circuit *circuit, int qubit);
void qc_z(t_q_circuit *circuit, int qubit);
void qc_cnot(t_q_circuit *circuit, int control, int target);

/* Advanced Gates */
void qc_phase(t_q_circuit *circuit, int qubit, double angle);
void qc_rx(t_q_circuit *circuit, int qubit, double angle);
void qc_ry(t_q_circuit *circuit, int qubit, double angle);
void qc_rz(t_q_circuit *circuit, int qubit, double angle);

/* Circuit Operations */
void qc_barrier(t_q_circuit *circuit);
void qc_reset(t_q_circuit *circuit, int qubit);

/* Measurement */
int qc_measure(t_q_circuit *circuit, int qubit);
void qc_measure_all(t_q_circuit *circuit, int *results);

/* Circuit Execution */
void qc_run(t_q_circuit *circuit);
void qc_run_shots(t_q_circuit *circuit, int shots, int *results);

/* State Access */
int qc_find_most_likely_state(t_q_circuit *circuit);
double qc_get_probability(t_q_circuit *circuit, int state);
void qc_print_state(t_q_circuit *circuit, int solution_index);
void qc_print_circuit(t_q_circuit *circuit);

/* Built-in Algorithms */
void qc_grover_search(t_q_circuit *circuit, int solution_state);
void qc_quantum_fourier_transform(t_q_circuit *circuit);
void qc_bernstein_vazirani(t_q_circuit *circuit, int hidden_string);
void qc_ghz_state(t_q_circuit *circuit);

/* Utility Functions */
int qc_get_num_qubits(t_q_circuit *circuit);
int qc_get_num_gates(t_q_circuit *circuit);
void qc_optimize(t_q_circuit *circuit);

/* ======================================================================== */
/* ADDITIONS (no counterpart declared in the reference header)               */
/* ======================================================================== */

/* Controlled phase diag(1, e^{i angle}) on target where control = 1.  The
 * reference defines it (src/qcs.c:431) but does not declare it. */
void qc_cphase(t_q_circuit *circuit, int control, int target, double angle);
/* SWAP of two qubits (QFT with final swaps). */
void qc_swap(t_q_circuit *circuit, int qubit_a, int qubit_b);

/* Circuit options for qc_create_ex (bit flags). */
#define QCS_PLAIN_FP64 0x0      /* complex128, gate by gate (default) */
#define QCS_FUSED_FP32 0x1      /* complex64 state; gates are recorded and,
                                   at the next observation, fused into
                                   dense k-qubit unitaries (composed in FP64)
                                   applied on the FP32 matrix cores; FP64
                                   renormalisation after each flush */
#define QCS_PLAIN_FP32 0x2      /* complex64, gate by gate */
#define QCS_REFERENCE_COMPAT 0x4 /* the reference's as-written maps (plain
                                    forms only) */
#define QCS_RENORM_AT_BARRIER 0x8 /* fused form: flush + FP64 renormalise at
                                     every qc_barrier */
#define QCS_FUSE_TILES 0x10     /* fused form: tile programs -- several
                                   fused blocks per HBM pass (num_qubits >=
                                   16); default: one HBM pass per block */
#define QCS_FUSE_SPLIT_BF16 0x40 /* fused form: tile programs (as
                                    QCS_FUSE_TILES) whose blocks run on the
                                    bf16 matrix cores as split products --
                                    the FP32 matrix in three bf16 parts, the
                                    state in two, FP32 accumulation -- so the
                                    state is rounded to 16 significant bits
                                    per block instead of 24 */
#define QCS_FUSE_K(k) ((k) << 8) /* fused form: max qubits per fused block,
                                    1..5 (default 5; at most 4 with
                                    QCS_FUSE_TILES) */
#define QCS_DEFERRED 0x20       /* no device state: gate calls are only
                                   recorded and the circuit runs in a batch
                                   (qc_batch_*), complex128 gate by gate or,
                                   with QCS_FUSED_FP32, fused complex64;
                                   1..22 qubits */
t_q_circuit *qc_create_ex(int num_qubits, int options);
int qc_get_options(t_q_circuit *circuit);

/* Apply every recorded gate (fused form) and wait for the device. */
void qc_flush(t_q_circuit *circuit);
/* Sticky error code of the circuit's device state (0 = none). */
int qc_last_error(t_q_circuit *circuit);

/* 64-bit index variants (states of 31..34 qubits). */
double qc_get_probability_l(t_q_circuit *circuit, long state);
long qc_find_most_likely_state_l(t_q_circuit *circuit);
void qc_grover_search_l(t_q_circuit *circuit, long solution_state);
/* Grover with an explicit iteration count (e.g. a truncated run). */
void qc_grover_search_iters(t_q_circuit *circuit, long solution_state,
                            long iterations);
/* Sampling without the 2^n histogram: indices[s] = sampled basis index of
 * shot s, or -1 for a dropped shot (same uniforms and rule as
 * qc_run_shots). */
void qc_sample(t_q_circuit *circuit, int shots, long *indices);

/* Reset the state to the basis state |index> (recorded gates are dropped;
 * the history is kept).  On a QCS_DEFERRED circuit (index 0 only) it drops
 * the recorded gates, so the circuit can be recorded again. */
void qc_set_basis_state(t_q_circuit *circuit, long index);

/* Amplitude access: 2*count doubles {re, im} from index `offset`. */
int qc_get_amplitudes(t_q_circuit *circuit, long offset, long count,
                      double *out);
int qc_set_amplitudes(t_q_circuit *circuit, long offset, long count,
                      const double *in);
/* Squared norm of the state (FP64 device reduction). */
double qc_norm_sq(t_q_circuit *circuit);
/* 1 - |<a|b>|^2/(|a|^2 |b|^2) of two circuits' states (FP64 on the device);
 * *max_abs (if not NULL) receives the max-abs amplitude difference. */
double qc_compare(t_q_circuit *a, t_q_circuit *b, double *max_abs);
/* Number of fused blocks / HBM passes applied so far (fused form). */
long qc_get_num_fused_blocks(t_q_circuit *circuit);
long qc_get_num_fused_passes(t_q_circuit *circuit);
/* The circuit's device state (a qcs_hip_state *, see qcs_hip.h), after
 * applying every recorded gate; NULL when the circuit has no state. */
void *qc_get_device_state(t_q_circuit *circuit);

/* Batched execution of QCS_DEFERRED circuits: every circuit starts from
 * |0...0>, runs its recorded gates and is sampled, all in one persistent
 * device launch (small states stay in on-chip memory).  A complex128 circuit
 * gives the amplitudes of qc_create + the same gates (equal values; only
 * the sign of a zero may differ) and the samples of qc_sample; a
 * QCS_FUSED_FP32 one those of the fused form
 * within its tolerance.  Not thread-safe, like the rest of the library. */
typedef struct s_q_batch t_q_batch;
#define QCS_BATCH_GRAPH 0x1       /* capture the launch sequence as a graph
                                     and replay it while the batch's shape
                                     (counts and sizes) is unchanged */
#define QCS_BATCH_KEEP_STATES 0x2 /* keep the final states for
                                     qc_batch_get_amplitudes */
#define QCS_BATCH_REPLAN 0x4      /* plan every run from scratch; by default
                                     a run of the same circuits (unchanged
                                     since, same shot count) as the last run
                                     reuses that run's plan and device
                                     descriptors and only draws new
                                     uniforms */
t_q_batch *qc_batch_create(int options);
void qc_batch_destroy(t_q_batch *batch);
/* indices[i * shots + s]: shot s of circuits[i] (-1: dropped).  Uniforms
 * come from rand() circuit after circuit, shot after shot -- the stream
 * qc_sample on each circuit in turn would use.  Returns 0 or a negative
 * QCS_HIP_E* code (see qcs_hip.h). */
int qc_batch_sample(t_q_batch *batch, t_q_circuit **circuits, int count,
                    int shots, long *indices);
/* qc_run_shots for every circuit: results[i] holds 2^num_qubits ints of
 * circuits[i], zeroed here. */
int qc_batch_run_shots(t_q_batch *batch, t_q_circuit **circuits, int count,
                       int shots, int **results);
/* 2*count doubles {re, im} of circuit i's final state in the last run
 * (QCS_BATCH_KEEP_STATES). */
int qc_batch_get_amplitudes(t_q_batch *batch, int circuit, long offset,
                            long count, double *out);
/* Device time of the last run in ms: the batch kernel (-1 when the run
 * replayed a graph) and the whole launch sequence.  Returns 1 when the last
 * run replayed a captured graph, else 0. */
int qc_batch_last_timing(t_q_batch *batch, double *kernel_ms,
                         double *total_ms);
/* Plan of the circuits with `num_qubits` qubits: LDS tile bits (complex128
 * <= 13, complex64 <= 14; a state of at most that many qubits stays on
 * chip) and, for fused circuits, the fusion width 1..5.  0 keeps the
 * default (the largest tile; the circuit's QCS_FUSE_K or 5). */
int qc_batch_set_plan(t_q_batch *batch, int fused, int num_qubits,
                      int tile_bits, int fuse_k);
/* threads per work-group of the batch kernel: 256 or 512 (default). */
int qc_batch_set_threads(t_q_batch *batch, int threads);

/* Autotuning, cached per device in a text file ($QCS_TUNE_CACHE, a path or
 * "off"; else $XDG_CACHE_HOME/qcs_amd/tune.txt or
 * ~/.cache/qcs_amd/tune.txt).
 * qc_autotune: fused circuit, one pass per block -- times the fused-pass
 * kernel variants of every width k on the circuit's state (identity blocks,
 * the state is left as it is), keeps the fastest per (n, k) and picks the
 * fusion width for the recorded gates (blocks x time per pass).  Fused
 * circuit as tile programs (QCS_FUSE_TILES / QCS_FUSE_SPLIT_BF16, n >= 16)
 * -- times the recorded gates' own pass plan with identity blocks at fusion
 * widths 3 and 4, 512 or 1024 threads per work-group, one or two
 * work-groups per CU and both forms of k = 4 blocks (QCS_HIP_TUNE_TILE_K4_FORM),
 * and keeps the one with the least total pass time.
 * Returns the fusion width or a negative QCS_HIP_E* code. */
int qc_autotune(t_q_circuit *circuit);
/* qc_batch_autotune: for every (precision, n) of the given deferred
 * circuits whose state is streamed through on-chip tiles, times candidate
 * tile sizes (and fusion widths) on a sample of them with zero shots (no
 * rand() draws), then the work-group size; sets the batch's plan.  Returns
 * 0 or a negative QCS_HIP_E* code. */
int qc_batch_autotune(t_q_batch *batch, t_q_circuit **circuits, int count);

#endif
