/*
 * qcs_hip.h -- C ABI of the MI355X (gfx950) state-vector backend.
 *
 * Plain C89-compatible declarations: opaque handle, ints, longs, doubles and
 * pointers only.  Every function returns 0 on success or a negative
 * QCS_HIP_E* code; qcs_hip_last_error() gives the message of the most recent
 * failure on the calling thread.  Nothing here falls back to host loops: when
 * the device is missing the call fails with QCS_HIP_ENODEV.
 *
 * Each entry point replaces one host-side seam of the reference library
 * (SegusFaultise/qcs); the reference file:line it replaces is cited on it.
 *
 * Indices are `long` (64-bit on LP64), so states up to 2^34 amplitudes work.
 * Amplitudes cross the boundary as interleaved {re, im} doubles, the layout of
 * the reference's `struct t_complex` (src/internal.h:6-9).
 * A 2x2 gate crosses as 8 doubles, row-major {m00re, m00im, m01re, m01im,
 * m10re, m10im, m11re, m11im}: the `data[0..3]` order of the reference's
 * `struct t_q_matrix` (src/internal.h:55-59).
 */
#ifndef QCS_HIP_H
#define QCS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qcs_hip_state qcs_hip_state;

/* precision of the device state */
#define QCS_HIP_C128 0 /* complex128: {double re, im} per amplitude */
#define QCS_HIP_C64 1  /* complex64:  {float re, im} per amplitude */

/* state flags */
#define QCS_HIP_COMPAT 1 /* keep the reference's double buffer (vector +
                            scratch_vector, src/q_state.c:49-66) and apply the
                            as-written controlled-gate / Grover maps */

/* error codes */
#define QCS_HIP_OK 0
#define QCS_HIP_EINVAL -1 /* bad argument (qubit, size, pointer) */
#define QCS_HIP_ENODEV -2 /* no HIP device / runtime failure at init */
#define QCS_HIP_ENOMEM -3 /* device allocation failed */
#define QCS_HIP_EDEVICE -4 /* a HIP call or kernel launch failed */
#define QCS_HIP_EUNSUP -5 /* combination not supported (e.g. fusion in
                             compat mode) */

/* ---- device / error -------------------------------------------------- */
const char *qcs_hip_last_error(void);
/* 1 if a HIP device is usable, 0 otherwise (never aborts) */
int qcs_hip_device_available(void);
/* hipDeviceSynchronize on the current device */
int qcs_hip_device_sync(void);
/* "<gcn arch>:<compute units>" of the current device (e.g. "gfx950:256"),
 * the key of cached autotuning results; no reference counterpart */
int qcs_hip_device_name(char *buf, int len);

/* ---- state lifetime ---------------------------------------------------
 * Replaces q_state_init / q_state_free (src/q_state.c:32-115): allocates
 * 2^num_qubits amplitudes in HBM (a second buffer only with QCS_HIP_COMPAT),
 * sets |0...0>, binds one HIP stream to the state.  Unlike the reference's
 * OpenCL backend (process-global buffers, src/gpu_complex.c:6-20) every state
 * owns its device memory. */
int qcs_hip_state_create(qcs_hip_state **out, int num_qubits, int precision,
                         int flags);
void qcs_hip_state_destroy(qcs_hip_state *s);
int qcs_hip_state_num_qubits(const qcs_hip_state *s);
int qcs_hip_state_precision(const qcs_hip_state *s);
int qcs_hip_state_flags(const qcs_hip_state *s);
/* |basis>; the compat scratch buffer is zeroed as q_state_init does
 * (src/q_state.c:93-99) */
int qcs_hip_state_set_basis(qcs_hip_state *s, long basis);
/* wait for every op queued on the state's stream */
int qcs_hip_sync(qcs_hip_state *s);

/* ---- host <-> device copies (observation, tests) -----------------------
 * `out` / `in` hold 2*count doubles {re, im}; c64 states convert. */
int qcs_hip_read(qcs_hip_state *s, long offset, long count, double *out);
int qcs_hip_write(qcs_hip_state *s, long offset, long count, const double *in);
/* compat mode only: the scratch buffer (reference `scratch_vector`) */
int qcs_hip_read_scratch(qcs_hip_state *s, long offset, long count,
                         double *out);

/* ---- gates -------------------------------------------------------------
 * qcs_hip_apply_1q replaces q_apply_1q_gate / q_apply_1q_gate_gpu
 * (src/q_gates.c:25-149, src/gpu_complex.c:521-695): one in-place pair sweep
 * (out-of-place plus buffer swap in compat mode).  Arithmetic per output is
 * c_add(c_mul(g0, v0), c_mul(g1, v1)) without FMA contraction
 * (src/complex.c:23-57), so c128 results equal the reference's sequential
 * loop bit for bit. */
int qcs_hip_apply_1q(qcs_hip_state *s, int target, const double m[8]);
/* qcs_hip_apply_cu replaces q_apply_2q_gate / q_apply_2q_gate_gpu
 * (src/q_gates.c:158-298, src/gpu_complex.c:704-921): controlled 2x2 on
 * `target` where bit `control` is set.  Textbook mode touches only the
 * control=1 half (CNOT: swap, CP: the c=1,t=1 quarter).  Compat mode
 * reproduces the as-written sequential map out[i0] = m00*v0 + m01*v1,
 * out[i1] = v1 (src/q_gates.c:279-292). */
int qcs_hip_apply_cu(qcs_hip_state *s, int control, int target,
                     const double m[8]);
/* SWAP of two qubits (no reference counterpart; used for QFT final swaps).
 * Textbook mode only. */
int qcs_hip_apply_swap(qcs_hip_state *s, int qubit_a, int qubit_b);

/* ---- Grover ------------------------------------------------------------
 * Replace q_apply_phase_flip (src/q_gates.c:305-317) and q_apply_diffusion
 * (src/q_gates.c:323-356).  Textbook: a[m] = -a[m]; a = 2*(sum a)/N - a.
 * Compat: the as-written buffer swaps and unit-magnitude mean. */
int qcs_hip_phase_flip(qcs_hip_state *s, long index);
int qcs_hip_diffusion(qcs_hip_state *s);
/* `iterations` rounds of (phase_flip(marked); diffusion) on the device with
 * no host round trip (textbook: one fused read+write pass per round).  The
 * loop body of qc_grover_search (src/qcs.c:412-418). */
int qcs_hip_grover(qcs_hip_state *s, long marked, long iterations);

/* ---- reductions and observation (FP64 on the device) -------------------
 * norm_sq = sum |a|^2 (q_state_normalize's sum, src/q_utils.c:102-110) */
int qcs_hip_norm_sq(qcs_hip_state *s, double *norm_sq);
/* q_state_normalize (src/q_utils.c:102-118): if norm_sq > 1e-12 and
 * norm_sq != 1.0 scale by 1/sqrt(norm_sq).  `norm_sq_out` may be NULL. */
int qcs_hip_normalize(qcs_hip_state *s, double *norm_sq_out);
/* sums of |a|^2 over bit `qubit` = 0 and = 1 (qc_measure's p0 sweep,
 * src/qcs.c:253-259) */
int qcs_hip_qubit_probs(qcs_hip_state *s, int qubit, double *p0, double *p1);
/* qc_measure's collapse (src/qcs.c:261-283): with u = rand()/RAND_MAX
 * supplied by the host, outcome = (u <= p0) ? 0 : 1; zero the other half and
 * normalize.  Returns the outcome in *outcome. */
int qcs_hip_measure(qcs_hip_state *s, int qubit, double u, int *outcome);
/* qc_measure_all's loop (src/qcs.c:296-302): qcs_hip_measure of qubits[0],
 * qubits[1], ... in order with the uniforms u[0], u[1], ..., the outcome of
 * each decided on the device, so the whole sequence needs one host
 * synchronisation instead of one per qubit.  The same reductions give the
 * same p0 and p1 as qcs_hip_measure. */
int qcs_hip_measure_seq(qcs_hip_state *s, int count, const int *qubits,
                        const double *u, int *outcomes);
/* |a_index|^2 computed as re*re + im*im (qc_get_probability,
 * src/qcs.c:391-395) */
int qcs_hip_probability(qcs_hip_state *s, long index, double *p);
/* first index of the largest |a|^2; 0 when all are zero
 * (qc_find_most_likely_state, src/qcs.c:464-478) */
int qcs_hip_argmax(qcs_hip_state *s, long *index);
/* qc_run_shots (src/qcs.c:575-607): for each uniform u[s] (host rand()
 * stream) the first index i with u < sum_{j<=i} |a_j|^2, or -1 when u is not
 * below the total (the reference drops such shots).  FP64 chunk sums, a
 * prefix scan of the chunk sums and a per-shot search; the CDF is never
 * materialised. */
int qcs_hip_sample(qcs_hip_state *s, long shots, const double *u,
                   long *index_out);
/* qcs_hip_sample_noalloc: the same samples as qcs_hip_sample for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  It
 * calls no hipMalloc, hipFree or hipHostMalloc: `u_dev` (shots doubles) and
 * `index_out_dev` (shots longs) are device-accessible buffers of the caller,
 * and the FP64 chunk sums live in one file-scope device array per GPU of
 * 2 * 2^(QCS_HIP_NOALLOC_MAX_QUBITS - 12) doubles (4 MiB), held until the
 * process ends.  States of more than QCS_HIP_NOALLOC_MAX_QUBITS qubits are
 * rejected with QCS_HIP_EINVAL.  The work is queued on the state's stream and
 * the call returns without waiting: qcs_hip_sync(s) before reading
 * index_out_dev.  Every call on one GPU shares that array, so calls whose
 * device work overlaps (states on different streams, or several host
 * threads) race; the caller orders them, e.g. qcs_hip_sync between calls. */
#define QCS_HIP_NOALLOC_MAX_QUBITS 30
int qcs_hip_sample_noalloc(qcs_hip_state *s, long shots, const double *u_dev,
                           long *index_out_dev);

/* ---- fused precision form ----------------------------------------------
 * A gate for the fusion pipeline (textbook semantics). */
#define QCS_HIP_G_1Q 0   /* 2x2 `m` on `target` */
#define QCS_HIP_G_CU 1   /* 2x2 `m` on `target` where bit `control` = 1 */
#define QCS_HIP_G_SWAP 2 /* swap qubits `target` and `control` */
typedef struct qcs_hip_gate {
  int kind;
  int target;
  int control;
  int reserved;
  double m[8];
} qcs_hip_gate;

/* One fused block: `num_gates` consecutive entries of the gate array,
 * starting at `first_gate`, acting only on the k qubits `targets`
 * (ascending, 1 <= k <= 5). */
#define QCS_HIP_MAX_FUSE 5
typedef struct qcs_hip_block {
  int k;
  int first_gate;
  int num_gates;
  int reserved;
  int targets[8];
} qcs_hip_block;

/* Compose each block's 2^k x 2^k unitary on the device in FP64, round it to
 * FP32 and apply it to the state: c64 states as a complex GEMM on the FP32
 * matrix cores (four real products, FP32 accumulation); c128 states with an
 * FP64 dense kernel.  Textbook mode only.  No reference counterpart (the
 * reference applies every gate eagerly, src/qcs.c:159-164). */
int qcs_hip_apply_fused(qcs_hip_state *s, const qcs_hip_block *blocks,
                        int num_blocks, const qcs_hip_gate *gates,
                        int num_gates);
/* Apply one dense k-qubit unitary U (2^k x 2^k complex, row-major,
 * interleaved {re, im} doubles) on ascending `targets`. */
int qcs_hip_apply_unitary(qcs_hip_state *s, int k, const int *targets,
                          const double *u);

/* Tile programs: several fused blocks per HBM pass.  A pass names
 * `tile_bits` index bits (ascending in `bits`, and containing bits 0..3,
 * so a tile is read in runs of at least 128 bytes);
 * the state is streamed as 2^tile_bits-amplitude tiles and each of the
 * pass's `num_blocks` consecutive blocks (whose targets lie in `bits`) is
 * applied to every tile in LDS before the tile is written back.  complex64
 * states with num_qubits >= tile_bits + 3; tile_bits is 13; blocks of a
 * pass act on at most QCS_HIP_TILE_KMAX qubits and a pass holds at most
 * QCS_HIP_TILE_MAX_BLOCKS blocks (their operands stay in LDS;
 * QCS_HIP_TILE_MAX_BLOCKS_SPLIT with QCS_HIP_TILE_MATH_SPLIT_BF16).  The
 * call returns once the passes are queued on the state's stream (the plan
 * arrays may be freed at once); a fault in them is reported by the next call
 * that waits for the stream (qcs_hip_sync, reads, the next plan).  No
 * reference counterpart. */
#define QCS_HIP_TILE_BITS 13
#define QCS_HIP_TILE_LOW_BITS 4
#define QCS_HIP_TILE_KMAX 4
#define QCS_HIP_TILE_MAX_BLOCKS 22
#define QCS_HIP_TILE_MAX_BLOCKS_SPLIT 15
typedef struct qcs_hip_pass {
  int tile_bits;
  int first_block;
  int num_blocks;
  int reserved;
  int bits[16];
} qcs_hip_pass;
int qcs_hip_apply_passes(qcs_hip_state *s, const qcs_hip_pass *passes,
                         int num_passes, const qcs_hip_block *blocks,
                         int num_blocks, const qcs_hip_gate *gates,
                         int num_gates);
/* qcs_hip_apply_passes with the FP64 renormalisation of qcs_hip_normalize
 * (q_state_normalize's rule, decided on the device): the fused flush.  No
 * counterpart in the reference (its FP64 gates keep the norm).  In the
 * three-product form with two passes or more, the pass before the last sums
 * |x|^2 of the amplitudes it writes (FP64, one partial per work-group, a
 * fixed order) and the last pass scales its first block by 1/sqrt(norm_sq):
 * the state is renormalised ahead of the last pass, its norm after it is 1
 * to that pass's FP32 rounding, and no kernel reads the state only for the
 * norm or the rescale.  With one such pass, it sums the norm of what it
 * writes and the rescale follows; in the other forms
 * qcs_hip_normalize(s, NULL) follows the passes. */
int qcs_hip_apply_passes_normalize(qcs_hip_state *s, const qcs_hip_pass *passes,
                                   int num_passes, const qcs_hip_block *blocks,
                                   int num_blocks, const qcs_hip_gate *gates,
                                   int num_gates);

/* Launch-shape knobs of the state's kernels (for the autotuner).
 * QCS_HIP_TUNE_TILE_THREADS: 512 or 1024 threads per tile-program
 * work-group (default 1024); QCS_HIP_TUNE_TILE_GRID: work-groups of the
 * persistent tile-program grid when one fits per CU (multiple of 8, default
 * 256; doubled when two fit). */
#define QCS_HIP_TUNE_TILE_THREADS 0
#define QCS_HIP_TUNE_TILE_GRID 1
/* QCS_HIP_TUNE_TILE_MATH: arithmetic of the tile-program blocks, one of: */
#define QCS_HIP_TUNE_TILE_MATH 2
#define QCS_HIP_TILE_MATH_FP32 0       /* FP32 matrix-core GEMM (default) */
#define QCS_HIP_TILE_MATH_SPLIT_BF16 1 /* the FP32 matrix as three bf16 parts,
                                          each B value as two, summed on the
                                          bf16 matrix cores with FP32
                                          accumulation: the state is rounded
                                          to 16 significant bits per block
                                          (1 - F per block ~1e-11) */
/* QCS_HIP_TUNE_TILE_K4_FORM: how FP32 tile programs on 1024 threads run
 * blocks of 4 qubits, one of (default QCS_HIP_TILE_K4_GAUSS): */
#define QCS_HIP_TUNE_TILE_K4_FORM 3
#define QCS_HIP_TILE_K4_CHAIN 0  /* one 32x32 matrix-core chain per wave
                                    (v_mfma_f32_32x32x2f32) */
#define QCS_HIP_TILE_K4_HALVES 1 /* two 16-column halves on
                                    v_mfma_f32_16x16x4f32, software pipelined:
                                    a half's LDS reads and stores, and the next
                                    block's, overlap the other half's chain */
#define QCS_HIP_TILE_K4_GAUSS 2  /* the pipelined halves with three real
                                    products per complex product (Gauss):
                                    Ur xr - Ui xi and (Ur + Ui)(xr + xi) -
                                    Ur xr - Ui xi, 3/4 of the matrix-core
                                    work of the other forms */
/* QCS_HIP_TUNE_TILE_DIRECT: 1 (default) or 0.  With 1, the three-product
 * form reads a pass's first block's operands straight from HBM into
 * registers and writes its last block's results straight back, wherever
 * both blocks leave the 5 lowest qubits to the lanes (no LDS round trip at
 * the tile boundaries); the results agree to FP32 rounding either way. */
#define QCS_HIP_TUNE_TILE_DIRECT 4
/* QCS_HIP_TUNE_BLOCK_VARIANT(k): launch variant of the one-block-per-pass
 * matrix-core kernel for blocks of k qubits (1..5), one of: */
#define QCS_HIP_TUNE_BLOCK_VARIANT(k) (8 + (k))
#define QCS_HIP_BLOCK_TB12_256 0      /* 4096-amplitude tiles, 256 threads */
#define QCS_HIP_BLOCK_TB12_256_PIPE 1 /* same, next tile's loads pipelined */
#define QCS_HIP_BLOCK_TB13_256_PIPE 2 /* 8192-amplitude tiles, 256 threads */
#define QCS_HIP_BLOCK_TB13_512_PIPE 3 /* 8192-amplitude tiles, 512 threads
                                         (default) */
int qcs_hip_set_tuning(qcs_hip_state *s, int knob, int value);
/* the current value of a knob (what the autotuner chose, or the default) */
int qcs_hip_get_tuning(qcs_hip_state *s, int knob, int *value);

/* ---- comparison ----------------------------------------------------------
 * FP64 on the device: one_minus_f = 1 - |<a|b>|^2 / (|a|^2 |b|^2) and the
 * max-abs amplitude difference.  States may differ in precision but not in
 * size.  Either output may be NULL. */
int qcs_hip_compare(qcs_hip_state *a, qcs_hip_state *b, double *one_minus_f,
                    double *max_abs);

/* ---- kernel timing ----------------------------------------------------
 * When enabled, every kernel the state launches is bracketed by HIP events
 * recorded on the state's stream; totals are kept per kernel class. */
#define QCS_HIP_K_1Q 0
#define QCS_HIP_K_CU 1
#define QCS_HIP_K_FUSED 2
#define QCS_HIP_K_COMPOSE 3
#define QCS_HIP_K_REDUCE 4
#define QCS_HIP_K_OTHER 5
#define QCS_HIP_K_COUNT 6
int qcs_hip_timing_enable(qcs_hip_state *s, int enable);
/* per class: total ms, launches, algorithmic bytes, flops since enable */
int qcs_hip_timing_read(qcs_hip_state *s, int kernel_class, double *total_ms,
                        long *launches, double *bytes, double *flops);
/* per class: the bytes of the 128-byte lines the launches read and write
 * since enable (HBM moves whole lines: a gate whose changed amplitudes
 * share lines with unchanged ones, e.g. a controlled gate whose control is
 * qubit 0, 1 or 2 on a complex128 state, moves more than its algorithmic
 * bytes).  No counterpart in the reference (kernel timing). */
int qcs_hip_timing_moved(qcs_hip_state *s, int kernel_class, double *moved_bytes);

/* ---- batched circuits ----------------------------------------------------
 * Many independent small circuits in one persistent launch.  Each circuit
 * starts from |0...0>, runs its gates (textbook semantics) and draws
 * `shots` samples exactly as qcs_hip_sample would on a state of its own.
 * A work-group takes the next circuit from a device queue (largest first),
 * keeps the state in LDS when it fits (n <= tile_bits) and otherwise
 * streams it through LDS tile by tile, one pass per entry of its pass list.
 *   QCS_HIP_C128: the gates are applied one by one with the arithmetic of
 *     qcs_hip_apply_1q / qcs_hip_apply_cu / qcs_hip_apply_swap, leaving out
 *     only products by exact-zero matrix entries, so the final amplitudes
 *     equal those of a qcs_hip_state of its own (equal values; a zero may
 *     differ in sign) and so do the samples.  A pass's first_block / num_blocks count gates; a pass is a
 *     run of consecutive gates whose qubits lie in the pass's bits.
 *   QCS_HIP_C64: fused precision form.  A pass's blocks (composed on the
 *     device in FP64 like qcs_hip_apply_fused) are applied on the FP32
 *     matrix cores; the state is then renormalised in FP64 (norm_sq > 1e-12
 *     and != 1) before sampling.
 * Pass bits are ascending (complex64: including bit 0; tiles that begin
 * with bits 0..5 read HBM in >= 1 KiB runs); a circuit with
 * n <= tile_bits has tile_bits == n.  No reference counterpart: the
 * reference runs one circuit at a time (qc_run_shots, src/qcs.c:575-607). */
#define QCS_HIP_BATCH_MAX_QUBITS 22
#define QCS_HIP_BATCH_C128_TILE_BITS 13 /* largest LDS tile, complex128 */
#define QCS_HIP_BATCH_C64_TILE_BITS 14  /* largest LDS tile, complex64 */
typedef struct qcs_hip_batch qcs_hip_batch;
typedef struct qcs_hip_batch_circuit {
  int num_qubits;  /* 1 .. QCS_HIP_BATCH_MAX_QUBITS */
  int precision;   /* QCS_HIP_C128 or QCS_HIP_C64 */
  int tile_bits;   /* bits per LDS tile (== num_qubits: LDS-resident) */
  int first_pass;  /* passes[first_pass .. first_pass + num_passes) */
  int num_passes;  /* >= 1 */
  int reserved[3];
} qcs_hip_batch_circuit;
/* flags of qcs_hip_batch_create */
#define QCS_HIP_BATCH_KEEP_STATES 1 /* keep every final state in HBM for
                                       qcs_hip_batch_read_state */
#define QCS_HIP_BATCH_GRAPH 2       /* capture the launch sequence as a
                                       hipGraph and replay it while the
                                       batch's shape is unchanged */
int qcs_hip_batch_create(qcs_hip_batch **out, int flags);
void qcs_hip_batch_destroy(qcs_hip_batch *b);
/* QCS_HIP_TUNE_BATCH_THREADS: 256 or 512 threads per work-group
 * (default 512). */
#define QCS_HIP_TUNE_BATCH_THREADS 0
int qcs_hip_batch_set_tuning(qcs_hip_batch *b, int knob, int value);
/* Run `num_circuits` circuits; u holds num_circuits * shots uniforms
 * (circuit-major, the order the host's rand() stream would feed
 * qc_run_shots circuit after circuit) and index_out receives the sampled
 * basis index of every shot (-1: dropped, as qcs_hip_sample). */
int qcs_hip_batch_run(qcs_hip_batch *b, const qcs_hip_batch_circuit *circuits,
                      int num_circuits, const qcs_hip_pass *passes,
                      int num_passes, const qcs_hip_block *blocks,
                      int num_blocks, const qcs_hip_gate *gates, int num_gates,
                      long shots, const double *u, long *index_out);
/* Repeat the last successful qcs_hip_batch_run with new uniforms: the
 * descriptors and composed blocks already on the device are reused, only
 * num_circuits * shots uniforms travel.  shots must equal the last run's;
 * fails with QCS_HIP_EINVAL when there is no such run (none yet, the last
 * one failed, or the work-group size changed since).  No reference
 * counterpart: for callers that sample unchanged circuits again. */
int qcs_hip_batch_rerun(qcs_hip_batch *b, long shots, const double *u,
                        long *index_out);
/* amplitudes [offset, offset + count) of circuit i's final state after the
 * last run (QCS_HIP_BATCH_KEEP_STATES only) as {re, im} doubles */
int qcs_hip_batch_read_state(qcs_hip_batch *b, int circuit, long offset,
                             long count, double *out);
/* device time of the last run's batch kernel (HIP events) and of the whole
 * launch sequence, in ms; whether the last run replayed a captured graph */
int qcs_hip_batch_last_timing(qcs_hip_batch *b, double *kernel_ms,
                              double *total_ms, int *graph_replayed);

#ifdef __cplusplus
}
#endif

#endif /* QCS_HIP_H */
