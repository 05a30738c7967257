/*
 * qcs.c -- C89 host side of the MI355X state-vector engine.
 *
 * Implements include/qcs.h on top of the HIP backend's C ABI (qcs_hip.h).
 * The circuit keeps the reference's gate history (src/qcs.c:13-23,132-152)
 * for printing, counting and qc_optimize; the state itself never leaves HBM.
 *
 * Plain forms (QCS_PLAIN_FP64, QCS_PLAIN_FP32): every gate call enqueues one
 * kernel on the state's stream right away, like the reference's eager
 * qc_* wrappers (src/qcs.c:159-164).  Fused form (QCS_FUSED_FP32): gate calls
 * append to a recorded list; any observation flushes it through the fusion
 * planner and the FP32 matrix-core kernels, then renormalises in FP64.
 *
 * Randomness: measurement and sampling draw u = rand()/(double)RAND_MAX on
 * the host in the reference's order (one per qc_measure, src/qcs.c:261; one
 * per shot, src/qcs.c:596), so srand() seeds give the same uniform stream.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qcs_internal.h"

#define QCS_PI 3.14159265358979323846

static int is_fused(const t_q_circuit *c) {
  return (c->options & QCS_FUSED_FP32) != 0;
}

static int is_compat(const t_q_circuit *c) {
  return (c->options & QCS_REFERENCE_COMPAT) != 0;
}

/* deferred circuits hold no device state: gates are only recorded, for a
 * batched run (batch.c) */
static int is_deferred(const t_q_circuit *c) {
  return (c->options & QCS_DEFERRED) != 0;
}

/* gate calls may proceed: a device state, or a deferred recording */
static int can_apply(const t_q_circuit *c) {
  return c != NULL && (c->state != NULL || is_deferred(c));
}

/* record a backend failure: message on stderr, sticky error code */
static int check(t_q_circuit *c, int rc, const char *what) {
  if (rc != 0) {
    fprintf(stderr, "Error: %s failed: %s\n", what, qcs_hip_last_error());
    if (c) c->error = rc;
  }
  return rc;
}

/* ---------------------------------------------------------------------- */
/* gate matrices (src/q_gates.c:362-491): {m00, m01, m10, m11} re/im       */
/* ---------------------------------------------------------------------- */

static void set_m(double m[8], double a, double b, double c, double d,
                  double e, double f, double g, double h) {
  m[0] = a; m[1] = b; m[2] = c; m[3] = d;
  m[4] = e; m[5] = f; m[6] = g; m[7] = h;
}

static void gate_h(double m[8]) {
  double r = 1.0 / sqrt(2.0);
  set_m(m, r, 0.0, r, 0.0, r, 0.0, -r, 0.0);
}
static void gate_x(double m[8]) { set_m(m, 0, 0, 1, 0, 1, 0, 0, 0); }
static void gate_y(double m[8]) { set_m(m, 0, 0, 0, -1, 0, 1, 0, 0); }
static void gate_z(double m[8]) { set_m(m, 1, 0, 0, 0, 0, 0, -1, 0); }
static void gate_p(double m[8], double a) {
  set_m(m, 1, 0, 0, 0, 0, 0, cos(a), sin(a));
}
static void gate_rx(double m[8], double a) {
  double c = cos(a / 2.0), s = sin(a / 2.0);
  set_m(m, c, 0, 0, -s, 0, -s, c, 0);
}
static void gate_ry(double m[8], double a) {
  double c = cos(a / 2.0), s = sin(a / 2.0);
  set_m(m, c, 0, -s, 0, s, 0, c, 0);
}
static void gate_rz(double m[8], double a) {
  double c = cos(a / 2.0), s = sin(a / 2.0);
  set_m(m, c, -s, 0, 0, 0, 0, c, s);
}

/* ---------------------------------------------------------------------- */
/* history (src/qcs.c:109-152)                                             */
/* ---------------------------------------------------------------------- */

static char *dup_str(const char *s) {
  size_t n = strlen(s) + 1;
  char *d = (char *)malloc(n);
  if (d) memcpy(d, s, n);
  return d;
}

static void add_history(t_q_circuit *c, const char *name, int target,
                        int control, double param) {
  if (c == NULL) return;
  if (c->history_size >= c->history_capacity) {
    int cap = c->history_capacity * 2;
    char **g = (char **)realloc(c->gate_history, cap * sizeof(char *));
    int *t = (int *)realloc(c->target_qubits, cap * sizeof(int));
    int *k = (int *)realloc(c->control_qubits, cap * sizeof(int));
    double *p = (double *)realloc(c->parameters, cap * sizeof(double));
    if (g) c->gate_history = g;
    if (t) c->target_qubits = t;
    if (k) c->control_qubits = k;
    if (p) c->parameters = p;
    if (!g || !t || !k || !p) return;
    c->history_capacity = cap;
  }
  c->gate_history[c->history_size] = dup_str(name);
  c->target_qubits[c->history_size] = target;
  c->control_qubits[c->history_size] = control;
  c->parameters[c->history_size] = param;
  c->history_size++;
  c->num_gates++;
}

/* ---------------------------------------------------------------------- */
/* recording and flushing (fused form)                                     */
/* ---------------------------------------------------------------------- */

static long g_version_stamp = 0;

/* a stamp no other change has had: the counter is shared by every circuit,
   which separate threads may record into at once */
void qcs_touch(t_q_circuit *c) {
  if (c) c->version = __atomic_add_fetch(&g_version_stamp, 1, __ATOMIC_RELAXED);
}

static void record(t_q_circuit *c, int kind, int target, int control,
                   const double m[8]) {
  qcs_hip_gate *g;
  qcs_touch(c);
  if (c->pending_size >= c->pending_capacity) {
    int cap = c->pending_capacity ? c->pending_capacity * 2 : 256;
    qcs_hip_gate *p =
        (qcs_hip_gate *)realloc(c->pending, (size_t)cap * sizeof(qcs_hip_gate));
    if (!p) {
      fprintf(stderr, "Error: out of host memory recording a gate.\n");
      c->error = QCS_HIP_ENOMEM;
      return;
    }
    c->pending = p;
    c->pending_capacity = cap;
  }
  g = &c->pending[c->pending_size++];
  memset(g, 0, sizeof(*g));
  g->kind = kind;
  g->target = target;
  g->control = control;
  if (m) memcpy(g->m, m, sizeof(g->m));
}

/* tile programs (several blocks per HBM pass) apply to complex64 fused
 * states large enough for the tile kernel when QCS_FUSE_TILES asks for
 * them */
static int use_tiles(const t_q_circuit *c) {
  return is_fused(c) && (c->options & (QCS_FUSE_TILES | QCS_FUSE_SPLIT_BF16)) &&
         c->num_qubits >= QCS_HIP_TILE_BITS + 3;
}

static int fuse_k(const t_q_circuit *c) {
  int k = c->tuned_k > 0 ? c->tuned_k : (c->options >> 8) & 0xff;
  if (k < 1 || k > QCS_HIP_MAX_FUSE) k = use_tiles(c) ? QCS_HIP_TILE_KMAX : QCS_HIP_MAX_FUSE;
  if (use_tiles(c) && k > QCS_HIP_TILE_KMAX) k = QCS_HIP_TILE_KMAX;
  if (k > c->num_qubits) k = c->num_qubits;
  return k;
}

static void flush_pending(t_q_circuit *c) {
  qcs_hip_block *blocks = NULL;
  qcs_hip_gate *ordered = NULL;
  qcs_hip_pass *passes = NULL;
  int nb = 0, np = 0, rc;
  if (!c->state || c->pending_size == 0) return;
  if (use_tiles(c))
    rc = qcs_pass_plan(c->pending, c->pending_size, c->num_qubits,
                       QCS_HIP_TILE_BITS, QCS_HIP_TILE_LOW_BITS, fuse_k(c),
                       (c->options & QCS_FUSE_SPLIT_BF16)
                           ? QCS_HIP_TILE_MAX_BLOCKS_SPLIT
                           : QCS_HIP_TILE_MAX_BLOCKS,
                       &passes, &np, &blocks, &nb,
                       &ordered);
  else
    rc = qcs_fusion_plan(c->pending, c->pending_size, c->num_qubits,
                         fuse_k(c), &blocks, &nb, &ordered);
  if (rc != 0) {
    fprintf(stderr, "Error: fusion planning failed.\n");
    c->error = QCS_HIP_ENOMEM;
  } else {
    /* then the FP64 norm reduction + renormalisation, decided on the device
       (tile programs: the norm summed by their last pass) */
    if (passes)
      rc = check(c, qcs_hip_apply_passes_normalize(c->state, passes, np, blocks, nb,
                                                   ordered, c->pending_size),
                 "tile programs");
    else if ((rc = check(c, qcs_hip_apply_fused(c->state, blocks, nb, ordered,
                                                c->pending_size),
                         "fused apply")) == 0)
      rc = check(c, qcs_hip_normalize(c->state, NULL), "renormalise");
    if (rc == 0) {
      c->fused_blocks += nb;
      c->fused_passes += passes ? np : nb;
    }
  }
  free(passes);
  free(blocks);
  free(ordered);
  c->pending_size = 0;
  qcs_touch(c);
}

/* ---------------------------------------------------------------------- */
/* gate dispatch                                                            */
/* ---------------------------------------------------------------------- */

static void apply_1q(t_q_circuit *c, int q, const double m[8]) {
  if (!can_apply(c) || q < 0 || q >= c->num_qubits) {
    fprintf(stderr, "Error: Invalid arguments for 1-qubit gate application.\n");
    return;
  }
  if (is_fused(c) || is_deferred(c)) {
    record(c, QCS_HIP_G_1Q, q, -1, m);
    return;
  }
  check(c, qcs_hip_apply_1q(c->state, q, m), "1-qubit gate");
}

static void apply_2q(t_q_circuit *c, int control, int target,
                     const double m[8]) {
  if (!can_apply(c) || control < 0 || target < 0 ||
      control >= c->num_qubits || target >= c->num_qubits ||
      control == target) {
    fprintf(stderr, "Error: Invalid arguments for 2-qubit gate application.\n");
    return;
  }
  if (is_fused(c) || is_deferred(c)) {
    record(c, QCS_HIP_G_CU, target, control, m);
    return;
  }
  check(c, qcs_hip_apply_cu(c->state, control, target, m), "2-qubit gate");
}

/* ---------------------------------------------------------------------- */
/* creation / destruction (src/qcs.c:30-101)                               */
/* ---------------------------------------------------------------------- */

t_q_circuit *qc_create_ex(int num_qubits, int options) {
  t_q_circuit *c;
  int precision, flags = 0, rc;
  if ((options & QCS_FUSED_FP32) && (options & QCS_REFERENCE_COMPAT)) {
    fprintf(stderr, "Error: QCS_REFERENCE_COMPAT applies to the plain forms only.\n");
    return NULL;
  }
  if ((options & QCS_FUSE_SPLIT_BF16) && !(options & QCS_FUSED_FP32)) {
    fprintf(stderr, "Error: QCS_FUSE_SPLIT_BF16 applies to the fused form only.\n");
    return NULL;
  }
  c = (t_q_circuit *)malloc(sizeof(t_q_circuit));
  if (!c) return NULL;
  memset(c, 0, sizeof(*c));
  qcs_touch(c);
  c->num_qubits = num_qubits;
  c->options = options;
  c->history_capacity = 100;
  c->gate_history = (char **)malloc(c->history_capacity * sizeof(char *));
  c->target_qubits = (int *)malloc(c->history_capacity * sizeof(int));
  c->control_qubits = (int *)malloc(c->history_capacity * sizeof(int));
  c->parameters = (double *)malloc(c->history_capacity * sizeof(double));
  if (!c->gate_history || !c->target_qubits || !c->control_qubits ||
      !c->parameters) {
    qc_destroy(c);
    return NULL;
  }
  if (num_qubits <= 0) {
    /* q_state_init's message; the circuit exists without a state
       (src/q_state.c:38-41, src/qcs.c:54) */
    fprintf(stderr, "Error: Number of qubits must be positive.\n");
    return c;
  }
  if (options & QCS_DEFERRED) {
    if ((options & (QCS_PLAIN_FP32 | QCS_REFERENCE_COMPAT | QCS_FUSE_TILES |
                    QCS_FUSE_SPLIT_BF16)) ||
        num_qubits > QCS_HIP_BATCH_MAX_QUBITS) {
      fprintf(stderr, "Error: QCS_DEFERRED takes 1..%d qubits, complex128 "
                      "plain or QCS_FUSED_FP32.\n", QCS_HIP_BATCH_MAX_QUBITS);
      qc_destroy(c);
      return NULL;
    }
    return c;
  }
  precision = (options & (QCS_FUSED_FP32 | QCS_PLAIN_FP32)) ? QCS_HIP_C64
                                                            : QCS_HIP_C128;
  if (options & QCS_REFERENCE_COMPAT) flags |= QCS_HIP_COMPAT;
  rc = qcs_hip_state_create(&c->state, num_qubits, precision, flags);
  if (rc != 0) {
    fprintf(stderr, "Error: could not create a %d-qubit device state: %s\n",
            num_qubits, qcs_hip_last_error());
    qc_destroy(c);
    return NULL;
  }
  if ((options & QCS_FUSE_SPLIT_BF16) &&
      qcs_hip_set_tuning(c->state, QCS_HIP_TUNE_TILE_MATH,
                         QCS_HIP_TILE_MATH_SPLIT_BF16) != 0) {
    fprintf(stderr, "Error: %s\n", qcs_hip_last_error());
    qc_destroy(c);
    return NULL;
  }
  return c;
}

t_q_circuit *qc_create(int num_qubits) {
  return qc_create_ex(num_qubits, QCS_PLAIN_FP64);
}

int qc_get_options(t_q_circuit *circuit) {
  return circuit ? circuit->options : 0;
}

void qc_destroy(t_q_circuit *circuit) {
  int i;
  if (!circuit) return;
  if (circuit->state) qcs_hip_state_destroy(circuit->state);
  if (circuit->gate_history) {
    for (i = 0; i < circuit->history_size; i++) free(circuit->gate_history[i]);
    free(circuit->gate_history);
  }
  free(circuit->target_qubits);
  free(circuit->control_qubits);
  free(circuit->parameters);
  free(circuit->pending);
  free(circuit);
}

/* ---------------------------------------------------------------------- */
/* gates (src/qcs.c:159-229,431-437,485-515)                                */
/* ---------------------------------------------------------------------- */

void qc_h(t_q_circuit *circuit, int qubit) {
  double m[8];
  gate_h(m);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "H", qubit, -1, 0.0);
}

void qc_x(t_q_circuit *circuit, int qubit) {
  double m[8];
  gate_x(m);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "X", qubit, -1, 0.0);
}

void qc_y(t_q_circuit *circuit, int qubit) {
  double m[8];
  gate_y(m);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "Y", qubit, -1, 0.0);
}

void qc_z(t_q_circuit *circuit, int qubit) {
  double m[8];
  gate_z(m);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "Z", qubit, -1, 0.0);
}

void qc_cnot(t_q_circuit *circuit, int control, int target) {
  double m[8];
  gate_x(m);
  apply_2q(circuit, control, target, m);
  add_history(circuit, "CNOT", target, control, 0.0);
}

void qc_phase(t_q_circuit *circuit, int qubit, double angle) {
  double m[8];
  gate_p(m, angle);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "P", qubit, -1, angle);
}

void qc_rx(t_q_circuit *circuit, int qubit, double angle) {
  double m[8];
  gate_rx(m, angle);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "RX", qubit, -1, angle);
}

void qc_ry(t_q_circuit *circuit, int qubit, double angle) {
  double m[8];
  gate_ry(m, angle);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "RY", qubit, -1, angle);
}

void qc_rz(t_q_circuit *circuit, int qubit, double angle) {
  double m[8];
  gate_rz(m, angle);
  apply_1q(circuit, qubit, m);
  add_history(circuit, "RZ", qubit, -1, angle);
}

void qc_cphase(t_q_circuit *circuit, int control, int target, double angle) {
  double m[8];
  gate_p(m, angle);
  apply_2q(circuit, control, target, m);
  add_history(circuit, "CPHASE", target, control, angle);
}

void qc_swap(t_q_circuit *circuit, int qubit_a, int qubit_b) {
  if (!can_apply(circuit) || qubit_a < 0 ||
      qubit_b < 0 || qubit_a >= circuit->num_qubits ||
      qubit_b >= circuit->num_qubits || qubit_a == qubit_b ||
      is_compat(circuit)) {
    fprintf(stderr, "Error: Invalid arguments for swap.\n");
  } else if (is_fused(circuit) || is_deferred(circuit)) {
    record(circuit, QCS_HIP_G_SWAP, qubit_a, qubit_b, NULL);
  } else {
    check(circuit, qcs_hip_apply_swap(circuit->state, qubit_a, qubit_b), "swap");
  }
  if (circuit) add_history(circuit, "SWAP", qubit_a, qubit_b, 0.0);
}

/* ---------------------------------------------------------------------- */
/* circuit operations                                                       */
/* ---------------------------------------------------------------------- */

void qc_flush(t_q_circuit *circuit) {
  if (!circuit || !circuit->state) return;
  flush_pending(circuit);
  check(circuit, qcs_hip_sync(circuit->state), "synchronise");
}

int qc_last_error(t_q_circuit *circuit) { return circuit ? circuit->error : 0; }

void qc_barrier(t_q_circuit *circuit) {
  if (circuit && (circuit->options & QCS_RENORM_AT_BARRIER))
    flush_pending(circuit);
  add_history(circuit, "BARRIER", -1, -1, 0.0);
}

/* src/qcs.c:237-284: prob_0 over bit = 0, u = rand()/RAND_MAX, outcome 0
 * iff u <= prob_0, collapse and renormalise (all on the device) */
int qc_measure(t_q_circuit *circuit, int qubit) {
  double u;
  int outcome = 0;
  if (circuit == NULL || circuit->state == NULL || qubit < 0 ||
      qubit >= circuit->num_qubits)
    return 0;
  flush_pending(circuit);
  u = rand() / (double)RAND_MAX;
  if (check(circuit, qcs_hip_measure(circuit->state, qubit, u, &outcome),
            "measure"))
    return 0;
  return outcome;
}

/* src/qcs.c:296-302: qc_measure of every qubit in order.  The n uniforms
 * are drawn first, in the same order, and the measurements run on the
 * device one after another with a single synchronisation at the end. */
void qc_measure_all(t_q_circuit *circuit, int *results) {
  int i, n;
  int *qubits;
  double *u;
  if (circuit == NULL || results == NULL) return;
  n = circuit->num_qubits;
  qubits = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  u = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (circuit->state == NULL || n <= 0 || !qubits || !u) {
    for (i = 0; i < n; i++) results[i] = qc_measure(circuit, i);
  } else {
    flush_pending(circuit);
    for (i = 0; i < n; i++) {
      qubits[i] = i;
      u[i] = rand() / (double)RAND_MAX;
    }
    if (check(circuit, qcs_hip_measure_seq(circuit->state, n, qubits, u, results),
              "measure"))
      for (i = 0; i < n; i++) results[i] = 0;
  }
  free(qubits);
  free(u);
  add_history(circuit, "MEASURE", -1, -1, 0.0);
}

void qc_reset(t_q_circuit *circuit, int qubit) {
  if (qc_measure(circuit, qubit) == 1) qc_x(circuit, qubit);
  add_history(circuit, "RESET", qubit, -1, 0.0);
}

void qc_run(t_q_circuit *circuit) {
  int *results;
  if (!circuit) return;
  results = (int *)malloc((size_t)(circuit->num_qubits > 0 ? circuit->num_qubits : 1) * sizeof(int));
  if (results) {
    qc_measure_all(circuit, results);
    free(results);
  }
}

/* u[s] for every shot from rand(), in the reference's order */
static double *draw_uniforms(int shots) {
  int s;
  double *u = (double *)malloc((size_t)shots * sizeof(double));
  if (!u) return NULL;
  for (s = 0; s < shots; s++) u[s] = rand() / (double)RAND_MAX;
  return u;
}

void qc_sample(t_q_circuit *circuit, int shots, long *indices) {
  double *u;
  if (!circuit || !circuit->state || shots <= 0 || !indices) return;
  flush_pending(circuit);
  u = draw_uniforms(shots);
  if (!u) return;
  check(circuit, qcs_hip_sample(circuit->state, shots, u, indices), "sample");
  free(u);
}

/* src/qcs.c:575-607: results holds 2^n ints, zeroed here; a shot whose u is
 * not below the total probability is dropped */
void qc_run_shots(t_q_circuit *circuit, int shots, int *results) {
  long *idx;
  long num_states;
  int s;
  if (!circuit || shots <= 0 || !results || !circuit->state) return;
  idx = (long *)malloc((size_t)shots * sizeof(long));
  if (!idx) return;
  num_states = 1L << circuit->num_qubits;
  for (s = 0; s < shots; s++) idx[s] = -1;
  qc_sample(circuit, shots, idx);
  memset(results, 0, (size_t)num_states * sizeof(int));
  for (s = 0; s < shots; s++)
    if (idx[s] >= 0 && idx[s] < num_states) results[idx[s]]++;
  free(idx);
}

/* ---------------------------------------------------------------------- */
/* state access                                                             */
/* ---------------------------------------------------------------------- */

long qc_find_most_likely_state_l(t_q_circuit *circuit) {
  long idx = 0;
  if (!circuit || !circuit->state) return 0;
  flush_pending(circuit);
  if (check(circuit, qcs_hip_argmax(circuit->state, &idx), "argmax")) return 0;
  return idx;
}

int qc_find_most_likely_state(t_q_circuit *circuit) {
  return (int)qc_find_most_likely_state_l(circuit);
}

double qc_get_probability_l(t_q_circuit *circuit, long state) {
  double p = 0.0;
  if (!circuit || !circuit->state) return 0.0;
  if (state < 0 || state >= (1L << circuit->num_qubits)) return 0.0;
  flush_pending(circuit);
  if (check(circuit, qcs_hip_probability(circuit->state, state, &p),
            "probability"))
    return 0.0;
  return p;
}

double qc_get_probability(t_q_circuit *circuit, int state) {
  return qc_get_probability_l(circuit, (long)state);
}

void qc_set_basis_state(t_q_circuit *circuit, long index) {
  if (!circuit) return;
  if (is_deferred(circuit)) {
    /* a deferred circuit always starts from |0...0>: drop its recorded
       gates so it can be recorded again (e.g. with new angles) */
    if (index != 0) {
      fprintf(stderr, "Error: a deferred circuit starts from |0...0>\n");
      return;
    }
    circuit->pending_size = 0;
    qcs_touch(circuit);
    return;
  }
  if (!circuit->state) return;
  if (index < 0 || index >= (1L << circuit->num_qubits)) {
    fprintf(stderr, "Error: Invalid state or basis index\n");
    return;
  }
  circuit->pending_size = 0;
  check(circuit, qcs_hip_state_set_basis(circuit->state, index), "set basis");
}

int qc_get_amplitudes(t_q_circuit *circuit, long offset, long count,
                      double *out) {
  if (!circuit || !circuit->state) return QCS_HIP_EINVAL;
  flush_pending(circuit);
  return check(circuit, qcs_hip_read(circuit->state, offset, count, out),
               "read amplitudes");
}

int qc_set_amplitudes(t_q_circuit *circuit, long offset, long count,
                      const double *in) {
  if (!circuit || !circuit->state) return QCS_HIP_EINVAL;
  flush_pending(circuit);
  return check(circuit, qcs_hip_write(circuit->state, offset, count, in),
               "write amplitudes");
}

double qc_norm_sq(t_q_circuit *circuit) {
  double ns = 0.0;
  if (!circuit || !circuit->state) return 0.0;
  flush_pending(circuit);
  check(circuit, qcs_hip_norm_sq(circuit->state, &ns), "norm");
  return ns;
}

double qc_compare(t_q_circuit *a, t_q_circuit *b, double *max_abs) {
  double omf = 1.0, mx = 0.0;
  if (!a || !b || !a->state || !b->state) return 1.0;
  flush_pending(a);
  flush_pending(b);
  check(a, qcs_hip_compare(a->state, b->state, &omf, &mx), "compare");
  if (max_abs) *max_abs = mx;
  return omf;
}

long qc_get_num_fused_blocks(t_q_circuit *circuit) {
  return circuit ? circuit->fused_blocks : 0;
}

long qc_get_num_fused_passes(t_q_circuit *circuit) {
  return circuit ? circuit->fused_passes : 0;
}

void *qc_get_device_state(t_q_circuit *circuit) {
  if (!circuit || !circuit->state) return NULL;
  flush_pending(circuit);
  return (void *)circuit->state;
}

/* src/q_state.c:139-168 */
void qc_print_state(t_q_circuit *circuit, int solution_index) {
  long i, size, max_print, last_index;
  double v[2];
  if (!circuit || !circuit->state) return;
  flush_pending(circuit);
  size = 1L << circuit->num_qubits;
  max_print = size > 8 ? 4 : size;
  printf("--- Quantum State (%d Qubits) ---\n", circuit->num_qubits);
  for (i = 0; i < max_print; i++) {
    v[0] = v[1] = 0.0;
    qcs_hip_read(circuit->state, i, 1, v);
    printf("|%ld>: %f + i%f%s\n", i, v[0], v[1],
           i == solution_index ? " <-- SOLUTION" : "");
  }
  if (solution_index >= max_print && solution_index < size - 1) {
    v[0] = v[1] = 0.0;
    qcs_hip_read(circuit->state, solution_index, 1, v);
    printf("...\n");
    printf("|%d>: %f + i%f <-- SOLUTION\n", solution_index, v[0], v[1]);
  }
  if (size > max_print) {
    printf("...\n");
    last_index = size - 1;
    v[0] = v[1] = 0.0;
    qcs_hip_read(circuit->state, last_index, 1, v);
    printf("|%ld>: %f + i%f%s\n", last_index, v[0], v[1],
           last_index == solution_index ? " <-- SOLUTION" : "");
  }
  printf("----------------------------------\n");
}

/* src/qcs.c:308-374 */
void qc_print_circuit(t_q_circuit *circuit) {
  int q, g, has_measurement = 0;
  if (!circuit) return;
  printf("\n╔══════════════════════════════════════════╗\n");
  printf("║              QUANTUM CIRCUIT             ║\n");
  printf("╚══════════════════════════════════════════╝\n");
  printf("Qubits: %d | Gates: %d\n\n", circuit->num_qubits, circuit->num_gates);
  for (g = 0; g < circuit->history_size; g++)
    if (strcmp(circuit->gate_history[g], "MEASURE") == 0) {
      has_measurement = 1;
      break;
    }
  for (q = 0; q < circuit->num_qubits; q++) {
    printf("q%-2d ", q);
    for (g = 0; g < circuit->history_size; g++) {
      const char *gate = circuit->gate_history[g];
      int target = circuit->target_qubits[g];
      int control = circuit->control_qubits[g];
      if (strcmp(gate, "H") == 0 && target == q) {
        printf("─H─");
      } else if (strcmp(gate, "X") == 0 && target == q) {
        printf("─X─");
      } else if (strcmp(gate, "CNOT") == 0) {
        if (control == q) printf("─∙─");
        else if (target == q) printf("─⊕─");
        else printf("───");
      } else {
        printf("───");
      }
      if (g < circuit->history_size - 1) printf("─");
    }
    if (!has_measurement) printf("─M─");
    printf("─┐\n");
  }
  printf("\nGATE SEQUENCE: ");
  for (g = 0; g < circuit->history_size; g++) {
    const char *gate = circuit->gate_history[g];
    if (strcmp(gate, "CNOT") == 0)
      printf("CNOT(%d,%d) ", circuit->control_qubits[g], circuit->target_qubits[g]);
    else if (strcmp(gate, "MEASURE") == 0)
      printf("MEASURE ");
    else
      printf("%s(%d) ", gate, circuit->target_qubits[g]);
  }
  printf("\n\n");
}

/* ---------------------------------------------------------------------- */
/* algorithms                                                               */
/* ---------------------------------------------------------------------- */

/* q_grover_iterations (src/q_utils.c:126-134): floor(pi/4 * sqrt(2^n)).
 * The reference forms 2^n with an int shift; textbook mode uses the exact
 * power, compat mode reproduces what the shift gives on x86-64 for n >= 31
 * (n = 31: INT_MIN, so no iterations; n >= 32: shift count taken mod 32). */
static long grover_iterations(int n, int compat) {
  double N;
  if (compat && n >= 31) {
    if (n == 31) return 0;
    N = (double)(1 << (n & 31));
  } else {
    N = ldexp(1.0, n);
  }
  return (long)floor((QCS_PI / 4.0) * sqrt(N));
}

void qc_grover_search_iters(t_q_circuit *circuit, long solution_state,
                            long iterations) {
  int q;
  long i, size;
  int valid;
  if (!circuit) return;
  for (q = 0; q < circuit->num_qubits; q++) qc_h(circuit, q);
  if (!circuit->state) return;
  flush_pending(circuit);
  size = 1L << circuit->num_qubits;
  valid = solution_state >= 0 && solution_state < size;
  if (is_compat(circuit)) {
    for (i = 0; i < iterations; i++) {
      if (valid)
        check(circuit, qcs_hip_phase_flip(circuit->state, solution_state),
              "phase flip");
      else
        fprintf(stderr, "Error: Invalid state or index for phase flip.\n");
      add_history(circuit, "ORACLE", (int)solution_state, -1, 0.0);
      check(circuit, qcs_hip_diffusion(circuit->state), "diffusion");
      add_history(circuit, "DIFFUSION", -1, -1, 0.0);
    }
    return;
  }
  if (!valid)
    for (i = 0; i < iterations; i++)
      fprintf(stderr, "Error: Invalid state or index for phase flip.\n");
  check(circuit, qcs_hip_grover(circuit->state, valid ? solution_state : -1,
                                iterations),
        "grover");
  for (i = 0; i < iterations; i++) {
    add_history(circuit, "ORACLE", (int)solution_state, -1, 0.0);
    add_history(circuit, "DIFFUSION", -1, -1, 0.0);
  }
}

void qc_grover_search_l(t_q_circuit *circuit, long solution_state) {
  if (!circuit) return;
  qc_grover_search_iters(circuit, solution_state,
                         grover_iterations(circuit->num_qubits, is_compat(circuit)));
}

void qc_grover_search(t_q_circuit *circuit, int solution_state) {
  qc_grover_search_l(circuit, (long)solution_state);
}

/* src/qcs.c:443-457: H(i), then CP(control j, target i, pi / 2^(j-i)) for
 * j > i, no final swaps.  The reference's divisor is (double)(1 << (j-i));
 * equal to 2^(j-i) up to j-i = 30, beyond which compat mode keeps the x86-64
 * int-shift value and textbook mode the exact power. */
void qc_quantum_fourier_transform(t_q_circuit *circuit) {
  const double m_pi = QCS_PI;
  int n, i, j;
  if (!circuit) return;
  n = circuit->num_qubits;
  for (i = 0; i < n; i++) {
    qc_h(circuit, i);
    for (j = i + 1; j < n; j++) {
      double angle;
      if (j - i <= 30)
        angle = m_pi / (double)(1 << (j - i));
      else if (is_compat(circuit))
        angle = m_pi / (double)(int)(1u << ((j - i) & 31));
      else
        angle = ldexp(m_pi, -(j - i));
      qc_cphase(circuit, j, i, angle);
    }
  }
}

/* src/qcs.c:614-644 */
void qc_bernstein_vazirani(t_q_circuit *circuit, int hidden_string) {
  int n, i;
  if (!circuit) return;
  n = circuit->num_qubits - 1;
  if (n <= 0) {
    fprintf(stderr, "Bernstein-Vazirani requires at least 2 qubits (1 input + "
                    "1 ancilla).\n");
    return;
  }
  qc_x(circuit, n);
  qc_h(circuit, n);
  for (i = 0; i < n; i++) qc_h(circuit, i);
  qc_barrier(circuit);
  for (i = 0; i < n; i++)
    if ((hidden_string >> i) & 1) qc_cnot(circuit, i, n);
  qc_barrier(circuit);
  for (i = 0; i < n; i++) qc_h(circuit, i);
}

/* src/qcs.c:521-535 */
void qc_ghz_state(t_q_circuit *circuit) {
  int n, i;
  if (!circuit) return;
  n = circuit->num_qubits;
  if (n < 2) {
    fprintf(stderr, "Error: GHZ state requires at least 2 qubits.\n");
    return;
  }
  qc_h(circuit, 0);
  for (i = 0; i < n - 1; i++) qc_cnot(circuit, i, i + 1);
}

/* ---------------------------------------------------------------------- */
/* utilities                                                                */
/* ---------------------------------------------------------------------- */

/* src/qcs.c getters; a NULL circuit reads as 0 (the reference dereferences it) */
int qc_get_num_qubits(t_q_circuit *circuit) { return circuit ? circuit->num_qubits : 0; }
int qc_get_num_gates(t_q_circuit *circuit) { return circuit ? circuit->num_gates : 0; }

/* src/qcs.c:650-698: drop adjacent identical H/X/Y/Z pairs on one qubit and
 * adjacent identical CNOT pairs from the history, restarting the scan after
 * each removal.  Like the reference it edits the log only; the state is not
 * touched. */
/* history entry that qc_optimize cancels against an identical neighbour:
 * H, X, Y, Z and CNOT (src/qcs.c:671-678) */
static int history_involution(const t_q_circuit *c, int i) {
  const char *g = c->gate_history[i];
  return strcmp(g, "H") == 0 || strcmp(g, "X") == 0 || strcmp(g, "Y") == 0 ||
         strcmp(g, "Z") == 0 || strcmp(g, "CNOT") == 0;
}

static int history_same(const t_q_circuit *c, int i, int j) {
  return strcmp(c->gate_history[i], c->gate_history[j]) == 0 &&
         c->target_qubits[i] == c->target_qubits[j] &&
         (strcmp(c->gate_history[i], "CNOT") != 0 ||
          c->control_qubits[i] == c->control_qubits[j]);
}

/* recorded gate qc_optimize may drop in pairs: the same self-inverse
 * matrices.  A deferred complex128 circuit keeps H pairs, whose product is
 * not exactly the identity in floating point, so that it still matches the
 * eager plain form amplitude for amplitude. */
static int pending_involution(const t_q_circuit *c, const qcs_hip_gate *g) {
  double m[8];
  if (g->kind == QCS_HIP_G_CU) {
    gate_x(m);
    return memcmp(g->m, m, sizeof(m)) == 0;
  }
  if (g->kind != QCS_HIP_G_1Q) return 0;
  gate_x(m);
  if (memcmp(g->m, m, sizeof(m)) == 0) return 1;
  gate_y(m);
  if (memcmp(g->m, m, sizeof(m)) == 0) return 1;
  gate_z(m);
  if (memcmp(g->m, m, sizeof(m)) == 0) return 1;
  gate_h(m);
  return is_fused(c) && memcmp(g->m, m, sizeof(m)) == 0;
}

/* src/qcs.c:650-698: drop adjacent identical H/X/Y/Z/CNOT pairs until none
 * is left.  Pairs of self-inverse entries reduce to a unique normal form,
 * so one stack pass gives the reference's result in linear time.  Gates
 * recorded and not yet applied (fused and deferred forms) are reduced the
 * same way, so the cancelled pairs are never computed. */
void qc_optimize(t_q_circuit *circuit) {
  int i, top = 0;
  if (!circuit) return;
  for (i = 0; i < circuit->history_size; i++) {
    if (top > 0 && history_involution(circuit, i) &&
        history_same(circuit, top - 1, i)) {
      free(circuit->gate_history[top - 1]);
      free(circuit->gate_history[i]);
      top--;
      circuit->num_gates -= 2;
      continue;
    }
    circuit->gate_history[top] = circuit->gate_history[i];
    circuit->target_qubits[top] = circuit->target_qubits[i];
    circuit->control_qubits[top] = circuit->control_qubits[i];
    circuit->parameters[top] = circuit->parameters[i];
    top++;
  }
  circuit->history_size = top;
  top = 0;
  for (i = 0; i < circuit->pending_size; i++) {
    const qcs_hip_gate *g = &circuit->pending[i];
    if (top > 0 && pending_involution(circuit, g)) {
      const qcs_hip_gate *p = &circuit->pending[top - 1];
      if (p->kind == g->kind && p->target == g->target &&
          (g->kind == QCS_HIP_G_1Q || p->control == g->control) &&
          memcmp(p->m, g->m, sizeof(g->m)) == 0) {
        top--;
        continue;
      }
    }
    circuit->pending[top++] = *g;
  }
  if (top != circuit->pending_size) qcs_touch(circuit);
  circuit->pending_size = top;
}
