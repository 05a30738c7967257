/* Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_sanitizers.py builds this with the oracle and a sanitized
 * copy of qcs_amd/src/*.c; no GPU is used):
 *   - the oracle's gates, measurement, sampling and algorithms, sequential
 *     and with the race-free OpenMP loops, on 1..9 qubits;
 *   - the planners of fusion.c on random gate lists (every gate placed once,
 *     blocks and passes within their limits) and on rejected inputs;
 *   - the qc_* API where no device is present: every entry point with a NULL
 *     circuit or batch.
 * Prints "sanitize ok" and exits 0 when every check holds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qcs.h"
#include "qcs_hip.h"
#include "qcs_internal.h"
#include "qcs_oracle.h"

static int failures = 0;
#define CHECK(cond, what)                                              \
  do {                                                                 \
    if (!(cond)) {                                                     \
      fprintf(stderr, "FAIL %s (%s:%d)\n", what, __FILE__, __LINE__);  \
      failures++;                                                      \
    }                                                                  \
  } while (0)

/* splitmix64: a fixed stream of test inputs */
static unsigned long long rng_state = 0x1234567ULL;
static unsigned long long next_u64(void) {
  unsigned long long z = (rng_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static int below(int n) { return (int)(next_u64() % (unsigned long long)n); }
static double unit(void) { return (double)(next_u64() >> 11) / 9007199254740992.0; }

static void random_ops(int n, int ng, int *op, int *tq, int *cq, double *par) {
  int i;
  for (i = 0; i < ng; i++) {
    op[i] = n > 1 ? below(11) : below(8);
    tq[i] = below(n);
    cq[i] = tq[i];
    if (op[i] >= OQ_CNOT)
      while (cq[i] == tq[i]) cq[i] = below(n);
    par[i] = 6.283185307179586 * unit() - 3.141592653589793;
  }
}

static void oracle_checks(void) {
  int n, compat;
  for (n = 1; n <= 9; n++)
    for (compat = 0; compat <= 1; compat++) {
      enum { NG = 40 };
      int op[NG], tq[NG], cq[NG], i;
      double par[NG], m[8], u[16];
      long out[16];
      oq_state *s = oq_create(n, compat), *t = oq_create(n, compat);
      CHECK(s && t, "oq_create");
      if (!s || !t) return;
      random_ops(n, NG, op, tq, cq, par);
      oq_run(s, NG, op, tq, cq, par);
      oq_set_threads(2);
      oq_run_omp(t, NG, op, tq, cq, par, 0);
      if (!compat) {
        /* the race-free OpenMP loops give the sequential result */
        double d = 0.0;
        long k;
        for (k = 0; k < 2 * oq_size(s); k++) d = fmax(d, fabs(oq_vector(s)[k] - oq_vector(t)[k]));
        CHECK(d < 1e-12, "oq_run_omp equals oq_run");
        CHECK(fabs(oq_norm_sq(s) - 1.0) < 1e-9, "unitary gates keep the norm");
      }
      for (i = 0; i < 16; i++) u[i] = unit();
      oq_sample(s, 16, u, out);
      for (i = 0; i < 16; i++) CHECK(out[i] >= -1 && out[i] < oq_size(s), "sample index");
      (void)oq_probability(s, oq_argmax(s));
      (void)oq_measure(s, below(n), unit());
      (void)oq_normalize(s);
      oq_gate(OQ_RY, 0.3, m);
      oq_apply_1q(t, 0, m);
      oq_apply_1q_omp(t, n - 1, m, 0);
      if (n > 1) {
        oq_gate(OQ_CPHASE, 0.7, m);
        oq_apply_cu(t, 0, n - 1, m);
        oq_apply_cu_omp(t, n - 1, 0, m, compat ? 1 : 0);
        oq_apply_swap(t, 0, n - 1);
      }
      oq_phase_flip(t, oq_size(t) - 1);
      oq_diffusion(t);
      oq_qft(t);
      oq_grover(t, oq_size(t) / 2, oq_grover_iterations(n, compat) % 4);
      oq_destroy(s);
      oq_destroy(t);
    }
}

/* gates of a random circuit in the planners' form */
static void random_gates(int n, int ng, qcs_hip_gate *g) {
  int i, j;
  for (i = 0; i < ng; i++) {
    memset(&g[i], 0, sizeof(g[i]));
    g[i].kind = n > 1 ? below(3) : QCS_HIP_G_1Q;
    g[i].target = below(n);
    g[i].control = -1;
    if (g[i].kind != QCS_HIP_G_1Q) {
      g[i].control = below(n);
      while (g[i].control == g[i].target) g[i].control = below(n);
    }
    for (j = 0; j < 8; j++) g[i].m[j] = unit();
  }
}

static int same_gate(const qcs_hip_gate *a, const qcs_hip_gate *b) {
  return a->kind == b->kind && a->target == b->target && a->control == b->control &&
         memcmp(a->m, b->m, sizeof(a->m)) == 0;
}

/* every input gate appears exactly once in `ordered` */
static int permutation_of(const qcs_hip_gate *gates, const qcs_hip_gate *ordered, int ng) {
  char *used = (char *)calloc((size_t)ng + 1, 1);
  int i, j, ok = 1;
  for (i = 0; i < ng && ok; i++) {
    int found = 0;
    for (j = 0; j < ng && !found; j++)
      if (!used[j] && same_gate(&gates[i], &ordered[j])) used[j] = found = 1;
    ok = found;
  }
  free(used);
  return ok;
}

static int in_block(const qcs_hip_block *b, int q) {
  int i;
  for (i = 0; i < b->k; i++)
    if (b->targets[i] == q) return 1;
  return 0;
}

static void planner_checks(void) {
  int trial;
  for (trial = 0; trial < 300; trial++) {
    const int n = 1 + below(20), ng = below(120), kmax = 1 + below(5);
    qcs_hip_gate *g = (qcs_hip_gate *)malloc(sizeof(qcs_hip_gate) * (size_t)(ng + 1));
    qcs_hip_gate *ord = NULL;
    qcs_hip_block *blocks = NULL;
    qcs_hip_pass *passes = NULL;
    int nb = 0, np = 0, b, i;
    random_gates(n, ng, g);
    if (qcs_fusion_plan(g, ng, n, kmax, &blocks, &nb, &ord) == 0) {
      CHECK(permutation_of(g, ord, ng), "fusion plan places every gate once");
      for (b = 0; b < nb; b++) {
        /* (a two-qubit gate makes a block of 2 even at kmax = 1) */
        CHECK(blocks[b].k >= 1 && blocks[b].k <= (kmax > 2 ? kmax : 2), "fusion block width");
        for (i = blocks[b].first_gate; i < blocks[b].first_gate + blocks[b].num_gates; i++) {
          CHECK(in_block(&blocks[b], ord[i].target), "gate target in its block");
          if (ord[i].kind != QCS_HIP_G_1Q)
            CHECK(in_block(&blocks[b], ord[i].control), "gate control in its block");
        }
      }
    } else {
      CHECK(0, "qcs_fusion_plan on valid input");
    }
    free(blocks);
    free(ord);
    blocks = NULL;
    ord = NULL;
    if (n >= 5) {
      const int tb = 5 + below(n - 4 < 9 ? n - 4 : 9), low = below(tb + 1);
      const int kp = 1 + below(4), mb = 1 + below(22);
      if (qcs_pass_plan(g, ng, n, tb, low, kp, mb, &passes, &np, &blocks, &nb, &ord) == 0) {
        int next = 0, p;
        CHECK(permutation_of(g, ord, ng), "pass plan places every gate once");
        for (p = 0; p < np; p++) {
          CHECK(passes[p].first_block == next, "passes consecutive");
          CHECK(passes[p].num_blocks >= 1 && passes[p].num_blocks <= mb, "blocks per pass");
          CHECK(passes[p].tile_bits == tb, "pass width");
          for (i = 1; i < tb; i++) CHECK(passes[p].bits[i - 1] < passes[p].bits[i], "bits ascending");
          for (b = passes[p].first_block; b < passes[p].first_block + passes[p].num_blocks; b++) {
            int q;
            for (q = 0; q < blocks[b].k; q++) {
              int hit = 0;
              for (i = 0; i < tb; i++) hit |= passes[p].bits[i] == blocks[b].targets[q];
              CHECK(hit, "block targets inside the pass");
            }
          }
          next += passes[p].num_blocks;
        }
        CHECK(next == nb, "every block in a pass");
      } else {
        /* rejected only when a two-qubit gate cannot fit beside the low
           qubits */
        CHECK(tb - low < 2, "qcs_pass_plan on valid input");
      }
      free(passes);
      free(blocks);
      free(ord);
      passes = NULL;
      blocks = NULL;
      ord = NULL;
    }
    if (qcs_gate_pass_plan(g, ng, n, 1 + below(14), below(6), &passes, &np) == 0) {
      int p, total = 0;
      for (p = 0; p < np; p++) total += passes[p].num_blocks;
      CHECK(total == ng, "gate passes hold every gate");
    } else {
      CHECK(0, "qcs_gate_pass_plan on valid input");
    }
    free(passes);
    passes = NULL;
    if (qcs_gate_hoist_exact(g, ng, n, &ord) == 0) {
      CHECK(permutation_of(g, ord, ng), "hoist keeps every gate");
      for (i = 0; i < ng; i++) (void)qcs_gate_is_exact(&ord[i]);
    } else {
      CHECK(0, "qcs_gate_hoist_exact on valid input");
    }
    free(ord);
    free(g);
  }
  /* rejected inputs return -1 and allocate nothing */
  {
    qcs_hip_gate g[2];
    qcs_hip_gate *ord = NULL;
    qcs_hip_block *blocks = NULL;
    qcs_hip_pass *passes = NULL;
    int nb = 0, np = 0;
    random_gates(4, 2, g);
    CHECK(qcs_pass_plan(g, -1, 20, 13, 4, 4, 22, &passes, &np, &blocks, &nb, &ord) == -1,
          "negative gate count");
    CHECK(qcs_pass_plan(g, 2, 8, 13, 4, 4, 22, &passes, &np, &blocks, &nb, &ord) == -1,
          "state smaller than a tile");
    CHECK(qcs_pass_plan(g, 2, 20, 13, 4, 0, 22, &passes, &np, &blocks, &nb, &ord) == -1,
          "zero fusion width");
    CHECK(qcs_gate_pass_plan(g, -1, 20, 13, 4, &passes, &np) == -1, "negative gate count");
    CHECK(!passes && !blocks && !ord, "nothing allocated on rejection");
  }
}

static void null_api_checks(void) {
  t_q_circuit *c = qc_create(5);
  t_q_batch *b = qc_batch_create(0);
  int res[8];
  long idx[4];
  double amp[4], mx = 0.0;
  if (c) {
    /* a device is present: one small circuit end to end */
    qc_h(c, 0);
    qc_cnot(c, 0, 1);
    qc_measure_all(c, res);
    qc_destroy(c);
  }
  if (b) qc_batch_destroy(b);
  qc_destroy(NULL);
  qc_h(NULL, 0);
  qc_x(NULL, 0);
  qc_y(NULL, 0);
  qc_z(NULL, 0);
  qc_cnot(NULL, 0, 1);
  qc_phase(NULL, 0, 0.1);
  qc_rx(NULL, 0, 0.1);
  qc_ry(NULL, 0, 0.1);
  qc_rz(NULL, 0, 0.1);
  qc_cphase(NULL, 0, 1, 0.1);
  qc_swap(NULL, 0, 1);
  qc_barrier(NULL);
  qc_reset(NULL, 0);
  (void)qc_measure(NULL, 0);
  qc_measure_all(NULL, res);
  qc_run(NULL);
  qc_run_shots(NULL, 4, res);
  (void)qc_find_most_likely_state(NULL);
  (void)qc_get_probability(NULL, 0);
  qc_grover_search(NULL, 1);
  qc_quantum_fourier_transform(NULL);
  qc_bernstein_vazirani(NULL, 1);
  qc_ghz_state(NULL);
  CHECK(qc_get_num_qubits(NULL) == 0, "num_qubits of NULL");
  CHECK(qc_get_num_gates(NULL) == 0, "num_gates of NULL");
  qc_optimize(NULL);
  qc_flush(NULL);
  (void)qc_get_options(NULL);
  (void)qc_last_error(NULL);
  qc_sample(NULL, 4, idx);
  qc_set_basis_state(NULL, 0);
  CHECK(qc_get_amplitudes(NULL, 0, 2, amp) != 0, "amplitudes of NULL");
  CHECK(qc_set_amplitudes(NULL, 0, 2, amp) != 0, "set amplitudes of NULL");
  (void)qc_norm_sq(NULL);
  (void)qc_compare(NULL, NULL, &mx);
  (void)qc_get_num_fused_blocks(NULL);
  (void)qc_get_num_fused_passes(NULL);
  CHECK(qc_get_device_state(NULL) == NULL, "device state of NULL");
  /* a non-positive size gives a circuit without a state, as in the
     reference: every call on it takes the no-state path */
  {
    t_q_circuit *e = qc_create(0);
    if (e) {
      qc_h(e, 0);
      qc_cnot(e, 0, 1);
      (void)qc_measure(e, 0);
      (void)qc_get_probability(e, 0);
      CHECK(qc_get_num_qubits(e) == 0, "num_qubits of an empty circuit");
      qc_destroy(e);
    }
  }
  qc_batch_destroy(NULL);
  CHECK(qc_batch_sample(NULL, NULL, 0, 1, idx) != 0, "batch sample of NULL");
  CHECK(qc_batch_set_threads(NULL, 256) != 0, "batch threads of NULL");
  CHECK(qc_autotune(NULL) < 0, "autotune of NULL");
}

int main(void) {
  oracle_checks();
  planner_checks();
  null_api_checks();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("sanitize ok\n");
  return 0;
}
