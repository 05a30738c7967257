/*
 * A C89 program written against the reference's qcs.h only (plus one batch
 * call at the end), linked with -lqcs.  It prints one "ok <name>" line per
 * check and exits non-zero on the first failure.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "qcs.h"

static int failures = 0;

static void check(int cond, const char *name) {
  if (cond) {
    printf("ok %s\n", name);
  } else {
    printf("FAIL %s\n", name);
    failures++;
  }
}

int main(void) {
  t_q_circuit *c;
  int results[1 << 6];
  int i, total;

  /* Bell pair: |00> + |11> */
  c = qc_create(2);
  qc_h(c, 0);
  qc_cnot(c, 0, 1);
  check(fabs(qc_get_probability(c, 0) - 0.5) < 1e-12 &&
            fabs(qc_get_probability(c, 3) - 0.5) < 1e-12,
        "bell_probabilities");
  check(qc_get_num_gates(c) == 2, "gate_count");
  qc_destroy(c);

  /* rotations: RX(pi) maps |0> to -i|1> */
  c = qc_create(1);
  qc_rx(c, 0, 3.14159265358979323846);
  check(fabs(qc_get_probability(c, 1) - 1.0) < 1e-12, "rx_pi");
  qc_destroy(c);

  /* Grover on 6 qubits finds the marked state */
  c = qc_create(6);
  qc_grover_search(c, 42);
  check(qc_find_most_likely_state(c) == 42 && qc_get_probability(c, 42) > 0.9,
        "grover_6");
  qc_destroy(c);

  /* shots: H on qubit 0, 1000 shots split about evenly */
  srand(7);
  c = qc_create(1);
  qc_h(c, 0);
  qc_run_shots(c, 1000, results);
  check(results[0] + results[1] == 1000 && results[0] > 400 && results[0] < 600,
        "run_shots");
  qc_destroy(c);

  /* measurement collapses: X then measure gives 1 */
  c = qc_create(3);
  qc_x(c, 1);
  check(qc_measure(c, 1) == 1 && qc_measure(c, 0) == 0, "measure");
  qc_destroy(c);

  /* QFT of |0...0> is uniform */
  c = qc_create(5);
  qc_quantum_fourier_transform(c);
  check(fabs(qc_get_probability(c, 17) - 1.0 / 32.0) < 1e-12, "qft_uniform");
  qc_destroy(c);

  /* GHZ and Bernstein-Vazirani */
  c = qc_create(4);
  qc_ghz_state(c);
  check(fabs(qc_get_probability(c, 15) - 0.5) < 1e-12, "ghz");
  qc_destroy(c);
  c = qc_create(6);
  qc_bernstein_vazirani(c, 13);
  check(qc_find_most_likely_state(c) == 13, "bernstein_vazirani");
  qc_destroy(c);

  /* optimize cancels H H */
  c = qc_create(2);
  qc_h(c, 0);
  qc_h(c, 0);
  qc_optimize(c);
  check(qc_get_num_gates(c) == 0, "optimize");
  qc_destroy(c);

  /* bad arguments: message on stderr, no crash */
  c = qc_create(2);
  qc_h(c, 5);
  qc_cnot(c, 1, 1);
  check(qc_measure(c, 9) == 0, "bad_arguments");
  qc_destroy(c);

  /* batched: 64 deferred circuits in one launch, each X on qubit i % 6 */
  {
    t_q_circuit *cs[64];
    long idx[64 * 4];
    t_q_batch *b = qc_batch_create(0);
    int ok = b != NULL;
    for (i = 0; i < 64 && ok; i++) {
      cs[i] = qc_create_ex(6, QCS_DEFERRED);
      qc_x(cs[i], i % 6);
    }
    if (ok) ok = qc_batch_sample(b, cs, 64, 4, idx) == 0;
    for (i = 0; i < 64 * 4 && ok; i++)
      if (idx[i] != 1L << ((i / 4) % 6)) ok = 0;
    for (i = 0; i < 64 && b; i++) qc_destroy(cs[i]);
    qc_batch_destroy(b);
    check(ok, "batch");
  }

  total = failures;
  printf("%d failure(s)\n", total);
  return total == 0 ? 0 : 1;
}
