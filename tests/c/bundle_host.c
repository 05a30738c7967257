/*
 * Host-only use of the API (a deferred circuit: gates are recorded, no
 * device state), built either against the single-header qcs.h (this file
 * compiled with QCS_IMPLEMENTATION) or against include/qcs.h and -lqcs; the two builds
 * must print the same thing.  Calls qcs_bundle_other() from a second
 * translation unit that includes qcs.h without the implementation.
 */
#include "qcs.h"

#include <stdio.h>

int qcs_bundle_other(t_q_circuit *c);

int main(void) {
  t_q_circuit *c = qc_create_ex(6, QCS_DEFERRED);
  int i, other;
  if (!c) return 2;
  for (i = 0; i < 6; ++i) qc_h(c, i);
  for (i = 0; i + 1 < 6; ++i) qc_cnot(c, i, i + 1);
  qc_rz(c, 3, 0.25);
  qc_x(c, 2);
  qc_x(c, 2);
  qc_phase(c, 5, 1.5);
  qc_optimize(c);
  other = qcs_bundle_other(c);
  printf("gates %d qubits %d other %d options %d\n", qc_get_num_gates(c),
         qc_get_num_qubits(c), other, qc_get_options(c));
  qc_print_circuit(c);
  qc_destroy(c);
  return 0;
}
