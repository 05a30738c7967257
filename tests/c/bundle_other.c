/* Second translation unit of bundle_host.c: includes qcs.h (no
 * implementation) and calls the API. */
#include "qcs.h"

int qcs_bundle_other(t_q_circuit *c) {
  qc_ry(c, 1, 0.5);
  return qc_get_num_gates(c);
}
