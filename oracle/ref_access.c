/*
 * ref_access.c -- compiles the reference's circuit layer (src/qcs.c of
 * SegusFaultise/qcs, unchanged, from the reference tree given with -I) into
 * oracle/_ref/ and adds accessors for its amplitude buffers, which the
 * public API does not expose (it returns |a|^2 only, src/qcs.c:391-395).
 * TEST INFRASTRUCTURE ONLY: used to write golden vectors and as the timed
 * CPU baseline; never part of the product.
 */
#include "qcs.c"

/* current amplitudes (struct t_q_state.vector) as 2*size doubles */
double *qcsref_vector(t_q_circuit *c) { return (double *)c->state->vector; }
/* the second buffer (struct t_q_state.scratch_vector) */
double *qcsref_scratch(t_q_circuit *c) {
  return (double *)c->state->scratch_vector;
}
long qcsref_size(t_q_circuit *c) { return c->state->size; }
