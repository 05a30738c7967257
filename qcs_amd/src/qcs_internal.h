/* Internal declarations of the C89 host library. */
#ifndef QCS_INTERNAL_H
#define QCS_INTERNAL_H

#include "qcs.h"
#include "qcs_hip.h"

struct t_q_circuit {
  int num_qubits;
  int num_gates;
  qcs_hip_state *state;
  int options;
  int error;
  /* gate history, same columns as the reference (src/qcs.c:13-23) */
  char **gate_history;
  int *target_qubits;
  int *control_qubits;
  double *parameters;
  int history_size;
  int history_capacity;
  /* gates recorded for the fused form, applied at the next flush */
  qcs_hip_gate *pending;
  int pending_size;
  int pending_capacity;
  long fused_blocks;
  long fused_passes;
  int tuned_k; /* fusion width chosen by qc_autotune (0: none) */
  long version; /* changes whenever the recorded gates change (a stamp
                   unique across circuits, for batch plan reuse) */
};

/* mark a circuit's recorded gates as changed (qcs.c) */
void qcs_touch(t_q_circuit *c);

/* Greedy fusion planner (fusion.c).
 * Input: `ng` gates in circuit order on `n` qubits.  Output: `*nblocks`
 * blocks of at most `kmax` qubits each and the gates re-ordered so that each
 * block's gates are contiguous (block b owns ordered[first .. first+num)).
 * Arrays are malloc'ed; the caller frees them.  Returns 0 or -1 (no memory /
 * bad input). */
int qcs_fusion_plan(const qcs_hip_gate *gates, int ng, int n, int kmax,
                    qcs_hip_block **blocks, int *nblocks,
                    qcs_hip_gate **ordered);

/* Pass planner for tile programs (fusion.c): passes of `tile_bits` qubits
 * that always include qubits 0..low_bits-1, each holding blocks of at most
 * kmax qubits; gates re-ordered so every block's gates are contiguous and
 * blocks of a pass consecutive. */
int qcs_pass_plan(const qcs_hip_gate *gates, int ng, int n, int tile_bits,
                  int low_bits, int kmax, int max_blocks, qcs_hip_pass **passes,
                  int *npasses, qcs_hip_block **blocks, int *nblocks,
                  qcs_hip_gate **ordered);

/* Pass planner for gate-by-gate application (batched complex128 circuits):
 * the gates stay in circuit order; each pass is a maximal run of
 * consecutive gates whose qubits, together with qubits 0..low_bits-1, fit in
 * `tile_bits` qubits (filled up with the lowest unused qubits; low_bits is
 * cut to tile_bits - 2 when the state is larger than a tile, so every gate
 * fits a pass of its own).  A pass's
 * first_block / num_blocks index gates.  With n <= tile_bits there is one
 * pass of n bits (also when ng == 0). */
int qcs_gate_pass_plan(const qcs_hip_gate *gates, int ng, int n, int tile_bits,
                       int low_bits, qcs_hip_pass **passes, int *npasses);

/* 1 if the gate only permutes amplitudes and multiplies them by +-1 or
 * +-i (X, Y, Z, CNOT, CZ, SWAP, ...): it then commutes bit-exactly with
 * any gate on other qubits. */
int qcs_gate_is_exact(const qcs_hip_gate *g);
/* Reorder a gate list, moving every exact gate (qcs_gate_is_exact) to just
 * after the latest earlier gate sharing a qubit with it; the other gates
 * keep their order.  The result gives the same amplitudes as the original
 * order (equal values).  *ordered is malloc'ed.  Returns 0 or -1. */
int qcs_gate_hoist_exact(const qcs_hip_gate *gates, int ng, int n,
                         qcs_hip_gate **ordered);

#endif
