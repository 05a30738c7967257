/*
 * fusion.c -- greedy gate-fusion planner for the fused precision form.
 *
 * The recorded gates form a DAG through their qubits (each qubit orders the
 * gates touching it).  A gate is *ready* when it is the earliest unplaced
 * gate on every qubit it touches.  A block starts from the earliest unplaced
 * gate and repeatedly
 *   1. absorbs ready gates whose qubits already lie in the block's set S;
 *   2. otherwise grows S by the ready gate that adds the fewest qubits while
 *      keeping |S| <= kmax (ties: earliest gate);
 * and closes when neither applies.  Gates are absorbed in a topological
 * order, so multiplying them in absorption order gives the block's unitary,
 * and blocks emitted in order reproduce the circuit.  For the 30-qubit
 * depth-20 layered Rz-Ry-Rz + CNOT-ladder circuit this packs 2,410 gates
 * into 99 blocks at kmax = 5 (144 at 4, 290 at 3).
 */
#include <stdlib.h>
#include <string.h>

#include "qcs_internal.h"

static int gate_qubits(const qcs_hip_gate *g, int q[2]) {
  q[0] = g->target;
  if (g->kind == QCS_HIP_G_1Q) return 1;
  q[1] = g->control;
  return 2;
}

typedef struct {
  const qcs_hip_gate *gates;
  int ng, n;
  int *start;   /* CSR offsets of the per-qubit gate lists, n + 1 */
  int *list;    /* gate indices per qubit in circuit order */
  int *ptr;     /* per qubit: position of the next unplaced gate */
  char *done;
} planner;

/* earliest unplaced gate on qubit q, or -1 */
static int frontier(const planner *p, int q) {
  int pos = p->start[q] + p->ptr[q];
  return pos < p->start[q + 1] ? p->list[pos] : -1;
}

static int ready(const planner *p, int g) {
  int q[2], nq, i;
  nq = gate_qubits(&p->gates[g], q);
  for (i = 0; i < nq; i++)
    if (frontier(p, q[i]) != g) return 0;
  return 1;
}

static void place(planner *p, int g) {
  int q[2], nq, i;
  nq = gate_qubits(&p->gates[g], q);
  for (i = 0; i < nq; i++) p->ptr[q[i]]++;
  p->done[g] = 1;
}

/* number of qubits of gate g not in the set `in` */
static int added(const planner *p, int g, const char *in) {
  int q[2], nq, i, a = 0;
  nq = gate_qubits(&p->gates[g], q);
  for (i = 0; i < nq; i++)
    if (!in[q[i]]) a++;
  return a;
}

int qcs_fusion_plan(const qcs_hip_gate *gates, int ng, int n, int kmax,
                    qcs_hip_block **blocks_out, int *nblocks_out,
                    qcs_hip_gate **ordered_out) {
  planner p;
  qcs_hip_block *blocks = NULL;
  qcs_hip_gate *ordered = NULL;
  char *in = NULL;
  int *members = NULL;
  int nb = 0, placed = 0, first = 0, i, q, rc = -1;

  *blocks_out = NULL;
  *ordered_out = NULL;
  *nblocks_out = 0;
  if (ng < 0 || n <= 0 || kmax < 1) return -1;
  if (kmax > QCS_HIP_MAX_FUSE) kmax = QCS_HIP_MAX_FUSE;
  memset(&p, 0, sizeof(p));
  p.gates = gates;
  p.ng = ng;
  p.n = n;
  p.start = (int *)calloc((size_t)n + 1, sizeof(int));
  p.list = (int *)malloc((size_t)(2 * ng + 1) * sizeof(int));
  p.ptr = (int *)calloc((size_t)n, sizeof(int));
  p.done = (char *)calloc((size_t)ng + 1, 1);
  in = (char *)calloc((size_t)n, 1);
  members = (int *)malloc((size_t)(n + 1) * sizeof(int));
  blocks = (qcs_hip_block *)malloc((size_t)(ng + 1) * sizeof(qcs_hip_block));
  ordered = (qcs_hip_gate *)malloc((size_t)(ng + 1) * sizeof(qcs_hip_gate));
  if (!p.start || !p.list || !p.ptr || !p.done || !in || !members || !blocks ||
      !ordered)
    goto out;

  /* per-qubit gate lists (CSR) */
  for (i = 0; i < ng; i++) {
    int qs[2], nq, j;
    nq = gate_qubits(&gates[i], qs);
    for (j = 0; j < nq; j++) {
      if (qs[j] < 0 || qs[j] >= n) goto out;
      p.start[qs[j] + 1]++;
    }
    if (nq == 2 && qs[0] == qs[1]) goto out;
  }
  for (q = 0; q < n; q++) p.start[q + 1] += p.start[q];
  {
    int *fill = (int *)calloc((size_t)n, sizeof(int));
    if (!fill) goto out;
    for (i = 0; i < ng; i++) {
      int qs[2], nq, j;
      nq = gate_qubits(&gates[i], qs);
      for (j = 0; j < nq; j++) p.list[p.start[qs[j]] + fill[qs[j]]++] = i;
    }
    free(fill);
  }

  while (placed < ng) {
    qcs_hip_block *b = &blocks[nb];
    int size = 0, nmem = 0, first_gate = placed;
    int qs[2], nq, j;
    while (p.done[first]) first++;
    /* seed: the earliest unplaced gate is always ready */
    nq = gate_qubits(&gates[first], qs);
    for (j = 0; j < nq; j++)
      if (!in[qs[j]]) {
        in[qs[j]] = 1;
        members[nmem++] = qs[j];
        size++;
      }
    place(&p, first);
    ordered[placed++] = gates[first];
    for (;;) {
      int progress = 0, best = -1, best_add = 0;
      /* 1. ready gates inside S */
      for (j = 0; j < nmem; j++) {
        int g;
        while ((g = frontier(&p, members[j])) >= 0 && ready(&p, g) &&
               added(&p, g, in) == 0) {
          place(&p, g);
          ordered[placed++] = gates[g];
          progress = 1;
        }
      }
      if (progress) continue;
      /* 2. grow S by the cheapest ready gate */
      for (q = 0; q < n; q++) {
        int g = frontier(&p, q), a;
        if (g < 0 || !ready(&p, g)) continue;
        a = added(&p, g, in);
        if (size + a > kmax) continue;
        if (best < 0 || a < best_add || (a == best_add && g < best)) {
          best = g;
          best_add = a;
        }
      }
      if (best < 0) break;
      nq = gate_qubits(&gates[best], qs);
      for (j = 0; j < nq; j++)
        if (!in[qs[j]]) {
          in[qs[j]] = 1;
          members[nmem++] = qs[j];
          size++;
        }
      place(&p, best);
      ordered[placed++] = gates[best];
    }
    /* emit: targets ascending */
    memset(b, 0, sizeof(*b));
    b->k = size;
    b->first_gate = first_gate;
    b->num_gates = placed - first_gate;
    {
      int k = 0;
      for (q = 0; q < n && k < size; q++)
        if (in[q]) b->targets[k++] = q;
    }
    for (j = 0; j < nmem; j++) in[members[j]] = 0;
    nb++;
  }
  rc = 0;
out:
  free(p.start);
  free(p.list);
  free(p.ptr);
  free(p.done);
  free(in);
  free(members);
  if (rc != 0) {
    free(blocks);
    free(ordered);
    return rc;
  }
  *blocks_out = blocks;
  *ordered_out = ordered;
  *nblocks_out = nb;
  return 0;
}
