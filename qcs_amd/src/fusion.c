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

/*
 * Pass planner for tile programs.  The whole circuit is first fused into
 * blocks of <= kmax qubits (qcs_fusion_plan); the blocks, ordered by their
 * dependencies (a block follows the previous block on each of its qubits),
 * are then dealt into passes.  A pass starts with S = the `low_bits` lowest
 * qubits and takes, among the blocks whose predecessors are all placed and
 * whose targets fit S within `tile_bits` qubits, the one adding the fewest
 * qubits to S (ties: the earliest), until none fits or it holds max_blocks;
 * S is then filled up to `tile_bits` with the lowest qubits not in it.
 * Fusing first keeps blocks whole across pass boundaries: for the 30-qubit
 * depth-20 layered circuit (tile_bits 13, low_bits 4, kmax 4) the 2,410
 * gates make 144 blocks (143 of four qubits) in 31 passes, where fusing each
 * pass's gates separately made 196 blocks (26 of them smaller) in 26 passes.
 */

static int popcount64(unsigned long x) {
  int c = 0;
  while (x) {
    x &= x - 1;
    c++;
  }
  return c;
}

int qcs_pass_plan(const qcs_hip_gate *gates, int ng, int n, int tile_bits,
                  int low_bits, int kmax, int max_blocks,
                  qcs_hip_pass **passes_out,
                  int *npasses_out, qcs_hip_block **blocks_out,
                  int *nblocks_out, qcs_hip_gate **ordered_out) {
  qcs_hip_block *gb = NULL, *blocks = NULL;
  qcs_hip_gate *gord = NULL, *ordered = NULL;
  qcs_hip_pass *passes = NULL;
  unsigned long *tmask = NULL;
  int *npred = NULL, *succ = NULL, *nsucc = NULL, *ready = NULL, *last = NULL;
  int m = 0, nready = 0, np = 0, nb = 0, placed = 0, i, j, q, rc = -1;

  *passes_out = NULL;
  *blocks_out = NULL;
  *ordered_out = NULL;
  *npasses_out = *nblocks_out = 0;
  if (ng < 0 || n < tile_bits || tile_bits > 16 || low_bits < 0 || low_bits > tile_bits ||
      kmax < 1 || max_blocks < 1 || n > 62)
    return -1;
  /* every block must fit a pass beside the low qubits */
  if (kmax > tile_bits - low_bits) kmax = tile_bits - low_bits;
  if (kmax < 1) return -1;
  if (qcs_fusion_plan(gates, ng, n, kmax, &gb, &m, &gord) != 0) return -1;
  tmask = (unsigned long *)calloc((size_t)m + 1, sizeof(unsigned long));
  npred = (int *)calloc((size_t)m + 1, sizeof(int));
  nsucc = (int *)calloc((size_t)m + 1, sizeof(int));
  succ = (int *)malloc(((size_t)m + 1) * QCS_HIP_MAX_FUSE * sizeof(int));
  ready = (int *)malloc(((size_t)m + 1) * sizeof(int));
  last = (int *)malloc((size_t)n * sizeof(int));
  passes = (qcs_hip_pass *)calloc((size_t)m + 1, sizeof(qcs_hip_pass));
  blocks = (qcs_hip_block *)malloc(((size_t)m + 1) * sizeof(qcs_hip_block));
  ordered = (qcs_hip_gate *)malloc(((size_t)ng + 1) * sizeof(qcs_hip_gate));
  if (!tmask || !npred || !nsucc || !succ || !ready || !last || !passes || !blocks ||
      !ordered)
    goto out;
  /* dependencies: the previous block on each qubit (once per pair) */
  for (q = 0; q < n; q++) last[q] = -1;
  for (i = 0; i < m; i++) {
    for (j = 0; j < gb[i].k; j++) {
      const int t = gb[i].targets[j], pr = last[t];
      int k, dup = 0;
      tmask[i] |= 1UL << t;
      if (pr >= 0) {
        for (k = 0; k < nsucc[pr]; k++) dup |= succ[pr * QCS_HIP_MAX_FUSE + k] == i;
        if (!dup) {
          succ[pr * QCS_HIP_MAX_FUSE + nsucc[pr]++] = i;
          npred[i]++;
        }
      }
      last[t] = i;
    }
    if (npred[i] == 0) ready[nready++] = i;
  }
  while (nb < m) {
    unsigned long S = 0;
    qcs_hip_pass *ps = &passes[np];
    int count = 0, k;
    for (q = 0; q < low_bits; q++) S |= 1UL << q;
    ps->first_block = nb;
    while (count < max_blocks) {
      int best = -1, best_add = 0, r;
      for (r = 0; r < nready; r++) {
        const int b = ready[r];
        const int add = popcount64(tmask[b] & ~S);
        if (popcount64(S | tmask[b]) > tile_bits) continue;
        if (best < 0 || add < best_add || (add == best_add && b < ready[best])) {
          best = r;
          best_add = add;
        }
      }
      if (best < 0) break;
      {
        const int b = ready[best];
        ready[best] = ready[--nready];
        S |= tmask[b];
        blocks[nb] = gb[b];
        blocks[nb].first_gate = placed;
        memcpy(&ordered[placed], &gord[gb[b].first_gate],
               (size_t)gb[b].num_gates * sizeof(qcs_hip_gate));
        placed += gb[b].num_gates;
        nb++;
        count++;
        for (r = 0; r < nsucc[b]; r++) {
          const int s2 = succ[b * QCS_HIP_MAX_FUSE + r];
          if (--npred[s2] == 0) ready[nready++] = s2;
        }
      }
    }
    if (count == 0) goto out; /* cannot happen: a ready block always fits */
    for (q = 0; q < n && popcount64(S) < tile_bits; q++) S |= 1UL << q;
    ps->tile_bits = tile_bits;
    ps->num_blocks = count;
    for (q = 0, k = 0; q < n && k < tile_bits; q++)
      if ((S >> q) & 1UL) ps->bits[k++] = q;
    np++;
  }
  rc = placed == ng ? 0 : -1;
out:
  free(gb);
  free(gord);
  free(tmask);
  free(npred);
  free(nsucc);
  free(succ);
  free(ready);
  free(last);
  if (rc != 0) {
    free(passes);
    free(blocks);
    free(ordered);
    return rc;
  }
  *passes_out = passes;
  *npasses_out = np;
  *blocks_out = blocks;
  *nblocks_out = nb;
  *ordered_out = ordered;
  return 0;
}

/* close the pass in `ps`: fill S up to tile_bits with the lowest qubits not
 * in it, write the bits ascending and clear `in` */
static void close_gate_pass(qcs_hip_pass *ps, char *in, int n, int tile_bits,
                            int size) {
  int q, k;
  for (q = 0; q < n && size < tile_bits; q++)
    if (!in[q]) {
      in[q] = 1;
      size++;
    }
  ps->tile_bits = tile_bits;
  for (q = 0, k = 0; q < n; q++)
    if (in[q]) {
      if (k < tile_bits) ps->bits[k++] = q;
      in[q] = 0;
    }
}

int qcs_gate_pass_plan(const qcs_hip_gate *gates, int ng, int n, int tile_bits,
                       int low_bits, qcs_hip_pass **passes_out,
                       int *npasses_out) {
  qcs_hip_pass *passes;
  char *in;
  int np = 0, size = 0, i, q;

  *passes_out = NULL;
  *npasses_out = 0;
  if (ng < 0 || n < 1 || tile_bits < 1 || tile_bits > 16 || low_bits < 0)
    return -1;
  if (tile_bits > n) tile_bits = n;
  if (low_bits > tile_bits) low_bits = tile_bits;
  /* leave room for a 2-qubit gate outside the low qubits */
  if (tile_bits < n && low_bits > tile_bits - 2)
    low_bits = tile_bits >= 2 ? tile_bits - 2 : 0;
  for (i = 0; i < ng; i++) {
    int qs[2], nq = gate_qubits(&gates[i], qs), j;
    for (j = 0; j < nq; j++)
      if (qs[j] < 0 || qs[j] >= n) return -1;
    if (nq == 2 && qs[0] == qs[1]) return -1;
  }
  passes = (qcs_hip_pass *)calloc((size_t)ng + 1, sizeof(qcs_hip_pass));
  in = (char *)calloc((size_t)n, 1);
  if (!passes || !in) {
    free(passes);
    free(in);
    return -1;
  }
  for (q = 0; q < low_bits; q++) in[q] = 1;
  size = low_bits;
  passes[0].first_block = 0;
  for (i = 0; i < ng; i++) {
    int qs[2], nq = gate_qubits(&gates[i], qs), add = 0, j;
    for (j = 0; j < nq; j++) add += !in[qs[j]];
    if (size + add > tile_bits) {
      close_gate_pass(&passes[np], in, n, tile_bits, size);
      np++;
      passes[np].first_block = i;
      for (q = 0; q < low_bits; q++) in[q] = 1;
      size = low_bits;
    }
    for (j = 0; j < nq; j++)
      if (!in[qs[j]]) {
        in[qs[j]] = 1;
        size++;
      }
    passes[np].num_blocks++;
  }
  close_gate_pass(&passes[np], in, n, tile_bits, size);
  np++;
  free(in);
  *passes_out = passes;
  *npasses_out = np;
  return 0;
}

/* a gate whose matrix is a permutation with entries in {+-1, +-i}: applying
 * it only moves amplitudes and flips signs / swaps re and im, with no
 * rounding, so it commutes bit-exactly with any gate on other qubits */
static int is_unit(double re, double im) {
  return (im == 0.0 && (re == 1.0 || re == -1.0)) ||
         (re == 0.0 && (im == 1.0 || im == -1.0));
}

static int is_zero(double re, double im) { return re == 0.0 && im == 0.0; }

int qcs_gate_is_exact(const qcs_hip_gate *g) {
  const double *m = g->m;
  if (g->kind == QCS_HIP_G_SWAP) return 1;
  if (is_zero(m[2], m[3]) && is_zero(m[4], m[5]))
    return is_unit(m[0], m[1]) && is_unit(m[6], m[7]);
  if (is_zero(m[0], m[1]) && is_zero(m[6], m[7]))
    return is_unit(m[2], m[3]) && is_unit(m[4], m[5]);
  return 0;
}

int qcs_gate_hoist_exact(const qcs_hip_gate *gates, int ng, int n,
                         qcs_hip_gate **ordered_out) {
  int *bucket = NULL, *last = NULL, *start = NULL, *fill = NULL;
  qcs_hip_gate *out = NULL;
  int nbk = 1, i, q, rc = -1;
  *ordered_out = NULL;
  if (ng < 0 || n < 1) return -1;
  bucket = (int *)malloc(((size_t)ng + 1) * sizeof(int));
  last = (int *)calloc((size_t)n, sizeof(int));
  out = (qcs_hip_gate *)malloc(((size_t)ng + 1) * sizeof(qcs_hip_gate));
  if (!bucket || !last || !out) goto done;
  /* bucket 0 holds exact gates before the first other gate; every other gate
   * opens a bucket; an exact gate joins the latest bucket that touches one
   * of its qubits */
  for (i = 0; i < ng; i++) {
    int qs[2], nq = gate_qubits(&gates[i], qs), j, b = 0;
    for (j = 0; j < nq; j++) {
      if (qs[j] < 0 || qs[j] >= n) goto done;
      if (last[qs[j]] > b) b = last[qs[j]];
    }
    if (!qcs_gate_is_exact(&gates[i])) b = nbk++;
    bucket[i] = b;
    for (j = 0; j < nq; j++) last[qs[j]] = b;
  }
  start = (int *)calloc((size_t)nbk + 1, sizeof(int));
  fill = (int *)calloc((size_t)nbk, sizeof(int));
  if (!start || !fill) goto done;
  for (i = 0; i < ng; i++) start[bucket[i] + 1]++;
  for (q = 0; q < nbk; q++) start[q + 1] += start[q];
  /* stable within a bucket: the opening gate first, then exact gates in
   * circuit order */
  for (i = 0; i < ng; i++) out[start[bucket[i]] + fill[bucket[i]]++] = gates[i];
  rc = 0;
done:
  free(bucket);
  free(last);
  free(start);
  free(fill);
  if (rc != 0) {
    free(out);
    return rc;
  }
  *ordered_out = out;
  return 0;
}
