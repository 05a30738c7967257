/*
 * batch.c -- batched execution of deferred circuits (qc_batch_*).
 *
 * A QCS_DEFERRED circuit records its gates and owns no device state.  A
 * batch run plans every circuit on the host and hands the plans to the
 * backend's persistent kernel in one call (qcs_hip_batch_run):
 *   - complex128 circuits: the gates in order (exact gates such as CNOT
 *     moved up, qcs_gate_hoist_exact), cut into passes of consecutive gates
 *     whose qubits fit one LDS tile (qcs_gate_pass_plan); a circuit of at
 *     most tile_bits qubits is one pass and stays on chip;
 *   - QCS_FUSED_FP32 circuits: fused blocks of <= k qubits (qcs_fusion_plan,
 *     or qcs_pass_plan when the state is larger than a tile).
 * Uniforms come from rand() circuit after circuit, shot after shot, which is
 * the stream qc_sample on each circuit in turn would draw (src/qcs.c:596).
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qcs_internal.h"

/* lowest qubits every streamed tile holds: 2^4 complex128 or 2^5 complex64
 * amplitudes = 256-byte runs in HBM */
#define QCS_BATCH_LOW_BITS_C128 4
#define QCS_BATCH_LOW_BITS_C64 5

struct s_q_batch {
  qcs_hip_batch *dev;
  int options;
  /* per qubit count: tile bits [plain, fused] and fusion width; 0 = default */
  signed char tile_bits[2][QCS_HIP_BATCH_MAX_QUBITS + 1];
  signed char fuse_k[QCS_HIP_BATCH_MAX_QUBITS + 1];
};

t_q_batch *qc_batch_create(int options) {
  t_q_batch *b;
  int flags = 0;
  if (options & ~(QCS_BATCH_GRAPH | QCS_BATCH_KEEP_STATES)) {
    fprintf(stderr, "Error: unknown batch options %d.\n", options);
    return NULL;
  }
  b = (t_q_batch *)calloc(1, sizeof(t_q_batch));
  if (!b) return NULL;
  b->options = options;
  if (options & QCS_BATCH_GRAPH) flags |= QCS_HIP_BATCH_GRAPH;
  if (options & QCS_BATCH_KEEP_STATES) flags |= QCS_HIP_BATCH_KEEP_STATES;
  if (qcs_hip_batch_create(&b->dev, flags) != 0) {
    fprintf(stderr, "Error: could not create a batch: %s\n", qcs_hip_last_error());
    free(b);
    return NULL;
  }
  return b;
}

void qc_batch_destroy(t_q_batch *batch) {
  if (!batch) return;
  qcs_hip_batch_destroy(batch->dev);
  free(batch);
}

int qc_batch_set_plan(t_q_batch *batch, int fused, int num_qubits,
                      int tile_bits, int fuse_k) {
  int maxtb = fused ? QCS_HIP_BATCH_C64_TILE_BITS : QCS_HIP_BATCH_C128_TILE_BITS;
  if (!batch || num_qubits < 1 || num_qubits > QCS_HIP_BATCH_MAX_QUBITS ||
      tile_bits < 0 || tile_bits > maxtb || fuse_k < 0 ||
      fuse_k > QCS_HIP_MAX_FUSE)
    return QCS_HIP_EINVAL;
  batch->tile_bits[fused ? 1 : 0][num_qubits] = (signed char)tile_bits;
  if (fused) batch->fuse_k[num_qubits] = (signed char)fuse_k;
  return 0;
}

int qc_batch_set_threads(t_q_batch *batch, int threads) {
  if (!batch) return QCS_HIP_EINVAL;
  return qcs_hip_batch_set_tuning(batch->dev, QCS_HIP_TUNE_BATCH_THREADS, threads);
}

static int plan_tile_bits(const t_q_batch *b, int fused, int n) {
  int tb = b->tile_bits[fused][n];
  int maxtb = fused ? QCS_HIP_BATCH_C64_TILE_BITS : QCS_HIP_BATCH_C128_TILE_BITS;
  if (tb <= 0) tb = maxtb;
  if (tb > n) tb = n;
  /* a streamed complex64 state needs tiles of >= 5 bits */
  if (fused && tb < n && tb < 5) tb = n < 5 ? n : 5;
  return tb;
}

static int plan_fuse_k(const t_q_batch *b, const t_q_circuit *c, int tb) {
  int k = b->fuse_k[c->num_qubits];
  if (k <= 0) k = (c->options >> 8) & 0xff;
  if (k < 1 || k > QCS_HIP_MAX_FUSE) k = QCS_HIP_MAX_FUSE;
  if (k > tb) k = tb;
  return k;
}

/* growable arrays of the concatenated plans */
typedef struct {
  qcs_hip_batch_circuit *circ;
  qcs_hip_pass *pass;
  qcs_hip_block *block;
  qcs_hip_gate *gate;
  int np, nb, ng, cap_p, cap_b, cap_g;
} batch_plan;

static int reserve(void **p, int *cap, int need, size_t elem) {
  void *q;
  int c;
  if (need <= *cap) return 0;
  c = *cap ? *cap : 64;
  while (c < need) {
    if (c > INT_MAX / 2) return -1;
    c *= 2;
  }
  q = realloc(*p, (size_t)c * elem);
  if (!q) return -1;
  *p = q;
  *cap = c;
  return 0;
}

static int append_circuit(batch_plan *P, const t_q_batch *b, t_q_circuit *c,
                          int i) {
  const int n = c->num_qubits, fused = (c->options & QCS_FUSED_FP32) != 0;
  const int tb = plan_tile_bits(b, fused, n);
  qcs_hip_pass *passes = NULL;
  qcs_hip_block *blocks = NULL;
  qcs_hip_gate *ordered = NULL;
  int np = 0, nb = 0, rc = 0, j, q;
  const qcs_hip_gate *gates = c->pending;
  int ng = c->pending_size;

  if (fused) {
    const int k = plan_fuse_k(b, c, tb);
    if (n <= tb) {
      rc = qcs_fusion_plan(gates, ng, n, k, &blocks, &nb, &ordered);
      if (rc == 0) {
        passes = (qcs_hip_pass *)calloc(1, sizeof(qcs_hip_pass));
        if (!passes) rc = -1;
        else {
          np = 1;
          passes[0].tile_bits = n;
          passes[0].first_block = 0;
          passes[0].num_blocks = nb;
          for (q = 0; q < n; q++) passes[0].bits[q] = q;
        }
      }
    } else {
      /* room for a k-qubit block beside the low qubits */
      int low = QCS_BATCH_LOW_BITS_C64;
      if (low > tb - 2) low = tb - 2;
      rc = qcs_pass_plan(gates, ng, n, tb, low, k, INT_MAX, &passes, &np,
                         &blocks, &nb, &ordered);
    }
    gates = ordered;
  } else {
    /* exact gates (CNOT, X, Z, SWAP, ...) move up to their qubits' previous
       gate, which gives the same amplitudes and far fewer passes (the
       layered circuit at 20 qubits: 2.05 instead of 5.5 per layer) */
    rc = qcs_gate_hoist_exact(gates, ng, n, &ordered);
    if (rc == 0) {
      gates = ordered;
      rc = qcs_gate_pass_plan(gates, ng, n, tb, QCS_BATCH_LOW_BITS_C128,
                              &passes, &np);
    }
  }
  if (rc != 0) goto out;
  rc = -1;
  if (np == 0) { /* no gates: one empty pass still prepares and samples */
    free(passes);
    passes = (qcs_hip_pass *)calloc(1, sizeof(qcs_hip_pass));
    if (!passes) goto out;
    np = 1;
    passes[0].tile_bits = tb;
    for (q = 0; q < tb; q++) passes[0].bits[q] = q;
  }
  if (reserve((void **)&P->pass, &P->cap_p, P->np + np, sizeof(qcs_hip_pass)) ||
      reserve((void **)&P->block, &P->cap_b, P->nb + nb, sizeof(qcs_hip_block)) ||
      reserve((void **)&P->gate, &P->cap_g, P->ng + ng, sizeof(qcs_hip_gate)))
    goto out;
  memset(&P->circ[i], 0, sizeof(P->circ[i]));
  P->circ[i].num_qubits = n;
  P->circ[i].precision = fused ? QCS_HIP_C64 : QCS_HIP_C128;
  P->circ[i].tile_bits = tb;
  P->circ[i].first_pass = P->np;
  P->circ[i].num_passes = np;
  for (j = 0; j < np; j++) {
    qcs_hip_pass *d = &P->pass[P->np + j];
    *d = passes[j];
    d->first_block += fused ? P->nb : P->ng;
  }
  for (j = 0; j < nb; j++) {
    qcs_hip_block *d = &P->block[P->nb + j];
    *d = blocks[j];
    d->first_gate += P->ng;
  }
  if (ng > 0) memcpy(&P->gate[P->ng], gates, (size_t)ng * sizeof(qcs_hip_gate));
  P->np += np;
  P->nb += nb;
  P->ng += ng;
  rc = 0;
out:
  free(passes);
  free(blocks);
  free(ordered);
  return rc;
}

int qc_batch_sample(t_q_batch *batch, t_q_circuit **circuits, int count,
                    int shots, long *indices) {
  batch_plan P;
  double *u = NULL;
  long total, s;
  int i, rc = QCS_HIP_ENOMEM;
  if (!batch || count < 0 || shots < 0 || (count > 0 && !circuits) ||
      (count > 0 && shots > 0 && !indices))
    return QCS_HIP_EINVAL;
  for (i = 0; i < count; i++)
    if (!circuits[i] || !(circuits[i]->options & QCS_DEFERRED) ||
        circuits[i]->num_qubits < 1) {
      fprintf(stderr, "Error: batch circuit %d is not a QCS_DEFERRED circuit.\n", i);
      return QCS_HIP_EINVAL;
    }
  if (count == 0) return 0;
  memset(&P, 0, sizeof(P));
  P.circ = (qcs_hip_batch_circuit *)calloc((size_t)count, sizeof(qcs_hip_batch_circuit));
  total = (long)count * (long)shots;
  if (total > 0) u = (double *)malloc((size_t)total * sizeof(double));
  if (!P.circ || (total > 0 && !u)) goto out;
  for (i = 0; i < count; i++)
    if (append_circuit(&P, batch, circuits[i], i) != 0) {
      fprintf(stderr, "Error: planning batch circuit %d failed.\n", i);
      goto out;
    }
  for (s = 0; s < total; s++) u[s] = rand() / (double)RAND_MAX;
  rc = qcs_hip_batch_run(batch->dev, P.circ, count, P.pass, P.np, P.block, P.nb,
                         P.gate, P.ng, shots, u, indices);
  if (rc != 0)
    fprintf(stderr, "Error: batch run failed: %s\n", qcs_hip_last_error());
out:
  free(P.circ);
  free(P.pass);
  free(P.block);
  free(P.gate);
  free(u);
  return rc;
}

int qc_batch_run_shots(t_q_batch *batch, t_q_circuit **circuits, int count,
                       int shots, int **results) {
  long *idx = NULL;
  int i, s, rc;
  if (!batch || count < 0 || shots < 0 || (count > 0 && (!circuits || !results)))
    return QCS_HIP_EINVAL;
  for (i = 0; i < count; i++)
    if (!results[i]) return QCS_HIP_EINVAL;
  if (count > 0 && shots > 0) {
    idx = (long *)malloc((size_t)count * (size_t)shots * sizeof(long));
    if (!idx) return QCS_HIP_ENOMEM;
  }
  rc = qc_batch_sample(batch, circuits, count, shots, idx);
  if (rc == 0)
    for (i = 0; i < count; i++) {
      long num_states = 1L << circuits[i]->num_qubits;
      memset(results[i], 0, (size_t)num_states * sizeof(int));
      for (s = 0; s < shots; s++) {
        long v = idx[(size_t)i * (size_t)shots + (size_t)s];
        if (v >= 0 && v < num_states) results[i][v]++;
      }
    }
  free(idx);
  return rc;
}

int qc_batch_get_amplitudes(t_q_batch *batch, int circuit, long offset,
                            long count, double *out) {
  if (!batch) return QCS_HIP_EINVAL;
  return qcs_hip_batch_read_state(batch->dev, circuit, offset, count, out);
}

int qc_batch_last_timing(t_q_batch *batch, double *kernel_ms,
                         double *total_ms) {
  int replayed = 0;
  if (!batch) return 0;
  if (qcs_hip_batch_last_timing(batch->dev, kernel_ms, total_ms, &replayed) != 0)
    return 0;
  return replayed;
}
