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
 * A run of the same circuits (same pointers and recorded-gate versions) with
 * the same shot count as the batch's last run reuses that run's device
 * descriptors (qcs_hip_batch_rerun): only the new uniforms travel.
 */
#define _POSIX_C_SOURCE 200112L
#include <limits.h>
#include <time.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

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
  /* the last successful run: its circuits and their versions */
  t_q_circuit **last_circ;
  long *last_version;
  int last_count, last_cap, last_shots, last_valid;
  /* the concatenated plan and the uniforms of a run, kept between runs so
     their pages are not faulted in again every time */
  qcs_hip_pass *cat_pass;
  qcs_hip_block *cat_block;
  qcs_hip_gate *cat_gate;
  double *cat_u;
  size_t cat_pass_cap, cat_block_cap, cat_gate_cap, cat_u_cap;
};

/* *p holds at least `need` elements of `size` bytes afterwards (NULL if out
   of memory, the old buffer freed); its contents are not kept */
static void *reserve(void **p, size_t *cap, size_t need, size_t size) {
  if (need < 1) need = 1;
  if (*p && *cap >= need) return *p;
  free(*p);
  *cap = 0;
  *p = malloc(need * size);
  if (*p) *cap = need;
  return *p;
}

t_q_batch *qc_batch_create(int options) {
  t_q_batch *b;
  int flags = 0;
  if (options & ~(QCS_BATCH_GRAPH | QCS_BATCH_KEEP_STATES | QCS_BATCH_REPLAN)) {
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
  free(batch->last_circ);
  free(batch->last_version);
  free(batch->cat_pass);
  free(batch->cat_block);
  free(batch->cat_gate);
  free(batch->cat_u);
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
  batch->last_valid = 0;
  return 0;
}

int qc_batch_set_threads(t_q_batch *batch, int threads) {
  if (!batch) return QCS_HIP_EINVAL;
  batch->last_valid = 0;
  return qcs_hip_batch_set_tuning(batch->dev, QCS_HIP_TUNE_BATCH_THREADS, threads);
}

/* 1 when (circuits, shots) is the batch's last successful run unchanged */
static int same_as_last(const t_q_batch *b, t_q_circuit **circuits, int count,
                        int shots) {
  int i;
  if (!b->last_valid || (b->options & QCS_BATCH_REPLAN) ||
      count != b->last_count || shots != b->last_shots)
    return 0;
  for (i = 0; i < count; i++)
    if (circuits[i] != b->last_circ[i] ||
        circuits[i]->version != b->last_version[i])
      return 0;
  return 1;
}

static void remember(t_q_batch *b, t_q_circuit **circuits, int count,
                     int shots) {
  int i;
  b->last_valid = 0;
  if (b->options & QCS_BATCH_REPLAN) return;
  if (count > b->last_cap) {
    t_q_circuit **lc = (t_q_circuit **)realloc(b->last_circ,
                                               (size_t)count * sizeof(*lc));
    long *lv;
    if (!lc) return;
    b->last_circ = lc;
    lv = (long *)realloc(b->last_version, (size_t)count * sizeof(*lv));
    if (!lv) return;
    b->last_version = lv;
    b->last_cap = count;
  }
  for (i = 0; i < count; i++) {
    b->last_circ[i] = circuits[i];
    b->last_version[i] = circuits[i]->version;
  }
  b->last_count = count;
  b->last_shots = shots;
  b->last_valid = 1;
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

/* the plan of one circuit, built independently of the others */
typedef struct {
  qcs_hip_pass *passes;
  qcs_hip_block *blocks;
  qcs_hip_gate *ordered; /* NULL: the circuit's recorded gates as they are */
  int np, nb, tb, rc;
} circuit_plan;

static void plan_circuit(const t_q_batch *b, const t_q_circuit *c,
                         circuit_plan *cp) {
  const int n = c->num_qubits, fused = (c->options & QCS_FUSED_FP32) != 0;
  const int tb = plan_tile_bits(b, fused, n);
  const qcs_hip_gate *gates = c->pending;
  const int ng = c->pending_size;
  int rc, q;
  memset(cp, 0, sizeof(*cp));
  cp->tb = tb;
  if (fused) {
    const int k = plan_fuse_k(b, c, tb);
    if (n <= tb) {
      rc = qcs_fusion_plan(gates, ng, n, k, &cp->blocks, &cp->nb, &cp->ordered);
      if (rc == 0) {
        cp->passes = (qcs_hip_pass *)calloc(1, sizeof(qcs_hip_pass));
        if (!cp->passes) rc = -1;
        else {
          cp->np = 1;
          cp->passes[0].tile_bits = n;
          cp->passes[0].num_blocks = cp->nb;
          for (q = 0; q < n; q++) cp->passes[0].bits[q] = q;
        }
      }
    } else {
      /* room for a k-qubit block beside the low qubits */
      int low = QCS_BATCH_LOW_BITS_C64;
      if (low > tb - 2) low = tb - 2;
      rc = qcs_pass_plan(gates, ng, n, tb, low, k, INT_MAX, &cp->passes,
                         &cp->np, &cp->blocks, &cp->nb, &cp->ordered);
    }
  } else {
    /* exact gates (CNOT, X, Z, SWAP, ...) move up to their qubits' previous
       gate, which gives the same amplitudes and far fewer passes (the
       layered circuit at 20 qubits: 2.05 instead of 5.5 per layer) */
    rc = qcs_gate_hoist_exact(gates, ng, n, &cp->ordered);
    if (rc == 0)
      rc = qcs_gate_pass_plan(cp->ordered, ng, n, tb, QCS_BATCH_LOW_BITS_C128,
                              &cp->passes, &cp->np);
  }
  if (rc == 0 && cp->np == 0) {
    /* no gates: one empty pass still prepares and samples the state */
    free(cp->passes);
    cp->passes = (qcs_hip_pass *)calloc(1, sizeof(qcs_hip_pass));
    if (!cp->passes) rc = -1;
    else {
      cp->np = 1;
      cp->passes[0].tile_bits = tb;
      for (q = 0; q < tb; q++) cp->passes[0].bits[q] = q;
    }
  }
  cp->rc = rc;
}

/* QCS_BATCH_PROFILE=1 prints the host phases of every run to stderr */
static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static int profiling(void) {
  const char *e = getenv("QCS_BATCH_PROFILE");
  return e && e[0] && e[0] != '0';
}

static void free_plan(circuit_plan *cp) {
  free(cp->passes);
  free(cp->blocks);
  free(cp->ordered);
}

/* circuits are planned on up to QCS_BATCH_MAX_THREADS host threads
 * (QCS_HOST_THREADS overrides the online core count) */
#define QCS_BATCH_MAX_THREADS 64

typedef struct {
  const t_q_batch *b;
  t_q_circuit **circuits;
  circuit_plan *plans;
  int count, next;
  pthread_mutex_t lock;
} plan_job;

static void *plan_worker(void *arg) {
  plan_job *job = (plan_job *)arg;
  for (;;) {
    int i;
    pthread_mutex_lock(&job->lock);
    i = job->next++;
    pthread_mutex_unlock(&job->lock);
    if (i >= job->count) break;
    plan_circuit(job->b, job->circuits[i], &job->plans[i]);
  }
  return NULL;
}

static int host_threads(int count) {
  long t = sysconf(_SC_NPROCESSORS_ONLN);
  const char *env = getenv("QCS_HOST_THREADS");
  if (env && atoi(env) > 0) t = atoi(env);
  if (t < 1) t = 1;
  if (t > QCS_BATCH_MAX_THREADS) t = QCS_BATCH_MAX_THREADS;
  if (t > count / 16 + 1) t = count / 16 + 1;
  return (int)t;
}

/* the concatenated plan: circuit i's passes / blocks / gates start at
   pass_off[i] / block_off[i] / gate_off[i] */
typedef struct {
  t_q_circuit **circuits;
  const circuit_plan *plans;
  const long *pass_off, *block_off, *gate_off;
  qcs_hip_batch_circuit *circ;
  qcs_hip_pass *pass;
  qcs_hip_block *block;
  qcs_hip_gate *gate;
  int count, parts;
} concat_job;

typedef struct {
  concat_job *job;
  int part;
} concat_arg;

static void *concat_worker(void *arg) {
  const concat_arg *a = (const concat_arg *)arg;
  const concat_job *job = a->job;
  const int i0 = (int)((long)job->count * a->part / job->parts);
  const int i1 = (int)((long)job->count * (a->part + 1) / job->parts);
  int i, j;
  for (i = i0; i < i1; i++) {
    const circuit_plan *cp = &job->plans[i];
    const t_q_circuit *c = job->circuits[i];
    const int fused = (c->options & QCS_FUSED_FP32) != 0;
    const long np = job->pass_off[i], nb = job->block_off[i], ng = job->gate_off[i];
    job->circ[i].num_qubits = c->num_qubits;
    job->circ[i].precision = fused ? QCS_HIP_C64 : QCS_HIP_C128;
    job->circ[i].tile_bits = cp->tb;
    job->circ[i].first_pass = (int)np;
    job->circ[i].num_passes = cp->np;
    for (j = 0; j < cp->np; j++) {
      job->pass[np + j] = cp->passes[j];
      job->pass[np + j].first_block += (int)(fused ? nb : ng);
    }
    for (j = 0; j < cp->nb; j++) {
      job->block[nb + j] = cp->blocks[j];
      job->block[nb + j].first_gate += (int)ng;
    }
    if (c->pending_size > 0)
      memcpy(&job->gate[ng], cp->ordered ? cp->ordered : c->pending,
             (size_t)c->pending_size * sizeof(qcs_hip_gate));
  }
  return NULL;
}

static void concat_all(concat_job *job) {
  pthread_t th[QCS_BATCH_MAX_THREADS];
  concat_arg args[QCS_BATCH_MAX_THREADS];
  int started[QCS_BATCH_MAX_THREADS];
  int nt = host_threads(job->count), t;
  job->parts = nt;
  for (t = 0; t < nt; t++) {
    args[t].job = job;
    args[t].part = t;
  }
  for (t = 1; t < nt; t++)
    started[t] = pthread_create(&th[t], NULL, concat_worker, &args[t]) == 0;
  concat_worker(&args[0]);
  for (t = 1; t < nt; t++) {
    if (started[t]) pthread_join(th[t], NULL);
    else concat_worker(&args[t]); /* no thread: do that part here */
  }
}

/* the shots' uniforms, rand() / RAND_MAX in the order qc_sample on each
   circuit in turn would draw them; drawn on a thread of their own while the
   circuits are planned (the planners call no rand()) */
typedef struct {
  double *u;
  long total;
} batch_uniform_job;

static void *draw_batch_uniforms(void *arg) {
  const batch_uniform_job *job = (const batch_uniform_job *)arg;
  long s;
  for (s = 0; s < job->total; s++) job->u[s] = rand() / (double)RAND_MAX;
  return NULL;
}

static void plan_all(const t_q_batch *b, t_q_circuit **circuits, int count,
                     circuit_plan *plans) {
  plan_job job;
  pthread_t th[QCS_BATCH_MAX_THREADS];
  int nt = host_threads(count), started = 0, i;
  job.b = b;
  job.circuits = circuits;
  job.plans = plans;
  job.count = count;
  job.next = 0;
  pthread_mutex_init(&job.lock, NULL);
  for (i = 1; i < nt; i++)
    if (pthread_create(&th[started], NULL, plan_worker, &job) == 0) started++;
  plan_worker(&job); /* the calling thread works too */
  for (i = 0; i < started; i++) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&job.lock);
}

int qc_batch_sample(t_q_batch *batch, t_q_circuit **circuits, int count,
                    int shots, long *indices) {
  circuit_plan *plans = NULL;
  qcs_hip_batch_circuit *circ = NULL;
  qcs_hip_pass *pass = NULL;
  qcs_hip_block *block = NULL;
  qcs_hip_gate *gate = NULL;
  long *offs = NULL;
  concat_job job;
  double *u = NULL;
  long total, s, np = 0, nb = 0, ng = 0;
  int i, rc = QCS_HIP_ENOMEM, drawing = 0;
  double t0, t1, t2, t3;
  batch_uniform_job uj;
  pthread_t uth;
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
  total = (long)count * (long)shots;
  if (same_as_last(batch, circuits, count, shots)) {
    if (total > 0) {
      u = (double *)reserve((void **)&batch->cat_u, &batch->cat_u_cap, (size_t)total,
                            sizeof(double));
      if (!u) return QCS_HIP_ENOMEM;
    }
    for (s = 0; s < total; s++) u[s] = rand() / (double)RAND_MAX;
    t0 = now_ms();
    rc = qcs_hip_batch_rerun(batch->dev, shots, u, indices);
    if (profiling())
      fprintf(stderr, "qc_batch_sample: rerun of the last plan %.1f ms\n",
              now_ms() - t0);
    if (rc != 0) {
      batch->last_valid = 0;
      fprintf(stderr, "Error: batch run failed: %s\n", qcs_hip_last_error());
    }
    return rc;
  }
  batch->last_valid = 0;
  plans = (circuit_plan *)calloc((size_t)count, sizeof(circuit_plan));
  circ = (qcs_hip_batch_circuit *)calloc((size_t)count, sizeof(qcs_hip_batch_circuit));
  offs = (long *)malloc(3 * (size_t)count * sizeof(long));
  if (total > 0)
    u = (double *)reserve((void **)&batch->cat_u, &batch->cat_u_cap, (size_t)total,
                          sizeof(double));
  if (!plans || !circ || !offs || (total > 0 && !u)) goto out;
  t0 = now_ms();
  uj.u = u;
  uj.total = total;
  drawing = total > 0 && pthread_create(&uth, NULL, draw_batch_uniforms, &uj) == 0;
  plan_all(batch, circuits, count, plans);
  t1 = now_ms();
  for (i = 0; i < count; i++) {
    if (plans[i].rc != 0) {
      fprintf(stderr, "Error: planning batch circuit %d failed.\n", i);
      goto out;
    }
    offs[i] = np;
    offs[count + i] = nb;
    offs[2 * count + i] = ng;
    np += plans[i].np;
    nb += plans[i].nb;
    ng += circuits[i]->pending_size;
  }
  if (np > INT_MAX || nb > INT_MAX || ng > INT_MAX) {
    rc = QCS_HIP_EINVAL;
    goto out;
  }
  pass = (qcs_hip_pass *)reserve((void **)&batch->cat_pass, &batch->cat_pass_cap,
                                 (size_t)np, sizeof(qcs_hip_pass));
  block = (qcs_hip_block *)reserve((void **)&batch->cat_block, &batch->cat_block_cap,
                                   (size_t)nb, sizeof(qcs_hip_block));
  gate = (qcs_hip_gate *)reserve((void **)&batch->cat_gate, &batch->cat_gate_cap,
                                 (size_t)ng, sizeof(qcs_hip_gate));
  if (!pass || !block || !gate) goto out;
  /* concatenate on host threads: circuit i's ranges start at its offsets */
  job.circuits = circuits;
  job.plans = plans;
  job.pass_off = offs;
  job.block_off = offs + count;
  job.gate_off = offs + 2 * (size_t)count;
  job.circ = circ;
  job.pass = pass;
  job.block = block;
  job.gate = gate;
  job.count = count;
  concat_all(&job);
  t2 = now_ms();
  if (drawing) {
    pthread_join(uth, NULL);
    drawing = 0;
  } else {
    draw_batch_uniforms(&uj);
  }
  t3 = now_ms();
  rc = qcs_hip_batch_run(batch->dev, circ, count, pass, (int)np, block, (int)nb,
                         gate, (int)ng, shots, u, indices);
  if (profiling())
    fprintf(stderr, "qc_batch_sample: plan %.1f ms (%d threads), concatenate %.1f ms, "
                    "uniforms (the rest of their draw) %.1f ms, run %.1f ms\n", t1 - t0,
            host_threads(count),
            t2 - t1, t3 - t2, now_ms() - t3);
  if (rc != 0)
    fprintf(stderr, "Error: batch run failed: %s\n", qcs_hip_last_error());
  else
    remember(batch, circuits, count, shots);
out:
  if (drawing) pthread_join(uth, NULL);
  if (plans)
    for (i = 0; i < count; i++) free_plan(&plans[i]);
  free(plans);
  free(offs);
  free(circ);
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
