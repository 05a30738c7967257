/*
 * tune.c -- autotuner of the fused and batched forms, with a result cache.
 *
 * Fused form (qc_autotune), one pass per block: for every fusion width k
 * the one-block-per-pass kernel has launch variants (tile size, threads,
 * pipelining); each is timed on the circuit's own state with identity blocks
 * (an identity block leaves every amplitude unchanged: products by 1 and by
 * 0 are exact), the fastest is kept per (device, n, k), and the fusion width
 * for the recorded gates is the k that minimises (blocks the planner makes
 * at width k) x (time of one pass at width k).
 *
 * Fused form as tile programs (QCS_FUSE_TILES, QCS_FUSE_SPLIT_BF16): the
 * circuit's own pass plan is timed, with identity blocks, at each fusion
 * width, work-group size and grid; the cost is the sum of its pass times.
 *
 * Batched form (qc_batch_autotune): for every (precision, n) whose states
 * are streamed through LDS, candidate tile sizes (and fusion widths for
 * complex64) are timed on a sample of the batch's own circuits (zero shots:
 * no uniforms are drawn, so rand() is untouched) and the fastest becomes
 * the batch's plan for that (precision, n); the work-group size is chosen
 * the same way.
 *
 * Results are cached per device ("<arch>:<compute units>") in a text file:
 * $QCS_TUNE_CACHE (a path; "off" disables the cache), else
 * $XDG_CACHE_HOME/qcs_amd/tune.txt, else $HOME/.cache/qcs_amd/tune.txt.
 * Lines are
 *   <device> fused <n> <k> <variant> <microseconds per pass>
 *   <device> tiles[split] <n> <structure hash> <k> <threads> <grid> <k4 form> <ms>
 *   <device> batch <precision> <n> <tile bits> <k> <milliseconds>
 *   <device> batchthreads <threads> <milliseconds>
 * and the last matching line wins.
 */
#define _POSIX_C_SOURCE 200112L
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "qcs_internal.h"

#define TUNE_LINE 256

static int cache_path(char *buf, size_t len) {
  const char *p = getenv("QCS_TUNE_CACHE");
  const char *base;
  if (p && *p) {
    if (strcmp(p, "off") == 0) return -1;
    if (strlen(p) + 1 > len) return -1;
    strcpy(buf, p);
    return 0;
  }
  base = getenv("XDG_CACHE_HOME");
  if (base && *base) {
    if ((size_t)snprintf(buf, len, "%s/qcs_amd/tune.txt", base) >= len) return -1;
    return 0;
  }
  base = getenv("HOME");
  if (!base || !*base) return -1;
  if ((size_t)snprintf(buf, len, "%s/.cache/qcs_amd/tune.txt", base) >= len) return -1;
  return 0;
}

/* mkdir -p of the directory part of `path` */
static void make_parent(const char *path) {
  char tmp[1024];
  size_t i, n = strlen(path);
  if (n >= sizeof(tmp)) return;
  strcpy(tmp, path);
  for (i = 1; i < n; i++)
    if (tmp[i] == '/') {
      tmp[i] = '\0';
      if (mkdir(tmp, 0755) != 0 && errno != EEXIST) return;
      tmp[i] = '/';
    }
}

static void cache_append(const char *line) {
  char path[1024];
  FILE *f;
  if (cache_path(path, sizeof(path)) != 0) return;
  make_parent(path);
  f = fopen(path, "a");
  if (!f) return;
  fputs(line, f);
  fclose(f);
}

/* last line with these leading fields; returns 1 and fills `rest` */
static int cache_find(const char *prefix, char *rest, size_t len) {
  char path[1024], line[TUNE_LINE];
  size_t np = strlen(prefix);
  int found = 0;
  FILE *f;
  if (cache_path(path, sizeof(path)) != 0) return 0;
  f = fopen(path, "r");
  if (!f) return 0;
  while (fgets(line, sizeof(line), f))
    if (strncmp(line, prefix, np) == 0 && line[np] == ' ') {
      size_t m = strlen(line + np + 1);
      if (m + 1 > len) continue;
      strcpy(rest, line + np + 1);
      found = 1;
    }
  fclose(f);
  return found;
}

static int device_key(char *buf, int len) {
  return qcs_hip_device_name(buf, len);
}

/* ---------------------------------------------------------------------- */
/* fused form                                                               */
/* ---------------------------------------------------------------------- */

/* mean device time (ms) of one identity k-block pass with `variant` */
static int time_variant(qcs_hip_state *s, int n, int k, int variant,
                        double *ms_out) {
  double u[2 * 32 * 32];
  int targets[5], w, rep, q, rc;
  int R = 1 << k;
  double ms = 0.0, bytes = 0.0, flops = 0.0;
  long launches = 0;
  memset(u, 0, sizeof(u));
  for (q = 0; q < R; q++) u[2 * (q * R + q)] = 1.0;
  rc = qcs_hip_set_tuning(s, QCS_HIP_TUNE_BLOCK_VARIANT(k), variant);
  if (rc) return rc;
  /* one untimed warm-up pass, then three target windows twice each */
  for (q = 0; q < k; q++) targets[q] = q;
  if ((rc = qcs_hip_apply_unitary(s, k, targets, u))) return rc;
  if ((rc = qcs_hip_timing_enable(s, 1))) return rc;
  for (rep = 0; rep < 2; rep++)
    for (w = 0; w < 3; w++) {
      int lo = w == 0 ? 0 : (w == 1 ? (n - k) / 2 : n - k);
      for (q = 0; q < k; q++) targets[q] = lo + q;
      if ((rc = qcs_hip_apply_unitary(s, k, targets, u))) return rc;
    }
  rc = qcs_hip_timing_read(s, QCS_HIP_K_FUSED, &ms, &launches, &bytes, &flops);
  qcs_hip_timing_enable(s, 0);
  if (rc) return rc;
  *ms_out = launches > 0 ? ms / (double)launches : 1e30;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* tile programs                                                            */
/* ---------------------------------------------------------------------- */

/* FNV-1a over the kind and qubits of the recorded gates (not the angles):
 * circuits with the same structure get the same pass plans */
static unsigned long structure_hash(const t_q_circuit *c) {
  unsigned long h = 1469598103934665603UL;
  int i, j;
  for (i = 0; i < c->pending_size; i++) {
    int v[3];
    v[0] = c->pending[i].kind;
    v[1] = c->pending[i].target;
    v[2] = c->pending[i].control;
    for (j = 0; j < 3; j++) {
      h ^= (unsigned long)(unsigned)v[j];
      h *= 1099511628211UL;
    }
  }
  return h;
}

/* device time (ms) of the circuit's own tile-program plan at fusion width k
 * under the state's current launch shape: the sum over its passes, every
 * block composed from no gates (the identity, so the state is unchanged) */
static int time_tile_plan(t_q_circuit *c, int k, int max_blocks, double *ms_out) {
  qcs_hip_block *blocks = NULL;
  qcs_hip_gate *ordered = NULL;
  qcs_hip_pass *passes = NULL;
  int nb = 0, np = 0, b, rc;
  double ms = 0.0, bytes = 0.0, flops = 0.0;
  long launches = 0;
  if (qcs_pass_plan(c->pending, c->pending_size, c->num_qubits, QCS_HIP_TILE_BITS,
                    QCS_HIP_TILE_LOW_BITS, k, max_blocks, &passes, &np, &blocks, &nb,
                    &ordered) != 0)
    return QCS_HIP_ENOMEM;
  for (b = 0; b < nb; b++) {
    blocks[b].first_gate = 0;
    blocks[b].num_gates = 0;
  }
  rc = qcs_hip_timing_enable(c->state, 1);
  if (!rc) rc = qcs_hip_apply_passes(c->state, passes, np, blocks, nb, ordered, c->pending_size);
  if (!rc)
    rc = qcs_hip_timing_read(c->state, QCS_HIP_K_FUSED, &ms, &launches, &bytes, &flops);
  qcs_hip_timing_enable(c->state, 0);
  free(passes);
  free(blocks);
  free(ordered);
  *ms_out = ms;
  return rc;
}

/* Tile-program circuits: fusion width k in {3, 4}, work-group size 512 or
 * 1024, a persistent grid of one or two work-groups per CU and (FP32 math on
 * 1024 threads) the form of k = 4 blocks, timed on the
 * circuit's own pass plan (the sum of its pass times, not blocks x a block's
 * time: a pass holds several blocks, and what a block costs depends on its
 * width and on the pass around it).  Cached per (device, n, arithmetic,
 * gate structure). */
static int set_tile_shape(qcs_hip_state *s, int threads, int grid, int form) {
  int rc;
  if ((rc = qcs_hip_set_tuning(s, QCS_HIP_TUNE_TILE_THREADS, threads)) ||
      (rc = qcs_hip_set_tuning(s, QCS_HIP_TUNE_TILE_GRID, grid)) ||
      (rc = qcs_hip_set_tuning(s, QCS_HIP_TUNE_TILE_K4_FORM, form)))
    return rc;
  return 0;
}

static int autotune_tiles(t_q_circuit *c, const char *dev) {
  char key[200], rest[TUNE_LINE], line[TUNE_LINE];
  const int split = (c->options & QCS_FUSE_SPLIT_BF16) != 0;
  const int max_blocks = split ? QCS_HIP_TILE_MAX_BLOCKS_SPLIT : QCS_HIP_TILE_MAX_BLOCKS;
  const char *colon = strchr(dev, ':');
  const int cus = colon ? atoi(colon + 1) : 256;
  int k, th, g, form, best_k = QCS_HIP_TILE_KMAX, best_th = 1024, best_grid = cus;
  int best_form = QCS_HIP_TILE_K4_CHAIN, rc;
  double best = 1e30, ms;
  if (c->pending_size == 0) {
    c->tuned_k = best_k;
    return best_k;
  }
  sprintf(key, "%s tiles%s %d %lx", dev, split ? "split" : "", c->num_qubits,
          structure_hash(c));
  if (cache_find(key, rest, sizeof(rest)) &&
      sscanf(rest, "%d %d %d %d", &best_k, &best_th, &best_grid, &best_form) == 4 &&
      best_k >= 1 && best_k <= QCS_HIP_TILE_KMAX &&
      set_tile_shape(c->state, best_th, best_grid, best_form) == 0) {
    c->tuned_k = best_k;
    return best_k;
  }
  /* one untimed run loads the kernels and warms the caches */
  if ((rc = time_tile_plan(c, QCS_HIP_TILE_KMAX, max_blocks, &ms))) return rc;
  for (th = 512; th <= 1024; th *= 2)
    for (g = 1; g <= 2; g++)
      for (k = 3; k <= QCS_HIP_TILE_KMAX; k++)
        /* the pipelined k = 4 forms exist for FP32 math on 1024 threads */
        for (form = QCS_HIP_TILE_K4_CHAIN;
             form <= ((th == 1024 && !split) ? QCS_HIP_TILE_K4_GAUSS : QCS_HIP_TILE_K4_CHAIN);
             form++) {
          if ((rc = set_tile_shape(c->state, th, g * cus, form))) return rc;
          if ((rc = time_tile_plan(c, k, max_blocks, &ms))) return rc;
          if (ms < best) {
            best = ms;
            best_k = k;
            best_th = th;
            best_grid = g * cus;
            best_form = form;
          }
        }
  if ((rc = set_tile_shape(c->state, best_th, best_grid, best_form))) return rc;
  sprintf(line, "%s %d %d %d %d %.3f\n", key, best_k, best_th, best_grid, best_form, best);
  cache_append(line);
  c->tuned_k = best_k;
  return best_k;
}

int qc_autotune(t_q_circuit *c) {
  char dev[96], key[160], rest[TUNE_LINE];
  double t_pass[QCS_HIP_MAX_FUSE + 1];
  int n, k, kmax, best_k = 0, rc;
  double best_cost = 0.0;
  if (!c || !c->state || !(c->options & QCS_FUSED_FP32)) return QCS_HIP_EINVAL;
  n = c->num_qubits;
  if ((c->options & (QCS_FUSE_TILES | QCS_FUSE_SPLIT_BF16)) &&
      n >= QCS_HIP_TILE_BITS + 3) {
    if ((rc = device_key(dev, sizeof(dev)))) return rc;
    return autotune_tiles(c, dev);
  }
  kmax = n < QCS_HIP_MAX_FUSE ? n : QCS_HIP_MAX_FUSE;
  if ((c->options & (QCS_FUSE_TILES | QCS_FUSE_SPLIT_BF16)) && kmax > QCS_HIP_TILE_KMAX)
    kmax = QCS_HIP_TILE_KMAX;
  if ((rc = device_key(dev, sizeof(dev)))) return rc;
  for (k = 1; k <= kmax; k++) {
    int variant = QCS_HIP_BLOCK_TB13_512_PIPE, v;
    double us = 0.0;
    sprintf(key, "%s fused %d %d", dev, n, k);
    if (cache_find(key, rest, sizeof(rest)) &&
        sscanf(rest, "%d %lf", &variant, &us) == 2 &&
        variant >= QCS_HIP_BLOCK_TB12_256 && variant <= QCS_HIP_BLOCK_TB13_512_PIPE) {
      t_pass[k] = us * 1e-3;
    } else {
      double best = 1e30;
      for (v = QCS_HIP_BLOCK_TB12_256; v <= QCS_HIP_BLOCK_TB13_512_PIPE; v++) {
        double ms;
        if ((rc = time_variant(c->state, n, k, v, &ms))) return rc;
        if (ms < best) {
          best = ms;
          variant = v;
        }
      }
      t_pass[k] = best;
      {
        char line[TUNE_LINE];
        sprintf(line, "%s %d %.3f\n", key, variant, best * 1e3);
        cache_append(line);
      }
    }
    if ((rc = qcs_hip_set_tuning(c->state, QCS_HIP_TUNE_BLOCK_VARIANT(k), variant)))
      return rc;
  }
  /* fusion width for the recorded gates: blocks x time per pass */
  for (k = 1; k <= kmax; k++) {
    qcs_hip_block *blocks = NULL;
    qcs_hip_gate *ordered = NULL;
    int nb = 0;
    double cost;
    if (c->pending_size == 0) break;
    if (qcs_fusion_plan(c->pending, c->pending_size, n, k, &blocks, &nb,
                        &ordered) != 0)
      return QCS_HIP_ENOMEM;
    free(blocks);
    free(ordered);
    cost = nb * t_pass[k];
    if (best_k == 0 || cost < best_cost) {
      best_k = k;
      best_cost = cost;
    }
  }
  if (best_k == 0) {
    /* nothing recorded: the width with the cheapest gate-equivalent pass */
    best_k = kmax;
  }
  c->tuned_k = best_k;
  return best_k;
}

/* ---------------------------------------------------------------------- */
/* batched form                                                             */
/* ---------------------------------------------------------------------- */

static double batch_time(t_q_batch *b, t_q_circuit **cs, int count) {
  double kernel_ms = -1.0, total_ms = -1.0;
  if (qc_batch_sample(b, cs, count, 0, NULL) != 0) return 1e30;
  qc_batch_last_timing(b, &kernel_ms, &total_ms);
  return kernel_ms > 0.0 ? kernel_ms : total_ms;
}

#define TUNE_SAMPLE 4

int qc_batch_autotune(t_q_batch *b, t_q_circuit **circuits, int count) {
  char dev[96], key[160], rest[TUNE_LINE], line[TUNE_LINE];
  t_q_circuit *sample[64];
  int fused, n, i, rc;
  if (!b || count < 0 || (count > 0 && !circuits)) return QCS_HIP_EINVAL;
  for (i = 0; i < count; i++)
    if (!circuits[i] || !(circuits[i]->options & QCS_DEFERRED)) return QCS_HIP_EINVAL;
  if ((rc = device_key(dev, sizeof(dev)))) return rc;
  for (fused = 0; fused <= 1; fused++) {
    const int maxtb = fused ? QCS_HIP_BATCH_C64_TILE_BITS : QCS_HIP_BATCH_C128_TILE_BITS;
    for (n = maxtb + 1; n <= QCS_HIP_BATCH_MAX_QUBITS; n++) {
      int ns = 0, total = 0, tb, k, best_tb = 0, best_k = 0, tb_lo;
      double best = 1e30;
      for (i = 0; i < count; i++)
        if (circuits[i]->num_qubits == n &&
            ((circuits[i]->options & QCS_FUSED_FP32) != 0) == fused)
          total++;
      if (total == 0) continue;
      sprintf(key, "%s batch %d %d", dev, fused, n);
      if (cache_find(key, rest, sizeof(rest)) &&
          sscanf(rest, "%d %d", &best_tb, &best_k) == 2) {
        qc_batch_set_plan(b, fused, n, best_tb, best_k);
        continue;
      }
      /* an evenly spread sample of this group */
      for (i = 0; i < count && ns < TUNE_SAMPLE; i++)
        if (circuits[i]->num_qubits == n &&
            ((circuits[i]->options & QCS_FUSED_FP32) != 0) == fused) {
          sample[ns++] = circuits[i];
          i += total / TUNE_SAMPLE;
        }
      tb_lo = maxtb - (fused ? 2 : 1);
      for (tb = tb_lo; tb <= maxtb; tb++)
        for (k = fused ? 3 : 0; k <= (fused ? QCS_HIP_MAX_FUSE : 0); k++) {
          double t;
          if (qc_batch_set_plan(b, fused, n, tb, k) != 0) continue;
          t = batch_time(b, sample, ns);
          if (t < best) {
            best = t;
            best_tb = tb;
            best_k = k;
          }
        }
      if (best_tb == 0) continue;
      qc_batch_set_plan(b, fused, n, best_tb, best_k);
      sprintf(line, "%s %d %d %.3f\n", key, best_tb, best_k, best);
      cache_append(line);
    }
  }
  /* work-group size on a sample of the whole batch */
  sprintf(key, "%s batchthreads", dev);
  {
    int threads = 0;
    if (cache_find(key, rest, sizeof(rest)) && sscanf(rest, "%d", &threads) == 1 &&
        qc_batch_set_threads(b, threads) == 0)
      return 0;
  }
  if (count > 0) {
    int ns = 0, th, best_th = 512;
    double best = 1e30;
    for (i = 0; i < count && ns < 64; i += count / 64 + 1) sample[ns++] = circuits[i];
    for (th = 256; th <= 512; th *= 2) {
      double t;
      if (qc_batch_set_threads(b, th) != 0) continue;
      t = batch_time(b, sample, ns);
      if (t < best) {
        best = t;
        best_th = th;
      }
    }
    qc_batch_set_threads(b, best_th);
    sprintf(line, "%s %d %.3f\n", key, best_th, best);
    cache_append(line);
  }
  return 0;
}
