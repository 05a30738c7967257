/*
 * qcs_oracle.c -- CPU FP64 oracle for the state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker or as the
 * timed CPU port; the product path (qcs_amd/) never calls it.
 *
 * A restatement of the reference's sequential FP64 path (SegusFaultise/qcs,
 * built without a PARALLEL_MODE), written from its algorithm, with two
 * semantics for the parts where the reference is not the textbook map:
 *   compat = 1  the reference's sequential maps as written,
 *   compat = 0  textbook controlled-U and Grover.
 * The state keeps the reference's two buffers (vector + scratch_vector,
 * src/q_state.c:49-66) and pointer swaps, so compat Grover sees the same
 * stale buffer as the reference.  Complex arithmetic follows
 * src/complex.c:23-57 in the same association (c_add(c_mul(a,b),c_mul(c,d)),
 * re = ar*br - ai*bi, im = ar*bi + ai*br); built with gcc -std=c89, which does
 * not contract into FMA, exactly like the reference's own build.
 *
 * Also holds the OpenMP restatement of the reference's CPU-parallel loops
 * for the CPU baseline: the 1q loop (src/q_gates.c:40-68), where one variant
 * keeps the shared inner index of the original (a data race) and one makes it
 * private, and the 2q loop (src/q_gates.c:174-208), where the original's
 * (c=1, t=1) iteration rewrites the partner that the (c=1, t=0) iteration
 * wrote (a write-write race between threads) and the fixed variants write
 * each entry from one iteration only.  The race-fixed variants equal the
 * sequential functions above bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "qcs_oracle.h"

#define ORACLE_PI 3.14159265358979323846

/* ---- complex helpers: src/complex.c:23-57,89-99 ---------------------- */

typedef struct {
  double re, im;
} ocx;

static ocx cx_add(ocx a, ocx b) {
  ocx r;
  r.re = a.re + b.re;
  r.im = a.im + b.im;
  return r;
}

static ocx cx_mul(ocx a, ocx b) {
  ocx r;
  r.re = (a.re * b.re) - (a.im * b.im);
  r.im = (a.re * b.im) + (a.im * b.re);
  return r;
}

static double cx_norm_sq(ocx a) { return (a.re * a.re) + (a.im * a.im); }

static ocx mat(const double m[8], int k) {
  ocx r;
  r.re = m[2 * k];
  r.im = m[2 * k + 1];
  return r;
}

/* ---- state: src/q_state.c:32-115 --------------------------------------- */

struct oq_state {
  int n;
  long size;
  int compat;
  ocx *vec;
  ocx *scr;
};

oq_state *oq_create(int n, int compat) {
  oq_state *s;
  long i;
  if (n <= 0 || n > 30) return NULL;
  s = (oq_state *)malloc(sizeof(oq_state));
  if (!s) return NULL;
  s->n = n;
  s->size = 1L << n;
  s->compat = compat;
  s->vec = (ocx *)malloc((size_t)s->size * sizeof(ocx));
  s->scr = (ocx *)malloc((size_t)s->size * sizeof(ocx));
  if (!s->vec || !s->scr) {
    oq_destroy(s);
    return NULL;
  }
  for (i = 0; i < s->size; i++) {
    s->vec[i].re = s->vec[i].im = 0.0;
    s->scr[i].re = s->scr[i].im = 0.0;
  }
  s->vec[0].re = 1.0;
  return s;
}

void oq_destroy(oq_state *s) {
  if (!s) return;
  free(s->vec);
  free(s->scr);
  free(s);
}

long oq_size(const oq_state *s) { return s->size; }
double *oq_vector(oq_state *s) { return (double *)s->vec; }
double *oq_scratch(oq_state *s) { return (double *)s->scr; }

static void swap_buffers(oq_state *s) {
  ocx *t = s->vec;
  s->vec = s->scr;
  s->scr = t;
}

/* ---- gates -------------------------------------------------------------- */

/* src/q_gates.c:25-36,130-148: memcpy, block/step pair loop, swap */
void oq_apply_1q(oq_state *s, int t, const double m[8]) {
  long step, block, i, j;
  ocx g0, g1, g2, g3;
  if (t < 0 || t >= s->n) return;
  step = 1L << t;
  block = 1L << (t + 1);
  g0 = mat(m, 0);
  g1 = mat(m, 1);
  g2 = mat(m, 2);
  g3 = mat(m, 3);
  memcpy(s->scr, s->vec, (size_t)s->size * sizeof(ocx));
  for (i = 0; i < s->size; i += block) {
    for (j = i; j < i + step; j++) {
      ocx v0 = s->vec[j], v1 = s->vec[j + step];
      s->scr[j] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
      s->scr[j + step] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
    }
  }
  swap_buffers(s);
}

/* src/q_gates.c:158-170,276-297.  compat: the loop as written, whose else
 * branch at i|t overwrites the partner written at i, leaving out[i1] = v1.
 * textbook: out[i1] = m10*v0 + m11*v1. */
void oq_apply_cu(oq_state *s, int c, int t, const double m[8]) {
  long cb, tb, i;
  ocx g0, g1, g2, g3;
  if (c < 0 || t < 0 || c >= s->n || t >= s->n || c == t) return;
  cb = 1L << c;
  tb = 1L << t;
  g0 = mat(m, 0);
  g1 = mat(m, 1);
  g2 = mat(m, 2);
  g3 = mat(m, 3);
  memcpy(s->scr, s->vec, (size_t)s->size * sizeof(ocx));
  if (s->compat) {
    for (i = 0; i < s->size; i++) {
      if ((i & cb) != 0 && (i & tb) == 0) {
        ocx v0 = s->vec[i], v1 = s->vec[i | tb];
        s->scr[i] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
        s->scr[i | tb] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
      } else {
        s->scr[i] = s->vec[i];
      }
    }
  } else {
    for (i = 0; i < s->size; i++) {
      if ((i & cb) != 0 && (i & tb) == 0) {
        ocx v0 = s->vec[i], v1 = s->vec[i | tb];
        s->scr[i] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
        s->scr[i | tb] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
      }
    }
  }
  swap_buffers(s);
}

/* SWAP of qubits a, b (no reference counterpart): exchange (1,0) <-> (0,1) */
void oq_apply_swap(oq_state *s, int a, int b) {
  long ab, bb, i;
  if (a < 0 || b < 0 || a >= s->n || b >= s->n || a == b) return;
  ab = 1L << a;
  bb = 1L << b;
  memcpy(s->scr, s->vec, (size_t)s->size * sizeof(ocx));
  for (i = 0; i < s->size; i++)
    if ((i & ab) && !(i & bb)) {
      long o = (i ^ ab) | bb;
      s->scr[i] = s->vec[o];
      s->scr[o] = s->vec[i];
    }
  swap_buffers(s);
}

/* ---- Grover: src/q_gates.c:305-356 --------------------------------------- */

void oq_phase_flip(oq_state *s, long index) {
  if (index < 0 || index >= s->size) return;
  if (s->compat) {
    /* swap, then vector[index] = -scratch[index]; the rest of the new
       vector is the previous scratch buffer */
    swap_buffers(s);
    s->vec[index].re = -s->scr[index].re;
    s->vec[index].im = -s->scr[index].im;
  } else {
    s->vec[index].re = -s->vec[index].re;
    s->vec[index].im = -s->vec[index].im;
  }
}

void oq_diffusion(oq_state *s) {
  long i;
  ocx mean;
  mean.re = mean.im = 0.0;
  for (i = 0; i < s->size; i++) mean = cx_add(mean, s->vec[i]);
  if (s->compat) {
    /* mean normalised to unit magnitude, 2*mean - old */
    double mag = sqrt(cx_norm_sq(mean));
    ocx two;
    if (mag > 1e-10) {
      mean.re /= mag;
      mean.im /= mag;
    }
    two.re = 2.0 * mean.re;
    two.im = 2.0 * mean.im;
    swap_buffers(s);
    for (i = 0; i < s->size; i++) {
      s->vec[i].re = two.re - s->scr[i].re;
      s->vec[i].im = two.im - s->scr[i].im;
    }
  } else {
    /* textbook inversion about the mean: 2*(sum/N) - a */
    double N = (double)s->size;
    double tr = 2.0 * (mean.re / N), ti = 2.0 * (mean.im / N);
    for (i = 0; i < s->size; i++) {
      s->vec[i].re = tr - s->vec[i].re;
      s->vec[i].im = ti - s->vec[i].im;
    }
  }
}

/* q_grover_iterations, src/q_utils.c:126-134 (textbook: exact 2^n) */
long oq_grover_iterations(int n, int compat) {
  double N;
  if (compat && n >= 31) {
    if (n == 31) return 0;
    N = (double)(1 << (n & 31));
  } else {
    N = ldexp(1.0, n);
  }
  return (long)floor((ORACLE_PI / 4.0) * sqrt(N));
}

/* ---- normalise / measure / observe --------------------------------------- */

/* q_state_normalize, src/q_utils.c:102-118 */
double oq_normalize(oq_state *s) {
  long i;
  double ns = 0.0;
  for (i = 0; i < s->size; i++) ns += cx_norm_sq(s->vec[i]);
  if (ns > 1e-12 && ns != 1.0) {
    double inv = 1.0 / sqrt(ns);
    for (i = 0; i < s->size; i++) {
      s->vec[i].re *= inv;
      s->vec[i].im *= inv;
    }
  }
  return ns;
}

double oq_norm_sq(const oq_state *s) {
  long i;
  double ns = 0.0;
  for (i = 0; i < s->size; i++) ns += cx_norm_sq(s->vec[i]);
  return ns;
}

/* qc_measure, src/qcs.c:250-283, with u = rand()/RAND_MAX supplied */
int oq_measure(oq_state *s, int q, double u) {
  long i, qb;
  double p0 = 0.0;
  if (q < 0 || q >= s->n) return 0;
  qb = 1L << q;
  for (i = 0; i < s->size; i++)
    if ((i & qb) == 0) p0 += cx_norm_sq(s->vec[i]);
  if (u <= p0) {
    for (i = 0; i < s->size; i++)
      if ((i & qb) != 0) s->vec[i].re = s->vec[i].im = 0.0;
    oq_normalize(s);
    return 0;
  }
  for (i = 0; i < s->size; i++)
    if ((i & qb) == 0) s->vec[i].re = s->vec[i].im = 0.0;
  oq_normalize(s);
  return 1;
}

/* qc_run_shots' sampling loop, src/qcs.c:589-604: first i with u < running
 * sum of |a|^2, -1 when none */
void oq_sample(const oq_state *s, int shots, const double *u, long *out) {
  int k;
  long i;
  double *p = (double *)malloc((size_t)s->size * sizeof(double));
  if (!p) return;
  for (i = 0; i < s->size; i++) p[i] = cx_norm_sq(s->vec[i]);
  for (k = 0; k < shots; k++) {
    double cum = 0.0;
    out[k] = -1;
    for (i = 0; i < s->size; i++) {
      cum += p[i];
      if (u[k] < cum) {
        out[k] = i;
        break;
      }
    }
  }
  free(p);
}

/* qc_get_probability, src/qcs.c:391-395 */
double oq_probability(const oq_state *s, long i) {
  if (i < 0 || i >= s->size) return 0.0;
  return cx_norm_sq(s->vec[i]);
}

/* qc_find_most_likely_state, src/qcs.c:464-478 */
long oq_argmax(const oq_state *s) {
  long i, best = 0;
  double bp = 0.0;
  for (i = 0; i < s->size; i++) {
    double p = cx_norm_sq(s->vec[i]);
    if (p > bp) {
      bp = p;
      best = i;
    }
  }
  return best;
}

/* ---- gate matrices: src/q_gates.c:362-491 --------------------------------- */

static void put(double m[8], double a, double b, double c, double d, double e,
                double f, double g, double h) {
  m[0] = a; m[1] = b; m[2] = c; m[3] = d;
  m[4] = e; m[5] = f; m[6] = g; m[7] = h;
}

void oq_gate(int kind, double angle, double m[8]) {
  double c = cos(angle / 2.0), s = sin(angle / 2.0), r = 1.0 / sqrt(2.0);
  switch (kind) {
    case OQ_H: put(m, r, 0, r, 0, r, 0, -r, 0); break;
    case OQ_X: put(m, 0, 0, 1, 0, 1, 0, 0, 0); break;
    case OQ_Y: put(m, 0, 0, 0, -1, 0, 1, 0, 0); break;
    case OQ_Z: put(m, 1, 0, 0, 0, 0, 0, -1, 0); break;
    case OQ_P: put(m, 1, 0, 0, 0, 0, 0, cos(angle), sin(angle)); break;
    case OQ_RX: put(m, c, 0, 0, -s, 0, -s, c, 0); break;
    case OQ_RY: put(m, c, 0, -s, 0, s, 0, c, 0); break;
    case OQ_RZ: put(m, c, -s, 0, 0, 0, 0, c, s); break;
    default: put(m, 1, 0, 0, 0, 0, 0, 1, 0); break;
  }
}

/* ---- circuit runner ------------------------------------------------------
 * Applies `ng` ops: op[i] is an OQ_* kind, tq/cq target and control qubit,
 * par the angle.  CNOT / CPHASE go through oq_apply_cu, as qc_cnot and
 * qc_cphase do (src/qcs.c:184-190,431-437). */
void oq_run(oq_state *s, int ng, const int *op, const int *tq, const int *cq,
            const double *par) {
  int i;
  double m[8];
  for (i = 0; i < ng; i++) {
    switch (op[i]) {
      case OQ_CNOT:
        oq_gate(OQ_X, 0.0, m);
        oq_apply_cu(s, cq[i], tq[i], m);
        break;
      case OQ_CPHASE:
        oq_gate(OQ_P, par[i], m);
        oq_apply_cu(s, cq[i], tq[i], m);
        break;
      case OQ_SWAP:
        oq_apply_swap(s, tq[i], cq[i]);
        break;
      default:
        oq_gate(op[i], par[i], m);
        oq_apply_1q(s, tq[i], m);
        break;
    }
  }
}

/* qc_quantum_fourier_transform, src/qcs.c:443-457 (no final swaps) */
void oq_qft(oq_state *s) {
  int i, j;
  double m[8];
  for (i = 0; i < s->n; i++) {
    oq_gate(OQ_H, 0.0, m);
    oq_apply_1q(s, i, m);
    for (j = i + 1; j < s->n; j++) {
      oq_gate(OQ_P, ORACLE_PI / (double)(1 << (j - i)), m);
      oq_apply_cu(s, j, i, m);
    }
  }
}

/* qc_grover_search, src/qcs.c:402-419 */
void oq_grover(oq_state *s, long marked, long iterations) {
  int q;
  long i;
  double m[8];
  oq_gate(OQ_H, 0.0, m);
  for (q = 0; q < s->n; q++) oq_apply_1q(s, q, m);
  for (i = 0; i < iterations; i++) {
    oq_phase_flip(s, marked);
    oq_diffusion(s);
  }
}

/* ---- OpenMP restatement of the CPU-parallel 1q loop (src/q_gates.c:40-68) */

void oq_apply_1q_omp(oq_state *s, int t, const double m[8], int racy) {
  long step, block, i;
  ocx g0, g1, g2, g3;
  if (t < 0 || t >= s->n) return;
  step = 1L << t;
  block = 1L << (t + 1);
  g0 = mat(m, 0);
  g1 = mat(m, 1);
  g2 = mat(m, 2);
  g3 = mat(m, 3);
  memcpy(s->scr, s->vec, (size_t)s->size * sizeof(ocx));
  if (racy) {
    /* the original's inner index is declared at function scope and so is
       shared by every thread of the parallel loop */
    long j;
#pragma omp parallel for
    for (i = 0; i < s->size; i += block) {
      for (j = i; j < i + step; j++) {
        ocx v0 = s->vec[j], v1 = s->vec[j + step];
        s->scr[j] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
        s->scr[j + step] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
      }
    }
  } else {
#pragma omp parallel for
    for (i = 0; i < s->size; i += block) {
      long j;
      for (j = i; j < i + step; j++) {
        ocx v0 = s->vec[j], v1 = s->vec[j + step];
        s->scr[j] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
        s->scr[j + step] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
      }
    }
  }
  swap_buffers(s);
}

/* ---- OpenMP restatement of the CPU-parallel 2q loop (src/q_gates.c:174-208)
 * mode 0: race-fixed, textbook: the (c=1, t=0) iteration writes both entries
 *         of its pair and the (c=1, t=1) iteration writes nothing;
 * mode 1: race-fixed, as written: the result the sequential loop leaves
 *         (out[i1] = v1, its later else branch), the (c=1, t=0) iteration
 *         writing only i0;
 * mode 2: as written, racy: both iterations write i1, possibly on two
 *         threads. */
void oq_apply_cu_omp(oq_state *s, int c, int t, const double m[8], int mode) {
  long cb, tb, i;
  ocx g0, g1, g2, g3;
  if (c < 0 || t < 0 || c >= s->n || t >= s->n || c == t) return;
  cb = 1L << c;
  tb = 1L << t;
  g0 = mat(m, 0);
  g1 = mat(m, 1);
  g2 = mat(m, 2);
  g3 = mat(m, 3);
  memcpy(s->scr, s->vec, (size_t)s->size * sizeof(ocx));
  if (mode == 2) {
#pragma omp parallel for
    for (i = 0; i < s->size; i++) {
      if ((i & cb) != 0 && (i & tb) == 0) {
        ocx v0 = s->vec[i], v1 = s->vec[i | tb];
        s->scr[i] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
        s->scr[i | tb] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
      } else {
        s->scr[i] = s->vec[i];
      }
    }
  } else if (mode == 1) {
#pragma omp parallel for
    for (i = 0; i < s->size; i++) {
      if ((i & cb) != 0 && (i & tb) == 0) {
        ocx v0 = s->vec[i], v1 = s->vec[i | tb];
        s->scr[i] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
      } else {
        s->scr[i] = s->vec[i];
      }
    }
  } else {
#pragma omp parallel for
    for (i = 0; i < s->size; i++) {
      if ((i & cb) != 0 && (i & tb) == 0) {
        ocx v0 = s->vec[i], v1 = s->vec[i | tb];
        s->scr[i] = cx_add(cx_mul(g0, v0), cx_mul(g1, v1));
        s->scr[i | tb] = cx_add(cx_mul(g2, v0), cx_mul(g3, v1));
      } else if ((i & cb) == 0) {
        s->scr[i] = s->vec[i];
      }
    }
  }
  swap_buffers(s);
}

/* oq_run with the OpenMP loops: racy = 1 runs both loops as the reference's
 * OpenMP build writes them; racy = 0 runs the race-fixed loops in the state's
 * semantics (textbook or as written).  SWAP stays sequential (no reference
 * counterpart). */
void oq_run_omp(oq_state *s, int ng, const int *op, const int *tq, const int *cq,
                const double *par, int racy) {
  int i;
  double m[8];
  const int mode = racy ? 2 : (s->compat ? 1 : 0);
  for (i = 0; i < ng; i++) {
    switch (op[i]) {
      case OQ_CNOT:
        oq_gate(OQ_X, 0.0, m);
        oq_apply_cu_omp(s, cq[i], tq[i], m, mode);
        break;
      case OQ_CPHASE:
        oq_gate(OQ_P, par[i], m);
        oq_apply_cu_omp(s, cq[i], tq[i], m, mode);
        break;
      case OQ_SWAP:
        oq_apply_swap(s, tq[i], cq[i]);
        break;
      default:
        oq_gate(op[i], par[i], m);
        oq_apply_1q_omp(s, tq[i], m, racy);
        break;
    }
  }
}

/* threads of the OpenMP loops (the runtime this library was linked with) */
void oq_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}
