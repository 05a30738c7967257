// FP64 reductions and the observation path: norm / renormalise, qubit
// probabilities + collapse (measurement), argmax, shot sampling, Grover's
// oracle and diffusion, and the state comparison (1 - F, max-abs).
//
// Every reduction runs on a fixed grid of <= kRedBlocks work-groups that
// write per-block partials; one work-group then folds the partials in a fixed
// order, so results are bit-reproducible run to run.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "device.h"
#include "internal.h"

namespace qcs {

static unsigned red_blocks(uint64_t n) {
  const uint64_t b = ceil_div(n, uint64_t(kThreads) * 4);
  return unsigned(b < uint64_t(kRedBlocks) ? (b ? b : 1) : kRedBlocks);
}

// ---- generic partial + final ---------------------------------------------------

// Sum |a|^2, split by bit q of the index when q >= 0 (slot 0: bit = 0,
// slot 1: bit = 1); q < 0 puts everything in slot 0.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_norm_partial(
    const typename Amp<T>::type *a, uint64_t n, int q, double *partial) {
  __shared__ double lds[kThreads / 64];
  double s0 = 0.0, s1 = 0.0;
  const uint64_t qb = q >= 0 ? (uint64_t(1) << q) : 0;
  // XCD-ordered virtual blocks (each XCD reads one contiguous range), four
  // loads in flight per thread
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads)), G = gridDim.x;
  auto at = [&](uint64_t vb) { return xcd_order(vb, nvb) * kThreads + threadIdx.x; };
  uint64_t vb = blockIdx.x;
  for (; vb + 3 * G < nvb; vb += 4 * G) {
    const uint64_t i0 = at(vb), i1 = at(vb + G), i2 = at(vb + 2 * G), i3 = at(vb + 3 * G);
    // xcd_order can put the grid's one partial virtual block at any of the
    // four slots, so each load is guarded (a skipped term adds 0.0: the
    // association of the sum is unchanged)
    const double p0 = i0 < n ? norm2(ld_stream(a + i0)) : 0.0;
    const double p1 = i1 < n ? norm2(ld_stream(a + i1)) : 0.0;
    const double p2 = i2 < n ? norm2(ld_stream(a + i2)) : 0.0;
    const double p3 = i3 < n ? norm2(ld_stream(a + i3)) : 0.0;
    if (i0 & qb) s1 += p0; else s0 += p0;
    if (i1 & qb) s1 += p1; else s0 += p1;
    if (i2 & qb) s1 += p2; else s0 += p2;
    if (i3 & qb) s1 += p3; else s0 += p3;
  }
  for (; vb < nvb; vb += G) {
    const uint64_t i = at(vb);
    if (i >= n) continue;
    const double p = norm2(ld_stream(a + i));
    if (i & qb) s1 += p; else s0 += p;
  }
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * kRedSlots + 0] = s0;
    partial[blockIdx.x * kRedSlots + 1] = s1;
  }
}

// Sum |a|^2 over the half of the state whose bit q is set (bq = 2^q) or
// clear (bq = 0): half-index j is amplitude insert0(j, q) | bq, so the half
// is read in runs of 2^q amplitudes (the caller uses it where a run is at
// least a 128-byte line).  Same loop shape as k_norm_partial; slot 0 only.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_half_norm_partial(
    const typename Amp<T>::type *a, uint64_t nh, int q, uint64_t bq, double *partial) {
  __shared__ double lds[kThreads / 64];
  double s0 = 0.0;
  const uint64_t nvb = ceil_div(nh, uint64_t(kThreads)), G = gridDim.x;
  auto at = [&](uint64_t vb) {
    return insert0(xcd_order(vb, nvb) * kThreads + threadIdx.x, q) | bq;
  };
  auto inside = [&](uint64_t vb) { return xcd_order(vb, nvb) * kThreads + threadIdx.x < nh; };
  uint64_t vb = blockIdx.x;
  for (; vb + 3 * G < nvb; vb += 4 * G) {
    // each of the four may be the partial virtual block (see k_norm_partial)
    const double p0 = inside(vb) ? norm2(ld_stream(a + at(vb))) : 0.0;
    const double p1 = inside(vb + G) ? norm2(ld_stream(a + at(vb + G))) : 0.0;
    const double p2 = inside(vb + 2 * G) ? norm2(ld_stream(a + at(vb + 2 * G))) : 0.0;
    const double p3 = inside(vb + 3 * G) ? norm2(ld_stream(a + at(vb + 3 * G))) : 0.0;
    s0 += p0;
    s0 += p1;
    s0 += p2;
    s0 += p3;
  }
  for (; vb < nvb; vb += G)
    if (inside(vb)) s0 += norm2(ld_stream(a + at(vb)));
  s0 = block_sum(s0, lds);
  if (threadIdx.x == 0) partial[blockIdx.x * kRedSlots + 0] = s0;
}

// Fold `nb` partials of `nslots` slots: slots in max_mask take the maximum,
// the others the sum.  Results go to out[0..nslots).
__global__ __launch_bounds__(kThreads) void k_final(const double *partial,
                                                    int nb, int nslots,
                                                    unsigned max_mask,
                                                    double *out) {
  __shared__ double lds[kThreads];
  for (int slot = 0; slot < nslots; ++slot) {
    const bool is_max = (max_mask >> slot) & 1u;
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) {
      const double x = partial[b * kRedSlots + slot];
      v = is_max ? fmax(v, x) : v + x;
    }
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        const double y = lds[threadIdx.x + w];
        lds[threadIdx.x] = is_max ? fmax(lds[threadIdx.x], y)
                                  : lds[threadIdx.x] + y;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = lds[0];
    __syncthreads();
  }
}

// a *= factor
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scale(typename Amp<T>::type *a,
                                                    uint64_t n, double f) {
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t i = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    if (i >= n) continue;
    typename Amp<T>::type v = ld_stream(a + i);
    v.x = T(double(v.x) * f);
    v.y = T(double(v.y) * f);
    st_stream(a + i, v);
  }
}

// a *= 1/sqrt(ns) when ns > 1e-12 and ns != 1 (q_state_normalize's guard,
// src/q_utils.c:112-118), with ns read from device memory.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scale_dev(
    typename Amp<T>::type *a, uint64_t n, const double *ns) {
  const double v = *ns;
  if (!(v > 1e-12 && v != 1.0)) return;
  const double f = 1.0 / sqrt(v);
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t i = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    if (i >= n) continue;
    typename Amp<T>::type x = ld_stream(a + i);
    x.x = T(double(x.x) * f);
    x.y = T(double(x.y) * f);
    st_stream(a + i, x);
  }
}

// qc_measure's collapse: zero the amplitudes whose bit q differs from
// `outcome`, scale the others by f when do_scale.
template <typename T>
__device__ __forceinline__ void collapse_body(typename Amp<T>::type *a, uint64_t n, int q,
                                              int outcome, double f, int do_scale);

template <typename T>
__global__ __launch_bounds__(kThreads) void k_collapse(
    typename Amp<T>::type *a, uint64_t n, int q, int outcome, double f,
    int do_scale) {
  collapse_body<T>(a, n, q, outcome, f, do_scale);
}

// qc_measure's decision on the device (src/qcs.c:261-283): m[0], m[1] hold
// p0 and p1; outcome = (u <= p0) ? 0 : 1, and the kept half, whose norm is
// p0 or p1, is rescaled by 1/sqrt of it when that is > 1e-12 and != 1 (the
// rule of q_state_normalize).  m[2..4] = outcome, factor, rescale flag; the
// outcome also goes to *record.
__global__ void k_measure_decide(double *m, double u, double *record) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double p0 = m[0], p1 = m[1];
  const int o = (u <= p0) ? 0 : 1;
  const double ns = o == 0 ? p0 : p1;
  const int sc = (ns > 1e-12 && ns != 1.0) ? 1 : 0;
  m[2] = double(o);
  m[3] = sc ? 1.0 / sqrt(ns) : 1.0;
  m[4] = double(sc);
  *record = double(o);
}

// k_collapse with the outcome, factor and flag k_measure_decide wrote
template <typename T>
__global__ __launch_bounds__(kThreads) void k_collapse_dev(typename Amp<T>::type *a,
                                                           uint64_t n, int q,
                                                           const double *m) {
  collapse_body<T>(a, n, q, int(m[2]), m[3], int(m[4]));
}

template <typename T>
__device__ __forceinline__ void collapse_body(typename Amp<T>::type *a, uint64_t n, int q,
                                              int outcome, double f, int do_scale) {
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t i = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    if (i >= n) continue;
    const int bit = int((i >> q) & 1);
    typename Amp<T>::type v;
    if (bit != outcome) {
      v.x = T(0);
      v.y = T(0);
      st_stream(a + i, v);
    } else if (do_scale) {
      v = ld_stream(a + i);
      v.x = T(double(v.x) * f);
      v.y = T(double(v.y) * f);
      st_stream(a + i, v);
    }
  }
}

// ---- argmax -------------------------------------------------------------------

__device__ __forceinline__ bool better(double p, uint64_t i, double bp,
                                       uint64_t bi) {
  return p > bp || (p == bp && i < bi);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_argmax_partial(
    const typename Amp<T>::type *a, uint64_t n, double *partial) {
  __shared__ double lp[kThreads];
  __shared__ uint64_t li[kThreads];
  double bp = -1.0;
  uint64_t bi = ~uint64_t(0);
  // XCD-ordered virtual blocks, four loads in flight; a thread does not
  // visit its indices in increasing order, so ties go to the smaller index
  // explicitly (the first maximum in index order wins; NaN is skipped)
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads)), G = gridDim.x;
  auto at = [&](uint64_t vb) { return xcd_order(vb, nvb) * kThreads + threadIdx.x; };
  auto take = [&](double p, uint64_t i) {
    if (better(p, i, bp, bi)) {
      bp = p;
      bi = i;
    }
  };
  uint64_t vb = blockIdx.x;
  for (; vb + 3 * G < nvb; vb += 4 * G) {
    const uint64_t i0 = at(vb), i1 = at(vb + G), i2 = at(vb + 2 * G), i3 = at(vb + 3 * G);
    // any of the four may be the partial virtual block (see k_norm_partial)
    const double p0 = i0 < n ? norm2(ld_stream(a + i0)) : -1.0;
    const double p1 = i1 < n ? norm2(ld_stream(a + i1)) : -1.0;
    const double p2 = i2 < n ? norm2(ld_stream(a + i2)) : -1.0;
    const double p3 = i3 < n ? norm2(ld_stream(a + i3)) : -1.0;
    if (i0 < n) take(p0, i0);
    if (i1 < n) take(p1, i1);
    if (i2 < n) take(p2, i2);
    if (i3 < n) take(p3, i3);
  }
  for (; vb < nvb; vb += G) {
    const uint64_t i = at(vb);
    if (i < n) take(norm2(ld_stream(a + i)), i);
  }
  lp[threadIdx.x] = bp;
  li[threadIdx.x] = bi;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double p = lp[threadIdx.x + w];
      const uint64_t i = li[threadIdx.x + w];
      if (better(p, i, lp[threadIdx.x], li[threadIdx.x])) {
        lp[threadIdx.x] = p;
        li[threadIdx.x] = i;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x * kRedSlots + 0] = lp[0];
    partial[blockIdx.x * kRedSlots + 1] = __longlong_as_double((long long)li[0]);
  }
}

__global__ __launch_bounds__(kThreads) void k_argmax_final(
    const double *partial, int nb, double *out) {
  __shared__ double lp[kThreads];
  __shared__ uint64_t li[kThreads];
  double bp = -1.0;
  uint64_t bi = ~uint64_t(0);
  for (int b = threadIdx.x; b < nb; b += kThreads) {
    const double p = partial[b * kRedSlots + 0];
    const uint64_t i = uint64_t(__double_as_longlong(partial[b * kRedSlots + 1]));
    if (better(p, i, bp, bi)) {
      bp = p;
      bi = i;
    }
  }
  lp[threadIdx.x] = bp;
  li[threadIdx.x] = bi;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double p = lp[threadIdx.x + w];
      const uint64_t i = li[threadIdx.x + w];
      if (better(p, i, lp[threadIdx.x], li[threadIdx.x])) {
        lp[threadIdx.x] = p;
        li[threadIdx.x] = i;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = lp[0];
    out[1] = __longlong_as_double((long long)li[0]);
  }
}

// ---- sampling -------------------------------------------------------------------
//
// The CDF of qc_run_shots (src/qcs.c:597-604) is defined blockwise:
//   chunk c covers indices [c*L, (c+1)*L), L = min(kChunk, N), read as rows of
//   64; inside a row the 64 lanes take an inclusive wave scan; rows are
//   chained by a sequential running sum.  chunk_tot[c] is the final running
//   sum.  E[c] (exclusive) and P[c] = E[c] + chunk_tot[c] come from a
//   sequential chain per scan thread.  cdf(i) = E[c] + local(i).
// The sampling kernel recomputes local(i) with the same operations, so
// cdf(last of chunk c) == P[c] bit for bit.


template <typename T>
__global__ __launch_bounds__(kThreads) void k_chunk_tot(
    const typename Amp<T>::type *a, uint64_t nchunks, uint64_t len,
    double *tot) {
  const int lane = threadIdx.x & 63;
  const uint64_t c = uint64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (c >= nchunks) return;
  const typename Amp<T>::type *p = a + c * len;
  double running = 0.0;
  for (uint64_t r = 0; r < len; r += 64) {
    const double x = (r + lane < len) ? norm2(p[r + lane]) : 0.0;
    const double s = wave_incl_scan(x, lane);
    running = running + __shfl(s, 63, 64);
  }
  if (lane == 0) tot[c] = running;
}

// One work-group: excl[c] (in place over tot) and inc[c] = excl[c] + tot[c].
// The additions run in a fixed order (each thread's segment left to right,
// then the segment sums left to right), which the batch kernel's sampling
// shares; the loads are issued 16 at a time ahead of the additions, so the
// sequential chains wait on arithmetic rather than on memory.
__global__ __launch_bounds__(1024) void k_chunk_scan(double *tot, double *inc,
                                                     uint64_t nchunks) {
  constexpr int U = 16;
  __shared__ double base[1024];
  const uint64_t per = ceil_div(nchunks, 1024);
  const uint64_t b = uint64_t(threadIdx.x) * per;
  const uint64_t e = b + per < nchunks ? b + per : nchunks;
  double s = 0.0;
  uint64_t c = b;
  for (; c + U <= e; c += U) {
    double v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) v[j] = tot[c + j];
#pragma unroll
    for (int j = 0; j < U; ++j) s = s + v[j];
  }
  for (; c < e; ++c) s = s + tot[c];
  base[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int t0 = 0; t0 < 1024; t0 += U) {
      double v[U];
#pragma unroll
      for (int j = 0; j < U; ++j) v[j] = base[t0 + j];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        base[t0 + j] = run;
        run = run + v[j];
      }
    }
  }
  __syncthreads();
  double run = base[threadIdx.x];
  c = b;
  for (; c + U <= e; c += U) {
    double v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) v[j] = tot[c + j];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      tot[c + j] = run;
      run = run + v[j];
      inc[c + j] = run;
    }
  }
  for (; c < e; ++c) {
    const double t = tot[c];
    tot[c] = run;
    run = run + t;
    inc[c] = run;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_sample(
    const typename Amp<T>::type *a, uint64_t nchunks, uint64_t len,
    const double *excl, const double *inc, const double *u, uint64_t shots,
    long *out) {
  const int lane = threadIdx.x & 63;
  const uint64_t s = uint64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  if (s >= shots) return;
  const double us = u[s];
  // first chunk with inc > u (inc is non-decreasing)
  uint64_t lo = 0, hi = nchunks;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (inc[mid] > us) hi = mid; else lo = mid + 1;
  }
  if (lo >= nchunks) {  // u not below the total: dropped, as the reference
    if (lane == 0) out[s] = -1;
    return;
  }
  const uint64_t c = lo;
  const double e = excl[c];
  const typename Amp<T>::type *p = a + c * len;
  double running = 0.0;
  long found = -1, last_nz = -1;
  for (uint64_t r = 0; r < len; r += 64) {
    const double x = (r + lane < len) ? norm2(p[r + lane]) : 0.0;
    const double sc = wave_incl_scan(x, lane);
    const double local = running + sc;
    const double cdf = e + local;
    const unsigned long long hit = __ballot(cdf > us && x > 0.0);
    const unsigned long long nz = __ballot(x > 0.0);
    if (nz) last_nz = long(c * len + r) + 63 - __clzll(nz);
    if (hit) {
      found = long(c * len + r) + __ffsll(hit) - 1;
      break;
    }
    running = running + __shfl(sc, 63, 64);
  }
  // only reachable at an ulp-level seam of the chunk scan: take the last
  // non-zero amplitude seen in the chunk
  if (found < 0) found = last_nz;
  if (lane == 0) out[s] = found;
}

// ---- Grover (textbook) ----------------------------------------------------
//
// One round: a[m] = -a[m]; mu = sum(a)/N; a = 2*mu - a.  Written as
// a'_i = 2*mu - s_i*a_i with s_m = -1.  k_grover_update also accumulates
// sum_i s_i*a'_i, the flipped sum the next round needs, so a round is one
// read+write pass plus a tiny fold.

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gsum_partial(
    const typename Amp<T>::type *a, uint64_t n, uint64_t m, double *partial) {
  __shared__ double lds[kThreads / 64];
  double sr = 0.0, si = 0.0;
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t i = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    if (i >= n) continue;
    const typename Amp<T>::type v = ld_stream(a + i);
    const double sg = (i == m) ? -1.0 : 1.0;
    sr += sg * double(v.x);
    si += sg * double(v.y);
  }
  sr = block_sum(sr, lds);
  si = block_sum(si, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * kRedSlots + 0] = sr;
    partial[blockIdx.x * kRedSlots + 1] = si;
  }
}

// fold partials to 2*mu = 2*sum/N in out[0..1]
__global__ __launch_bounds__(kThreads) void k_gfinal(const double *partial,
                                                     int nb, double ninv,
                                                     double *out) {
  __shared__ double lds[kThreads / 64];
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nb; b += kThreads) {
    sr += partial[b * kRedSlots + 0];
    si += partial[b * kRedSlots + 1];
  }
  sr = block_sum(sr, lds);
  si = block_sum(si, lds);
  if (threadIdx.x == 0) {
    out[0] = 2.0 * (sr * ninv);
    out[1] = 2.0 * (si * ninv);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gupdate(typename Amp<T>::type *a,
                                                      uint64_t n, uint64_t m,
                                                      const double *two_mu,
                                                      double *partial) {
  __shared__ double lds[kThreads / 64];
  const double tr = two_mu[0], ti = two_mu[1];
  double sr = 0.0, si = 0.0;
  // XCD-ordered virtual blocks: each XCD sweeps one contiguous range
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t i = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    if (i >= n) continue;
    typename Amp<T>::type v = ld_stream(a + i);
    const double sg = (i == m) ? -1.0 : 1.0;
    v.x = T(tr - sg * double(v.x));
    v.y = T(ti - sg * double(v.y));
    st_stream(a + i, v);
    sr += sg * double(v.x);
    si += sg * double(v.y);
  }
  sr = block_sum(sr, lds);
  si = block_sum(si, lds);
  if (threadIdx.x == 0) {
    partial[blockIdx.x * kRedSlots + 0] = sr;
    partial[blockIdx.x * kRedSlots + 1] = si;
  }
}

// ---- Grover (compat, as written in src/q_gates.c:305-356) -----------------

// after the host swapped the buffers: vector[m] = -scratch[m]
template <typename T>
__global__ void k_flip_compat(typename Amp<T>::type *vec,
                              const typename Amp<T>::type *old, uint64_t m) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    typename Amp<T>::type v = old[m];
    v.x = -v.x;
    v.y = -v.y;
    vec[m] = v;
  }
}

// mean = sum a; if |mean| > 1e-10 then mean /= |mean|; out = 2*mean
__global__ __launch_bounds__(kThreads) void k_cfinal(const double *partial,
                                                     int nb, double *out) {
  __shared__ double lds[kThreads / 64];
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nb; b += kThreads) {
    sr += partial[b * kRedSlots + 0];
    si += partial[b * kRedSlots + 1];
  }
  sr = block_sum(sr, lds);
  si = block_sum(si, lds);
  if (threadIdx.x == 0) {
    const double mag = sqrt(sr * sr + si * si);
    if (mag > 1e-10) {
      sr /= mag;
      si /= mag;
    }
    out[0] = 2.0 * sr;
    out[1] = 2.0 * si;
  }
}

// vector[i] = two_mean - scratch[i]
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cdiff(
    const typename Amp<T>::type *old, typename Amp<T>::type *vec, uint64_t n,
    const double *two_mean) {
  const double tr = two_mean[0], ti = two_mean[1];
  const uint64_t stride = uint64_t(gridDim.x) * kThreads;
  for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += stride) {
    const typename Amp<T>::type v = old[i];
    typename Amp<T>::type o;
    o.x = T(tr - double(v.x));
    o.y = T(ti - double(v.y));
    vec[i] = o;
  }
}

// a[m] = -a[m]  (textbook oracle, one amplitude)
template <typename T>
__global__ void k_negate(typename Amp<T>::type *a, uint64_t m) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    typename Amp<T>::type v = a[m];
    v.x = -v.x;
    v.y = -v.y;
    a[m] = v;
  }
}

// ---- comparison --------------------------------------------------------------

template <typename TA, typename TB>
__global__ __launch_bounds__(kThreads) void k_cmp_partial(
    const typename Amp<TA>::type *a, const typename Amp<TB>::type *b,
    uint64_t n, double *partial) {
  __shared__ double lds[kThreads / 64];
  __shared__ double lmax[kThreads];
  const uint64_t stride = uint64_t(gridDim.x) * kThreads;
  double dr = 0.0, di = 0.0, na = 0.0, nb = 0.0, mx = 0.0;
  for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += stride) {
    const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
    dr += ar * br + ai * bi;  // conj(a) * b
    di += ar * bi - ai * br;
    na += ar * ar + ai * ai;
    nb += br * br + bi * bi;
    const double ex = ar - br, ey = ai - bi;
    const double e = sqrt(ex * ex + ey * ey);
    mx = e > mx ? e : mx;
  }
  dr = block_sum(dr, lds);
  di = block_sum(di, lds);
  na = block_sum(na, lds);
  nb = block_sum(nb, lds);
  lmax[threadIdx.x] = mx;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) lmax[threadIdx.x] = fmax(lmax[threadIdx.x], lmax[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double *o = partial + blockIdx.x * kRedSlots;
    o[0] = dr;
    o[1] = di;
    o[2] = na;
    o[3] = nb;
    o[4] = lmax[0];
  }
}

// ---- host launchers ----------------------------------------------------------------

static double *result_slot(qcs_hip_state *s) {
  return s->red + size_t(kRedBlocks) * kRedSlots;
}

// copy `count` doubles from the device result slots to the pinned buffer and
// wait for the stream
static int fetch_results(qcs_hip_state *s, int count) {
  QCS_HIP_TRY(hipMemcpyAsync(s->host, result_slot(s), sizeof(double) * count,
                             hipMemcpyDeviceToHost, s->stream));
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

static unsigned stream_grid(uint64_t n) {
  const uint64_t b = ceil_div(n, kThreads);
  return unsigned(b < 8192 ? b : 8192);
}

// element-wise sweeps: one amplitude per thread over XCD-ordered virtual
// blocks, up to 2^20 work-groups (the gate kernels' launch shape)
static unsigned elem_grid(uint64_t n) {
  const uint64_t b = ceil_div(n, kThreads), cap = uint64_t(1) << 20;
  return unsigned(b < cap ? b : cap);
}

template <typename T>
static int norm_partials(qcs_hip_state *s, int q, unsigned nb) {
  typedef typename Amp<T>::type A;
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_REDUCE, double(s->size) * sizeof(A), 3.0 * s->size,
               &tl);
  hipLaunchKernelGGL(k_norm_partial<T>, dim3(nb), dim3(kThreads), 0, s->stream,
                     static_cast<const A *>(s->vec), s->size, q, s->red);
  int rc = launch_check("k_norm_partial");
  if (rc) return rc;
  return timing_end(s, &tl);
}

static int fold(qcs_hip_state *s, unsigned nb, int nslots, unsigned max_mask) {
  hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s->stream, s->red,
                     int(nb), nslots, max_mask, result_slot(s));
  return launch_check("k_final");
}

int reduce_norm(qcs_hip_state *s, double *out_host) {
  const unsigned nb = red_blocks(s->size);
  int rc = s->precision == QCS_HIP_C128 ? norm_partials<double>(s, -1, nb)
                                        : norm_partials<float>(s, -1, nb);
  if (rc) return rc;
  if ((rc = fold(s, nb, 1, 0u))) return rc;
  if ((rc = fetch_results(s, 1))) return rc;
  *out_host = s->host[0];
  return 0;
}

template <typename T>
static int scale_impl(qcs_hip_state *s, double f, const double *dev_ns) {
  typedef typename Amp<T>::type A;
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_OTHER, 2.0 * double(s->size) * sizeof(A),
               2.0 * s->size, &tl);
  if (dev_ns)
    hipLaunchKernelGGL(k_scale_dev<T>, dim3(elem_grid(s->size)),
                       dim3(kThreads), 0, s->stream, static_cast<A *>(s->vec),
                       s->size, dev_ns);
  else
    hipLaunchKernelGGL(k_scale<T>, dim3(elem_grid(s->size)), dim3(kThreads),
                       0, s->stream, static_cast<A *>(s->vec), s->size, f);
  int rc = launch_check("k_scale");
  if (rc) return rc;
  return timing_end(s, &tl);
}

int scale_state(qcs_hip_state *s, double f) {
  return s->precision == QCS_HIP_C128 ? scale_impl<double>(s, f, nullptr)
                                      : scale_impl<float>(s, f, nullptr);
}

int normalize_device(qcs_hip_state *s) {
  const unsigned nb = red_blocks(s->size);
  int rc = s->precision == QCS_HIP_C128 ? norm_partials<double>(s, -1, nb)
                                        : norm_partials<float>(s, -1, nb);
  if (rc) return rc;
  if ((rc = fold(s, nb, 1, 0u))) return rc;
  return s->precision == QCS_HIP_C128
             ? scale_impl<double>(s, 1.0, result_slot(s))
             : scale_impl<float>(s, 1.0, result_slot(s));
}

int fold_norm_partials(qcs_hip_state *s, int nb, const double **ns_dev) {
  *ns_dev = result_slot(s);
  return fold(s, unsigned(nb), 1, 0u);
}

int normalize_from_partials(qcs_hip_state *s, int nb) {
  int rc;
  if ((rc = fold(s, unsigned(nb), 1, 0u))) return rc;
  return s->precision == QCS_HIP_C128
             ? scale_impl<double>(s, 1.0, result_slot(s))
             : scale_impl<float>(s, 1.0, result_slot(s));
}

int qubit_probs(qcs_hip_state *s, int q, double *p0, double *p1) {
  const unsigned nb = red_blocks(s->size);
  int rc = s->precision == QCS_HIP_C128 ? norm_partials<double>(s, q, nb)
                                        : norm_partials<float>(s, q, nb);
  if (rc) return rc;
  if ((rc = fold(s, nb, 2, 0u))) return rc;
  if ((rc = fetch_results(s, 2))) return rc;
  *p0 = s->host[0];
  *p1 = s->host[1];
  return 0;
}

template <typename T>
static int half_partials(qcs_hip_state *s, int q, int bit, unsigned nb) {
  typedef typename Amp<T>::type A;
  const uint64_t nh = s->size / 2;
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_REDUCE, double(nh) * sizeof(A), 3.0 * double(nh), &tl);
  hipLaunchKernelGGL(k_half_norm_partial<T>, dim3(nb), dim3(kThreads), 0, s->stream,
                     static_cast<const A *>(s->vec), nh, q,
                     bit ? (uint64_t(1) << q) : uint64_t(0), s->red);
  int rc = launch_check("k_half_norm_partial");
  if (rc) return rc;
  return timing_end(s, &tl);
}

bool half_prob_pays(const qcs_hip_state *s, int q) {
  return (size_t(1) << q) * size_t(s->amp_bytes) >= 128;
}

int half_prob(qcs_hip_state *s, int q, int bit, double *p) {
  const unsigned nb = red_blocks(s->size / 2);
  int rc = s->precision == QCS_HIP_C128 ? half_partials<double>(s, q, bit, nb)
                                        : half_partials<float>(s, q, bit, nb);
  if (rc) return rc;
  if ((rc = fold(s, nb, 1, 0u))) return rc;
  if ((rc = fetch_results(s, 1))) return rc;
  *p = s->host[0];
  return 0;
}

template <typename T>
static int collapse_impl(qcs_hip_state *s, int q, int outcome, double f,
                         int do_scale) {
  typedef typename Amp<T>::type A;
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_OTHER, 1.5 * double(s->size) * sizeof(A), 0.0, &tl);
  hipLaunchKernelGGL(k_collapse<T>, dim3(elem_grid(s->size)), dim3(kThreads),
                     0, s->stream, static_cast<A *>(s->vec), s->size, q,
                     outcome, f, do_scale);
  int rc = launch_check("k_collapse");
  if (rc) return rc;
  return timing_end(s, &tl);
}

int collapse(qcs_hip_state *s, int q, int outcome, double f, int do_scale) {
  return s->precision == QCS_HIP_C128
             ? collapse_impl<double>(s, q, outcome, f, do_scale)
             : collapse_impl<float>(s, q, outcome, f, do_scale);
}

// Measurements of `count` qubits in order, with no host round trip between
// them: per qubit the p0 / p1 reductions qcs_hip_measure uses (the half
// passes, or one pass for both where the halves' runs are shorter than a
// line; here p1 is always formed, as the device decides the outcome),
// k_measure_decide, and the collapse reading its decision.  Result slots:
// 8..12 the current measurement, 16.. the outcomes.
template <typename T>
static int measure_seq_impl(qcs_hip_state *s, int count, const int *qubits,
                            const double *u, int *outcomes) {
  typedef typename Amp<T>::type A;
  double *m = result_slot(s) + 8, *rec = result_slot(s) + 16;
  constexpr int kMaxRec = 48;
  int rc;
  for (int i0 = 0; i0 < count; i0 += kMaxRec) {
    const int c = std::min(kMaxRec, count - i0);
    for (int i = 0; i < c; ++i) {
      const int q = qubits[i0 + i];
      if (half_prob_pays(s, q)) {
        const unsigned nb = red_blocks(s->size / 2);
        for (int bit = 0; bit < 2; ++bit) {
          if ((rc = half_partials<T>(s, q, bit, nb))) return rc;
          hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s->stream, s->red,
                             int(nb), 1, 0u, m + bit);
          if ((rc = launch_check("k_final"))) return rc;
        }
      } else {
        const unsigned nb = red_blocks(s->size);
        if ((rc = norm_partials<T>(s, q, nb))) return rc;
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kThreads), 0, s->stream, s->red,
                           int(nb), 2, 0u, m);
        if ((rc = launch_check("k_final"))) return rc;
      }
      hipLaunchKernelGGL(k_measure_decide, dim3(1), dim3(64), 0, s->stream, m,
                         u[i0 + i], rec + i);
      if ((rc = launch_check("k_measure_decide"))) return rc;
      TimedLaunch tl;
      timing_begin(s, QCS_HIP_K_OTHER, 1.5 * double(s->size) * sizeof(A), 0.0, &tl);
      hipLaunchKernelGGL(k_collapse_dev<T>, dim3(elem_grid(s->size)), dim3(kThreads), 0,
                         s->stream, static_cast<A *>(s->vec), s->size, q, m);
      if ((rc = launch_check("k_collapse_dev"))) return rc;
      if ((rc = timing_end(s, &tl))) return rc;
    }
    QCS_HIP_TRY(hipMemcpyAsync(s->host, rec, sizeof(double) * size_t(c),
                               hipMemcpyDeviceToHost, s->stream));
    QCS_HIP_TRY(hipStreamSynchronize(s->stream));
    for (int i = 0; i < c; ++i) outcomes[i0 + i] = int(s->host[i]);
  }
  return 0;
}

int measure_seq(qcs_hip_state *s, int count, const int *qubits, const double *u,
                int *outcomes) {
  return s->precision == QCS_HIP_C128
             ? measure_seq_impl<double>(s, count, qubits, u, outcomes)
             : measure_seq_impl<float>(s, count, qubits, u, outcomes);
}

template <typename T>
static int argmax_impl(qcs_hip_state *s, long *idx) {
  typedef typename Amp<T>::type A;
  const unsigned nb = red_blocks(s->size);
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_REDUCE, double(s->size) * sizeof(A), 3.0 * s->size,
               &tl);
  hipLaunchKernelGGL(k_argmax_partial<T>, dim3(nb), dim3(kThreads), 0,
                     s->stream, static_cast<const A *>(s->vec), s->size, s->red);
  int rc = launch_check("k_argmax_partial");
  if (rc) return rc;
  if ((rc = timing_end(s, &tl))) return rc;
  hipLaunchKernelGGL(k_argmax_final, dim3(1), dim3(kThreads), 0, s->stream,
                     s->red, int(nb), result_slot(s));
  if ((rc = launch_check("k_argmax_final"))) return rc;
  if ((rc = fetch_results(s, 2))) return rc;
  long long bits;
  static_assert(sizeof(bits) == sizeof(double), "double is 64-bit");
  __builtin_memcpy(&bits, &s->host[1], sizeof(bits));
  *idx = (s->host[0] > 0.0) ? long(bits) : 0;
  return 0;
}

int argmax(qcs_hip_state *s, long *idx) {
  return s->precision == QCS_HIP_C128 ? argmax_impl<double>(s, idx)
                                      : argmax_impl<float>(s, idx);
}

template <typename T>
static int sample_impl(qcs_hip_state *s, long shots, const double *u,
                       long *out) {
  typedef typename Amp<T>::type A;
  const uint64_t len = s->size < uint64_t(kChunk) ? s->size : uint64_t(kChunk);
  const uint64_t nchunks = s->size / len;
  if (s->chunk_cap < nchunks) {
    if (s->chunk_tot) (void)hipFree(s->chunk_tot);
    if (s->chunk_inc) (void)hipFree(s->chunk_inc);
    s->chunk_tot = s->chunk_inc = nullptr;
    s->chunk_cap = 0;
    QCS_HIP_TRY(hipMalloc(&s->chunk_tot, nchunks * sizeof(double)));
    QCS_HIP_TRY(hipMalloc(&s->chunk_inc, nchunks * sizeof(double)));
    s->chunk_cap = nchunks;
  }
  if (s->shots_cap < size_t(shots)) {
    if (s->u_dev) (void)hipFree(s->u_dev);
    if (s->idx_dev) (void)hipFree(s->idx_dev);
    s->u_dev = nullptr;
    s->idx_dev = nullptr;
    s->shots_cap = 0;
    QCS_HIP_TRY(hipMalloc(&s->u_dev, size_t(shots) * sizeof(double)));
    QCS_HIP_TRY(hipMalloc(&s->idx_dev, size_t(shots) * sizeof(long)));
    s->shots_cap = size_t(shots);
  }
  QCS_HIP_TRY(hipMemcpyAsync(s->u_dev, u, size_t(shots) * sizeof(double),
                             hipMemcpyHostToDevice, s->stream));
  const A *a = static_cast<const A *>(s->vec);
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_REDUCE, double(s->size) * sizeof(A), 3.0 * s->size,
               &tl);
  hipLaunchKernelGGL(k_chunk_tot<T>, dim3(unsigned(ceil_div(nchunks, kThreads / 64))),
                     dim3(kThreads), 0, s->stream, a, nchunks, len, s->chunk_tot);
  int rc = launch_check("k_chunk_tot");
  if (rc) return rc;
  if ((rc = timing_end(s, &tl))) return rc;
  hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, s->stream,
                     s->chunk_tot, s->chunk_inc, nchunks);
  if ((rc = launch_check("k_chunk_scan"))) return rc;
  timing_begin(s, QCS_HIP_K_OTHER, double(shots) * len * sizeof(A), 0.0, &tl);
  // one wave per shot; at most 2^24 shots per dispatch keeps the
  // work-item count below 2^32
  for (uint64_t s0 = 0; s0 < uint64_t(shots); s0 += uint64_t(1) << 24) {
    const uint64_t ns = std::min<uint64_t>(uint64_t(shots) - s0, uint64_t(1) << 24);
    hipLaunchKernelGGL(k_sample<T>, dim3(unsigned(ceil_div(ns, kThreads / 64))),
                       dim3(kThreads), 0, s->stream, a, nchunks, len,
                       s->chunk_tot, s->chunk_inc, s->u_dev + s0, ns,
                       s->idx_dev + s0);
    if ((rc = launch_check("k_sample"))) return rc;
  }
  if ((rc = timing_end(s, &tl))) return rc;
  QCS_HIP_TRY(hipMemcpyAsync(out, s->idx_dev, size_t(shots) * sizeof(long),
                             hipMemcpyDeviceToHost, s->stream));
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

int sample(qcs_hip_state *s, long shots, const double *u, long *out) {
  return s->precision == QCS_HIP_C128 ? sample_impl<double>(s, shots, u, out)
                                      : sample_impl<float>(s, shots, u, out);
}

// ---- sampling with no allocation ---------------------------------------------
//
// qcs_hip_sample_noalloc keeps the chunk totals (tot) and inclusive prefixes
// (inc) in one file-scope device array per GPU instead of buffers grown per
// state.  Its device address is looked up once per device with that device
// current; host code passes that address to the kernels, never the array's
// name (in host code the name is a host-side stand-in).  Every entry
// [0, nchunks) of both halves is written by k_chunk_tot / k_chunk_scan before
// any kernel reads it, so nothing needs clearing between calls.
constexpr uint64_t kNoAllocChunks = uint64_t(1) << (QCS_HIP_NOALLOC_MAX_QUBITS - 12);
static_assert(kChunk == 4096, "kNoAllocChunks assumes 4096-amplitude chunks");
static __device__ __align__(256) double scratch[2 * kNoAllocChunks];

constexpr int kMaxDevices = 64;
static std::once_flag g_scratch_once[kMaxDevices];
static double *g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];

static int noalloc_scratch(int device, double **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(QCS_HIP_EUNSUP, "sample_noalloc: device %d beyond %d", device, kMaxDevices);
  std::call_once(g_scratch_once[device], [device] {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? static_cast<double *>(p) : nullptr;
  });
  if (g_scratch_err[device] != hipSuccess)
    return check_hip(g_scratch_err[device], "hipGetSymbolAddress(scratch)");
  *out = g_scratch_ptr[device];
  return 0;
}

template <typename T>
static int sample_noalloc_impl(qcs_hip_state *s, double *scr, long shots,
                               const double *u_dev, long *idx_dev) {
  typedef typename Amp<T>::type A;
  const uint64_t len = s->size < uint64_t(kChunk) ? s->size : uint64_t(kChunk);
  const uint64_t nchunks = s->size / len;
  double *tot = scr, *inc = scr + kNoAllocChunks;
  const A *a = static_cast<const A *>(s->vec);
  hipLaunchKernelGGL(k_chunk_tot<T>, dim3(unsigned(ceil_div(nchunks, kThreads / 64))),
                     dim3(kThreads), 0, s->stream, a, nchunks, len, tot);
  int rc = launch_check("k_chunk_tot");
  if (rc) return rc;
  hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, s->stream, tot, inc, nchunks);
  if ((rc = launch_check("k_chunk_scan"))) return rc;
  for (uint64_t s0 = 0; s0 < uint64_t(shots); s0 += uint64_t(1) << 24) {
    const uint64_t ns = std::min<uint64_t>(uint64_t(shots) - s0, uint64_t(1) << 24);
    hipLaunchKernelGGL(k_sample<T>, dim3(unsigned(ceil_div(ns, kThreads / 64))),
                       dim3(kThreads), 0, s->stream, a, nchunks, len, tot, inc,
                       u_dev + s0, ns, idx_dev + s0);
    if ((rc = launch_check("k_sample"))) return rc;
  }
  return 0;
}

int sample_noalloc(qcs_hip_state *s, long shots, const double *u_dev, long *idx_dev) {
  if (s->n > QCS_HIP_NOALLOC_MAX_QUBITS)
    return fail(QCS_HIP_EINVAL,
                "sample_noalloc: %d qubits exceed the scratch array's %d (%llu chunk sums)",
                s->n, QCS_HIP_NOALLOC_MAX_QUBITS, (unsigned long long)kNoAllocChunks);
  double *scr = nullptr;
  int rc = noalloc_scratch(s->device, &scr);
  if (rc) return rc;
  return s->precision == QCS_HIP_C128
             ? sample_noalloc_impl<double>(s, scr, shots, u_dev, idx_dev)
             : sample_noalloc_impl<float>(s, scr, shots, u_dev, idx_dev);
}

template <typename T>
static int grover_impl(qcs_hip_state *s, long marked, long iters) {
  typedef typename Amp<T>::type A;
  A *a = static_cast<A *>(s->vec);
  const unsigned nb = red_blocks(s->size);
  const uint64_t m = marked < 0 ? ~uint64_t(0) : uint64_t(marked);
  const double ninv = 1.0 / double(s->size);
  double *two_mu = result_slot(s);
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_REDUCE, double(s->size) * sizeof(A), 2.0 * s->size,
               &tl);
  hipLaunchKernelGGL(k_gsum_partial<T>, dim3(nb), dim3(kThreads), 0, s->stream,
                     a, s->size, m, s->red);
  int rc = launch_check("k_gsum_partial");
  if (rc) return rc;
  if ((rc = timing_end(s, &tl))) return rc;
  for (long it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_gfinal, dim3(1), dim3(kThreads), 0, s->stream, s->red,
                       int(nb), ninv, two_mu);
    if ((rc = launch_check("k_gfinal"))) return rc;
    timing_begin(s, QCS_HIP_K_OTHER, 2.0 * double(s->size) * sizeof(A),
                 6.0 * s->size, &tl);
    hipLaunchKernelGGL(k_gupdate<T>, dim3(nb), dim3(kThreads), 0, s->stream, a,
                       s->size, m, two_mu, s->red);
    if ((rc = launch_check("k_gupdate"))) return rc;
    if ((rc = timing_end(s, &tl))) return rc;
  }
  return 0;
}

int grover_textbook(qcs_hip_state *s, long marked, long iters) {
  return s->precision == QCS_HIP_C128 ? grover_impl<double>(s, marked, iters)
                                      : grover_impl<float>(s, marked, iters);
}

template <typename T>
static int flip_compat_impl(qcs_hip_state *s, long index) {
  typedef typename Amp<T>::type A;
  void *t = s->vec;  // src/q_gates.c:311-316
  s->vec = s->scratch;
  s->scratch = t;
  hipLaunchKernelGGL(k_flip_compat<T>, dim3(1), dim3(64), 0, s->stream,
                     static_cast<A *>(s->vec), static_cast<const A *>(s->scratch),
                     uint64_t(index));
  return launch_check("k_flip_compat");
}

int phase_flip_compat(qcs_hip_state *s, long index) {
  return s->precision == QCS_HIP_C128 ? flip_compat_impl<double>(s, index)
                                      : flip_compat_impl<float>(s, index);
}

template <typename T>
static int diffusion_compat_impl(qcs_hip_state *s) {
  typedef typename Amp<T>::type A;
  const unsigned nb = red_blocks(s->size);
  hipLaunchKernelGGL(k_gsum_partial<T>, dim3(nb), dim3(kThreads), 0, s->stream,
                     static_cast<const A *>(s->vec), s->size, ~uint64_t(0),
                     s->red);
  int rc = launch_check("k_gsum_partial");
  if (rc) return rc;
  double *two_mean = result_slot(s);
  hipLaunchKernelGGL(k_cfinal, dim3(1), dim3(kThreads), 0, s->stream, s->red,
                     int(nb), two_mean);
  if ((rc = launch_check("k_cfinal"))) return rc;
  void *t = s->vec;  // src/q_gates.c:346-348
  s->vec = s->scratch;
  s->scratch = t;
  hipLaunchKernelGGL(k_cdiff<T>, dim3(stream_grid(s->size)), dim3(kThreads), 0,
                     s->stream, static_cast<const A *>(s->scratch),
                     static_cast<A *>(s->vec), s->size, two_mean);
  return launch_check("k_cdiff");
}

int diffusion_compat(qcs_hip_state *s) {
  return s->precision == QCS_HIP_C128 ? diffusion_compat_impl<double>(s)
                                      : diffusion_compat_impl<float>(s);
}

int negate_one(qcs_hip_state *s, long index) {
  if (s->precision == QCS_HIP_C128)
    hipLaunchKernelGGL(k_negate<double>, dim3(1), dim3(64), 0, s->stream,
                       static_cast<double2 *>(s->vec), uint64_t(index));
  else
    hipLaunchKernelGGL(k_negate<float>, dim3(1), dim3(64), 0, s->stream,
                       static_cast<float2 *>(s->vec), uint64_t(index));
  return launch_check("k_negate");
}

template <typename TA, typename TB>
static int cmp_impl(qcs_hip_state *a, qcs_hip_state *b, unsigned nb) {
  typedef typename Amp<TA>::type AA;
  typedef typename Amp<TB>::type AB;
  hipLaunchKernelGGL((k_cmp_partial<TA, TB>), dim3(nb), dim3(kThreads), 0,
                     a->stream, static_cast<const AA *>(a->vec),
                     static_cast<const AB *>(b->vec), a->size, a->red);
  return launch_check("k_cmp_partial");
}

int compare(qcs_hip_state *a, qcs_hip_state *b, double *omf, double *maxabs) {
  // order b's pending work before a's stream reads it
  QCS_HIP_TRY(hipStreamSynchronize(b->stream));
  const unsigned nb = red_blocks(a->size);
  int rc;
  const bool a64 = a->precision == QCS_HIP_C128, b64 = b->precision == QCS_HIP_C128;
  if (a64 && b64) rc = cmp_impl<double, double>(a, b, nb);
  else if (a64) rc = cmp_impl<double, float>(a, b, nb);
  else if (b64) rc = cmp_impl<float, double>(a, b, nb);
  else rc = cmp_impl<float, float>(a, b, nb);
  if (rc) return rc;
  if ((rc = fold(a, nb, 5, 1u << 4))) return rc;
  if ((rc = fetch_results(a, 5))) return rc;
  const double dr = a->host[0], di = a->host[1], na = a->host[2],
               nbv = a->host[3];
  if (omf) *omf = (na > 0.0 && nbv > 0.0) ? 1.0 - (dr * dr + di * di) / (na * nbv) : 1.0;
  if (maxabs) *maxabs = a->host[4];
  return 0;
}

}  // namespace qcs
