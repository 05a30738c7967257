// Plain-form gate kernels: one HBM sweep per gate.
//
// 1-qubit gate on target t: pair p in [0, N/2) -> i0 = p with a 0 inserted at
// bit t, i1 = i0 | 2^t (the reference's block/step loop, src/q_gates.c:133-143,
// flattened).  Consecutive threads take consecutive pairs, so for t >= 6 both
// halves of a wave's pairs are single 1 KiB runs (c128); for low t the two
// 16-byte loads of a thread sit in the same cache lines and L2 serves the
// second touch, so HBM still moves each byte once.
//
// Controlled gates select amplitudes by index-bit masks: index q in
// [0, N/4) gets zeros inserted at the control and target bits, then the
// control bit is set; only the control=1 half is read and written.
#include "device.h"
#include "internal.h"

namespace qcs {

// One pair (or amplitude) per thread; non-temporal accesses and an
// XCD-aware work-group order (device.h) keep the in-place sweeps near the
// copy bandwidth: 5.6-6.6 TB/s for a complex128 1-qubit gate at 30 qubits,
// against 5.1-5.9 TB/s for two pairs per thread with cached accesses.
// A dispatch holds fewer than 2^32 work-items, so the grid is capped at
// kMaxGrid work-groups (a multiple of 8) and each work-group walks the
// virtual blocks vb = blockIdx.x + i * gridDim.x; the XCD order applies to
// the virtual block index, which shares blockIdx.x's XCD.
constexpr int kUnroll = 1;
constexpr uint64_t kMaxGrid = uint64_t(1) << 20;

// ---- 1-qubit gates --------------------------------------------------------

// Dense 2x2 on (a[i0], a[i1]); `in` may equal `out` (in place).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_1q_dense(
    const typename Amp<T>::type *in, typename Amp<T>::type *out,
    uint64_t npairs, int t, Mat2<T> m) {
  typedef typename Amp<T>::type A;
  const uint64_t tb = uint64_t(1) << t;
  const uint64_t nvb = ceil_div(npairs, uint64_t(kThreads) * kUnroll);
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
  const uint64_t base = xcd_order(vb, nvb) * (kThreads * kUnroll) + threadIdx.x;
  A v0[kUnroll], v1[kUnroll];
  uint64_t i0[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t p = base + uint64_t(u) * kThreads;
    i0[u] = insert0(p, t);
    if (p < npairs) {
      v0[u] = ld_stream(in + i0[u]);
      v1[u] = ld_stream(in + (i0[u] | tb));
    }
  }
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t p = base + uint64_t(u) * kThreads;
    if (p >= npairs) continue;
    A o0, o1;
    row2(m.v[0], m.v[1], m.v[2], m.v[3], v0[u].x, v0[u].y, v1[u].x, v1[u].y,
         o0.x, o0.y);
    row2(m.v[4], m.v[5], m.v[6], m.v[7], v0[u].x, v0[u].y, v1[u].x, v1[u].y,
         o1.x, o1.y);
    st_stream(out + i0[u], o0);
    st_stream(out + (i0[u] | tb), o1);
  }
  }
}

// diag(1, m11) on target t (Z, P): only the bit-t=1 half changes.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_1q_phase(
    typename Amp<T>::type *a, uint64_t nhalf, int t, T mr, T mi) {
  typedef typename Amp<T>::type A;
  const uint64_t tb = uint64_t(1) << t;
  const uint64_t nvb = ceil_div(nhalf, uint64_t(kThreads) * kUnroll);
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
  const uint64_t base = xcd_order(vb, nvb) * (kThreads * kUnroll) + threadIdx.x;
  A v[kUnroll];
  uint64_t idx[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t p = base + uint64_t(u) * kThreads;
    idx[u] = insert0(p, t) | tb;
    if (p < nhalf) v[u] = ld_stream(a + idx[u]);
  }
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t p = base + uint64_t(u) * kThreads;
    if (p >= nhalf) continue;
    A o;
    cmul(mr, mi, v[u].x, v[u].y, o.x, o.y);
    st_stream(a + idx[u], o);
  }
  }
}

// ---- controlled gates (textbook) --------------------------------------------

// index of quarter q with zeros inserted at bits lo < hi, then `set` OR-ed in
template <int KIND, typename T>
__global__ __launch_bounds__(kThreads) void k_cu(typename Amp<T>::type *a,
                                                 uint64_t nquarter, int lo,
                                                 int hi, uint64_t setb,
                                                 uint64_t tb, Mat2<T> m) {
  // KIND 0: dense 2x2 on (i0, i1); 1: swap (CNOT); 2: phase on i1 only (CP)
  typedef typename Amp<T>::type A;
  const uint64_t nvb = ceil_div(nquarter, uint64_t(kThreads) * kUnroll);
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
  const uint64_t base = xcd_order(vb, nvb) * (kThreads * kUnroll) + threadIdx.x;
  A v0[kUnroll], v1[kUnroll];
  uint64_t i0[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t q = base + uint64_t(u) * kThreads;
    i0[u] = insert0(insert0(q, lo), hi) | setb;
    if (q < nquarter) {
      if (KIND != 2) v0[u] = ld_stream(a + i0[u]);
      v1[u] = ld_stream(a + (i0[u] | tb));
    }
  }
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t q = base + uint64_t(u) * kThreads;
    if (q >= nquarter) continue;
    if (KIND == 0) {
      A o0, o1;
      row2(m.v[0], m.v[1], m.v[2], m.v[3], v0[u].x, v0[u].y, v1[u].x, v1[u].y,
           o0.x, o0.y);
      row2(m.v[4], m.v[5], m.v[6], m.v[7], v0[u].x, v0[u].y, v1[u].x, v1[u].y,
           o1.x, o1.y);
      st_stream(a + i0[u], o0);
      st_stream(a + (i0[u] | tb), o1);
    } else if (KIND == 1) {
      st_stream(a + i0[u], v1[u]);
      st_stream(a + (i0[u] | tb), v0[u]);
    } else {
      A o;
      cmul(m.v[6], m.v[7], v1[u].x, v1[u].y, o.x, o.y);
      st_stream(a + (i0[u] | tb), o);
    }
  }
  }
}

// Controlled gates and SWAP whose lowest fixed bit is so low that the
// amplitudes they change come in runs shorter than 64 bytes (control bit 0
// or 1 for complex128, 0..2 for complex64): the half-state kernels above
// would move each cache line for a fraction of its bytes (4.6x slower for a
// CNOT on control 0 at 30 qubits).  These sweep pairs along one bit the way
// k_1q_dense does -- consecutive threads take consecutive pairs, so each
// line is read and written whole (by at most two instructions) -- and leave
// the pairs the gate does not change as they are:
//   KIND 0 (controlled-U) and 1 (CNOT): pairs along the target, changed where
//     the control bit (cb) is set; both members stored.
//   KIND 2 (controlled phase): pairs along the upper bit, the upper member
//     multiplied where the lower bit (cb) is set; with BOTH the lower member
//     is stored too (when the upper bit's runs are short as well), else it is
//     neither read nor written.
//   KIND 3 (SWAP of bits lo < hi): pairs along hi; the pair whose lower member
//     has bit lo set exchanges that member with the upper member of the pair
//     in lane ^ 2^lo (lo < 6 here: the partner is in the same wave); both
//     members of every pair are stored.
template <int KIND, bool BOTH, typename T>
__global__ __launch_bounds__(kThreads) void k_pair_sweep(typename Amp<T>::type *a,
                                                         uint64_t npairs, int pb,
                                                         uint64_t cb, Mat2<T> m) {
  typedef typename Amp<T>::type A;
  const uint64_t tb = uint64_t(1) << pb;
  const uint64_t nvb = ceil_div(npairs, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t p = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    const uint64_t i0 = insert0(p, pb), i1 = i0 | tb;
    const bool on = (i0 & cb) != 0;
    if (KIND == 3) {
      // every lane of the wave takes part in the exchange (npairs is a
      // multiple of 64 whenever lo < hi < n and the state has >= 7 qubits;
      // smaller states run a single partial wave whose partners are active)
      A v0, v1;
      if (p < npairs) {
        v0 = ld_stream(a + i0);
        v1 = ld_stream(a + i1);
      }
      const int src = int((threadIdx.x & 63u) ^ unsigned(cb));
      A w0, w1;
      w0.x = __shfl(v0.x, src);
      w0.y = __shfl(v0.y, src);
      w1.x = __shfl(v1.x, src);
      w1.y = __shfl(v1.y, src);
      if (p < npairs) {
        st_stream(a + i0, on ? w1 : v0);  // (lo, hi) = (1, 0) <- (0, 1)
        st_stream(a + i1, on ? v1 : w0);  // (0, 1) <- (1, 0)
      }
      continue;
    }
    if (p >= npairs) continue;
    if (KIND == 2) {
      const A v1 = ld_stream(a + i1);
      A o = v1;
      if (on) cmul(m.v[6], m.v[7], v1.x, v1.y, o.x, o.y);
      if (BOTH) st_stream(a + i0, ld_stream(a + i0));
      st_stream(a + i1, o);
      continue;
    }
    const A v0 = ld_stream(a + i0), v1 = ld_stream(a + i1);
    A o0 = v0, o1 = v1;
    if (on) {
      if (KIND == 0) {
        row2(m.v[0], m.v[1], m.v[2], m.v[3], v0.x, v0.y, v1.x, v1.y, o0.x, o0.y);
        row2(m.v[4], m.v[5], m.v[6], m.v[7], v0.x, v0.y, v1.x, v1.y, o1.x, o1.y);
      } else {
        o0 = v1;
        o1 = v0;
      }
    }
    st_stream(a + i0, o0);
    st_stream(a + i1, o1);
  }
}

// One-qubit and controlled gates whose target bit t is so low that a pair's
// members share a 128-byte line (2^t * s < 128: t <= 2 for complex128, <= 3
// for complex64).  Taking the pairs along t, each load or store instruction
// would cover every line of its span in part, and the other part with the
// next instruction; measured at 30 qubits, complex128 1-qubit gates on
// qubits 0..2 took 6.5-6.75 ms against 5.9-6.0 ms on qubits 3..27.  Here
// lane l of a wave holds amplitude base + l, so every access instruction
// moves 64 consecutive amplitudes, and the partner's value comes from lane
// l ^ 2^t (t < 6) by a cross-lane read.  The result is the one k_1q_dense
// (or k_pair_sweep) computes, with the same operand order:
//   KIND 0: m applied where the control bits cm are all set (cm = 0: always);
//   KIND 1: the members exchanged where they are (CNOT).
// Every amplitude is read and written once (2 N s).
template <int KIND, typename T>
__global__ __launch_bounds__(kThreads) void k_lane_pairs(typename Amp<T>::type *a,
                                                         uint64_t n, int t, uint64_t cm,
                                                         Mat2<T> m) {
  typedef typename Amp<T>::type A;
  const uint64_t nvb = ceil_div(n, uint64_t(kThreads));
  const int src = int((threadIdx.x & 63u) ^ (1u << t));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t e = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    // every lane takes part in the exchange; a lane past the end (states of
    // fewer than 8 qubits) is never a partner of one inside, since 2^t < n
    A v;
    v.x = 0;
    v.y = 0;
    if (e < n) v = ld_stream(a + e);
    A w;
    w.x = __shfl(v.x, src);
    w.y = __shfl(v.y, src);
    if (e >= n) continue;
    A o = v;
    if ((e & cm) == cm) {
      if (KIND == 1)
        o = w;
      else if ((e >> t) & 1)
        row2(m.v[4], m.v[5], m.v[6], m.v[7], w.x, w.y, v.x, v.y, o.x, o.y);
      else
        row2(m.v[0], m.v[1], m.v[2], m.v[3], v.x, v.y, w.x, w.y, o.x, o.y);
    }
    st_stream(a + e, o);
  }
}

// true when a pair along bit t lies within one 128-byte line
static bool lane_pairs(int t, size_t amp_bytes) { return (size_t(1) << t) * amp_bytes < 128; }

// true when the runs of a half-state sweep with lowest fixed bit `b` are
// shorter than a 64-byte line
static bool short_runs(int b, size_t amp_bytes) { return (size_t(1) << b) * amp_bytes < 64; }

// HBM bytes a kernel moves when it reads and writes the amplitudes whose
// index bits in `fixed` are all 1: HBM moves whole 128-byte lines, so only
// the fixed bits above a line's index bits thin out the lines touched
static double line_bytes(double n, size_t amp_bytes, uint64_t fixed) {
  const int in_line = __builtin_ctzll(uint64_t(128) / amp_bytes);
  return 2.0 * n * double(amp_bytes) /
         double(uint64_t(1) << __builtin_popcountll(fixed >> in_line));
}

// Compat controlled gate, out of place over all N (src/q_gates.c:277-292 as
// written): out[i] = m00*v[i] + m01*v[i|t] where bit c = 1 and bit t = 0,
// out[i] = v[i] everywhere else (iteration i|t overwrites the partner).  One
// work-item per (bit t = 0, bit t = 1) pair, so each amplitude is read and
// written once.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cu_compat(
    const typename Amp<T>::type *in, typename Amp<T>::type *out, uint64_t npairs,
    uint64_t cb, int t, Mat2<T> m) {
  typedef typename Amp<T>::type A;
  const uint64_t tb = uint64_t(1) << t;
  const uint64_t nvb = ceil_div(npairs, uint64_t(kThreads));
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    const uint64_t p = xcd_order(vb, nvb) * kThreads + threadIdx.x;
    if (p >= npairs) continue;
    const uint64_t i0 = insert0(p, t);
    const A v0 = ld_stream(in + i0), v1 = ld_stream(in + (i0 | tb));
    A o = v0;
    if (i0 & cb)
      row2(m.v[0], m.v[1], m.v[2], m.v[3], v0.x, v0.y, v1.x, v1.y, o.x, o.y);
    st_stream(out + i0, o);
    st_stream(out + (i0 | tb), v1);
  }
}

// SWAP of qubits a < b: exchange amplitudes with (bit a, bit b) = (1,0), (0,1)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_swapq(typename Amp<T>::type *a,
                                                    uint64_t nquarter, int lo,
                                                    int hi) {
  typedef typename Amp<T>::type A;
  const uint64_t bl = uint64_t(1) << lo, bh = uint64_t(1) << hi;
  const uint64_t nvb = ceil_div(nquarter, uint64_t(kThreads) * kUnroll);
  for (uint64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
  const uint64_t base = xcd_order(vb, nvb) * (kThreads * kUnroll) + threadIdx.x;
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const uint64_t q = base + uint64_t(u) * kThreads;
    if (q >= nquarter) continue;
    const uint64_t i = insert0(insert0(q, lo), hi);
    const A x = ld_stream(a + (i | bl));
    const A y = ld_stream(a + (i | bh));
    st_stream(a + (i | bl), y);
    st_stream(a + (i | bh), x);
  }
  }
}

// Fill with zeros and optionally set a[basis] = 1.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fill(typename Amp<T>::type *a,
                                                   uint64_t n, uint64_t basis,
                                                   int one) {
  typedef typename Amp<T>::type A;
  const uint64_t stride = uint64_t(gridDim.x) * kThreads;
  for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += stride) {
    A v;
    v.x = (one && i == basis) ? T(1) : T(0);
    v.y = T(0);
    a[i] = v;
  }
}

// ---- host launchers ------------------------------------------------------------

template <typename T>
static Mat2<T> to_mat(const double m[8]) {
  Mat2<T> r;
  for (int i = 0; i < 8; ++i) r.v[i] = T(m[i]);
  return r;
}

bool is_phase1(const double m[8]) {
  // diag(1, x): m00 = 1, m01 = m10 = 0
  return m[0] == 1.0 && m[1] == 0.0 && m[2] == 0.0 && m[3] == 0.0 &&
         m[4] == 0.0 && m[5] == 0.0;
}

bool is_x(const double m[8]) {
  return m[0] == 0.0 && m[1] == 0.0 && m[2] == 1.0 && m[3] == 0.0 &&
         m[4] == 1.0 && m[5] == 0.0 && m[6] == 0.0 && m[7] == 0.0;
}

static unsigned grid_for(uint64_t items) {
  const uint64_t g = ceil_div(items, uint64_t(kThreads) * kUnroll);
  return unsigned(g < kMaxGrid ? g : kMaxGrid);
}

template <typename T>
static int do_1q(qcs_hip_state *s, int t, const double md[8]) {
  typedef typename Amp<T>::type A;
  const Mat2<T> m = to_mat<T>(md);
  const bool compat = (s->flags & QCS_HIP_COMPAT) != 0;
  A *cur = static_cast<A *>(s->vec);
  const double N = double(s->size), sb = double(sizeof(A));
  TimedLaunch tl;
  if (!compat && is_phase1(md)) {
    const uint64_t nhalf = s->size / 2;
    timing_begin(s, QCS_HIP_K_1Q, N * sb, 3.0 * N, &tl,
                 line_bytes(N, sizeof(A), uint64_t(1) << t));
    hipLaunchKernelGGL(k_1q_phase<T>, dim3(grid_for(nhalf)), dim3(kThreads), 0,
                       s->stream, cur, nhalf, t, m.v[6], m.v[7]);
  } else if (!compat && lane_pairs(t, sizeof(A))) {
    timing_begin(s, QCS_HIP_K_1Q, 2.0 * N * sb, 14.0 * N, &tl);
    hipLaunchKernelGGL((k_lane_pairs<0, T>), dim3(grid_for(s->size)), dim3(kThreads), 0,
                       s->stream, cur, s->size, t, uint64_t(0), m);
  } else {
    A *out = compat ? static_cast<A *>(s->scratch) : cur;
    const uint64_t npairs = s->size / 2;
    timing_begin(s, QCS_HIP_K_1Q, 2.0 * N * sb, 14.0 * N, &tl);
    hipLaunchKernelGGL(k_1q_dense<T>, dim3(grid_for(npairs)), dim3(kThreads), 0,
                       s->stream, cur, out, npairs, t, m);
    if (compat) {  // src/q_gates.c:146-148
      s->scratch = s->vec;
      s->vec = out;
    }
  }
  int rc = launch_check("k_1q");
  if (rc) return rc;
  return timing_end(s, &tl);
}

template <typename T>
static int do_cu(qcs_hip_state *s, int c, int t, const double md[8]) {
  typedef typename Amp<T>::type A;
  const Mat2<T> m = to_mat<T>(md);
  A *cur = static_cast<A *>(s->vec);
  const uint64_t cb = uint64_t(1) << c, tb = uint64_t(1) << t;
  const double N = double(s->size), sb = double(sizeof(A));
  TimedLaunch tl;
  if (s->flags & QCS_HIP_COMPAT) {
    A *out = static_cast<A *>(s->scratch);
    timing_begin(s, QCS_HIP_K_CU, 2.0 * N * sb, 0.0, &tl);
    hipLaunchKernelGGL(k_cu_compat<T>, dim3(grid_for(s->size / 2)), dim3(kThreads),
                       0, s->stream, cur, out, s->size / 2, cb, t, m);
    s->scratch = s->vec;
    s->vec = out;
  } else {
    const int lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t nq = s->size / 4;
    const dim3 g(grid_for(nq)), b(kThreads);
    const bool x = is_x(md), ph = !x && is_phase1(md);
    const double mv = line_bytes(N, sizeof(A), ph ? (cb | tb) : cb);
    // the lowest bit fixed in the changed amplitudes: the control, or for a
    // controlled phase the lower of the two
    if (short_runs(ph ? lo : c, sizeof(A))) {
      timing_begin(s, QCS_HIP_K_CU, (ph ? 0.5 : 1.0) * N * sb, 0.0, &tl, mv);
      const uint64_t np = s->size / 2;
      const dim3 gp(grid_for(np));
      if (!ph && lane_pairs(t, sizeof(A))) {
        const dim3 ge(grid_for(s->size));
        if (x)
          hipLaunchKernelGGL((k_lane_pairs<1, T>), ge, b, 0, s->stream, cur, s->size, t, cb, m);
        else
          hipLaunchKernelGGL((k_lane_pairs<0, T>), ge, b, 0, s->stream, cur, s->size, t, cb, m);
      } else if (x)
        hipLaunchKernelGGL((k_pair_sweep<1, false, T>), gp, b, 0, s->stream, cur, np, t, cb, m);
      else if (!ph)
        hipLaunchKernelGGL((k_pair_sweep<0, false, T>), gp, b, 0, s->stream, cur, np, t, cb, m);
      else if (short_runs(hi, sizeof(A)))
        hipLaunchKernelGGL((k_pair_sweep<2, true, T>), gp, b, 0, s->stream, cur, np, hi,
                           uint64_t(1) << lo, m);
      else
        hipLaunchKernelGGL((k_pair_sweep<2, false, T>), gp, b, 0, s->stream, cur, np, hi,
                           uint64_t(1) << lo, m);
    } else if (x) {
      timing_begin(s, QCS_HIP_K_CU, N * sb, 0.0, &tl, mv);
      hipLaunchKernelGGL((k_cu<1, T>), g, b, 0, s->stream, cur, nq, lo, hi, cb,
                         tb, m);
    } else if (ph) {
      timing_begin(s, QCS_HIP_K_CU, 0.5 * N * sb, 1.5 * N, &tl, mv);
      hipLaunchKernelGGL((k_cu<2, T>), g, b, 0, s->stream, cur, nq, lo, hi,
                         cb | tb, uint64_t(0), m);
    } else {
      timing_begin(s, QCS_HIP_K_CU, N * sb, 7.0 * N, &tl, mv);
      hipLaunchKernelGGL((k_cu<0, T>), g, b, 0, s->stream, cur, nq, lo, hi, cb,
                         tb, m);
    }
  }
  int rc = launch_check("k_cu");
  if (rc) return rc;
  return timing_end(s, &tl);
}

int launch_1q(qcs_hip_state *s, int target, const double m[8]) {
  return s->precision == QCS_HIP_C128 ? do_1q<double>(s, target, m)
                                      : do_1q<float>(s, target, m);
}

int launch_cu(qcs_hip_state *s, int control, int target, const double m[8]) {
  return s->precision == QCS_HIP_C128 ? do_cu<double>(s, control, target, m)
                                      : do_cu<float>(s, control, target, m);
}

int launch_swap(qcs_hip_state *s, int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const uint64_t nq = s->size / 4;
  TimedLaunch tl;
  const bool sweep = short_runs(lo, size_t(s->amp_bytes));
  // the exchanged amplitudes have bit lo = 1 or bit hi = 1 but not both:
  // every line holds some when lo is inside a line
  const bool lo_in_line = (size_t(1) << lo) * size_t(s->amp_bytes) < 128;
  timing_begin(s, QCS_HIP_K_CU, double(s->size) * double(s->amp_bytes), 0.0, &tl,
               (lo_in_line ? 2.0 : 1.0) * double(s->size) * double(s->amp_bytes));
  const uint64_t np = s->size / 2;
  if (s->precision == QCS_HIP_C128) {
    if (sweep)
      hipLaunchKernelGGL((k_pair_sweep<3, false, double>), dim3(grid_for(np)), dim3(kThreads),
                         0, s->stream, static_cast<double2 *>(s->vec), np, hi,
                         uint64_t(1) << lo, Mat2<double>());
    else
      hipLaunchKernelGGL(k_swapq<double>, dim3(grid_for(nq)), dim3(kThreads), 0,
                         s->stream, static_cast<double2 *>(s->vec), nq, lo, hi);
  } else {
    if (sweep)
      hipLaunchKernelGGL((k_pair_sweep<3, false, float>), dim3(grid_for(np)), dim3(kThreads),
                         0, s->stream, static_cast<float2 *>(s->vec), np, hi,
                         uint64_t(1) << lo, Mat2<float>());
    else
      hipLaunchKernelGGL(k_swapq<float>, dim3(grid_for(nq)), dim3(kThreads), 0,
                         s->stream, static_cast<float2 *>(s->vec), nq, lo, hi);
  }
  int rc = launch_check("k_swapq");
  if (rc) return rc;
  return timing_end(s, &tl);
}

int launch_fill_basis(qcs_hip_state *s, void *buf, uint64_t basis, int one) {
  const uint64_t blocks = ceil_div(s->size, kThreads);
  const unsigned g = unsigned(blocks < 65536 ? blocks : 65536);
  if (s->precision == QCS_HIP_C128)
    hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kThreads), 0, s->stream,
                       static_cast<double2 *>(buf), s->size, basis, one);
  else
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kThreads), 0, s->stream,
                       static_cast<float2 *>(buf), s->size, basis, one);
  return launch_check("k_fill");
}

}  // namespace qcs
