// Batched circuits (qcs_hip_batch_*): thousands of small independent
// circuits in one persistent launch.
//
// Work distribution.  The host orders the circuits by estimated cost,
// largest first; each work-group of a persistent grid (as many work-groups
// as are resident at once) takes the next circuit from a device counter, so
// the long circuits start first and the short ones fill in behind them.
// Every wave reaches the exit test once the counter passes the end.
//
// State placement.  A circuit whose 2^n amplitudes fit in the LDS tile
// (n <= tile_bits: up to 2^13 complex128 or 2^14 complex64 amplitudes, 128
// KiB) lives in LDS from |0...0> to its last shot; HBM sees only the
// uniforms in and the sampled indices out.  A larger circuit keeps its state
// in an HBM slot and streams it through LDS pass by pass (a set S of
// tile_bits index bits with the ops on those bits) and tile by tile (the
// 2^(n - tile_bits) tiles of amplitudes that agree outside S).  In the first
// pass only tile 0 holds a non-zero amplitude; the others are written as
// zeros without running the ops (every op is linear).
//
// complex128 (plain FP64): the ops are the circuit's gates in order, each
// one sweep over the tile with the arithmetic and gate classes of the
// per-gate kernels in gates.hip, so the amplitudes equal those of a state of
// its own bit for bit.  complex64 (fused FP32): the ops are composed k-qubit
// blocks applied on the matrix cores with the operand layout of
// k_fused_mfma (blockpass.hip) on an XOR-swizzled tile (the layout of
// tiles.hip, with one more fold for 14-bit tiles), then FP64
// renormalisation.
//
// Sampling is qcs_hip_sample's algorithm inside one work-group: FP64 sums of
// 4096-amplitude chunks (rows of 64, a wave prefix scan per row, rows
// chained sequentially), a sequential scan of the chunk sums, and per shot a
// binary search over the chunks and a scan inside the chunk -- the same
// operations in the same order as reduce.hip, so the same state gives the
// same indices.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <atomic>
#include <thread>
#include <vector>

#include "device.h"
#include "internal.h"

namespace qcs {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float bf4v __attribute__((ext_vector_type(4)));

constexpr int kBMaxTb = 14;
constexpr int kBMaxN = QCS_HIP_BATCH_MAX_QUBITS;
constexpr size_t kBAlign = 256;

enum : int {
  kOp1Dense = 0,  // 2x2 on (i0, i1)
  kOp1Phase = 1,  // diag(1, m11): bit-1 half only
  kOpCuDense = 2, // 2x2 on control=1 pairs
  kOpCnot = 3,    // exchange on control=1 pairs
  kOpCPhase = 4,  // m11 on control=1, target=1
  kOpSwap = 5,    // exchange (1,0) <-> (0,1) of qubits t, c
  // register groups only: the dense 2x2 without the products by exact-zero
  // entries, which add only signed zeros -- the results compare equal to
  // kOp1Dense's (they may differ in the sign of a zero)
  kOp1Diag = 6,   // m01 = m10 = 0 (RZ)
  kOp1Real = 7,   // every imaginary part 0 (H, RY)
  kOp1Anti = 8,   // m00 = m11 = 0 (X, Y)
  // three gates on one position: diag(a, conj a), then the real rotation
  // [[c, -s], [s, c]], then diag(b, conj b) (RZ, RY, RZ), in the op's m as
  // {Re a, Im a, c, -s, Re b, Im b}; applied one after the other with the
  // arithmetic of kOp1Diag, kOp1Real, kOp1Diag
  kOp1Euler = 9
};

// Register groups (complex128): consecutive gates of a pass acting on at
// most kGR tile bits are applied by one sweep in which every thread holds
// the kGA amplitudes that differ only in those bits.
constexpr int kGR = 4;
constexpr int kGA = 1 << kGR;

struct BCirc {
  int n;
  int prec;
  int tb;
  int npass;
  int first_pass;
  int reserved;
  uint64_t state_off;  // byte offset of the HBM slot, kNoSlot when none
  uint64_t shot_off;   // first shot in u / out
};
constexpr uint64_t kNoSlot = ~uint64_t(0);

// One queue item: a circuit, or one of the `nparts` parts of a streamed
// complex128 circuit that several work-groups share (part j takes tiles
// j, j + nparts, ... of every pass; the parts meet between passes).
struct BItem {
  int circ;
  int part;
  int nparts;
  int reserved;
};
// spins of a part's wait at a pass boundary before it gives up (each spin
// sleeps ~1000 clocks: several seconds, far beyond any pass)
constexpr unsigned kPartSpinLimit = 1u << 23;

struct BPass {
  int first_op;
  int nops;
  int sbits[kBMaxTb];  // tile bit j -> index bit, ascending
};

struct BGate {
  int cls;
  int t;  // tile positions
  int c;
  int reserved;
  double m[8];
};

struct BGroup {
  int first_gate;      // gates with t, c as positions 0..kGR-1 in qpos
  int ngates;
  int qpos[kGR];       // tile positions of the group, ascending
  int rowslot[kGA];    // LDS slot (dswz) of the row pattern j
  int outslot[kGA];    // slot register j is stored to: rowslot[j], or a
                       // permutation of it that applies the group's
                       // trailing CNOTs and SWAPs
};

struct BBlock {
  int k;
  int wslot;
  int tpos[6];        // tile positions of the targets, ascending
  int rowslot[32];    // byte offset 8 * bswz(row bits) of row s
};

// XOR swizzle of a complex64 tile element (float2 slot); linear over GF(2)
// and a bijection of every [0, 2^tb), tb >= 5.
__host__ __device__ __forceinline__ int bswz(int e) {
  return e ^ ((e >> 5) & 31) ^ ((e >> 10) & 15);
}

// XOR swizzle of a complex128 tile element (double2 slot): bits 4..7 fold
// onto bits 0..3, so the 16 lanes of a ds_read_b128 quarter-wave that step
// through a group's columns hit distinct banks for most bit layouts.  Linear
// over GF(2), a bijection of every [0, 2^tb), the identity below 16.
__host__ __device__ __forceinline__ int dswz(int e) {
  return e ^ ((e >> 4) & 15);
}

__device__ __forceinline__ int bins0(int x, int b) {
  return ((x >> b) << (b + 1)) | (x & ((1 << b) - 1));
}

// A pass's tile positions -> index bits, ascending, held in registers (the
// same in every thread) so that the address arithmetic of a tile reads no
// memory; entries at j >= tb are never used.
struct TileBits {
  int sb[kBMaxTb];
};

__device__ __forceinline__ TileBits load_tile_bits(const int *src) {
  TileBits b;
#pragma unroll
  for (int j = 0; j < kBMaxTb; ++j) b.sb[j] = src[j];
  return b;
}

// first index of tile t: t's bits spread over the index bits outside S
__device__ __forceinline__ uint64_t tile_base(uint64_t t, const TileBits &b, int tb) {
#pragma unroll
  for (int j = 0; j < kBMaxTb; ++j)
    if (j < tb) t = insert0(t, b.sb[j]);
  return t;
}

// Index bits outside the tile, of an n-qubit state.  Tile first indices
// are t's bits spread over these, so the first index of tile t + d is
// ((base(t) | ~M) + base(d)) & M: the carries run through the bits of M.
__device__ __forceinline__ uint64_t outside_mask(const TileBits &b, int tb, int n) {
  uint64_t m = (uint64_t(1) << n) - 1;
#pragma unroll
  for (int j = 0; j < kBMaxTb; ++j)
    if (j < tb) m &= ~(uint64_t(1) << b.sb[j]);
  return m;
}

__device__ __forceinline__ uint64_t masked_add(uint64_t x, uint64_t y, uint64_t m) {
  return ((x | ~m) + y) & m;
}

// index bits of tile element e: bit j of e goes to index bit sb[j], for
// j in [J0, min(J1, tb)).  With e the same in every thread (the element
// bits above the thread bits) this is scalar arithmetic.
template <int J0, int J1>
__device__ __forceinline__ uint64_t deposit(uint32_t e, const TileBits &b, int tb) {
  uint64_t a = 0;
#pragma unroll
  for (int j = J0; j < J1; ++j)
    if (j < tb && ((e >> j) & 1u)) a |= uint64_t(1) << b.sb[j];
  return a;
}

// ---- complex128: gates on an LDS tile -------------------------------------------

// (tiles of fewer than kGR bits; there dswz is the identity)
template <int NTH>
__device__ void tile_gate(double2 *L, int tb, const BGate *__restrict__ gp) {
  const int cls = gp->cls, t = gp->t, c = gp->c;
  double m[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) m[i] = gp->m[i];
  const int tid = threadIdx.x;
  if (cls == kOp1Dense || cls == kOp1Phase) {
    const int tbit = 1 << t, np = 1 << (tb - 1);
    if (cls == kOp1Dense) {
      for (int p = tid; p < np; p += NTH) {
        const int i0 = bins0(p, t), i1 = i0 | tbit;
        const double2 v0 = L[i0], v1 = L[i1];
        double2 o0, o1;
        row2(m[0], m[1], m[2], m[3], v0.x, v0.y, v1.x, v1.y, o0.x, o0.y);
        row2(m[4], m[5], m[6], m[7], v0.x, v0.y, v1.x, v1.y, o1.x, o1.y);
        L[i0] = o0;
        L[i1] = o1;
      }
    } else {
      for (int p = tid; p < np; p += NTH) {
        const int i1 = bins0(p, t) | tbit;
        const double2 v = L[i1];
        double2 o;
        cmul(m[6], m[7], v.x, v.y, o.x, o.y);
        L[i1] = o;
      }
    }
    return;
  }
  const int lo = t < c ? t : c, hi = t < c ? c : t;
  const int nq = 1 << (tb - 2);
  if (cls == kOpSwap) {
    const int bl = 1 << lo, bh = 1 << hi;
    for (int q = tid; q < nq; q += NTH) {
      const int i = bins0(bins0(q, lo), hi);
      const double2 x = L[i | bl], y = L[i | bh];
      L[i | bl] = y;
      L[i | bh] = x;
    }
    return;
  }
  const int cb = 1 << c, tbit = 1 << t;
  for (int q = tid; q < nq; q += NTH) {
    const int i0 = bins0(bins0(q, lo), hi) | cb, i1 = i0 | tbit;
    if (cls == kOpCuDense) {
      const double2 v0 = L[i0], v1 = L[i1];
      double2 o0, o1;
      row2(m[0], m[1], m[2], m[3], v0.x, v0.y, v1.x, v1.y, o0.x, o0.y);
      row2(m[4], m[5], m[6], m[7], v0.x, v0.y, v1.x, v1.y, o1.x, o1.y);
      L[i0] = o0;
      L[i1] = o1;
    } else if (cls == kOpCnot) {
      const double2 v0 = L[i0], v1 = L[i1];
      L[i0] = v1;
      L[i1] = v0;
    } else {
      const double2 v = L[i1];
      double2 o;
      cmul(m[6], m[7], v.x, v.y, o.x, o.y);
      L[i1] = o;
    }
  }
}

// ---- complex128: register groups ----------------------------------------------
//
// v[j] holds the amplitude whose group bits are the bits of j (bit i of j =
// tile bit qpos[i]); a gate on group positions t (and c) pairs registers at
// compile-time indices, with the arithmetic of gates.hip.

// The gate class and positions are template parameters, so each variant is
// straight-line register code; r_gate dispatches once per gate.
template <int Q, int CLS>
__device__ __forceinline__ void r_1q(double (&vr)[kGA], double (&vi)[kGA], const double *m) {
#pragma unroll
  for (int j = 0; j < kGA; ++j) {
    if (j & (1 << Q)) continue;
    const int k = j | (1 << Q);
    const double x0 = vr[j], y0 = vi[j], x1 = vr[k], y1 = vi[k];
    // the products and sums of cmul / row2 (gates.hip), every product of
    // the pair before any sum: each result is then defined after the last
    // use of the inputs and takes an input's register, so every gate
    // variant leaves each amplitude in the register it came in (no copies
    // at the join after the dispatch switch)
    if constexpr (CLS == kOp1Dense) {
      const double a0 = m[0] * x0, a1 = m[1] * y0, a2 = m[0] * y0, a3 = m[1] * x0;
      const double b0 = m[4] * x0, b1 = m[5] * y0, b2 = m[4] * y0, b3 = m[5] * x0;
      const double c0 = m[2] * x1, c1 = m[3] * y1, c2 = m[2] * y1, c3 = m[3] * x1;
      const double d0 = m[6] * x1, d1 = m[7] * y1, d2 = m[6] * y1, d3 = m[7] * x1;
      vr[j] = (a0 - a1) + (c0 - c1);
      vi[j] = (a2 + a3) + (c2 + c3);
      vr[k] = (b0 - b1) + (d0 - d1);
      vi[k] = (b2 + b3) + (d2 + d3);
    } else if constexpr (CLS == kOp1Phase) {
      const double p0 = m[6] * x1, p1 = m[7] * y1, p2 = m[6] * y1, p3 = m[7] * x1;
      vr[k] = p0 - p1;
      vi[k] = p2 + p3;
    } else if constexpr (CLS == kOp1Diag) {
      const double a0 = m[0] * x0, a1 = m[1] * y0, a2 = m[0] * y0, a3 = m[1] * x0;
      const double d0 = m[6] * x1, d1 = m[7] * y1, d2 = m[6] * y1, d3 = m[7] * x1;
      vr[j] = a0 - a1;
      vi[j] = a2 + a3;
      vr[k] = d0 - d1;
      vi[k] = d2 + d3;
    } else if constexpr (CLS == kOp1Real) {
      const double a0 = m[0] * x0, a1 = m[0] * y0, b0 = m[4] * x0, b1 = m[4] * y0;
      const double c0 = m[2] * x1, c1 = m[2] * y1, d0 = m[6] * x1, d1 = m[6] * y1;
      vr[j] = a0 + c0;
      vi[j] = a1 + c1;
      vr[k] = b0 + d0;
      vi[k] = b1 + d1;
    } else if constexpr (CLS == kOp1Euler) {
      // the entries the three matrices repeat are negations and copies of
      // the ones held: -m7 = m1 and m6 = m0 (diagonal), m4 = -m2 and m6 = m0
      // (rotation), so every product below is the one the separate ops form
      double u0 = x0, v0 = y0, u1 = x1, v1 = y1;
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const double ar = m[4 * d], ai = m[4 * d + 1], nai = -ai;
        if (d == 1) {
          const double c = m[2], s = m[3], ns = -s;
          const double a0 = c * u0, a1 = c * v0, b0 = ns * u0, b1 = ns * v0;
          const double c0 = s * u1, c1 = s * v1, d0 = c * u1, d1 = c * v1;
          u0 = a0 + c0;
          v0 = a1 + c1;
          u1 = b0 + d0;
          v1 = b1 + d1;
        }
        const double a0 = ar * u0, a1 = ai * v0, a2 = ar * v0, a3 = ai * u0;
        const double d0 = ar * u1, d1 = nai * v1, d2 = ar * v1, d3 = nai * u1;
        u0 = a0 - a1;
        v0 = a2 + a3;
        u1 = d0 - d1;
        v1 = d2 + d3;
      }
      vr[j] = u0;
      vi[j] = v0;
      vr[k] = u1;
      vi[k] = v1;
    } else {  // kOp1Anti
      const double c0 = m[2] * x1, c1 = m[3] * y1, c2 = m[2] * y1, c3 = m[3] * x1;
      const double b0 = m[4] * x0, b1 = m[5] * y0, b2 = m[4] * y0, b3 = m[5] * x0;
      vr[j] = c0 - c1;
      vi[j] = c2 + c3;
      vr[k] = b0 - b1;
      vi[k] = b2 + b3;
    }
  }
}

// Exchange two register values in place: each keeps its register, so a
// permutation gate (CNOT, SWAP) leaves every group amplitude in the register
// it had, and the gate variants of the dispatch switch agree on where each
// amplitude lives (otherwise the join after the switch needs register
// copies on every gate).  v_swap_b32 is an ALU exchange of two VGPRs.
__device__ __forceinline__ void swap_in_place(double &a, double &b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  uint32_t alo = uint32_t(ua), ahi = uint32_t(ua >> 32), blo = uint32_t(ub), bhi = uint32_t(ub >> 32);
  asm volatile("v_swap_b32 %0, %1" : "+v"(alo), "+v"(blo));
  asm volatile("v_swap_b32 %0, %1" : "+v"(ahi), "+v"(bhi));
  a = __builtin_bit_cast(double, (uint64_t(ahi) << 32) | alo);
  b = __builtin_bit_cast(double, (uint64_t(bhi) << 32) | blo);
}

// Controlled gates: pairs (j, j | T) with bit C of j set.  SWAP: pairs
// (j, j ^ C ^ T) with (bit C, bit T) of j = (1, 0).
template <int C, int T, int CLS>
__device__ __forceinline__ void r_cu(double (&vr)[kGA], double (&vi)[kGA], const double *m) {
#pragma unroll
  for (int j = 0; j < kGA; ++j) {
    if (!(j & (1 << C)) || (j & (1 << T))) continue;
    if constexpr (CLS == kOpSwap) {
      const int k = (j ^ (1 << C)) | (1 << T);
      swap_in_place(vr[j], vr[k]);
      swap_in_place(vi[j], vi[k]);
    } else {
      const int k = j | (1 << T);
      const double x0 = vr[j], y0 = vi[j], x1 = vr[k], y1 = vi[k];
      if constexpr (CLS == kOpCuDense) {
        const double a0 = m[0] * x0, a1 = m[1] * y0, a2 = m[0] * y0, a3 = m[1] * x0;
        const double b0 = m[4] * x0, b1 = m[5] * y0, b2 = m[4] * y0, b3 = m[5] * x0;
        const double c0 = m[2] * x1, c1 = m[3] * y1, c2 = m[2] * y1, c3 = m[3] * x1;
        const double d0 = m[6] * x1, d1 = m[7] * y1, d2 = m[6] * y1, d3 = m[7] * x1;
        vr[j] = (a0 - a1) + (c0 - c1);
        vi[j] = (a2 + a3) + (c2 + c3);
        vr[k] = (b0 - b1) + (d0 - d1);
        vi[k] = (b2 + b3) + (d2 + d3);
      } else if constexpr (CLS == kOpCnot) {
        swap_in_place(vr[j], vr[k]);
        swap_in_place(vi[j], vi[k]);
      } else {  // kOpCPhase
        const double p0 = m[6] * x1, p1 = m[7] * y1, p2 = m[6] * y1, p3 = m[7] * x1;
        vr[k] = p0 - p1;
        vi[k] = p2 + p3;
      }
    }
  }
}

template <int CLS>
__device__ __forceinline__ void r_1q_at(double (&vr)[kGA], double (&vi)[kGA], int t, const double *m) {
  switch (t) {
    case 0: r_1q<0, CLS>(vr, vi, m); break;
    case 1: r_1q<1, CLS>(vr, vi, m); break;
    case 2: r_1q<2, CLS>(vr, vi, m); break;
    default: r_1q<3, CLS>(vr, vi, m); break;
  }
}

template <int CLS>
__device__ __forceinline__ void r_cu_at(double (&vr)[kGA], double (&vi)[kGA], int c, int t,
                                        const double *m) {
  switch (c * kGR + t) {
    case 1: r_cu<0, 1, CLS>(vr, vi, m); break;
    case 2: r_cu<0, 2, CLS>(vr, vi, m); break;
    case 3: r_cu<0, 3, CLS>(vr, vi, m); break;
    case 4: r_cu<1, 0, CLS>(vr, vi, m); break;
    case 6: r_cu<1, 2, CLS>(vr, vi, m); break;
    case 7: r_cu<1, 3, CLS>(vr, vi, m); break;
    case 8: r_cu<2, 0, CLS>(vr, vi, m); break;
    case 9: r_cu<2, 1, CLS>(vr, vi, m); break;
    case 11: r_cu<2, 3, CLS>(vr, vi, m); break;
    case 12: r_cu<3, 0, CLS>(vr, vi, m); break;
    case 13: r_cu<3, 1, CLS>(vr, vi, m); break;
    default: r_cu<3, 2, CLS>(vr, vi, m); break;
  }
}

__device__ __forceinline__ void r_gate(double (&vr)[kGA], double (&vi)[kGA], const BGate *__restrict__ g) {
  double m[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) m[i] = g->m[i];
  const int t = g->t, c = g->c;
  switch (g->cls) {
    case kOp1Dense: r_1q_at<kOp1Dense>(vr, vi, t, m); break;
    case kOp1Phase: r_1q_at<kOp1Phase>(vr, vi, t, m); break;
    case kOp1Diag: r_1q_at<kOp1Diag>(vr, vi, t, m); break;
    case kOp1Real: r_1q_at<kOp1Real>(vr, vi, t, m); break;
    case kOp1Anti: r_1q_at<kOp1Anti>(vr, vi, t, m); break;
    case kOp1Euler: r_1q_at<kOp1Euler>(vr, vi, t, m); break;
    case kOpCuDense: r_cu_at<kOpCuDense>(vr, vi, c, t, m); break;
    case kOpCnot: r_cu_at<kOpCnot>(vr, vi, c, t, m); break;
    case kOpCPhase: r_cu_at<kOpCPhase>(vr, vi, c, t, m); break;
    default: r_cu_at<kOpSwap>(vr, vi, c, t, m); break;
  }
}

// One group on a tile of 2^tb (>= kGR) amplitudes: thread `col` takes the
// column of group-bit patterns at the other tile bits = col.
template <int NTH>
__device__ void tile_group(double2 *L, int tb, const BGroup *__restrict__ gp,
                           const BGate *__restrict__ gates) {
  const int ncols = 1 << (tb - kGR);
  const int q0 = gp->qpos[0], q1 = gp->qpos[1], q2 = gp->qpos[2], q3 = gp->qpos[3];
  const int ng = gp->ngates;
  const BGate *g0 = gates + gp->first_gate;
  int rs[kGA];
#pragma unroll
  for (int j = 0; j < kGA; ++j) rs[j] = gp->rowslot[j];
  for (int col = threadIdx.x; col < ncols; col += NTH) {
    const int base = dswz(bins0(bins0(bins0(bins0(col, q0), q1), q2), q3));
    double vr[kGA], vi[kGA];
#pragma unroll
    for (int j = 0; j < kGA; ++j) {
      const double2 x = L[base ^ rs[j]];
      vr[j] = x.x;
      vi[j] = x.y;
    }
    for (int g = 0; g < ng; ++g) r_gate(vr, vi, g0 + g);
#pragma unroll
    for (int j = 0; j < kGA; ++j) L[base ^ gp->outslot[j]] = make_double2(vr[j], vi[j]);
  }
}

// complex128 tile <-> HBM: element e <-> G[base | deposit(e)].  A thread
// moves the elements e = tid (mod NTH), through LDS slot dswz(e) (the same
// slots in both directions, so a store phase followed by the next tile's
// load needs no barrier between them).  Loads go out up to 16 per thread at
// a time: a 13-bit tile (128 KiB) in one round trip at 512 threads (the
// register-group values are dead here, so the registers are free).
template <int U, int NTH, bool LOAD>
__device__ __forceinline__ int io_rounds_c128(double2 *L, double2 *G, uint64_t dt,
                                              const TileBits &b, int tb, int e0) {
  constexpr int LT = __builtin_ctz(NTH);
  const int tid = threadIdx.x, tile = 1 << tb;
  constexpr int LU = __builtin_ctz(U);
  static_assert(LT + LU <= kBMaxTb, "tile bit index in range");
  // element e0 + u * NTH + tid: bits below LT are tid's (in dt), bits LT ..
  // LT + LU - 1 are u's, the bits above are e0's (a multiple of U * NTH)
  uint64_t bit[LU];
#pragma unroll
  for (int k = 0; k < LU; ++k) bit[k] = uint64_t(1) << (b.sb[LT + k] & 63);
  for (; e0 + U * NTH <= tile; e0 += U * NTH) {
    const uint64_t d0 = dt | deposit<LT + LU, kBMaxTb>(uint32_t(e0), b, tb);
    uint64_t a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t du = 0;
#pragma unroll
      for (int k = 0; k < LU; ++k)
        if ((u >> k) & 1) du |= bit[k];
      a[u] = d0 | du;
    }
    if (LOAD) {
      double2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = G[a[u]];
#pragma unroll
      for (int u = 0; u < U; ++u) L[dswz(e0 + u * NTH + tid)] = v[u];
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) G[a[u]] = L[dswz(e0 + u * NTH + tid)];
    }
  }
  return e0;
}

template <int NTH, bool LOAD>
__device__ void tile_io_c128(double2 *L, double2 *G, uint64_t base, const TileBits &b, int tb) {
  constexpr int LT = __builtin_ctz(NTH);
  const int tid = threadIdx.x, tile = 1 << tb;
  const uint64_t dt = base | deposit<0, LT>(uint32_t(tid), b, tb);
  int e0 = io_rounds_c128<16, NTH, LOAD>(L, G, dt, b, tb, 0);
  e0 = io_rounds_c128<4, NTH, LOAD>(L, G, dt, b, tb, e0);
  for (; e0 < tile; e0 += NTH) {
    const int e = e0 + tid;
    if (e < tile) {
      const uint64_t a = dt | deposit<LT, kBMaxTb>(uint32_t(e0), b, tb);
      if (LOAD) L[dswz(e)] = G[a];
      else G[a] = L[dswz(e)];
    }
  }
}

// ---- complex64: fused blocks on an LDS tile --------------------------------------

// One composed k-qubit block on the swizzled tile: rows are the 2^k patterns
// of the target bits, columns the other tile bits; each wave takes 32
// columns at a time, W (A operand) in registers, the tile column as B.
template <int K, int NTH>
__device__ void tile_block(char *Lb, int tb, const BBlock *__restrict__ bp,
                           const float *__restrict__ wbuf) {
  constexpr int R = 1 << K, RR = 2 * R, MT = (RR + 31) / 32, NW = NTH / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  const float *W = wbuf + size_t(bp->wslot) * kWSlotFloats;
  float wa[MT][R];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int row = mt * 32 + l32;
#pragma unroll
    for (int s = 0; s < R; ++s)
      wa[mt][s] = (row < RR) ? W[row * RR + 2 * s + h] : 0.0f;
  }
  int tpos[K], rs[R];
#pragma unroll
  for (int q = 0; q < K; ++q) tpos[q] = bp->tpos[q];
#pragma unroll
  for (int s = 0; s < R; ++s) rs[s] = bp->rowslot[s];
  // slots of the output rows of this lane: accumulators 4q, 4q+1 hold row
  // mt*16 + 4q + 2h, 4q+2, 4q+3 the next row (lane-dependent, so loaded per
  // lane rather than picked from a private array)
  int os0[MT][4], os1[MT][4];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r0 = mt * 16 + 4 * q + 2 * h;
      os0[mt][q] = r0 < R ? bp->rowslot[r0] : 0;
      os1[mt][q] = r0 + 1 < R ? bp->rowslot[r0 + 1] : 0;
    }
  const int ncols = 1 << (tb - K);
  for (int c0 = wave * 32; c0 < ncols; c0 += NW * 32) {
    const int cc = c0 + l32;
    const bool ok = cc < ncols;
    int col = ok ? cc : 0;
#pragma unroll
    for (int q = 0; q < K; ++q) col = bins0(col, tpos[q]);
    const int cs = 8 * bswz(col), csh = cs + 4 * h;
    float b[R];
#pragma unroll
    for (int s = 0; s < R; ++s)
      b[s] = ok ? *reinterpret_cast<const float *>(Lb + (csh ^ rs[s])) : 0.0f;
    floatx16 acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      acc[mt] = floatx16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                         0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < R; ++s)
        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[mt][s], b[s], acc[mt], 0, 0, 0);
    }
    if (!ok) continue;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ra = mt * 16 + 4 * q;
        if (ra >= R) continue;
        const int r0 = ra + 2 * h;
        if (r0 < R)
          *reinterpret_cast<float2 *>(Lb + (cs ^ os0[mt][q])) =
              make_float2(acc[mt][4 * q], acc[mt][4 * q + 1]);
        if (r0 + 1 < R)
          *reinterpret_cast<float2 *>(Lb + (cs ^ os1[mt][q])) =
              make_float2(acc[mt][4 * q + 2], acc[mt][4 * q + 3]);
      }
  }
}

template <int NTH>
__device__ void tile_block_any(char *Lb, int tb, const BBlock *__restrict__ bp,
                               const float *__restrict__ wbuf) {
  switch (bp->k) {
    case 1: tile_block<1, NTH>(Lb, tb, bp, wbuf); break;
    case 2: tile_block<2, NTH>(Lb, tb, bp, wbuf); break;
    case 3: tile_block<3, NTH>(Lb, tb, bp, wbuf); break;
    case 4: tile_block<4, NTH>(Lb, tb, bp, wbuf); break;
    default: tile_block<5, NTH>(Lb, tb, bp, wbuf); break;
  }
}

// complex64 tile <-> HBM in element pairs (e, e + 1), e even: tile bit 0 is
// index bit 0, so a pair is one 16-byte access, and swizzled slots
// bswz(e), bswz(e) ^ 1.
template <int NTH, bool LOAD>
__device__ __forceinline__ void pair_io_c64(char *Lb, float2 *G, uint64_t a,
                                            int p) {
  float2 *s0 = reinterpret_cast<float2 *>(Lb + 8 * bswz(2 * p));
  float2 *s1 = reinterpret_cast<float2 *>(Lb + 8 * (bswz(2 * p) ^ 1));
  if (LOAD) {
    const bf4v v = *reinterpret_cast<const bf4v *>(G + a);
    *s0 = make_float2(v.x, v.y);
    *s1 = make_float2(v.z, v.w);
  } else {
    const float2 x = *s0, y = *s1;
    const bf4v w = {x.x, x.y, y.x, y.y};
    *reinterpret_cast<bf4v *>(G + a) = w;
  }
}

// A thread moves the pairs p = tid (mod NTH) through the same LDS slots in
// both directions; loads go out up to 16 per thread at a time (a 14-bit
// tile in one round trip at 512 threads).
template <int U, int NTH, bool LOAD>
__device__ __forceinline__ int io_rounds_c64(char *Lb, float2 *G, uint64_t dt,
                                             const TileBits &b, int tb, int p0) {
  constexpr int LT = __builtin_ctz(NTH);
  const int tid = threadIdx.x, npair = 1 << (tb - 1);
  constexpr int LU = __builtin_ctz(U);
  static_assert(1 + LT + LU <= kBMaxTb, "tile bit index in range");
  // pair p0 + u * NTH + tid holds elements with bits 1 .. LT from tid (in
  // dt), bits LT + 1 .. LT + LU from u, the bits above from p0 (a multiple
  // of U * NTH)
  uint64_t bit[LU];
#pragma unroll
  for (int k = 0; k < LU; ++k) bit[k] = uint64_t(1) << (b.sb[1 + LT + k] & 63);
  for (; p0 + U * NTH <= npair; p0 += U * NTH) {
    const uint64_t d0 = dt | deposit<1 + LT + LU, kBMaxTb>(uint32_t(p0) << 1, b, tb);
    uint64_t a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint64_t du = 0;
#pragma unroll
      for (int k = 0; k < LU; ++k)
        if ((u >> k) & 1) du |= bit[k];
      a[u] = d0 | du;
    }
    if (LOAD) {
      bf4v v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const bf4v *>(G + a[u]);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int pp = p0 + u * NTH + tid;
        float2 *s0 = reinterpret_cast<float2 *>(Lb + 8 * bswz(2 * pp));
        float2 *s1 = reinterpret_cast<float2 *>(Lb + 8 * (bswz(2 * pp) ^ 1));
        *s0 = make_float2(v[u].x, v[u].y);
        *s1 = make_float2(v[u].z, v[u].w);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) pair_io_c64<NTH, false>(Lb, G, a[u], p0 + u * NTH + tid);
    }
  }
  return p0;
}

template <int NTH, bool LOAD>
__device__ void tile_io_c64(char *Lb, float2 *G, uint64_t base, const TileBits &b, int tb) {
  constexpr int LT = __builtin_ctz(NTH);
  const int tid = threadIdx.x, npair = 1 << (tb - 1);
  // tile bits 1..LT from tid (those below tb)
  const uint64_t dt = base | deposit<1, 1 + LT>(uint32_t(tid) << 1, b, tb);
  int p0 = io_rounds_c64<16, NTH, LOAD>(Lb, G, dt, b, tb, 0);
  p0 = io_rounds_c64<4, NTH, LOAD>(Lb, G, dt, b, tb, p0);
  for (; p0 < npair; p0 += NTH) {
    const int p = p0 + tid;
    if (p < npair)
      pair_io_c64<NTH, LOAD>(Lb, G, dt | deposit<1 + LT, kBMaxTb>(uint32_t(p0) << 1, b, tb), p);
  }
}

// ---- FP64 norm and sampling ------------------------------------------------------

// Sum over the work-group, broadcast to every thread.
template <int NTH>
__device__ double wg_sum(double v, double *red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < NTH / 64; ++w) r += red[w];
  __syncthreads();
  return r;
}

// qcs_hip_sample on one state; prob(i) = |a_i|^2 in FP64.
template <int NTH, class P>
__device__ void sample_state(const P &prob, int n, double *tot, double *inc,
                             const double *__restrict__ u,
                             long *__restrict__ out, uint64_t shots) {
  constexpr int NW = NTH / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t N = uint64_t(1) << n;
  const uint64_t len = N < uint64_t(kChunk) ? N : uint64_t(kChunk);
  const uint64_t nch = N / len;
  for (uint64_t ch = wave; ch < nch; ch += NW) {
    double running = 0.0;
    for (uint64_t r = 0; r < len; r += 64) {
      const double x = (r + lane < len) ? prob(ch * len + r + lane) : 0.0;
      const double s = wave_incl_scan(x, lane);
      running = running + __shfl(s, 63, 64);
    }
    if (lane == 0) tot[ch] = running;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (uint64_t ch = 0; ch < nch; ++ch) {
      const double t = tot[ch];
      tot[ch] = run;
      run = run + t;
      inc[ch] = run;
    }
  }
  __syncthreads();
  for (uint64_t s = wave; s < shots; s += NW) {
    const double us = u[s];
    uint64_t lo = 0, hi = nch;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (inc[mid] > us) hi = mid; else lo = mid + 1;
    }
    if (lo >= nch) {
      if (lane == 0) out[s] = -1;
      continue;
    }
    const uint64_t c = lo;
    const double e = tot[c];
    double running = 0.0;
    long found = -1, last_nz = -1;
    for (uint64_t r = 0; r < len; r += 64) {
      const double x = (r + lane < len) ? prob(c * len + r + lane) : 0.0;
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
    if (found < 0) found = last_nz;
    if (lane == 0) out[s] = found;
  }
  __syncthreads();
}

// ---- one circuit per work-group, or one part of a shared circuit ------------------

// Pass boundary of the work-groups that share a circuit.  Every thread's
// tile stores are made visible device-wide (agent-scope release: across
// the XCDs' L2s) before the work-group adds itself to the circuit's
// counter; thread 0 then waits for `target` arrivals and acquires.  The wait
// is bounded: on timeout, or when another part has timed out, the error flag
// is set and every wave of the work-group returns false.
__device__ bool parts_barrier(unsigned *cnt, unsigned target, unsigned *err) {
  __shared__ int ok_s;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    atomicAdd(cnt, 1u);
    for (unsigned spins = 0;
         __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          ++spins > kPartSpinLimit) {
        atomicOr(err, 1u);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(16);
    }
    __threadfence();
    ok_s = ok;
  }
  __syncthreads();
  return ok_s != 0;
}

template <int NTH>
__device__ void run_c128(const BCirc c, int part, int nparts, unsigned *sync,
                         unsigned *err, const BPass *__restrict__ passes,
                         const BGate *__restrict__ gates,
                         const BGroup *__restrict__ groups, char *states,
                         char *lds, double *tot, double *inc, double *red,
                         const double *__restrict__ u, long *__restrict__ out,
                         uint64_t shots, int keep) {
  double2 *L = reinterpret_cast<double2 *>(lds);
  const int n = c.n, tb = c.tb, tile = 1 << tb, tid = threadIdx.x;
  const bool resident = tb == n;
  double2 *G = c.state_off == kNoSlot
                   ? nullptr
                   : reinterpret_cast<double2 *>(states + c.state_off);
  const uint64_t ntiles = uint64_t(1) << (n - tb);
  for (int p = 0; p < c.npass; ++p) {
    const BPass &ps = passes[c.first_pass + p];
    const TileBits tbits = load_tile_bits(ps.sbits);
    // a pass reads tiles the previous pass wrote, possibly in another part
    if (nparts > 1 && p > 0 && !parts_barrier(sync, unsigned(p) * unsigned(nparts), err))
      return;
    const uint64_t outside = outside_mask(tbits, tb, n);
    const uint64_t bstep = tile_base(uint64_t(nparts), tbits, tb);
    uint64_t base = tile_base(uint64_t(part), tbits, tb);
    for (uint64_t t = uint64_t(part); t < ntiles;
         t += uint64_t(nparts), base = masked_add(base, bstep, outside)) {
      if (p == 0 || resident) {
        if (p == 0)
          for (int e = tid; e < tile; e += NTH)
            L[e] = make_double2((t == 0 && e == 0) ? 1.0 : 0.0, 0.0);
      } else {
        tile_io_c128<NTH, true>(L, G, base, tbits, tb);
      }
      __syncthreads();
      if (p > 0 || t == 0) {
        for (int o = 0; o < ps.nops; ++o) {
          if (tb >= kGR)
            tile_group<NTH>(L, tb, &groups[ps.first_op + o], gates);
          else
            tile_gate<NTH>(L, tb, &gates[ps.first_op + o]);
          __syncthreads();
        }
      }
      if (!resident) {
        tile_io_c128<NTH, false>(L, G, base, tbits, tb);
        // the next tile of this pass loads into the slots each thread has
        // just stored from; the first pass's tiles are set up by element
        // instead, and the next pass reads what other threads stored
        if (p == 0 || t + uint64_t(nparts) >= ntiles) __syncthreads();
      }
    }
  }
  // the last pass done by every part; part 0 samples the state
  if (nparts > 1 && !parts_barrier(sync, unsigned(c.npass) * unsigned(nparts), err)) return;
  if (part != 0) return;
  if (resident && keep && G) {
    for (int e = tid; e < tile; e += NTH) G[e] = L[dswz(e)];
    __syncthreads();
  }
  (void)red;
  if (shots == 0) return;
  if (resident) {
    sample_state<NTH>([&](uint64_t i) { return norm2(L[dswz(int(i))]); }, n, tot, inc,
                      u + c.shot_off, out + c.shot_off, shots);
  } else {
    const double2 *Gc = G;
    sample_state<NTH>([&](uint64_t i) { return norm2(Gc[i]); }, n, tot, inc,
                      u + c.shot_off, out + c.shot_off, shots);
  }
}

template <int NTH>
__device__ void run_c64(const BCirc c, const BPass *__restrict__ passes,
                        const BBlock *__restrict__ blocks,
                        const float *__restrict__ wbuf, char *states, char *lds,
                        double *tot, double *inc, double *red,
                        const double *__restrict__ u, long *__restrict__ out,
                        uint64_t shots, int keep) {
  const int n = c.n, tb = c.tb, tile = 1 << tb, tid = threadIdx.x;
  const bool resident = tb == n;
  float2 *L = reinterpret_cast<float2 *>(lds);
  float2 *G = c.state_off == kNoSlot
                  ? nullptr
                  : reinterpret_cast<float2 *>(states + c.state_off);
  const uint64_t ntiles = uint64_t(1) << (n - tb);
  // tiles of fewer than 32 elements are not swizzled (bswz needs tb >= 5)
  for (int p = 0; p < c.npass; ++p) {
    const BPass &ps = passes[c.first_pass + p];
    const TileBits tbits = load_tile_bits(ps.sbits);
    const uint64_t outside = outside_mask(tbits, tb, n);
    const uint64_t bstep = outside & (~outside + 1);  // tile 1
    uint64_t base = 0;
    for (uint64_t t = 0; t < ntiles; ++t, base = masked_add(base, bstep, outside)) {
      if (p == 0 || resident) {
        if (p == 0)
          for (int e = tid; e < tile; e += NTH)
            L[e] = make_float2((t == 0 && e == 0) ? 1.0f : 0.0f, 0.0f);
      } else {
        tile_io_c64<NTH, true>(lds, G, base, tbits, tb);
      }
      __syncthreads();
      if (p > 0 || t == 0) {
        for (int o = 0; o < ps.nops; ++o) {
          tile_block_any<NTH>(lds, tb, &blocks[ps.first_op + o], wbuf);
          __syncthreads();
        }
      }
      if (!resident) {
        tile_io_c64<NTH, false>(lds, G, base, tbits, tb);
        // as in run_c128: within a pass each thread reloads its own slots
        if (p == 0 || t + 1 >= ntiles) __syncthreads();
      }
    }
  }
  // FP64 renormalisation, q_state_normalize's rule (src/q_utils.c:112-118)
  const uint64_t N = uint64_t(1) << n;
  double ns = 0.0;
  if (resident)
    for (int e = tid; e < tile; e += NTH) ns += norm2(L[e]);
  else
    for (uint64_t i = tid; i < N; i += NTH) ns += norm2(G[i]);
  ns = wg_sum<NTH>(ns, red);
  if (ns > 1e-12 && ns != 1.0) {
    const double f = 1.0 / sqrt(ns);
    if (resident) {
      for (int e = tid; e < tile; e += NTH) {
        float2 x = L[e];
        x.x = float(double(x.x) * f);
        x.y = float(double(x.y) * f);
        L[e] = x;
      }
    } else {
      for (uint64_t i = tid; i < N; i += NTH) {
        float2 x = G[i];
        x.x = float(double(x.x) * f);
        x.y = float(double(x.y) * f);
        G[i] = x;
      }
    }
    __syncthreads();
  }
  if (resident && keep && G) {
    for (int e = tid; e < tile; e += NTH) G[e] = L[bswz(e)];
    __syncthreads();
  }
  if (shots == 0) return;
  if (resident) {
    sample_state<NTH>([&](uint64_t i) { return norm2(L[bswz(int(i))]); }, n,
                      tot, inc, u + c.shot_off, out + c.shot_off, shots);
  } else {
    const float2 *Gc = G;
    sample_state<NTH>([&](uint64_t i) { return norm2(Gc[i]); }, n, tot, inc,
                      u + c.shot_off, out + c.shot_off, shots);
  }
}

// The persistent kernel.  `queue[0]` starts at 0 (k_batch_reset); each
// work-group takes queue items until the counter passes nitems.  The parts
// of a shared circuit are consecutive items, at most half the grid, and
// every work-group is resident, so the parts a waiting part needs are
// always taken by work-groups that make progress.  queue[1] is the error
// flag of the part barriers; sync[c] counts circuit c's pass arrivals.
template <int NTH>
__global__ __launch_bounds__(NTH) void k_batch(
    const BCirc *__restrict__ circs, const BItem *__restrict__ items, int nitems,
    const BPass *__restrict__ passes, const BGate *__restrict__ gates,
    const BGroup *__restrict__ groups, const BBlock *__restrict__ blocks,
    const float *__restrict__ wbuf, char *states, const double *__restrict__ u,
    long *__restrict__ out, uint64_t shots, unsigned *queue, unsigned *sync,
    int tile_bytes, int nch_cap, int keep) {
  extern __shared__ double4 smem4[];
  char *lds = reinterpret_cast<char *>(smem4);
  double *tot = reinterpret_cast<double *>(lds + tile_bytes);
  double *inc = tot + nch_cap;
  __shared__ double red[NTH / 64];
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = int(atomicAdd(queue, 1u));
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();
    if (i >= nitems) break;
    const BItem it = items[i];
    const BCirc c = circs[it.circ];
    if (c.prec == QCS_HIP_C128)
      run_c128<NTH>(c, it.part, it.nparts, sync + it.circ, queue + 1, passes, gates,
                    groups, states, lds, tot, inc, red, u, out, shots, keep);
    else
      run_c64<NTH>(c, passes, blocks, wbuf, states, lds, tot, inc, red, u,
                   out, shots, keep);
  }
}

// queue counter, error flag and the circuits' pass counters to zero (a
// kernel rather than a memset, so a captured graph replays it reliably)
__global__ void k_batch_reset(unsigned *queue, unsigned *sync, int nc) {
  if (blockIdx.x == 0 && threadIdx.x < 2) queue[threadIdx.x] = 0u;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x)
    sync[i] = 0u;
}

// ---- host side ---------------------------------------------------------------------

// byte offsets of the regions of the descriptor block (device and pinned
// staging alike)
struct BatchLayout {
  size_t circ, order, sync, pass, gate, group, block, hgate, hblock, u, total;
};

// one host thread's share of a run's device ops (gates, register groups,
// blocks) and the passes it built
struct BatchPart {
  std::vector<BGate> g;
  std::vector<BGroup> grp;
  std::vector<BBlock> blk;
  std::vector<int> own;
};

}  // namespace qcs

struct qcs_hip_batch {
  int flags = 0;
  int device = 0;
  int nth = 512;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  // device buffers, grown on demand
  void *d_desc = nullptr;  // descriptors, compose inputs and uniforms
  size_t desc_cap = 0;
  float *d_w = nullptr;
  size_t w_cap = 0;  // blocks
  char *d_states = nullptr;
  size_t states_cap = 0;
  long *d_out = nullptr;
  size_t out_cap = 0;
  unsigned *d_queue = nullptr;  // [0] queue counter, [1] part-barrier error flag
  unsigned *h_flag = nullptr;   // pinned copy of the error flag
  // pinned staging (same layout as d_desc) and output
  char *h_stage = nullptr;
  size_t stage_cap = 0;
  long *h_out = nullptr;
  size_t hout_cap = 0;
  long generation = 0;  // bumped whenever a buffer moves
  std::vector<qcs::BatchPart> parts;  // host op lists, reused between runs
  // captured launch sequence
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<long> graph_key;
  hipGraph_t rerun_graph = nullptr;  // the same run with new uniforms only
  hipGraphExec_t rerun_exec = nullptr;
  // launch of the last successful run, repeated by qcs_hip_batch_rerun
  bool has_last = false;
  qcs::BatchLayout last_layout = {};
  int last_nc = 0, last_nitems = 0, last_tile_bytes = 0, last_nch_cap = 0, last_grid = 0;
  long last_shots = 0;
  // last run, for read_state and timing
  std::vector<int> last_n, last_prec;
  std::vector<uint64_t> last_off;
  float kernel_ms = -1.0f, total_ms = -1.0f;
  int replayed = 0;
};

namespace qcs {

static size_t align_up(size_t x) { return (x + kBAlign - 1) / kBAlign * kBAlign; }

static void drop_graph(qcs_hip_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->rerun_exec) (void)hipGraphExecDestroy(b->rerun_exec);
  if (b->rerun_graph) (void)hipGraphDestroy(b->rerun_graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->rerun_exec = nullptr;
  b->rerun_graph = nullptr;
  b->graph_key.clear();
}

static int grow_device(qcs_hip_batch *b, void **p, size_t *cap, size_t need) {
  if (*cap >= need) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  ++b->generation;
  QCS_HIP_TRY(hipMalloc(p, need));
  *cap = need;
  return 0;
}

static int grow_pinned(qcs_hip_batch *b, void **p, size_t *cap, size_t need) {
  if (*cap >= need) return 0;
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  *cap = 0;
  ++b->generation;
  QCS_HIP_TRY(hipHostMalloc(p, need, hipHostMallocDefault));
  *cap = need;
  return 0;
}

static int max_tb(int prec) {
  return prec == QCS_HIP_C128 ? QCS_HIP_BATCH_C128_TILE_BITS
                              : QCS_HIP_BATCH_C64_TILE_BITS;
}

// every argument check of qcs_hip_batch_run; fills the op counts
static int host_threads() {
  int t = int(std::thread::hardware_concurrency());
  const char *env = std::getenv("QCS_HOST_THREADS");
  if (env && std::atoi(env) > 0) t = std::atoi(env);
  return std::max(1, std::min(t, 64));
}

// circuit i of a batch: qubits, precision, tile size, its passes and their
// gates or blocks (0, or QCS_HIP_EINVAL with this thread's message set)
static int validate_circuit(int i, const qcs_hip_batch_circuit *cs,
                            const qcs_hip_pass *passes, int np,
                            const qcs_hip_block *blocks, int nb,
                            const qcs_hip_gate *gates, int ng) {
  const qcs_hip_batch_circuit &c = cs[i];
  if (c.num_qubits < 1 || c.num_qubits > kBMaxN)
    return fail(QCS_HIP_EINVAL, "circuit %d: num_qubits %d outside [1,%d]", i,
                c.num_qubits, kBMaxN);
  if (c.precision != QCS_HIP_C128 && c.precision != QCS_HIP_C64)
    return fail(QCS_HIP_EINVAL, "circuit %d: unknown precision %d", i, c.precision);
  const int n = c.num_qubits, tb = c.tile_bits;
  if (tb < 1 || tb > n || tb > max_tb(c.precision))
    return fail(QCS_HIP_EINVAL, "circuit %d: tile_bits %d outside [1,%d]", i, tb,
                std::min(n, max_tb(c.precision)));
  if (c.precision == QCS_HIP_C64 && tb < 5 && tb != n)
    return fail(QCS_HIP_EINVAL, "circuit %d: complex64 tiles of < 5 bits must hold the state", i);
  if (c.num_passes < 1 || c.first_pass < 0 || c.first_pass + c.num_passes > np)
    return fail(QCS_HIP_EINVAL, "circuit %d: pass range outside the pass array", i);
  for (int p = c.first_pass; p < c.first_pass + c.num_passes; ++p) {
    const qcs_hip_pass &ps = passes[p];
    if (ps.tile_bits != tb)
      return fail(QCS_HIP_EINVAL, "pass %d: tile_bits %d != circuit's %d", p,
                  ps.tile_bits, tb);
    for (int j = 0; j < tb; ++j) {
      if (ps.bits[j] < 0 || ps.bits[j] >= n || (j > 0 && ps.bits[j] <= ps.bits[j - 1]))
        return fail(QCS_HIP_EINVAL, "pass %d: bits not ascending in range", p);
    }
    if (c.precision == QCS_HIP_C64 && ps.bits[0] != 0)
      return fail(QCS_HIP_EINVAL, "pass %d: complex64 tiles must hold bit 0", p);
    unsigned long long mask = 0;
    for (int j = 0; j < tb; ++j) mask |= 1ull << ps.bits[j];
    if (ps.first_block < 0 || ps.num_blocks < 0)
      return fail(QCS_HIP_EINVAL, "pass %d: bad op range", p);
    if (c.precision == QCS_HIP_C128) {
      if (ps.first_block + ps.num_blocks > ng)
        return fail(QCS_HIP_EINVAL, "pass %d: gate range outside the gate array", p);
      for (int g = ps.first_block; g < ps.first_block + ps.num_blocks; ++g) {
        const qcs_hip_gate &gt = gates[g];
        if (gt.kind < QCS_HIP_G_1Q || gt.kind > QCS_HIP_G_SWAP || gt.target < 0 ||
            gt.target >= n || !((mask >> gt.target) & 1))
          return fail(QCS_HIP_EINVAL, "pass %d: gate %d outside the pass bits", p, g);
        if (gt.kind != QCS_HIP_G_1Q &&
            (gt.control < 0 || gt.control >= n || gt.control == gt.target ||
             !((mask >> gt.control) & 1)))
          return fail(QCS_HIP_EINVAL, "pass %d: gate %d: bad control", p, g);
      }
    } else {
      if (ps.first_block + ps.num_blocks > nb)
        return fail(QCS_HIP_EINVAL, "pass %d: block range outside the block array", p);
      int rc = validate_blocks(n, blocks + ps.first_block, ps.num_blocks, gates, ng);
      if (rc) return rc;
      for (int bi = ps.first_block; bi < ps.first_block + ps.num_blocks; ++bi) {
        if (blocks[bi].k > tb)
          return fail(QCS_HIP_EINVAL, "block %d: k > tile_bits", bi);
        for (int q = 0; q < blocks[bi].k; ++q)
          if (!((mask >> blocks[bi].targets[q]) & 1))
            return fail(QCS_HIP_EINVAL, "block %d: target outside the pass bits", bi);
      }
    }
  }
  return 0;
}

static int validate_batch(const qcs_hip_batch_circuit *cs, int nc,
                          const qcs_hip_pass *passes, int np,
                          const qcs_hip_block *blocks, int nb,
                          const qcs_hip_gate *gates, int ng, long shots,
                          const double *u, long *out) {
  if (nc < 0 || np < 0 || nb < 0 || ng < 0 || shots < 0)
    return fail(QCS_HIP_EINVAL, "negative count");
  if ((nc > 0 && !cs) || (np > 0 && !passes) || (nb > 0 && !blocks) ||
      (ng > 0 && !gates))
    return fail(QCS_HIP_EINVAL, "null array");
  if (shots > 0 && nc > 0 && (!u || !out))
    return fail(QCS_HIP_EINVAL, "null uniforms or output");
  if (shots > (long(1) << 40) / (nc > 0 ? nc : 1))
    return fail(QCS_HIP_EINVAL, "too many shots");
  // circuits in chunks on host threads; the first failing circuit (in
  // order) is checked again on this thread, whose message the caller reads
  constexpr int kChunkCircuits = 32;
  const int nt = std::max(1, std::min(nc / kChunkCircuits, host_threads()));
  std::atomic<int> next{0}, first_bad{nc};
  auto work = [&]() {
    for (;;) {
      const int c0 = next.fetch_add(kChunkCircuits);
      if (c0 >= nc || c0 > first_bad.load()) break;
      for (int i = c0; i < std::min(nc, c0 + kChunkCircuits); ++i)
        if (validate_circuit(i, cs, passes, np, blocks, nb, gates, ng)) {
          int cur = first_bad.load();
          while (i < cur && !first_bad.compare_exchange_weak(cur, i)) {
          }
          break;
        }
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
  }
  const int bad = first_bad.load();
  return bad < nc ? validate_circuit(bad, cs, passes, np, blocks, nb, gates, ng) : 0;
}

static int position_in(const qcs_hip_pass &ps, int q) {
  for (int j = 0; j < ps.tile_bits; ++j)
    if (ps.bits[j] == q) return j;
  return -1;
}

// grouped: the gate runs in a register group, where the dense class splits
// further (diagonal, real, anti-diagonal); elsewhere it keeps the classes of
// the per-gate kernels
static int gate_class(const qcs_hip_gate &g, bool grouped) {
  const double *m = g.m;
  if (g.kind == QCS_HIP_G_SWAP) return kOpSwap;
  if (g.kind == QCS_HIP_G_CU)
    return is_x(m) ? kOpCnot : (is_phase1(m) ? kOpCPhase : kOpCuDense);
  if (is_phase1(m)) return kOp1Phase;
  if (!grouped) return kOp1Dense;
  if (m[2] == 0.0 && m[3] == 0.0 && m[4] == 0.0 && m[5] == 0.0) return kOp1Diag;
  if (m[1] == 0.0 && m[3] == 0.0 && m[5] == 0.0 && m[7] == 0.0) return kOp1Real;
  if (m[0] == 0.0 && m[1] == 0.0 && m[6] == 0.0 && m[7] == 0.0) return kOp1Anti;
  return kOp1Dense;
}

static BGate make_bgate(const qcs_hip_pass &ps, const qcs_hip_gate &g) {
  BGate o;
  std::memset(&o, 0, sizeof(o));
  o.t = position_in(ps, g.target);
  o.c = g.kind == QCS_HIP_G_1Q ? 0 : position_in(ps, g.control);
  std::memcpy(o.m, g.m, sizeof(o.m));
  o.cls = gate_class(g, false);
  return o;
}

static bool same_bits(double a, double b) {
  uint64_t x, y;
  std::memcpy(&x, &a, sizeof(x));
  std::memcpy(&y, &b, sizeof(y));
  return x == y;
}
// diag(a, conj a) and [[c, -s], [s, c]], entry for entry (bit patterns): the
// repeated entries of an RZ and an RY
static bool z_like(const BGate &o) {
  return o.cls == kOp1Diag && same_bits(o.m[6], o.m[0]) && same_bits(o.m[7], -o.m[1]);
}
static bool y_like(const BGate &o) {
  return o.cls == kOp1Real && same_bits(o.m[6], o.m[0]) && same_bits(o.m[4], -o.m[2]);
}

static bool is_perm(const BGate &o) { return o.cls == kOpCnot || o.cls == kOpSwap; }
static unsigned op_positions(const BGate &o) {
  return (1u << o.t) | (o.cls >= kOpCuDense && o.cls <= kOpSwap ? 1u << o.c : 0u);
}

// The register index map of a permutation op: after it, register x holds
// what register perm_src(o, x) held.
static int perm_src(const BGate &o, int x) {
  const int bc = (x >> o.c) & 1, bt = (x >> o.t) & 1;
  if (o.cls == kOpCnot) return bc ? x ^ (1 << o.t) : x;
  // SWAP: exchange bits c and t
  return bc == bt ? x : x ^ (1 << o.c) ^ (1 << o.t);
}

// CNOTs and SWAPs move amplitudes without arithmetic, so they commute
// exactly with ops on other positions: each is moved past the later ops it
// shares no position with, and the run that ends the group is applied by
// storing the registers to permuted slots, outslot[F(j)] = rowslot[j] with
// F(j) the register whose value ends at row j.  Returns the ops left.
static std::vector<BGate> fold_trailing_perms(std::vector<BGate> ops, const int *rowslot,
                                              int *outslot) {
  const size_t n = ops.size();
  for (bool moved = true; moved;) {
    moved = false;
    for (size_t i = n; i-- > 1;) {
      BGate &a = ops[i - 1], &b = ops[i];
      if (is_perm(a) && !is_perm(b) && !(op_positions(a) & op_positions(b))) {
        std::swap(a, b);
        moved = true;
      }
    }
  }
  size_t k = n;
  while (k > 0 && is_perm(ops[k - 1])) --k;
  for (int j = 0; j < kGA; ++j) {
    int x = j;
    for (size_t i = n; i-- > k;) x = perm_src(ops[i], x);
    outslot[x] = rowslot[j];
  }
  ops.resize(k);
  return ops;
}

// One group's ops, with every run z-like, y-like, z-like on one position
// made one kOp1Euler op (one dispatch on the device instead of three).
static void emit_group_ops(const std::vector<BGate> &ops, std::vector<BGate> *bg) {
  for (size_t x = 0; x < ops.size();) {
    if (x + 2 < ops.size() && z_like(ops[x]) && y_like(ops[x + 1]) && z_like(ops[x + 2]) &&
        ops[x + 1].t == ops[x].t && ops[x + 2].t == ops[x].t) {
      BGate e;
      std::memset(&e, 0, sizeof(e));
      e.cls = kOp1Euler;
      e.t = ops[x].t;
      e.m[0] = ops[x].m[0];
      e.m[1] = ops[x].m[1];
      e.m[2] = ops[x + 1].m[0];
      e.m[3] = ops[x + 1].m[2];
      e.m[4] = ops[x + 2].m[0];
      e.m[5] = ops[x + 2].m[1];
      bg->push_back(e);
      x += 3;
    } else {
      bg->push_back(ops[x++]);
    }
  }
}

// Register groups of one complex128 pass (tile_bits >= kGR): maximal runs of
// consecutive gates on at most kGR tile positions, padded with the lowest
// unused positions.
static void make_groups(const qcs_hip_pass &ps, const qcs_hip_gate *gates,
                        std::vector<BGroup> *groups, std::vector<BGate> *bg) {
  int pos[64];
  for (int j = 0; j < ps.tile_bits; ++j) pos[ps.bits[j]] = j;
  const int end = ps.first_block + ps.num_blocks;
  int g = ps.first_block;
  while (g < end) {
    unsigned mask = 0;
    const int start = g;
    for (; g < end; ++g) {
      unsigned gm = 1u << pos[gates[g].target];
      if (gates[g].kind != QCS_HIP_G_1Q) gm |= 1u << pos[gates[g].control];
      if (__builtin_popcount(mask | gm) > kGR) break;
      mask |= gm;
    }
    for (int q = 0; __builtin_popcount(mask) < kGR; ++q) mask |= 1u << q;
    BGroup grp;
    std::memset(&grp, 0, sizeof(grp));
    int local[32];
    for (int q = 0, i = 0; q < 32; ++q)
      if ((mask >> q) & 1u) {
        local[q] = i;
        grp.qpos[i++] = q;
      }
    for (int j = 0; j < kGA; ++j) {
      int off = 0;
      for (int i = 0; i < kGR; ++i)
        if ((j >> i) & 1) off |= 1 << grp.qpos[i];
      grp.rowslot[j] = dswz(off);
    }
    std::vector<BGate> ops;
    ops.reserve(size_t(g - start));
    for (int x = start; x < g; ++x) {
      BGate o;
      std::memset(&o, 0, sizeof(o));
      o.t = local[pos[gates[x].target]];
      o.c = gates[x].kind == QCS_HIP_G_1Q ? 0 : local[pos[gates[x].control]];
      std::memcpy(o.m, gates[x].m, sizeof(o.m));
      o.cls = gate_class(gates[x], true);
      ops.push_back(o);
    }
    grp.first_gate = int(bg->size());
    emit_group_ops(fold_trailing_perms(std::move(ops), grp.rowslot, grp.outslot), bg);
    grp.ngates = int(bg->size()) - grp.first_gate;
    groups->push_back(grp);
  }
}

static BBlock make_bblock(const qcs_hip_pass &ps, const qcs_hip_block &bl,
                          int wslot, bool swizzle) {
  BBlock o;
  std::memset(&o, 0, sizeof(o));
  o.k = bl.k;
  o.wslot = wslot;
  for (int q = 0; q < bl.k; ++q) o.tpos[q] = position_in(ps, bl.targets[q]);
  for (int r = 0; r < (1 << bl.k); ++r) {
    int off = 0;
    for (int q = 0; q < bl.k; ++q)
      if ((r >> q) & 1) off |= 1 << o.tpos[q];
    o.rowslot[r] = 8 * (swizzle ? bswz(off) : off);
  }
  return o;
}

template <int NTH>
static int launch_batch(qcs_hip_batch *b, const BatchLayout &L, int nitems,
                        uint64_t shots, int tile_bytes, int nch_cap, int grid) {
  char *d = static_cast<char *>(b->d_desc);
  hipLaunchKernelGGL(k_batch<NTH>, dim3(unsigned(grid)), dim3(NTH),
                     size_t(tile_bytes) + 2 * sizeof(double) * size_t(nch_cap),
                     b->stream, reinterpret_cast<const BCirc *>(d + L.circ),
                     reinterpret_cast<const BItem *>(d + L.order), nitems,
                     reinterpret_cast<const BPass *>(d + L.pass),
                     reinterpret_cast<const BGate *>(d + L.gate),
                     reinterpret_cast<const BGroup *>(d + L.group),
                     reinterpret_cast<const BBlock *>(d + L.block), b->d_w,
                     b->d_states, reinterpret_cast<const double *>(d + L.u),
                     b->d_out, shots, b->d_queue, reinterpret_cast<unsigned *>(d + L.sync),
                     tile_bytes, nch_cap, (b->flags & QCS_HIP_BATCH_KEEP_STATES) ? 1 : 0);
  return launch_check("k_batch");
}

template <int NTH>
static int occupancy(int lds_bytes, int *per_cu) {
  const void *fn = reinterpret_cast<const void *>(&k_batch<NTH>);
  QCS_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  lds_bytes));
  QCS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fn, NTH,
                                                           size_t(lds_bytes)));
  return 0;
}

static int batch_occupancy(int nth, int lds_bytes, int *per_cu) {
  switch (nth) {
    case 256: return occupancy<256>(lds_bytes, per_cu);
    default: return occupancy<512>(lds_bytes, per_cu);
  }
}

static int batch_launch(qcs_hip_batch *b, int nth, const BatchLayout &L, int nitems,
                        uint64_t shots, int tile_bytes, int nch_cap, int grid) {
  switch (nth) {
    case 256: return launch_batch<256>(b, L, nitems, shots, tile_bytes, nch_cap, grid);
    default: return launch_batch<512>(b, L, nitems, shots, tile_bytes, nch_cap, grid);
  }
}

// The launch sequence of one run on b->stream: uploads, composition, queue
// reset, the persistent kernel, the download of the sampled indices.  A
// rerun uploads only the uniforms and keeps the composed blocks.
static int enqueue_run(qcs_hip_batch *b, const BatchLayout &L, int nc, int nb,
                       uint64_t shots, int tile_bytes, int nch_cap, int grid,
                       bool timed, bool rerun) {
  char *d = static_cast<char *>(b->d_desc);
  const size_t ubytes = sizeof(double) * size_t(nc) * size_t(shots);
  if (!rerun)
    QCS_HIP_TRY(hipMemcpyAsync(d, b->h_stage, L.total, hipMemcpyHostToDevice,
                               b->stream));
  else if (ubytes)
    QCS_HIP_TRY(hipMemcpyAsync(d + L.u, b->h_stage + L.u, ubytes,
                               hipMemcpyHostToDevice, b->stream));
  int rc;
  if (!rerun && nb > 0 &&
      (rc = launch_compose(b->stream,
                           reinterpret_cast<const qcs_hip_gate *>(d + L.hgate),
                           reinterpret_cast<const qcs_hip_block *>(d + L.hblock),
                           nb, b->d_w, nullptr)))
    return rc;
  hipLaunchKernelGGL(k_batch_reset, dim3(unsigned(std::min<uint64_t>(ceil_div(nc, 256), 64))),
                     dim3(256), 0, b->stream, b->d_queue,
                     reinterpret_cast<unsigned *>(d + L.sync), nc);
  if ((rc = launch_check("k_batch_reset"))) return rc;
  if (timed) QCS_HIP_TRY(hipEventRecord(b->ev[1], b->stream));
  if ((rc = batch_launch(b, b->nth, L, b->last_nitems, shots, tile_bytes, nch_cap, grid)))
    return rc;
  if (timed) QCS_HIP_TRY(hipEventRecord(b->ev[2], b->stream));
  if (ubytes)
    QCS_HIP_TRY(hipMemcpyAsync(b->h_out, b->d_out, ubytes, hipMemcpyDeviceToHost,
                               b->stream));
  QCS_HIP_TRY(hipMemcpyAsync(b->h_flag, b->d_queue + 1, sizeof(unsigned),
                             hipMemcpyDeviceToHost, b->stream));
  return 0;
}

// Captures the launch sequence as a graph (graph mode: on first use, then
// replays it) or enqueues it directly with the kernel timed, waits, and
// records the device times.
static int submit(qcs_hip_batch *b, bool rerun, int nb, bool build_graph) {
  const BatchLayout &L = b->last_layout;
  const int nc = b->last_nc;
  const uint64_t shots = uint64_t(b->last_shots);
  const bool use_graph = (b->flags & QCS_HIP_BATCH_GRAPH) != 0;
  hipGraph_t *graph = rerun ? &b->rerun_graph : &b->graph;
  hipGraphExec_t *exec = rerun ? &b->rerun_exec : &b->exec;
  int rc;
  QCS_HIP_TRY(hipEventRecord(b->ev[0], b->stream));
  b->replayed = 0;
  if (use_graph) {
    if (!*exec || build_graph) {
      if (*exec) (void)hipGraphExecDestroy(*exec);
      if (*graph) (void)hipGraphDestroy(*graph);
      *exec = nullptr;
      *graph = nullptr;
      QCS_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_run(b, L, nc, nb, shots, b->last_tile_bytes, b->last_nch_cap,
                       b->last_grid, false, rerun);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
      *graph = g;
      QCS_HIP_TRY(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    } else {
      b->replayed = 1;
    }
    QCS_HIP_TRY(hipGraphLaunch(*exec, b->stream));
  } else {
    if ((rc = enqueue_run(b, L, nc, nb, shots, b->last_tile_bytes, b->last_nch_cap,
                          b->last_grid, true, rerun)))
      return rc;
  }
  QCS_HIP_TRY(hipEventRecord(b->ev[3], b->stream));
  QCS_HIP_TRY(hipStreamSynchronize(b->stream));
  if (*b->h_flag != 0u)
    return fail(QCS_HIP_EDEVICE, "batch: the work-groups sharing a circuit did not all "
                                 "reach a pass boundary in time");
  float ms = 0.0f;
  QCS_HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[3]));
  b->total_ms = ms;
  b->kernel_ms = -1.0f;
  if (!use_graph) {
    QCS_HIP_TRY(hipEventElapsedTime(&ms, b->ev[1], b->ev[2]));
    b->kernel_ms = ms;
  }
  return 0;
}

// host threads for building descriptors: QCS_HOST_THREADS, else the
// hardware's, at most 64
static double host_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int batch_run(qcs_hip_batch *b, const qcs_hip_batch_circuit *cs, int nc,
              const qcs_hip_pass *passes, int np, const qcs_hip_block *blocks,
              int nb, const qcs_hip_gate *gates, int ng, long shots,
              const double *u, long *out) {
  const double t0 = host_ms();
  int rc = validate_batch(cs, nc, passes, np, blocks, nb, gates, ng, shots, u,
                          out);
  if (rc) return rc;
  b->has_last = false;
  if (nc == 0) return 0;
  // a pass may serve several circuits, all of one precision and tile size;
  // the first circuit that names it owns it
  std::vector<int> owner(static_cast<size_t>(np), -1);
  for (int i = 0; i < nc; ++i)
    for (int p = cs[i].first_pass; p < cs[i].first_pass + cs[i].num_passes; ++p) {
      const int j = owner[size_t(p)];
      if (j < 0) owner[size_t(p)] = i;
      else if (cs[j].precision != cs[i].precision || cs[j].tile_bits != cs[i].tile_bits)
        return fail(QCS_HIP_EINVAL, "pass %d shared by circuits of different kinds", p);
    }
  QCS_HIP_TRY(hipSetDevice(b->device));

  const double t1 = host_ms();
  // device descriptors: passes, then their gates / register groups / blocks,
  // built for contiguous ranges of circuits on host threads and joined
  std::vector<BPass> hp(static_cast<size_t>(np));
  const int nparts = std::max(1, std::min(nc / 64 + 1, host_threads()));
  // the parts' op lists live in the batch between runs, so their memory is
  // reused rather than faulted in again on every planned run
  std::vector<BatchPart> &parts = b->parts;
  if (parts.size() < size_t(nparts)) parts.resize(size_t(nparts));
  for (int t = 0; t < nparts; ++t) {
    parts[size_t(t)].g.clear();
    parts[size_t(t)].grp.clear();
    parts[size_t(t)].blk.clear();
    parts[size_t(t)].own.clear();
  }
  auto build_part = [&](int t) {
    BatchPart &pt = parts[size_t(t)];
    const int i0 = int(int64_t(nc) * t / nparts), i1 = int(int64_t(nc) * (t + 1) / nparts);
    for (int i = i0; i < i1; ++i) {
      const qcs_hip_batch_circuit &c = cs[i];
      for (int p = c.first_pass; p < c.first_pass + c.num_passes; ++p) {
        if (owner[size_t(p)] != i) continue;
        const qcs_hip_pass &ps = passes[p];
        BPass &dp = hp[size_t(p)];
        std::memset(&dp, 0, sizeof(dp));
        for (int j = 0; j < ps.tile_bits; ++j) dp.sbits[j] = ps.bits[j];
        if (c.precision == QCS_HIP_C128 && c.tile_bits >= kGR) {
          dp.first_op = int(pt.grp.size());
          make_groups(ps, gates, &pt.grp, &pt.g);
          dp.nops = int(pt.grp.size()) - dp.first_op;
        } else if (c.precision == QCS_HIP_C128) {
          dp.first_op = int(pt.g.size());
          for (int g = ps.first_block; g < ps.first_block + ps.num_blocks; ++g)
            pt.g.push_back(make_bgate(ps, gates[g]));
          dp.nops = ps.num_blocks;
        } else {
          dp.first_op = int(pt.blk.size());
          for (int bi = ps.first_block; bi < ps.first_block + ps.num_blocks; ++bi)
            pt.blk.push_back(make_bblock(ps, blocks[bi], bi, ps.tile_bits >= 5));
          dp.nops = ps.num_blocks;
        }
        pt.own.push_back(p);
      }
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 1; t < nparts; ++t) th.emplace_back(build_part, t);
    build_part(0);
    for (auto &x : th) x.join();
  }
  const double tb1 = host_ms();
  // each part's ops go to a contiguous range of the joined arrays: offsets
  std::vector<size_t> og(size_t(nparts) + 1, 0), ogrp(size_t(nparts) + 1, 0),
      ob(size_t(nparts) + 1, 0);
  for (int t = 0; t < nparts; ++t) {
    og[size_t(t) + 1] = og[size_t(t)] + parts[size_t(t)].g.size();
    ogrp[size_t(t) + 1] = ogrp[size_t(t)] + parts[size_t(t)].grp.size();
    ob[size_t(t) + 1] = ob[size_t(t)] + parts[size_t(t)].blk.size();
  }
  const size_t n_gate = og[size_t(nparts)], n_group = ogrp[size_t(nparts)],
               n_block = ob[size_t(nparts)];

  const double tb2 = host_ms();
  // composition inputs: the blocks and only the gates they use (complex128
  // circuits never reach the composer), gates renumbered block by block
  std::vector<qcs_hip_block> cblk(blocks, blocks + nb);
  size_t n_cgate = 0;
  for (int bi = 0; bi < nb; ++bi) {
    cblk[size_t(bi)].first_gate = int(n_cgate);
    n_cgate += size_t(blocks[bi].num_gates);
  }

  // state slots and per-circuit cost
  const bool keep = (b->flags & QCS_HIP_BATCH_KEEP_STATES) != 0;
  std::vector<uint64_t> slot(size_t(nc), kNoSlot);
  std::vector<double> cost(size_t(nc), 0.0);
  size_t states_bytes = 0;
  int max_tile_bytes = 0, max_n = 1;
  for (int i = 0; i < nc; ++i) {
    const qcs_hip_batch_circuit &c = cs[i];
    const size_t amp = c.precision == QCS_HIP_C128 ? 16 : 8;
    const bool resident = c.tile_bits == c.num_qubits;
    if (!resident || keep) {
      slot[i] = states_bytes;
      states_bytes = align_up(states_bytes + (amp << c.num_qubits));
    }
    max_tile_bytes = std::max(max_tile_bytes, int(amp << c.tile_bits));
    max_n = std::max(max_n, c.num_qubits);
    double ops = 0.0;
    for (int p = c.first_pass; p < c.first_pass + c.num_passes; ++p) {
      if (c.precision == QCS_HIP_C128) {
        ops += passes[p].num_blocks + (resident ? 0.0 : 4.0);
      } else {
        for (int bi = passes[p].first_block;
             bi < passes[p].first_block + passes[p].num_blocks; ++bi)
          ops += 0.25 * double(1 << blocks[bi].k);
        ops += resident ? 0.0 : 2.0;
      }
    }
    cost[i] = double(amp) * double(uint64_t(1) << c.num_qubits) * (ops + 1.0);
  }
  const int nch_cap = int(std::max<uint64_t>(
      1, (uint64_t(1) << max_n) / std::min<uint64_t>(uint64_t(1) << max_n, kChunk)));
  const int lds_bytes = max_tile_bytes + int(2 * sizeof(double)) * nch_cap;

  // persistent grid: every work-group resident at once
  int per_cu = 0;
  if ((rc = batch_occupancy(b->nth, lds_bytes, &per_cu))) return rc;
  if (per_cu < 1) return fail(QCS_HIP_EDEVICE, "batch kernel does not fit on a CU");
  const int grid = std::min(nc, per_cu * b->num_cus);
  // A streamed complex128 circuit that costs several times a work-group's
  // share of the batch is shared by that many work-groups (at most 32, its
  // tile count, and half the grid), so no single circuit sets the launch
  // time.  QCS_BATCH_MAX_PARTS lowers the cap (1: no sharing).
  std::vector<int> shares(size_t(nc), 1);
  {
    int cap = 32;
    if (const char *e = std::getenv("QCS_BATCH_MAX_PARTS")) cap = std::max(1, std::atoi(e));
    cap = std::min(cap, std::max(1, grid / 2));
    double total = 0.0;
    for (int i = 0; i < nc; ++i) total += cost[size_t(i)];
    const double share = total / double(std::max(grid, 1));
    for (int i = 0; i < nc; ++i) {
      const qcs_hip_batch_circuit &c = cs[i];
      if (c.precision != QCS_HIP_C128 || c.tile_bits >= c.num_qubits) continue;
      const uint64_t ntiles = uint64_t(1) << (c.num_qubits - c.tile_bits);
      const int w = int(std::min<double>({cost[size_t(i)] / share, double(cap), double(ntiles)}));
      if (w >= 2) shares[size_t(i)] = w;
    }
  }
  // queue items: circuits by cost, largest first; the parts of a shared
  // circuit are consecutive
  std::vector<BItem> items;
  {
    std::vector<int> idx(static_cast<size_t>(nc));
    for (int i = 0; i < nc; ++i) idx[size_t(i)] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int c) { return cost[size_t(a)] > cost[size_t(c)]; });
    items.reserve(size_t(nc));
    for (int i : idx)
      for (int j = 0; j < shares[size_t(i)]; ++j) items.push_back(BItem{i, j, shares[size_t(i)], 0});
  }
  const int nitems = int(items.size());

  // layout of the descriptor block (device and pinned staging alike)
  BatchLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + (bytes ? bytes : 1));
    return at;
  };
  L.circ = take(sizeof(BCirc) * size_t(nc));
  L.order = take(sizeof(BItem) * size_t(nitems));
  L.sync = take(sizeof(unsigned) * size_t(nc));
  L.pass = take(sizeof(BPass) * size_t(np));
  L.gate = take(sizeof(BGate) * n_gate);
  L.group = take(sizeof(BGroup) * n_group);
  L.block = take(sizeof(BBlock) * n_block);
  L.hgate = take(sizeof(qcs_hip_gate) * n_cgate);
  L.hblock = take(sizeof(qcs_hip_block) * size_t(nb));
  L.u = take(sizeof(double) * size_t(nc) * size_t(shots));
  L.total = off;

  const double tb3 = host_ms();
  // buffers
  if ((rc = grow_device(b, &b->d_desc, &b->desc_cap, L.total))) return rc;
  if ((rc = grow_pinned(b, reinterpret_cast<void **>(&b->h_stage), &b->stage_cap,
                        L.total)))
    return rc;
  {
    void *p = b->d_w;
    if ((rc = grow_device(b, &p, &b->w_cap, sizeof(float) * kWSlotFloats *
                                                 size_t(nb > 0 ? nb : 1))))
      return rc;
    b->d_w = static_cast<float *>(p);
    p = b->d_states;
    if ((rc = grow_device(b, &p, &b->states_cap, states_bytes ? states_bytes : kBAlign)))
      return rc;
    b->d_states = static_cast<char *>(p);
    const size_t obytes = sizeof(long) * std::max<size_t>(1, size_t(nc) * size_t(shots));
    p = b->d_out;
    if ((rc = grow_device(b, &p, &b->out_cap, obytes))) return rc;
    b->d_out = static_cast<long *>(p);
    p = b->h_out;
    if ((rc = grow_pinned(b, &p, &b->hout_cap, obytes))) return rc;
    b->h_out = static_cast<long *>(p);
  }

  const double tb4 = host_ms();
  // descriptors into the pinned staging block
  char *h = b->h_stage;
  BCirc *hc = reinterpret_cast<BCirc *>(h + L.circ);
  for (int i = 0; i < nc; ++i) {
    const qcs_hip_batch_circuit &c = cs[i];
    BCirc &d = hc[i];
    std::memset(&d, 0, sizeof(d));
    d.n = c.num_qubits;
    d.prec = c.precision;
    d.tb = c.tile_bits;
    d.npass = c.num_passes;
    d.first_pass = c.first_pass;
    d.state_off = slot[i];
    d.shot_off = uint64_t(i) * uint64_t(shots);
  }
  // every part writes its ops straight into the staging block (and fixes up
  // the op indices of the passes it owns), and a share of the composer's
  // gates; parts own disjoint passes and ranges
  auto fill_part = [&](int t) {
    BatchPart &pt = parts[size_t(t)];
    const int o_g = int(og[size_t(t)]), o_grp = int(ogrp[size_t(t)]), o_b = int(ob[size_t(t)]);
    for (int p : pt.own) {
      const int i = owner[size_t(p)];
      if (cs[i].precision == QCS_HIP_C128 && cs[i].tile_bits >= kGR)
        hp[size_t(p)].first_op += o_grp;
      else if (cs[i].precision == QCS_HIP_C128)
        hp[size_t(p)].first_op += o_g;
      else
        hp[size_t(p)].first_op += o_b;
    }
    for (BGroup &grp : pt.grp) grp.first_gate += o_g;
    if (!pt.g.empty())
      std::memcpy(h + L.gate + sizeof(BGate) * size_t(o_g), pt.g.data(),
                  sizeof(BGate) * pt.g.size());
    if (!pt.grp.empty())
      std::memcpy(h + L.group + sizeof(BGroup) * size_t(o_grp), pt.grp.data(),
                  sizeof(BGroup) * pt.grp.size());
    if (!pt.blk.empty())
      std::memcpy(h + L.block + sizeof(BBlock) * size_t(o_b), pt.blk.data(),
                  sizeof(BBlock) * pt.blk.size());
    const int b0 = int(int64_t(nb) * t / nparts), b1 = int(int64_t(nb) * (t + 1) / nparts);
    qcs_hip_gate *cg = reinterpret_cast<qcs_hip_gate *>(h + L.hgate);
    for (int bi = b0; bi < b1; ++bi)
      if (blocks[bi].num_gates > 0)
        std::memcpy(cg + cblk[size_t(bi)].first_gate, gates + blocks[bi].first_gate,
                    sizeof(qcs_hip_gate) * size_t(blocks[bi].num_gates));
  };
  {
    std::vector<std::thread> th;
    for (int t = 1; t < nparts; ++t) th.emplace_back(fill_part, t);
    fill_part(0);
    for (auto &x : th) x.join();
  }
  if (np > 0) std::memcpy(h + L.pass, hp.data(), sizeof(BPass) * hp.size());
  if (nb > 0) std::memcpy(h + L.hblock, cblk.data(), sizeof(qcs_hip_block) * size_t(nb));
  if (shots > 0)
    std::memcpy(h + L.u, u, sizeof(double) * size_t(nc) * size_t(shots));
  std::memcpy(h + L.order, items.data(), sizeof(BItem) * items.size());

  const double t2 = host_ms();
  std::vector<long> key = {b->generation, nc, nitems, np, nb, long(n_cgate), long(n_gate),
                           long(n_group), long(n_block), shots, b->nth,
                           max_tile_bytes, nch_cap, grid, long(L.total),
                           long(states_bytes)};
  const bool new_graph = (b->flags & QCS_HIP_BATCH_GRAPH) && b->graph_key != key;
  if (new_graph) drop_graph(b);
  b->last_layout = L;
  b->last_nc = nc;
  b->last_nitems = nitems;
  b->last_shots = shots;
  b->last_tile_bytes = max_tile_bytes;
  b->last_nch_cap = nch_cap;
  b->last_grid = grid;
  if ((rc = submit(b, false, nb, new_graph))) return rc;
  if (b->flags & QCS_HIP_BATCH_GRAPH) b->graph_key = key;
  const double t3 = host_ms();
  if (shots > 0)
    std::memcpy(out, b->h_out, sizeof(long) * size_t(nc) * size_t(shots));
  const char *prof = std::getenv("QCS_BATCH_PROFILE");
  if (prof && prof[0] && prof[0] != '0')
    std::fprintf(stderr,
                 "qcs_hip_batch_run: validate %.1f ms, descriptors %.1f ms (parts %.1f, "
                 "offsets %.1f, compose inputs + layout %.1f, buffers %.1f, staging %.1f), "
                 "device %.1f ms (kernel %.1f), wall to sync %.1f ms, %.1f MB staged\n",
                 t1 - t0, t2 - t1, tb1 - t1, tb2 - tb1, tb3 - tb2, tb4 - tb3, t2 - tb4,
                 b->total_ms, b->kernel_ms, t3 - t2, double(L.total) / 1e6);
  b->last_n.assign(size_t(nc), 0);
  b->last_prec.assign(size_t(nc), 0);
  b->last_off = slot;
  for (int i = 0; i < nc; ++i) {
    b->last_n[size_t(i)] = cs[i].num_qubits;
    b->last_prec[size_t(i)] = cs[i].precision;
  }
  b->has_last = true;
  return 0;
}

int batch_rerun(qcs_hip_batch *b, long shots, const double *u, long *out) {
  if (!b->has_last) return fail(QCS_HIP_EINVAL, "no successful run to repeat");
  if (shots != b->last_shots)
    return fail(QCS_HIP_EINVAL, "shots %ld differ from the last run's %ld", shots,
                b->last_shots);
  const size_t n = size_t(b->last_nc) * size_t(shots);
  if (n > 0 && (!u || !out)) return fail(QCS_HIP_EINVAL, "null uniforms or output");
  QCS_HIP_TRY(hipSetDevice(b->device));
  if (n > 0) std::memcpy(b->h_stage + b->last_layout.u, u, sizeof(double) * n);
  int rc = submit(b, true, 0, false);
  if (rc) {
    b->has_last = false;
    return rc;
  }
  if (n > 0) std::memcpy(out, b->h_out, sizeof(long) * n);
  return 0;
}

}  // namespace qcs

using namespace qcs;

extern "C" {

int qcs_hip_batch_create(qcs_hip_batch **out, int flags) {
  if (!out) return fail(QCS_HIP_EINVAL, "null output handle");
  *out = nullptr;
  if (flags & ~(QCS_HIP_BATCH_KEEP_STATES | QCS_HIP_BATCH_GRAPH))
    return fail(QCS_HIP_EINVAL, "unknown flags %d", flags);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(QCS_HIP_ENODEV, "no HIP device available");
  qcs_hip_batch *b = new (std::nothrow) qcs_hip_batch();
  if (!b) return fail(QCS_HIP_ENOMEM, "host allocation failed");
  b->flags = flags;
  int rc = 0;
  hipError_t e;
  if ((e = hipGetDevice(&b->device)) != hipSuccess) rc = check_hip(e, "hipGetDevice");
  if (!rc && (e = hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount,
                                        b->device)) != hipSuccess)
    rc = check_hip(e, "hipDeviceGetAttribute");
  if (!rc && (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
    rc = check_hip(e, "hipStreamCreate");
  for (int i = 0; i < 4 && !rc; ++i)
    if ((e = hipEventCreate(&b->ev[i])) != hipSuccess) rc = check_hip(e, "hipEventCreate");
  if (!rc && (e = hipMalloc(reinterpret_cast<void **>(&b->d_queue), kBAlign)) != hipSuccess)
    rc = check_hip(e, "hipMalloc(queue)");
  if (!rc && (e = hipHostMalloc(reinterpret_cast<void **>(&b->h_flag), kBAlign,
                                hipHostMallocDefault)) != hipSuccess)
    rc = check_hip(e, "hipHostMalloc(flag)");
  if (!rc) *b->h_flag = 0u;
  if (rc) {
    qcs_hip_batch_destroy(b);
    return rc;
  }
  *out = b;
  return 0;
}

void qcs_hip_batch_destroy(qcs_hip_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_w) (void)hipFree(b->d_w);
  if (b->d_states) (void)hipFree(b->d_states);
  if (b->d_out) (void)hipFree(b->d_out);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->h_stage) (void)hipHostFree(b->h_stage);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_flag) (void)hipHostFree(b->h_flag);
  for (int i = 0; i < 4; ++i)
    if (b->ev[i]) (void)hipEventDestroy(b->ev[i]);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int qcs_hip_batch_set_tuning(qcs_hip_batch *b, int knob, int value) {
  if (!b) return fail(QCS_HIP_EINVAL, "null batch");
  if (knob == QCS_HIP_TUNE_BATCH_THREADS) {
    if (value != 256 && value != 512)
      return fail(QCS_HIP_EINVAL, "batch threads %d: use 256 or 512", value);
    if (value != b->nth) {
      drop_graph(b);
      b->has_last = false;  // the grid was sized for the old work-group
    }
    b->nth = value;
    return 0;
  }
  return fail(QCS_HIP_EINVAL, "unknown batch tuning knob %d", knob);
}

int qcs_hip_batch_run(qcs_hip_batch *b, const qcs_hip_batch_circuit *circuits,
                      int num_circuits, const qcs_hip_pass *passes,
                      int num_passes, const qcs_hip_block *blocks,
                      int num_blocks, const qcs_hip_gate *gates, int num_gates,
                      long shots, const double *u, long *index_out) {
  if (!b) return fail(QCS_HIP_EINVAL, "null batch");
  return batch_run(b, circuits, num_circuits, passes, num_passes, blocks,
                   num_blocks, gates, num_gates, shots, u, index_out);
}

int qcs_hip_batch_rerun(qcs_hip_batch *b, long shots, const double *u,
                        long *index_out) {
  if (!b) return fail(QCS_HIP_EINVAL, "null batch");
  return batch_rerun(b, shots, u, index_out);
}

int qcs_hip_batch_read_state(qcs_hip_batch *b, int circuit, long offset,
                             long count, double *out) {
  if (!b) return fail(QCS_HIP_EINVAL, "null batch");
  if (!(b->flags & QCS_HIP_BATCH_KEEP_STATES))
    return fail(QCS_HIP_EUNSUP, "batch was created without QCS_HIP_BATCH_KEEP_STATES");
  if (circuit < 0 || size_t(circuit) >= b->last_n.size())
    return fail(QCS_HIP_EINVAL, "circuit %d not in the last run", circuit);
  const int n = b->last_n[size_t(circuit)];
  const long N = long(1) << n;
  if (offset < 0 || count < 0 || offset > N || count > N - offset)
    return fail(QCS_HIP_EINVAL, "range [%ld, %ld+%ld) outside 2^%d amplitudes",
                offset, offset, count, n);
  if (count > 0 && !out) return fail(QCS_HIP_EINVAL, "null buffer");
  if (count == 0) return 0;
  QCS_HIP_TRY(hipSetDevice(b->device));
  const char *src = b->d_states + b->last_off[size_t(circuit)];
  if (b->last_prec[size_t(circuit)] == QCS_HIP_C128) {
    QCS_HIP_TRY(hipMemcpy(out, src + 16 * size_t(offset), 16 * size_t(count),
                          hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<float> tmp(2 * size_t(count));
  QCS_HIP_TRY(hipMemcpy(tmp.data(), src + 8 * size_t(offset), 8 * size_t(count),
                        hipMemcpyDeviceToHost));
  for (size_t i = 0; i < tmp.size(); ++i) out[i] = double(tmp[i]);
  return 0;
}

int qcs_hip_batch_last_timing(qcs_hip_batch *b, double *kernel_ms,
                              double *total_ms, int *graph_replayed) {
  if (!b) return fail(QCS_HIP_EINVAL, "null batch");
  if (kernel_ms) *kernel_ms = b->kernel_ms;
  if (total_ms) *total_ms = b->total_ms;
  if (graph_replayed) *graph_replayed = b->replayed;
  return 0;
}

}  // extern "C"
