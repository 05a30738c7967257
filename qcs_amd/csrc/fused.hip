// Fused precision form.
//
// A fused block is a run of gates acting on k <= 5 qubits T (ascending).
// 1. k_compose (FP64, one work-group per block) builds U = G_last ... G_1 by
//    applying each gate to the columns of the 2^k identity with the
//    reference's complex arithmetic, then writes
//      - W: the real 2^(k+1) x 2^(k+1) FP32 form of U with interleaved
//        {re, im} rows/columns: W[2i][2j] = Ur, W[2i][2j+1] = -Ui,
//        W[2i+1][2j] = Ui, W[2i+1][2j+1] = Ur  (the four real products),
//      - U itself in FP64 for the VALU kernel.
// 2. k_fused_mfma applies W to a complex64 state as a GEMM on the FP32 matrix
//    cores.  A work-group owns a tile of 4096 amplitudes: the 2^k rows are the
//    target-bit patterns, the 2^(12-k) columns the lowest 12-k non-target
//    bits.  The tile's 12 index bits S = T + those columns always include bit
//    0, so each thread moves 16-byte runs of two amplitudes and every wave
//    touches whole cache lines.  Global -> LDS in address order, MFMA
//    32x32x2 f32 with W held in registers as the A operand and the LDS tile
//    as B, accumulators back to LDS, LDS -> global in address order.  With
//    interleaved rows a complex64 amplitude is two consecutive real rows,
//    so the B operand of K-step s is {Re, Im} of complex row s and each
//    32-row accumulator block holds whole complex outputs.
#include <cstring>

#include "device.h"
#include "internal.h"

namespace qcs {

constexpr int kTileBits = 12;               // amplitudes per MFMA tile = 4096
constexpr int kWFloats = 64 * 64;           // W slot per block (k = 5)
constexpr int kUDoubles = 32 * 32 * 2;      // U slot per block (k = 5)

struct FusedGeom {
  int sbits[kTileBits];  // tile bit j -> index bit (ascending)
  int wpos[kTileBits];   // tile bit j -> LDS offset (float2 units)
};

typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---- composition (FP64) -------------------------------------------------------

__global__ __launch_bounds__(64) void k_compose(const qcs_hip_gate *gates,
                                                const qcs_hip_block *blocks,
                                                float *wbuf, double *ubuf) {
  __shared__ double cr[32][33];
  __shared__ double ci[32][33];
  const qcs_hip_block blk = blocks[blockIdx.x];
  const int k = blk.k, R = 1 << k;
  const int j = threadIdx.x;  // column of U owned by this thread
  if (j < R)
    for (int i = 0; i < R; ++i) {
      cr[j][i] = (i == j) ? 1.0 : 0.0;
      ci[j][i] = 0.0;
    }
  for (int g = blk.first_gate; g < blk.first_gate + blk.num_gates; ++g) {
    const qcs_hip_gate gt = gates[g];
    int lt = -1, lc = -1;
    for (int q = 0; q < k; ++q) {
      if (blk.targets[q] == gt.target) lt = q;
      if (blk.targets[q] == gt.control) lc = q;
    }
    if (j >= R || lt < 0) continue;
    const int tb = 1 << lt;
    if (gt.kind == QCS_HIP_G_SWAP) {
      if (lc < 0) continue;
      const int cb = 1 << lc;
      for (int i = 0; i < R; ++i) {
        if ((i & tb) && !(i & cb)) {
          const int o = i ^ tb ^ cb;
          const double xr = cr[j][i], xi = ci[j][i];
          cr[j][i] = cr[j][o];
          ci[j][i] = ci[j][o];
          cr[j][o] = xr;
          ci[j][o] = xi;
        }
      }
      continue;
    }
    const int cb = (gt.kind == QCS_HIP_G_CU && lc >= 0) ? (1 << lc) : 0;
    const double *m = gt.m;
    for (int i0 = 0; i0 < R; ++i0) {
      if ((i0 & tb) || ((i0 & cb) != cb)) continue;
      const int i1 = i0 | tb;
      const double v0r = cr[j][i0], v0i = ci[j][i0];
      const double v1r = cr[j][i1], v1i = ci[j][i1];
      double o0r, o0i, o1r, o1i;
      row2(m[0], m[1], m[2], m[3], v0r, v0i, v1r, v1i, o0r, o0i);
      row2(m[4], m[5], m[6], m[7], v0r, v0i, v1r, v1i, o1r, o1i);
      cr[j][i0] = o0r;
      ci[j][i0] = o0i;
      cr[j][i1] = o1r;
      ci[j][i1] = o1i;
    }
  }
  __syncthreads();
  const int RR = 2 * R;
  float *W = wbuf + size_t(blockIdx.x) * kWFloats;
  for (int idx = threadIdx.x; idx < RR * RR; idx += 64) {
    const int row = idx / RR, col = idx % RR;
    const int i = row >> 1, p = row & 1, jj = col >> 1, q = col & 1;
    const double ur = cr[jj][i], ui = ci[jj][i];
    const double v = (p == q) ? ur : (p == 0 ? -ui : ui);
    W[idx] = float(v);
  }
  double *U = ubuf + size_t(blockIdx.x) * kUDoubles;
  for (int idx = threadIdx.x; idx < R * R; idx += 64) {
    const int i = idx / R, jj = idx % R;
    U[2 * idx] = cr[jj][i];
    U[2 * idx + 1] = ci[jj][i];
  }
}

// ---- MFMA apply (complex64, n >= 12) ------------------------------------------

template <int K>
__global__ __launch_bounds__(kThreads) void k_fused_mfma(float2 *a,
                                                         const float *W,
                                                         FusedGeom g,
                                                         uint64_t ntiles) {
  constexpr int R = 1 << K;            // complex rows
  constexpr int RR = 2 * R;            // real rows / columns of W
  constexpr int C = 1 << (kTileBits - K);  // columns per tile
  constexpr int CS = C + 1;            // LDS row stride (float2), padded
  constexpr int MT = (RR + 31) / 32;   // 32-row blocks of the output
  constexpr int CT = C / 128;          // 32-column MFMA tiles per wave
  static_assert(CT >= 1, "at least one column tile per wave");
  __shared__ float2 lds[R * CS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = lane >> 5, l32 = lane & 31;

  // A operand: lane holds W[mt*32 + l32][2s + h]
  float wa[MT][R];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int row = mt * 32 + l32;
#pragma unroll
    for (int s = 0; s < R; ++s)
      wa[mt][s] = (row < RR) ? W[row * RR + 2 * s + h] : 0.0f;
  }

  // thread offsets of tile element e = 2*tid (tile bits 1..8)
  uint64_t addr_t = 0;
  int pos_t = 0;
#pragma unroll
  for (int j = 1; j <= 8; ++j)
    if ((tid >> (j - 1)) & 1) {
      addr_t |= uint64_t(1) << g.sbits[j];
      pos_t += g.wpos[j];
    }
  const int w0 = g.wpos[0];
  uint64_t addr_it[8];
  int pos_it[8];
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    uint64_t ai = 0;
    int pi = 0;
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
      if ((it >> jj) & 1) {
        ai |= uint64_t(1) << g.sbits[9 + jj];
        pi += g.wpos[9 + jj];
      }
    addr_it[it] = ai | addr_t;
    pos_it[it] = pi + pos_t;
  }

  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint64_t base = tile;
#pragma unroll
    for (int j = 0; j < kTileBits; ++j) base = insert0(base, g.sbits[j]);

    float4 v[8];
#pragma unroll
    for (int it = 0; it < 8; ++it)
      v[it] = *reinterpret_cast<const float4 *>(a + (base | addr_it[it]));
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      lds[pos_it[it]] = make_float2(v[it].x, v[it].y);
      lds[pos_it[it] + w0] = make_float2(v[it].z, v[it].w);
    }
    __syncthreads();

    const float *ldsf = reinterpret_cast<const float *>(lds);
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int col = (wave * CT + ct) * 32 + l32;
      float b[R];
#pragma unroll
      for (int s = 0; s < R; ++s) b[s] = ldsf[(s * CS + col) * 2 + h];
      floatx16 acc[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        acc[mt] = floatx16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                           0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[mt][s], b[s],
                                                         acc[mt], 0, 0, 0);
      }
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r0 = mt * 16 + 4 * q + 2 * h;
          if (r0 < R)
            lds[r0 * CS + col] = make_float2(acc[mt][4 * q], acc[mt][4 * q + 1]);
          if (r0 + 1 < R)
            lds[(r0 + 1) * CS + col] =
                make_float2(acc[mt][4 * q + 2], acc[mt][4 * q + 3]);
        }
    }
    __syncthreads();

#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const float2 x = lds[pos_it[it]], y = lds[pos_it[it] + w0];
      *reinterpret_cast<float4 *>(a + (base | addr_it[it])) =
          make_float4(x.x, x.y, y.x, y.y);
    }
    __syncthreads();
  }
}

// ---- VALU apply (any precision, any n >= k) -------------------------------------

template <typename T, int K>
__global__ __launch_bounds__(kThreads) void k_fused_valu(
    typename Amp<T>::type *a, const double *U, int t0, int t1, int t2, int t3,
    int t4, uint64_t ncols) {
  typedef typename Amp<T>::type A;
  constexpr int R = 1 << K;
  __shared__ T ur[R * R], ui[R * R];
  for (int i = threadIdx.x; i < R * R; i += kThreads) {
    ur[i] = T(U[2 * i]);
    ui[i] = T(U[2 * i + 1]);
  }
  __syncthreads();
  const int tq[5] = {t0, t1, t2, t3, t4};
  const uint64_t c = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (c >= ncols) return;
  uint64_t base = c;
#pragma unroll
  for (int q = 0; q < K; ++q) base = insert0(base, tq[q]);
  uint64_t off[R];
  A x[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    uint64_t o = base;
#pragma unroll
    for (int q = 0; q < K; ++q)
      if ((j >> q) & 1) o |= uint64_t(1) << tq[q];
    off[j] = o;
    x[j] = a[o];
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    T sr = T(0), si = T(0);
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const T mr = ur[i * R + j], mi = ui[i * R + j];
      sr += mr * x[j].x - mi * x[j].y;
      si += mr * x[j].y + mi * x[j].x;
    }
    A o;
    o.x = sr;
    o.y = si;
    a[off[i]] = o;  // each thread owns its column: no cross-thread hazard
  }
}

// ---- host side -------------------------------------------------------------------

static int grow(void **p, size_t *cap, size_t need) {
  if (*cap >= need) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  QCS_HIP_TRY(hipMalloc(p, need));
  *cap = need;
  return 0;
}

static FusedGeom make_geom(int n, int k, const int *targets) {
  FusedGeom g;
  const int lc = kTileBits - k;
  const int cs = (1 << lc) + 1;
  int is_t[64] = {0};
  for (int q = 0; q < k; ++q) is_t[targets[q]] = 1;
  int nl = 0, j = 0;
  for (int b = 0; b < n && j < kTileBits; ++b) {
    if (is_t[b]) {
      int r = 0;
      while (targets[r] != b) ++r;
      g.sbits[j] = b;
      g.wpos[j] = (1 << r) * cs;
      ++j;
    } else if (nl < lc) {
      g.sbits[j] = b;
      g.wpos[j] = 1 << nl;
      ++nl;
      ++j;
    }
  }
  return g;
}

template <int K>
static void launch_mfma_k(qcs_hip_state *s, const float *W, const FusedGeom &g,
                          uint64_t ntiles) {
  const uint64_t grid = ntiles < 2048 ? ntiles : 2048;
  hipLaunchKernelGGL(k_fused_mfma<K>, dim3(unsigned(grid)), dim3(kThreads), 0,
                     s->stream, static_cast<float2 *>(s->vec), W, g, ntiles);
}

template <typename T, int K>
static void launch_valu_k(qcs_hip_state *s, const double *U, const int *t) {
  const uint64_t ncols = s->size >> K;
  int tq[5] = {0, 0, 0, 0, 0};
  for (int q = 0; q < K; ++q) tq[q] = t[q];
  hipLaunchKernelGGL((k_fused_valu<T, K>), dim3(unsigned(ceil_div(ncols, kThreads))),
                     dim3(kThreads), 0, s->stream,
                     static_cast<typename Amp<T>::type *>(s->vec), U, tq[0],
                     tq[1], tq[2], tq[3], tq[4], ncols);
}

template <typename T>
static void launch_valu(qcs_hip_state *s, int k, const double *U, const int *t) {
  switch (k) {
    case 1: launch_valu_k<T, 1>(s, U, t); break;
    case 2: launch_valu_k<T, 2>(s, U, t); break;
    case 3: launch_valu_k<T, 3>(s, U, t); break;
    case 4: launch_valu_k<T, 4>(s, U, t); break;
    default: launch_valu_k<T, 5>(s, U, t); break;
  }
}

// one HBM pass applying a composed block (W for MFMA, U for the VALU kernel)
static int apply_block(qcs_hip_state *s, int k, const int *targets,
                       const float *W, const double *U) {
  const double N = double(s->size);
  const double bytes = 2.0 * N * double(s->amp_bytes);
  const double flops = 8.0 * double(1 << k) * N;
  TimedLaunch tl;
  int rc;
  timing_begin(s, QCS_HIP_K_FUSED, bytes, flops, &tl);
  if (s->precision == QCS_HIP_C64 && s->n >= kTileBits) {
    const FusedGeom g = make_geom(s->n, k, targets);
    const uint64_t ntiles = s->size >> kTileBits;
    switch (k) {
      case 1: launch_mfma_k<1>(s, W, g, ntiles); break;
      case 2: launch_mfma_k<2>(s, W, g, ntiles); break;
      case 3: launch_mfma_k<3>(s, W, g, ntiles); break;
      case 4: launch_mfma_k<4>(s, W, g, ntiles); break;
      default: launch_mfma_k<5>(s, W, g, ntiles); break;
    }
    if ((rc = launch_check("k_fused_mfma"))) return rc;
  } else {
    if (s->precision == QCS_HIP_C128) launch_valu<double>(s, k, U, targets);
    else launch_valu<float>(s, k, U, targets);
    if ((rc = launch_check("k_fused_valu"))) return rc;
  }
  return timing_end(s, &tl);
}

static int ensure_wbuf(qcs_hip_state *s, size_t nb) {
  if (s->wbuf_blocks >= nb) return 0;
  if (s->wbuf) (void)hipFree(s->wbuf);
  if (s->ubuf) (void)hipFree(s->ubuf);
  s->wbuf = nullptr;
  s->ubuf = nullptr;
  s->wbuf_blocks = 0;
  QCS_HIP_TRY(hipMalloc(&s->wbuf, sizeof(float) * kWFloats * nb));
  QCS_HIP_TRY(hipMalloc(&s->ubuf, sizeof(double) * kUDoubles * nb));
  s->wbuf_blocks = nb;
  return 0;
}

// Upload the gate list and blocks, compose every block's unitary on the
// device (W into wbuf slot b, U into ubuf slot b).  `extra` bytes of pinned
// staging after the gates and blocks are left for the caller's own uploads.
int compose_blocks(qcs_hip_state *s, const qcs_hip_block *blocks, int nb,
                   const qcs_hip_gate *gates, int ng, size_t extra,
                   char **extra_stage) {
  // the previous flush may still read the device gate list
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  const size_t gbytes = sizeof(qcs_hip_gate) * size_t(ng > 0 ? ng : 1);
  const size_t bbytes = sizeof(qcs_hip_block) * size_t(nb);
  int rc;
  if ((rc = grow(&s->gates_dev, &s->gates_cap, gbytes))) return rc;
  if ((rc = grow(&s->blocks_dev, &s->blocks_cap, bbytes))) return rc;
  if ((rc = ensure_wbuf(s, size_t(nb)))) return rc;
  if ((rc = ensure_pinned(s, gbytes + bbytes + extra))) return rc;
  char *stage = reinterpret_cast<char *>(s->host);
  if (extra_stage) *extra_stage = stage + gbytes + bbytes;
  if (ng > 0) std::memcpy(stage, gates, gbytes);
  std::memcpy(stage + gbytes, blocks, bbytes);
  if (ng > 0)
    QCS_HIP_TRY(hipMemcpyAsync(s->gates_dev, stage, gbytes,
                               hipMemcpyHostToDevice, s->stream));
  QCS_HIP_TRY(hipMemcpyAsync(s->blocks_dev, stage + gbytes, bbytes,
                             hipMemcpyHostToDevice, s->stream));
  TimedLaunch tl;
  timing_begin(s, QCS_HIP_K_COMPOSE, 0.0, 0.0, &tl);
  hipLaunchKernelGGL(k_compose, dim3(nb), dim3(64), 0, s->stream,
                     static_cast<const qcs_hip_gate *>(s->gates_dev),
                     static_cast<const qcs_hip_block *>(s->blocks_dev), s->wbuf,
                     s->ubuf);
  if ((rc = launch_check("k_compose"))) return rc;
  return timing_end(s, &tl);
}

int fused_apply(qcs_hip_state *s, const qcs_hip_block *blocks, int nb,
                const qcs_hip_gate *gates, int ng) {
  if (nb <= 0) return 0;
  int rc = compose_blocks(s, blocks, nb, gates, ng, 0, nullptr);
  if (rc) return rc;
  for (int b = 0; b < nb; ++b) {
    rc = apply_block(s, blocks[b].k, blocks[b].targets,
                     s->wbuf + size_t(b) * kWFloats,
                     s->ubuf + size_t(b) * kUDoubles);
    if (rc) return rc;
  }
  return 0;
}

int fused_apply_unitary(qcs_hip_state *s, int k, const int *targets,
                        const double *u) {
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  int rc;
  if ((rc = ensure_wbuf(s, 1))) return rc;
  const size_t wbytes = sizeof(float) * kWFloats, ubytes = sizeof(double) * kUDoubles;
  if ((rc = ensure_pinned(s, wbytes + ubytes))) return rc;
  float *W = reinterpret_cast<float *>(s->host);
  double *U = reinterpret_cast<double *>(reinterpret_cast<char *>(s->host) + wbytes);
  const int R = 1 << k, RR = 2 * R;
  for (int row = 0; row < RR; ++row)
    for (int col = 0; col < RR; ++col) {
      const int i = row >> 1, p = row & 1, j = col >> 1, q = col & 1;
      const double ur = u[2 * (i * R + j)], ui = u[2 * (i * R + j) + 1];
      W[row * RR + col] = float((p == q) ? ur : (p == 0 ? -ui : ui));
    }
  std::memcpy(U, u, sizeof(double) * 2 * R * R);
  QCS_HIP_TRY(hipMemcpyAsync(s->wbuf, W, wbytes, hipMemcpyHostToDevice, s->stream));
  QCS_HIP_TRY(hipMemcpyAsync(s->ubuf, U, ubytes, hipMemcpyHostToDevice, s->stream));
  if ((rc = apply_block(s, k, targets, s->wbuf, s->ubuf))) return rc;
  // keep the staging buffer alive until the copies have been consumed
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

}  // namespace qcs
