// Fused precision form.
//
// A fused block is a run of gates acting on k <= 5 qubits T (ascending).
// 1. k_compose (FP64, one work-group per block) builds U = G_last ... G_1 by
//    applying each gate to the columns of the 2^k identity with the
//    reference's complex arithmetic, then writes
//      - W: the real 2^(k+1) x 2^(k+1) FP32 form of U with interleaved
//        {re, im} rows/columns: W[2i][2j] = Ur, W[2i][2j+1] = -Ui,
//        W[2i+1][2j] = Ui, W[2i+1][2j+1] = Ur  (the four real products),
//        followed, for k <= 4, by the 2^k x 2^k planes Ur + Ui and Ur - Ui,
//      - U itself in FP64 for the VALU kernel.
// 2. Each block is then applied in one HBM pass: complex64 states with
//    n >= 12 by the matrix-core kernel of blockpass.hip (k_fused_mfma),
//    other states by k_fused_valu below.  Tile programs (tiles.hip) apply
//    several composed blocks per pass instead.
#include <cstring>

#include "device.h"
#include "internal.h"

namespace qcs {

constexpr int kWFloats = 64 * 64;           // W slot per block (k = 5)
constexpr int kUDoubles = 32 * 32 * 2;      // U slot per block (k = 5)

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
  // blocks of up to 4 qubits also get Ur + Ui and Ur - Ui, each rounded
  // once, after the real form (the three-product form of tile programs
  // reads them)
  if (RR * RR + 2 * R * R <= kWFloats)
    for (int idx = threadIdx.x; idx < R * R; idx += 64) {
      const double ur = cr[idx % R][idx / R], ui = ci[idx % R][idx / R];
      W[RR * RR + idx] = float(ur + ui);
      W[RR * RR + R * R + idx] = float(ur - ui);
    }
  if (!ubuf) return;
  double *U = ubuf + size_t(blockIdx.x) * kUDoubles;
  for (int idx = threadIdx.x; idx < R * R; idx += 64) {
    const int i = idx / R, jj = idx % R;
    U[2 * idx] = cr[jj][i];
    U[2 * idx + 1] = ci[jj][i];
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
  if (s->precision == QCS_HIP_C64 && s->n >= 12) {
    if ((rc = launch_block_mfma(s, k, targets, W, s->block_variant[k]))) return rc;
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

int launch_compose(hipStream_t stream, const qcs_hip_gate *gates,
                   const qcs_hip_block *blocks, int nb, float *wbuf,
                   double *ubuf) {
  if (nb <= 0) return 0;
  hipLaunchKernelGGL(k_compose, dim3(nb), dim3(64), 0, stream, gates, blocks,
                     wbuf, ubuf);
  return launch_check("k_compose");
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
