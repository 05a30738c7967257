// One fused block per HBM pass (complex64, FP32 MFMA).
//
// A composed block W (real 2^(k+1) x 2^(k+1) form of the k-qubit unitary,
// rows/columns interleaved {re, im}) is applied to a complex64 state as a
// GEMM on the matrix cores (v_mfma_f32_32x32x2f32).  A work-group owns a
// tile of 2^TB amplitudes: the 2^k rows are the target-bit patterns, the
// 2^(TB-k) columns the lowest TB-k non-target bits.  The tile's index bits S
// always include bit 0, so every thread moves 16-byte runs of two amplitudes;
// for the usual windows of k consecutive qubits the runs are 2^(TB-k) * 8
// bytes long.  Global -> LDS in address order, MFMA with W held in registers
// as the A operand and the LDS tile as B, accumulators back to LDS, LDS ->
// global in address order.  With interleaved rows a complex64 amplitude is
// two consecutive real rows, so the B operand of K-step s is {Re, Im} of
// complex row s and each 32-row accumulator block holds whole complex
// outputs.
//
// Launch variants (chosen per block by the autotuner, see qcs_hip_set_tuning):
//   TB 12 / 256 threads, or TB 13 / 256 or 512 threads; with or without
//   software pipelining (the next tile's loads issued before the current
//   tile's MFMA work).  All use non-temporal global accesses and, when the
//   state has >= 8 tiles, an XCD-aware tile order (work-group i runs on XCD
//   i % 8, so XCD x takes the contiguous x-th eighth of the tiles).
#include "device.h"
#include "internal.h"

namespace qcs {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kBpMaxBits = 13;

struct BlockGeom {
  int sbits[kBpMaxBits];  // tile bit j -> index bit (ascending)
  int wpos[kBpMaxBits];   // tile bit j -> LDS offset (float2 units)
};

__device__ __forceinline__ float4 bp_ld(const float2 *p) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void bp_st(float2 *p, float4 x) {
  const f4v v = {x.x, x.y, x.z, x.w};
  __builtin_nontemporal_store(v, reinterpret_cast<f4v *>(p));
}

template <int K, int TB, int NTH, bool PIPE>
__global__ __launch_bounds__(NTH) void k_fused_mfma(float2 *__restrict__ a,
                                                    const float *__restrict__ W,
                                                    BlockGeom g, uint64_t ntiles) {
  constexpr int R = 1 << K, RR = 2 * R;
  constexpr int C = 1 << (TB - K);            // columns per tile
  constexpr int CS = C + 1;                   // LDS row stride (float2), padded
  constexpr int MT = (RR + 31) / 32;          // 32-row blocks of the output
  constexpr int NW = NTH / 64;
  constexpr int CT = C / (32 * NW);           // 32-column MFMA tiles per wave
  constexpr int TBITS = __builtin_ctz(NTH) + 1;
  constexpr int IT = (1 << TB) / (2 * NTH);
  constexpr int IB = TB - TBITS;
  static_assert(CT >= 1, "at least one column tile per wave");
  // real and imaginary parts in separate planes: a half-wave reads one part
  // of 32 consecutive columns, 32 distinct 4-byte banks
  __shared__ float lre[R * CS], lim[R * CS];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = lane >> 5, l32 = lane & 31;

  float wa[MT][R];  // A operand: lane holds W[mt*32 + l32][2s + h]
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int row = mt * 32 + l32;
#pragma unroll
    for (int s = 0; s < R; ++s) wa[mt][s] = (row < RR) ? W[row * RR + 2 * s + h] : 0.0f;
  }

  uint64_t addr_t = 0;
  int pos_t = 0;
#pragma unroll
  for (int j = 1; j < TBITS; ++j)
    if ((tid >> (j - 1)) & 1) {
      addr_t |= uint64_t(1) << g.sbits[j];
      pos_t += g.wpos[j];
    }
  const int w0 = g.wpos[0];
  uint64_t addr_it[IT];
  int pos_it[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    uint64_t ai = 0;
    int pi = 0;
#pragma unroll
    for (int jj = 0; jj < IB; ++jj)
      if ((it >> jj) & 1) {
        ai |= uint64_t(1) << g.sbits[TBITS + jj];
        pi += g.wpos[TBITS + jj];
      }
    addr_it[it] = ai | addr_t;
    pos_it[it] = pi + pos_t;
  }

  // tile sequence of this work-group
  uint64_t first, step, end;
  if ((ntiles & 7) == 0 && (gridDim.x & 7) == 0) {
    const uint64_t G8 = gridDim.x >> 3, T8 = ntiles >> 3;
    const uint64_t x = blockIdx.x & 7, j = blockIdx.x >> 3;
    first = x * T8 + j;
    step = G8;
    end = (x + 1) * T8;
  } else {
    first = blockIdx.x;
    step = gridDim.x;
    end = ntiles;
  }
  auto tile_base = [&](uint64_t t) {
    uint64_t b = t;
#pragma unroll
    for (int j = 0; j < TB; ++j) b = insert0(b, g.sbits[j]);
    return b;
  };

  float4 v[IT];
  uint64_t tile = first;
  uint64_t base = tile_base(tile);
  if (PIPE && tile < end) {
#pragma unroll
    for (int it = 0; it < IT; ++it) v[it] = bp_ld(a + (base | addr_it[it]));
  }
  const float *bpl = h ? lim : lre;  // this lane's part of the B operand
  for (; tile < end; tile += step) {
    if (!PIPE) {
#pragma unroll
      for (int it = 0; it < IT; ++it) v[it] = bp_ld(a + (base | addr_it[it]));
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      lre[pos_it[it]] = v[it].x;
      lim[pos_it[it]] = v[it].y;
      lre[pos_it[it] + w0] = v[it].z;
      lim[pos_it[it] + w0] = v[it].w;
    }
    __syncthreads();
    const uint64_t next = tile + step;
    const uint64_t nbase = tile_base(next);
    if (PIPE && next < end) {
#pragma unroll
      for (int it = 0; it < IT; ++it) v[it] = bp_ld(a + (nbase | addr_it[it]));
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int col = (wave * CT + ct) * 32 + l32;
      float b[R];
#pragma unroll
      for (int s = 0; s < R; ++s) b[s] = bpl[s * CS + col];
      floatx16 acc[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) {
        acc[mt] = floatx16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                           0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < R; ++s)
          acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[mt][s], b[s], acc[mt], 0, 0, 0);
      }
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r0 = mt * 16 + 4 * q + 2 * h;
          if (r0 < R) {
            lre[r0 * CS + col] = acc[mt][4 * q];
            lim[r0 * CS + col] = acc[mt][4 * q + 1];
          }
          if (r0 + 1 < R) {
            lre[(r0 + 1) * CS + col] = acc[mt][4 * q + 2];
            lim[(r0 + 1) * CS + col] = acc[mt][4 * q + 3];
          }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int p = pos_it[it];
      bp_st(a + (base | addr_it[it]),
            make_float4(lre[p], lim[p], lre[p + w0], lim[p + w0]));
    }
    __syncthreads();
    base = nbase;
  }
}

// ---- host side ------------------------------------------------------------------

static BlockGeom make_geom(int n, int k, const int *targets, int tb) {
  BlockGeom g = {};
  const int lc = tb - k;
  const int cs = (1 << lc) + 1;
  int is_t[64] = {0};
  for (int q = 0; q < k; ++q) is_t[targets[q]] = 1;
  int nl = 0, j = 0;
  for (int b = 0; b < n && j < tb; ++b) {
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

template <int K, int TB, int NTH, bool PIPE>
static void launch_k(qcs_hip_state *s, const float *W, const BlockGeom &g,
                     unsigned grid_cap) {
  const uint64_t ntiles = s->size >> TB;
  const uint64_t grid = ntiles < grid_cap ? ntiles : grid_cap;
  hipLaunchKernelGGL((k_fused_mfma<K, TB, NTH, PIPE>), dim3(unsigned(grid)),
                     dim3(NTH), 0, s->stream, static_cast<float2 *>(s->vec), W, g,
                     ntiles);
}

template <int K>
static void launch_variant(qcs_hip_state *s, const float *W, const int *targets,
                           int variant) {
  switch (variant) {
    case QCS_HIP_BLOCK_TB12_256:
      launch_k<K, 12, 256, false>(s, W, make_geom(s->n, K, targets, 12), 2048);
      break;
    case QCS_HIP_BLOCK_TB12_256_PIPE:
      launch_k<K, 12, 256, true>(s, W, make_geom(s->n, K, targets, 12), 1024);
      break;
    case QCS_HIP_BLOCK_TB13_256_PIPE:
      launch_k<K, 13, 256, true>(s, W, make_geom(s->n, K, targets, 13), 512);
      break;
    default:
      launch_k<K, 13, 512, true>(s, W, make_geom(s->n, K, targets, 13), 512);
      break;
  }
}

// variant usable for this state size (13-bit tiles need n >= 13)
int block_variant_for(const qcs_hip_state *s, int variant) {
  if (s->n < 13 && (variant == QCS_HIP_BLOCK_TB13_256_PIPE ||
                    variant == QCS_HIP_BLOCK_TB13_512_PIPE))
    return QCS_HIP_BLOCK_TB12_256_PIPE;
  return variant;
}

int launch_block_mfma(qcs_hip_state *s, int k, const int *targets,
                      const float *W, int variant) {
  variant = block_variant_for(s, variant);
  switch (k) {
    case 1: launch_variant<1>(s, W, targets, variant); break;
    case 2: launch_variant<2>(s, W, targets, variant); break;
    case 3: launch_variant<3>(s, W, targets, variant); break;
    case 4: launch_variant<4>(s, W, targets, variant); break;
    default: launch_variant<5>(s, W, targets, variant); break;
  }
  return launch_check("k_fused_mfma");
}

}  // namespace qcs
