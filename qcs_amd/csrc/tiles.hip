// Tile programs: several fused blocks per HBM pass (complex64, FP32 MFMA).
//
// A pass owns a set S of TB = 13 index bits (ascending; always the lowest 4,
// so a tile is read as >= 128-byte contiguous runs).  The state splits into
// 2^(n-13) tiles of 8192 amplitudes; a work-group streams tiles through LDS
// and, per tile, applies every block of the pass in order -- each block a
// composed 2^k x 2^k unitary (k <= 4) on k of the tile's bits, applied as a
// real GEMM on the matrix cores (v_mfma_f32_32x32x2f32, the operand layout
// and accumulation order of k_fused_mfma) -- before the tile goes back to
// HBM.  HBM traffic per pass is one read and one write of the state, however
// many blocks it holds.
//
// LDS layout.  The tile is two planes of 8192 floats, real parts then
// imaginary parts; tile element e (bit j of e = index bit S[j]) sits at slot
// swz(e) = e ^ (e>>5 & 31) ^ (e>>10 & 7) of both.  swz is linear over GF(2),
// so the slot of (column c, row s) of a block is swz(col) ^ swz(row): the
// host precomputes the 16 row terms of every block and each B-operand read
// or result write costs one XOR.  The fold puts index bits 5..12 onto the
// low slot bits, so the 32 columns a half-wave reads (the 5 lowest
// non-target tile bits) are 32 distinct slots mod 32; in separate planes
// those are 32 distinct 4-byte banks (interleaved {re, im} would put a
// half-wave's 32 real parts 8 bytes apart, two per bank).  After the tile
// come the A operands of every block of the pass, staged once per
// work-group in lane order; they stay in LDS because on gfx9 every vector
// memory load retires through one in-order counter, so a global load inside
// the tile loop would wait behind the next tile's prefetch.
//
// Work-group tile order is XCD aware: the hardware places work-group i on
// XCD i % 8, so XCD x is given the contiguous x-th eighth of the tiles.
#include <cstring>

#include "device.h"
#include "internal.h"

namespace qcs {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kPassMaxBits = 14;
constexpr int kTileKMax = QCS_HIP_TILE_KMAX;  // qubits per block in a pass
constexpr int kTileRMax = 1 << kTileKMax;
constexpr int kTileWFloats = 64 * 64;          // W slot per block in wbuf
constexpr int kTileBitsFixed = QCS_HIP_TILE_BITS;
constexpr size_t kTileLdsTile = sizeof(float2) << kTileBitsFixed;
constexpr size_t kTileLdsPerBlock = sizeof(float) * kTileRMax * 64;
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr int kTileMaxBlocks = int((kLdsPerCu - kTileLdsTile) / kTileLdsPerBlock);
static_assert(kTileMaxBlocks >= QCS_HIP_TILE_MAX_BLOCKS,
              "QCS_HIP_TILE_MAX_BLOCKS blocks must fit in LDS beside a tile");

struct TilePass {
  int tb;
  int nblocks;
  int first_block;
  int reserved;
  int sbits[kPassMaxBits];
};

struct TileBlock {
  int k;
  int wslot;
  int tpos[kTileKMax];        // tile-bit positions of the targets, ascending
  int rowslot[kTileRMax];     // slot swz(row bits of s) of row s
};

__host__ __device__ __forceinline__ int swz(int e) {
  return e ^ ((e >> 5) & 31) ^ ((e >> 10) & 7);
}

__device__ __forceinline__ float4 ld_nt(const float2 *p) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void st_nt(float2 *p, float4 x) {
  const f4v v = {x.x, x.y, x.z, x.w};
  __builtin_nontemporal_store(v, reinterpret_cast<f4v *>(p));
}

// insert a zero bit at position b (32-bit)
__device__ __forceinline__ int ins0(int x, int b) {
  return ((x >> b) << (b + 1)) | (x & ((1 << b) - 1));
}

// One block of a pass on the tile, its width K known at compile time.  The
// B operand rows and the target positions are uniform (scalar loads); the
// output rows of a lane depend on its half-wave, so their slots are computed
// here once per block instead of being read per lane from global memory --
// a vector load inside the tile loop would wait behind the next tile's
// prefetch, which shares the in-order vmcnt counter.
template <int K, int NTH, int TILE>
__device__ __forceinline__ void tp_block(char *ldsb, const TileBlock *__restrict__ bl,
                                         const float *wlb) {
  constexpr int R = 1 << K, NW = NTH / 64, NCOLS = TILE >> K;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l32 = lane & 31;
  float *re = reinterpret_cast<float *>(ldsb);
  float *im = re + TILE;
  const float *bplane = h ? im : re;  // the lane's part of the B operand
  float wa[R];
#pragma unroll
  for (int s = 0; s < R; ++s) wa[s] = wlb[s * 64 + lane];
  int tpos[K], rs[R];
#pragma unroll
  for (int q = 0; q < K; ++q) tpos[q] = bl->tpos[q];
#pragma unroll
  for (int s = 0; s < R; ++s) rs[s] = bl->rowslot[s];
  // output rows of this lane: accumulators 4q, 4q+1 hold row 4q + 2h,
  // 4q+2, 4q+3 row 4q + 2h + 1; both candidates are uniform values
  int os0[4], os1[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int a0 = rs[4 * q < R ? 4 * q : 0], a1 = rs[4 * q + 1 < R ? 4 * q + 1 : 0];
    const int b0 = rs[4 * q + 2 < R ? 4 * q + 2 : 0], b1 = rs[4 * q + 3 < R ? 4 * q + 3 : 0];
    os0[q] = __builtin_amdgcn_readfirstlane(a0) ^ (h ? (b0 ^ a0) : 0);
    os1[q] = __builtin_amdgcn_readfirstlane(a1) ^ (h ? (b1 ^ a1) : 0);
  }
  for (int c0 = wave * 32; c0 < NCOLS; c0 += NW * 32) {
    int col = c0 + l32;
#pragma unroll
    for (int q = 0; q < K; ++q) col = ins0(col, tpos[q]);
    const int cs = swz(col);  // slot of (col, row 0)
    float bv[R];
#pragma unroll
    for (int s = 0; s < R; ++s) bv[s] = bplane[cs ^ rs[s]];
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[s], bv[s], acc, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (4 * q >= R) continue;
      const int r0 = 4 * q + 2 * h;
      if (r0 < R) {
        re[cs ^ os0[q]] = acc[4 * q];
        im[cs ^ os0[q]] = acc[4 * q + 1];
      }
      if (r0 + 1 < R) {
        re[cs ^ os1[q]] = acc[4 * q + 2];
        im[cs ^ os1[q]] = acc[4 * q + 3];
      }
    }
  }
}

// The descriptor pointers are __restrict__ so the compiler may read them
// through the scalar cache (they are uniform and never written by the
// kernel); vector loads of them would wait behind the tile prefetch.
template <int TB, int NTH>
__global__ __launch_bounds__(NTH) void k_tile_program(
    float2 *__restrict__ a, const TilePass *__restrict__ passes, int pass_index,
    const TileBlock *__restrict__ blocks, const float *__restrict__ wbuf,
    uint64_t ntiles) {
  constexpr int TILE = 1 << TB;
  constexpr int TBITS = __builtin_ctz(NTH) + 1;  // tile bits covered by 2*tid
  constexpr int IT = TILE / (2 * NTH);           // float4 loads per thread
  constexpr int IB = TB - TBITS;                 // tile bits covered by `it`
  constexpr int RM = kTileRMax;
  extern __shared__ float4 smem[];
  char *ldsb = reinterpret_cast<char *>(smem);
  float *wl = reinterpret_cast<float *>(ldsb + (sizeof(float2) << TB));

  const TilePass &ps = passes[pass_index];
  const int nb = ps.nblocks;
  const int tid = threadIdx.x;

  // stage the A operands of every block: wl[(b*RM + s)*64 + lane] holds
  // W_b[l32][2s + h] (zero outside the block's 2^(k+1) x 2^(k+1) matrix)
  for (int idx = tid; idx < nb * RM * 64; idx += NTH) {
    const int b = idx / (RM * 64), rem = idx % (RM * 64);
    const int s = rem >> 6, ln = rem & 63;
    const TileBlock &bl = blocks[ps.first_block + b];
    const int R = 1 << bl.k, RR = 2 * R, row = ln & 31, hh = ln >> 5;
    wl[idx] = (row < RR && s < R)
                  ? wbuf[size_t(bl.wslot) * kTileWFloats + row * RR + 2 * s + hh]
                  : 0.0f;
  }

  // element e = 2*tid + 2*NTH*it and its neighbour e+1 (tile bit 0)
  uint64_t addr_t = 0;
  int e_t = 0;
#pragma unroll
  for (int j = 1; j < TBITS; ++j)
    if ((tid >> (j - 1)) & 1) {
      addr_t |= uint64_t(1) << ps.sbits[j];
      e_t |= 1 << j;
    }
  const int slot_t = swz(e_t);
  auto it_addr = [&](int it) {
    uint64_t ai = 0;
#pragma unroll
    for (int jj = 0; jj < IB; ++jj)
      if ((it >> jj) & 1) ai |= uint64_t(1) << ps.sbits[TBITS + jj];
    return ai;
  };

  const uint64_t G8 = gridDim.x >> 3, T8 = ntiles >> 3;
  const uint64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  uint64_t tile = xcd * T8 + slot;
  const uint64_t tend = (xcd + 1) * T8;
  auto tile_base = [&](uint64_t t) {
    uint64_t b = t;
#pragma unroll
    for (int j = 0; j < TB; ++j) b = insert0(b, ps.sbits[j]);
    return b;
  };

  float4 v[IT];
  uint64_t base = tile_base(tile);
  if (tile < tend) {
#pragma unroll
    for (int it = 0; it < IT; ++it) v[it] = ld_nt(a + (base | it_addr(it) | addr_t));
  }
  __syncthreads();  // A operands staged

  for (; tile < tend; tile += G8) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int s0 = slot_t ^ swz(it << TBITS);  // slot of e; e+1 is s0 ^ 1
      float *re = reinterpret_cast<float *>(ldsb), *im = re + TILE;
      re[s0] = v[it].x;
      im[s0] = v[it].y;
      re[s0 ^ 1] = v[it].z;
      im[s0 ^ 1] = v[it].w;
    }
    __syncthreads();
    const uint64_t next = tile + G8;
    const uint64_t nbase = tile_base(next);
    if (next < tend) {
#pragma unroll
      for (int it = 0; it < IT; ++it)
        v[it] = ld_nt(a + (nbase | it_addr(it) | addr_t));
    }

    for (int b = 0; b < nb; ++b) {
      const TileBlock *bl = &blocks[ps.first_block + b];
      const float *wlb = wl + b * RM * 64;
      switch (bl->k) {
        case 1: tp_block<1, NTH, TILE>(ldsb, bl, wlb); break;
        case 2: tp_block<2, NTH, TILE>(ldsb, bl, wlb); break;
        case 3: tp_block<3, NTH, TILE>(ldsb, bl, wlb); break;
        default: tp_block<4, NTH, TILE>(ldsb, bl, wlb); break;
      }
      __syncthreads();
    }

#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int s0 = slot_t ^ swz(it << TBITS);
      const float *re = reinterpret_cast<const float *>(ldsb), *im = re + TILE;
      const float2 x = make_float2(re[s0], im[s0]);
      const float2 y = make_float2(re[s0 ^ 1], im[s0 ^ 1]);
      st_nt(a + (base | it_addr(it) | addr_t), make_float4(x.x, x.y, y.x, y.y));
    }
    __syncthreads();
    base = nbase;
  }
}

// ---- host side ------------------------------------------------------------------

static int grow_dev(void **p, size_t *cap, size_t need) {
  if (*cap >= need) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  QCS_HIP_TRY(hipMalloc(p, need));
  *cap = need;
  return 0;
}

template <int NTH>
static int launch_tp(qcs_hip_state *s, const TilePass *dpass, int pi,
                     const TileBlock *dblocks, uint64_t ntiles, int nblocks) {
  static bool attr_set = false;
  if (!attr_set) {
    QCS_HIP_TRY(hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_tile_program<kTileBitsFixed, NTH>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(kLdsPerCu)));
    attr_set = true;
  }
  const size_t lds = kTileLdsTile + kTileLdsPerBlock * size_t(nblocks);
  const uint64_t per_cu = kLdsPerCu / lds;  // work-groups that fit per CU
  uint64_t grid = uint64_t(s->tile_grid) * (per_cu > 1 ? 2 : 1);
  if (grid > ntiles) grid = ntiles;
  hipLaunchKernelGGL((k_tile_program<kTileBitsFixed, NTH>), dim3(unsigned(grid)),
                     dim3(NTH), lds, s->stream, static_cast<float2 *>(s->vec),
                     dpass, pi, dblocks, s->wbuf, ntiles);
  return launch_check("k_tile_program");
}

int tile_max_blocks() { return QCS_HIP_TILE_MAX_BLOCKS; }

int tiles_apply(qcs_hip_state *s, const qcs_hip_pass *passes, int np,
                const qcs_hip_block *blocks, int nb, const qcs_hip_gate *gates,
                int ng) {
  if (np <= 0) return 0;
  const size_t pbytes = sizeof(TilePass) * size_t(np);
  const size_t tbytes = sizeof(TileBlock) * size_t(nb > 0 ? nb : 1);
  int rc;
  if ((rc = grow_dev(&s->tpass_dev, &s->tpass_cap, pbytes))) return rc;
  if ((rc = grow_dev(&s->tblock_dev, &s->tblock_cap, tbytes))) return rc;
  char *stage = nullptr;
  if ((rc = compose_blocks(s, blocks, nb, gates, ng, pbytes + tbytes, &stage)))
    return rc;
  TilePass *tp = reinterpret_cast<TilePass *>(stage);
  TileBlock *tb = reinterpret_cast<TileBlock *>(stage + pbytes);
  for (int p = 0; p < np; ++p) {
    std::memset(&tp[p], 0, sizeof(TilePass));
    tp[p].tb = passes[p].tile_bits;
    tp[p].nblocks = passes[p].num_blocks;
    tp[p].first_block = passes[p].first_block;
    for (int j = 0; j < passes[p].tile_bits; ++j) tp[p].sbits[j] = passes[p].bits[j];
    for (int b = passes[p].first_block;
         b < passes[p].first_block + passes[p].num_blocks; ++b) {
      std::memset(&tb[b], 0, sizeof(TileBlock));
      tb[b].k = blocks[b].k;
      tb[b].wslot = b;
      for (int q = 0; q < blocks[b].k; ++q) {
        int pos = 0;
        while (passes[p].bits[pos] != blocks[b].targets[q]) ++pos;
        tb[b].tpos[q] = pos;
      }
      for (int r = 0; r < kTileRMax; ++r) {
        int off = 0;
        for (int q = 0; q < blocks[b].k; ++q)
          if ((r >> q) & 1) off |= 1 << tb[b].tpos[q];
        tb[b].rowslot[r] = swz(off);
      }
    }
  }
  QCS_HIP_TRY(hipMemcpyAsync(s->tpass_dev, tp, pbytes, hipMemcpyHostToDevice,
                             s->stream));
  QCS_HIP_TRY(hipMemcpyAsync(s->tblock_dev, tb, tbytes, hipMemcpyHostToDevice,
                             s->stream));
  const TilePass *dp = static_cast<const TilePass *>(s->tpass_dev);
  const TileBlock *db = static_cast<const TileBlock *>(s->tblock_dev);
  const double N = double(s->size);
  for (int p = 0; p < np; ++p) {
    const uint64_t ntiles = s->size >> passes[p].tile_bits;
    double flops = 0.0;
    for (int b = passes[p].first_block;
         b < passes[p].first_block + passes[p].num_blocks; ++b)
      flops += 8.0 * double(1 << blocks[b].k) * N;
    TimedLaunch tl;
    timing_begin(s, QCS_HIP_K_FUSED, 2.0 * N * double(s->amp_bytes), flops, &tl);
    if (s->tile_threads == 1024)
      rc = launch_tp<1024>(s, dp, p, db, ntiles, passes[p].num_blocks);
    else
      rc = launch_tp<512>(s, dp, p, db, ntiles, passes[p].num_blocks);
    if (rc) return rc;
    if ((rc = timing_end(s, &tl))) return rc;
  }
  // the staging buffer holds the descriptors until the copies have run
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

}  // namespace qcs
