// Tile programs: several fused blocks per HBM pass (complex64, matrix cores).
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
// LDS layout.  The tile is two planes of 8192 floats, real parts at byte 0
// and imaginary parts at PLANE = 32 KiB; tile element e (bit j of e = index
// bit S[j]) sits at slot swz(e) = e ^ (e>>5 & 31) ^ (e>>10 & 7) of both.
// swz is linear over GF(2), so the byte address of (column c, row s) of a
// block is 4 swz(c) ^ 4 swz(row bits of s): the host precomputes the row
// terms of every block, and since the kernel has no static LDS (checked on
// the host) its dynamic LDS starts at address 0, so each B-operand read or
// result write costs one XOR and no base add.  The fold puts index bits
// 5..12 onto the low slot bits, so the 32 columns a half-wave reads (the 5
// lowest non-target tile bits) are 32 distinct slots mod 32; in separate
// planes those are 32 distinct 4-byte banks.
//
// After the tile come, per block of the pass, its A operands in lane order
// and a 256-byte table: the per-lane part of the column address (column l32
// with zero bits at the targets, through swz) and the byte offsets of the
// lane's output rows.  Column c of a wave's group splits into that lane part
// and a wave-uniform part computed on the scalar unit.  The operands stay
// in LDS because on gfx9 every vector memory load retires through one
// in-order counter, so a global load inside the tile loop would wait behind
// the next tile's prefetch.
//
// Vector ALU work does not run alongside FP32 MFMA on gfx950 (their times
// add; bf16 MFMA overlaps it only in part), so the block code is written for
// few VALU instructions: table lookups instead of index arithmetic, uniform
// values in scalar registers, one XOR per LDS access.  A column group's
// B-operand reads are all issued before its MFMA chain: left to itself the
// compiler reuses one register for every operand, and each MFMA then waits
// a full LDS round trip.
//
// Split-bf16 math (QCS_HIP_TILE_MATH_SPLIT_BF16).  gfx950's bf16 matrix rate
// is 16x its FP32 rate, so a block can instead run as a sum of bf16 GEMMs
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation): the FP32 matrix is split
// into three bf16 parts W = Wh + Wm + Wl (exact: 3 x 8 significant bits),
// each B value into two, x = xh + xl (|x - xh - xl| <= 2^-16 |x|), and
//   W x ~ Wl xh + Wm xh + Wh xl + Wh xh
// leaves out Wm xl and Wl xl (<= 2^-16 |W||x|).  The matrix is carried at
// FP32 accuracy, so its error cannot add up coherently over repeated blocks;
// what remains is a rounding of the state to 16 bits per block, which adds
// up like a random walk.  The A operands of a block are staged as three
// bf16 parts in lane order (6 KiB per block), so a pass holds at most
// QCS_HIP_TILE_MAX_BLOCKS_SPLIT blocks.
//
// Work-group tile order is XCD aware: the hardware places work-group i on
// XCD i % 8, so XCD x is given the contiguous x-th eighth of the tiles.
#include <cstring>
#include <mutex>

#include "device.h"
#include "internal.h"

namespace qcs {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kPassMaxBits = 14;
constexpr int kTileKMax = QCS_HIP_TILE_KMAX;  // qubits per block in a pass
constexpr int kTileRMax = 1 << kTileKMax;
constexpr int kTileWFloats = 64 * 64;          // W slot per block in wbuf
constexpr int kTileBitsFixed = QCS_HIP_TILE_BITS;
constexpr int kTilePlane = 4 << kTileBitsFixed;  // bytes per plane of the tile
constexpr size_t kTileLdsTile = sizeof(float2) << kTileBitsFixed;
constexpr size_t kLdsPerCu = 160 * 1024;
// per block: A operands, then the lane table
constexpr int kTileAFp32 = 4 * 64 * 16;             // [4][lane] float4
constexpr int kTileASplit = 3 * 2 * 64 * 16;        // [part][instr][lane] 8 x bf16
constexpr int kTileAux = 256;                        // lanecol[32], outoff[2][8]
constexpr int kTileStrideFp32 = kTileAFp32 + kTileAux;
constexpr int kTileStrideSplit = kTileASplit + kTileAux;
static_assert(kTileLdsTile + size_t(kTileStrideFp32) * QCS_HIP_TILE_MAX_BLOCKS <= kLdsPerCu,
              "QCS_HIP_TILE_MAX_BLOCKS blocks must fit in LDS beside a tile");
static_assert(kTileLdsTile + size_t(kTileStrideSplit) * QCS_HIP_TILE_MAX_BLOCKS_SPLIT <=
                  kLdsPerCu,
              "QCS_HIP_TILE_MAX_BLOCKS_SPLIT blocks must fit in LDS beside a tile");

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
  int tpos[kTileKMax];       // tile-bit positions of the targets, ascending
  int rowoff[kTileRMax];     // 4 swz(row bits of s): byte offset of row s
  int lanecol[32];           // 4 swz(l32 with zero bits at the targets)
  int outoff[2][8];          // [h][2q + j]: rowoff of output row 4q + 2h + j
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
__host__ __device__ __forceinline__ int ins0(int x, int b) {
  return ((x >> b) << (b + 1)) | (x & ((1 << b) - 1));
}

// LDS by absolute byte address (the dynamic LDS starts at 0)
template <typename T>
__device__ __forceinline__ T lds_load(int addr) {
  typedef __attribute__((address_space(3))) T lds_t;
  return *reinterpret_cast<lds_t *>(static_cast<size_t>(static_cast<unsigned>(addr)));
}
template <typename T>
__device__ __forceinline__ void lds_store(int addr, T v) {
  typedef __attribute__((address_space(3))) T lds_t;
  *reinterpret_cast<lds_t *>(static_cast<size_t>(static_cast<unsigned>(addr))) = v;
}

// x0, x1 -> packed bf16 pairs hi (round to nearest even) and lo (the rounded
// remainder): x = hi + lo within 2^-16 |x|
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned &hi,
                                           unsigned &lo) {
  const f2v x = {x0, x1};
  const unsigned h = __builtin_bit_cast(unsigned, __builtin_convertvector(x, bf16x2));
  const f2v r = {x0 - __uint_as_float(h << 16), x1 - __uint_as_float(h & 0xffff0000u)};
  hi = h;
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector(r, bf16x2));
}

// What every block kernel needs of its block: uniform row offsets and target
// positions (scalar registers), the lane's column part and output offsets.
template <int K>
struct BlockGeom {
  static constexpr int R = 1 << K;
  int tpos[K];
  int ro[R];
  int lp, lph;  // lane part of the column address: real plane, B plane
  int os[8];    // output row offsets of this lane
  __device__ __forceinline__ BlockGeom(const TileBlock *__restrict__ bl, int aux) {
    const int lane = threadIdx.x & 63, h = lane >> 5, l32 = lane & 31;
#pragma unroll
    for (int q = 0; q < K; ++q) tpos[q] = bl->tpos[q];
#pragma unroll
    for (int s = 0; s < R; ++s) ro[s] = bl->rowoff[s];
    lp = lds_load<int>(aux + 4 * l32);
    lph = lp | (h ? kTilePlane : 0);
    const i32x4 o0 = lds_load<i32x4>(aux + 128 + 32 * h);
    const i32x4 o1 = lds_load<i32x4>(aux + 128 + 32 * h + 16);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      os[j] = o0[j];
      os[4 + j] = o1[j];
    }
  }
  // wave-uniform part of the address of column group `grp` (32 columns)
  __device__ __forceinline__ int group_part(int grp) const {
    int up = grp * 32;
#pragma unroll
    for (int q = 0; q < K; ++q) up = ins0(up, tpos[q]);
    return 4 * swz(up);
  }
  // result rows of a 32x32 accumulator tile (rows 4q + 2h + j of the block)
  __device__ __forceinline__ void store(int cb, const floatx16 &acc) const {
    const int h = (threadIdx.x >> 5) & 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (4 * q >= R) continue;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (4 * q + 2 * h + j < R) {
          const int o = cb ^ os[2 * q + j];
          lds_store<float>(o, acc[4 * q + 2 * j]);
          lds_store<float>(o + kTilePlane, acc[4 * q + 2 * j + 1]);
        }
      }
    }
  }
};

// One block in FP32: lane (l32, h) supplies A[l32][2s + h] (output row
// l32/2, part l32&1; input row s, part h) and B[2s + h][column l32] -- the
// lane's column of plane h, row s -- for s = 0 .. R-1, one MFMA each.
template <int K, int NTH, int TILE>
__device__ __forceinline__ void tp_block(const TileBlock *__restrict__ bl, int blk) {
  constexpr int R = 1 << K, NW = NTH / 64, GROUPS = (TILE >> K) / 32 / NW;
  static_assert(GROUPS >= 1, "a wave takes whole column groups");
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float wa[R];
#pragma unroll
  for (int v = 0; v < (R + 3) / 4; ++v) {
    const f4v x = lds_load<f4v>(blk + (v * 64 + lane) * 16);
    wa[4 * v] = x.x;
    if (4 * v + 1 < R) wa[4 * v + 1] = x.y;
    if (4 * v + 2 < R) wa[4 * v + 2] = x.z;
    if (4 * v + 3 < R) wa[4 * v + 3] = x.w;
  }
  const BlockGeom<K> geo(bl, blk + kTileAFp32);
#pragma unroll 1
  for (int g = 0; g < GROUPS; ++g) {
    const int up = geo.group_part(wave + g * NW);
    const int cbh = geo.lph ^ up;
    float bv[R];
#pragma unroll
    for (int s = 0; s < R; ++s) bv[s] = lds_load<float>(cbh ^ geo.ro[s]);
    // all reads first, one wait, then the chain: interleaved, each MFMA
    // would wait out an LDS round trip for its operand
    __builtin_amdgcn_sched_barrier(0);
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[s], bv[s], acc, 0, 0, 0);
    geo.store(geo.lp ^ up, acc);
  }
}

// One block in split-bf16 form (see the file comment).  Lane (l32, h)
// supplies K-index 8h + j of instruction c, which is input row s = 8c + j,
// part h of the B operand -- the same rows and plane as the FP32 form reads
// -- and the output rows are those of the FP32 form.  A wave takes G column
// groups at once: G independent accumulation chains, so one chain's split
// overlaps another's matrix-core work.
template <int K, int NTH, int TILE>
__device__ __forceinline__ void tp_block_split(const TileBlock *__restrict__ bl, int blk) {
  constexpr int R = 1 << K, NC = R > 8 ? 2 : 1, NW = NTH / 64;
  constexpr int GROUPS = (TILE >> K) / 32 / NW;
  constexpr int G = GROUPS >= 2 ? 2 : 1;
  static_assert(GROUPS >= 1 && GROUPS % G == 0, "column groups must split evenly");
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  bf16x8 wh[NC], wm[NC], wl[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    wh[c] = lds_load<bf16x8>(blk + ((0 * 2 + c) * 64 + lane) * 16);
    wm[c] = lds_load<bf16x8>(blk + ((1 * 2 + c) * 64 + lane) * 16);
    wl[c] = lds_load<bf16x8>(blk + ((2 * 2 + c) * 64 + lane) * 16);
  }
  const BlockGeom<K> geo(bl, blk + kTileASplit);
#pragma unroll 1
  for (int g0 = 0; g0 < GROUPS; g0 += G) {
    int up[G];
    float bv[G][8 * NC];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      up[g] = geo.group_part(wave + (g0 + g) * NW);
      const int cbh = geo.lph ^ up[g];
#pragma unroll
      for (int s = 0; s < 8 * NC; ++s)
        bv[g][s] = s < R ? lds_load<float>(cbh ^ geo.ro[s < R ? s : 0]) : 0.0f;
    }
    __builtin_amdgcn_sched_barrier(0);  // all B reads before the splits
    floatx16 acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g)
      acc[g] = floatx16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                        0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      bf16x8 xh[G], xl[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        u32x4 hi, lo;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          unsigned a, b;
          split_pair(bv[g][8 * c + 2 * p], bv[g][8 * c + 2 * p + 1], a, b);
          hi[p] = a;
          lo[p] = b;
        }
        xh[g] = __builtin_bit_cast(bf16x8, hi);
        xl[g] = __builtin_bit_cast(bf16x8, lo);
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl[c], xh[g], acc[g], 0, 0, 0);
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wm[c], xh[g], acc[g], 0, 0, 0);
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh[c], xl[g], acc[g], 0, 0, 0);
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh[c], xh[g], acc[g], 0, 0, 0);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) geo.store(geo.lp ^ up[g], acc[g]);
  }
}

// The descriptor pointers are __restrict__ so the compiler may read them
// through the scalar cache (they are uniform and never written by the
// kernel); vector loads of them would wait behind the tile prefetch.
template <int TB, int NTH, bool SPLIT>
__global__ __launch_bounds__(NTH) void k_tile_program(
    float2 *__restrict__ a, const TilePass *__restrict__ passes, int pass_index,
    const TileBlock *__restrict__ blocks, const float *__restrict__ wbuf,
    uint64_t ntiles) {
  constexpr int TILE = 1 << TB;
  constexpr int TBITS = __builtin_ctz(NTH) + 1;  // tile bits covered by 2*tid
  constexpr int IT = TILE / (2 * NTH);           // float4 loads per thread
  constexpr int IB = TB - TBITS;                 // tile bits covered by `it`
  constexpr int STRIDE = SPLIT ? kTileStrideSplit : kTileStrideFp32;
  constexpr int ABYTES = SPLIT ? kTileASplit : kTileAFp32;
  constexpr int OPS = int(sizeof(float2)) << TB;  // first block's LDS address
  static_assert(TB == kTileBitsFixed, "the plane offset assumes 13-bit tiles");
  extern __shared__ float4 smem[];
  float *re = reinterpret_cast<float *>(smem);
  float *im = re + TILE;

  const TilePass &ps = passes[pass_index];
  const int nb = ps.nblocks;
  const int tid = threadIdx.x;

  // stage the A operands and lane tables of every block
  for (int idx = tid; idx < nb * 64 * 16; idx += NTH) {
    const int b = idx / (64 * 16), ln = (idx / 16) % 64, s = idx % 16;
    const TileBlock &bl = blocks[ps.first_block + b];
    const int R = 1 << bl.k, RR = 2 * R, row = ln & 31, hh = ln >> 5;
    const float w = (row < RR && s < R)
                        ? wbuf[size_t(bl.wslot) * kTileWFloats + row * RR + 2 * s + hh]
                        : 0.0f;
    const int blk = OPS + b * STRIDE;
    if (!SPLIT) {
      // [s / 4][lane] float4, element s % 4
      lds_store<float>(blk + ((s >> 2) * 64 + ln) * 16 + (s & 3) * 4, w);
    } else {
      // part P of W[l32][2s + h], s = 8c + j: [P][c][lane] x 8 bf16
      const __bf16 ph = __bf16(w);
      const float r1 = w - float(ph);
      const __bf16 pm = __bf16(r1);
      const __bf16 pl = __bf16(r1 - float(pm));
      const int c = s >> 3, j = s & 7;
      lds_store<__bf16>(blk + ((0 * 2 + c) * 64 + ln) * 16 + 2 * j, ph);
      lds_store<__bf16>(blk + ((1 * 2 + c) * 64 + ln) * 16 + 2 * j, pm);
      lds_store<__bf16>(blk + ((2 * 2 + c) * 64 + ln) * 16 + 2 * j, pl);
    }
  }
  for (int idx = tid; idx < nb * 48; idx += NTH) {
    const int b = idx / 48, e = idx % 48;
    const TileBlock &bl = blocks[ps.first_block + b];
    const int v = e < 32 ? bl.lanecol[e] : bl.outoff[(e - 32) >> 3][(e - 32) & 7];
    lds_store<int>(OPS + b * STRIDE + ABYTES + 4 * e, v);
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

  // The next tile is loaded into registers while the current one is
  // computed.  At a tile boundary the finished tile is read out of LDS, the
  // prefetched one written in, and only then are the stores issued: stores
  // count against the same in-order vector-memory counter as loads, so a
  // wait for the prefetch placed after the stores would also wait for the
  // stores to reach memory.
  float4 v[IT];
  uint64_t base = tile_base(tile);
  if (tile < tend) {
#pragma unroll
    for (int it = 0; it < IT; ++it) v[it] = ld_nt(a + (base | it_addr(it) | addr_t));
  }
  __syncthreads();  // operands staged
  if (tile < tend) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int s0 = slot_t ^ swz(it << TBITS);  // slot of e; e+1 is s0 ^ 1
      re[s0] = v[it].x;
      im[s0] = v[it].y;
      re[s0 ^ 1] = v[it].z;
      im[s0 ^ 1] = v[it].w;
    }
  }
  __syncthreads();

  for (; tile < tend; tile += G8) {
    const uint64_t next = tile + G8;
    const uint64_t nbase = tile_base(next);
    if (next < tend) {
#pragma unroll
      for (int it = 0; it < IT; ++it)
        v[it] = ld_nt(a + (nbase | it_addr(it) | addr_t));
    }

    for (int b = 0; b < nb; ++b) {
      const TileBlock *bl = &blocks[ps.first_block + b];
      const int blk = OPS + b * STRIDE;
      if (SPLIT) {
        switch (bl->k) {
          case 1: tp_block_split<1, NTH, TILE>(bl, blk); break;
          case 2: tp_block_split<2, NTH, TILE>(bl, blk); break;
          case 3: tp_block_split<3, NTH, TILE>(bl, blk); break;
          default: tp_block_split<4, NTH, TILE>(bl, blk); break;
        }
      } else {
        switch (bl->k) {
          case 1: tp_block<1, NTH, TILE>(bl, blk); break;
          case 2: tp_block<2, NTH, TILE>(bl, blk); break;
          case 3: tp_block<3, NTH, TILE>(bl, blk); break;
          default: tp_block<4, NTH, TILE>(bl, blk); break;
        }
      }
      __syncthreads();
    }

    float4 o[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int s0 = slot_t ^ swz(it << TBITS);
      o[it] = make_float4(re[s0], im[s0], re[s0 ^ 1], im[s0 ^ 1]);
    }
    __syncthreads();  // the finished tile is out of LDS
    if (next < tend) {
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int s0 = slot_t ^ swz(it << TBITS);
        re[s0] = v[it].x;
        im[s0] = v[it].y;
        re[s0 ^ 1] = v[it].z;
        im[s0 ^ 1] = v[it].w;
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the stores after the LDS writes
#pragma unroll
    for (int it = 0; it < IT; ++it) st_nt(a + (base | it_addr(it) | addr_t), o[it]);
    __syncthreads();  // the next tile is in LDS
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

template <int NTH, bool SPLIT>
static int launch_tp(qcs_hip_state *s, const TilePass *dpass, int pi,
                     const TileBlock *dblocks, uint64_t ntiles, int nblocks) {
  // once per kernel instance (states may launch from separate threads)
  static std::once_flag once;
  static int attr_rc = 0;
  std::call_once(once, [] {
    const void *fn = reinterpret_cast<const void *>(&k_tile_program<kTileBitsFixed, NTH, SPLIT>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(kLdsPerCu));
    if (e != hipSuccess) {
      attr_rc = check_hip(e, "hipFuncSetAttribute");
      return;
    }
    // LDS is addressed by absolute byte address: the dynamic LDS must start
    // at 0, i.e. the kernel has no static LDS
    hipFuncAttributes fa;
    if ((e = hipFuncGetAttributes(&fa, fn)) != hipSuccess) {
      attr_rc = check_hip(e, "hipFuncGetAttributes");
      return;
    }
    if (fa.sharedSizeBytes != 0)
      attr_rc = fail(QCS_HIP_EDEVICE, "k_tile_program has %zu bytes of static LDS",
                     size_t(fa.sharedSizeBytes));
  });
  if (attr_rc) return attr_rc;
  const size_t lds =
      kTileLdsTile + size_t(SPLIT ? kTileStrideSplit : kTileStrideFp32) * size_t(nblocks);
  const uint64_t per_cu = kLdsPerCu / lds;  // work-groups that fit per CU
  uint64_t grid = uint64_t(s->tile_grid) * (per_cu > 1 ? 2 : 1);
  if (grid > ntiles) grid = ntiles;
  hipLaunchKernelGGL((k_tile_program<kTileBitsFixed, NTH, SPLIT>), dim3(unsigned(grid)),
                     dim3(NTH), lds, s->stream, static_cast<float2 *>(s->vec),
                     dpass, pi, dblocks, s->wbuf, ntiles);
  return launch_check("k_tile_program");
}

int tile_max_blocks(const qcs_hip_state *s) {
  return s->tile_math == QCS_HIP_TILE_MATH_SPLIT_BF16 ? QCS_HIP_TILE_MAX_BLOCKS_SPLIT
                                                      : QCS_HIP_TILE_MAX_BLOCKS;
}

// the block's descriptor: target positions, row offsets, lane tables
static void make_tile_block(const qcs_hip_pass &ps, const qcs_hip_block &blk, int wslot,
                            TileBlock *tb) {
  std::memset(tb, 0, sizeof(TileBlock));
  const int k = blk.k, R = 1 << k;
  tb->k = k;
  tb->wslot = wslot;
  for (int q = 0; q < k; ++q) {
    int pos = 0;
    while (ps.bits[pos] != blk.targets[q]) ++pos;
    tb->tpos[q] = pos;
  }
  for (int r = 0; r < R; ++r) {
    int off = 0;
    for (int q = 0; q < k; ++q)
      if ((r >> q) & 1) off |= 1 << tb->tpos[q];
    tb->rowoff[r] = 4 * swz(off);
  }
  for (int l = 0; l < 32; ++l) {
    int c = l;
    for (int q = 0; q < k; ++q) c = ins0(c, tb->tpos[q]);
    tb->lanecol[l] = 4 * swz(c);
  }
  for (int h = 0; h < 2; ++h)
    for (int q = 0; q < 4; ++q)
      for (int j = 0; j < 2; ++j) {
        const int r = 4 * q + 2 * h + j;
        tb->outoff[h][2 * q + j] = r < R ? tb->rowoff[r] : 0;
      }
}

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
         b < passes[p].first_block + passes[p].num_blocks; ++b)
      make_tile_block(passes[p], blocks[b], b, &tb[b]);
  }
  QCS_HIP_TRY(hipMemcpyAsync(s->tpass_dev, tp, pbytes, hipMemcpyHostToDevice,
                             s->stream));
  QCS_HIP_TRY(hipMemcpyAsync(s->tblock_dev, tb, tbytes, hipMemcpyHostToDevice,
                             s->stream));
  const TilePass *dp = static_cast<const TilePass *>(s->tpass_dev);
  const TileBlock *db = static_cast<const TileBlock *>(s->tblock_dev);
  const bool split = s->tile_math == QCS_HIP_TILE_MATH_SPLIT_BF16;
  const double N = double(s->size);
  for (int p = 0; p < np; ++p) {
    const uint64_t ntiles = s->size >> passes[p].tile_bits;
    double flops = 0.0;
    for (int b = passes[p].first_block;
         b < passes[p].first_block + passes[p].num_blocks; ++b)
      flops += 8.0 * double(1 << blocks[b].k) * N;
    TimedLaunch tl;
    timing_begin(s, QCS_HIP_K_FUSED, 2.0 * N * double(s->amp_bytes), flops, &tl);
    const int nbp = passes[p].num_blocks;
    if (s->tile_threads == 1024)
      rc = split ? launch_tp<1024, true>(s, dp, p, db, ntiles, nbp)
                 : launch_tp<1024, false>(s, dp, p, db, ntiles, nbp);
    else
      rc = split ? launch_tp<512, true>(s, dp, p, db, ntiles, nbp)
                 : launch_tp<512, false>(s, dp, p, db, ntiles, nbp);
    if (rc) return rc;
    if ((rc = timing_end(s, &tl))) return rc;
  }
  // the staging buffer holds the descriptors until the copies have run
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

}  // namespace qcs
