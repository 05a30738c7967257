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
// block is 4 swz(c) ^ 4 swz(row bits of s), and since the kernel has no
// static LDS (checked on the host) its dynamic LDS starts at address 0:
// each B-operand read or result write costs one XOR and no base add.
//
// Work split.  A block's column bits (the tile bits that are not targets)
// are split on the host into 5 lane bits, the wave bits (the wave index) and
// group bits (the wave's column groups), each part a table of swz terms
// (TileBlock).  Each wave then touches only the amplitudes whose wave bits
// hold its index, so consecutive blocks given the same wave bits need no
// barrier between them: the host takes a pass's blocks in runs whose targets
// leave enough tile bits free and gives each run one choice of wave bits,
// the one with the fewest LDS bank conflicts over its blocks (the 32 columns
// a half-wave reads fall in distinct banks when the lane bits land on
// distinct low slot bits; a shared choice can leave a block 2-way
// conflicted, which costs less than a barrier).
//
// After the tile come, per block of the pass, its A operands in lane order
// and the lane table (the lane part of the column address).  The FP32 form
// reads a block's operands and lane part, and its scalars (row offsets of
// rows 1, 2, 4, 8 -- the rest follow by linearity -- the wave part,
// the width and the barrier flag), while the previous block's matrix-core
// chain runs.  The operands stay in LDS because on gfx9 every vector memory
// load retires through one in-order counter, so a global load inside the
// tile loop would wait behind the next tile's prefetch.
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
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

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
typedef int i32x2 __attribute__((ext_vector_type(2)));

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
constexpr int kTileAux = 128;                        // lanecol[32]
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
  int direct;  // 1: the first block's operands come from HBM and the last
               // block's results go to HBM (the three-product kernel)
  int sbits[kPassMaxBits];
};

// A block's columns (the tile bits that are not targets) split into three
// sets of bit positions, chosen on the host: 5 lane bits (bit i of the
// lane's l32), the wave bits (bit i of the wave index) and the group bits
// (bit i of the wave's column-group counter j).  By linearity of swz the
// byte address of (row r, lane, wave, j) is
//   rowoff[r] ^ lanecol[l32] ^ wpart[wave] ^ jpart[j]   (+ PLANE for imag).
struct TileBlock {
  int k;
  int wslot;
  int sync;                  // 1: barrier after this block (the next block's
                             //    waves own other amplitudes)
  int early;                 // 1: k = 4 in two pipelined column halves, and
                             //    the next block is too, in the same run and
                             //    with the same half bit: its first half's
                             //    B operands are read during this block
  int rowoff[kTileRMax];     // 4 swz(row bits of s): byte offset of row s
  int wpart[16];             // 4 swz(wave index on the wave bits)
  int jpart[16];             // 4 swz(j on the group bits)
  int lanecol[32];           // 4 swz(l32 on the lane bits)
  // the same tables as tile element indices (no swizzle, not in bytes), for
  // the first and last blocks of a direct pass, whose operands and results
  // move between registers and HBM
  int elane[32];             // l32 on the lane bits
  int ewave[16];             // wave index on the wave bits
  int erow[kTileKMax];       // target q
  // row bit q of the kernel's row patterns is bit tidx[q] of the composed
  // unitary's (whose bit j is the j-th lowest target): the pipelined k = 4
  // forms put the target that minimises bank conflicts on row bit 0
  int tidx[kTileKMax];
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

// wait for every outstanding LDS access and scalar load of the wave
__device__ __forceinline__ void lgkm_wait_all() {
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0); vmcnt and expcnt not waited for
  __builtin_amdgcn_sched_barrier(0);
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

// The uniform part of a block (scalar registers), read one block ahead so
// that the scalar-cache latency is hidden behind the current block.  Row
// offsets are linear in the row bits (swz is), so the offsets of rows 1, 2,
// 4 and 8 give all of them.
struct BlockScalars {
  int k, sync, up;  // up: the wave's part of the column address
  int early;
  int rb[kTileKMax];
  __device__ __forceinline__ int ro(int s) const {
    int o = 0;
#pragma unroll
    for (int q = 0; q < kTileKMax; ++q)
      if ((s >> q) & 1) o ^= rb[q];
    return o;
  }
};

__device__ __forceinline__ BlockScalars load_scalars(const TileBlock *__restrict__ bl,
                                                     int wave) {
  // (readfirstlane: the compiler otherwise carries them in vector registers
  // from one block to the next)
  BlockScalars d;
  d.k = __builtin_amdgcn_readfirstlane(bl->k);
  d.sync = __builtin_amdgcn_readfirstlane(bl->sync);
  d.early = __builtin_amdgcn_readfirstlane(bl->early);
  d.up = __builtin_amdgcn_readfirstlane(bl->wpart[wave]);
#pragma unroll
  for (int q = 0; q < kTileKMax; ++q)
    d.rb[q] = __builtin_amdgcn_readfirstlane(bl->rowoff[1 << q]);
  return d;
}

// The per-lane values of an FP32 block kept in LDS -- its A operands and the
// lane part of the column address -- read one block ahead, while the
// current block's matrix-core chain runs.
struct BlockRegs {
  f4v a[4];
  int lp;   // lane table entry l32
  int lp2;  // lane table entry l32 ^ 16 (the 16-column form of k <= 3 blocks)
};

__device__ __forceinline__ void load_regs(BlockRegs &r, int blk) {
  const int lane = threadIdx.x & 63, l32 = lane & 31;
#pragma unroll
  for (int v = 0; v < 4; ++v) r.a[v] = lds_load<f4v>(blk + (v * 64 + lane) * 16);
  r.lp = lds_load<int>(blk + kTileAFp32 + 4 * l32);
  r.lp2 = lds_load<int>(blk + kTileAFp32 + 4 * (l32 ^ 16));
}

// The pipelined k = 4 form reads the lane's column in each half directly:
// lp = lane table entry (lane & 15), lp2 = entry (lane & 15) | 16.
__device__ __forceinline__ void load_regs4p(BlockRegs &r, int blk) {
  const int lane = threadIdx.x & 63, c = lane & 15;
#pragma unroll
  for (int v = 0; v < 4; ++v) r.a[v] = lds_load<f4v>(blk + (v * 64 + lane) * 16);
  r.lp = lds_load<int>(blk + kTileAFp32 + 4 * c);
  r.lp2 = lds_load<int>(blk + kTileAFp32 + 4 * (c | 16));
}

// What every block kernel needs of its block: uniform row offsets and the
// wave's column part (scalar registers), the lane's column part and output
// offsets.
template <int K>
struct BlockGeom {
  static constexpr int R = 1 << K;
  int ro[R];
  int upw;      // wave part of the column address
  int lph;      // lane part of the column address in the lane's B plane
  int lpo;      // lane part of the output address: lanes of the upper half
                // hold rows 4q + 2 + j, i.e. rowoff[4q + j] ^ rowoff[2]
  __device__ __forceinline__ BlockGeom(const BlockScalars &d, int lp) {
    const int h = (threadIdx.x >> 5) & 1;
#pragma unroll
    for (int s = 0; s < R; ++s) ro[s] = d.ro(s);
    upw = d.up;
    lph = lp | (h ? kTilePlane : 0);
    lpo = lp ^ (h ? d.rb[1] : 0);
  }
  // from the lane table at LDS address aux
  static __device__ __forceinline__ BlockGeom from_lds(const BlockScalars &d, int aux) {
    return BlockGeom(d, lds_load<int>(aux + 4 * (threadIdx.x & 31)));
  }
  // wave-uniform part of the address of the wave's column group j
  template <int GROUPS>
  __device__ __forceinline__ int group_part(const TileBlock *__restrict__ bl, int j) const {
    if (GROUPS == 1) return upw;
    return upw ^ bl->jpart[j];
  }
  // result rows of a 32x32 accumulator tile (rows 4q + 2h + j of the block);
  // cb = lpo ^ the wave-uniform column part
  __device__ __forceinline__ void store(int cb, const floatx16 &acc) const {
    const int h = (threadIdx.x >> 5) & 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (4 * q >= R) continue;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (4 * q + 2 * h + j < R) {
          const int o = cb ^ ro[4 * q + j];
          lds_store<float>(o, acc[4 * q + 2 * j]);
          lds_store<float>(o + kTilePlane, acc[4 * q + 2 * j + 1]);
        }
      }
    }
  }
};

// One block in FP32: lane (l32, h) supplies A[l32][2s + h] (output row
// l32/2, part l32&1; input row s, part h) and B[2s + h][column l32] -- the
// lane's column of plane h, row s -- for s = 0 .. R-1, one MFMA each.  The
// next block's scalars and per-lane values (nd, nr) are read once the first
// column group's chain is issued.
template <int K, int NTH, int TILE>
__device__ __forceinline__ void tp_block(const TileBlock *__restrict__ bl,
                                         const BlockScalars &d, const BlockRegs &r,
                                         const TileBlock *__restrict__ nbl, int nblk,
                                         int wave, BlockScalars &nd, BlockRegs &nr) {
  constexpr int R = 1 << K, NW = NTH / 64, GROUPS = (TILE >> K) / 32 / NW;
  static_assert(GROUPS >= 1 && GROUPS <= 16, "a wave takes 1..16 column groups");
  float wa[R];
#pragma unroll
  for (int s = 0; s < R; ++s) wa[s] = r.a[s >> 2][s & 3];
  const BlockGeom<K> geo(d, r.lp);
#pragma unroll 1
  for (int g = 0; g < GROUPS; ++g) {
    const int up = geo.template group_part<GROUPS>(bl, g);
    const int cbh = geo.lph ^ up;
    // Nothing else may be outstanding on the LDS/scalar counter while the
    // B reads return (scalar loads return out of order, so the compiler
    // would wait for all of them before the first MFMA); the reads go out
    // in the order the chain uses them, so each MFMA waits only for its own.
    lgkm_wait_all();
    float bv[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      bv[s] = lds_load<float>(cbh ^ geo.ro[s]);
      __builtin_amdgcn_sched_barrier(0);
    }
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < R; ++s)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[s], bv[s], acc, 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
    if (g == 0) {
      nd = load_scalars(nbl, wave);
      load_regs(nr, nblk);
    }
    geo.store(geo.lpo ^ up, acc);
  }
}

// One block of k <= 3 qubits in FP32 on v_mfma_f32_16x16x4f32.  The 32x32
// form pads a block's 2^(k+1) real output rows to 32, wasting half the
// matrix-core work of a k = 3 block and three quarters of a k = 2 one.  Here
// a wave's 32-column group runs as two 16-column halves (lane bit 4 of the
// column picks the half; lc0 and lc1 are the lane's column in each), and
// lane l = 16q + c supplies A[l & 15][q] and B[q][c] with K-index q meaning
// input row 2sp + (q & 1), part q >> 1, for instruction sp = 0 .. R/2 - 1.
// The lane receives rows 4q + v of the 16x16 result, laid out as output row
// (q & 1) | (v >> 1) << 1 | (q >> 1) << 2, part v & 1 (the host stages A in
// that row order).  So each half-wave reads, and writes, one plane at two
// rows that differ in the block's first target bit: it is conflict-free when
// that bit's bank bit is none of the four low lane bits' (the host picks the
// lanes so).  The next block's values are read before the first column
// group when there are several (after the chain when there is one).
template <int K, int NTH, int TILE>
__device__ __forceinline__ void tp_block16(const TileBlock *__restrict__ bl,
                                           const BlockScalars &d, const BlockRegs &r,
                                           const TileBlock *__restrict__ nbl, int nblk,
                                           int wave, BlockScalars &nd, BlockRegs &nr) {
  constexpr int R = 1 << K, NI = R / 2, NW = NTH / 64, GROUPS = (TILE >> K) / 32 / NW;
  static_assert(K >= 1 && K <= 3, "the 16-column form takes k = 1..3");
  static_assert(GROUPS >= 1 && GROUPS <= 16, "a wave takes 1..16 column groups");
  const int lane = threadIdx.x & 63, q = lane >> 4;
  float wa[NI];
  int ro2[NI];
#pragma unroll
  for (int sp = 0; sp < NI; ++sp) {
    wa[sp] = r.a[0][sp];
    ro2[sp] = d.ro(2 * sp);
  }
  const int lc0 = (lane & 16) ? r.lp2 : r.lp, lc1 = (lane & 16) ? r.lp : r.lp2;
  const int rq = (q & 1) ? d.rb[0] : 0;
  const int bq = rq ^ ((q >> 1) ? kTilePlane : 0);  // B: row 2sp + (q & 1), part q >> 1
  const int oq = rq ^ ((q >> 1) ? d.rb[2] : 0);     // result: rows (q & 1) | (q >> 1) << 2
  if (GROUPS > 1) {
    nd = load_scalars(nbl, wave);
    load_regs(nr, nblk);
  }
#pragma unroll 1
  for (int g = 0; g < GROUPS; ++g) {
    const int up = GROUPS == 1 ? d.up : d.up ^ bl->jpart[g];
    const int b0 = lc0 ^ bq ^ up, b1 = lc1 ^ bq ^ up;
    float bv0[NI], bv1[NI];
#pragma unroll
    for (int sp = 0; sp < NI; ++sp) {
      bv0[sp] = lds_load<float>(b0 ^ ro2[sp]);
      bv1[sp] = lds_load<float>(b1 ^ ro2[sp]);
    }
    __builtin_amdgcn_sched_barrier(0);
    f4v acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sp = 0; sp < NI; ++sp) {
      acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[sp], bv0[sp], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[sp], bv1[sp], acc1, 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (GROUPS == 1) {
      nd = load_scalars(nbl, wave);
      load_regs(nr, nblk);
    }
    const int o0 = lc0 ^ oq ^ up, o1 = lc1 ^ oq ^ up;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = (q & 1) | ((v >> 1) << 1) | ((q >> 1) << 2);
      if (row < R) {
        const int ov = ((v >> 1) ? d.rb[1] : 0) ^ ((v & 1) ? kTilePlane : 0);
        lds_store<float>(o0 ^ ov, acc0[v]);
        lds_store<float>(o1 ^ ov, acc1[v]);
      }
    }
  }
}

// A k = 4 block on v_mfma_f32_16x16x4f32 as two 16-column halves, software
// pipelined (1024 threads: one 32-column group per wave).  The 32x32 chain
// of tp_block leaves the matrix cores idle while a wave stores its results
// and reads the next block's operands, and the waves of a SIMD do that at
// the same time.  Here a wave keeps its own matrix-core queue filled:
//   half-0 chain | previous block's half-1 stores, half-1 B reads, next
//                |   block's scalars
//   half-1 chain | next block's A operands and lane table, half-0 stores,
//                |   and (early) the next block's half-0 B reads
// The half-1 results are stored at the start of the next block, behind its
// half-0 chain, when that block is a k = 4 block of the same run with the
// same half bit (lane bit 4): its first half then reads only amplitudes that
// this block's first half wrote, or that no block of the run wrote, so its
// B reads go out before this block's second half is stored.  Otherwise the
// half-1 results are stored (and, at the end of a run, the work-group
// waits) before the next block's B reads.  Output rows of
// row block rb are those of tp_block16 with target bit 3 = rb; the A operands
// of instruction sp of row block rb sit at index 8 rb + sp.
struct K4State {
  float bv0[8];  // first-half B operands of the current block, read ahead
  f4v acc[2];    // second-half results of the previous block, not yet stored
  int po, pr1, pr3;  // their base address, row offsets of target bits 1, 3
  int pend;
};

// One column half's two 16x16 accumulator tiles (row blocks 0 and 1) to LDS.
// Lane (q, c) holds rows (q & 1) | (v >> 1) << 1 | (q >> 1) << 2 | rb << 3,
// part v & 1; `o` is the lane's column ^ the wave part ^ the row offsets q
// selects.  Every term is below the plane offset, so re and im pairs are one
// two-address store.
__device__ __forceinline__ void store_k4_half(int o, int r1, int r3, const f4v &a0,
                                              const f4v &a1) {
#pragma unroll
  for (int rbk = 0; rbk < 2; ++rbk) {
#pragma unroll
    for (int vh = 0; vh < 2; ++vh) {
      const int addr = o ^ (vh ? r1 : 0) ^ (rbk ? r3 : 0);
      lds_store<float>(addr, (rbk ? a1 : a0)[2 * vh]);
      lds_store<float>(addr + kTilePlane, (rbk ? a1 : a0)[2 * vh + 1]);
    }
  }
}

// First-half B operands of a pipelined k = 4 block into st.bv0.
// Lane (q, c) of a 16x16x4 instruction: q & 1 selects the row of target
// bit 0, q >> 1 the plane (B) or the row of target bit 2 (results); as
// all-ones masks, so a row offset is selected with one AND.
__device__ __forceinline__ int q1_mask() { return -int((threadIdx.x >> 4) & 1); }
__device__ __forceinline__ int q2_mask() { return -int((threadIdx.x >> 5) & 1); }

// First-half B operands of a pipelined k = 4 block into st.bv0.
__device__ __forceinline__ void read_k4_b0(const BlockScalars &d, const BlockRegs &r,
                                           K4State &st) {
  const int bq = (d.rb[0] & q1_mask()) ^ (kTilePlane & q2_mask()) ^ d.up;
#pragma unroll
  for (int sp = 0; sp < 8; ++sp) st.bv0[sp] = lds_load<float>(r.lp ^ bq ^ d.ro(2 * sp));
}

template <int NTH, int TILE>
__device__ __forceinline__ void tp_block4p(const BlockScalars &d, const BlockRegs &r,
                                           const TileBlock *__restrict__ nbl, int nblk,
                                           int wave, BlockScalars &nd, BlockRegs &nr,
                                           K4State &st) {
  static_assert((TILE >> 4) / 32 / (NTH / 64) == 1, "one 32-column group per wave");
  float wa[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) wa[s] = r.a[s >> 2][s & 3];
  int ro2[8];
#pragma unroll
  for (int sp = 0; sp < 8; ++sp) ro2[sp] = d.ro(2 * sp);
  const int lc0 = r.lp, lc1 = r.lp2;  // the lane's column in each half
  const int rq = d.rb[0] & q1_mask();
  const int bq = rq ^ (kTilePlane & q2_mask()) ^ d.up;  // B: row 2sp + (q & 1), part q >> 1
  const int oq = rq ^ (d.rb[2] & q2_mask()) ^ d.up;     // results: rows (q & 1) | (q >> 1) << 2
  __builtin_amdgcn_sched_barrier(0);
  const f4v z = {0.f, 0.f, 0.f, 0.f};
  f4v a00 = z, a01 = z;
#pragma unroll
  for (int sp = 0; sp < 8; ++sp) {
    a00 = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[sp], st.bv0[sp], a00, 0, 0, 0);
    a01 = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[8 + sp], st.bv0[sp], a01, 0, 0, 0);
  }
  __builtin_amdgcn_sched_barrier(0);
  if (st.pend) store_k4_half(st.po, st.pr1, st.pr3, st.acc[0], st.acc[1]);
  float bv1[8];
#pragma unroll
  for (int sp = 0; sp < 8; ++sp) bv1[sp] = lds_load<float>(lc1 ^ bq ^ ro2[sp]);
  nd = load_scalars(nbl, wave);
  __builtin_amdgcn_sched_barrier(0);
  // the second half accumulates straight into the pending registers (the
  // previous block's are stored above)
  st.acc[0] = z;
  st.acc[1] = z;
#pragma unroll
  for (int sp = 0; sp < 8; ++sp) {
    st.acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[sp], bv1[sp], st.acc[0], 0, 0, 0);
    st.acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(wa[8 + sp], bv1[sp], st.acc[1], 0, 0, 0);
  }
  __builtin_amdgcn_sched_barrier(0);
  load_regs4p(nr, nblk);
  store_k4_half(lc0 ^ oq, d.rb[1], d.rb[3], a00, a01);
  st.po = lc1 ^ oq;
  st.pr1 = d.rb[1];
  st.pr3 = d.rb[3];
  st.pend = d.early;
  // Unless the next block may read ahead, this block's second half is
  // stored now and, at the end of a run, the work-group waits; either way
  // the next block's first-half inputs are final below.  (After a pass's
  // last block the reads serve nothing: the next tile's first block reads
  // again once the tile is in LDS.)
  if (!d.early) {
    store_k4_half(st.po, st.pr1, st.pr3, st.acc[0], st.acc[1]);
    if (d.sync) __syncthreads();
  }
  read_k4_b0(nd, nr, st);
}

// The three-product form of a pipelined k = 4 block.  With U = Ur + i Ui
// and a column x = xr + i xi, and the products
//   P1 = (Ur + Ui) xr,   P2 = (Ur - Ui) xi,   P3 = -Ui (xr + xi),
// U x = (P1 + P3) + i (P2 - P3): three real 16x16 products (four 16x16x4
// instructions each) where the other forms do four.  The matrix cores also
// do the additions: P3's chain runs first, and the chains of P1 and -P2
// start from its result, so a block yields Re(U x) and -Im(U x) with no
// vector ALU work beyond xr + xi.  Inside a pass the tile's imaginary plane
// therefore holds -Im after every block: a block reads v = s Im(x), with
// s = +1 for the pass's first block (the tile as loaded) and -1 after, and
// its operands are (Ur + Ui), -s (Ur - Ui) and -Ui, with xr + xi = xr + s v.
// The tile is written back with its imaginary plane negated.
//
// Lane (q, c) supplies B[q][c] = row q | t << 2 of column c for instruction
// t, and receives rows 4q + v of each 16x16 result, which the host stages to
// mean output row q | v << 2: a lane reads and writes the rows that lane
// bits 4 and 5 (q) select on target bits 0 and 1.  The operands are a[p][t]
// for product p in the order above; in place of a fourth, the lane's
// address in each column half (lane table entry ^ the row offsets q
// selects), staged per block, so a block needs one XOR with the wave part
// per half.
//
// Schedule.  Both halves' B operands of a block are read at the end of the
// block before it, so its two chains (24 instructions) issue back to back
// with no LDS wait between them; the next block's lane addresses are read
// at the top of the block, its scalars during the first chain and its A
// operands once the second chain is issued.  Under the `early` rule (the
// next block is in the same run, with the same half bit: its first half
// reads only what this block's first half wrote) the next block's
// first-half reads go out before this block's second half is stored.
__device__ __forceinline__ void load_a4g(BlockRegs &r, int blk) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int v = 0; v < 3; ++v) r.a[v] = lds_load<f4v>(blk + (v * 64 + lane) * 16);
}
__device__ __forceinline__ i32x2 load_lanes4g(int blk) {
  return lds_load<i32x2>(blk + (3 * 64 + (threadIdx.x & 63)) * 16);
}
__device__ __forceinline__ void load_regs4g(BlockRegs &r, int blk) {
  load_a4g(r, blk);
  const i32x2 l = load_lanes4g(blk);
  r.lp = l[0];
  r.lp2 = l[1];
}

// both halves' B operands of a block, read ahead
struct K4GState {
  float b0[8], b1[8];
};

// B operands of one column half (l: the lane's address in it): real parts
// of rows q | t << 2 in b[t], imaginary parts in b[4 + t]
__device__ __forceinline__ void read_k4g(const BlockScalars &d, int l, float *b) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int o = l ^ d.ro(t << 2);
    b[t] = lds_load<float>(o);
    b[4 + t] = lds_load<float>(o + kTilePlane);
  }
}

__device__ __forceinline__ void read_k4g_both(const BlockScalars &d, const BlockRegs &r,
                                              K4GState &st) {
  read_k4g(d, r.lp ^ d.up, st.b0);
  read_k4g(d, r.lp2 ^ d.up, st.b1);
}

// the three products of one column half (b: real parts of the B rows, then
// s times the imaginary parts): re = Re(U x), nim = -Im(U x)
__device__ __forceinline__ void k4g_chain(const BlockRegs &r, const float *b, float sgn,
                                          f4v &re, f4v &nim) {
  f4v p3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 4; ++t)
    p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(r.a[2][t], __builtin_fmaf(sgn, b[4 + t], b[t]),
                                              p3, 0, 0, 0);
  re = p3;
  nim = p3;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    re = __builtin_amdgcn_mfma_f32_16x16x4f32(r.a[0][t], b[t], re, 0, 0, 0);
    nim = __builtin_amdgcn_mfma_f32_16x16x4f32(r.a[1][t], b[4 + t], nim, 0, 0, 0);
  }
}

// rows q | v << 2 of one half to LDS; r2, r3: row offsets of target bits 2, 3
__device__ __forceinline__ void store_k4g(int o, int r2, int r3, const f4v &re,
                                          const f4v &im) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int addr = o ^ ((v & 1) ? r2 : 0) ^ ((v & 2) ? r3 : 0);
    lds_store<float>(addr, re[v]);
    lds_store<float>(addr + kTilePlane, im[v]);
  }
}

template <int NTH, int TILE>
__device__ __forceinline__ void tp_block4g(const BlockScalars &d, const BlockRegs &r,
                                           float sgn, const TileBlock *__restrict__ nbl,
                                           int nblk, int wave, BlockScalars &nd, BlockRegs &nr,
                                           K4GState &st) {
  static_assert((TILE >> 4) / 32 / (NTH / 64) == 1, "one 32-column group per wave");
  const int l0 = r.lp ^ d.up, l1 = r.lp2 ^ d.up;  // the lane's address in each half
  const i32x2 nl = load_lanes4g(nblk);
  __builtin_amdgcn_sched_barrier(0);
  f4v re0, nim0, re1, nim1;
  k4g_chain(r, st.b0, sgn, re0, nim0);
  __builtin_amdgcn_sched_barrier(0);
  nd = load_scalars(nbl, wave);
  __builtin_amdgcn_sched_barrier(0);
  k4g_chain(r, st.b1, sgn, re1, nim1);
  __builtin_amdgcn_sched_barrier(0);
  load_a4g(nr, nblk);
  nr.lp = nl[0];
  nr.lp2 = nl[1];
  store_k4g(l0, d.rb[2], d.rb[3], re0, nim0);
  if (d.early) {
    read_k4g(nd, nr.lp ^ nd.up, st.b0);
    store_k4g(l1, d.rb[2], d.rb[3], re1, nim1);
    read_k4g(nd, nr.lp2 ^ nd.up, st.b1);
  } else {
    // (after a pass's last block the reads serve nothing: the next tile's
    // first block reads again once the tile is in LDS)
    store_k4g(l1, d.rb[2], d.rb[3], re1, nim1);
    if (d.sync) __syncthreads();
    read_k4g_both(nd, nr, st);
  }
}

// The last block of a direct pass: the three-product block with its results
// written to HBM (rows q | v << 2 of the lane's column, both halves in one
// 16-byte store, at `base` | the lane's offset o | the row offsets of target
// bits 2 and 3), the imaginary parts negated back; the next tile's first
// block's scalars, A operands and lane addresses are read as usual.
template <int NTH, int TILE>
__device__ __forceinline__ void tp_block4g_out(const BlockScalars &d, const BlockRegs &r,
                                               float sgn, const TileBlock *__restrict__ nbl,
                                               int nblk, int wave, BlockScalars &nd,
                                               BlockRegs &nr, const K4GState &st, float2 *a,
                                               uint64_t base, uint64_t o, uint64_t m2,
                                               uint64_t m3, bool acc_norm, double &nacc) {
  static_assert((TILE >> 4) / 32 / (NTH / 64) == 1, "one 32-column group per wave");
  const i32x2 nl = load_lanes4g(nblk);
  __builtin_amdgcn_sched_barrier(0);
  f4v re0, nim0, re1, nim1;
  k4g_chain(r, st.b0, sgn, re0, nim0);
  __builtin_amdgcn_sched_barrier(0);
  nd = load_scalars(nbl, wave);
  __builtin_amdgcn_sched_barrier(0);
  k4g_chain(r, st.b1, sgn, re1, nim1);
  __builtin_amdgcn_sched_barrier(0);
  load_a4g(nr, nblk);
  nr.lp = nl[0];
  nr.lp2 = nl[1];
#pragma unroll
  for (int v = 0; v < 4; ++v)
    st_nt(a + (base | o | ((v & 1) ? m2 : 0) | ((v & 2) ? m3 : 0)),
          make_float4(re0[v], -nim0[v], re1[v], -nim1[v]));
  if (acc_norm) {
    // |x|^2 of the values written, in FP64, in a fixed order
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double x0 = re0[v], y0 = nim0[v], x1 = re1[v], y1 = nim1[v];
      nacc = nacc + (x0 * x0 + y0 * y0);
      nacc = nacc + (x1 * x1 + y1 * y1);
    }
  }
}

// One FP32 block of the pass (index b), with the next block's values read
// into (nd, nr); the block after the last is the pass's first (next tile).
template <int NTH, int TILE>
__device__ __forceinline__ void tp_step(const TileBlock *__restrict__ blocks, int b, int nb,
                                        int ops, int wave, const BlockScalars &d,
                                        const BlockRegs &r, BlockScalars &nd, BlockRegs &nr) {
  const int bn = b + 1 < nb ? b + 1 : 0;
  const TileBlock *bl = blocks + b, *nbl = blocks + bn;
  const int nblk = ops + bn * kTileStrideFp32;
  switch (d.k) {
    case 1: tp_block16<1, NTH, TILE>(bl, d, r, nbl, nblk, wave, nd, nr); break;
    case 2: tp_block16<2, NTH, TILE>(bl, d, r, nbl, nblk, wave, nd, nr); break;
    case 3: tp_block16<3, NTH, TILE>(bl, d, r, nbl, nblk, wave, nd, nr); break;
    default: tp_block<4, NTH, TILE>(bl, d, r, nbl, nblk, wave, nd, nr); break;
  }
  // consecutive blocks whose waves own the same amplitudes (same wave bits)
  // need no barrier between them: each wave's LDS accesses are in order
  if (d.sync) __syncthreads();
}

// One block in split-bf16 form (see the file comment).  Lane (l32, h)
// supplies K-index 8h + j of instruction c, which is input row s = 8c + j,
// part h of the B operand -- the same rows and plane as the FP32 form reads
// -- and the output rows are those of the FP32 form.  A wave takes G column
// groups at once: G independent accumulation chains, so one chain's split
// overlaps another's matrix-core work.
template <int K, int NTH, int TILE>
__device__ __forceinline__ void tp_block_split(const TileBlock *__restrict__ bl,
                                               const BlockScalars &d, int blk) {
  constexpr int R = 1 << K, NC = R > 8 ? 2 : 1, NW = NTH / 64;
  constexpr int GROUPS = (TILE >> K) / 32 / NW;
  constexpr int G = GROUPS >= 2 ? 2 : 1;
  static_assert(GROUPS >= 1 && GROUPS <= 16 && GROUPS % G == 0,
                "column groups must split evenly");
  const int lane = threadIdx.x & 63;
  bf16x8 wh[NC], wm[NC], wl[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    wh[c] = lds_load<bf16x8>(blk + ((0 * 2 + c) * 64 + lane) * 16);
    wm[c] = lds_load<bf16x8>(blk + ((1 * 2 + c) * 64 + lane) * 16);
    wl[c] = lds_load<bf16x8>(blk + ((2 * 2 + c) * 64 + lane) * 16);
  }
  const BlockGeom<K> geo = BlockGeom<K>::from_lds(d, blk + kTileASplit);
#pragma unroll 1
  for (int g0 = 0; g0 < GROUPS; g0 += G) {
    int up[G];
    float bv[G][8 * NC];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      up[g] = geo.template group_part<GROUPS>(bl, g0 + g);
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
    for (int g = 0; g < G; ++g) geo.store(geo.lpo ^ up[g], acc[g]);
  }
}

// The work-group's FP64 sum of v to norm[blockIdx.x * kRedSlots]: lanes by
// the wave's fixed xor tree, then the waves in order (the tile's LDS is free
// once its last tile is written)
template <int NTH>
__device__ __forceinline__ void wg_norm_out(double v, double *norm) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds_store<double>(8 * int(threadIdx.x >> 6), v);
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NTH / 64; ++w) t = t + lds_load<double>(8 * w);
    norm[size_t(blockIdx.x) * kRedSlots] = t;
  }
}

// a kernel row pattern as a row of the composed unitary (TileBlock::tidx)
__device__ __forceinline__ int sigma4(const TileBlock &bl, int r) {
  int x = 0;
#pragma unroll
  for (int q = 0; q < kTileKMax; ++q)
    if ((r >> q) & 1) x |= 1 << bl.tidx[q];
  return x;
}

// The descriptor pointers are __restrict__ so the compiler may read them
// through the scalar cache (they are uniform and never written by the
// kernel); vector loads of them would wait behind the tile prefetch.
// K4: how FP32 k = 4 blocks run -- 0: the 32x32 chain (k <= 3 blocks on
// the 16-column form), 1: pipelined halves, 2: pipelined three-product
// halves (1 and 2 run every block as a k = 4 block), 3: 2 on a direct pass
// (the first block's operands loaded from HBM into registers, the last
// block's results stored from registers to HBM: no tile fill or read-out
// through LDS)
template <int TB, int NTH, bool SPLIT, int K4>
__global__ __launch_bounds__(NTH) void k_tile_program(
    float2 *__restrict__ a, const TilePass *__restrict__ passes, int pass_index,
    const TileBlock *__restrict__ blocks, const float *__restrict__ wbuf,
    uint64_t ntiles, double *__restrict__ norm, const double *__restrict__ fsum) {
  constexpr int TILE = 1 << TB;
  constexpr bool DIR = K4 == 3;
  constexpr int KF = DIR ? 2 : K4;  // the block form
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
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

  // fsum (three-product form only): the FP64 norm_sq of the pass's input;
  // the first block's products are scaled by 1/sqrt(norm_sq) under
  // q_state_normalize's rule, so the pass also renormalises
  double fscale = 1.0;
  if (!SPLIT && KF == 2 && fsum) {
    const double v = *fsum;
    if (v > 1e-12 && v != 1.0) fscale = 1.0 / sqrt(v);
  }
  // stage the A operands and lane tables of every block
  for (int idx = tid; idx < nb * 64 * 16; idx += NTH) {
    const int b = idx / (64 * 16), ln = (idx / 16) % 64, s = idx % 16;
    const TileBlock &bl = blocks[ps.first_block + b];
    const int R = 1 << bl.k, RR = 2 * R, row = ln & 31, hh = ln >> 5;
    const float *W = wbuf + size_t(bl.wslot) * kTileWFloats;
    float w = (row < RR && s < R) ? W[row * RR + 2 * s + hh] : 0.0f;
    if (!SPLIT && KF == 2) {
      // three-product form: a[p][t] = A[i][qq] of product p (Ur + Ui,
      // -s (Ur - Ui) with s = +1 for the pass's first block and -1 after,
      // -Ui), instruction t; D row i = 4 qi + vi means output row qi | vi << 2
      // and K-index qq input row qq | t << 2.  a[3][0] and a[3][1]: the lane's
      // address in column half 0 and 1, lane table entry (ln & 15) | 16 h ^
      // the row offsets of target bits 0 and 1 on lane bits 4 and 5.
      const int i = ln & 15, qq = ln >> 4, p = s >> 2, t = s & 3;
      const int ro = sigma4(bl, (i >> 2) | ((i & 3) << 2)), si = sigma4(bl, qq | (t << 2));
      if (p < 3) {
        const bool in = ro < R && si < R;
        w = !in ? 0.0f
            : p == 0 ? W[RR * RR + ro * R + si]
            : p == 1 ? (b == 0 ? -1.0f : 1.0f) * W[RR * RR + R * R + ro * R + si]
                     : -W[(2 * ro + 1) * RR + 2 * si];
        if (b == 0) w = float(double(w) * fscale);
      } else {
        const int lrows = ((ln & 16) ? bl.rowoff[1] : 0) ^ ((ln & 32) ? bl.rowoff[2] : 0);
        w = t < 2 ? __int_as_float(bl.lanecol[(ln & 15) | (t << 4)] ^ lrows) : 0.0f;
      }
    } else if (!SPLIT && (bl.k <= 3 || KF == 1)) {
      // the 16-column form (tp_block16, tp_block4p): instruction sp = s & 7
      // of row block rb = s >> 3, A[i][qq] with i = 4 qi + vi holding output
      // row (qi & 1) | (vi >> 1) << 1 | (qi >> 1) << 2 | rb << 3, part vi & 1,
      // and qq meaning input row 2 sp + (qq & 1), part qq >> 1
      const int i = ln & 15, qq = ln >> 4, qi = i >> 2, vi = i & 3, rb = s >> 3, sp = s & 7;
      const int ro = sigma4(bl, (qi & 1) | ((vi >> 1) << 1) | ((qi >> 1) << 2) | (rb << 3));
      const int si = sigma4(bl, 2 * sp + (qq & 1));
      w = (ro < R && si < R) ? W[(2 * ro + (vi & 1)) * RR + 2 * si + (qq >> 1)] : 0.0f;
    }
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
  for (int idx = tid; idx < nb * 32; idx += NTH) {
    const int b = idx / 32, e = idx % 32;
    lds_store<int>(OPS + b * STRIDE + ABYTES + 4 * e, blocks[ps.first_block + b].lanecol[e]);
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
  uint64_t ib[IB];  // index bit of each bit of `it`
#pragma unroll
  for (int jj = 0; jj < IB; ++jj) ib[jj] = uint64_t(1) << ps.sbits[TBITS + jj];
  auto it_addr = [&](int it) {
    uint64_t ai = 0;
#pragma unroll
    for (int jj = 0; jj < IB; ++jj)
      if ((it >> jj) & 1) ai |= ib[jj];
    return ai;
  };

  const uint64_t G8 = gridDim.x >> 3, T8 = ntiles >> 3;
  const uint64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  uint64_t tile = xcd * T8 + slot;
  const uint64_t tend = (xcd + 1) * T8;
  // tile t's base address deposits t on the index bits outside the tile
  // (cmask); the next tile's, t + G8, is found by adding the deposited G8 in
  // the deposited space: ((x | ~cmask) + dinc) & cmask
  auto deposit = [&](uint64_t t) {
    uint64_t b = t;
#pragma unroll
    for (int j = 0; j < TB; ++j) b = insert0(b, ps.sbits[j]);
    return b;
  };
  uint64_t tmask = 0;
#pragma unroll
  for (int j = 0; j < TB; ++j) tmask |= uint64_t(1) << ps.sbits[j];
  const uint64_t cmask = ((ntiles << TB) - 1) & ~tmask;
  const uint64_t dinc = deposit(G8);

  if constexpr (DIR) {
    static_assert(!SPLIT && NTH == 1024, "direct passes: the three-product form");
    // lane (q, c) of wave w holds, in column half h, rows q | t << 2 (t =
    // 0..3) of one column of the first and of the last block: tile element
    // = the lane part (c on tile positions 1..4, h on position 0) ^ the rows
    // of target bits 0, 1 (q) ^ the wave part ^ the rows of target bits 2, 3
    // (t); as index bits, the element's bits deposited on the pass's tile
    // bits
    auto dep = [&](int e) {
      uint64_t x = 0;
#pragma unroll
      for (int j = 0; j < TB; ++j)
        if ((e >> j) & 1) x |= uint64_t(1) << ps.sbits[j];
      return x;
    };
    const TileBlock &B0 = blocks[ps.first_block], &BL = blocks[ps.first_block + nb - 1];
    const int lane = tid & 63, c16 = lane & 15, q = lane >> 4;
    const int fq = ((q & 1) ? B0.erow[0] : 0) ^ ((q & 2) ? B0.erow[1] : 0) ^ B0.ewave[wave];
    const int lq = ((q & 1) ? BL.erow[0] : 0) ^ ((q & 2) ? BL.erow[1] : 0) ^ BL.ewave[wave];
    // (half 1 is half 0 + 1: the half bit is tile position 0, index bit 0)
    const uint64_t fo = dep(B0.elane[c16] ^ fq), lo = dep(BL.elane[c16] ^ lq);
    const uint64_t fr2 = dep(B0.erow[2]), fr3 = dep(B0.erow[3]);
    const uint64_t lr2 = dep(BL.erow[2]), lr3 = dep(BL.erow[3]);
    // the first block's operands of the tile at b: row t -> (re, im) of
    // half 0, (re, im) of half 1
    auto load_ops = [&](float4 (&x)[4], uint64_t b) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
        x[t] = ld_nt(a + (b | fo | ((t & 1) ? fr2 : 0) | ((t & 2) ? fr3 : 0)));
    };
    auto step_base = [&](uint64_t b) { return ((b | ~cmask) + dinc) & cmask; };
    uint64_t base = deposit(tile);
    uint64_t nbase = step_base(base), nbase2 = step_base(nbase);
    // two tiles in flight, in two register sets taken in turn; past the
    // end a load reads the current tile again, unused
    float4 xa[4], xb[4];
    if (tile < tend) {
      load_ops(xa, base);
      load_ops(xb, tile + G8 < tend ? nbase : base);
    }
    __syncthreads();  // operands staged
    BlockScalars sa = load_scalars(&blocks[ps.first_block], wave), sb;
    BlockRegs ra;
    load_regs4g(ra, OPS);
    const TileBlock *fb = &blocks[ps.first_block];
    double nacc = 0.0;  // the pass's |x|^2 when `norm` is given
    auto step = [&](float4 (&cur)[4]) {
      K4GState st;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        st.b0[t] = cur[t].x;
        st.b0[4 + t] = cur[t].y;
        st.b1[t] = cur[t].z;
        st.b1[4 + t] = cur[t].w;
      }
      load_ops(cur, tile + 2 * G8 < tend ? nbase2 : base);
      for (int b = 0; b + 1 < nb; ++b) {
        tp_block4g<NTH, TILE>(sa, ra, b == 0 ? 1.0f : -1.0f, fb + b + 1,
                              OPS + (b + 1) * kTileStrideFp32, wave, sb, ra, st);
        sa = sb;
      }
      tp_block4g_out<NTH, TILE>(sa, ra, nb == 1 ? 1.0f : -1.0f, fb, OPS, wave, sb, ra, st, a,
                                base, lo, lr2, lr3, norm != nullptr, nacc);
      sa = sb;
      // the next tile's first block writes the LDS tile its last block read
      __syncthreads();
      tile += G8;
      base = nbase;
      nbase = nbase2;
      nbase2 = step_base(nbase2);
    };
    while (tile < tend) {
      step(xa);
      if (tile >= tend) break;
      step(xb);
    }
    if (norm) wg_norm_out<NTH>(nacc, norm);
    return;
  }

  // The next tile is loaded into registers while the current one is
  // computed.  At a tile boundary the finished tile is read out of LDS, the
  // prefetched one written in, and only then are the stores issued: stores
  // count against the same in-order vector-memory counter as loads, so a
  // wait for the prefetch placed after the stores would also wait for the
  // stores to reach memory.
  // DEPTH tiles are in flight: the three-product form loads tile t + 2
  // while it computes tile t (two register sets taken in turn), so a pass
  // whose tiles are scattered over many pages still streams at the HBM rate.
  constexpr int DEPTH = (!SPLIT && KF == 2) ? 2 : 1;
  float4 v0[IT], v1[IT];
  uint64_t base = deposit(tile);
  auto step_base = [&](uint64_t b) { return ((b | ~cmask) + dinc) & cmask; };
  uint64_t nbase = step_base(base), nbase2 = step_base(nbase);
  if (tile < tend) {
#pragma unroll
    for (int it = 0; it < IT; ++it) v0[it] = ld_nt(a + (base | it_addr(it) | addr_t));
  }
  __syncthreads();  // operands staged
  BlockScalars sa, sb;
  BlockRegs ra, rb;
  if (!SPLIT) {
    sa = load_scalars(&blocks[ps.first_block], wave);
    if (KF == 2)
      load_regs4g(ra, OPS);
    else if (KF == 1)
      load_regs4p(ra, OPS);
    else
      load_regs(ra, OPS);
  }
  if (tile < tend) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int s0 = slot_t ^ swz(it << TBITS);  // slot of e; e+1 is s0 ^ 1
      re[s0] = v0[it].x;
      im[s0] = v0[it].y;
      re[s0 ^ 1] = v0[it].z;
      im[s0 ^ 1] = v0[it].w;
    }
  }
  __syncthreads();
  if (DEPTH == 2 && tile < tend) {
    const uint64_t lb = tile + G8 < tend ? nbase : base;
#pragma unroll
    for (int it = 0; it < IT; ++it) v0[it] = ld_nt(a + (lb | it_addr(it) | addr_t));
  }

  double nacc = 0.0;  // the pass's |x|^2 when `norm` is given
  // one tile: load tile + DEPTH into `ld`, apply the blocks, write the tile
  // back and move tile + 1 from `fl` into LDS
  auto tile_step = [&](float4 (&ld)[IT], float4 (&fl)[IT]) {
    const uint64_t next = tile + G8;
    {
      // issued on every path (past the last tile it reads this tile again,
      // unused): the wait for `fl` before the fill then counts the same
      // outstanding loads on every path, and does not wait for `ld`
      const uint64_t lb = tile + uint64_t(DEPTH) * G8 < tend ? (DEPTH == 2 ? nbase2 : nbase)
                                                             : base;
#pragma unroll
      for (int it = 0; it < IT; ++it) ld[it] = ld_nt(a + (lb | it_addr(it) | addr_t));
    }

    if (SPLIT) {
      BlockScalars cur = load_scalars(&blocks[ps.first_block], wave);
      // the first block's scalars are waited for here, once per tile: the
      // compiler otherwise waits at the top of every block, for the next
      // block's loads as well
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0) only
      for (int b = 0; b < nb; ++b) {
        const TileBlock *bl = &blocks[ps.first_block + b];
        const BlockScalars nxt = load_scalars(bl + min(1, nb - 1 - b), wave);
        const int blk = OPS + b * STRIDE;
        switch (cur.k) {
          case 1: tp_block_split<1, NTH, TILE>(bl, cur, blk); break;
          case 2: tp_block_split<2, NTH, TILE>(bl, cur, blk); break;
          case 3: tp_block_split<3, NTH, TILE>(bl, cur, blk); break;
          default: tp_block_split<4, NTH, TILE>(bl, cur, blk); break;
        }
        if (cur.sync) __syncthreads();
        cur = nxt;
      }
    } else {
      // two register sets taken in turn, so the values read one block ahead
      // need no copies; (sa, ra) hold the pass's first block on entry
      const TileBlock *fb = &blocks[ps.first_block];
      if constexpr (KF != 0) {
        // every block is a k = 4 block (the host pads smaller ones); one set
        // of per-lane values, read in place once a block's last chain is
        // issued; the scalars alternate between two sets
        if constexpr (KF == 2) {
          K4GState st;
          read_k4g_both(sa, ra, st);
          for (int b = 0; b < nb; ++b) {
            const int bn = b + 1 < nb ? b + 1 : 0;
            tp_block4g<NTH, TILE>(sa, ra, b == 0 ? 1.0f : -1.0f, fb + bn,
                                  OPS + bn * kTileStrideFp32, wave, sb, ra, st);
            sa = sb;
          }
        } else {
          K4State st;
          st.pend = 0;
          read_k4_b0(sa, ra, st);
          for (int b = 0; b < nb; ++b) {
            const int bn = b + 1 < nb ? b + 1 : 0;
            tp_block4p<NTH, TILE>(sa, ra, fb + bn, OPS + bn * kTileStrideFp32, wave, sb, ra,
                                  st);
            sa = sb;
          }
        }
      } else {
        for (int b = 0;;) {
          tp_step<NTH, TILE>(fb, b, nb, OPS, wave, sa, ra, sb, rb);
          if (++b == nb) {
            sa = sb;
            ra = rb;
            break;
          }
          tp_step<NTH, TILE>(fb, b, nb, OPS, wave, sb, rb, sa, ra);
          if (++b == nb) break;
        }
      }
    }

    float4 o[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int s0 = slot_t ^ swz(it << TBITS);
      if constexpr (KF == 2)  // the three-product form leaves -Im in the tile
        o[it] = make_float4(re[s0], -im[s0], re[s0 ^ 1], -im[s0 ^ 1]);
      else
        o[it] = make_float4(re[s0], im[s0], re[s0 ^ 1], im[s0 ^ 1]);
    }
    if (norm) {
      // |x|^2 of the values written, in FP64, in a fixed order
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const double x0 = o[it].x, y0 = o[it].y, x1 = o[it].z, y1 = o[it].w;
        nacc = nacc + (x0 * x0 + y0 * y0);
        nacc = nacc + (x1 * x1 + y1 * y1);
      }
    }
    // each thread refills the slots it has just read out: no other thread
    // touches them in between, so no barrier is needed here
    if (next < tend) {
#pragma unroll
      for (int it = 0; it < IT; ++it) {
        const int s0 = slot_t ^ swz(it << TBITS);
        re[s0] = fl[it].x;
        im[s0] = fl[it].y;
        re[s0 ^ 1] = fl[it].z;
        im[s0 ^ 1] = fl[it].w;
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the stores after the LDS writes
#pragma unroll
    for (int it = 0; it < IT; ++it) st_nt(a + (base | it_addr(it) | addr_t), o[it]);
    __syncthreads();  // the next tile is in LDS
    tile = next;
    base = nbase;
    nbase = nbase2;
    nbase2 = step_base(nbase2);
  };
  if constexpr (DEPTH == 2) {
    // v0 holds tile + 1 on entry; the sets alternate
    while (tile < tend) {
      tile_step(v1, v0);
      if (tile >= tend) break;
      tile_step(v0, v1);
    }
  } else {
    while (tile < tend) tile_step(v0, v0);
  }
  if (norm) wg_norm_out<NTH>(nacc, norm);
}

// ---- host side ------------------------------------------------------------------

// a device buffer of at least `need` bytes; the old one is freed only once
// the state's stream has drained (a previous call's passes may still read it)
static int grow_dev(qcs_hip_state *s, void **p, size_t *cap, size_t need) {
  if (*cap >= need) return 0;
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  QCS_HIP_TRY(hipMalloc(p, need));
  *cap = need;
  return 0;
}

template <int NTH, bool SPLIT, int K4>
static int launch_tp(qcs_hip_state *s, const TilePass *dpass, int pi,
                     const TileBlock *dblocks, uint64_t ntiles, int nblocks,
                     double *norm = nullptr, int *grid_out = nullptr,
                     const double *fsum = nullptr) {
  // once per kernel instance (states may launch from separate threads)
  static std::once_flag once;
  static int attr_rc = 0;
  std::call_once(once, [] {
    const void *fn =
        reinterpret_cast<const void *>(&k_tile_program<kTileBitsFixed, NTH, SPLIT, K4>);
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
  // the pass's norm: one partial per work-group in the reduction buffer
  if (norm && grid > uint64_t(kRedBlocks)) norm = nullptr;
  if (grid_out) *grid_out = norm ? int(grid) : 0;
  hipLaunchKernelGGL((k_tile_program<kTileBitsFixed, NTH, SPLIT, K4>), dim3(unsigned(grid)),
                     dim3(NTH), lds, s->stream, static_cast<float2 *>(s->vec),
                     dpass, pi, dblocks, s->wbuf, ntiles, norm, fsum);
  return launch_check("k_tile_program");
}

int tile_max_blocks(const qcs_hip_state *s) {
  return s->tile_math == QCS_HIP_TILE_MATH_SPLIT_BF16 ? QCS_HIP_TILE_MAX_BLOCKS_SPLIT
                                                      : QCS_HIP_TILE_MAX_BLOCKS;
}

// Low slot bit (LDS bank bit) that tile position p lands on under swz.
static int bank_bit(int p) { return p < 5 ? p : p < 10 ? p - 5 : p - 10; }

// Lane bits from `avail`: one position per bank bit where there is one
// (lowest first), the rest the lowest positions left.  Returns the bank
// conflict factor of the 32 columns a half-wave reads: 2^(5 - rank of their
// bank bits), 1 when they fall in 32 distinct banks.
static int pick_lanes(int avail, int lanes[5]) {
  int n = 0, used = 0;
  for (int c = 0; c < 5; ++c)
    for (int p = 0; p < kTileBitsFixed; ++p)
      if (((avail >> p) & 1) && bank_bit(p) == c) {
        lanes[n++] = p;
        used |= 1 << p;
        break;
      }
  const int distinct = n;
  for (int p = 0; p < kTileBitsFixed && n < 5; ++p)
    if (((avail >> p) & 1) && !((used >> p) & 1)) lanes[n++] = p;
  return 1 << (5 - distinct);
}

// Lane bits of a k <= 3 block in the 16-column form (tp_block16): lane bits
// 0..3 on distinct bank bits other than tb0, the bank bit of the block's
// first target (a half-wave reads 16 columns at two rows that differ in that
// target), lane bit 4 (the half a column falls in) on tb0 where it can, the
// rest the lowest positions left.  Returns the half-wave's conflict factor,
// 2^(5 - rank of the bank bits of lane bits 0..3 and tb0).
static int pick_lanes16(int avail, int tb0, int lanes[5], int half = -1) {
  if (half >= 0 && ((avail >> half) & 1)) {
    // lane bit 4 on `half`, lane bits 0..3 from the other positions
    const int f = pick_lanes16(avail & ~(1 << half), tb0, lanes);
    lanes[4] = half;
    return f;
  }
  int n = 0, used = 0;
  for (int c = 0; c < 5 && n < 4; ++c) {
    if (c == tb0) continue;
    for (int p = 0; p < kTileBitsFixed; ++p)
      if (((avail >> p) & 1) && bank_bit(p) == c) {
        lanes[n++] = p;
        used |= 1 << p;
        break;
      }
  }
  const int distinct = n + 1;
  for (int p = 0; p < kTileBitsFixed && n < 4; ++p)
    if (((avail >> p) & 1) && !((used >> p) & 1)) {
      lanes[n++] = p;
      used |= 1 << p;
    }
  lanes[4] = -1;
  for (int p = 0; p < kTileBitsFixed && lanes[4] < 0; ++p)
    if (((avail >> p) & 1) && !((used >> p) & 1) && bank_bit(p) == tb0) lanes[4] = p;
  for (int p = 0; p < kTileBitsFixed && lanes[4] < 0; ++p)
    if (((avail >> p) & 1) && !((used >> p) & 1)) lanes[4] = p;
  return 1 << (5 - distinct);
}

static int target_mask(const qcs_hip_pass &ps, const qcs_hip_block &blk) {
  int m = 0;
  for (int q = 0; q < blk.k; ++q) {
    int pos = 0;
    while (ps.bits[pos] != blk.targets[q]) ++pos;
    m |= 1 << pos;
  }
  return m;
}

// value whose bit i sits at position pos[i], as a tile element index
static int scatter_elem(int v, const int *pos, int npos) {
  int e = 0;
  for (int i = 0; i < npos; ++i)
    if ((v >> i) & 1) e |= 1 << pos[i];
  return e;
}

// value whose bit i sits at position pos[i], through swz, as a byte offset
static int scatter_off(int v, const int *pos, int npos) {
  int e = 0;
  for (int i = 0; i < npos; ++i)
    if ((v >> i) & 1) e |= 1 << pos[i];
  return 4 * swz(e);
}

// The block's descriptor for the given wave bits (a mask of `nwb` tile
// positions, none of them a target): row offsets, lane, wave and group
// tables.  Lane bits are conflict-free where the remaining positions allow,
// else the lowest ones; form16: the block runs in the 16-column form.
static void make_tile_block(const qcs_hip_pass &ps, const qcs_hip_block &blk, int wslot,
                            int wmask, int nwb, bool sync, bool form16, int half,
                            TileBlock *tb, bool direct = false, bool choose_t0 = false) {
  std::memset(tb, 0, sizeof(TileBlock));
  const int k = blk.k, R = 1 << k;
  tb->k = k;
  tb->wslot = wslot;
  tb->sync = sync ? 1 : 0;
  int tpos[kTileKMax], wpos[kTileBitsFixed], lanes[5], jpos[kTileBitsFixed];
  const int tmask = target_mask(ps, blk);
  for (int q = 0, p = 0; p < kTileBitsFixed; ++p)
    if ((tmask >> p) & 1) tpos[q++] = p;
  int nw = 0;
  for (int p = 0; p < kTileBitsFixed; ++p)
    if ((wmask >> p) & 1) wpos[nw++] = p;
  int avail = ((1 << kTileBitsFixed) - 1) & ~tmask & ~wmask;
  // the target on row bit 0 (lane bit 4's rows): the one with the fewest
  // bank conflicts of a half-wave's accesses; the others keep their order
  int tidx[kTileKMax];
  for (int q = 0; q < k; ++q) tidx[q] = q;
  if (choose_t0) {
    int best = 0, best_f = 1 << 30;
    for (int q = 0; q < k; ++q) {
      int f, tmp[5];
      if (direct) {
        f = bank_bit(tpos[q]) == 0 ? 1 : 2;  // lane bits on banks 1..4
      } else {
        f = pick_lanes16(avail, bank_bit(tpos[q]), tmp, half);
      }
      if (f < best_f) {
        best_f = f;
        best = q;
      }
    }
    for (int q = 0, j = 1; q < k; ++q)
      if (q != best) tidx[j++] = q;
    tidx[0] = best;
    int t2[kTileKMax];
    for (int q = 0; q < k; ++q) t2[q] = tpos[tidx[q]];
    for (int q = 0; q < k; ++q) tpos[q] = t2[q];
  }
  for (int q = 0; q < kTileKMax; ++q) tb->tidx[q] = q < k ? tidx[q] : q;
  if (direct) {
    // the half bit on tile position 0 and lane bits 0..3 on positions 1..4
    // (the lowest index bits): a lane's two halves of a row are one 16-byte
    // access, and 16 lanes cover 256 contiguous bytes of HBM (8 lanes, one
    // 128-byte line, measured 2 % slower)
    for (int i = 0; i < 4; ++i) lanes[i] = i + 1;
    lanes[4] = 0;
  } else if (form16) {
    (void)pick_lanes16(avail, bank_bit(tpos[0]), lanes, half);
  } else {
    (void)pick_lanes(avail, lanes);
  }
  for (int i = 0; i < 5; ++i) avail &= ~(1 << lanes[i]);
  int nj = 0;
  for (int p = 0; p < kTileBitsFixed; ++p)
    if ((avail >> p) & 1) jpos[nj++] = p;
  for (int r = 0; r < R; ++r) tb->rowoff[r] = scatter_off(r, tpos, k);
  for (int l = 0; l < 32; ++l) tb->lanecol[l] = scatter_off(l, lanes, 5);
  for (int w = 0; w < (1 << nwb); ++w) tb->wpart[w] = scatter_off(w, wpos, nwb);
  for (int l = 0; l < 32; ++l) tb->elane[l] = scatter_elem(l, lanes, 5);
  for (int w = 0; w < (1 << nwb); ++w) tb->ewave[w] = scatter_elem(w, wpos, nwb);
  for (int q = 0; q < k; ++q) tb->erow[q] = 1 << tpos[q];
  for (int j = 0; j < (1 << nj); ++j) tb->jpart[j] = scatter_off(j, jpos, nj);
}

// Descriptors of one pass.  Blocks are taken in runs that share their wave
// bits, and only the last block of a run is followed by a barrier.  The
// wave bits of each block (and the half bit of the pipelined k = 4 forms)
// trade barriers against LDS bank conflicts, by dynamic programming over the
// pass.  With FP32 math, blocks of k <= 3 run in the 16-column form, whose
// conflicts also depend on the first target.
static bool layout_pass(const qcs_hip_pass &ps, const qcs_hip_block *blocks, int nwb,
                        bool split, bool k4p, TileBlock *tb, bool direct = false) {
  // conflict factor of every set of free positions (and, for the 16-column
  // form, bank bit of the first target), and the candidate wave bits of each
  // width, computed once per process
  constexpr int kSets = 1 << kTileBitsFixed;
  static std::once_flag once;
  static unsigned char factor[kSets], factor16[5][kSets];
  static std::vector<int> masks[5];
  std::call_once(once, [] {
    int lanes[5];
    for (int a = 0; a < kSets; ++a) {
      // 5 lanes; the 16-column form needs 4 (and a half bit, which the
      // pipelined forms take out of the set first)
      const int np = __builtin_popcount(unsigned(a));
      factor[a] = np >= 5 ? pick_lanes(a, lanes) : 0;
      for (int c = 0; c < 5; ++c) factor16[c][a] = np >= 4 ? pick_lanes16(a, c, lanes) : 0;
    }
    for (int m = 0; m < kSets; ++m)
      if (__builtin_popcount(unsigned(m)) <= 4) masks[__builtin_popcount(unsigned(m))].push_back(m);
  });
  auto form16 = [&](int i) { return !split && (blocks[i].k <= 3 || k4p); };
  const int nblk = ps.num_blocks;
  std::vector<int> tm(static_cast<size_t>(nblk));
  for (int i = 0; i < nblk; ++i) tm[i] = target_mask(ps, blocks[ps.first_block + i]);
  // Wave bits and half bit of every block by dynamic programming over the
  // pass's blocks.  State (j, h): the block runs on wave bits cand[j] and,
  // in the pipelined k = 4 forms, half bit h (lane bit 4; one state h = 0
  // otherwise).  A block continues the previous block's run when both have
  // the same wave bits (no barrier between them), and reads ahead when it
  // also has the same half bit.  Costs: kRunCost per barrier, kNoAheadCost
  // per block of a run that cannot read ahead, kConflictCost per extra way
  // of a block's bank conflicts (a 4-way conflicted block makes LDS the
  // bottleneck of the CU; a 2-way one mostly hides under the matrix cores).
  const std::vector<int> &cand = masks[nwb];
  const int nm = int(cand.size()), nh = k4p ? kTileBitsFixed : 1;
  // (weights measured on the 30-qubit layered circuit: these beat sharing
  // wave and half bits over long runs by ~3%, and one barrier per block)
  constexpr int kRunCost = 6, kNoAheadCost = 1, kConflictCost = 4;
  constexpr int kInf = 1 << 29;
  // conflict cost of block i on wave bits mk with half bit h (-1 if h is
  // not free)
  auto conflict = [&](int i, int mk, int h) {
    const int t = tm[i];
    int a = (kSets - 1) & ~t & ~mk;
    int f;
    if (direct && (i == 0 || i + 1 == nblk)) {
      // the direct lanes of make_tile_block (banks 1..4): conflict-free when
      // a target on bank 0 can take row bit 0
      f = 2;
      for (int p = 0; p < kTileBitsFixed; ++p)
        if (((t >> p) & 1) && bank_bit(p) == 0) f = 1;
      return kConflictCost * (f - 1);
    }
    if (form16(ps.first_block + i)) {
      if (k4p) {
        if (!((a >> h) & 1)) return -1;
        a &= ~(1 << h);
      }
      if (k4p) {
        // make_tile_block picks the first target
        f = 1 << 30;
        for (int p = 0; p < kTileBitsFixed; ++p)
          if ((t >> p) & 1) f = std::min(f, int(factor16[bank_bit(p)][a]));
      } else {
        f = factor16[bank_bit(__builtin_ctz(unsigned(t)))][a];
      }
    } else {
      f = factor[a];
    }
    return kConflictCost * (f - 1);
  };
  const size_t ns = size_t(nm) * size_t(nh);
  // from[i][s]: -1 when block i continues state s of block i - 1, else the
  // state of block i - 1 it follows (a new half bit or a new run)
  std::vector<int> cost(ns * size_t(nblk), kInf), from(ns * size_t(nblk), -1);
  std::vector<int> best_of_mask(static_cast<size_t>(nm));
  std::vector<int> arg_of_mask(static_cast<size_t>(nm));
  int prev_best = 0, prev_arg = -1;
  for (int i = 0; i < nblk; ++i) {
    int best = kInf, arg = -1;
    int *c = cost.data() + ns * size_t(i);
    int *fr = from.data() + ns * size_t(i);
    const int *pc = i > 0 ? cost.data() + ns * size_t(i - 1) : nullptr;
    if (pc)
      for (int j = 0; j < nm; ++j) {
        best_of_mask[j] = kInf;
        for (int h = 0; h < nh; ++h)
          if (pc[j * nh + h] < best_of_mask[j]) {
            best_of_mask[j] = pc[j * nh + h];
            arg_of_mask[j] = j * nh + h;
          }
      }
    // a direct pass's first and last blocks keep tile positions 0..4 for
    // the half bit (position 0) and the lanes: no wave bit there
    const bool edge = direct && (i == 0 || i + 1 == nblk);
    for (int j = 0; j < nm; ++j) {
      if (cand[j] & tm[i]) continue;
      if (edge && (cand[j] & 31)) continue;
      for (int h = 0; h < nh; ++h) {
        if (edge && h != 0) continue;
        const int cf = conflict(i, cand[j], h);
        if (cf < 0) continue;
        const int sidx = j * nh + h;
        int v = prev_best + kRunCost, f = prev_arg;  // a new run
        if (pc && best_of_mask[j] + kNoAheadCost < v) {
          v = best_of_mask[j] + kNoAheadCost;  // same run, no read-ahead
          f = arg_of_mask[j];
        }
        if (pc && pc[sidx] <= v) {
          v = pc[sidx];  // same run, reads ahead
          f = -1;
        }
        c[sidx] = v + cf;
        fr[sidx] = f;
        if (c[sidx] < best) {
          best = c[sidx];
          arg = sidx;
        }
      }
    }
    if (arg < 0) return false;  // no layout under the direct constraints
    prev_best = best;
    prev_arg = arg;
  }
  std::vector<int> wmask(static_cast<size_t>(nblk)), whalf(static_cast<size_t>(nblk));
  for (int i = nblk - 1, sidx = prev_arg; i >= 0; --i) {
    wmask[i] = cand[sidx / nh];
    whalf[i] = k4p ? sidx % nh : -1;
    const int f = from[ns * size_t(i) + size_t(sidx)];
    if (f >= 0) sidx = f;
  }
  for (int i = 0; i < nblk; ++i) {
    const int bi = ps.first_block + i;
    const bool last = i + 1 == nblk || wmask[i + 1] != wmask[i];
    make_tile_block(ps, blocks[bi], bi, wmask[i], nwb, last, form16(bi), whalf[i], &tb[bi],
                    direct && (i == 0 || i + 1 == nblk), k4p);
  }
  // a k = 4 block reads the next block's first half ahead when that block is
  // also k = 4, in this run, with the same half bit
  for (int i = 0; i + 1 < nblk; ++i) {
    const int bi = ps.first_block + i;
    tb[bi].early = (k4p && blocks[bi].k == 4 && blocks[bi + 1].k == 4 && !tb[bi].sync &&
                    tb[bi].lanecol[16] == tb[bi + 1].lanecol[16])
                       ? 1
                       : 0;
  }
  return true;
}

// Blocks of fewer than 4 qubits widened to 4 (their unitary composes to
// U x I on the added targets), for the kernels that run every block as a
// k = 4 block.  The added targets are taken from the neighbouring blocks'
// targets where possible, then from the pass bits, highest first, so no run
// of blocks needs more tile bits than before.
static void pad_blocks(const qcs_hip_pass *passes, int np, const qcs_hip_block *blocks,
                       int nb, std::vector<qcs_hip_block> *out) {
  out->assign(blocks, blocks + nb);
  for (int p = 0; p < np; ++p) {
    const int b0 = passes[p].first_block, b1 = b0 + passes[p].num_blocks;
    for (int b = b0; b < b1; ++b) {
      qcs_hip_block &bl = (*out)[size_t(b)];
      if (bl.k >= QCS_HIP_TILE_KMAX) continue;
      int cand[3 * QCS_HIP_TILE_KMAX + QCS_HIP_TILE_BITS], nc = 0;
      for (int nbh : {b - 1, b + 1})
        if (nbh >= b0 && nbh < b1)
          for (int q = 0; q < blocks[nbh].k; ++q) cand[nc++] = blocks[nbh].targets[q];
      for (int j = passes[p].tile_bits - 1; j >= 0; --j) cand[nc++] = passes[p].bits[j];
      for (int c = 0; c < nc && bl.k < QCS_HIP_TILE_KMAX; ++c) {
        bool have = false;
        for (int q = 0; q < bl.k; ++q) have = have || bl.targets[q] == cand[c];
        if (!have) bl.targets[bl.k++] = cand[c];
      }
      std::sort(bl.targets, bl.targets + bl.k);
    }
  }
}

int tiles_apply(qcs_hip_state *s, const qcs_hip_pass *passes, int np,
                const qcs_hip_block *blocks, int nb, const qcs_hip_gate *gates,
                int ng, int *norm_partials) {
  if (norm_partials) *norm_partials = 0;
  if (np <= 0) return 0;
  const size_t pbytes = sizeof(TilePass) * size_t(np);
  const size_t tbytes = sizeof(TileBlock) * size_t(nb > 0 ? nb : 1);
  int rc;
  if ((rc = grow_dev(s, &s->tpass_dev, &s->tpass_cap, pbytes))) return rc;
  if ((rc = grow_dev(s, &s->tblock_dev, &s->tblock_cap, tbytes))) return rc;
  const int nwb = s->tile_threads == 1024 ? 4 : 3;  // wave-index bits
  const bool split_math = s->tile_math == QCS_HIP_TILE_MATH_SPLIT_BF16;
  // k = 4 blocks as pipelined 16-column halves: FP32 math, 1024 threads;
  // that kernel runs every block as a k = 4 block, so smaller blocks get
  // more targets (their unitary composes to U x I on them), taken from the
  // neighbouring blocks' targets where possible so no run needs more bits
  const bool k4p = !split_math && s->tile_threads == 1024 &&
                   (s->tile_k4 == QCS_HIP_TILE_K4_HALVES || s->tile_k4 == QCS_HIP_TILE_K4_GAUSS);
  // the three-product kernel's direct form (first block's operands from HBM,
  // last block's results to HBM)
  const bool direct_ok = k4p && s->tile_threads == 1024 &&
                         s->tile_k4 == QCS_HIP_TILE_K4_GAUSS && s->tile_direct != 0;
  std::vector<qcs_hip_block> padded;
  const qcs_hip_block *lay = blocks;
  if (k4p) {
    pad_blocks(passes, np, blocks, nb, &padded);
    lay = padded.data();
  }
  char *stage = nullptr;
  if ((rc = compose_blocks(s, lay, nb, gates, ng, pbytes + tbytes, &stage)))
    return rc;
  TilePass *tp = reinterpret_cast<TilePass *>(stage);
  TileBlock *tb = reinterpret_cast<TileBlock *>(stage + pbytes);
  for (int p = 0; p < np; ++p) {
    std::memset(&tp[p], 0, sizeof(TilePass));
    tp[p].tb = passes[p].tile_bits;
    tp[p].nblocks = passes[p].num_blocks;
    tp[p].first_block = passes[p].first_block;
    for (int j = 0; j < passes[p].tile_bits; ++j) tp[p].sbits[j] = passes[p].bits[j];
    // direct when the first and last blocks leave tile positions 0..4 (the
    // lowest index bits, in every pass) to the half bit and the lanes
    const int nbp = passes[p].num_blocks, fb = passes[p].first_block;
    const bool dir = direct_ok && nbp > 0 && !(target_mask(passes[p], lay[fb]) & 31) &&
                     !(target_mask(passes[p], lay[fb + nbp - 1]) & 31);
    if (dir && layout_pass(passes[p], lay, nwb, split_math, k4p, tb, true))
      tp[p].direct = 1;
    else
      layout_pass(passes[p], lay, nwb, split_math, k4p, tb);
  }
  QCS_HIP_TRY(hipMemcpyAsync(s->tpass_dev, tp, pbytes, hipMemcpyHostToDevice,
                             s->stream));
  QCS_HIP_TRY(hipMemcpyAsync(s->tblock_dev, tb, tbytes, hipMemcpyHostToDevice,
                             s->stream));
  const TilePass *dp = static_cast<const TilePass *>(s->tpass_dev);
  const TileBlock *db = static_cast<const TileBlock *>(s->tblock_dev);
  const bool split = s->tile_math == QCS_HIP_TILE_MATH_SPLIT_BF16;
  const double N = double(s->size);
  int last_pass = -1, prev_pass = -1;
  for (int p = 0; p < np; ++p)
    if (passes[p].num_blocks > 0) {
      prev_pass = last_pass;
      last_pass = p;
    }
  // Renormalisation (when the caller asks for the norm).  In the
  // three-product form with two passes or more, the pass before the last
  // sums |x|^2 of what it writes (one FP64 partial per work-group), the
  // partials are folded on the device, and the last pass scales its first
  // block by 1/sqrt(norm_sq): no pass of its own reads the state to rescale
  // it.  Otherwise the last pass sums the norm of its output for the caller.
  const bool gauss = !split && k4p && s->tile_k4 == QCS_HIP_TILE_K4_GAUSS &&
                     s->tile_threads == 1024;
  const bool scale_ahead = norm_partials && gauss && prev_pass >= 0;
  const double *ns_dev = nullptr;  // the folded norm_sq, once available
  for (int p = 0; p < np; ++p) {
    // the kernel runs its pass's first block unconditionally: an empty pass
    // (which the planner does not make) would change nothing anyway
    if (passes[p].num_blocks <= 0) continue;
    const uint64_t ntiles = s->size >> passes[p].tile_bits;
    double flops = 0.0;
    for (int b = passes[p].first_block;
         b < passes[p].first_block + passes[p].num_blocks; ++b)
      flops += 8.0 * double(1 << blocks[b].k) * N;
    TimedLaunch tl;
    timing_begin(s, QCS_HIP_K_FUSED, 2.0 * N * double(s->amp_bytes), flops, &tl);
    const int nbp = passes[p].num_blocks;
    double *nrm = nullptr;
    const double *fs = nullptr;
    if (norm_partials && p == last_pass) {
      if (ns_dev) fs = ns_dev;
      else nrm = s->red;
    } else if (scale_ahead && p == prev_pass) {
      nrm = s->red;
    }
    int parts = 0;
    if (s->tile_threads == 1024)
      rc = split ? launch_tp<1024, true, 0>(s, dp, p, db, ntiles, nbp)
           : !k4p ? launch_tp<1024, false, 0>(s, dp, p, db, ntiles, nbp)
           : s->tile_k4 == QCS_HIP_TILE_K4_GAUSS
               ? (tp[p].direct
                      ? launch_tp<1024, false, 3>(s, dp, p, db, ntiles, nbp, nrm, &parts, fs)
                      : launch_tp<1024, false, 2>(s, dp, p, db, ntiles, nbp, nrm, &parts, fs))
               : launch_tp<1024, false, 1>(s, dp, p, db, ntiles, nbp);
    else
      rc = split ? launch_tp<512, true, 0>(s, dp, p, db, ntiles, nbp)
                 : launch_tp<512, false, 0>(s, dp, p, db, ntiles, nbp);
    if (rc) return rc;
    if ((rc = timing_end(s, &tl))) return rc;
    if (norm_partials && p == last_pass) *norm_partials = fs ? -1 : parts;
    if (scale_ahead && p == prev_pass && parts > 0 &&
        (rc = fold_norm_partials(s, parts, &ns_dev)))
      return rc;
  }
  // No wait for the passes: the pinned staging buffer that the descriptor
  // copies read is written again only after a stream synchronisation (every
  // host writer of it waits first), and the device buffers only after one
  // (grow_dev, compose_blocks), so the caller prepares its next work while
  // the passes run.  A fault in them surfaces at the next synchronisation.
  return 0;
}

}  // namespace qcs
