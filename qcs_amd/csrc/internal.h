// Internal declarations of the gfx950 state-vector backend.
//
// Layout in HBM: one array of 2^n amplitudes, interleaved {re, im}
// (double2 for complex128, float2 for complex64), index bit q = qubit q
// (qubit 0 is the least significant bit, as `1L << qubit` in the reference,
// src/q_gates.c:28).  Every update is in place; compat mode adds the
// reference's second buffer and swaps the two pointers like
// src/q_gates.c:146-148.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/qcs_hip.h"

namespace qcs {

constexpr int kThreads = 256;     // work-group size of the streaming kernels
constexpr int kRedBlocks = 1024;  // fixed grid of the FP64 reductions
constexpr int kRedSlots = 8;      // doubles per reduction partial
constexpr long kChunk = 4096;     // amplitudes per sampling chunk
constexpr int kMaxQubits = 34;

struct TimedLaunch {
  hipEvent_t start;
  hipEvent_t stop;
  int cls;
  double bytes;   // algorithmic: the amplitudes the operation reads and writes
  double flops;
  double moved;   // the bytes of every 128-byte line those amplitudes lie in
};

}  // namespace qcs

struct qcs_hip_state {
  int n = 0;
  int precision = QCS_HIP_C128;
  int flags = 0;
  int device = 0;
  uint64_t size = 0;
  size_t amp_bytes = 16;
  void *vec = nullptr;      // current amplitudes
  void *scratch = nullptr;  // compat: the reference's scratch_vector
  hipStream_t stream = nullptr;
  double *red = nullptr;    // kRedBlocks * kRedSlots partials + result slots
  double *host = nullptr;   // pinned staging for small read-backs
  size_t host_bytes = 0;
  // fused-form device buffers (grown on demand)
  void *gates_dev = nullptr;
  size_t gates_cap = 0;
  void *blocks_dev = nullptr;
  size_t blocks_cap = 0;
  float *wbuf = nullptr;    // per block: real 2^(k+1) x 2^(k+1) FP32 matrix
  double *ubuf = nullptr;   // per block: complex 2^k x 2^k FP64 matrix
  size_t wbuf_blocks = 0;
  // sampling buffers
  double *chunk_tot = nullptr;  // per chunk totals, then exclusive prefix
  double *chunk_inc = nullptr;  // inclusive prefix per chunk
  size_t chunk_cap = 0;
  double *u_dev = nullptr;
  long *idx_dev = nullptr;
  size_t shots_cap = 0;
  // tile-program buffers and tuning
  void *tpass_dev = nullptr;
  size_t tpass_cap = 0;
  void *tblock_dev = nullptr;
  size_t tblock_cap = 0;
  int block_variant[QCS_HIP_MAX_FUSE + 1] = {
      QCS_HIP_BLOCK_TB13_512_PIPE, QCS_HIP_BLOCK_TB13_512_PIPE,
      QCS_HIP_BLOCK_TB13_512_PIPE, QCS_HIP_BLOCK_TB13_512_PIPE,
      QCS_HIP_BLOCK_TB13_512_PIPE, QCS_HIP_BLOCK_TB13_512_PIPE};
  int tile_threads = 1024;  // work-group size of the tile-program kernel
  int tile_grid = 256;      // persistent grid of the tile-program kernel
  int tile_math = QCS_HIP_TILE_MATH_FP32;  // block arithmetic of tile programs
  int tile_k4 = QCS_HIP_TILE_K4_GAUSS;     // form of k = 4 blocks in tile programs
  int tile_direct = 1;  // three-product tile programs: first block from HBM,
                        // last block to HBM, where the pass allows
  // kernel timing
  int timing = 0;
  std::vector<qcs::TimedLaunch> pending_events;
  std::vector<hipEvent_t> free_events;
  double t_ms[QCS_HIP_K_COUNT] = {0};
  long t_launches[QCS_HIP_K_COUNT] = {0};
  double t_bytes[QCS_HIP_K_COUNT] = {0};
  double t_flops[QCS_HIP_K_COUNT] = {0};
  double t_moved[QCS_HIP_K_COUNT] = {0};
};

namespace qcs {

// error reporting (thread-local message)
int fail(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);
#define QCS_HIP_TRY(expr)                              \
  do {                                                 \
    hipError_t qcs_e_ = (expr);                        \
    if (qcs_e_ != hipSuccess) return qcs::check_hip(qcs_e_, #expr); \
  } while (0)

// fused blocks of an n-qubit state: k, ascending targets, gate ranges and
// gates inside their block (0 or QCS_HIP_EINVAL with the message set)
int validate_blocks(int n, const qcs_hip_block *blocks, int num_blocks,
                    const qcs_hip_gate *gates, int num_gates);

// launch bracketing for the kernel timers; `moved` (default: `bytes`) for a
// kernel whose amplitudes share 128-byte lines with ones it does not change
void timing_begin(qcs_hip_state *s, int cls, double bytes, double flops,
                  TimedLaunch *tl, double moved = -1.0);
int timing_end(qcs_hip_state *s, TimedLaunch *tl);
// after a launch: check for a launch error
int launch_check(const char *what);

// buffer helpers
int ensure_pinned(qcs_hip_state *s, size_t bytes);

// kernels in gates.hip
int launch_1q(qcs_hip_state *s, int target, const double m[8]);
int launch_cu(qcs_hip_state *s, int control, int target, const double m[8]);
int launch_swap(qcs_hip_state *s, int a, int b);
int launch_fill_basis(qcs_hip_state *s, void *buf, uint64_t basis, int one);
// gate classes of the plain kernels: diag(1, x) (only the bit-1 half
// changes) and X (a pure exchange)
bool is_phase1(const double m[8]);
bool is_x(const double m[8]);

// kernels in reduce.hip
int reduce_norm(qcs_hip_state *s, double *out_host);  // host result
int normalize_device(qcs_hip_state *s);               // no host sync
int scale_state(qcs_hip_state *s, double factor);
int qubit_probs(qcs_hip_state *s, int q, double *p0, double *p1);
// probability of bit q == bit, reading only that half of the state; worth it
// where the half's runs of 2^q amplitudes are at least a 128-byte line
bool half_prob_pays(const qcs_hip_state *s, int q);
int half_prob(qcs_hip_state *s, int q, int bit, double *p);
int collapse(qcs_hip_state *s, int q, int outcome, double factor, int do_scale);
// measurements of qubits[0..count) in order with the uniforms u, decided on
// the device, one host synchronisation
int measure_seq(qcs_hip_state *s, int count, const int *qubits, const double *u,
                int *outcomes);
int argmax(qcs_hip_state *s, long *idx);
int sample(qcs_hip_state *s, long shots, const double *u, long *out);
// qcs_hip_sample's kernels on caller-supplied device buffers, with the chunk
// sums in the file-scope scratch array of reduce.hip (no allocation)
int sample_noalloc(qcs_hip_state *s, long shots, const double *u_dev, long *idx_dev);
int grover_textbook(qcs_hip_state *s, long marked, long iters);
int phase_flip_compat(qcs_hip_state *s, long index);
int diffusion_compat(qcs_hip_state *s);
int negate_one(qcs_hip_state *s, long index);
int compare(qcs_hip_state *a, qcs_hip_state *b, double *omf, double *maxabs);

// kernels in fused.hip
int fused_apply(qcs_hip_state *s, const qcs_hip_block *blocks, int nb,
                const qcs_hip_gate *gates, int ng);
int fused_apply_unitary(qcs_hip_state *s, int k, const int *targets,
                        const double *u);
int compose_blocks(qcs_hip_state *s, const qcs_hip_block *blocks, int nb,
                   const qcs_hip_gate *gates, int ng, size_t extra,
                   char **extra_stage);
// k_compose over nb device blocks: W slots of 64 * 64 floats per block, U
// slots (skipped when ubuf is null) of 32 * 32 complex doubles
constexpr int kWSlotFloats = 64 * 64;
int launch_compose(hipStream_t stream, const qcs_hip_gate *gates,
                   const qcs_hip_block *blocks, int nb, float *wbuf,
                   double *ubuf);

// kernels in tiles.hip
int tile_max_blocks(const qcs_hip_state *s);

// kernels in blockpass.hip
int launch_block_mfma(qcs_hip_state *s, int k, const int *targets,
                      const float *W, int variant);
int block_variant_for(const qcs_hip_state *s, int variant);
// norm_partials (may be null): on return -1 when the passes renormalised the
// state themselves (the last pass scaled by the norm of its input, under
// q_state_normalize's rule), else the number of FP64 partial sums of |x|^2
// the last pass left in s->red (one per work-group), 0 when it did not
int tiles_apply(qcs_hip_state *s, const qcs_hip_pass *passes, int np,
                const qcs_hip_block *blocks, int nb, const qcs_hip_gate *gates,
                int ng, int *norm_partials = nullptr);
// renormalise from `nb` partial sums already in s->red (FP64, the rule of
// normalize_device)
int normalize_from_partials(qcs_hip_state *s, int nb);
// fold `nb` partial sums in s->red into the device result slot; *ns_dev
// points at it (the FP64 norm_sq, valid in stream order)
int fold_norm_partials(qcs_hip_state *s, int nb, const double **ns_dev);

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) {
  return (a + b - 1) / b;
}

}  // namespace qcs
