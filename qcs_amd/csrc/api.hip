// extern "C" entry points of the backend (declared in include/qcs_hip.h):
// argument validation, state lifetime, copies, kernel timers, dispatch.
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <new>

#include "internal.h"

namespace qcs {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                       ? QCS_HIP_ENOMEM
                       : (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
                             ? QCS_HIP_ENODEV
                             : QCS_HIP_EDEVICE;
  return fail(code, "%s: %s", what, hipGetErrorString(e));
}

// Fault injection for tests of the error path: QCS_HIP_INJECT_FAULT=launch:N
// makes the N-th launch check of the process (1-based) report a failure, and
// alloc:N the N-th state allocation.  It is decided on the host: the kernel
// still runs (or the allocation is never made), so the device never faults.
enum { kInjectNone = 0, kInjectLaunch = 1, kInjectAlloc = 2 };
static int g_inject_kind = kInjectNone;
static long g_inject_at = 0;
static std::atomic<long> g_inject_count[3];
static std::once_flag g_inject_once;

static bool inject_fault(int kind) {
  std::call_once(g_inject_once, [] {
    const char *e = std::getenv("QCS_HIP_INJECT_FAULT");
    if (!e) return;
    if (std::strncmp(e, "launch:", 7) == 0) g_inject_kind = kInjectLaunch;
    else if (std::strncmp(e, "alloc:", 6) == 0) g_inject_kind = kInjectAlloc;
    else return;
    g_inject_at = std::atol(std::strchr(e, ':') + 1);
  });
  if (kind != g_inject_kind || g_inject_at <= 0) return false;
  return g_inject_count[kind].fetch_add(1) + 1 == g_inject_at;
}

int launch_check(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess && inject_fault(kInjectLaunch))
    return fail(QCS_HIP_EDEVICE, "%s: injected failure (QCS_HIP_INJECT_FAULT)", what);
  return check_hip(e, what);
}

static hipEvent_t take_event(qcs_hip_state *s) {
  if (!s->free_events.empty()) {
    hipEvent_t e = s->free_events.back();
    s->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void timing_begin(qcs_hip_state *s, int cls, double bytes, double flops,
                  TimedLaunch *tl, double moved) {
  tl->start = tl->stop = nullptr;
  tl->cls = cls;
  tl->bytes = bytes;
  tl->flops = flops;
  tl->moved = moved < 0.0 ? bytes : moved;
  if (!s->timing) return;
  tl->start = take_event(s);
  tl->stop = take_event(s);
  if (tl->start) (void)hipEventRecord(tl->start, s->stream);
}

int timing_end(qcs_hip_state *s, TimedLaunch *tl) {
  if (!s->timing || !tl->start || !tl->stop) return 0;
  QCS_HIP_TRY(hipEventRecord(tl->stop, s->stream));
  s->pending_events.push_back(*tl);
  return 0;
}

// fold finished event pairs into the per-class totals
static int drain_events(qcs_hip_state *s) {
  for (TimedLaunch &tl : s->pending_events) {
    QCS_HIP_TRY(hipEventSynchronize(tl.stop));
    float ms = 0.f;
    QCS_HIP_TRY(hipEventElapsedTime(&ms, tl.start, tl.stop));
    s->t_ms[tl.cls] += ms;
    s->t_launches[tl.cls] += 1;
    s->t_bytes[tl.cls] += tl.bytes;
    s->t_flops[tl.cls] += tl.flops;
    s->t_moved[tl.cls] += tl.moved;
    s->free_events.push_back(tl.start);
    s->free_events.push_back(tl.stop);
  }
  s->pending_events.clear();
  return 0;
}

int ensure_pinned(qcs_hip_state *s, size_t bytes) {
  if (s->host_bytes >= bytes) return 0;
  if (s->host) {
    QCS_HIP_TRY(hipStreamSynchronize(s->stream));
    (void)hipHostFree(s->host);
  }
  s->host = nullptr;
  s->host_bytes = 0;
  size_t want = bytes < 4096 ? 4096 : bytes;
  QCS_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&s->host), want,
                            hipHostMallocDefault));
  s->host_bytes = want;
  return 0;
}

static void free_state(qcs_hip_state *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (TimedLaunch &tl : s->pending_events) {
    (void)hipEventDestroy(tl.start);
    (void)hipEventDestroy(tl.stop);
  }
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  if (s->vec) (void)hipFree(s->vec);
  if (s->scratch) (void)hipFree(s->scratch);
  if (s->red) (void)hipFree(s->red);
  if (s->host) (void)hipHostFree(s->host);
  if (s->gates_dev) (void)hipFree(s->gates_dev);
  if (s->blocks_dev) (void)hipFree(s->blocks_dev);
  if (s->wbuf) (void)hipFree(s->wbuf);
  if (s->ubuf) (void)hipFree(s->ubuf);
  if (s->chunk_tot) (void)hipFree(s->chunk_tot);
  if (s->chunk_inc) (void)hipFree(s->chunk_inc);
  if (s->u_dev) (void)hipFree(s->u_dev);
  if (s->idx_dev) (void)hipFree(s->idx_dev);
  if (s->tpass_dev) (void)hipFree(s->tpass_dev);
  if (s->tblock_dev) (void)hipFree(s->tblock_dev);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

static int check_state(const qcs_hip_state *s) {
  if (!s) return fail(QCS_HIP_EINVAL, "null state");
  return 0;
}

static int check_qubit(const qcs_hip_state *s, int q, const char *what) {
  if (q < 0 || q >= s->n)
    return fail(QCS_HIP_EINVAL, "%s: qubit %d out of range [0,%d)", what, q,
                s->n);
  return 0;
}

static int check_range(const qcs_hip_state *s, long off, long count,
                       const void *p) {
  if (!p && count > 0) return fail(QCS_HIP_EINVAL, "null buffer");
  if (off < 0 || count < 0 || uint64_t(off) > s->size ||
      uint64_t(count) > s->size - uint64_t(off))
    return fail(QCS_HIP_EINVAL, "range [%ld, %ld+%ld) outside 2^%d amplitudes",
                off, off, count, s->n);
  return 0;
}

// copy between a device buffer and host doubles through the pinned buffer,
// converting complex64 <-> double pairs on the host
static int copy_out(qcs_hip_state *s, const void *dev, long off, long count,
                    double *out) {
  const size_t chunk_amps = size_t(1) << 20;
  int rc;
  if ((rc = ensure_pinned(s, chunk_amps * 16))) return rc;
  for (long done = 0; done < count;) {
    const long m = long(std::min<size_t>(chunk_amps, size_t(count - done)));
    const char *src = static_cast<const char *>(dev) + size_t(off + done) * s->amp_bytes;
    QCS_HIP_TRY(hipMemcpyAsync(s->host, src, size_t(m) * s->amp_bytes,
                               hipMemcpyDeviceToHost, s->stream));
    QCS_HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->precision == QCS_HIP_C128) {
      std::memcpy(out + 2 * done, s->host, size_t(m) * 16);
    } else {
      const float *f = reinterpret_cast<const float *>(s->host);
      for (long i = 0; i < 2 * m; ++i) out[2 * done + i] = double(f[i]);
    }
    done += m;
  }
  return 0;
}

int validate_blocks(int n, const qcs_hip_block *blocks, int num_blocks,
                    const qcs_hip_gate *gates, int num_gates) {
  if (num_blocks < 0 || num_gates < 0 || (num_blocks > 0 && !blocks) ||
      (num_gates > 0 && !gates))
    return fail(QCS_HIP_EINVAL, "bad block/gate arrays");
  for (int b = 0; b < num_blocks; ++b) {
    const qcs_hip_block &bl = blocks[b];
    if (bl.k < 1 || bl.k > QCS_HIP_MAX_FUSE || bl.k > n)
      return fail(QCS_HIP_EINVAL, "block %d: k=%d outside [1,%d]", b, bl.k,
                  QCS_HIP_MAX_FUSE);
    for (int q = 0; q < bl.k; ++q) {
      if (bl.targets[q] < 0 || bl.targets[q] >= n ||
          (q > 0 && bl.targets[q] <= bl.targets[q - 1]))
        return fail(QCS_HIP_EINVAL, "block %d: targets not ascending in range", b);
    }
    if (bl.first_gate < 0 || bl.num_gates < 0 ||
        bl.first_gate + bl.num_gates > num_gates)
      return fail(QCS_HIP_EINVAL, "block %d: gate range outside the gate array", b);
    for (int g = bl.first_gate; g < bl.first_gate + bl.num_gates; ++g) {
      const qcs_hip_gate &gt = gates[g];
      bool t_in = false, c_in = gt.kind == QCS_HIP_G_1Q;
      for (int q = 0; q < bl.k; ++q) {
        if (bl.targets[q] == gt.target) t_in = true;
        if (bl.targets[q] == gt.control) c_in = true;
      }
      if (gt.kind < QCS_HIP_G_1Q || gt.kind > QCS_HIP_G_SWAP || !t_in || !c_in ||
          (gt.kind != QCS_HIP_G_1Q && gt.control == gt.target))
        return fail(QCS_HIP_EINVAL, "block %d: gate %d acts outside the block", b, g);
    }
  }
  return 0;
}

}  // namespace qcs

using namespace qcs;

extern "C" {

const char *qcs_hip_last_error(void) { return g_err; }

int qcs_hip_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

int qcs_hip_device_sync(void) {
  QCS_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int qcs_hip_device_name(char *buf, int len) {
  if (!buf || len < 1) return fail(QCS_HIP_EINVAL, "bad buffer");
  int dev = 0;
  QCS_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  QCS_HIP_TRY(hipGetDeviceProperties(&prop, dev));
  char arch[64];
  std::snprintf(arch, sizeof(arch), "%s", prop.gcnArchName);
  for (char *c = arch; *c; ++c)
    if (*c == ':' || *c == ' ') *c = '\0';  // drop target features
  std::snprintf(buf, size_t(len), "%s:%d", arch, prop.multiProcessorCount);
  return 0;
}

int qcs_hip_state_create(qcs_hip_state **out, int num_qubits, int precision,
                         int flags) {
  if (!out) return fail(QCS_HIP_EINVAL, "null output handle");
  *out = nullptr;
  if (num_qubits <= 0 || num_qubits > kMaxQubits)
    return fail(QCS_HIP_EINVAL, "num_qubits %d outside [1,%d]", num_qubits,
                kMaxQubits);
  if (precision != QCS_HIP_C128 && precision != QCS_HIP_C64)
    return fail(QCS_HIP_EINVAL, "unknown precision %d", precision);
  if (flags & ~QCS_HIP_COMPAT) return fail(QCS_HIP_EINVAL, "unknown flags %d", flags);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(QCS_HIP_ENODEV, "no HIP device available");
  qcs_hip_state *s = new (std::nothrow) qcs_hip_state();
  if (!s) return fail(QCS_HIP_ENOMEM, "host allocation failed");
  s->n = num_qubits;
  s->precision = precision;
  s->flags = flags;
  s->size = uint64_t(1) << num_qubits;
  s->amp_bytes = precision == QCS_HIP_C128 ? 16 : 8;
  int rc = 0;
  hipError_t e;
  if ((e = hipGetDevice(&s->device)) != hipSuccess) rc = check_hip(e, "hipGetDevice");
  if (!rc && (e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess)
    rc = check_hip(e, "hipStreamCreate");
  if (!rc && inject_fault(kInjectAlloc))
    rc = fail(QCS_HIP_ENOMEM, "hipMalloc(state): injected failure (QCS_HIP_INJECT_FAULT)");
  if (!rc && (e = hipMalloc(&s->vec, s->size * s->amp_bytes)) != hipSuccess)
    rc = check_hip(e, "hipMalloc(state)");
  if (!rc && (flags & QCS_HIP_COMPAT) &&
      (e = hipMalloc(&s->scratch, s->size * s->amp_bytes)) != hipSuccess)
    rc = check_hip(e, "hipMalloc(scratch)");
  if (!rc && (e = hipMalloc(&s->red, sizeof(double) * (size_t(kRedBlocks) * kRedSlots + 64))) != hipSuccess)
    rc = check_hip(e, "hipMalloc(reduction)");
  if (!rc) rc = ensure_pinned(s, 4096);
  if (!rc) rc = launch_fill_basis(s, s->vec, 0, 1);
  if (!rc && s->scratch) rc = launch_fill_basis(s, s->scratch, 0, 0);
  if (!rc && (e = hipStreamSynchronize(s->stream)) != hipSuccess)
    rc = check_hip(e, "hipStreamSynchronize");
  if (rc) {
    free_state(s);
    return rc;
  }
  *out = s;
  return 0;
}

void qcs_hip_state_destroy(qcs_hip_state *s) { free_state(s); }

int qcs_hip_state_num_qubits(const qcs_hip_state *s) { return s ? s->n : 0; }
int qcs_hip_state_precision(const qcs_hip_state *s) { return s ? s->precision : -1; }
int qcs_hip_state_flags(const qcs_hip_state *s) { return s ? s->flags : 0; }

int qcs_hip_state_set_basis(qcs_hip_state *s, long basis) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (basis < 0 || uint64_t(basis) >= s->size)
    return fail(QCS_HIP_EINVAL, "basis %ld outside 2^%d", basis, s->n);
  if ((rc = launch_fill_basis(s, s->vec, uint64_t(basis), 1))) return rc;
  if (s->scratch && (rc = launch_fill_basis(s, s->scratch, 0, 0))) return rc;
  return 0;
}

int qcs_hip_sync(qcs_hip_state *s) {
  int rc;
  if ((rc = check_state(s))) return rc;
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

int qcs_hip_read(qcs_hip_state *s, long offset, long count, double *out) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_range(s, offset, count, out))) return rc;
  return copy_out(s, s->vec, offset, count, out);
}

int qcs_hip_read_scratch(qcs_hip_state *s, long offset, long count,
                         double *out) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_range(s, offset, count, out))) return rc;
  if (!s->scratch) return fail(QCS_HIP_EUNSUP, "state has no scratch buffer (not compat)");
  return copy_out(s, s->scratch, offset, count, out);
}

int qcs_hip_write(qcs_hip_state *s, long offset, long count, const double *in) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_range(s, offset, count, in))) return rc;
  const size_t chunk_amps = size_t(1) << 20;
  if ((rc = ensure_pinned(s, chunk_amps * 16))) return rc;
  for (long done = 0; done < count;) {
    const long m = long(std::min<size_t>(chunk_amps, size_t(count - done)));
    QCS_HIP_TRY(hipStreamSynchronize(s->stream));  // staging buffer free
    if (s->precision == QCS_HIP_C128) {
      std::memcpy(s->host, in + 2 * done, size_t(m) * 16);
    } else {
      float *f = reinterpret_cast<float *>(s->host);
      for (long i = 0; i < 2 * m; ++i) f[i] = float(in[2 * done + i]);
    }
    char *dst = static_cast<char *>(s->vec) + size_t(offset + done) * s->amp_bytes;
    QCS_HIP_TRY(hipMemcpyAsync(dst, s->host, size_t(m) * s->amp_bytes,
                               hipMemcpyHostToDevice, s->stream));
    done += m;
  }
  QCS_HIP_TRY(hipStreamSynchronize(s->stream));
  return 0;
}

int qcs_hip_apply_1q(qcs_hip_state *s, int target, const double m[8]) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_qubit(s, target, "apply_1q"))) return rc;
  if (!m) return fail(QCS_HIP_EINVAL, "null gate matrix");
  return launch_1q(s, target, m);
}

int qcs_hip_apply_cu(qcs_hip_state *s, int control, int target,
                     const double m[8]) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_qubit(s, control, "apply_cu")) ||
      (rc = check_qubit(s, target, "apply_cu")))
    return rc;
  if (control == target) return fail(QCS_HIP_EINVAL, "apply_cu: control == target");
  if (!m) return fail(QCS_HIP_EINVAL, "null gate matrix");
  return launch_cu(s, control, target, m);
}

int qcs_hip_apply_swap(qcs_hip_state *s, int a, int b) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_qubit(s, a, "apply_swap")) ||
      (rc = check_qubit(s, b, "apply_swap")))
    return rc;
  if (a == b) return 0;
  if (s->flags & QCS_HIP_COMPAT) return fail(QCS_HIP_EUNSUP, "swap in compat mode");
  return launch_swap(s, a, b);
}

int qcs_hip_phase_flip(qcs_hip_state *s, long index) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (index < 0 || uint64_t(index) >= s->size)
    return fail(QCS_HIP_EINVAL, "phase flip index %ld outside 2^%d", index, s->n);
  if (s->flags & QCS_HIP_COMPAT) return phase_flip_compat(s, index);
  return negate_one(s, index);
}

int qcs_hip_diffusion(qcs_hip_state *s) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (s->flags & QCS_HIP_COMPAT) return diffusion_compat(s);
  return grover_textbook(s, -1, 1);
}

int qcs_hip_grover(qcs_hip_state *s, long marked, long iterations) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (iterations <= 0) return 0;
  if (marked >= 0 && uint64_t(marked) >= s->size)
    return fail(QCS_HIP_EINVAL, "marked index %ld outside 2^%d", marked, s->n);
  if (s->flags & QCS_HIP_COMPAT) {
    for (long i = 0; i < iterations; ++i) {
      if (marked >= 0 && (rc = phase_flip_compat(s, marked))) return rc;
      if ((rc = diffusion_compat(s))) return rc;
    }
    return 0;
  }
  return grover_textbook(s, marked, iterations);
}

int qcs_hip_norm_sq(qcs_hip_state *s, double *norm_sq) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (!norm_sq) return fail(QCS_HIP_EINVAL, "null output");
  return reduce_norm(s, norm_sq);
}

int qcs_hip_normalize(qcs_hip_state *s, double *norm_sq_out) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (!norm_sq_out) return normalize_device(s);
  double ns = 0.0;
  if ((rc = reduce_norm(s, &ns))) return rc;
  *norm_sq_out = ns;
  if (ns > 1e-12 && ns != 1.0) return scale_state(s, 1.0 / std::sqrt(ns));
  return 0;
}

int qcs_hip_qubit_probs(qcs_hip_state *s, int qubit, double *p0, double *p1) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_qubit(s, qubit, "qubit_probs"))) return rc;
  double a = 0.0, b = 0.0;
  if ((rc = qubit_probs(s, qubit, &a, &b))) return rc;
  if (p0) *p0 = a;
  if (p1) *p1 = b;
  return 0;
}

int qcs_hip_measure(qcs_hip_state *s, int qubit, double u, int *outcome) {
  int rc;
  if ((rc = check_state(s)) || (rc = check_qubit(s, qubit, "measure"))) return rc;
  // p0 over the bit = 0 half (src/qcs.c:253-259); p1, the norm the kept
  // half has after zeroing, only when the outcome is 1.  Where the halves'
  // runs are shorter than a line, one pass over the whole state gives both.
  double p0 = 0.0, p1 = 0.0;
  const bool halves = half_prob_pays(s, qubit);
  if ((rc = halves ? half_prob(s, qubit, 0, &p0) : qubit_probs(s, qubit, &p0, &p1))) return rc;
  const int o = (u <= p0) ? 0 : 1;  // src/qcs.c:263
  if (o == 1 && halves && (rc = half_prob(s, qubit, 1, &p1))) return rc;
  // after zeroing, q_state_normalize sums only the kept half
  const double ns = o == 0 ? p0 : p1;
  const int do_scale = (ns > 1e-12 && ns != 1.0) ? 1 : 0;
  const double f = do_scale ? 1.0 / std::sqrt(ns) : 1.0;
  if ((rc = collapse(s, qubit, o, f, do_scale))) return rc;
  if (outcome) *outcome = o;
  return 0;
}

int qcs_hip_measure_seq(qcs_hip_state *s, int count, const int *qubits,
                        const double *u, int *outcomes) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (count <= 0) return 0;
  if (!qubits || !u || !outcomes) return fail(QCS_HIP_EINVAL, "null buffer");
  for (int i = 0; i < count; ++i)
    if ((rc = check_qubit(s, qubits[i], "measure_seq"))) return rc;
  return measure_seq(s, count, qubits, u, outcomes);
}

int qcs_hip_probability(qcs_hip_state *s, long index, double *p) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (!p) return fail(QCS_HIP_EINVAL, "null output");
  if (index < 0 || uint64_t(index) >= s->size) {
    *p = 0.0;  // src/qcs.c:392-393
    return 0;
  }
  double v[2];
  if ((rc = copy_out(s, s->vec, index, 1, v))) return rc;
  *p = v[0] * v[0] + v[1] * v[1];
  return 0;
}

int qcs_hip_argmax(qcs_hip_state *s, long *index) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (!index) return fail(QCS_HIP_EINVAL, "null output");
  return argmax(s, index);
}

int qcs_hip_sample(qcs_hip_state *s, long shots, const double *u,
                   long *index_out) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (shots <= 0) return 0;
  if (!u || !index_out) return fail(QCS_HIP_EINVAL, "null buffer");
  return sample(s, shots, u, index_out);
}

// a buffer the device can address (device, managed or pinned host memory);
// an unregistered host pointer would fault the kernel instead of failing
static int check_device_buffer(const void *p, const char *what) {
  if (!p) return fail(QCS_HIP_EINVAL, "%s: null buffer", what);
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();  // not a launch error of a later call
    return fail(QCS_HIP_EINVAL, "%s: %p is not device-accessible memory", what, p);
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged &&
      !(attr.type == hipMemoryTypeHost && attr.devicePointer == p))
    return fail(QCS_HIP_EINVAL, "%s: %p is not device-accessible memory", what, p);
  return 0;
}

int qcs_hip_sample_noalloc(qcs_hip_state *s, long shots, const double *u_dev,
                           long *index_out_dev) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (shots <= 0) return 0;
  if ((rc = check_device_buffer(u_dev, "sample_noalloc uniforms")) ||
      (rc = check_device_buffer(index_out_dev, "sample_noalloc indices")))
    return rc;
  return sample_noalloc(s, shots, u_dev, index_out_dev);
}

int qcs_hip_apply_fused(qcs_hip_state *s, const qcs_hip_block *blocks,
                        int num_blocks, const qcs_hip_gate *gates,
                        int num_gates) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (s->flags & QCS_HIP_COMPAT)
    return fail(QCS_HIP_EUNSUP, "fusion needs textbook semantics");
  if ((rc = validate_blocks(s->n, blocks, num_blocks, gates, num_gates))) return rc;
  return fused_apply(s, blocks, num_blocks, gates, num_gates);
}

int qcs_hip_apply_unitary(qcs_hip_state *s, int k, const int *targets,
                          const double *u) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (s->flags & QCS_HIP_COMPAT)
    return fail(QCS_HIP_EUNSUP, "fusion needs textbook semantics");
  if (k < 1 || k > QCS_HIP_MAX_FUSE || k > s->n || !targets || !u)
    return fail(QCS_HIP_EINVAL, "apply_unitary: bad arguments");
  for (int q = 0; q < k; ++q)
    if (targets[q] < 0 || targets[q] >= s->n || (q > 0 && targets[q] <= targets[q - 1]))
      return fail(QCS_HIP_EINVAL, "apply_unitary: targets not ascending in range");
  return fused_apply_unitary(s, k, targets, u);
}

static int check_passes(qcs_hip_state *s, const qcs_hip_pass *passes, int num_passes,
                        const qcs_hip_block *blocks, int num_blocks,
                        const qcs_hip_gate *gates, int num_gates) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (s->flags & QCS_HIP_COMPAT)
    return fail(QCS_HIP_EUNSUP, "fusion needs textbook semantics");
  if (s->precision != QCS_HIP_C64)
    return fail(QCS_HIP_EUNSUP, "tile programs need a complex64 state");
  if (num_passes < 0 || (num_passes > 0 && !passes))
    return fail(QCS_HIP_EINVAL, "bad pass array");
  // blocks and gates are checked as for qcs_hip_apply_fused
  int covered = 0;
  for (int p = 0; p < num_passes; ++p) {
    const qcs_hip_pass &ps = passes[p];
    if (ps.tile_bits != QCS_HIP_TILE_BITS || s->n < ps.tile_bits + 3)
      return fail(QCS_HIP_EINVAL, "pass %d: tile_bits %d (need %d and n >= %d)",
                  p, ps.tile_bits, QCS_HIP_TILE_BITS, QCS_HIP_TILE_BITS + 3);
    for (int j = 0; j < ps.tile_bits; ++j) {
      if (ps.bits[j] < 0 || ps.bits[j] >= s->n || (j > 0 && ps.bits[j] <= ps.bits[j - 1]))
        return fail(QCS_HIP_EINVAL, "pass %d: bits not ascending in range", p);
      if (j < QCS_HIP_TILE_LOW_BITS && ps.bits[j] != j)
        return fail(QCS_HIP_EINVAL, "pass %d: bits must start with 0..%d", p,
                    QCS_HIP_TILE_LOW_BITS - 1);
    }
    if (ps.first_block != covered || ps.num_blocks < 0 ||
        ps.first_block + ps.num_blocks > num_blocks)
      return fail(QCS_HIP_EINVAL, "pass %d: blocks must follow the previous pass", p);
    if (ps.num_blocks > tile_max_blocks(s))
      return fail(QCS_HIP_EINVAL, "pass %d: %d blocks (at most %d)", p,
                  ps.num_blocks, tile_max_blocks(s));
    for (int b = ps.first_block; b < ps.first_block + ps.num_blocks; ++b) {
      if (blocks[b].k > QCS_HIP_TILE_KMAX)
        return fail(QCS_HIP_EINVAL, "pass %d: block %d has k=%d (at most %d)", p,
                    b, blocks[b].k, QCS_HIP_TILE_KMAX);
      for (int q = 0; q < blocks[b].k && q < QCS_HIP_MAX_FUSE; ++q) {
        bool in = false;
        for (int j = 0; j < ps.tile_bits; ++j) in = in || ps.bits[j] == blocks[b].targets[q];
        if (!in) return fail(QCS_HIP_EINVAL, "pass %d: block %d targets a bit outside the tile", p, b);
      }
    }
    covered += ps.num_blocks;
  }
  if (covered != num_blocks)
    return fail(QCS_HIP_EINVAL, "passes cover %d of %d blocks", covered, num_blocks);
  return validate_blocks(s->n, blocks, num_blocks, gates, num_gates);
}

int qcs_hip_apply_passes(qcs_hip_state *s, const qcs_hip_pass *passes,
                         int num_passes, const qcs_hip_block *blocks,
                         int num_blocks, const qcs_hip_gate *gates,
                         int num_gates) {
  int rc;
  if ((rc = check_passes(s, passes, num_passes, blocks, num_blocks, gates, num_gates)))
    return rc;
  return tiles_apply(s, passes, num_passes, blocks, num_blocks, gates, num_gates);
}

int qcs_hip_apply_passes_normalize(qcs_hip_state *s, const qcs_hip_pass *passes,
                                   int num_passes, const qcs_hip_block *blocks,
                                   int num_blocks, const qcs_hip_gate *gates,
                                   int num_gates) {
  int rc, nparts = 0;
  if ((rc = check_passes(s, passes, num_passes, blocks, num_blocks, gates, num_gates)))
    return rc;
  if ((rc = tiles_apply(s, passes, num_passes, blocks, num_blocks, gates, num_gates,
                        &nparts)))
    return rc;
  if (nparts < 0) return 0;  // renormalised by the last pass
  // the last pass summed |x|^2 of what it wrote: no separate read of the state
  return nparts > 0 ? normalize_from_partials(s, nparts) : normalize_device(s);
}

int qcs_hip_set_tuning(qcs_hip_state *s, int knob, int value) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (knob == QCS_HIP_TUNE_TILE_THREADS) {
    if (value != 512 && value != 1024)
      return fail(QCS_HIP_EINVAL, "tile threads must be 512 or 1024");
    s->tile_threads = value;
    return 0;
  }
  if (knob >= QCS_HIP_TUNE_BLOCK_VARIANT(1) &&
      knob <= QCS_HIP_TUNE_BLOCK_VARIANT(QCS_HIP_MAX_FUSE)) {
    if (value < QCS_HIP_BLOCK_TB12_256 || value > QCS_HIP_BLOCK_TB13_512_PIPE)
      return fail(QCS_HIP_EINVAL, "unknown block variant %d", value);
    s->block_variant[knob - QCS_HIP_TUNE_BLOCK_VARIANT(0)] = value;
    return 0;
  }
  if (knob == QCS_HIP_TUNE_TILE_MATH) {
    if (value != QCS_HIP_TILE_MATH_FP32 && value != QCS_HIP_TILE_MATH_SPLIT_BF16)
      return fail(QCS_HIP_EINVAL, "unknown tile math %d", value);
    s->tile_math = value;
    return 0;
  }
  if (knob == QCS_HIP_TUNE_TILE_K4_FORM) {
    if (value != QCS_HIP_TILE_K4_CHAIN && value != QCS_HIP_TILE_K4_HALVES &&
        value != QCS_HIP_TILE_K4_GAUSS)
      return fail(QCS_HIP_EINVAL, "unknown k = 4 tile form %d", value);
    s->tile_k4 = value;
    return 0;
  }
  if (knob == QCS_HIP_TUNE_TILE_DIRECT) {
    if (value != 0 && value != 1) return fail(QCS_HIP_EINVAL, "tile direct must be 0 or 1");
    s->tile_direct = value;
    return 0;
  }
  if (knob == QCS_HIP_TUNE_TILE_GRID) {
    if (value < 8 || value > 65536 || (value & 7))
      return fail(QCS_HIP_EINVAL, "tile grid must be a multiple of 8 in [8, 65536]");
    s->tile_grid = value;
    return 0;
  }
  return fail(QCS_HIP_EINVAL, "unknown tuning knob %d", knob);
}

int qcs_hip_get_tuning(qcs_hip_state *s, int knob, int *value) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (!value) return fail(QCS_HIP_EINVAL, "null output");
  if (knob == QCS_HIP_TUNE_TILE_THREADS) *value = s->tile_threads;
  else if (knob == QCS_HIP_TUNE_TILE_GRID) *value = s->tile_grid;
  else if (knob == QCS_HIP_TUNE_TILE_MATH) *value = s->tile_math;
  else if (knob == QCS_HIP_TUNE_TILE_K4_FORM) *value = s->tile_k4;
  else if (knob == QCS_HIP_TUNE_TILE_DIRECT) *value = s->tile_direct;
  else if (knob >= QCS_HIP_TUNE_BLOCK_VARIANT(1) &&
           knob <= QCS_HIP_TUNE_BLOCK_VARIANT(QCS_HIP_MAX_FUSE))
    *value = s->block_variant[knob - QCS_HIP_TUNE_BLOCK_VARIANT(0)];
  else
    return fail(QCS_HIP_EINVAL, "unknown tuning knob %d", knob);
  return 0;
}

int qcs_hip_compare(qcs_hip_state *a, qcs_hip_state *b, double *one_minus_f,
                    double *max_abs) {
  int rc;
  if ((rc = check_state(a)) || (rc = check_state(b))) return rc;
  if (a->n != b->n) return fail(QCS_HIP_EINVAL, "compare: sizes differ");
  return compare(a, b, one_minus_f, max_abs);
}

int qcs_hip_timing_enable(qcs_hip_state *s, int enable) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if ((rc = drain_events(s))) return rc;
  for (int c = 0; c < QCS_HIP_K_COUNT; ++c) {
    s->t_ms[c] = 0.0;
    s->t_launches[c] = 0;
    s->t_bytes[c] = 0.0;
    s->t_flops[c] = 0.0;
    s->t_moved[c] = 0.0;
  }
  s->timing = enable ? 1 : 0;
  return 0;
}

int qcs_hip_timing_read(qcs_hip_state *s, int kernel_class, double *total_ms,
                        long *launches, double *bytes, double *flops) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (kernel_class < 0 || kernel_class >= QCS_HIP_K_COUNT)
    return fail(QCS_HIP_EINVAL, "unknown kernel class %d", kernel_class);
  if ((rc = drain_events(s))) return rc;
  if (total_ms) *total_ms = s->t_ms[kernel_class];
  if (launches) *launches = s->t_launches[kernel_class];
  if (bytes) *bytes = s->t_bytes[kernel_class];
  if (flops) *flops = s->t_flops[kernel_class];
  return 0;
}

int qcs_hip_timing_moved(qcs_hip_state *s, int kernel_class, double *moved_bytes) {
  int rc;
  if ((rc = check_state(s))) return rc;
  if (kernel_class < 0 || kernel_class >= QCS_HIP_K_COUNT)
    return fail(QCS_HIP_EINVAL, "unknown kernel class %d", kernel_class);
  if (!moved_bytes) return fail(QCS_HIP_EINVAL, "null output");
  if ((rc = drain_events(s))) return rc;
  *moved_bytes = s->t_moved[kernel_class];
  return 0;
}

}  // extern "C"
