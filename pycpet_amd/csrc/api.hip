// C-ABI layer of libcpet_hip.so: contexts, staging, argument validation and
// the reference-compatible entry points (names and argument meaning of
// CPET/utils/math_module.c, bound by CPET/utils/c_ops.py).

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "cpet_internal.h"

namespace cpet {

// Implemented in field.hip / topo.hip.
int field_launch(cpet_ctx* ctx, bool grid_mode, int nx, int ny, int nz,
                 const float* ax, const float* ay, const float* az, long long M,
                 const float* pts, const float4* ch, const double4* chd, int npad,
                 void* E_out, void* phi_out, int soften, uint32_t flags,
                 int forced_k, int forced_splits);
int topo_run(cpet_ctx* ctx, int N, int L, const float* starts,
             const int32_t* n_iter, float h, const float* dims, const float4* ch,
             int npad, double* feats, int32_t* steps, int32_t* exited, int nb_d,
             const double* edges_d, int nb_c, const double* edges_c,
             uint32_t* hist, uint32_t flags, int src);

namespace {
thread_local char g_err[512] = "";
std::mutex g_default_mu;
cpet_ctx* g_default = nullptr;

// Packed charge j holds input charge (j * a) mod N, with a the first odd
// number from 40503 up that is coprime to N: a fixed permutation that spreads
// the input order (in which neighbouring rows are usually neighbouring atoms)
// over the 64-charge tiles.  A tile is then a sample of the whole frame, so
// its FP32 partial sums do not grow monotonically over a run of similar,
// same-direction terms, whose roundings would be correlated.  Sums are
// independent of the order only up to rounding; the permutation depends on
// N alone, so a frame packs the same way in every entry point and batch.
__device__ __forceinline__ long long shuffle_src(long long j, long long n) {
  long long a = 40503;
  for (;;) {  // gcd(a, n) == 1 ?
    long long u = a, v = n;
    while (v) {
      const long long t = u % v;
      u = v;
      v = t;
    }
    if (u == 1) break;
    a += 2;
  }
  return (j * a) % n;
}

__global__ __launch_bounds__(256) void pack_charges_kernel(
    int N, int npad, const float* __restrict__ x, const float* __restrict__ q,
    float4* __restrict__ out, double4* __restrict__ out64) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad) return;
  float4 c;
  if (j < N) {
    const long long g = shuffle_src(j, N);
    c = make_float4(x[3 * g], x[3 * g + 1], x[3 * g + 2], q[g]);
  } else {
    c = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.f);
  }
  out[j] = c;
  if (out64) out64[j] = make_double4(c.x, c.y, c.z, c.w);
}

// Dipoles: out[2j] = (x, y, z, 0), out[2j+1] = (mu, 0); padding dipoles have
// mu = 0 at 1e18 (their terms are exactly 0).
__global__ __launch_bounds__(256) void pack_dipoles_kernel(
    int N, int npad, const float* __restrict__ x, const float* __restrict__ mu,
    float4* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad) return;
  if (j < N) {
    out[2 * j] = make_float4(x[3 * j], x[3 * j + 1], x[3 * j + 2], 0.f);
    out[2 * j + 1] = make_float4(mu[3 * j], mu[3 * j + 1], mu[3 * j + 2], 0.f);
  } else {
    out[2 * j] = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.f);
    out[2 * j + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Batched pack: thread j -> frame f (binary search on chg_off), row j - chg_off[f].
__global__ __launch_bounds__(256) void pack_batch_kernel(
    int F, int total, const int32_t* __restrict__ chg_off, const int64_t* __restrict__ x_off,
    const float* __restrict__ x, const float* __restrict__ q, float4* __restrict__ out,
    double4* __restrict__ out64) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  int lo = 0, hi = F;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (chg_off[mid] <= j)
      lo = mid;
    else
      hi = mid;
  }
  const long long loc = j - chg_off[lo];
  const long long n = x_off[lo + 1] - x_off[lo];
  float4 c;
  if (loc < n) {
    const long long g = x_off[lo] + shuffle_src(loc, n);
    c = make_float4(x[3 * g], x[3 * g + 1], x[3 * g + 2], q[g]);
  } else {
    c = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.f);
  }
  out[j] = c;
  if (out64) out64[j] = make_double4(c.x, c.y, c.z, c.w);
}

bool finite3(const float* v) {
  return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}
}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int ctx_activate(cpet_ctx* ctx) {
  CPET_HIP_TRY(hipSetDevice(ctx->device));
  return CPET_OK;
}

static_assert(kSlotCount <= 32, "arena slot table too small");

void arena_reset(cpet_ctx* ctx) {
  ctx->arena_used = 0;
  for (int i = 0; i < 32; ++i) {
    ctx->arena_slot[i] = nullptr;
    ctx->arena_slot_bytes[i] = 0;
  }
}

int ctx_buf(cpet_ctx* ctx, int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= kSlotCount) {
    set_error("internal: bad scratch slot %d", slot);
    return CPET_ERR_ARG;
  }
  if (bytes == 0) bytes = 256;
  if (ctx->arena) {
    if (ctx->arena_slot[slot] && ctx->arena_slot_bytes[slot] >= bytes) {
      *out = ctx->arena_slot[slot];
      return CPET_OK;
    }
    const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
    if (off > ctx->arena_bytes || bytes > ctx->arena_bytes - off) {
      set_error("this call needs more than the %zu bytes of the fixed scratch array "
                "(%zu in use, %zu more asked for)", ctx->arena_bytes, off, bytes);
      return CPET_ERR_BUDGET;
    }
    ctx->arena_slot[slot] = ctx->arena + off;
    ctx->arena_slot_bytes[slot] = bytes;
    ctx->arena_used = off + bytes;
    *out = ctx->arena + off;
    return CPET_OK;
  }
  DevBuf& b = ctx->bufs[slot];
  if (b.bytes < bytes) {
    // tables cached in this buffer are lost with it
    for (auto& t : ctx->tables)
      if (t.dst && b.ptr && t.dst >= (const char*)b.ptr && t.dst < (const char*)b.ptr + b.bytes)
        t.dst = nullptr;
    if (b.ptr) {
      // the old buffer may still be read by queued work
      CPET_HIP_TRY(hipStreamSynchronize(ctx->stream));
      CPET_HIP_TRY(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;  // grow with headroom
    want = (want + 4095) & ~(size_t)4095;
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
      return CPET_ERR_NOMEM;
    }
    b.bytes = want;
  }
  *out = b.ptr;
  return CPET_OK;
}

int upload_table(cpet_ctx* ctx, int key, void* dst, const void* src, size_t bytes) {
  auto& t = ctx->tables[key];
  const char* d = (const char*)dst;
  if (t.dst == d && t.data.size() == bytes &&
      (bytes == 0 || std::memcmp(t.data.data(), src, bytes) == 0))
    return CPET_OK;
  for (int k = 0; k < kTabCount; ++k) {
    auto& o = ctx->tables[k];
    if (k != key && o.dst && o.dst < d + bytes && d < o.dst + o.data.size()) o.dst = nullptr;
  }
  // a copy from host memory must not be recorded into a graph: the graph
  // would read this host buffer again at every replay
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CPET_HIP_TRY(hipStreamIsCapturing(ctx->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) {
    set_error("batch tables changed while the stream is being captured: run the step once "
              "eagerly on this context, with the same shapes, before capturing it");
    return CPET_ERR_ARG;
  }
  t.dst = nullptr;
  t.data.assign((const char*)src, (const char*)src + bytes);
  if (bytes)
    CPET_HIP_TRY(hipMemcpyAsync(dst, t.data.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
  t.dst = d;
  return CPET_OK;
}

int stage_in(cpet_ctx* ctx, int slot, const void* src, size_t bytes, bool dev,
             const void** out) {
  if (dev) {
    *out = src;
    return CPET_OK;
  }
  void* d = nullptr;
  int rc = ctx_buf(ctx, slot, bytes, &d);
  if (rc) return rc;
  if (bytes)
    CPET_HIP_TRY(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice,
                                ctx->stream));
  *out = d;
  return CPET_OK;
}

int pack_charges(cpet_ctx* ctx, int N, const float* x, const float* q,
                 bool dev, const float4** out, const double4** out64, int* npad) {
  const int np = round_up(N > 0 ? N : 1, kTile);
  const void* dx = nullptr;
  const void* dq = nullptr;
  int rc = stage_in(ctx, kSlotX, x, (size_t)N * 3 * sizeof(float), dev, &dx);
  if (rc) return rc;
  rc = stage_in(ctx, kSlotQ, q, (size_t)N * sizeof(float), dev, &dq);
  if (rc) return rc;
  void* d = nullptr;
  rc = ctx_buf(ctx, kSlotCharges, (size_t)np * sizeof(float4), &d);
  if (rc) return rc;
  void* d64 = nullptr;
  if (out64) {
    rc = ctx_buf(ctx, kSlotCharges64, (size_t)np * sizeof(double4), &d64);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(pack_charges_kernel, dim3((np + 255) / 256), dim3(256), 0,
                     ctx->stream, N, np, (const float*)dx, (const float*)dq,
                     (float4*)d, (double4*)d64);
  CPET_HIP_TRY(hipGetLastError());
  *out = (const float4*)d;
  if (out64) *out64 = (const double4*)d64;
  *npad = np;
  return CPET_OK;
}

int pack_charges_batch(cpet_ctx* ctx, int which, int F, const int32_t* nch, const float* x,
                       const float* q, bool dev, const float4** out,
                       const double4** out64, std::vector<int32_t>& chg_off,
                       const int32_t** d_chg_off) {
  chg_off.assign(F + 1, 0);
  std::vector<int64_t> x_off(F + 1, 0);
  for (int f = 0; f < F; ++f) {
    if (nch[f] <= 0) {
      set_error("frame %d has no charges", f);
      return CPET_ERR_ARG;
    }
    x_off[f + 1] = x_off[f] + nch[f];
    const long long next = (long long)chg_off[f] + round_up(nch[f], kTile);
    if (next > 0x7fffffffLL) {
      set_error("batch too large (%lld padded charges)", next);
      return CPET_ERR_ARG;
    }
    chg_off[f + 1] = (int32_t)next;
  }
  const int total = chg_off[F];
  const long long nraw = x_off[F];
  const void* dx = nullptr;
  const void* dq = nullptr;
  int rc = stage_in(ctx, kSlotX, x, (size_t)nraw * 3 * sizeof(float), dev, &dx);
  if (rc) return rc;
  rc = stage_in(ctx, kSlotQ, q, (size_t)nraw * sizeof(float), dev, &dq);
  if (rc) return rc;
  void* tab = nullptr;
  const size_t tab_bytes = (size_t)(F + 1) * (sizeof(int32_t) + sizeof(int64_t)) + 256;
  if ((rc = ctx_buf(ctx, which ? kSlotBatchTables2 : kSlotBatchTables, tab_bytes, &tab)))
    return rc;
  int64_t* d_xoff = (int64_t*)tab;
  int32_t* d_coff = (int32_t*)(d_xoff + F + 1);
  if ((rc = upload_table(ctx, which ? kTabPackX2 : kTabPackX, d_xoff, x_off.data(),
                         (F + 1) * sizeof(int64_t))))
    return rc;
  if ((rc = upload_table(ctx, which ? kTabPackC2 : kTabPackC, d_coff, chg_off.data(),
                         (F + 1) * sizeof(int32_t))))
    return rc;
  void* d = nullptr;
  if ((rc = ctx_buf(ctx, kSlotCharges, (size_t)total * sizeof(float4), &d))) return rc;
  void* d64 = nullptr;
  if (out64 && (rc = ctx_buf(ctx, kSlotCharges64, (size_t)total * sizeof(double4), &d64)))
    return rc;
  hipLaunchKernelGGL(pack_batch_kernel, dim3((total + 255) / 256), dim3(256), 0, ctx->stream,
                     F, total, d_coff, d_xoff, (const float*)dx, (const float*)dq, (float4*)d,
                     (double4*)d64);
  CPET_HIP_TRY(hipGetLastError());
  *out = (const float4*)d;
  if (out64) *out64 = (const double4*)d64;
  *d_chg_off = d_coff;
  return CPET_OK;
}

int default_ctx(cpet_ctx** out) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    int dev = 0;
    if (const char* s = std::getenv("CPET_DEVICE")) dev = std::atoi(s);
    cpet_ctx* c = nullptr;
    int rc = cpet_ctx_create(dev, &c);
    if (rc) return rc;
    g_default = c;
  }
  *out = g_default;
  return CPET_OK;
}

int prof_mark(cpet_ctx* ctx, int i) {
  if (!ctx->profile) return CPET_OK;
  if (!ctx->ev[i]) CPET_HIP_TRY(hipEventCreate(&ctx->ev[i]));
  CPET_HIP_TRY(hipEventRecord(ctx->ev[i], ctx->stream));
  ctx->marks = i + 1;
  return CPET_OK;
}

int prof_collect(cpet_ctx* ctx, int n) {
  if (!ctx->profile || n < 2) return CPET_OK;
  CPET_HIP_TRY(hipEventSynchronize(ctx->ev[n - 1]));
  for (int i = 0; i + 1 < n && i < 4; ++i)
    CPET_HIP_TRY(hipEventElapsedTime(&ctx->phase_ms[i], ctx->ev[i], ctx->ev[i + 1]));
  ctx->marks = 0;
  return CPET_OK;
}

// Copies device results back for host-pointer calls.
static int copy_out(cpet_ctx* ctx, void* host, const void* dev, size_t bytes) {
  if (bytes)
    CPET_HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost,
                                ctx->stream));
  return CPET_OK;
}

}  // namespace cpet

using namespace cpet;

extern "C" {

const char* cpet_last_error(void) { return g_err; }

int cpet_device_count(int* n) {
  if (!n) {
    set_error("n is NULL");
    return CPET_ERR_ARG;
  }
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return CPET_OK;
}

int cpet_ctx_create(int device, cpet_ctx** out) {
  if (!out) {
    set_error("out is NULL");
    return CPET_ERR_ARG;
  }
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_error("no GPU visible (hipGetDeviceCount returned %d devices)", n);
    return CPET_ERR_NODEV;
  }
  if (device < 0 || device >= n) {
    set_error("device %d out of range [0,%d)", device, n);
    return CPET_ERR_ARG;
  }
  cpet_ctx* c = new cpet_ctx();
  c->device = device;
  c->bufs.resize(kSlotCount);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount,
                              device);
  if (e != hipSuccess) {
    set_error("context setup failed: %s", hipGetErrorString(e));
    delete c;
    return CPET_ERR_HIP;
  }
  c->own_stream = true;
  c->own = c->stream;
  *out = c;
  return CPET_OK;
}

int cpet_ctx_destroy(cpet_ctx* ctx) {
  if (!ctx) return CPET_OK;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream && ctx->own)
      (void)hipStreamSynchronize(ctx->own);
    (void)hipStreamSynchronize(ctx->stream);  // the caller's stream, or the null stream
    for (auto& b : ctx->bufs)
      if (b.ptr) (void)hipFree(b.ptr);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto& ev : ctx->ev)
      if (ev) (void)hipEventDestroy(ev);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
  }
  if (ctx == g_default) {
    std::lock_guard<std::mutex> lk(g_default_mu);
    g_default = nullptr;
  }
  delete ctx;
  return CPET_OK;
}

int cpet_ctx_set_stream(cpet_ctx* ctx, void* hip_stream) {
  if (!ctx) {
    set_error("ctx is NULL");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  // NULL is the null (legacy default) stream, which is what torch reports
  // for its default stream: work queued there is ordered with torch's, so
  // torch may free and reuse its tensors as soon as its stream allows.
  ctx->stream = (hipStream_t)hip_stream;
  return CPET_OK;
}

int cpet_ctx_use_own_stream(cpet_ctx* ctx) {
  if (!ctx) {
    set_error("ctx is NULL");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->stream = ctx->own;
  return CPET_OK;
}

int cpet_ctx_get_stream(cpet_ctx* ctx, void** hip_stream) {
  if (!ctx || !hip_stream) {
    set_error("NULL argument");
    return CPET_ERR_ARG;
  }
  *hip_stream = (void*)ctx->stream;
  return CPET_OK;
}

int cpet_ctx_sync(cpet_ctx* ctx) {
  if (!ctx) {
    set_error("ctx is NULL");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  CPET_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return CPET_OK;
}

int cpet_ctx_set_profiling(cpet_ctx* ctx, int on) {
  if (!ctx) {
    set_error("ctx is NULL");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->profile = on != 0;
  return CPET_OK;
}

int cpet_ctx_phase_times(cpet_ctx* ctx, float* ms, int n) {
  if (!ctx || !ms || n < 0) {
    set_error("cpet_ctx_phase_times: bad arguments");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (ctx->marks > 1) {  // a CPET_F_NO_SYNC call left marks to collect
    int rc = ctx_activate(ctx);
    if (rc) return rc;
    if ((rc = prof_collect(ctx, ctx->marks))) return rc;
  }
  for (int i = 0; i < n && i < 4; ++i) ms[i] = ctx->phase_ms[i];
  return CPET_OK;
}

int cpet_ctx_refined_points(cpet_ctx* ctx, long long* n) {
  if (!ctx || !n) {
    set_error("cpet_ctx_refined_points: bad arguments");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  *n = 0;
  if (!ctx->fix_count) return CPET_OK;
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  int v = 0;
  CPET_HIP_TRY(hipMemcpyAsync(&v, ctx->fix_count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CPET_HIP_TRY(hipStreamSynchronize(ctx->stream));
  *n = v;
  return CPET_OK;
}

int cpet_set_tuning(cpet_ctx* ctx, int kind, int a, int b) {
  if (!ctx || kind < 0 || kind > 5 || a < 0 || b < 0) {
    set_error("cpet_set_tuning: bad arguments");
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->tune[kind][0] = a;
  ctx->tune[kind][1] = b;
  return CPET_OK;
}

static int field_common(cpet_ctx* ctx, bool grid_mode, int nx, int ny, int nz,
                        const float* ax, const float* ay, const float* az,
                        long long M, const float* pts, int N, const float* x,
                        const float* q, void* E_out, void* phi_out, int soften,
                        uint32_t flags) {
  if (!ctx) {
    set_error("ctx is NULL");
    return CPET_ERR_ARG;
  }
  if (M <= 0 || N <= 0) {
    set_error("need M > 0 and N > 0 (got M=%lld N=%d)", M, N);
    return CPET_ERR_ARG;
  }
  if (!x || !q || (!E_out && !phi_out) ||
      (grid_mode ? (!ax || !ay || !az) : !pts)) {
    set_error("NULL array argument");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  const size_t osz = (flags & CPET_F_OUT_F64) ? sizeof(double) : sizeof(float);
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const float4* ch = nullptr;
  const double4* chd = nullptr;
  int npad = 0;
  rc = pack_charges(ctx, N, x, q, dev, &ch, &chd, &npad);
  if (rc) return rc;
  const void *dax = nullptr, *day = nullptr, *daz = nullptr, *dpts = nullptr;
  if (grid_mode) {
    void* axes = nullptr;
    if (dev) {
      dax = ax;
      day = ay;
      daz = az;
    } else {
      rc = ctx_buf(ctx, kSlotAxes, (size_t)(nx + ny + nz) * sizeof(float), &axes);
      if (rc) return rc;
      float* a = (float*)axes;
      CPET_HIP_TRY(hipMemcpyAsync(a, ax, nx * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      CPET_HIP_TRY(hipMemcpyAsync(a + nx, ay, ny * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      CPET_HIP_TRY(hipMemcpyAsync(a + nx + ny, az, nz * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      dax = a;
      day = a + nx;
      daz = a + nx + ny;
    }
  } else {
    rc = stage_in(ctx, kSlotPts, pts, (size_t)M * 3 * sizeof(float), dev, &dpts);
    if (rc) return rc;
  }
  void* dE = E_out;
  void* dphi = phi_out;
  if (!dev) {
    if (E_out) {
      rc = ctx_buf(ctx, kSlotOut0, (size_t)M * 3 * osz, &dE);
      if (rc) return rc;
    }
    if (phi_out) {
      rc = ctx_buf(ctx, kSlotOut1, (size_t)M * osz, &dphi);
      if (rc) return rc;
    }
  }
  const int kind = grid_mode ? 0 : 1;
  rc = field_launch(ctx, grid_mode, nx, ny, nz, (const float*)dax,
                    (const float*)day, (const float*)daz, M,
                    (const float*)dpts, ch, chd, npad, E_out ? dE : nullptr,
                    phi_out ? dphi : nullptr, soften, flags, ctx->tune[kind][0],
                    ctx->tune[kind][1]);
  if (rc) return rc;
  if (!dev) {
    if (E_out && (rc = copy_out(ctx, E_out, dE, (size_t)M * 3 * osz))) return rc;
    if (phi_out && (rc = copy_out(ctx, phi_out, dphi, (size_t)M * osz))) return rc;
  }
  if (!dev || !(flags & CPET_F_NO_SYNC)) {
    CPET_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((rc = prof_collect(ctx, ctx->marks))) return rc;
  }
  return CPET_OK;
}

int cpet_field_points(cpet_ctx* ctx, int M, int N, const float* pts,
                      const float* x, const float* q, void* E_out,
                      void* phi_out, int soften, uint32_t flags) {
  return field_common(ctx, false, 0, 0, 0, nullptr, nullptr, nullptr, M, pts,
                      N, x, q, E_out, phi_out, soften, flags);
}

int cpet_field_grid(cpet_ctx* ctx, int nx, int ny, int nz, const float* ax,
                    const float* ay, const float* az, int N, const float* x,
                    const float* q, void* E_out, void* phi_out, int soften,
                    uint32_t flags) {
  if (nx <= 0 || ny <= 0 || nz <= 0) {
    set_error("grid dimensions must be positive (%d,%d,%d)", nx, ny, nz);
    return CPET_ERR_ARG;
  }
  const long long M = (long long)nx * ny * nz;
  return field_common(ctx, true, nx, ny, nz, ax, ay, az, M, nullptr, N, x, q,
                      E_out, phi_out, soften, flags);
}

int cpet_topo(cpet_ctx* ctx, int N, int L, const float* starts,
              const int32_t* n_iter, float h, const float* dims,
              const float* x, const float* q, double* feats, int32_t* steps,
              int32_t* exited, int nb_d, const double* edges_d, int nb_c,
              const double* edges_c, uint32_t* hist, uint32_t flags) {
  if (!ctx) {
    set_error("ctx is NULL");
    return CPET_ERR_ARG;
  }
  // dims is a parameter like h: a host array in every mode, so a
  // device-pointer call never waits on the stream (and can be captured)
  float hdims[3];
  int rc = topo_check_args(N, L, starts, n_iter, h, dims, x, q, feats, nb_d, edges_d, nb_c,
                           edges_c, hist, hdims);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  rc = ctx_activate(ctx);
  if (rc) return rc;
  return topo_call(ctx, N, L, starts, n_iter, h, hdims, x, q, feats, steps, exited, nb_d,
                   edges_d, nb_c, edges_c, hist, flags);
}

}  // extern "C"

namespace cpet {

int topo_check_args(int N, int L, const float* starts, const int32_t* n_iter, float h,
                    const float* dims, const float* x, const float* q, const double* feats,
                    int nb_d, const double* edges_d, int nb_c, const double* edges_c,
                    const uint32_t* hist, float* hdims) {
  if (N <= 0 || L <= 0) {
    set_error("need N > 0 and L > 0 (got N=%d L=%d)", N, L);
    return CPET_ERR_ARG;
  }
  if (!starts || !n_iter || !dims || !x || !q || !feats) {
    set_error("NULL array argument");
    return CPET_ERR_ARG;
  }
  if (!(std::isfinite(h) && h > 0.f)) {
    set_error("step size must be finite and > 0 (got %g)", (double)h);
    return CPET_ERR_ARG;
  }
  if (hist && (nb_d <= 0 || nb_c <= 0 || !edges_d || !edges_c)) {
    set_error("histogram requested without valid bins/edges");
    return CPET_ERR_ARG;
  }
  std::memcpy(hdims, dims, 3 * sizeof(float));
  if (!finite3(hdims) || !(hdims[0] > 0.f && hdims[1] > 0.f && hdims[2] > 0.f)) {
    set_error("box half-widths must be finite and > 0");
    return CPET_ERR_ARG;
  }
  return CPET_OK;
}

int topo_call(cpet_ctx* ctx, int N, int L, const float* starts, const int32_t* n_iter, float h,
              const float* hdims, const float* x, const float* q, double* feats,
              int32_t* steps, int32_t* exited, int nb_d, const double* edges_d, int nb_c,
              const double* edges_c, uint32_t* hist, uint32_t flags) {
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  const float4* ch = nullptr;
  int npad = 0;
  int rc = pack_charges(ctx, N, x, q, dev, &ch, nullptr, &npad);
  if (rc) return rc;
  return topo_run(ctx, N, L, starts, n_iter, h, hdims, ch, npad, feats, steps,
                  exited, nb_d, edges_d, nb_c, edges_c, hist, flags,
                  (flags & CPET_F_PREC_MASK) == CPET_F_PREC_F64
                      ? ((flags & CPET_F_TOPO_RK4) ? kSrcChargeF64Rk4 : kSrcChargeF64)
                      : ((flags & CPET_F_TOPO_RK4) ? kSrcChargeRk4 : kSrcCharge));
}

}  // namespace cpet

extern "C" {

int cpet_topo_dipole(cpet_ctx* ctx, int N, int L, const float* starts, const int32_t* n_iter,
                     float h, const float* dims, const float* x, const float* mu,
                     double* feats, int32_t* steps, int32_t* exited, uint32_t flags) {
  if (!ctx || N <= 0 || L <= 0 || !starts || !n_iter || !dims || !x || !mu || !feats) {
    set_error("cpet_topo_dipole: bad arguments (N=%d L=%d)", N, L);
    return CPET_ERR_ARG;
  }
  if (!(std::isfinite(h) && h > 0.f)) {
    set_error("step size must be finite and > 0 (got %g)", (double)h);
    return CPET_ERR_ARG;
  }
  float hdims[3];
  std::memcpy(hdims, dims, sizeof(hdims));
  if (!finite3(hdims) || !(hdims[0] > 0.f && hdims[1] > 0.f && hdims[2] > 0.f)) {
    set_error("box half-widths must be finite and > 0");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const int npad = round_up(N, kTile);
  const void *dx = nullptr, *dmu = nullptr;
  if ((rc = stage_in(ctx, kSlotX, x, (size_t)N * 3 * sizeof(float), dev, &dx))) return rc;
  if ((rc = stage_in(ctx, kSlotQ, mu, (size_t)N * 3 * sizeof(float), dev, &dmu))) return rc;
  void* d = nullptr;
  if ((rc = ctx_buf(ctx, kSlotCharges, (size_t)npad * 2 * sizeof(float4), &d))) return rc;
  hipLaunchKernelGGL(pack_dipoles_kernel, dim3((npad + 255) / 256), dim3(256), 0, ctx->stream, N,
                     npad, (const float*)dx, (const float*)dmu, (float4*)d);
  CPET_HIP_TRY(hipGetLastError());
  const int src = (flags & CPET_F_DIPOLE_PHYSICAL) ? kSrcDipolePhys : kSrcDipoleRef;
  return topo_run(ctx, N, L, starts, n_iter, h, hdims, (const float4*)d, npad, feats, steps,
                  exited, 0, nullptr, 0, nullptr, nullptr, flags, src);
}

// ---------------------------------------------------------------------------
// Reference-compatible entry points (host arrays, default context).

int compute_looped_field(int total_points, int n_charges, const float* x_0,
                         const float* x, const float* Q, float* E) {
  cpet_ctx* c = nullptr;
  int rc = default_ctx(&c);
  if (rc) return rc;
  return cpet_field_points(c, total_points, n_charges, x_0, x, Q, E, nullptr,
                           1, 0u);
}

int compute_batched_field(int total_points, int batch_size, int n_charges, const float* x_0,
                          const float* x, const float* Q, float* E) {
  // batch_size only chunks the reference's loop (math_module.c:375-377); a
  // non-positive value would never advance it, so it is rejected here
  if (batch_size <= 0) {
    set_error("compute_batched_field: batch_size must be > 0 (got %d)", batch_size);
    return CPET_ERR_ARG;
  }
  cpet_ctx* c = nullptr;
  int rc = default_ctx(&c);
  if (rc) return rc;
  return cpet_field_points(c, total_points, n_charges, x_0, x, Q, E, nullptr, 0, 0u);
}

int calc_esp_base(float* ESP, const float* x_init, int n_charges,
                  const float* x, const float* Q) {
  if (!ESP) {
    set_error("ESP is NULL");
    return CPET_ERR_ARG;
  }
  cpet_ctx* c = nullptr;
  int rc = default_ctx(&c);
  if (rc) return rc;
  float phi = 0.f;
  rc = cpet_field_points(c, 1, n_charges, x_init, x, Q, nullptr, &phi, 0, 0u);
  if (rc) return rc;
  ESP[0] += phi;  // the reference accumulates into ESP[0]
  return CPET_OK;
}

int calc_field(float* E, const float* x_init, int n_charges, const float* x,
               const float* Q) {
  if (!E) {
    set_error("E is NULL");
    return CPET_ERR_ARG;
  }
  cpet_ctx* c = nullptr;
  int rc = default_ctx(&c);
  if (rc) return rc;
  return cpet_field_points(c, 1, n_charges, x_init, x, Q, E, nullptr, 0, 0u);
}

int calc_field_base(float* E, const float* x_init, int n_charges,
                    const float* x, const float* Q) {
  if (!E) {
    set_error("E is NULL");
    return CPET_ERR_ARG;
  }
  float tmp[3] = {0.f, 0.f, 0.f};
  int rc = calc_field(tmp, x_init, n_charges, x, Q);
  if (rc) return rc;
  E[0] += tmp[0];  // the reference accumulates into E
  E[1] += tmp[1];
  E[2] += tmp[2];
  return CPET_OK;
}

int thread_operation(int n_charges, int n_iter, float step_size,
                     const float* x_0, const float* dimensions, const float* x,
                     const float* Q, float* ret) {
  if (!ret) {
    set_error("ret is NULL");
    return CPET_ERR_ARG;
  }
  cpet_ctx* c = nullptr;
  int rc = default_ctx(&c);
  if (rc) return rc;
  double f[2];
  const int32_t it = n_iter;
  rc = cpet_topo(c, n_charges, 1, x_0, &it, step_size, dimensions, x, Q, f,
                 nullptr, nullptr, 0, nullptr, 0, nullptr, nullptr, 0u);
  if (rc) return rc;
  ret[0] = (float)f[0];
  ret[1] = (float)f[1];
  return CPET_OK;
}

int thread_operation_dipole(int n_dipoles, int n_iter, float step_size, const float* x_0,
                            const float* dimensions, const float* x, const float* mu,
                            float* ret) {
  if (!ret) {
    set_error("ret is NULL");
    return CPET_ERR_ARG;
  }
  cpet_ctx* c = nullptr;
  int rc = default_ctx(&c);
  if (rc) return rc;
  double f[2];
  const int32_t it = n_iter;
  rc = cpet_topo_dipole(c, n_dipoles, 1, x_0, &it, step_size, dimensions, x, mu, f, nullptr,
                        nullptr, 0u);
  if (rc) return rc;
  ret[0] = (float)f[0];
  ret[1] = (float)f[1];
  return CPET_OK;
}

}  // extern "C"
