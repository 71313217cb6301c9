// Frame preparation kernels: atom filters, box-frame transform, compaction.
//
// Replaces the per-frame numpy steps of calculator.__init__
// (CPET/source/calculator.py:238-371):
//   filter_radius_whole_residue  CPET/utils/io.py:283-338
//     keep a whole residue (run of rows with equal (resid, resnum)) iff any
//     of its atoms has |x - center| < radius, the norm computed as numpy
//     does: sqrt((d0*d0 + d1*d1) + d2*d2) in the arithmetic type, no FMA;
//   box-frame transform           CPET/source/calculator.py:341-346,369-371
//     x <- (x - center) @ inv(T) - box_shift;
//   filter_in_box                 CPET/utils/io.py:422-475
//     drop rows with every (x - center)_i strictly inside (-d_i, d_i),
//     compared in float64 as numpy promotes against the float64 dims.
// Residue runs are numbered on the host (they come from string columns); the
// "any atom inside" reduction over a run, the masks and the stable
// compaction of the kept charges run here.

#include <cmath>

#include "cpet_internal.h"

namespace cpet {
namespace {

// Exactly rounded operations without contraction.  The contract(off) pragma
// keeps each product a separate rounded multiply (the default
// -ffp-contract=fast would fuse mul+add into an FMA across these inlined
// helpers), and __builtin_sqrt(f) is IEEE correctly rounded (HIP compiles
// float sqrt correctly rounded by default; the CUDA-style __fsqrt_rn maps to
// the approximate native sqrt in these headers).
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double sub_rn(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}
__device__ __forceinline__ float sqrt_rn(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double sqrt_rn(double a) { return __builtin_sqrt(a); }

__global__ __launch_bounds__(256) void fill_kernel(unsigned int* p, long long n, unsigned int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename TX, typename TS>
__device__ __forceinline__ TS load_as(const void* x, long long i) {
  return (TS)((const TX*)x)[i];
}

// TX: storage type of x; TS: arithmetic type of (x - center) and the norm.
template <typename TX, typename TS>
__global__ __launch_bounds__(256) void radius_mark_kernel(
    int N, const void* __restrict__ x, TS c0, TS c1, TS c2, TS radius,
    const int32_t* __restrict__ runs, unsigned int* __restrict__ run_any) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const TS d0 = sub_rn(load_as<TX, TS>(x, 3LL * i), c0);
  const TS d1 = sub_rn(load_as<TX, TS>(x, 3LL * i + 1), c1);
  const TS d2 = sub_rn(load_as<TX, TS>(x, 3LL * i + 2), c2);
  const TS r = sqrt_rn(add_rn(add_rn(mul_rn(d0, d0), mul_rn(d1, d1)), mul_rn(d2, d2)));
  if (r < radius) atomicOr(&run_any[runs[i]], 1u);
}

__global__ __launch_bounds__(256) void radius_keep_kernel(
    int N, const int32_t* __restrict__ runs, const unsigned int* __restrict__ run_any,
    uint8_t* __restrict__ keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) keep[i] = run_any[runs[i]] ? 1 : 0;
}

// out_j = ((x - c) @ Tinv)_j - s_j.  Each step runs in the type numpy
// would use: TS for x - c, TD for the 3-term dot product (evaluated as
// (d0*T0j + d1*T1j) + d2*T2j, no FMA), TH for the box shift; TO is stored.
template <typename TX, typename TS, typename TD, typename TH, typename TO>
__global__ __launch_bounds__(256) void transform_kernel(
    int N, const void* __restrict__ x, TS c0, TS c1, TS c2, const double* __restrict__ Tinv,
    int has_T, TH s0, TH s1, TH s2, int has_shift, TO* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const TD d[3] = {(TD)sub_rn(load_as<TX, TS>(x, 3LL * i), c0),
                   (TD)sub_rn(load_as<TX, TS>(x, 3LL * i + 1), c1),
                   (TD)sub_rn(load_as<TX, TS>(x, 3LL * i + 2), c2)};
  const TH s[3] = {s0, s1, s2};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    TD v = d[j];
    if (has_T) {
      const TD t0 = (TD)Tinv[0 * 3 + j], t1 = (TD)Tinv[1 * 3 + j], t2 = (TD)Tinv[2 * 3 + j];
      v = add_rn(add_rn(mul_rn(d[0], t0), mul_rn(d[1], t1)), mul_rn(d[2], t2));
    }
    TH w = (TH)v;
    if (has_shift) w = sub_rn(w, s[j]);
    out[3LL * i + j] = (TO)w;
  }
}

template <typename TX, typename TS>
__global__ __launch_bounds__(256) void inbox_keep_kernel(
    int N, const void* __restrict__ x, TS c0, TS c1, TS c2, double e0, double e1,
    double e2, uint8_t* __restrict__ keep) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double r0 = (double)sub_rn(load_as<TX, TS>(x, 3LL * i), c0);
  const double r1 = (double)sub_rn(load_as<TX, TS>(x, 3LL * i + 1), c1);
  const double r2 = (double)sub_rn(load_as<TX, TS>(x, 3LL * i + 2), c2);
  const bool inside = (r0 > -e0) && (r0 < e0) && (r1 > -e1) && (r1 < e1) &&
                      (r2 > -e2) && (r2 < e2);
  keep[i] = inside ? 0 : 1;
}

// ---- stable compaction ---------------------------------------------------

constexpr int kCompactBlock = 256;

__global__ __launch_bounds__(kCompactBlock) void compact_count_kernel(
    int N, const uint8_t* __restrict__ keep, unsigned int* __restrict__ block_counts) {
  __shared__ unsigned int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long b = __ballot(i < N && keep[i]);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cnt, (unsigned int)__popcll(b));
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = cnt;
}

// Exclusive scan of nb block counts (one 1024-thread block); the total goes
// to block_counts[nb].
__global__ __launch_bounds__(1024) void compact_scan_kernel(int nb, unsigned int* counts) {
  __shared__ unsigned int sums[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  unsigned int local = 0;
  for (int k = 0; k < per; ++k) {
    const int i = t * per + k;
    if (i < nb) local += counts[i];
  }
  sums[t] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  unsigned int run = sums[t] - local;
  for (int k = 0; k < per; ++k) {
    const int i = t * per + k;
    if (i < nb) {
      const unsigned int c = counts[i];
      counts[i] = run;
      run += c;
    }
  }
  if (t == 1023) counts[nb] = sums[1023];
}

__global__ __launch_bounds__(kCompactBlock) void compact_scatter_kernel(
    int N, const uint8_t* __restrict__ keep, const unsigned int* __restrict__ offsets,
    const float* __restrict__ x, const float* __restrict__ q, float* __restrict__ xo,
    float* __restrict__ qo, int32_t* __restrict__ idx_out) {
  __shared__ unsigned int wave_base[kCompactBlock / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool k = i < N && keep[i];
  const unsigned long long b = __ballot(k);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wave_base[w] = (unsigned int)__popcll(b);
  __syncthreads();
  unsigned int base = offsets[blockIdx.x];
  for (int j = 0; j < w; ++j) base += wave_base[j];
  if (k) {
    const unsigned int pos = base + (unsigned int)__popcll(b & ((1ull << lane) - 1ull));
    xo[3ull * pos] = x[3LL * i];
    xo[3ull * pos + 1] = x[3LL * i + 1];
    xo[3ull * pos + 2] = x[3LL * i + 2];
    if (q) qo[pos] = q[i];
    if (idx_out) idx_out[pos] = i;
  }
}

template <typename TS>
void launch_radius(hipStream_t st, int dt, int N, const void* x, const double* c, double radius,
                   const int32_t* runs, unsigned int* run_any) {
  const dim3 g((N + 255) / 256), b(256);
  if (dt == 0)
    hipLaunchKernelGGL((radius_mark_kernel<float, TS>), g, b, 0, st, N, x, (TS)c[0], (TS)c[1],
                       (TS)c[2], (TS)radius, runs, run_any);
  else
    hipLaunchKernelGGL((radius_mark_kernel<double, TS>), g, b, 0, st, N, x, (TS)c[0], (TS)c[1],
                       (TS)c[2], (TS)radius, runs, run_any);
}

struct XformArgs {
  int N;
  const void* x;
  const double* c;
  const double* Tinv;
  int has_T;
  const double* s;
  int has_shift;
  void* out;
};

template <typename TX, typename TS, typename TD, typename TH, typename TO>
void launch_xform(hipStream_t st, const XformArgs& a) {
  hipLaunchKernelGGL((transform_kernel<TX, TS, TD, TH, TO>), dim3((a.N + 255) / 256), dim3(256),
                     0, st, a.N, a.x, (TS)a.c[0], (TS)a.c[1], (TS)a.c[2], a.Tinv, a.has_T,
                     (TH)a.s[0], (TH)a.s[1], (TH)a.s[2], a.has_shift, (TO*)a.out);
}

// Resolves the five dtype codes (0 float, 1 double) one at a time.
template <typename... Ts>
struct XformDispatch {
  static void go(hipStream_t st, const int* codes, int n, const XformArgs& a) {
    if constexpr (sizeof...(Ts) == 5) {
      launch_xform<Ts...>(st, a);
    } else {
      if (codes[sizeof...(Ts)] == 0)
        XformDispatch<Ts..., float>::go(st, codes, n, a);
      else
        XformDispatch<Ts..., double>::go(st, codes, n, a);
    }
  }
};

template <typename TS>
void launch_inbox(hipStream_t st, int dt, int N, const void* x, const double* c, const double* e,
                  uint8_t* keep) {
  const dim3 g((N + 255) / 256), b(256);
  if (dt == 0)
    hipLaunchKernelGGL((inbox_keep_kernel<float, TS>), g, b, 0, st, N, x, (TS)c[0], (TS)c[1],
                       (TS)c[2], e[0], e[1], e[2], keep);
  else
    hipLaunchKernelGGL((inbox_keep_kernel<double, TS>), g, b, 0, st, N, x, (TS)c[0], (TS)c[1],
                       (TS)c[2], e[0], e[1], e[2], keep);
}

bool dtype_ok(int d) { return d == 0 || d == 1; }

}  // namespace
}  // namespace cpet

using namespace cpet;

extern "C" {

int cpet_filter_radius_residue(cpet_ctx* ctx, int N, const void* x, int x_dtype,
                               const double* center, int sub_dtype, double radius,
                               const int32_t* runs, int n_runs, uint8_t* keep,
                               uint32_t flags) {
  if (!ctx || N <= 0 || !x || !center || !runs || !keep || n_runs <= 0 ||
      !dtype_ok(x_dtype) || !dtype_ok(sub_dtype)) {
    set_error("cpet_filter_radius_residue: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const size_t xb = (size_t)N * 3 * (x_dtype ? sizeof(double) : sizeof(float));
  const void *dx = nullptr, *druns = nullptr;
  if ((rc = stage_in(ctx, kSlotFrame0, x, xb, dev, &dx))) return rc;
  if ((rc = stage_in(ctx, kSlotFrame1, runs, (size_t)N * sizeof(int32_t), dev, &druns))) return rc;
  void* scratch = nullptr;
  if ((rc = ctx_buf(ctx, kSlotFrame2, (size_t)n_runs * sizeof(unsigned int) + (size_t)N + 256,
                    &scratch)))
    return rc;
  unsigned int* run_any = (unsigned int*)scratch;
  uint8_t* dkeep = dev ? keep : (uint8_t*)(run_any + n_runs);
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(fill_kernel, dim3((n_runs + 255) / 256), dim3(256), 0, st,
                     run_any, (long long)n_runs, 0u);
  if (sub_dtype == 0)
    launch_radius<float>(st, x_dtype, N, dx, center, radius, (const int32_t*)druns, run_any);
  else
    launch_radius<double>(st, x_dtype, N, dx, center, radius, (const int32_t*)druns, run_any);
  hipLaunchKernelGGL(radius_keep_kernel, dim3((N + 255) / 256), dim3(256), 0, st, N,
                     (const int32_t*)druns, run_any, dkeep);
  CPET_HIP_TRY(hipGetLastError());
  if (!dev) CPET_HIP_TRY(hipMemcpyAsync(keep, dkeep, N, hipMemcpyDeviceToHost, st));
  if (!dev || !(flags & CPET_F_NO_SYNC)) CPET_HIP_TRY(hipStreamSynchronize(st));
  return CPET_OK;
}

int cpet_frame_transform(cpet_ctx* ctx, int N, const void* x, int x_dtype,
                         const double* center, int sub_dtype, const double* Tinv,
                         int dot_dtype, const double* box_shift, int shift_dtype, void* out,
                         int out_dtype, uint32_t flags) {
  if (!ctx || N <= 0 || !x || !center || !out || !dtype_ok(x_dtype) || !dtype_ok(sub_dtype) ||
      !dtype_ok(dot_dtype) || !dtype_ok(shift_dtype) || !dtype_ok(out_dtype)) {
    set_error("cpet_frame_transform: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const size_t xb = (size_t)N * 3 * (x_dtype ? sizeof(double) : sizeof(float));
  const size_t ob = (size_t)N * 3 * (out_dtype ? sizeof(double) : sizeof(float));
  const void* dx = nullptr;
  if ((rc = stage_in(ctx, kSlotFrame0, x, xb, dev, &dx))) return rc;
  void* params = nullptr;
  if ((rc = ctx_buf(ctx, kSlotFrame1, 16 * sizeof(double) + ob, &params))) return rc;
  double* dT = (double*)params;
  void* dout = dev ? out : (void*)(dT + 16);
  hipStream_t st = ctx->stream;
  if (Tinv) CPET_HIP_TRY(hipMemcpyAsync(dT, Tinv, 9 * sizeof(double), hipMemcpyHostToDevice, st));
  const double zero[3] = {0.0, 0.0, 0.0};
  XformArgs a;
  a.N = N;
  a.x = dx;
  a.c = center;
  a.Tinv = dT;
  a.has_T = Tinv != nullptr;
  a.s = box_shift ? box_shift : zero;
  a.has_shift = box_shift != nullptr;
  a.out = dout;
  const int codes[5] = {x_dtype, sub_dtype, dot_dtype, shift_dtype, out_dtype};
  XformDispatch<>::go(st, codes, 5, a);
  CPET_HIP_TRY(hipGetLastError());
  if (!dev) CPET_HIP_TRY(hipMemcpyAsync(out, dout, ob, hipMemcpyDeviceToHost, st));
  if (!dev || !(flags & CPET_F_NO_SYNC)) CPET_HIP_TRY(hipStreamSynchronize(st));
  return CPET_OK;
}

int cpet_filter_in_box(cpet_ctx* ctx, int N, const void* x, int x_dtype, const double* center,
                       int sub_dtype, const double* dims, uint8_t* keep, uint32_t flags) {
  if (!ctx || N <= 0 || !x || !center || !dims || !keep || !dtype_ok(x_dtype) ||
      !dtype_ok(sub_dtype)) {
    set_error("cpet_filter_in_box: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const size_t xb = (size_t)N * 3 * (x_dtype ? sizeof(double) : sizeof(float));
  const void* dx = nullptr;
  if ((rc = stage_in(ctx, kSlotFrame0, x, xb, dev, &dx))) return rc;
  void* kb = nullptr;
  if ((rc = ctx_buf(ctx, kSlotFrame2, (size_t)N + 256, &kb))) return rc;
  uint8_t* dkeep = dev ? keep : (uint8_t*)kb;
  hipStream_t st = ctx->stream;
  if (sub_dtype == 0)
    launch_inbox<float>(st, x_dtype, N, dx, center, dims, dkeep);
  else
    launch_inbox<double>(st, x_dtype, N, dx, center, dims, dkeep);
  CPET_HIP_TRY(hipGetLastError());
  if (!dev) CPET_HIP_TRY(hipMemcpyAsync(keep, dkeep, N, hipMemcpyDeviceToHost, st));
  if (!dev || !(flags & CPET_F_NO_SYNC)) CPET_HIP_TRY(hipStreamSynchronize(st));
  return CPET_OK;
}

int cpet_compact_charges(cpet_ctx* ctx, int N, const uint8_t* keep, const float* x,
                         const float* q, float* x_out, float* q_out, int32_t* idx_out,
                         int* count, uint32_t flags) {
  if (!ctx || N <= 0 || !keep || !x || !x_out || !count || (q && !q_out)) {
    set_error("cpet_compact_charges: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const void *dkeep = nullptr, *dx = nullptr, *dq = nullptr;
  if ((rc = stage_in(ctx, kSlotFrame0, keep, (size_t)N, dev, &dkeep))) return rc;
  if ((rc = stage_in(ctx, kSlotFrame1, x, (size_t)N * 3 * sizeof(float), dev, &dx))) return rc;
  if (q && (rc = stage_in(ctx, kSlotFrame2, q, (size_t)N * sizeof(float), dev, &dq))) return rc;
  const int nb = (N + kCompactBlock - 1) / kCompactBlock;
  void* work = nullptr;
  const size_t outb = (size_t)N * (3 * sizeof(float) + sizeof(float) + sizeof(int32_t));
  if ((rc = ctx_buf(ctx, kSlotOut2, ((size_t)nb + 1) * sizeof(unsigned int) + outb + 256, &work)))
    return rc;
  unsigned int* offs = (unsigned int*)work;
  char* ob = (char*)work + (((size_t)nb + 1) * sizeof(unsigned int) + 255) / 256 * 256;
  float* dxo = dev ? x_out : (float*)ob;
  float* dqo = dev ? q_out : (float*)(ob + (size_t)N * 3 * sizeof(float));
  int32_t* dio = dev ? idx_out : (int32_t*)(ob + (size_t)N * 4 * sizeof(float));
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(compact_count_kernel, dim3(nb), dim3(kCompactBlock), 0, st, N,
                     (const uint8_t*)dkeep, offs);
  hipLaunchKernelGGL(compact_scan_kernel, dim3(1), dim3(1024), 0, st, nb, offs);
  hipLaunchKernelGGL(compact_scatter_kernel, dim3(nb), dim3(kCompactBlock), 0, st, N,
                     (const uint8_t*)dkeep, offs, (const float*)dx, (const float*)dq, dxo,
                     q ? dqo : nullptr, (idx_out || !dev) ? dio : nullptr);
  CPET_HIP_TRY(hipGetLastError());
  unsigned int total = 0;
  CPET_HIP_TRY(hipMemcpyAsync(&total, offs + nb, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
  CPET_HIP_TRY(hipStreamSynchronize(st));
  *count = (int)total;
  if (!dev) {
    CPET_HIP_TRY(hipMemcpyAsync(x_out, dxo, (size_t)total * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
    if (q) CPET_HIP_TRY(hipMemcpyAsync(q_out, dqo, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, st));
    if (idx_out) CPET_HIP_TRY(hipMemcpyAsync(idx_out, dio, (size_t)total * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CPET_HIP_TRY(hipStreamSynchronize(st));
  }
  return CPET_OK;
}

}  // extern "C"
