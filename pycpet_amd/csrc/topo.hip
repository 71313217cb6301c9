// Streamline ("topology") kernels for gfx950.
//
// Replaces thread_operation (CPET/utils/math_module.c:523-591) fanned over a
// multiprocessing.Pool (CPET/source/calculator.py:675-712), and the torch
// topo_GPU path (CPET/source/calculator.py:793-978, CPET/utils/gpu.py).
//
// Semantics (per streamline, same as the reference):
//   x <- x + h E(x)/|E(x)| for up to n_iter steps (Euler on the unit field
//   direction, math_module.c:489-503), stopping after the first point that
//   is strictly outside +-dims (math_module.c:539-547; that point is the final
//   point xf); then x0+, x0++, xf+, xf++ (math_module.c:565-568),
//   kappa = |v' x v''| / |v'|^3 with v' = x+ - x, v'' = x++ - 2x+ + x
//   (math_module.c:108-121, 575-584), feature = [|x0 - xf|, mean kappa].
//
// Pipeline:
//   1. order: counting sort of streamline ids by n_iter (4096 keys),
//      longest first, so the persistent kernel's tail consists of short
//      streamlines.
//   2. propagate (persistent): a workgroup of S waves owns 64 streamline
//      slots; wave s sums the field of charge slice s for all 64 slots (FP32
//      per chunk of up to 64 charges, FP64 across chunks; the slices differ
//      by at most four charges and leave out the trailing padding), the S
//      partials are reduced in
//      a fixed order through LDS and every wave applies the identical FP64
//      position update.  A finished slot takes the next id from a global
//      queue.  Results do not depend on which slot runs a streamline.
//   3. epilogue: one lane per streamline in each of four waves that split
//      its sources, the four curvature-defining steps in FP64 (FP32 finite
//      differences at h = 1e-3 lose most digits), distance, and the 2-D
//      histogram with numpy.histogramdd bin rules.
//   Batches run the same three stages for many frames: a per-frame
//   longest-first order, one persistent launch whose workgroups take
//   frames by assignment, and a per-frame epilogue.

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "cpet_internal.h"

namespace cpet {
namespace {

constexpr int kNumBuckets = 4096;

// 1/sqrt(x) from the FP32 hardware estimate refined by one FP64 Newton step
// (relative error ~1e-7 -> ~1e-14), cheaper than an FP64 square root and
// divide.  Clamping y*y keeps x == 0 at +inf (as 1/sqrt(0)).  x must lie in
// the FP32 normal range (a pair's r^2; padding sources at 1e18 give 3e36).
__device__ __forceinline__ double rsqrt_mixed(double x) {
  const double y = (double)__builtin_amdgcn_rsqf((float)x);
  const double e = fma(-x, fmin(y * y, 1.7976931348623157e308), 1.0);
  return fma(0.5 * y, e, y);
}

// 1/sqrt(x) over the whole FP64 range: the FP64 hardware estimate and two
// Newton steps.  x == 0 and x == inf give NaN, as E/|E| does for E = 0 or
// an infinite E.
__device__ __forceinline__ double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}

__global__ __launch_bounds__(256) void fill_u32_kernel(unsigned int* p,
                                                       long long n,
                                                       unsigned int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- 1. longest-first order -------------------------------------------
// Keys are exact below 2048 steps and 16-step classes above (4096 keys), so
// the per-block histograms fit in LDS; only the order of processing depends
// on them, never a result.

__device__ __forceinline__ int bucket_of(int n) {
  n = n < 0 ? 0 : n;
  const int key = n < 2048 ? n : 2048 + min((n - 2048) >> 4, 2047);
  return kNumBuckets - 1 - key;  // descending n_iter
}

__global__ __launch_bounds__(1024) void bucket_count_kernel(
    int L, const int32_t* __restrict__ n_iter, unsigned int* counts) {
  __shared__ unsigned int lh[kNumBuckets];
  for (int b = threadIdx.x; b < kNumBuckets; b += blockDim.x) lh[b] = 0u;
  __syncthreads();
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < L; l += gridDim.x * blockDim.x)
    atomicAdd(&lh[bucket_of(n_iter[l])], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < kNumBuckets; b += blockDim.x)
    if (lh[b]) atomicAdd(&counts[b], lh[b]);
}

// Exclusive scan of kNumBuckets counts by one 1024-thread block.
__global__ __launch_bounds__(1024) void bucket_scan_kernel(unsigned int* counts) {
  __shared__ unsigned int sums[1024];
  constexpr int per = kNumBuckets / 1024;
  const int t = threadIdx.x;
  unsigned int local = 0;
  for (int i = 0; i < per; ++i) local += counts[t * per + i];
  sums[t] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  unsigned int run = sums[t] - local;  // exclusive prefix of this chunk
  for (int i = 0; i < per; ++i) {
    const unsigned int c = counts[t * per + i];
    counts[t * per + i] = run;
    run += c;
  }
}

// One element per thread: rank inside the block's bucket from an LDS
// atomic, one global reservation per (block, non-empty bucket).
__global__ __launch_bounds__(1024) void bucket_scatter_kernel(
    int L, const int32_t* __restrict__ n_iter, unsigned int* offsets,
    int32_t* __restrict__ order) {
  __shared__ unsigned int lh[kNumBuckets];
  __shared__ unsigned int base[kNumBuckets];
  for (int b = threadIdx.x; b < kNumBuckets; b += blockDim.x) lh[b] = 0u;
  __syncthreads();
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  int b = 0;
  unsigned int rank = 0;
  if (l < L) {
    b = bucket_of(n_iter[l]);
    rank = atomicAdd(&lh[b], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kNumBuckets; k += blockDim.x)
    if (lh[k]) base[k] = atomicAdd(&offsets[k], lh[k]);
  __syncthreads();
  if (l < L) order[base[b] + rank] = l;
}

// Per-frame longest-first order for batches: one counting sort over
// (frame, n_iter class) keys in global memory, frame-major, so each frame's
// streamlines land in its own range [sl_off[f], sl_off[f+1]) as local ids.
// 512 classes of 8 steps per frame (n_iter >= 4095 share the first).
constexpr int kBatchBuckets = 512;

__device__ __forceinline__ int batch_bucket_of(int n) {
  n = n < 0 ? 0 : (n > 4095 ? 4095 : n);
  return kBatchBuckets - 1 - (n >> 3);  // descending n_iter
}

// The frame holding streamline l: the largest f with off[f] <= l.  Every
// frame has at least one streamline, so off is strictly increasing.
__device__ __forceinline__ int frame_of(const int32_t* __restrict__ off, int F, int l) {
  int lo = 0, hi = F - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= l)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void batch_bucket_count_kernel(
    int Lt, int F, const int32_t* __restrict__ sl_off, const int32_t* __restrict__ n_iter,
    unsigned int* counts) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= Lt) return;
  const int f = frame_of(sl_off, F, l);
  atomicAdd(&counts[(size_t)f * kBatchBuckets + batch_bucket_of(n_iter[l])], 1u);
}

// Exclusive scan of n counts in place by one 1024-thread block.
__global__ __launch_bounds__(1024) void scan_u32_kernel(unsigned int* a, int n) {
  __shared__ unsigned int sums[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b0 = min(t * per, n), b1 = min(b0 + per, n);
  unsigned int local = 0;
  for (int i = b0; i < b1; ++i) local += a[i];
  sums[t] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  unsigned int run = sums[t] - local;
  for (int i = b0; i < b1; ++i) {
    const unsigned int c = a[i];
    a[i] = run;
    run += c;
  }
}

__global__ __launch_bounds__(256) void batch_bucket_scatter_kernel(
    int Lt, int F, const int32_t* __restrict__ sl_off, const int32_t* __restrict__ n_iter,
    unsigned int* offsets, int32_t* __restrict__ order) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= Lt) return;
  const int f = frame_of(sl_off, F, l);
  const unsigned int pos =
      atomicAdd(&offsets[(size_t)f * kBatchBuckets + batch_bucket_of(n_iter[l])], 1u);
  order[pos] = l - sl_off[f];
}

// ---- 2. persistent propagation ----------------------------------------

struct PropArgs {
  int npad;
  int L;
  const float* starts;
  const int32_t* n_iter;
  const int32_t* order;
  double h, d0, d1, d2;
  double4* res;     // (L) final point, steps and stop kind (ResultRec)
  unsigned int* counter;
  const int* far_unsafe;   // null (no far-field form), or 0: every charge is far
};

// Streams source records [r0, r1) (float4; r0 and r1 multiples of four) to
// body(c, f), four records at a time, body accumulating NC FP32 sums f[NC];
// the sums are FP32 over chunks of kRec records from r0 and FP64 across
// chunks (into e[NC]).  The records pass through two
// SGPR buffers of four float4 (one 64-byte scalar load each).  The load of
// the next group is issued only after the current group has arrived (an
// explicit lgkmcnt(0) wait: scalar loads return out of order, so any later
// wait covers every load in flight), and scheduling barriers keep it ahead
// of the current group's arithmetic, which then hides its latency.  The
// stream runs on across chunk boundaries; the last prefetch is clamped to
// the range's last group, so no load leaves the range.  A last partial chunk
// runs whole group pairs and then at most one single group.
template <int kRec, int NC, int kGroup = 4, class Body>
__device__ __forceinline__ void stream_recs(const float4* __restrict__ ch, int r0, int r1,
                                            Body body, double (&e)[NC]) {
  constexpr int kWaitLgkm0 = 0xc07f;  // s_waitcnt lgkmcnt(0), other counters untouched
  static_assert(kRec % (2 * kGroup) == 0, "a chunk holds whole group pairs");
#pragma unroll
  for (int c = 0; c < NC; ++c) e[c] = 0.0;
  if (r0 >= r1) return;
  const int last = r1 - kGroup;  // index of the range's last group
  int nx = r0;
  float4 a[kGroup], b[kGroup];
#pragma unroll
  for (int u = 0; u < kGroup; ++u) a[u] = ch[nx + u];
  // One group pair: a has been issued; b (the next group, which exists) is
  // loaded during a's arithmetic and the group after it into a during b's.
  auto pair = [&](float (&f)[NC]) __attribute__((always_inline)) {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    nx += kGroup;
#pragma unroll
    for (int u = 0; u < kGroup; ++u) b[u] = ch[nx + u];
    __builtin_amdgcn_sched_barrier(0);
    body(a, f);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    nx += kGroup;
    const int pf = min(nx, last);
#pragma unroll
    for (int u = 0; u < kGroup; ++u) a[u] = ch[pf + u];
    __builtin_amdgcn_sched_barrier(0);
    body(b, f);
  };
  int r = r0;
  for (; r + kRec <= r1; r += kRec) {
    float f[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
#pragma unroll
    for (int g = 0; g < kRec; g += 2 * kGroup) pair(f);
#pragma unroll
    for (int c = 0; c < NC; ++c) e[c] += (double)f[c];
  }
  if (r < r1) {
    float f[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) f[c] = 0.f;
    for (; r + 2 * kGroup <= r1; r += 2 * kGroup) pair(f);
    if (r < r1) {  // one group left, already loaded into a
      __builtin_amdgcn_sched_barrier(0);
      __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
      __builtin_amdgcn_sched_barrier(0);
      body(a, f);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) e[c] += (double)f[c];
  }
}

// Records of the sources [rec0, rec0 + nrec) that precede the trailing
// padding (rps records per source; padding sources sit at kPadCoord with a
// zero charge or moment, so their terms are exactly 0), rounded up to a
// multiple of four.  Padding is fewer than kTile sources at the end of a
// block of a multiple of kTile sources; lane l of the calling wave checks
// source nsrc - 64 + l.  Every lane of the wave must call it.
__device__ __forceinline__ int used_records(const float4* __restrict__ ch, int rec0, int nrec,
                                            int rps, int lane, int gran = 4) {
  const int nsrc = nrec / rps;
  const int sidx = nsrc - 64 + lane;
  bool real = true;
  if (sidx >= 0) {
    const float4 X = ch[rec0 + sidx * rps];
    bool pad = X.x == kPadCoord && X.y == kPadCoord && X.z == kPadCoord;
    if (rps == 1) {
      pad = pad && X.w == 0.f;
    } else {
      const float4 M = ch[rec0 + sidx * rps + 1];
      pad = pad && M.x == 0.f && M.y == 0.f && M.z == 0.f;
    }
    real = !pad;
  }
  const unsigned long long m = __ballot(real);
  const int used = m ? (nsrc - 64 + (63 - __clzll(m)) + 1) * rps : 0;
  return min((used + gran - 1) / gran * gran, nrec);
}

// Splits nrec / gran groups of gran records (from record rec0) evenly over
// S waves: wave w gets groups [w n / S, (w+1) n / S), so slices differ by at
// most one group whatever the source count.
__device__ __forceinline__ void wave_records(int rec0, int nrec, int wave, int S, int& r0,
                                             int& r1, int gran = 4) {
  const long long n = nrec / gran;
  r0 = rec0 + (int)(wave * n / S) * gran;
  r1 = rec0 + (int)((wave + 1) * n / S) * gran;
}

// Sources of the field a streamline follows.  A policy gives the FP32
// slice sum of the propagate kernel and the FP64 field of the epilogue.
//
// Charges: one float4 (x, y, z, q) per source (math_module.c:296-333).
struct ChargeSrc {
  static constexpr int kRecPerSrc = 1;
  static constexpr int kComp = 3;  // the slice sums are the field components
  // Unsoftened field of charge records [r0, r1) at one point (the k factor
  // is applied by the caller).
  // The FP32 terms of one group of four charge records, added into f.
  __device__ static void group(float px, float py, float pz, const float4 (&c)[4],
                               float (&f)[3]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dx = px - c[u].x, dy = py - c[u].y, dz = pz - c[u].z;
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      const float ri = __builtin_amdgcn_rsqf(r2);
      const float s = (c[u].w * ri) * (ri * ri);
      f[0] = fmaf(s, dx, f[0]);
      f[1] = fmaf(s, dy, f[1]);
      f[2] = fmaf(s, dz, f[2]);
    }
  }
  __device__ static void slice(float px, float py, float pz, const float4* __restrict__ ch,
                               int r0, int r1, double (&e)[3]) {
    stream_recs<kTile, 3>(ch, r0, r1, [&](const float4 (&c)[4], float (&f)[3]) {
      group(px, py, pz, c, f);
    }, e);
  }
  // Sources [g0, g1), both multiples of 16.
  __device__ static void field2_f64(const double (&p)[2][3], double (&e)[2][3],
                                    const float4* __restrict__ ch, int g0, int g1) {
    constexpr int kGroup = 8;  // one scalar-load wait per 8 charges
    for (int g = g0; g < g1; g += kGroup) {
      float4 cg[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) cg[u] = ch[g + u];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
      const float4 c = cg[u];
      const double cx = c.x, cy = c.y, cz = c.z, q = c.w;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double dx = p[k][0] - cx, dy = p[k][1] - cy, dz = p[k][2] - cz;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double ri = rsqrt_mixed(r2);
        const double s = (q * ri) * (ri * ri);
        e[k][0] = fma(s, dx, e[k][0]);
        e[k][1] = fma(s, dy, e[k][1]);
        e[k][2] = fma(s, dz, e[k][2]);
      }
      }
    }
  }
};

// Point dipoles (calc_field_base_dipole, math_module.c:335-371): two float4
// per source, (x, y, z, 0) and (mu_x, mu_y, mu_z, 0); the term is
//   r^-5 (c R - mu r^2),  R = p - x,
// with the reference's c = 3 mu_x R_x + mu_y R_y + mu_z R_z, or with
// PHYS the physical c = 3 mu . R.  Padding dipoles (mu = 0 at 1e18) add 0.
template <bool PHYS>
struct DipoleSrc {
  static constexpr int kRecPerSrc = 2;
  static constexpr int kComp = 3;
  __device__ static float coef(float mx, float my, float mz, float dx, float dy, float dz) {
    return PHYS ? 3.f * fmaf(mz, dz, fmaf(my, dy, mx * dx)) : fmaf(3.f * mx, dx, fmaf(my, dy, mz * dz));
  }
  // The FP32 terms of one group of four records (two dipoles), added into f.
  __device__ static void group(float px, float py, float pz, const float4 (&c)[4],
                               float (&f)[3]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const float4 X = c[2 * u], M = c[2 * u + 1];
      const float dx = px - X.x, dy = py - X.y, dz = pz - X.z;
      const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      const float ri = __builtin_amdgcn_rsqf(r2);
      const float ri2 = ri * ri;
      const float ri5 = (ri2 * ri2) * ri;
      const float cc = coef(M.x, M.y, M.z, dx, dy, dz);
      f[0] = fmaf(ri5, fmaf(cc, dx, -M.x * r2), f[0]);
      f[1] = fmaf(ri5, fmaf(cc, dy, -M.y * r2), f[1]);
      f[2] = fmaf(ri5, fmaf(cc, dz, -M.z * r2), f[2]);
    }
  }
  __device__ static void slice(float px, float py, float pz, const float4* __restrict__ ch,
                               int r0, int r1, double (&e)[3]) {
    stream_recs<2 * kTile, 3>(ch, r0, r1, [&](const float4 (&c)[4], float (&f)[3]) {
      group(px, py, pz, c, f);
    }, e);
  }
  // Sources [g0, g1), both multiples of 16.
  __device__ static void field2_f64(const double (&p)[2][3], double (&e)[2][3],
                                    const float4* __restrict__ ch, int g0, int g1) {
    constexpr int kGroup = 4;  // 4 dipoles = 8 float4 per scalar-load wait
    for (int g = g0; g < g1; g += kGroup) {
      float4 cg[2 * kGroup];
#pragma unroll
      for (int u = 0; u < 2 * kGroup; ++u) cg[u] = ch[2 * g + u];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
      const float4 X = cg[2 * u], M = cg[2 * u + 1];
      const double mx = M.x, my = M.y, mz = M.z;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const double dx = p[k][0] - X.x, dy = p[k][1] - X.y, dz = p[k][2] - X.z;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double ri = rsqrt_mixed(r2);
        const double ri2 = ri * ri;
        const double ri5 = (ri2 * ri2) * ri;
        const double cc = PHYS ? 3.0 * fma(mz, dz, fma(my, dy, mx * dx))
                               : fma(3.0 * mx, dx, fma(my, dy, mz * dz));
        e[k][0] = fma(ri5, fma(cc, dx, -mx * r2), e[k][0]);
        e[k][1] = fma(ri5, fma(cc, dy, -my * r2), e[k][1]);
        e[k][2] = fma(ri5, fma(cc, dz, -mz * r2), e[k][2]);
      }
      }
    }
  }
};

// Classical RK4 (weights 1, 2, 2, 1 over 6) on dx/ds = E/|E| with the
// sources of Base: an extension, selected by CPET_F_TOPO_RK4; the
// reference integrates with Euler.  Each step evaluates the field four
// times; positions, the stage points and the weighted direction sum are FP64.
template <class Base>
struct Rk4 : Base {};
template <class Src>
struct IsRk4 {
  static constexpr bool value = false;
};
template <class Base>
struct IsRk4<Rk4<Base>> {
  static constexpr bool value = true;
};

// Charges with the propagate sum in FP64 throughout (CPET_F_PREC_F64): the
// float32 charges are converted exactly, pair terms and sums are FP64, and
// 1/r comes from v_rsq_f64 with two Newton steps.  A precision reference on
// the GPU, not a fast path.  The epilogue is the charge policy's (FP64).
struct ChargeSrcF64 : ChargeSrc {
  __device__ static void slice64(double px, double py, double pz, const float4* __restrict__ ch,
                                 int r0, int r1, double (&e)[3]) {
    constexpr int kGroup = 4;
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int j = r0; j < r1; j += kGroup) {
      float4 c[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) c[u] = ch[j + u];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        const double dx = px - (double)c[u].x, dy = py - (double)c[u].y,
                     dz = pz - (double)c[u].z;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double ri = rsqrt_f64(r2);
        const double s = ((double)c[u].w * ri) * (ri * ri);
        ex = fma(s, dx, ex);
        ey = fma(s, dy, ey);
        ez = fma(s, dz, ez);
      }
    }
    e[0] = ex;
    e[1] = ey;
    e[2] = ez;
  }
};
// Policies whose FP32 slice is stream_recs over 4-record groups with a
// group() body (the one-wave walk streams such groups directly).
template <class Src>
struct HasGroup {
  static constexpr bool value = false;
};
template <>
struct HasGroup<ChargeSrc> {
  static constexpr bool value = true;
};
template <bool P>
struct HasGroup<DipoleSrc<P>> {
  static constexpr bool value = true;
};
template <class Base>
struct HasGroup<Rk4<Base>> : HasGroup<Base> {};

template <class Src>
struct IsF64 {
  static constexpr bool value = false;
};
template <>
struct IsF64<ChargeSrcF64> {
  static constexpr bool value = true;
};
template <class Base>
struct IsF64<Rk4<Base>> : IsF64<Base> {};

// Charges far from the walk region, in the expanded form of the pair term:
//   r^2 = |p|^2 + |c|^2 - 2 p.c,   w = r^-3,
//   S0 = sum q w,   S = sum (q c) w,   E / k = p S0 - S,
// four FP32 sums per tile (FP64 across tiles).  Per pair that is 3 FMA + 1
// add for r^2, the rsq, 2 multiplies for w and 4 FMA: 11 VALU operations
// against the direct form's 13 (3 subtractions, 3 for r^2, rsq, 3 for
// q r^-3, 3 FMA).  Two float4 records per charge: (x, y, z, |c|^2) and
// (q, q x, q y, q z), built by far_pack_kernel.
//
// The expansion loses accuracy when |p| or |c| is large next to r = |p - c|
// (the absolute rounding of r^2 is ~ u (|p| + |c|)^2).  far_pack_kernel
// therefore checks, per frame, that every charge has
//   |dims + h| + |c| <= kFarRatio * dist(c, box grown by h),
// the region every walk point and curvature step stays in; otherwise the
// frame runs the direct form.  Then the relative rounding of each r^2 stays
// below ~ kFarRatio^2 * 3 u, and on the 5,000-charge / +-1.5 A streamline
// workload (charges at least 3 A from the box) the field errors of the two
// forms against FP64 are the same size (max 4.9e-7 and 4.4e-7 of |E| in a
// float32 emulation over 60 points).
//
// The form is opt-in (cpet_set_tuning kind 5, a = 2): on that workload it
// runs 155 ms per frame against the direct form's 146 ms (16 waves x 256
// workgroups), because its two records per charge double the scalar loads
// and the SGPRs that hold them, which costs more than the two VALU
// operations it saves.
constexpr float kFarRatio = 3.0f;

struct ChargeSrcFar {
  static constexpr int kRecPerSrc = 2;
  static constexpr int kComp = 4;   // S0, Sx, Sy, Sz
  static constexpr int kGroup = 8;  // records per scalar-load group: 4 charges
  __device__ static void slice(float px, float py, float pz, const float4* __restrict__ ch,
                               int r0, int r1, double (&e)[4]) {
    // -2 p (exact) and |p|^2, once per evaluation
    const float mx = -2.f * px, my = -2.f * py, mz = -2.f * pz;
    const float pp = fmaf(pz, pz, fmaf(py, py, px * px));
    stream_recs<2 * kTile, 4, kGroup>(ch, r0, r1, [&](const float4 (&c)[kGroup], float (&f)[4]) {
#pragma unroll
      for (int u = 0; u < kGroup / 2; ++u) {
        const float4 X = c[2 * u], Q = c[2 * u + 1];
        float r2 = fmaf(mz, X.z, pp);
        r2 = fmaf(my, X.y, r2);
        r2 = fmaf(mx, X.x, r2);
        r2 = r2 + X.w;
        const float t = __builtin_amdgcn_rsqf(r2);
        const float w = (t * t) * t;
        f[0] = fmaf(Q.x, w, f[0]);
        f[1] = fmaf(Q.y, w, f[1]);
        f[2] = fmaf(Q.z, w, f[2]);
        f[3] = fmaf(Q.w, w, f[3]);
      }
    }, e);
  }
};

// Far-field records of padded charges [0, total) of the frame(s), and the
// per-frame safety flag (unsafe[f] = 1 when a charge breaks the distance
// condition above; cleared by the caller).  hd (F,4) = h, d0, d1, d2 and
// chg_off (F+1) the padded charge offsets of a batch; for one frame (F == 1)
// both are null and h1 = (h, d0, d1, d2).
// Padding charges (q = 0 at 1e18) give (1e18, 1e18, 1e18, 3e36) and zeros:
// their terms are exactly 0 and they are always far.  Every writer of a
// flag stores the same 1, so no atomics are needed.
__global__ __launch_bounds__(256) void far_pack_kernel(int total, const float4* __restrict__ ch,
                                                       int F, const int32_t* __restrict__ chg_off,
                                                       const double* __restrict__ hd, double4 h1,
                                                       float4* __restrict__ far,
                                                       int* __restrict__ unsafe) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  int fr = 0;
  if (F > 1) {
    int lo = 0, hi = F - 1;  // the last frame f with chg_off[f] <= j
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (chg_off[mid] <= j)
        lo = mid;
      else
        hi = mid - 1;
    }
    fr = lo;
  }
  const float4 c = ch[j];
  const float cc = fmaf(c.z, c.z, fmaf(c.y, c.y, c.x * c.x));
  far[2 * j] = make_float4(c.x, c.y, c.z, cc);
  far[2 * j + 1] = make_float4(c.w, c.w * c.x, c.w * c.y, c.w * c.z);
  // distance from c to the box grown by h, and the walk region's radius
  const double4 fh = hd ? make_double4(hd[4 * fr], hd[4 * fr + 1], hd[4 * fr + 2], hd[4 * fr + 3])
                       : h1;
  const double g0 = fh.y + fh.x, g1 = fh.z + fh.x, g2 = fh.w + fh.x;
  const double o0 = fmax(fabs((double)c.x) - g0, 0.0), o1 = fmax(fabs((double)c.y) - g1, 0.0),
               o2 = fmax(fabs((double)c.z) - g2, 0.0);
  const double dist = sqrt(o0 * o0 + o1 * o1 + o2 * o2);
  const double reach = sqrt(g0 * g0 + g1 * g1 + g2 * g2) +
                       sqrt((double)c.x * c.x + (double)c.y * c.y + (double)c.z * c.z);
  if (!(reach <= (double)kFarRatio * dist)) unsafe[fr] = 1;
}

enum SlotState : int { kActive = 0, kNeed = 1, kDone = 2 };

// Result record of one streamline, written once when it stops: the final
// point (FP64) and, in the bit pattern of the fourth double, the step count
// (low word) and the stop kind (high word; 1 = left the box).  One aligned
// 32-byte record per streamline, so each scattered store fills one whole
// 32-byte HBM sector instead of parts of three arrays' sectors; the
// epilogue, which reads the records in streamline order, writes the step
// counts and stop kinds out coalesced.
__device__ __forceinline__ double4 result_rec(double x, double y, double z, int steps, int out) {
  return make_double4(x, y, z, __hiloint2double(out, steps));
}

// One frame's streamline set as the slot loop sees it.  Streamline indices
// are local to the frame; outputs are indexed the same way.
struct FrameView {
  int L;
  const float* starts;     // (L,3)
  const int32_t* n_iter;   // (L)
  const int32_t* order;    // queue position -> streamline, or null (identity)
  double h, d0, d1, d2;
  double4* res;            // (L) result records
  unsigned int* counter;   // next queue position
  int r0, r1;              // this wave's source records (float4)
  int vrec0, vnrec, vgran;  // the frame's used records and split granule (S == 1 walk)
};

// A one-wave workgroup (S == 1) walks the kVirtWaves slices of the
// kVirtWaves-wave split in order and adds their partial sums in that order,
// the order in which the kVirtWaves-wave kernel reduces its waves' partials:
// the two give the same sums bit for bit.  Frames with few charges run
// faster as one-wave workgroups (no barriers, one FP64 update per slot
// instead of one per wave), so the launch shape can follow the charge count
// without changing any result.
constexpr int kVirtWaves = 16;
constexpr int kVirtMaxCharges = 128;  // batch default: one-wave walk up to this many
constexpr int kVirtMaxSingle = 16;    // single-frame default: the same, with
constexpr int kVirtMinStarts = 50000;  // at least this many starts

// Runs the workgroup's 64 slots until the frame's queue is exhausted.  Slot
// state is identical in every wave: each wave sums its charge slice for all
// 64 slots (Src::kComp sums per slot), the partials are reduced in a fixed
// order through LDS, and all waves apply the same FP64 update.  With S >= 4
// waves, waves 0 .. kComp-1 reduce one sum each into esum and the others
// wait at a second barrier: one pass over the partials instead of one per
// wave, which would be S times the LDS reads of the same sums; the second
// barrier also orders the next step's partial writes after every read, so
// one partial buffer suffices (S < 4 alternates two).  A finished slot asks
// the queue for the next streamline (one idle iteration).  first_q >= 0
// gives the slot an initial queue position, first_q < 0 starts it empty.
// part holds kPartBufs(S) * S * 4 * 64 doubles, esum 4 * 64.
// The one-wave walk, general form: the kVirtWaves slices in order, each
// summed by the policy's slice exactly as its wave of the kVirtWaves-wave
// split would, added into fc (which starts at 0) in that order.  An empty
// slice would add +0.0 there, so it is skipped.
template <class Src, int NC>
__device__ __forceinline__ void walk_slices(float px, float py, float pz, double dpx, double dpy,
                                            double dpz, const float4* __restrict__ ch,
                                            const FrameView& f, double (&fc)[NC]) {
  for (int v = 0; v < kVirtWaves; ++v) {
    int r0, r1;
    wave_records(f.vrec0, f.vnrec, v, kVirtWaves, r0, r1, f.vgran);
    if (r0 >= r1) continue;
    double pv[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) pv[c] = 0.0;
    if constexpr (IsF64<Src>::value)
      Src::slice64(dpx, dpy, dpz, ch, r0, r1, pv);
    else
      Src::slice(px, py, pz, ch, r0, r1, pv);
#pragma unroll
    for (int c = 0; c < NC; ++c) fc[c] += pv[c];
  }
}

// The one-wave walk when no slice holds more than one 4-record group (at
// most kVirtWaves groups): the non-empty slices are the groups themselves,
// in order, so the walk streams the groups with the next group's scalar
// loads in flight during the current group's arithmetic (as stream_recs
// does within a slice).  A slice's sum is 0.0 + x, x the FP64 value of its
// group's FP32 sum, which differs from x only for x = -0; fc starts at +0
// and holds sums of multiples of 2^-149, so it is never -0, and adding -0 or
// +0 leaves it the same: adding x directly gives the 16-wave reduction's fc.
template <class Src>
__device__ __forceinline__ void walk_groups(float px, float py, float pz,
                                            const float4* __restrict__ ch, int rec0, int nrec,
                                            double (&fc)[3]) {
  constexpr int kWaitLgkm0 = 0xc07f;  // s_waitcnt lgkmcnt(0)
  const int ng = nrec / 4;
  if (ng <= 0) return;
  const int last = rec0 + 4 * (ng - 1);
  int nx = rec0;
  float4 a[4], b[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) a[u] = ch[nx + u];
  auto one = [&](const float4 (&c)[4]) __attribute__((always_inline)) {
    float g[3] = {0.f, 0.f, 0.f};
    Src::group(px, py, pz, c, g);
#pragma unroll
    for (int k = 0; k < 3; ++k) fc[k] += (double)g[k];
  };
  for (int i = 0; i < ng; i += 2) {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    nx = min(nx + 4, last);
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = ch[nx + u];
    __builtin_amdgcn_sched_barrier(0);
    one(a);
    if (i + 1 >= ng) break;
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    nx = min(nx + 4, last);
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = ch[nx + u];
    __builtin_amdgcn_sched_barrier(0);
    one(b);
  }
  __builtin_amdgcn_s_waitcnt(kWaitLgkm0);  // no load left in flight
}

__host__ __device__ constexpr int kPartBufs(int S) { return S >= 4 ? 1 : 2; }

template <int S, class Src>
__device__ __forceinline__ void run_slots(const float4* __restrict__ ch, const FrameView& f,
                                          int first_q, double* part, double* esum,
                                          int (*nsid)[64], int lane, int wave) {
  constexpr bool rk4 = IsRk4<Src>::value;
  constexpr int NC = Src::kComp;
  constexpr int NB = kPartBufs(S);
  static_assert(S >= 4 ? NC <= S : true, "one reducing wave per sum");
  int st = kDone, sid = 0, step = 0, nmax = 0;
  double px = 0.0, py = 0.0, pz = 0.0;
  // RK4: the stage (0-3), its evaluation point and the weighted direction sum
  int stage = 0;
  double qx = 0.0, qy = 0.0, qz = 0.0, ax = 0.0, ay = 0.0, az = 0.0;

  auto finish = [&](int out) {
    if (wave == 0) f.res[sid] = result_rec(px, py, pz, step, out);
    st = kNeed;
  };
  auto load = [&](int q) {
    if (q >= f.L) {
      st = kDone;
      return;
    }
    sid = f.order ? f.order[q] : q;
    px = (double)f.starts[3 * sid];
    py = (double)f.starts[3 * sid + 1];
    pz = (double)f.starts[3 * sid + 2];
    nmax = f.n_iter[sid];
    step = 0;
    stage = 0;
    qx = px;
    qy = py;
    qz = pz;
    if (nmax <= 0)
      finish(0);  // the reference loop body never runs: xf = x0
    else
      st = kActive;
  };

  if (first_q >= 0)
    load(first_q);
  else
    st = kNeed;
  int par = 0;
  while (__ballot(st != kDone) != 0ull) {
    double fc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) fc[c] = 0.0;
    // the evaluation point as the FP32 slice sees it
    const float ex32 = (float)(rk4 ? qx : px), ey32 = (float)(rk4 ? qy : py),
                ez32 = (float)(rk4 ? qz : pz);
    if (st == kActive) {
      if constexpr (S == 1 && HasGroup<Src>::value) {
        if (f.vnrec <= kVirtWaves * 4) {
          walk_groups<Src>(ex32, ey32, ez32, ch, f.vrec0, f.vnrec, fc);
        } else {
          walk_slices<Src, NC>(ex32, ey32, ez32, rk4 ? qx : px, rk4 ? qy : py, rk4 ? qz : pz,
                               ch, f, fc);
        }
      } else if constexpr (S == 1) {
        walk_slices<Src, NC>(ex32, ey32, ez32, rk4 ? qx : px, rk4 ? qy : py, rk4 ? qz : pz, ch,
                             f, fc);
      } else {
        if constexpr (IsF64<Src>::value)
          Src::slice64(rk4 ? qx : px, rk4 ? qy : py, rk4 ? qz : pz, ch, f.r0, f.r1, fc);
        else
          Src::slice(ex32, ey32, ez32, ch, f.r0, f.r1, fc);
      }
    }
    const int b = NB == 1 ? 0 : par;
#pragma unroll
    for (int c = 0; c < NC; ++c) part[((b * S + wave) * NC + c) * 64 + lane] = fc[c];
    if (wave == 0 && st == kNeed)
      nsid[par][lane] = (int)atomicAdd(f.counter, 1u);
    __syncthreads();
    if constexpr (S >= 4) {
      // esum is read after the second barrier and rewritten only after the
      // next step's first barrier, which every reader has passed by then.
      if (wave < NC) {
        double e = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) e += part[(s * NC + wave) * 64 + lane];
        esum[wave * 64 + lane] = e;
      }
      __syncthreads();
    }
    if (st == kActive) {
      double ev[NC];
      if constexpr (S >= 4) {
#pragma unroll
        for (int c = 0; c < NC; ++c) ev[c] = esum[c * 64 + lane];
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) ev[c] = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
          for (int c = 0; c < NC; ++c) ev[c] += part[((b * S + s) * NC + c) * 64 + lane];
      }
      double ex, ey, ez;
      if constexpr (NC == 4) {
        // E / k = p S0 - S at the point the slice used
        ex = fma((double)ex32, ev[0], -ev[1]);
        ey = fma((double)ey32, ev[0], -ev[2]);
        ez = fma((double)ez32, ev[0], -ev[3]);
      } else {
        ex = ev[0];
        ey = ev[1];
        ez = ev[2];
      }
      ex *= kCoulomb;
      ey *= kCoulomb;
      ez *= kCoulomb;
      bool stepped = true;  // a whole step ended (RK4: after its fourth stage)
      if constexpr (rk4) {
        const double inv = rsqrt_f64(fma(ez, ez, fma(ey, ey, ex * ex)));
        const double ux = ex * inv, uy = ey * inv, uz = ez * inv;
        if (stage == 0) {
          ax = ux;
          ay = uy;
          az = uz;
        } else {
          const double w = stage == 3 ? 1.0 : 2.0;
          ax = fma(w, ux, ax);
          ay = fma(w, uy, ay);
          az = fma(w, uz, az);
        }
        if (stage < 3) {  // next stage point: x + h/2 u1, x + h/2 u2, x + h u3
          const double hs = stage == 2 ? f.h : 0.5 * f.h;
          qx = fma(hs, ux, px);
          qy = fma(hs, uy, py);
          qz = fma(hs, uz, pz);
          ++stage;
          stepped = false;
        } else {
          const double h6 = f.h / 6.0;
          px = fma(h6, ax, px);
          py = fma(h6, ay, py);
          pz = fma(h6, az, pz);
          qx = px;
          qy = py;
          qz = pz;
          stage = 0;
        }
      } else {
        // x += h E/|E|, with one reciprocal square root instead of an FP64
        // square root and three divides.
        const double hn = f.h * rsqrt_f64(fma(ez, ez, fma(ey, ey, ex * ex)));
        px = fma(hn, ex, px);
        py = fma(hn, ey, py);
        pz = fma(hn, ez, pz);
      }
      if (stepped) {
        ++step;
        const int out = (px < -f.d0 || px > f.d0 || py < -f.d1 || py > f.d1 ||
                         pz < -f.d2 || pz > f.d2)
                            ? 1
                            : 0;
        if (out || step >= nmax) finish(out);
      }
    } else if (st == kNeed) {
      load(nsid[par][lane]);
    }
    par ^= 1;
  }
}

// Single frame.  The charge pointer is a separate __restrict__ parameter so
// the compiler may read charges through the scalar cache despite the
// kernel's stores.
template <int S, class Src>
__global__ __launch_bounds__(64 * S, 8) void topo_propagate_kernel(
    const float4* __restrict__ ch, PropArgs a) {
  __shared__ double part[kPartBufs(S) * S * 4 * 64];
  __shared__ double esum[4 * 64];
  __shared__ int nsid[2][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  FrameView f;
  f.L = a.L;
  f.starts = a.starts;
  f.n_iter = a.n_iter;
  f.order = a.order;
  f.h = a.h;
  f.d0 = a.d0;
  f.d1 = a.d1;
  f.d2 = a.d2;
  f.res = a.res;
  f.counter = a.counter;
  if (a.far_unsafe && __builtin_amdgcn_readfirstlane(*a.far_unsafe) == 0)
    return;  // every charge is far: topo_far_kernel runs this launch's work
  const int nrec = __builtin_amdgcn_readfirstlane(
      used_records(ch, 0, a.npad * Src::kRecPerSrc, Src::kRecPerSrc, lane));
  wave_records(0, nrec, wave, S, f.r0, f.r1);
  f.vrec0 = 0;
  f.vnrec = nrec;
  f.vgran = 4;
  run_slots<S, Src>(ch, f, blockIdx.x * 64 + lane, part, esum, nsid, lane, wave);
}

// The far-field form (ChargeSrcFar) of the single-frame kernel, launched
// beside topo_propagate_kernel; exactly one of the two does the work, as the
// device flag decides.  Its own kernel, so that its four-charge scalar-load
// groups (32 SGPRs per buffer) get the SGPR budget of a 4-wave occupancy
// target: one 16-wave workgroup per CU runs 4 waves per SIMD anyway.
template <int S>
__global__ __launch_bounds__(64 * S, 4) void topo_far_kernel(const float4* __restrict__ far,
                                                           PropArgs a) {
  __shared__ double part[kPartBufs(S) * S * 4 * 64];
  __shared__ double esum[4 * 64];
  __shared__ int nsid[2][64];
  if (__builtin_amdgcn_readfirstlane(*a.far_unsafe) != 0) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  FrameView f;
  f.L = a.L;
  f.starts = a.starts;
  f.n_iter = a.n_iter;
  f.order = a.order;
  f.h = a.h;
  f.d0 = a.d0;
  f.d1 = a.d1;
  f.d2 = a.d2;
  f.res = a.res;
  f.counter = a.counter;
  constexpr int rps = ChargeSrcFar::kRecPerSrc, gran = ChargeSrcFar::kGroup;
  const int nrec = __builtin_amdgcn_readfirstlane(used_records(far, 0, a.npad * rps, rps, lane, gran));
  wave_records(0, nrec, wave, S, f.r0, f.r1, gran);
  f.vrec0 = 0;
  f.vnrec = nrec;
  f.vgran = gran;
  run_slots<S, ChargeSrcFar>(far, f, blockIdx.x * 64 + lane, part, esum, nsid, lane, wave);
}

// Many frames in one persistent launch.  Workgroups claim assignments
// (frame ids, several per frame) from a global counter; inside a frame the
// slots refill from that frame's own queue, so a frame's streamlines are
// shared by whichever workgroups joined it and a late joiner just finds the
// queue empty.  Each wave of a workgroup sums its slice of that frame's
// charges (charge reads stay wave-uniform).
struct BatchTopoArgs {
  int F;
  const int32_t* chg_off;     // (F+1) float4 offsets, multiples of kTile
  const int32_t* sl_off;      // (F+1) streamline offsets
  const float* starts;        // (sum L, 3)
  const int32_t* n_iter;      // (sum L)
  const double* hdims;        // (F,4): h, d0, d1, d2
  double4* res;               // (sum L) result records
  unsigned int* sl_counter;   // (F) per-frame queue positions, start at 0
  const int32_t* order;       // (sum L) per-frame longest-first local ids
  const int32_t* assign;      // frame id per assignment
  int n_assign;
  unsigned int* assign_counter;  // [0]: direct-form kernel, [1]: far-form kernel
  const int* far_unsafe;      // (F) 0: frame f's charges are far from its walk region, or null
};

// The batch body for one form: FAR = 0 runs the frames that use the direct
// pair term, FAR = 1 those whose charges are all far from their walk region
// (far_unsafe[f] == 0).  The two forms are separate kernels (with their own
// occupancy targets, see topo_far_kernel), each walking the whole
// assignment list with its own counter and skipping the other form's
// frames, so every frame runs in exactly one of them.
template <int S, bool FAR>
__device__ __forceinline__ void batch_body(const float4* __restrict__ src, BatchTopoArgs a) {
  __shared__ double part[kPartBufs(S) * S * 4 * 64];
  __shared__ double esum[4 * 64];
  __shared__ int nsid[2][64];
  __shared__ int cur;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned int* counter = a.assign_counter + (FAR ? 1 : 0);
  while (true) {
    __syncthreads();  // previous assignment's LDS reads are done
    if (threadIdx.x == 0) cur = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int as = __builtin_amdgcn_readfirstlane(cur);
    if (as >= a.n_assign) break;
    const int fr = a.assign[as];
    const bool far_frame =
        a.far_unsafe && __builtin_amdgcn_readfirstlane(a.far_unsafe[fr]) == 0;
    if (far_frame != FAR) continue;
    const int s0 = a.sl_off[fr];
    FrameView f;
    f.L = a.sl_off[fr + 1] - s0;
    f.starts = a.starts + 3LL * s0;
    f.n_iter = a.n_iter + s0;
    f.order = a.order + s0;
    f.h = a.hdims[4 * fr];
    f.d0 = a.hdims[4 * fr + 1];
    f.d1 = a.hdims[4 * fr + 2];
    f.d2 = a.hdims[4 * fr + 3];
    f.res = a.res + s0;
    f.counter = a.sl_counter + fr;
    const int c0 = a.chg_off[fr], nc = a.chg_off[fr + 1] - c0;
    if constexpr (FAR) {
      constexpr int rps = ChargeSrcFar::kRecPerSrc, gran = ChargeSrcFar::kGroup;
      const int nrec =
          __builtin_amdgcn_readfirstlane(used_records(src, rps * c0, rps * nc, rps, lane, gran));
      wave_records(rps * c0, nrec, wave, S, f.r0, f.r1, gran);
      f.vrec0 = rps * c0;
      f.vnrec = nrec;
      f.vgran = gran;
      run_slots<S, ChargeSrcFar>(src, f, -1, part, esum, nsid, lane, wave);
    } else {
      const int nrec = __builtin_amdgcn_readfirstlane(used_records(src, c0, nc, 1, lane));
      wave_records(c0, nrec, wave, S, f.r0, f.r1);
      f.vrec0 = c0;
      f.vnrec = nrec;
      f.vgran = 4;
      run_slots<S, ChargeSrc>(src, f, -1, part, esum, nsid, lane, wave);
    }
  }
}

template <int S>
__global__ __launch_bounds__(64 * S, 8) void topo_batch_kernel(const float4* __restrict__ ch,
                                                            BatchTopoArgs a) {
  batch_body<S, false>(ch, a);
}

template <int S>
__global__ __launch_bounds__(64 * S, 4) void topo_batch_far_kernel(
    const float4* __restrict__ far, BatchTopoArgs a) {
  batch_body<S, true>(far, a);
}

// ---- 3. FP64 epilogue -------------------------------------------------

struct EpiArgs {
  int L, npad;
  const float* starts;
  const double4* res;  // result records of the propagate kernel
  double h;
  double* feats;
  int32_t* steps;
  int32_t* exited;
  const double* ed;
  int nbd;
  const double* ec;
  int nbc;
  unsigned int* hist;
  int lds_hist;
};

// Epilogue blocks: kEpiWaves waves share kEpiLanes streamlines (one per
// lane); wave w sums sources [w npad / kEpiWaves, (w+1) npad / kEpiWaves) and
// the partial fields are added through LDS in wave order, so every wave
// holds the same FP64 field.  This gives the launch four times the waves of
// one lane per streamline, which the FP64 sum's latency needs.
constexpr int kEpiWaves = 4;
constexpr int kEpiLanes = 64;

struct EpiShared {
  double part[kEpiWaves][2][3][kEpiLanes];
};

// The unsoftened field (unscaled by k) of all sources at two points, FP64:
// wave w of the block sums sources [w npad / kEpiWaves, ...) and the partial
// fields are added through LDS in wave order.
template <class Src>
__device__ __forceinline__ void field2_split(const double (&p)[2][3], double (&e)[2][3],
                                             const float4* __restrict__ ch, int npad,
                                             EpiShared& sh, int wave, int lane) {
  const int per = npad / kEpiWaves;  // npad is a multiple of 64
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) e[k][c] = 0.0;
  Src::field2_f64(p, e, ch, wave * per, wave * per + per);
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) sh.part[wave][k][c][lane] = e[k][c];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < kEpiWaves; ++w) v += sh.part[w][k][c][lane];
      e[k][c] = v;
    }
  __syncthreads();  // part is reused by the next evaluation
}

// Unit field direction at two points, FP64.
template <class Src>
__device__ __forceinline__ void unit2_f64(const double (&p)[2][3], double (&u)[2][3],
                                          const float4* __restrict__ ch, int npad,
                                          EpiShared& sh, int wave, int lane) {
  double e[2][3];
  field2_split<Src>(p, e, ch, npad, sh, wave, lane);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double ex = kCoulomb * e[k][0], ey = kCoulomb * e[k][1],
                 ez = kCoulomb * e[k][2];
    const double n = sqrt(ex * ex + ey * ey + ez * ez);
    u[k][0] = ex / n;
    u[k][1] = ey / n;
    u[k][2] = ez / n;
  }
}

// One step (Euler, or RK4 for Rk4 sources) for two points at once, FP64
// throughout.
template <class Src>
__device__ __forceinline__ void step2_f64(const double (&p)[2][3], double (&o)[2][3],
                                          const float4* __restrict__ ch, int npad, double h,
                                          EpiShared& sh, int wave, int lane) {
  if constexpr (!IsRk4<Src>::value) {
    double e[2][3];
    field2_split<Src>(p, e, ch, npad, sh, wave, lane);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double ex = kCoulomb * e[k][0], ey = kCoulomb * e[k][1],
                   ez = kCoulomb * e[k][2];
      const double n = sqrt(ex * ex + ey * ey + ez * ez);
      o[k][0] = p[k][0] + h * ex / n;
      o[k][1] = p[k][1] + h * ey / n;
      o[k][2] = p[k][2] + h * ez / n;
    }
  } else {
    double u[2][3], q[2][3], acc[2][3];
    unit2_f64<Src>(p, u, ch, npad, sh, wave, lane);
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[k][c] = u[k][c];
        q[k][c] = p[k][c] + 0.5 * h * u[k][c];
      }
    unit2_f64<Src>(q, u, ch, npad, sh, wave, lane);
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[k][c] += 2.0 * u[k][c];
        q[k][c] = p[k][c] + 0.5 * h * u[k][c];
      }
    unit2_f64<Src>(q, u, ch, npad, sh, wave, lane);
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc[k][c] += 2.0 * u[k][c];
        q[k][c] = p[k][c] + h * u[k][c];
      }
    unit2_f64<Src>(q, u, ch, npad, sh, wave, lane);
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) o[k][c] = p[k][c] + (h / 6.0) * (acc[k][c] + u[k][c]);
  }
}

__device__ __forceinline__ double curve_f64(const double (&a)[3],
                                            const double (&b)[3]) {
  const double c0 = a[1] * b[2] - a[2] * b[1];
  const double c1 = a[2] * b[0] - a[0] * b[2];
  const double c2 = a[0] * b[1] - a[1] * b[0];
  const double nc = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  return nc / (na * na * na);
}

// Features (and histogram count) of streamline l of one frame.
// Every wave of the block calls this for its lane's streamline l (l < L
// or not: the steps have barriers); wave 0 stores the features.
template <class Src>
__device__ __forceinline__ void epilogue_one(const float4* __restrict__ ch, int npad,
                                             const float* starts, const double4* res,
                                             double h, double* feats, int32_t* steps,
                                             int32_t* exited, int l, int L,
                                             const EpiArgs& a, unsigned int* lh,
                                             unsigned int* hist, EpiShared& sh) {
  const int wave = threadIdx.x / kEpiLanes, lane = threadIdx.x % kEpiLanes;
  const bool valid = l < L;
  const double4 rec = valid ? res[l] : make_double4(0.0, 0.0, 0.0, 0.0);
  double p[2][3], q1[2][3], q2[2][3];
  for (int i = 0; i < 3; ++i) p[0][i] = valid ? (double)starts[3 * l + i] : 0.0;
  p[1][0] = rec.x;
  p[1][1] = rec.y;
  p[1][2] = rec.z;
  step2_f64<Src>(p, q1, ch, npad, h, sh, wave, lane);
  step2_f64<Src>(q1, q2, ch, npad, h, sh, wave, lane);
  if (!valid || wave != 0) return;
  steps[l] = __double2loint(rec.w);
  exited[l] = __double2hiint(rec.w);
  double kap[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    double v1[3], v2[3];
    for (int i = 0; i < 3; ++i) {
      v1[i] = q1[k][i] - p[k][i];
      v2[i] = (q2[k][i] - 2.0 * q1[k][i]) + p[k][i];
    }
    kap[k] = curve_f64(v1, v2);
  }
  const double d0 = p[0][0] - p[1][0], d1 = p[0][1] - p[1][1], d2 = p[0][2] - p[1][2];
  const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double curv = 0.5 * (kap[0] + kap[1]);
  feats[2 * l] = dist;
  feats[2 * l + 1] = curv;
  if (hist) {
    // bin the float32-rounded features, as the .top file stores them
    const int bd = hist_bin(a.ed, a.nbd, (double)(float)dist);
    const int bc = hist_bin(a.ec, a.nbc, (double)(float)curv);
    if (bd >= 0 && bc >= 0) {
      if (a.lds_hist)
        atomicAdd(&lh[bd * a.nbc + bc], 1u);
      else
        atomicAdd(&hist[bd * a.nbc + bc], 1u);
    }
  }
}

template <class Src>
__global__ __launch_bounds__(kEpiWaves * kEpiLanes) void topo_epilogue_kernel(
    const float4* __restrict__ ch, EpiArgs a) {
  extern __shared__ unsigned int lh[];
  __shared__ EpiShared sh;
  const int nbins = a.nbd * a.nbc;
  if (a.hist && a.lds_hist) {
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) lh[i] = 0u;
    __syncthreads();
  }
  const int l = blockIdx.x * kEpiLanes + threadIdx.x % kEpiLanes;
  epilogue_one<Src>(ch, a.npad, a.starts, a.res, a.h, a.feats, a.steps, a.exited, l, a.L, a, lh,
                    a.hist, sh);
  if (a.hist && a.lds_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
      if (lh[i]) atomicAdd(&a.hist[i], lh[i]);
  }
}

// Batch epilogue: blk[2b] = frame, blk[2b+1] = block within the frame, so a
// workgroup only ever reads one frame's charges.  Histograms are per frame
// (hist + frame * nbins) over shared edges.
__global__ __launch_bounds__(kEpiWaves * kEpiLanes) void topo_batch_epilogue_kernel(
    const float4* __restrict__ ch, EpiArgs a, const int32_t* __restrict__ blk,
    const int32_t* __restrict__ chg_off, const int32_t* __restrict__ sl_off,
    const double* __restrict__ hdims) {
  extern __shared__ unsigned int lh[];
  __shared__ EpiShared sh;
  const int nbins = a.nbd * a.nbc;
  const int fr = blk[2 * blockIdx.x], b = blk[2 * blockIdx.x + 1];
  if (a.hist && a.lds_hist) {
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) lh[i] = 0u;
    __syncthreads();
  }
  const int s0 = sl_off[fr];
  const int L = sl_off[fr + 1] - s0;
  const int l = b * kEpiLanes + threadIdx.x % kEpiLanes;
  unsigned int* hist = a.hist ? a.hist + (size_t)fr * nbins : nullptr;
  epilogue_one<ChargeSrc>(ch + chg_off[fr], chg_off[fr + 1] - chg_off[fr], a.starts + 3LL * s0,
                          a.res + s0, hdims[4 * fr], a.feats + 2LL * s0, a.steps + s0,
                          a.exited + s0, l, L, a, lh, hist, sh);
  if (a.hist && a.lds_hist) {
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
      if (lh[i]) atomicAdd(&hist[i], lh[i]);
  }
}

template <int S, class Src>
hipError_t launch_prop(hipStream_t st, int G, const float4* ch, const float4* far,
                       const PropArgs& a) {
  if constexpr (std::is_same<Src, ChargeSrc>::value && S >= 4) {
    if (far) {
      hipLaunchKernelGGL((topo_far_kernel<S>), dim3(G), dim3(64 * S), 0, st, far, a);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  hipLaunchKernelGGL((topo_propagate_kernel<S, Src>), dim3(G), dim3(64 * S), 0, st, ch, a);
  return hipGetLastError();
}

template <class Src>
hipError_t launch_prop_s(int S, hipStream_t st, int G, const float4* ch, const float4* far,
                         const PropArgs& a) {
  switch (S) {
    case 1: return launch_prop<1, Src>(st, G, ch, far, a);
    case 2: return launch_prop<2, Src>(st, G, ch, far, a);
    case 4: return launch_prop<4, Src>(st, G, ch, far, a);
    case 8: return launch_prop<8, Src>(st, G, ch, far, a);
    case 16: return launch_prop<16, Src>(st, G, ch, far, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int topo_run(cpet_ctx* ctx, int N, int L, const float* starts,
             const int32_t* n_iter, float h, const float* dims, const float4* ch,
             int npad, double* feats, int32_t* steps, int32_t* exited, int nb_d,
             const double* edges_d, int nb_c, const double* edges_c,
             uint32_t* hist, uint32_t flags, int src) {
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  hipStream_t st = ctx->stream;
  int rc;
  const void* dstarts = nullptr;
  const void* diter = nullptr;
  if ((rc = stage_in(ctx, kSlotPts, starts, (size_t)L * 3 * sizeof(float), dev, &dstarts))) return rc;
  if ((rc = stage_in(ctx, kSlotIter, n_iter, (size_t)L * sizeof(int32_t), dev, &diter))) return rc;

  // state: result records (L double4) | feats (L*2 double) | steps (L) | exited (L)
  void* state = nullptr;
  const size_t st_bytes = (size_t)L * (6 * sizeof(double) + 2 * sizeof(int32_t));
  if ((rc = ctx_buf(ctx, kSlotTopoState, st_bytes, &state))) return rc;
  double4* d_res = (double4*)state;
  double* d_feats = dev ? feats : (double*)(d_res + L);
  int32_t* d_steps = (int32_t*)((double*)(d_res + L) + 2 * (size_t)L);
  int32_t* d_exited = d_steps + L;
  if (dev && steps) d_steps = steps;
  if (dev && exited) d_exited = exited;

  // order (L ints) | bucket counts
  void* ord = nullptr;
  if ((rc = ctx_buf(ctx, kSlotTopoOrder, ((size_t)L + kNumBuckets) * sizeof(int32_t), &ord))) return rc;
  int32_t* d_order = (int32_t*)ord;
  unsigned int* d_buckets = (unsigned int*)(d_order + L);

  // misc: counter | hist | edges
  const int nbins = hist ? nb_d * nb_c : 0;
  const size_t misc_bytes = 256 + (size_t)nbins * sizeof(uint32_t) +
                            (hist ? (size_t)(nb_d + nb_c + 2) * sizeof(double) : 0) + 256;
  void* misc = nullptr;
  if ((rc = ctx_buf(ctx, kSlotTopoMisc, misc_bytes, &misc))) return rc;
  unsigned int* d_counter = (unsigned int*)misc;
  unsigned int* d_hist = nullptr;
  const double* d_ed = nullptr;
  const double* d_ec = nullptr;
  if (hist) {
    d_hist = dev ? (unsigned int*)hist : (unsigned int*)((char*)misc + 256);
    double* e = (double*)((char*)misc + 256 + (((size_t)nbins * sizeof(uint32_t) + 255) & ~(size_t)255));
    if (dev) {
      d_ed = edges_d;
      d_ec = edges_c;
    } else {
      CPET_HIP_TRY(hipMemcpyAsync(e, edges_d, (nb_d + 1) * sizeof(double), hipMemcpyHostToDevice, st));
      CPET_HIP_TRY(hipMemcpyAsync(e + nb_d + 1, edges_c, (nb_c + 1) * sizeof(double), hipMemcpyHostToDevice, st));
      d_ed = e;
      d_ec = e + nb_d + 1;
    }
  }

  // far-field records and their safety flag (plain charges, S >= 4)
  // 16 waves x 64 slots: fastest measured (S 4/8/16, G 256..2048); frames
  // of at most kVirtMaxSingle charges and at least kVirtMinStarts starts:
  // one-wave workgroups, the same results (kVirtWaves), 2.2x faster at 16
  // charges and 100,000 starts; slower at 64 charges, and at 20,000 starts,
  // where the longest streamline's step latency bounds the call (2.2 against
  // 1.95 ms)
  int S = ctx->tune[2][0];
  if (S <= 0) S = (N <= kVirtMaxSingle && L >= kVirtMinStarts) ? 1 : 16;
  const float4* d_far = nullptr;
  const int* d_far_unsafe = nullptr;
  if (src == kSrcCharge && S >= 4 && ctx->tune[5][0] == 2) {
    void* fb = nullptr;
    if ((rc = ctx_buf(ctx, kSlotFar, (size_t)npad * 2 * sizeof(float4) + 256, &fb))) return rc;
    int* flag = (int*)((char*)fb + (size_t)npad * 2 * sizeof(float4));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, st, (unsigned int*)flag, 1LL, 0u);
    hipLaunchKernelGGL(far_pack_kernel, dim3((npad + 255) / 256), dim3(256), 0, st, npad, ch, 1,
                       (const int32_t*)nullptr, (const double*)nullptr,
                       make_double4((double)h, (double)dims[0], (double)dims[1], (double)dims[2]),
                       (float4*)fb, flag);
    CPET_HIP_TRY(hipGetLastError());
    d_far = (const float4*)fb;
    d_far_unsafe = flag;
  }

  // 1. longest-first order
  if ((rc = prof_mark(ctx, 0))) return rc;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((kNumBuckets + 255) / 256), dim3(256), 0, st,
                     d_buckets, (long long)kNumBuckets, 0u);
  hipLaunchKernelGGL(bucket_count_kernel, dim3(std::min((L + 1023) / 1024, 256)), dim3(1024),
                     0, st, L, (const int32_t*)diter, d_buckets);
  hipLaunchKernelGGL(bucket_scan_kernel, dim3(1), dim3(1024), 0, st, d_buckets);
  hipLaunchKernelGGL(bucket_scatter_kernel, dim3((L + 1023) / 1024), dim3(1024), 0, st, L,
                     (const int32_t*)diter, d_buckets, d_order);
  CPET_HIP_TRY(hipGetLastError());

  // 2. propagate
  int G = ctx->tune[2][1];
  const int gmax = (L + 63) / 64;
  // One 16-wave workgroup per CU: the fewest slots that keep every SIMD fed
  // (4 waves each, with the charge loads prefetched), so the tail of long
  // streamlines still running after the queue drains is shortest.
  if (G <= 0)
    G = std::min(gmax, std::max(1, ctx->num_cus * (S >= 16 ? 1 : S == 1 ? 24 : 2)));
  G = std::min(G, gmax);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, st, d_counter, 1LL,
                     (unsigned int)(G * 64));
  if ((rc = prof_mark(ctx, 1))) return rc;
  PropArgs pa;
  pa.npad = npad;
  pa.L = L;
  pa.starts = (const float*)dstarts;
  pa.n_iter = (const int32_t*)diter;
  pa.order = d_order;
  pa.h = (double)h;
  pa.d0 = dims[0];
  pa.d1 = dims[1];
  pa.d2 = dims[2];
  pa.res = d_res;
  pa.counter = d_counter;
  pa.far_unsafe = d_far_unsafe;
  if (S != 1 && S != 2 && S != 4 && S != 8 && S != 16) {
    set_error("unsupported waves per workgroup %d (1,2,4,8,16)", S);
    return CPET_ERR_ARG;
  }
  hipError_t e =
      src == kSrcCharge         ? launch_prop_s<ChargeSrc>(S, st, G, ch, d_far, pa)
      : src == kSrcChargeRk4    ? launch_prop_s<Rk4<ChargeSrc>>(S, st, G, ch, nullptr, pa)
      : src == kSrcChargeF64    ? launch_prop_s<ChargeSrcF64>(S, st, G, ch, nullptr, pa)
      : src == kSrcChargeF64Rk4 ? launch_prop_s<Rk4<ChargeSrcF64>>(S, st, G, ch, nullptr, pa)
      : src == kSrcDipoleRef    ? launch_prop_s<DipoleSrc<false>>(S, st, G, ch, nullptr, pa)
                                : launch_prop_s<DipoleSrc<true>>(S, st, G, ch, nullptr, pa);
  if (e != hipSuccess) {
    set_error("topo propagate launch failed: %s", hipGetErrorString(e));
    return CPET_ERR_HIP;
  }

  // 3. epilogue (+ histogram)
  EpiArgs ea;
  ea.L = L;
  ea.npad = npad;
  ea.starts = (const float*)dstarts;
  ea.res = d_res;
  ea.h = (double)h;
  ea.feats = d_feats;
  ea.steps = d_steps;
  ea.exited = d_exited;
  ea.ed = d_ed;
  ea.nbd = nb_d;
  ea.ec = d_ec;
  ea.nbc = nb_c;
  ea.hist = d_hist;
  ea.lds_hist = (hist && nbins <= 8192) ? 1 : 0;
  if ((rc = prof_mark(ctx, 2))) return rc;
  if (hist)
    hipLaunchKernelGGL(fill_u32_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st,
                       d_hist, (long long)nbins, 0u);
  const size_t lds = ea.lds_hist ? (size_t)nbins * sizeof(unsigned int) : 0;
  // the FP64-sum variants share the charge epilogues (already FP64)
  if (src == kSrcCharge || src == kSrcChargeF64)
    hipLaunchKernelGGL(topo_epilogue_kernel<ChargeSrc>, dim3((L + kEpiLanes - 1) / kEpiLanes),
                       dim3(kEpiWaves * kEpiLanes), lds,
                       st, ch, ea);
  else if (src == kSrcChargeRk4 || src == kSrcChargeF64Rk4)
    hipLaunchKernelGGL(topo_epilogue_kernel<Rk4<ChargeSrc>>, dim3((L + kEpiLanes - 1) / kEpiLanes),
                       dim3(kEpiWaves * kEpiLanes), lds, st, ch, ea);
  else if (src == kSrcDipoleRef)
    hipLaunchKernelGGL(topo_epilogue_kernel<DipoleSrc<false>>, dim3((L + kEpiLanes - 1) / kEpiLanes),
                       dim3(kEpiWaves * kEpiLanes),
                       lds, st, ch, ea);
  else
    hipLaunchKernelGGL(topo_epilogue_kernel<DipoleSrc<true>>, dim3((L + kEpiLanes - 1) / kEpiLanes),
                       dim3(kEpiWaves * kEpiLanes),
                       lds, st, ch, ea);
  CPET_HIP_TRY(hipGetLastError());
  if ((rc = prof_mark(ctx, 3))) return rc;

  if (!dev) {
    CPET_HIP_TRY(hipMemcpyAsync(feats, d_feats, (size_t)L * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (steps) CPET_HIP_TRY(hipMemcpyAsync(steps, d_steps, (size_t)L * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (exited) CPET_HIP_TRY(hipMemcpyAsync(exited, d_exited, (size_t)L * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (hist) CPET_HIP_TRY(hipMemcpyAsync(hist, d_hist, (size_t)nbins * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  }
  if (!dev || !(flags & CPET_F_NO_SYNC)) {
    CPET_HIP_TRY(hipStreamSynchronize(st));
    if ((rc = prof_collect(ctx, 4))) return rc;
  }
  return CPET_OK;
}

// Many frames in one persistent launch (see topo_batch_kernel).
int topo_batch_run(cpet_ctx* ctx, int F, const int32_t* n_starts, const float* starts,
                   const int32_t* n_iter, const float* h, const float* dims, const float4* ch,
                   const int32_t* d_chg_off, const std::vector<int32_t>& chg_off, double* feats,
                   int32_t* steps, int32_t* exited, int nb_d, const double* edges_d, int nb_c,
                   const double* edges_c, uint32_t* hist, uint32_t flags) {
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  hipStream_t st = ctx->stream;
  int rc;
  std::vector<int32_t> sl_off(F + 1, 0);
  for (int f = 0; f < F; ++f) {
    if (n_starts[f] <= 0) {
      set_error("frame %d has no streamline starts", f);
      return CPET_ERR_ARG;
    }
    if (!(std::isfinite(h[f]) && h[f] > 0.f) || !(dims[3 * f] > 0.f && dims[3 * f + 1] > 0.f &&
                                                  dims[3 * f + 2] > 0.f)) {
      set_error("frame %d: step size and box half-widths must be finite and > 0", f);
      return CPET_ERR_ARG;
    }
    const long long next = (long long)sl_off[f] + n_starts[f];
    if (next > 0x7fffffffLL) {
      set_error("batch too large (%lld streamlines)", next);
      return CPET_ERR_ARG;
    }
    sl_off[f + 1] = (int32_t)next;
  }
  const int Lt = sl_off[F];
  const void* dstarts = nullptr;
  const void* diter = nullptr;
  if ((rc = stage_in(ctx, kSlotPts, starts, (size_t)Lt * 3 * sizeof(float), dev, &dstarts))) return rc;
  if ((rc = stage_in(ctx, kSlotIter, n_iter, (size_t)Lt * sizeof(int32_t), dev, &diter))) return rc;
  void* state = nullptr;
  // state: result records (Lt double4) | feats (Lt*2 double) | steps (Lt) | exited (Lt)
  if ((rc = ctx_buf(ctx, kSlotTopoState, (size_t)Lt * (6 * sizeof(double) + 2 * sizeof(int32_t)),
                    &state)))
    return rc;
  double4* d_res = (double4*)state;
  double* d_feats = dev ? feats : (double*)(d_res + Lt);
  int32_t* d_steps =
      (dev && steps) ? steps : (int32_t*)((double*)(d_res + Lt) + 2 * (size_t)Lt);
  int32_t* d_exited =
      (dev && exited) ? exited : (int32_t*)((double*)(d_res + Lt) + 2 * (size_t)Lt) + Lt;

  // 16 waves per workgroup, as the single-frame kernel; when every frame has
  // at most kVirtMaxCharges charges (padded), one-wave workgroups walking the
  // 16 slices in order (the same results, see kVirtWaves), 24 per CU (their
  // LDS allows about that many): 4.9x faster on 16-charge frames, 1.7x at
  // 64, 1.1x at 128, slower from 256 on (200 to 6 frames of 20,000 starts).
  int maxpad = 0;
  for (int f = 0; f < F; ++f) maxpad = std::max(maxpad, chg_off[f + 1] - chg_off[f]);
  const int S = ctx->tune[3][0] > 0 ? ctx->tune[3][0] : (maxpad <= kVirtMaxCharges ? 1 : 16);
  const int G0 = ctx->tune[3][1] > 0 ? ctx->tune[3][1]
               : S == 1            ? ctx->num_cus * 24
                                   : ctx->num_cus * (S >= 16 ? 1 : 2);
  // Assignments: frames by decreasing work, one entry per A streamlines.
  // A workgroup that takes an assignment runs its 64 slots on that frame
  // until the frame's queue is empty, then drains them; larger assignments
  // put fewer workgroups on a frame, each running more streamlines per slot,
  // so the drain is a smaller share of its work.  A grows from 64 to 256
  // while the batch still gives every workgroup four or more assignments.
  const int A = 64 * std::max(1, std::min(4, (int)(Lt / ((long long)G0 * 64 * 4))));
  std::vector<int> byc(F);
  for (int f = 0; f < F; ++f) byc[f] = f;
  std::sort(byc.begin(), byc.end(), [&](int a, int b) {
    return (double)n_starts[a] * (chg_off[a + 1] - chg_off[a]) >
           (double)n_starts[b] * (chg_off[b + 1] - chg_off[b]);
  });
  std::vector<int32_t> assign;
  for (int f : byc)
    for (int r = 0; r < (n_starts[f] + A - 1) / A; ++r) assign.push_back(f);
  std::vector<int32_t> blk;
  for (int f = 0; f < F; ++f)
    for (int b = 0; b < (n_starts[f] + kEpiLanes - 1) / kEpiLanes; ++b) {
      blk.push_back(f);
      blk.push_back(b);
    }
  std::vector<double> hd(4 * (size_t)F);
  for (int f = 0; f < F; ++f) {
    hd[4 * f] = (double)h[f];
    hd[4 * f + 1] = dims[3 * f];
    hd[4 * f + 2] = dims[3 * f + 1];
    hd[4 * f + 3] = dims[3 * f + 2];
  }
  const int nbins = hist ? nb_d * nb_c : 0;
  // read-only tables (cached uploads) in their own buffer ...
  const size_t o_hd = 0;
  const size_t o_sl = o_hd + hd.size() * sizeof(double);
  const size_t o_as = (o_sl + sl_off.size() * sizeof(int32_t) + 255) & ~(size_t)255;
  const size_t o_blk = (o_as + assign.size() * sizeof(int32_t) + 255) & ~(size_t)255;
  const size_t t_total = o_blk + blk.size() * sizeof(int32_t) + 256;
  void* tabs = nullptr;
  if ((rc = ctx_buf(ctx, kSlotTopoTables, t_total, &tabs))) return rc;
  char* tw = (char*)tabs;
  if ((rc = upload_table(ctx, kTabTopoHd, tw + o_hd, hd.data(), hd.size() * sizeof(double)))) return rc;
  if ((rc = upload_table(ctx, kTabTopoSl, tw + o_sl, sl_off.data(), sl_off.size() * sizeof(int32_t)))) return rc;
  if ((rc = upload_table(ctx, kTabTopoAs, tw + o_as, assign.data(), assign.size() * sizeof(int32_t)))) return rc;
  if ((rc = upload_table(ctx, kTabTopoBlk, tw + o_blk, blk.data(), blk.size() * sizeof(int32_t)))) return rc;
  // ... and the counters, histograms and edges the call writes
  const size_t o_cnt = 0;
  const size_t o_hist = ((F + 2) * sizeof(uint32_t) + 255) & ~(size_t)255;
  const size_t o_edges = (o_hist + (size_t)F * nbins * sizeof(uint32_t) + 255) & ~(size_t)255;
  const size_t o_bk = (o_edges + (hist ? (size_t)(nb_d + nb_c + 2) * sizeof(double) : 0) + 255) &
                      ~(size_t)255;
  const size_t o_ord = (o_bk + (size_t)F * kBatchBuckets * sizeof(uint32_t) + 255) & ~(size_t)255;
  const size_t total = o_ord + (size_t)Lt * sizeof(int32_t) + 256;
  void* work = nullptr;
  if ((rc = ctx_buf(ctx, kSlotBatchWork, total, &work))) return rc;
  char* w = (char*)work;
  unsigned int* d_cnt = (unsigned int*)(w + o_cnt);
  hipLaunchKernelGGL(fill_u32_kernel, dim3((F + 2 + 255) / 256), dim3(256), 0, st, d_cnt,
                     (long long)(F + 2), 0u);
  unsigned int* d_hist = nullptr;
  const double *d_ed = nullptr, *d_ec = nullptr;
  if (hist) {
    d_hist = dev ? (unsigned int*)hist : (unsigned int*)(w + o_hist);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(((long long)F * nbins + 255) / 256), dim3(256), 0,
                       st, d_hist, (long long)F * nbins, 0u);
    if (dev) {
      d_ed = edges_d;
      d_ec = edges_c;
    } else {
      double* e = (double*)(w + o_edges);
      CPET_HIP_TRY(hipMemcpyAsync(e, edges_d, (nb_d + 1) * sizeof(double), hipMemcpyHostToDevice, st));
      CPET_HIP_TRY(hipMemcpyAsync(e + nb_d + 1, edges_c, (nb_c + 1) * sizeof(double), hipMemcpyHostToDevice, st));
      d_ed = e;
      d_ec = e + nb_d + 1;
    }
  }

  // per-frame longest-first order
  unsigned int* d_bk = (unsigned int*)(w + o_bk);
  int32_t* d_ord = (int32_t*)(w + o_ord);
  const long long nbk = (long long)F * kBatchBuckets;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((nbk + 255) / 256)), dim3(256), 0, st, d_bk,
                     nbk, 0u);
  hipLaunchKernelGGL(batch_bucket_count_kernel, dim3((Lt + 255) / 256), dim3(256), 0, st, Lt, F,
                     (const int32_t*)(tw + o_sl), (const int32_t*)diter, d_bk);
  hipLaunchKernelGGL(scan_u32_kernel, dim3(1), dim3(1024), 0, st, d_bk, (int)nbk);
  hipLaunchKernelGGL(batch_bucket_scatter_kernel, dim3((Lt + 255) / 256), dim3(256), 0, st, Lt,
                     F, (const int32_t*)(tw + o_sl), (const int32_t*)diter, d_bk, d_ord);
  CPET_HIP_TRY(hipGetLastError());

  // far-field records of every frame and the per-frame safety flags
  const float4* d_far = nullptr;
  const int* d_far_unsafe = nullptr;
  if (S >= 4 && ctx->tune[5][0] == 2) {
    const int total = chg_off[F];
    void* fb = nullptr;
    if ((rc = ctx_buf(ctx, kSlotFar, (size_t)total * 2 * sizeof(float4) + (size_t)F * 4 + 256,
                      &fb)))
      return rc;
    int* flags_d = (int*)((char*)fb + (size_t)total * 2 * sizeof(float4));
    hipLaunchKernelGGL(fill_u32_kernel, dim3((F + 255) / 256), dim3(256), 0, st,
                       (unsigned int*)flags_d, (long long)F, 0u);
    hipLaunchKernelGGL(far_pack_kernel, dim3((total + 255) / 256), dim3(256), 0, st, total, ch, F,
                       d_chg_off, (const double*)(tw + o_hd), make_double4(0.0, 0.0, 0.0, 0.0),
                       (float4*)fb, flags_d);
    CPET_HIP_TRY(hipGetLastError());
    d_far = (const float4*)fb;
    d_far_unsafe = flags_d;
  }

  const int G = std::max(1, std::min<int>(G0, (int)assign.size()));
  BatchTopoArgs ba;
  ba.F = F;
  ba.chg_off = d_chg_off;
  ba.sl_off = (const int32_t*)(tw + o_sl);
  ba.starts = (const float*)dstarts;
  ba.n_iter = (const int32_t*)diter;
  ba.hdims = (const double*)(tw + o_hd);
  ba.res = d_res;
  ba.sl_counter = d_cnt;
  ba.order = d_ord;
  ba.assign = (const int32_t*)(tw + o_as);
  ba.n_assign = (int)assign.size();
  ba.assign_counter = d_cnt + F;
  ba.far_unsafe = d_far_unsafe;
  if ((rc = prof_mark(ctx, 0))) return rc;
  if (d_far) {  // the far-form frames (none when every flag is set)
    switch (S) {
      case 4: hipLaunchKernelGGL((topo_batch_far_kernel<4>), dim3(G), dim3(256), 0, st, d_far, ba); break;
      case 8: hipLaunchKernelGGL((topo_batch_far_kernel<8>), dim3(G), dim3(512), 0, st, d_far, ba); break;
      case 16: hipLaunchKernelGGL((topo_batch_far_kernel<16>), dim3(G), dim3(1024), 0, st, d_far, ba); break;
      default: break;
    }
    CPET_HIP_TRY(hipGetLastError());
  }
  switch (S) {
    case 1: hipLaunchKernelGGL((topo_batch_kernel<1>), dim3(G), dim3(64), 0, st, ch, ba); break;
    case 2: hipLaunchKernelGGL((topo_batch_kernel<2>), dim3(G), dim3(128), 0, st, ch, ba); break;
    case 4: hipLaunchKernelGGL((topo_batch_kernel<4>), dim3(G), dim3(256), 0, st, ch, ba); break;
    case 8: hipLaunchKernelGGL((topo_batch_kernel<8>), dim3(G), dim3(512), 0, st, ch, ba); break;
    case 16: hipLaunchKernelGGL((topo_batch_kernel<16>), dim3(G), dim3(1024), 0, st, ch, ba); break;
    default:
      set_error("unsupported waves per workgroup %d (1,2,4,8,16)", S);
      return CPET_ERR_ARG;
  }
  CPET_HIP_TRY(hipGetLastError());
  if ((rc = prof_mark(ctx, 1))) return rc;
  EpiArgs ea;
  ea.L = 0;
  ea.npad = 0;
  ea.starts = (const float*)dstarts;
  ea.res = d_res;
  ea.h = 0.0;
  ea.feats = d_feats;
  ea.steps = d_steps;
  ea.exited = d_exited;
  ea.ed = d_ed;
  ea.nbd = nb_d;
  ea.ec = d_ec;
  ea.nbc = nb_c;
  ea.hist = d_hist;
  ea.lds_hist = (hist && nbins <= 8192) ? 1 : 0;
  const size_t lds = ea.lds_hist ? (size_t)nbins * sizeof(unsigned int) : 0;
  hipLaunchKernelGGL(topo_batch_epilogue_kernel, dim3((unsigned)(blk.size() / 2)),
                     dim3(kEpiWaves * kEpiLanes), lds,
                     st, ch, ea, (const int32_t*)(tw + o_blk), d_chg_off,
                     (const int32_t*)(tw + o_sl), (const double*)(tw + o_hd));
  CPET_HIP_TRY(hipGetLastError());
  if ((rc = prof_mark(ctx, 2))) return rc;
  if (!dev) {
    CPET_HIP_TRY(hipMemcpyAsync(feats, d_feats, (size_t)Lt * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    if (steps) CPET_HIP_TRY(hipMemcpyAsync(steps, d_steps, (size_t)Lt * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (exited) CPET_HIP_TRY(hipMemcpyAsync(exited, d_exited, (size_t)Lt * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (hist) CPET_HIP_TRY(hipMemcpyAsync(hist, d_hist, (size_t)F * nbins * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  }
  if (!dev || !(flags & CPET_F_NO_SYNC)) {
    CPET_HIP_TRY(hipStreamSynchronize(st));
    if ((rc = prof_collect(ctx, 3))) return rc;
  }
  return CPET_OK;
}

}  // namespace cpet
