// Direct Coulomb field / electrostatic potential kernels for gfx950.
//
// Hot loop replaced: compute_looped_field (CPET/utils/math_module.c:405-451,
// E with an r^2 >= 1e-6 floor), calc_esp_base (:453-486, phi, no floor) and
// calc_field (:255-293, E, no floor).  The reference runs one FP32 point at
// a time on one CPU thread; here one GPU lane owns K points, every lane of a
// wave walks the same charge, and charges are read through the scalar unit
// (wave-uniform addresses -> s_load), so the vector ALUs only do pair math.
//
// Grid kernel (tensor-product mesh): a lane owns K consecutive z-points of
// one (x, y) line, so dx, dy and dx^2 + dy^2 are computed once per charge:
//   per pair (FP32 E):  dz, r2 = dz*dz + dxy2, rinv = rsq(r2),
//                       min(rinv, rsq(1e-6)), s = (q*re)*(re*re), 3 FMA
// Points kernel (arbitrary points): 3 subtracts + 3 ops for r2 instead.
//
// Precision is chosen per quantity (flags bits 8..10):
//   auto   E: FP32 pair terms, FP32 sum per 64-charge tile, FP64 across tiles,
//          plus A = sum |q| / r^2 (the sum of the term magnitudes).  A point
//          with u A > kAutoTol |E| (u = 2^-24; strong cancellation) is
//          appended to a list and recomputed in the "mixed" scheme by
//          field_fix_kernel.  The worst-case FP32 error is about 140 u A
//          (per term ~18 u through d, r^2, rsq and q r^-3, plus up to 63 u
//          of the tile sum, per component), far above what occurs: the
//          packed charges are shuffled (pack_charges), so a tile's terms are
//          a sample of the frame, and the error measured on realistic and
//          adversarial layouts (clustered frames, rock-salt lattices, shells,
//          same-sign clusters of charges, up to 16 coincident copies, on
//          both sides of a field zero) stays below ~2.5 u A at the accepted
//          points, so kAutoTol = 4e-6 keeps them within 1e-5 |E| (emulated
//          worst case 8.5e-6).  Many exactly coincident charges (64 copies
//          of one charge) give every copy the same rounding, up to ~4 u A
//          in all: merge coincident charges (sum their q) before the call,
//          or use the "mixed" mode, when a frame has them.
//          phi: "mixed" — FP64 displacement and r^2 (exact for FP32 inputs),
//          an FP32 rsq seed refined by one FP64 Newton step, FP64 sums.
//          (phi cancels strongly between charges of both signs, so FP32
//          pair rounding alone exceeds a 1e-5 pointwise bound near phi ~ 0.)
//   tiled / kahan / naive / f64 / mixed: the same scheme for both, no check.
// Small grids split the charges over blockIdx.y and reduce the FP64 partials
// in a fixed order in a second kernel.

#include <cmath>
#include <cstdio>
#include <vector>

#include "cpet_internal.h"

namespace cpet {
namespace {

// per-quantity precision
enum QPrec : int { kQNaive = 0, kQKahan = 1, kQTiled = 2, kQF64 = 3, kQMixed = 4 };

__host__ __device__ constexpr int qprec_e(int mode) {
  return mode == kModeNaive ? kQNaive
       : mode == kModeKahan ? kQKahan
       : mode == kModeF64   ? kQF64
       : mode == kModeMixed ? kQMixed
                            : kQTiled;  // auto, tiled
}
__host__ __device__ constexpr int qprec_p(int mode) {
  return mode == kModeNaive ? kQNaive
       : mode == kModeKahan ? kQKahan
       : mode == kModeF64   ? kQF64
       : mode == kModeTiled ? kQTiled
                            : kQMixed;  // auto, mixed
}
__host__ __device__ constexpr bool q_is32(int q) { return q <= kQTiled; }

__device__ __forceinline__ double rsqrt_f64(double x) {
  // Hardware estimate refined by two Newton-Raphson steps for 1/sqrt(x):
  // y <- y * (1.5 - 0.5 x y^2).  x == 0 keeps +inf like 1/sqrt(0).
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double y1 = y * fma(-hx * y, y, 1.5);
  y1 = y1 * fma(-hx * y1, y1, 1.5);
  return (x == 0.0) ? y : y1;
}

// One Newton-Raphson step in FP64 from an FP32 estimate of 1/sqrt(x):
// relative error ~1e-7 -> ~1e-14.  Clamping y*y keeps x == 0 (y = +inf)
// from producing 0*inf: the residual is then 1 and the result stays +inf.
__device__ __forceinline__ double rsqrt_refine(float seed, double x) {
  const double y = (double)seed;
  const double e = fma(-x, fmin(y * y, 1.7976931348623157e308), 1.0);
  return fma(0.5 * y, e, y);
}

// The same step for x > 0 (no charge on the point): one clamp fewer.
__device__ __forceinline__ double rsqrt_refine_pos(float seed, double x) {
  const double y = (double)seed;
  const double e = fma(-x, y * y, 1.0);
  return fma(0.5 * y, e, y);
}

// Relative error target of the auto-mode check (see above): points whose
// estimate u * A / |E| exceeds it are recomputed in FP64.
constexpr double kAutoTol = 4e-6;
constexpr double kUnit32 = 5.9604644775390625e-08;  // 2^-24

// Slots of the per-point FP64 accumulator: E x, y, z, phi, A (auto E only).
constexpr int kAcc = 5;

__device__ __forceinline__ void kahan_add(float t, float& s, float& c) {
  const float y = t - c;
  const float u = s + y;
  c = (u - s) - y;
  s = u;
}

// Accumulates the unscaled sums (sum q d / r^3, sum q / r and, for auto E,
// sum |q| / r^2) of charges [t0*kTile, t1*kTile) at K points into
// out[k][0..4] (FP64).  SHARED_XY: all K points share px[0], py[0] (grid
// z-line).  APART: no charge lies within 1e-3 A of a point (checked before
// the launch), so r^2 > 0 and the FP64 refinement needs no clamp for 1/0.
template <int K, bool SHARED_XY, int MODE, int WHAT, bool APART = false>
__device__ __forceinline__ void accumulate(
    const float (&px)[K], const float (&py)[K], const float (&pz)[K],
    const float4* __restrict__ ch, const double4* __restrict__ chd, int t0,
    int t1, double (&out)[K][kAcc]) {
  constexpr bool hasE = WHAT != kWhatPhi;
  constexpr bool hasP = WHAT == kWhatPhi || WHAT == kWhatESoftPhi;
  constexpr bool soft = WHAT == kWhatESoft || WHAT == kWhatESoftPhi;
  constexpr int QE = qprec_e(MODE), QP = qprec_p(MODE);
  constexpr bool e32 = hasE && q_is32(QE), p32 = hasP && q_is32(QP);
  constexpr bool e64 = hasE && !q_is32(QE), p64 = hasP && !q_is32(QP);
  constexpr bool need32 = e32 || p32;
  constexpr bool need64 = e64 || p64;
  constexpr bool full64 = MODE == kModeF64;
  constexpr bool naive = MODE == kModeNaive;
  constexpr bool kahan = MODE == kModeKahan;
  constexpr bool track = MODE == kModeAuto && e32;  // magnitude sum for the check
  constexpr int kGroup = (need64 || kahan) ? 4 : 8;

  const float rclamp = __builtin_amdgcn_rsqf(1e-6f);
  const double rclampd = rsqrt_f64((double)1e-6f);
  double pxd[K], pyd[K], pzd[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    pxd[k] = (double)px[SHARED_XY ? 0 : k];
    pyd[k] = (double)py[SHARED_XY ? 0 : k];
    pzd[k] = (double)pz[k];
#pragma unroll
    for (int c = 0; c < kAcc; ++c) out[k][c] = 0.0;
  }
  float na[K][4];  // naive: FP32 sums over the whole charge range
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) na[k][c] = 0.f;

  for (int t = t0; t < t1; ++t) {
    float ta[K][4], tc[K][4];  // per-tile FP32 sums (+ Kahan compensation)
    float tm[K];               // per-tile sum |q| / r^2 (track only)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      tm[k] = 0.f;
#pragma unroll
      for (int c = 0; c < 4; ++c) ta[k][c] = tc[k][c] = 0.f;
    }
    const int jb = t * kTile;
    // Charges are fetched kGroup at a time (the scalar loads of a group are
    // issued together), so the scalar-cache latency is paid once per group.
    for (int g = 0; g < kTile; g += kGroup) {
      float4 cg[kGroup];
      double4 cdg[need64 ? kGroup : 1];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        cg[u] = ch[jb + g + u];
        if constexpr (need64) cdg[u] = chd[jb + g + u];
      }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const float4 c = cg[u];
      float sdx = 0.f, sdy = 0.f, sdxy2 = 0.f;
      double sdxd = 0.0, sdyd = 0.0, sdxy2d = 0.0;
      double4 cd = make_double4(0.0, 0.0, 0.0, 0.0);
      if constexpr (need64) cd = cdg[u];
      if constexpr (SHARED_XY && need32) {
        sdx = px[0] - c.x;
        sdy = py[0] - c.y;
        sdxy2 = fmaf(sdy, sdy, sdx * sdx);
      }
      if constexpr (SHARED_XY && need64) {
        sdxd = pxd[0] - cd.x;
        sdyd = pyd[0] - cd.y;
        sdxy2d = fma(sdyd, sdyd, sdxd * sdxd);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        float ri = 0.f;
        if constexpr (need32) {
          const float dx = SHARED_XY ? sdx : px[k] - c.x;
          const float dy = SHARED_XY ? sdy : py[k] - c.y;
          const float dxy2 = SHARED_XY ? sdxy2 : fmaf(dy, dy, dx * dx);
          const float dz = pz[k] - c.z;
          const float r2 = fmaf(dz, dz, dxy2);
          ri = __builtin_amdgcn_rsqf(r2);
          float* acc = naive ? na[k] : ta[k];
          if constexpr (p32) {
            if constexpr (kahan)
              kahan_add(c.w * ri, acc[3], tc[k][3]);
            else
              acc[3] = fmaf(c.w, ri, acc[3]);
          }
          if constexpr (e32) {
            const float re = soft ? fminf(ri, rclamp) : ri;
            const float re2 = re * re;
            const float s = (c.w * re) * re2;
            if constexpr (track) tm[k] = fmaf(fabsf(c.w), re2, tm[k]);
            if constexpr (kahan) {
              kahan_add(s * dx, acc[0], tc[k][0]);
              kahan_add(s * dy, acc[1], tc[k][1]);
              kahan_add(s * dz, acc[2], tc[k][2]);
            } else {
              acc[0] = fmaf(s, dx, acc[0]);
              acc[1] = fmaf(s, dy, acc[1]);
              acc[2] = fmaf(s, dz, acc[2]);
            }
          }
        }
        if constexpr (need64) {
          const double dx = SHARED_XY ? sdxd : pxd[k] - cd.x;
          const double dy = SHARED_XY ? sdyd : pyd[k] - cd.y;
          const double dxy2 = SHARED_XY ? sdxy2d : fma(dy, dy, dx * dx);
          const double dz = pzd[k] - cd.z;
          const double r2 = fma(dz, dz, dxy2);
          double y;
          if constexpr (full64)
            y = rsqrt_f64(r2);
          else
            y = APART ? rsqrt_refine_pos(need32 ? ri : __builtin_amdgcn_rsqf((float)r2), r2)
                      : rsqrt_refine(need32 ? ri : __builtin_amdgcn_rsqf((float)r2), r2);
          if constexpr (p64) out[k][3] = fma(cd.w, y, out[k][3]);
          if constexpr (e64) {
            const double re = soft ? fmin(y, rclampd) : y;
            const double s = (cd.w * re) * (re * re);
            out[k][0] = fma(s, dx, out[k][0]);
            out[k][1] = fma(s, dy, out[k][1]);
            out[k][2] = fma(s, dz, out[k][2]);
          }
        }
      }
    }
    }
    if constexpr (!naive && need32) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if constexpr (e32) {
#pragma unroll
          for (int c = 0; c < 3; ++c)
            out[k][c] += kahan ? (double)ta[k][c] - (double)tc[k][c] : (double)ta[k][c];
        }
        if constexpr (p32)
          out[k][3] += kahan ? (double)ta[k][3] - (double)tc[k][3] : (double)ta[k][3];
        if constexpr (track) out[k][4] += (double)tm[k];
      }
    }
  }
  if constexpr (naive) {
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c) out[k][c] += (double)na[k][c];
  }
}

struct FieldOut {
  void* e;         // (M,3) float or double, or null
  void* phi;       // (M) float or double, or null
  double* part;    // (splits, M, kAcc) partial sums when splits > 1, else null
  int f64;         // outputs are double
  long long M;
  int* fix_count;  // auto-mode check: number of listed points, or null
  long long* fix_list;  // output indices of points to recompute
  const int* near;  // softened E: 0 when no charge is near a grid point, or null
};

// The r^2 >= 1e-6 floor of the volume field can only bind for a charge
// within 1e-3 A of a grid point.  kNearGrid adds 1 % for the FP32 rounding
// of p - x (|p| up to ~1e3 A), so below it the floored and unfloored sums
// are the same numbers.
constexpr double kNearGrid = 1.01e-3;

__device__ __forceinline__ double nearest_sq(const float* __restrict__ a, int n, float v) {
  double best = 1e300;
  for (int i = 0; i < n; ++i) {
    const double d = (double)a[i] - (double)v;
    best = fmin(best, d * d);
  }
  return best;
}

// *flag = 1 when some charge lies within kNearGrid of a grid point; the
// caller clears it first (near_clear_kernel).  One charge per thread with
// linear scans over the axes, so they need not be sorted; the nearest grid
// point of a rectilinear grid is the product of the nearest axis values.
// Every writer stores the same 1, so the stores need no atomics.
__global__ __launch_bounds__(256) void near_grid_kernel(int n, const float4* __restrict__ ch,
                                                        int nx, const float* __restrict__ ax,
                                                        int ny, const float* __restrict__ ay,
                                                        int nz, const float* __restrict__ az,
                                                        int* flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const float4 c = ch[j];
  const double d2 = nearest_sq(ax, nx, c.x) + nearest_sq(ay, ny, c.y) + nearest_sq(az, nz, c.z);
  if (d2 < kNearGrid * kNearGrid) *flag = 1;
}

__global__ void near_clear_kernel(int* flag) { *flag = 0; }

__global__ __launch_bounds__(256) void near_clear_n_kernel(int* flags, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = 0;
}

// Batches: flags[f] = 1 when a charge of frame f lies within kNearGrid of a
// point of frame f's grid.  chg_off (F+1) holds the padded charge offsets,
// geo (F,4) nx, ny, nz, nzc and ax_off (F) the start of each frame's axes.
__global__ __launch_bounds__(256) void near_batch_kernel(int n, const float4* __restrict__ ch,
                                                         int F, const int32_t* __restrict__ chg_off,
                                                         const int32_t* __restrict__ geo,
                                                         const int32_t* __restrict__ ax_off,
                                                         const float* __restrict__ axes,
                                                         int* flags) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  int lo = 0, hi = F - 1;  // the last frame f with chg_off[f] <= j
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chg_off[mid] <= j)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int nx = geo[4 * lo], ny = geo[4 * lo + 1], nz = geo[4 * lo + 2];
  const float* ax = axes + ax_off[lo];
  const float4 c = ch[j];
  const double d2 = nearest_sq(ax, nx, c.x) + nearest_sq(ax + nx, ny, c.y) +
                    nearest_sq(ax + nx + ny, nz, c.z);
  if (d2 < kNearGrid * kNearGrid) flags[lo] = 1;
}

template <int WHAT>
__device__ __forceinline__ void store_point(const FieldOut& o, long long m,
                                            const double (&acc)[kAcc], int split) {
  if (o.part) {
    double* p = o.part + ((long long)split * o.M + m) * kAcc;
#pragma unroll
    for (int c = 0; c < kAcc; ++c) p[c] = acc[c];
    return;
  }
  const double k = kCoulomb;
  if (WHAT != kWhatPhi && o.e) {
    if (o.f64) {
      double* e = (double*)o.e + 3 * m;
      e[0] = k * acc[0];
      e[1] = k * acc[1];
      e[2] = k * acc[2];
    } else {
      float* e = (float*)o.e + 3 * m;
      e[0] = (float)(k * acc[0]);
      e[1] = (float)(k * acc[1]);
      e[2] = (float)(k * acc[2]);
    }
  }
  if ((WHAT == kWhatPhi || WHAT == kWhatESoftPhi) && o.phi) {
    if (o.f64)
      ((double*)o.phi)[m] = k * acc[3];
    else
      ((float*)o.phi)[m] = (float)(k * acc[3]);
  }
}

// True when the FP32 error estimate u * A exceeds kAutoTol * |E|.
__device__ __forceinline__ bool needs_fix(const double (&acc)[kAcc]) {
  const double a = acc[4] * (kUnit32 / kAutoTol);
  return a * a > fma(acc[2], acc[2], fma(acc[1], acc[1], acc[0] * acc[0]));
}

// Appends the output index g of every active lane with flag set to the fix
// list: one atomic per wave, slots in lane order.
__device__ __forceinline__ void append_fix(const FieldOut& o, long long g, bool flag) {
  const unsigned long long mask = __ballot(flag);
  if (mask == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(o.fix_count, __popcll(mask));
  base = __shfl(base, leader);
  if (flag) o.fix_list[base + __popcll(mask & ((1ull << lane) - 1ull))] = g;
}

// Grid kernel: lane -> (line (i,j), chunk of K z-points).
template <int K, int MODE, int WHAT>
__global__ __launch_bounds__(256) void field_grid_kernel(
    int nx, int ny, int nz, int nzc, const float* __restrict__ ax,
    const float* __restrict__ ay, const float* __restrict__ az,
    const float4* __restrict__ ch, const double4* __restrict__ chd,
    int tiles_per_split, int ntiles, FieldOut o) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long lanes = (long long)nx * ny * nzc;
  if (gid >= lanes) return;
  const long long line = gid / nzc;
  const int chunk = (int)(gid - line * nzc);
  const int i = (int)(line / ny);
  const int j = (int)(line - (long long)i * ny);
  const int z0 = chunk * K;
  float px[K], py[K], pz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    px[k] = ax[i];
    py[k] = ay[j];
    pz[k] = az[min(z0 + k, nz - 1)];
  }
  const int split = blockIdx.y;
  const int t0 = split * tiles_per_split;
  const int t1 = min(t0 + tiles_per_split, ntiles);
  double acc[K][kAcc];
  if constexpr (WHAT == kWhatESoft) {
    if (o.near && *o.near == 0)  // no charge near a grid point: the floor never binds
      accumulate<K, true, MODE, kWhatEHard>(px, py, pz, ch, chd, t0, t1, acc);
    else
      accumulate<K, true, MODE, WHAT>(px, py, pz, ch, chd, t0, t1, acc);
  } else if constexpr (WHAT == kWhatPhi && !q_is32(qprec_p(MODE)) && MODE != kModeF64) {
    if (o.near && *o.near == 0)  // r^2 >= 1e-6 at every pair: no 1/0 to keep
      accumulate<K, true, MODE, WHAT, true>(px, py, pz, ch, chd, t0, t1, acc);
    else
      accumulate<K, true, MODE, WHAT>(px, py, pz, ch, chd, t0, t1, acc);
  } else {
    accumulate<K, true, MODE, WHAT>(px, py, pz, ch, chd, t0, t1, acc);
  }
  const long long base = line * nz + z0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool valid = z0 + k < nz;
    if (valid) store_point<WHAT>(o, base + k, acc[k], split);
    if constexpr (MODE == kModeAuto && WHAT != kWhatPhi)
      if (o.fix_count && !o.part) append_fix(o, base + k, valid && needs_fix(acc[k]));
  }
}

// Points kernel: lane owns points base + k*blockDim + tid.
template <int K, int MODE, int WHAT>
__global__ __launch_bounds__(256) void field_points_kernel(
    long long M, const float* __restrict__ pts, const float4* __restrict__ ch,
    const double4* __restrict__ chd, int tiles_per_split, int ntiles,
    FieldOut o) {
  const long long base = (long long)blockIdx.x * blockDim.x * K + threadIdx.x;
  float px[K], py[K], pz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    long long m = base + (long long)k * blockDim.x;
    m = m < M ? m : M - 1;
    px[k] = pts[3 * m];
    py[k] = pts[3 * m + 1];
    pz[k] = pts[3 * m + 2];
  }
  const int split = blockIdx.y;
  const int t0 = split * tiles_per_split;
  const int t1 = min(t0 + tiles_per_split, ntiles);
  double acc[K][kAcc];
  accumulate<K, false, MODE, WHAT>(px, py, pz, ch, chd, t0, t1, acc);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const long long m = base + (long long)k * blockDim.x;
    const bool valid = m < M;
    if (valid) store_point<WHAT>(o, m, acc[k], split);
    if constexpr (MODE == kModeAuto && WHAT != kWhatPhi)
      if (o.fix_count && !o.part) append_fix(o, m, valid && needs_fix(acc[k]));
  }
}

// Sums the per-split partials in split order and writes the outputs.
template <int WHAT>
__global__ __launch_bounds__(256) void field_reduce_kernel(long long M,
                                                           int splits,
                                                           FieldOut o) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double acc[kAcc] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < splits; ++s) {
    const double* p = o.part + ((long long)s * M + m) * kAcc;
#pragma unroll
    for (int c = 0; c < kAcc; ++c) acc[c] += p[c];
  }
  FieldOut d = o;
  d.part = nullptr;
  store_point<WHAT>(d, m, acc, 0);
  if (WHAT != kWhatPhi && o.fix_count) append_fix(o, m, needs_fix(acc));
}

__global__ void fix_reset_kernel(int* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0;
}

// Recomputes E at the listed points in the "mixed" scheme (FP64
// displacement and r^2, FP32 rsq seed + one FP64 Newton step, FP64 sums):
// one wave per point, lanes stride over the charges, then a fixed-order
// butterfly reduction.  Single frame (F == 0: grid axes or points) or a
// batch (F > 0: the batch tables; the frame of output index g is the last
// one whose first output index is <= g).
struct FixArgs {
  const int* count;
  const long long* list;
  const float4* ch;
  const float *ax, *ay, *az, *pts;
  int nx, ny, nz, npad;
  int F;
  const int32_t* geo;
  const int64_t* out_off;
  const int32_t* ax_off;
  const float* axes;
  const int32_t* chg_off;
  void* e;
  int f64;
};

// Grid point, charge range [c0, c1) of the listed output index g.
__device__ __forceinline__ void fix_point(const FixArgs& a, long long g, double& px, double& py,
                                          double& pz, int& c0, int& c1) {
  c0 = 0;
  c1 = a.npad;
  if (a.F > 0) {
    int lo = 0, hi = a.F - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.out_off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const long long m = g - a.out_off[lo];
    const int nx = a.geo[4 * lo], ny = a.geo[4 * lo + 1], nz = a.geo[4 * lo + 2];
    const long long line = m / nz;
    const int z = (int)(m - line * nz), i = (int)(line / ny), j = (int)(line - (long long)i * ny);
    const float* ax = a.axes + a.ax_off[lo];
    px = ax[i];
    py = ax[nx + j];
    pz = ax[nx + ny + z];
    c0 = a.chg_off[lo];
    c1 = a.chg_off[lo + 1];
  } else if (a.pts) {
    px = a.pts[3 * g];
    py = a.pts[3 * g + 1];
    pz = a.pts[3 * g + 2];
  } else {
    const long long line = g / a.nz;
    const int z = (int)(g - line * a.nz), i = (int)(line / a.ny),
              j = (int)(line - (long long)i * a.ny);
    px = a.ax[i];
    py = a.ay[j];
    pz = a.az[z];
  }
}

// One wave refines kFixPts listed points at once: lanes split the charges,
// every charge read from L2 serves all kFixPts points (the kernel is bound
// by that traffic: each point reads every charge), and the wave's partial
// sums are reduced across lanes in a fixed order.  Points of one group that
// lie in different frames of a batch (different charge ranges) are refined
// one at a time.
constexpr int kFixPts = 4;

template <bool SOFT>
__global__ __launch_bounds__(256) void field_fix_kernel(FixArgs a) {
  const int n = *a.count;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double rclampd = rsqrt_f64((double)1e-6f);
  auto pair = [&](const float4& c, double px, double py, double pz, double (&e)[3]) {
    const double dx = px - (double)c.x, dy = py - (double)c.y, dz = pz - (double)c.z;
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    const double y = rsqrt_refine(__builtin_amdgcn_rsqf((float)r2), r2);
    const double re = SOFT ? fmin(y, rclampd) : y;
    const double s = ((double)c.w * re) * (re * re);
    e[0] = fma(s, dx, e[0]);
    e[1] = fma(s, dy, e[1]);
    e[2] = fma(s, dz, e[2]);
  };
  for (long long t0 = ((long long)blockIdx.x * 4 + wave) * kFixPts; t0 < n;
       t0 += (long long)gridDim.x * 4 * kFixPts) {
    const int np = (int)min((long long)kFixPts, (long long)n - t0);
    long long g[kFixPts];
    double px[kFixPts], py[kFixPts], pz[kFixPts];
    int c0[kFixPts], c1[kFixPts];
    bool same = true;
#pragma unroll
    for (int p = 0; p < kFixPts; ++p) {
      g[p] = a.list[t0 + min(p, np - 1)];  // short groups repeat the last point
      fix_point(a, g[p], px[p], py[p], pz[p], c0[p], c1[p]);
      same = same && c0[p] == c0[0] && c1[p] == c1[0];
    }
    double e[kFixPts][3] = {};
    if (same) {
      for (int jq = c0[0] + lane; jq < c1[0]; jq += 64) {
        const float4 c = a.ch[jq];
#pragma unroll
        for (int p = 0; p < kFixPts; ++p) pair(c, px[p], py[p], pz[p], e[p]);
      }
    } else {
#pragma unroll
      for (int p = 0; p < kFixPts; ++p)
        for (int jq = c0[p] + lane; jq < c1[p]; jq += 64) pair(a.ch[jq], px[p], py[p], pz[p], e[p]);
    }
#pragma unroll
    for (int p = 0; p < kFixPts; ++p)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        e[p][0] += __shfl_xor(e[p][0], off);
        e[p][1] += __shfl_xor(e[p][1], off);
        e[p][2] += __shfl_xor(e[p][2], off);
      }
    if (lane < np) {  // lane p stores point p
      double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int p = 0; p < kFixPts; ++p)
        if (p == lane) {
          v[0] = e[p][0];
          v[1] = e[p][1];
          v[2] = e[p][2];
        }
      const long long gp = a.list[t0 + lane];
      const double k = kCoulomb;
      if (a.f64) {
        double* o = (double*)a.e + 3 * gp;
        o[0] = k * v[0];
        o[1] = k * v[1];
        o[2] = k * v[2];
      } else {
        float* o = (float*)a.e + 3 * gp;
        o[0] = (float)(k * v[0]);
        o[1] = (float)(k * v[1]);
        o[2] = (float)(k * v[2]);
      }
    }
  }
}

// Many grids in one launch: blk[2b] = frame, blk[2b+1] = block within the
// frame, so each workgroup reads one frame's charges (wave-uniform).
struct BatchFieldArgs {
  const int32_t* blk;      // (nblocks, 2)
  const int32_t* geo;      // (F, 4): nx, ny, nz, nzc
  const int64_t* out_off;  // (F) first output point of the frame
  const int32_t* ax_off;   // (F) offset of the frame's nx+ny+nz axis values
  const int32_t* chg_off;  // (F+1) padded charge offsets
  const float* axes;
  void* e;
  void* phi;
  int f64;
  int* fix_count;
  long long* fix_list;
  const int* near;  // (F) softened E: 0 when no charge is near the frame's grid, or null
};

template <int K, int MODE, int WHAT>
__global__ __launch_bounds__(256) void field_batch_kernel(const float4* __restrict__ ch,
                                                          const double4* __restrict__ chd,
                                                          BatchFieldArgs a) {
  const int fr = a.blk[2 * blockIdx.x], b = a.blk[2 * blockIdx.x + 1];
  const int nx = a.geo[4 * fr], ny = a.geo[4 * fr + 1], nz = a.geo[4 * fr + 2],
            nzc = a.geo[4 * fr + 3];
  const long long gid = (long long)b * blockDim.x + threadIdx.x;
  const long long lanes = (long long)nx * ny * nzc;
  if (gid >= lanes) return;
  const long long line = gid / nzc;
  const int chunk = (int)(gid - line * nzc);
  const int i = (int)(line / ny);
  const int j = (int)(line - (long long)i * ny);
  const int z0 = chunk * K;
  const float* ax = a.axes + a.ax_off[fr];
  float px[K], py[K], pz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    px[k] = ax[i];
    py[k] = ax[nx + j];
    pz[k] = ax[nx + ny + min(z0 + k, nz - 1)];
  }
  const int c0 = a.chg_off[fr];
  const int ntiles = (a.chg_off[fr + 1] - c0) / kTile;
  double acc[K][kAcc];
  const double4* cd = chd ? chd + c0 : nullptr;
  if constexpr (WHAT == kWhatESoft) {
    if (a.near && a.near[fr] == 0)  // no charge near this grid: the floor never binds
      accumulate<K, true, MODE, kWhatEHard>(px, py, pz, ch + c0, cd, 0, ntiles, acc);
    else
      accumulate<K, true, MODE, WHAT>(px, py, pz, ch + c0, cd, 0, ntiles, acc);
  } else {
    accumulate<K, true, MODE, WHAT>(px, py, pz, ch + c0, cd, 0, ntiles, acc);
  }
  FieldOut o;
  const size_t es = a.f64 ? sizeof(double) : sizeof(float);
  const long long off = a.out_off[fr];
  o.e = a.e ? (void*)((char*)a.e + (size_t)off * 3 * es) : nullptr;
  o.phi = a.phi ? (void*)((char*)a.phi + (size_t)off * es) : nullptr;
  o.part = nullptr;
  o.f64 = a.f64;
  o.M = (long long)nx * ny * nz;
  o.fix_count = a.fix_count;
  o.fix_list = a.fix_list;
  o.near = nullptr;
  const long long base = line * nz + z0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool valid = z0 + k < nz;
    if (valid) store_point<WHAT>(o, base + k, acc[k], 0);
    if constexpr (MODE == kModeAuto && WHAT != kWhatPhi)
      if (a.fix_count) append_fix(o, off + base + k, valid && needs_fix(acc[k]));
  }
}

template <int K, int MODE, int WHAT>
void launch_batch_t(hipStream_t st, int nblk, const float4* ch, const double4* chd,
                    const BatchFieldArgs& a) {
  hipLaunchKernelGGL((field_batch_kernel<K, MODE, WHAT>), dim3(nblk), dim3(256), 0, st, ch, chd,
                     a);
}

template <int WHAT>
bool launch_batch_w(int mode, int K, hipStream_t st, int nblk, const float4* ch,
                    const double4* chd, const BatchFieldArgs& a) {
#define CPET_BATCH_K(MODE_)                                                   \
  switch (K) {                                                                \
    case 1: launch_batch_t<1, MODE_, WHAT>(st, nblk, ch, chd, a); return true; \
    case 2: launch_batch_t<2, MODE_, WHAT>(st, nblk, ch, chd, a); return true; \
    case 4: launch_batch_t<4, MODE_, WHAT>(st, nblk, ch, chd, a); return true; \
    default: return false;                                                    \
  }
  switch (mode) {
    case kModeAuto: CPET_BATCH_K(kModeAuto)
    case kModeNaive: CPET_BATCH_K(kModeNaive)
    case kModeKahan: CPET_BATCH_K(kModeKahan)
    case kModeF64: CPET_BATCH_K(kModeF64)
    case kModeTiled: CPET_BATCH_K(kModeTiled)
    case kModeMixed: CPET_BATCH_K(kModeMixed)
  }
#undef CPET_BATCH_K
  return false;
}

// ---- dispatch ---------------------------------------------------------

struct Launch {
  hipStream_t st;
  dim3 grid;
  bool grid_mode;
  int nx, ny, nz, nzc;
  const float *ax, *ay, *az, *pts;
  long long M;
  const float4* ch;
  const double4* chd;
  int tps, ntiles;
  FieldOut o;
};

template <int K, int MODE, int WHAT>
void launch_t(const Launch& L) {
  if (L.grid_mode)
    hipLaunchKernelGGL((field_grid_kernel<K, MODE, WHAT>), L.grid, dim3(256), 0, L.st,
                       L.nx, L.ny, L.nz, L.nzc, L.ax, L.ay, L.az, L.ch, L.chd, L.tps,
                       L.ntiles, L.o);
  else
    hipLaunchKernelGGL((field_points_kernel<K, MODE, WHAT>), L.grid, dim3(256), 0, L.st,
                       L.M, L.pts, L.ch, L.chd, L.tps, L.ntiles, L.o);
}

// K choices: every listed K for the grid kernel in the default mode; 1, 2, 4
// otherwise (the arbitrary-points kernel and the study modes).
template <int MODE, int WHAT>
bool launch_k(int K, const Launch& L) {
  if (MODE == kModeAuto && L.grid_mode) {
    switch (K) {
      case 3: launch_t<3, MODE, WHAT>(L); return true;
      case 6: launch_t<6, MODE, WHAT>(L); return true;
      case 8: launch_t<8, MODE, WHAT>(L); return true;
      default: break;
    }
  }
  switch (K) {
    case 1: launch_t<1, MODE, WHAT>(L); return true;
    case 2: launch_t<2, MODE, WHAT>(L); return true;
    case 4: launch_t<4, MODE, WHAT>(L); return true;
    default: return false;
  }
}

template <int WHAT>
bool launch_m(int mode, int K, const Launch& L) {
  switch (mode) {
    case kModeAuto: return launch_k<kModeAuto, WHAT>(K, L);
    case kModeNaive: return launch_k<kModeNaive, WHAT>(K, L);
    case kModeKahan: return launch_k<kModeKahan, WHAT>(K, L);
    case kModeF64: return launch_k<kModeF64, WHAT>(K, L);
    case kModeTiled: return launch_k<kModeTiled, WHAT>(K, L);
    case kModeMixed: return launch_k<kModeMixed, WHAT>(K, L);
  }
  return false;
}

int what_of(bool want_e, bool want_phi, int soften) {
  if (want_e && want_phi) return soften ? kWhatESoftPhi : -1;
  if (want_phi) return kWhatPhi;
  return soften ? kWhatESoft : kWhatEHard;
}

// Number of charge splits so that a launch has about `target` waves.
int pick_splits(long long lanes, int ntiles, int num_cus, int forced) {
  if (forced > 0) return std::max(1, std::min(forced, ntiles));
  const long long waves = (lanes + 63) / 64;
  const long long target = (long long)num_cus * 16;  // 4 waves per SIMD
  if (waves >= target) return 1;
  long long s = (target + waves - 1) / waves;
  // keep at least 2 tiles per split so the partial traffic stays small
  s = std::min<long long>(s, std::max(1, ntiles / 2));
  return (int)std::max<long long>(1, s);
}

// Reserves the auto-mode check's counter (offset 0) and list (offset 256)
// for up to M points and queues the counter reset.
int fix_prepare(cpet_ctx* ctx, long long M, int** count, long long** list) {
  void* buf = nullptr;
  const int rc = ctx_buf(ctx, kSlotFix, 256 + (size_t)M * sizeof(long long), &buf);
  if (rc) return rc;
  *count = (int*)buf;
  *list = (long long*)((char*)buf + 256);
  hipLaunchKernelGGL(fix_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, *count);
  ctx->fix_count = *count;
  return CPET_OK;
}

void fix_launch(cpet_ctx* ctx, long long M, bool soft, const FixArgs& a) {
  const long long want = (M + 3) / 4;
  const int blocks = (int)std::max<long long>(1, std::min<long long>(want, (long long)ctx->num_cus * 8));
  if (soft)
    hipLaunchKernelGGL(field_fix_kernel<true>, dim3(blocks), dim3(256), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(field_fix_kernel<false>, dim3(blocks), dim3(256), 0, ctx->stream, a);
}

}  // namespace

bool field_mode_valid(int mode) { return mode >= kModeAuto && mode <= kModeMixed; }

// Shared implementation of the grid and points entry points.
int field_launch(cpet_ctx* ctx, bool grid_mode, int nx, int ny, int nz,
                 const float* ax, const float* ay, const float* az, long long M,
                 const float* pts, const float4* ch, const double4* chd, int npad,
                 void* E_out, void* phi_out, int soften, uint32_t flags,
                 int forced_k, int forced_splits) {
  const int what = what_of(E_out != nullptr, phi_out != nullptr, soften);
  if (what < 0) {
    set_error("E and phi together require soften=1 (volume semantics)");
    return CPET_ERR_ARG;
  }
  const int mode = (int)((flags & CPET_F_PREC_MASK) >> 8);
  if (!field_mode_valid(mode)) {
    set_error("unknown precision mode %d", mode);
    return CPET_ERR_ARG;
  }
  const int ntiles = npad / kTile;
  FieldOut o;
  o.e = E_out;
  o.phi = phi_out;
  o.part = nullptr;
  o.f64 = (flags & CPET_F_OUT_F64) ? 1 : 0;
  o.M = M;
  o.fix_count = nullptr;
  o.fix_list = nullptr;
  o.near = nullptr;
  ctx->fix_count = nullptr;

  int K;
  long long lanes;
  int nzc = 0;
  const bool heavy = mode != kModeAuto && mode != kModeTiled && mode != kModeNaive;
  if (grid_mode) {
    // K = 2 points per lane along z for E: measured fastest at every grid
    // size (21^3..101^3, 5k..30k charges, against K = 1, 3, 4, 6, 8), where
    // sharing dx, dy over more points saves less than the coarser waves
    // cost.  The potential alone shares four FP64 operations per charge
    // (dx, dy, dx^2 + dy^2) against ten per pair, and more points per lane
    // are faster there (101^3 x 30k: 10.9 / 10.2 / 9.9 ms at K = 2 / 4 / 8).
    K = forced_k > 0 ? forced_k : (what == kWhatPhi ? (nz >= 40 ? 8 : 4) : 2);
    nzc = (nz + K - 1) / K;
    lanes = (long long)nx * ny * nzc;
  } else {
    K = forced_k > 0 ? forced_k : (heavy ? 1 : 2);
    lanes = (M + K - 1) / K;
  }
  // Charge splits: at least enough waves to fill the GPU, and up to 16
  // splits of >= 4 tiles each, which measured faster at every grid size
  // (shorter, more numerous waves balance the SIMDs: 101^3 x 30k charges
  // 8.90 -> 8.48 ms, 41^3 x 20k 1.65 -> 0.45 ms); the partial sums
  // (splits x M x 40 B) stay under 1 GiB.
  int splits = pick_splits(lanes, ntiles, ctx->num_cus, forced_splits);
  if (forced_splits <= 0) {
    const long long cap = std::max(1LL, (1LL << 30) / std::max(1LL, M * (long long)(kAcc * sizeof(double))));
    splits = (int)std::min<long long>(std::max(splits, std::min(16, std::max(1, ntiles / 4))), cap);
  }
  const int tps = (ntiles + splits - 1) / splits;
  if (splits > 1) {
    void* part = nullptr;
    int rc = ctx_buf(ctx, kSlotPartial, (size_t)splits * M * kAcc * sizeof(double),
                     &part);
    if (rc) return rc;
    o.part = (double*)part;
  }
  const bool check = mode == kModeAuto && E_out != nullptr;
  if (check) {
    const int rc = fix_prepare(ctx, M, &o.fix_count, &o.fix_list);
    if (rc) return rc;
  }
  if (grid_mode && (what == kWhatESoft || what == kWhatPhi)) {
    void* nf = nullptr;
    const int rc = ctx_buf(ctx, kSlotNear, sizeof(int), &nf);
    if (rc) return rc;
    hipLaunchKernelGGL(near_clear_kernel, dim3(1), dim3(1), 0, ctx->stream, (int*)nf);
    hipLaunchKernelGGL(near_grid_kernel, dim3((npad + 255) / 256), dim3(256), 0, ctx->stream,
                       npad, ch, nx, ax, ny, ay, nz, az, (int*)nf);
    CPET_HIP_TRY(hipGetLastError());
    o.near = (const int*)nf;
  }
  const long long blocks_x = grid_mode ? (lanes + 255) / 256
                                       : (M + 256LL * K - 1) / (256LL * K);
  Launch L;
  L.st = ctx->stream;
  L.grid = dim3((unsigned)blocks_x, (unsigned)splits);
  L.grid_mode = grid_mode;
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.nzc = nzc;
  L.ax = ax;
  L.ay = ay;
  L.az = az;
  L.pts = pts;
  L.M = M;
  L.ch = ch;
  L.chd = chd;
  L.tps = tps;
  L.ntiles = ntiles;
  L.o = o;
  int prc = prof_mark(ctx, 0);
  if (prc) return prc;
  bool ok = false;
  switch (what) {
    case kWhatESoft: ok = launch_m<kWhatESoft>(mode, K, L); break;
    case kWhatPhi: ok = launch_m<kWhatPhi>(mode, K, L); break;
    case kWhatESoftPhi: ok = launch_m<kWhatESoftPhi>(mode, K, L); break;
    case kWhatEHard: ok = launch_m<kWhatEHard>(mode, K, L); break;
  }
  if (!ok) {
    set_error("unsupported points-per-lane %d for precision mode %d", K, mode);
    return CPET_ERR_ARG;
  }
  CPET_HIP_TRY(hipGetLastError());
  if ((prc = prof_mark(ctx, 1))) return prc;
  if (splits > 1) {
    const int blocks = (int)((M + 255) / 256);
    switch (what) {
      case kWhatESoft: hipLaunchKernelGGL((field_reduce_kernel<kWhatESoft>), dim3(blocks), dim3(256), 0, L.st, M, splits, o); break;
      case kWhatPhi: hipLaunchKernelGGL((field_reduce_kernel<kWhatPhi>), dim3(blocks), dim3(256), 0, L.st, M, splits, o); break;
      case kWhatESoftPhi: hipLaunchKernelGGL((field_reduce_kernel<kWhatESoftPhi>), dim3(blocks), dim3(256), 0, L.st, M, splits, o); break;
      case kWhatEHard: hipLaunchKernelGGL((field_reduce_kernel<kWhatEHard>), dim3(blocks), dim3(256), 0, L.st, M, splits, o); break;
    }
    CPET_HIP_TRY(hipGetLastError());
  }
  if ((prc = prof_mark(ctx, 2))) return prc;
  if (check) {
    FixArgs a = {};
    a.count = o.fix_count;
    a.list = o.fix_list;
    a.ch = ch;
    a.ax = grid_mode ? ax : nullptr;
    a.ay = ay;
    a.az = az;
    a.pts = grid_mode ? nullptr : pts;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.npad = npad;
    a.F = 0;
    a.e = E_out;
    a.f64 = o.f64;
    fix_launch(ctx, M, what != kWhatEHard, a);
    CPET_HIP_TRY(hipGetLastError());
  }
  if ((prc = prof_mark(ctx, 3))) return prc;
  return CPET_OK;
}

// Many grids in one launch (see field_batch_kernel).  Host tables: nx, ny,
// nz per frame (geo3, F*3), charges already packed (chg_off).
int field_batch_launch(cpet_ctx* ctx, int F, const int32_t* geo3, const float* axes_dev,
                       const float4* ch, const double4* chd, const int32_t* d_chg_off,
                       int ch_total, void* E_out, void* phi_out, int soften, uint32_t flags) {
  const int what = what_of(E_out != nullptr, phi_out != nullptr, soften);
  if (what < 0) {
    set_error("E and phi together require soften=1 (volume semantics)");
    return CPET_ERR_ARG;
  }
  const int mode = (int)((flags & CPET_F_PREC_MASK) >> 8);
  if (!field_mode_valid(mode)) {
    set_error("unknown precision mode %d", mode);
    return CPET_ERR_ARG;
  }
  int K = ctx->tune[4][0] > 0 ? ctx->tune[4][0] : 2;  // as the single-frame grid kernel
  std::vector<int32_t> geo(4 * (size_t)F), ax_off(F), blk;
  std::vector<int64_t> out_off(F);
  long long pts = 0;
  int axn = 0;
  for (int f = 0; f < F; ++f) {
    const int nx = geo3[3 * f], ny = geo3[3 * f + 1], nz = geo3[3 * f + 2];
    if (nx <= 0 || ny <= 0 || nz <= 0) {
      set_error("frame %d: grid dimensions must be positive", f);
      return CPET_ERR_ARG;
    }
    const int nzc = (nz + K - 1) / K;
    geo[4 * f] = nx;
    geo[4 * f + 1] = ny;
    geo[4 * f + 2] = nz;
    geo[4 * f + 3] = nzc;
    ax_off[f] = axn;
    axn += nx + ny + nz;
    out_off[f] = pts;
    pts += (long long)nx * ny * nz;
    const long long lanes = (long long)nx * ny * nzc;
    for (long long b = 0; b < (lanes + 255) / 256; ++b) {
      blk.push_back(f);
      blk.push_back((int32_t)b);
    }
  }
  const size_t o_geo = 0;
  const size_t o_ax = (o_geo + geo.size() * 4 + 255) & ~(size_t)255;
  const size_t o_out = (o_ax + ax_off.size() * 4 + 255) & ~(size_t)255;
  const size_t o_blk = (o_out + out_off.size() * 8 + 255) & ~(size_t)255;
  void* work = nullptr;
  int rc = ctx_buf(ctx, kSlotVolTables, o_blk + blk.size() * 4 + 256, &work);
  if (rc) return rc;
  char* w = (char*)work;
  hipStream_t st = ctx->stream;
  if ((rc = upload_table(ctx, kTabVolGeo, w + o_geo, geo.data(), geo.size() * 4))) return rc;
  if ((rc = upload_table(ctx, kTabVolAx, w + o_ax, ax_off.data(), ax_off.size() * 4))) return rc;
  if ((rc = upload_table(ctx, kTabVolOut, w + o_out, out_off.data(), out_off.size() * 8))) return rc;
  if ((rc = upload_table(ctx, kTabVolBlk, w + o_blk, blk.data(), blk.size() * 4))) return rc;
  BatchFieldArgs a;
  a.blk = (const int32_t*)(w + o_blk);
  a.geo = (const int32_t*)(w + o_geo);
  a.out_off = (const int64_t*)(w + o_out);
  a.ax_off = (const int32_t*)(w + o_ax);
  a.chg_off = d_chg_off;
  a.axes = axes_dev;
  a.e = E_out;
  a.phi = phi_out;
  a.f64 = (flags & CPET_F_OUT_F64) ? 1 : 0;
  a.fix_count = nullptr;
  a.fix_list = nullptr;
  a.near = nullptr;
  ctx->fix_count = nullptr;
  if (what == kWhatESoft && ch_total > 0) {
    void* nf = nullptr;
    if ((rc = ctx_buf(ctx, kSlotNear, (size_t)F * sizeof(int), &nf))) return rc;
    hipLaunchKernelGGL(near_clear_n_kernel, dim3((F + 255) / 256), dim3(256), 0, st, (int*)nf, F);
    hipLaunchKernelGGL(near_batch_kernel, dim3((ch_total + 255) / 256), dim3(256), 0, st,
                       ch_total, ch, F, d_chg_off, a.geo, a.ax_off, axes_dev, (int*)nf);
    CPET_HIP_TRY(hipGetLastError());
    a.near = (const int*)nf;
  }
  const bool check = mode == kModeAuto && E_out != nullptr;
  if (check && (rc = fix_prepare(ctx, pts, &a.fix_count, &a.fix_list))) return rc;
  const int nblk = (int)(blk.size() / 2);
  if ((rc = prof_mark(ctx, 0))) return rc;
  bool ok = false;
  switch (what) {
    case kWhatESoft: ok = launch_batch_w<kWhatESoft>(mode, K, st, nblk, ch, chd, a); break;
    case kWhatPhi: ok = launch_batch_w<kWhatPhi>(mode, K, st, nblk, ch, chd, a); break;
    case kWhatESoftPhi: ok = launch_batch_w<kWhatESoftPhi>(mode, K, st, nblk, ch, chd, a); break;
    case kWhatEHard: ok = launch_batch_w<kWhatEHard>(mode, K, st, nblk, ch, chd, a); break;
  }
  if (!ok) {
    set_error("unsupported points-per-lane %d for the batch kernel (1, 2, 4)", K);
    return CPET_ERR_ARG;
  }
  CPET_HIP_TRY(hipGetLastError());
  if ((rc = prof_mark(ctx, 1))) return rc;
  if (check) {
    FixArgs fa = {};
    fa.count = a.fix_count;
    fa.list = a.fix_list;
    fa.ch = ch;
    fa.F = F;
    fa.geo = a.geo;
    fa.out_off = a.out_off;
    fa.ax_off = a.ax_off;
    fa.axes = axes_dev;
    fa.chg_off = d_chg_off;
    fa.e = E_out;
    fa.f64 = a.f64;
    fix_launch(ctx, pts, what != kWhatEHard, fa);
    CPET_HIP_TRY(hipGetLastError());
  }
  if ((rc = prof_mark(ctx, 2))) return rc;
  return CPET_OK;
}

}  // namespace cpet
