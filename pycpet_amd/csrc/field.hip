// Direct Coulomb field / electrostatic potential kernels for gfx950.
//
// Hot loop replaced: compute_looped_field (CPET/utils/math_module.c:405-451,
// E with an r^2 >= 1e-6 floor), calc_esp_base (:453-486, phi, no floor) and
// calc_field (:255-293, E, no floor).  The reference runs one FP32 point at
// a time on one CPU thread; here one GPU lane owns K points, every lane of a
// wave walks the same charge, and charges are read through the scalar unit
// (wave-uniform addresses -> s_load), so the vector ALUs only do pair math:
//
//   grid kernel (points on a tensor-product mesh, K points along z per lane)
//     per charge: dx, dy, dxy2 = dy*dy + dx*dx                (4 ops / K)
//     per pair:   dz, r2 = dz*dz + dxy2, rinv = rsq(r2),
//                 [min(rinv, rsq(1e-6))], s = (q*re)*(re*re),
//                 3 FMA into E, [1 FMA into phi]               (~9-11 ops)
//   points kernel (arbitrary points): 3 subtracts + 3 ops for r2 instead.
//
// Accumulation: FP32 inside a 64-charge tile, FP64 across tiles (default),
// or naive FP32 / Kahan FP32 / FP64 for the precision study.  Small grids
// split the charges over blockIdx.y and reduce the FP64 partials in a fixed
// order in a second kernel.

#include <cmath>
#include <cstdio>

#include "cpet_internal.h"

namespace cpet {
namespace {

__device__ __forceinline__ double rsqrt_f64(double x) {
  // Hardware estimate refined by two Newton-Raphson steps for 1/sqrt(x):
  // y <- y * (1.5 - 0.5 x y^2).  x == 0 keeps +inf like 1/sqrt(0).
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double y1 = y * fma(-hx * y, y, 1.5);
  y1 = y1 * fma(-hx * y1, y1, 1.5);
  return (x == 0.0) ? y : y1;
}

// One FP32 pair term.  WHAT selects field (floored or not) and/or phi.
template <int WHAT>
__device__ __forceinline__ void pair_f32(float dx, float dy, float dz,
                                         float dxy2, float q, float rclamp,
                                         float& ex, float& ey, float& ez,
                                         float& ph) {
  (void)dy;
  const float r2 = fmaf(dz, dz, dxy2);
  const float ri = __builtin_amdgcn_rsqf(r2);
  if (WHAT == kWhatPhi || WHAT == kWhatESoftPhi) ph = fmaf(q, ri, ph);
  if (WHAT != kWhatPhi) {
    const float re = (WHAT == kWhatEHard) ? ri : fminf(ri, rclamp);
    const float s = (q * re) * (re * re);
    ex = fmaf(s, dx, ex);
    ey = fmaf(s, dy, ey);
    ez = fmaf(s, dz, ez);
  }
}

// Kahan-compensated accumulation of one term.
__device__ __forceinline__ void kahan_add(float t, float& s, float& c) {
  const float y = t - c;
  const float u = s + y;
  c = (u - s) - y;
  s = u;
}

template <int WHAT>
__device__ __forceinline__ void pair_kahan(float dx, float dy, float dz,
                                           float dxy2, float q, float rclamp,
                                           float (&acc)[4], float (&cmp)[4]) {
  const float r2 = fmaf(dz, dz, dxy2);
  const float ri = __builtin_amdgcn_rsqf(r2);
  if (WHAT == kWhatPhi || WHAT == kWhatESoftPhi) kahan_add(q * ri, acc[3], cmp[3]);
  if (WHAT != kWhatPhi) {
    const float re = (WHAT == kWhatEHard) ? ri : fminf(ri, rclamp);
    const float s = (q * re) * (re * re);
    kahan_add(s * dx, acc[0], cmp[0]);
    kahan_add(s * dy, acc[1], cmp[1]);
    kahan_add(s * dz, acc[2], cmp[2]);
  }
}

template <int WHAT>
__device__ __forceinline__ void pair_f64(double px, double py, double pz,
                                         float4 c, double (&acc)[4]) {
  const double dx = px - (double)c.x;
  const double dy = py - (double)c.y;
  const double dz = pz - (double)c.z;
  const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
  const double q = (double)c.w;
  if (WHAT == kWhatPhi || WHAT == kWhatESoftPhi) acc[3] += q * rsqrt_f64(r2);
  if (WHAT != kWhatPhi) {
    const double floor2 = (double)1e-6f;
    const double r2e = (WHAT == kWhatEHard) ? r2 : fmax(r2, floor2);
    const double re = rsqrt_f64(r2e);
    const double s = (q * re) * (re * re);
    acc[0] = fma(s, dx, acc[0]);
    acc[1] = fma(s, dy, acc[1]);
    acc[2] = fma(s, dz, acc[2]);
  }
}

// Accumulates the field of charges [t0*kTile, t1*kTile) at K points that
// share (px, py) and differ in z.  For the arbitrary-point kernel K points
// are passed with their own x/y through the generic path below.
template <int K, int PREC, int WHAT>
__device__ __forceinline__ void accumulate_zline(
    float px, float py, const float (&pz)[K], const float4* __restrict__ ch,
    int t0, int t1, float rclamp, double (&out)[K][4]) {
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) out[k][c] = 0.0;

  if constexpr (PREC == kPrecF64) {
    for (int j = t0 * kTile; j < t1 * kTile; ++j) {
      const float4 c = ch[j];
#pragma unroll
      for (int k = 0; k < K; ++k) pair_f64<WHAT>(px, py, pz[k], c, out[k]);
    }
  } else if constexpr (PREC == kPrecNaive) {
    float a[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c) a[k][c] = 0.f;
    for (int j = t0 * kTile; j < t1 * kTile; ++j) {
      const float4 c = ch[j];
      const float dx = px - c.x, dy = py - c.y;
      const float dxy2 = fmaf(dy, dy, dx * dx);
#pragma unroll
      for (int k = 0; k < K; ++k)
        pair_f32<WHAT>(dx, dy, pz[k] - c.z, dxy2, c.w, rclamp, a[k][0],
                       a[k][1], a[k][2], a[k][3]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c) out[k][c] = (double)a[k][c];
  } else if constexpr (PREC == kPrecKahan) {
    for (int t = t0; t < t1; ++t) {
      float a[K][4], cm[K][4];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) a[k][c] = cm[k][c] = 0.f;
      const float4* cp = ch + t * kTile;
#pragma unroll 4
      for (int j = 0; j < kTile; ++j) {
        const float4 c = cp[j];
        const float dx = px - c.x, dy = py - c.y;
        const float dxy2 = fmaf(dy, dy, dx * dx);
#pragma unroll
        for (int k = 0; k < K; ++k)
          pair_kahan<WHAT>(dx, dy, pz[k] - c.z, dxy2, c.w, rclamp, a[k], cm[k]);
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) out[k][c] += (double)a[k][c] - (double)cm[k][c];
    }
  } else {  // kPrecTiled
    for (int t = t0; t < t1; ++t) {
      float a[K][4];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) a[k][c] = 0.f;
      const float4* cp = ch + t * kTile;
#pragma unroll 16
      for (int j = 0; j < kTile; ++j) {
        const float4 c = cp[j];
        const float dx = px - c.x, dy = py - c.y;
        const float dxy2 = fmaf(dy, dy, dx * dx);
#pragma unroll
        for (int k = 0; k < K; ++k)
          pair_f32<WHAT>(dx, dy, pz[k] - c.z, dxy2, c.w, rclamp, a[k][0],
                         a[k][1], a[k][2], a[k][3]);
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) out[k][c] += (double)a[k][c];
    }
  }
}

// Per-point generic accumulation (each of the K points has its own x, y).
template <int K, int PREC, int WHAT>
__device__ __forceinline__ void accumulate_points(
    const float (&px)[K], const float (&py)[K], const float (&pz)[K],
    const float4* __restrict__ ch, int t0, int t1, float rclamp,
    double (&out)[K][4]) {
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) out[k][c] = 0.0;
  if constexpr (PREC == kPrecF64) {
    for (int j = t0 * kTile; j < t1 * kTile; ++j) {
      const float4 c = ch[j];
#pragma unroll
      for (int k = 0; k < K; ++k) pair_f64<WHAT>(px[k], py[k], pz[k], c, out[k]);
    }
  } else if constexpr (PREC == kPrecNaive) {
    float a[K][4];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c) a[k][c] = 0.f;
    for (int j = t0 * kTile; j < t1 * kTile; ++j) {
      const float4 c = ch[j];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float dx = px[k] - c.x, dy = py[k] - c.y;
        pair_f32<WHAT>(dx, dy, pz[k] - c.z, fmaf(dy, dy, dx * dx), c.w, rclamp,
                       a[k][0], a[k][1], a[k][2], a[k][3]);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c) out[k][c] = (double)a[k][c];
  } else if constexpr (PREC == kPrecKahan) {
    for (int t = t0; t < t1; ++t) {
      float a[K][4], cm[K][4];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) a[k][c] = cm[k][c] = 0.f;
      const float4* cp = ch + t * kTile;
#pragma unroll 4
      for (int j = 0; j < kTile; ++j) {
        const float4 c = cp[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float dx = px[k] - c.x, dy = py[k] - c.y;
          pair_kahan<WHAT>(dx, dy, pz[k] - c.z, fmaf(dy, dy, dx * dx), c.w,
                           rclamp, a[k], cm[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) out[k][c] += (double)a[k][c] - (double)cm[k][c];
    }
  } else {
    for (int t = t0; t < t1; ++t) {
      float a[K][4];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) a[k][c] = 0.f;
      const float4* cp = ch + t * kTile;
#pragma unroll 16
      for (int j = 0; j < kTile; ++j) {
        const float4 c = cp[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const float dx = px[k] - c.x, dy = py[k] - c.y;
          pair_f32<WHAT>(dx, dy, pz[k] - c.z, fmaf(dy, dy, dx * dx), c.w,
                         rclamp, a[k][0], a[k][1], a[k][2], a[k][3]);
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) out[k][c] += (double)a[k][c];
    }
  }
}

struct FieldOut {
  void* e;         // (M,3) float or double, or null
  void* phi;       // (M) float or double, or null
  double* part;    // (splits, M, 4) partial sums when splits > 1, else null
  int f64;         // outputs are double
  long long M;
};

template <int WHAT>
__device__ __forceinline__ void store_point(const FieldOut& o, long long m,
                                            const double (&acc)[4], int split) {
  if (o.part) {
    double* p = o.part + ((long long)split * o.M + m) * 4;
    p[0] = acc[0];
    p[1] = acc[1];
    p[2] = acc[2];
    p[3] = acc[3];
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

// Grid kernel: lane -> (line (i,j), chunk of K z-points).
template <int K, int PREC, int WHAT>
__global__ __launch_bounds__(256) void field_grid_kernel(
    int nx, int ny, int nz, int nzc, const float* __restrict__ ax,
    const float* __restrict__ ay, const float* __restrict__ az,
    const float4* __restrict__ ch, int tiles_per_split, int ntiles,
    FieldOut o) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long lanes = (long long)nx * ny * nzc;
  if (gid >= lanes) return;
  const long long line = gid / nzc;
  const int chunk = (int)(gid - line * nzc);
  const int i = (int)(line / ny);
  const int j = (int)(line - (long long)i * ny);
  const int z0 = chunk * K;
  const float px = ax[i], py = ay[j];
  float pz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) pz[k] = az[min(z0 + k, nz - 1)];
  const int split = blockIdx.y;
  const int t0 = split * tiles_per_split;
  const int t1 = min(t0 + tiles_per_split, ntiles);
  const float rclamp = __builtin_amdgcn_rsqf(1e-6f);
  double acc[K][4];
  accumulate_zline<K, PREC, WHAT>(px, py, pz, ch, t0, t1, rclamp, acc);
  const long long base = line * nz + z0;
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (z0 + k < nz) store_point<WHAT>(o, base + k, acc[k], split);
}

// Points kernel: lane owns points base + k*blockDim + tid.
template <int K, int PREC, int WHAT>
__global__ __launch_bounds__(256) void field_points_kernel(
    long long M, const float* __restrict__ pts, const float4* __restrict__ ch,
    int tiles_per_split, int ntiles, FieldOut o) {
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
  const float rclamp = __builtin_amdgcn_rsqf(1e-6f);
  double acc[K][4];
  accumulate_points<K, PREC, WHAT>(px, py, pz, ch, t0, t1, rclamp, acc);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const long long m = base + (long long)k * blockDim.x;
    if (m < M) store_point<WHAT>(o, m, acc[k], split);
  }
}

// Sums the per-split partials in split order and writes the outputs.
template <int WHAT>
__global__ __launch_bounds__(256) void field_reduce_kernel(long long M,
                                                           int splits,
                                                           FieldOut o) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < splits; ++s) {
    const double* p = o.part + ((long long)s * M + m) * 4;
    acc[0] += p[0];
    acc[1] += p[1];
    acc[2] += p[2];
    acc[3] += p[3];
  }
  FieldOut d = o;
  d.part = nullptr;
  store_point<WHAT>(d, m, acc, 0);
}

template <int WHAT>
void launch_reduce(hipStream_t st, long long M, int splits, const FieldOut& o) {
  const int blocks = (int)((M + 255) / 256);
  hipLaunchKernelGGL((field_reduce_kernel<WHAT>), dim3(blocks), dim3(256), 0,
                     st, M, splits, o);
}

// ---- dispatch tables --------------------------------------------------

template <int K, int PREC, int WHAT>
void launch_grid_t(hipStream_t st, dim3 grid, int nx, int ny, int nz, int nzc,
                   const float* ax, const float* ay, const float* az,
                   const float4* ch, int tps, int ntiles, const FieldOut& o) {
  hipLaunchKernelGGL((field_grid_kernel<K, PREC, WHAT>), grid, dim3(256), 0,
                     st, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o);
}

template <int K, int PREC, int WHAT>
void launch_points_t(hipStream_t st, dim3 grid, long long M, const float* pts,
                     const float4* ch, int tps, int ntiles, const FieldOut& o) {
  hipLaunchKernelGGL((field_points_kernel<K, PREC, WHAT>), grid, dim3(256), 0,
                     st, M, pts, ch, tps, ntiles, o);
}

template <int PREC, int WHAT>
bool launch_grid_k(int K, hipStream_t st, dim3 grid, int nx, int ny, int nz,
                   int nzc, const float* ax, const float* ay, const float* az,
                   const float4* ch, int tps, int ntiles, const FieldOut& o) {
  switch (K) {
    case 1: launch_grid_t<1, PREC, WHAT>(st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o); return true;
    case 2: launch_grid_t<2, PREC, WHAT>(st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o); return true;
    case 3: launch_grid_t<3, PREC, WHAT>(st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o); return true;
    case 4: launch_grid_t<4, PREC, WHAT>(st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o); return true;
    case 6: launch_grid_t<6, PREC, WHAT>(st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o); return true;
    case 8: launch_grid_t<8, PREC, WHAT>(st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o); return true;
    default: return false;
  }
}

template <int PREC, int WHAT>
bool launch_points_k(int K, hipStream_t st, dim3 grid, long long M,
                     const float* pts, const float4* ch, int tps, int ntiles,
                     const FieldOut& o) {
  switch (K) {
    case 1: launch_points_t<1, PREC, WHAT>(st, grid, M, pts, ch, tps, ntiles, o); return true;
    case 2: launch_points_t<2, PREC, WHAT>(st, grid, M, pts, ch, tps, ntiles, o); return true;
    case 4: launch_points_t<4, PREC, WHAT>(st, grid, M, pts, ch, tps, ntiles, o); return true;
    default: return false;
  }
}

template <int WHAT>
bool launch_grid_w(int prec, int K, hipStream_t st, dim3 grid, int nx, int ny,
                   int nz, int nzc, const float* ax, const float* ay,
                   const float* az, const float4* ch, int tps, int ntiles,
                   const FieldOut& o) {
  switch (prec) {
    case kPrecTiled: return launch_grid_k<kPrecTiled, WHAT>(K, st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o);
    case kPrecNaive: return launch_grid_k<kPrecNaive, WHAT>(K, st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o);
    case kPrecKahan: return launch_grid_k<kPrecKahan, WHAT>(K, st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o);
    case kPrecF64:   return launch_grid_k<kPrecF64, WHAT>(K, st, grid, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o);
  }
  return false;
}

template <int WHAT>
bool launch_points_w(int prec, int K, hipStream_t st, dim3 grid, long long M,
                     const float* pts, const float4* ch, int tps, int ntiles,
                     const FieldOut& o) {
  switch (prec) {
    case kPrecTiled: return launch_points_k<kPrecTiled, WHAT>(K, st, grid, M, pts, ch, tps, ntiles, o);
    case kPrecNaive: return launch_points_k<kPrecNaive, WHAT>(K, st, grid, M, pts, ch, tps, ntiles, o);
    case kPrecKahan: return launch_points_k<kPrecKahan, WHAT>(K, st, grid, M, pts, ch, tps, ntiles, o);
    case kPrecF64:   return launch_points_k<kPrecF64, WHAT>(K, st, grid, M, pts, ch, tps, ntiles, o);
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

}  // namespace

// Shared implementation of the grid and points entry points.
int field_launch(cpet_ctx* ctx, bool grid_mode, int nx, int ny, int nz,
                 const float* ax, const float* ay, const float* az, long long M,
                 const float* pts, const float4* ch, int npad, void* E_out,
                 void* phi_out, int soften, uint32_t flags, int forced_k,
                 int forced_splits) {
  const int what = what_of(E_out != nullptr, phi_out != nullptr, soften);
  if (what < 0) {
    set_error("E and phi together require soften=1 (volume semantics)");
    return CPET_ERR_ARG;
  }
  const int prec = (int)((flags & CPET_F_PREC_MASK) >> 8);
  const int ntiles = npad / kTile;
  FieldOut o;
  o.e = E_out;
  o.phi = phi_out;
  o.part = nullptr;
  o.f64 = (flags & CPET_F_OUT_F64) ? 1 : 0;
  o.M = M;

  int K;
  long long lanes;
  int nzc = 0;
  if (grid_mode) {
    if (forced_k > 0) {
      K = forced_k;
    } else if (prec == kPrecF64 || prec == kPrecKahan) {
      K = 2;
    } else {
      // choose K so the z-line splits evenly and enough waves remain
      K = 4;
      const long long l4 = (long long)nx * ny * ((nz + 3) / 4);
      if (l4 < (long long)ctx->num_cus * 64 * 4) K = 2;
      if (nz % 3 == 0 && nz % 4 != 0 && K == 4) K = 3;
    }
    nzc = (nz + K - 1) / K;
    lanes = (long long)nx * ny * nzc;
  } else {
    K = forced_k > 0 ? forced_k : (prec == kPrecF64 || prec == kPrecKahan ? 1 : 2);
    lanes = (M + K - 1) / K;
  }
  const int splits = pick_splits(lanes, ntiles, ctx->num_cus, forced_splits);
  const int tps = (ntiles + splits - 1) / splits;
  if (splits > 1) {
    void* part = nullptr;
    int rc = ctx_buf(ctx, kSlotPartial, (size_t)splits * M * 4 * sizeof(double),
                     &part);
    if (rc) return rc;
    o.part = (double*)part;
  }
  const long long blocks_x = grid_mode ? (lanes + 255) / 256
                                       : (M + 256LL * K - 1) / (256LL * K);
  dim3 g((unsigned)blocks_x, (unsigned)splits);
  bool ok = false;
  hipStream_t st = ctx->stream;
  int prc = prof_mark(ctx, 0);
  if (prc) return prc;
  switch (what) {
    case kWhatESoft:
      ok = grid_mode ? launch_grid_w<kWhatESoft>(prec, K, st, g, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o)
                     : launch_points_w<kWhatESoft>(prec, K, st, g, M, pts, ch, tps, ntiles, o);
      break;
    case kWhatPhi:
      ok = grid_mode ? launch_grid_w<kWhatPhi>(prec, K, st, g, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o)
                     : launch_points_w<kWhatPhi>(prec, K, st, g, M, pts, ch, tps, ntiles, o);
      break;
    case kWhatESoftPhi:
      ok = grid_mode ? launch_grid_w<kWhatESoftPhi>(prec, K, st, g, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o)
                     : launch_points_w<kWhatESoftPhi>(prec, K, st, g, M, pts, ch, tps, ntiles, o);
      break;
    case kWhatEHard:
      ok = grid_mode ? launch_grid_w<kWhatEHard>(prec, K, st, g, nx, ny, nz, nzc, ax, ay, az, ch, tps, ntiles, o)
                     : launch_points_w<kWhatEHard>(prec, K, st, g, M, pts, ch, tps, ntiles, o);
      break;
  }
  if (!ok) {
    set_error("unsupported points-per-lane %d", K);
    return CPET_ERR_ARG;
  }
  CPET_HIP_TRY(hipGetLastError());
  if ((prc = prof_mark(ctx, 1))) return prc;
  if (splits > 1) {
    switch (what) {
      case kWhatESoft: launch_reduce<kWhatESoft>(st, M, splits, o); break;
      case kWhatPhi: launch_reduce<kWhatPhi>(st, M, splits, o); break;
      case kWhatESoftPhi: launch_reduce<kWhatESoftPhi>(st, M, splits, o); break;
      case kWhatEHard: launch_reduce<kWhatEHard>(st, M, splits, o); break;
    }
    CPET_HIP_TRY(hipGetLastError());
  }
  if ((prc = prof_mark(ctx, 2))) return prc;
  return CPET_OK;
}

}  // namespace cpet
