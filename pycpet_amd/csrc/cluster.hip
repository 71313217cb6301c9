// The clustering step after "topo" on the GPU: per-frame normalised 2-D
// histograms of the streamline features and the pairwise chi-square
// distance matrix between them.
//
//   make_histograms            CPET/utils/calculator.py:596-718: per frame,
//       np.histogram2d(dist, curv, bins=[nd, nc], range=[[dmin, dmax],
//       [cmin, cmax]]) over shared edges, divided by its own total count;
//   distance_numpy             CPET/utils/calculator.py:975-978:
//       d(a, b) = sum over bins with a + b != 0 of (a - b)^2 / (a + b), / 2;
//   construct_distance_matrix  CPET/utils/calculator.py:1000-1015: all pairs,
//       zero diagonal.
//
// The bin rule is numpy's histogramdd one (searchsorted right on the edges,
// a value equal to the last edge in the last bin), so counts are exact.  The
// histogram values are count / total in FP64, correctly rounded as numpy's.
// The distance matrix is FP64; its sums run bin by bin in order, where
// numpy sums pairwise, so the two agree to FP64 rounding.

#include "cpet_internal.h"

namespace cpet {
namespace {

// counts[f][bd * nbc + bc] += 1 for every in-range feature row.  Rows of
// frame f are [off[f], off[f+1]).
__global__ __launch_bounds__(256) void frame_hist_kernel(
    int F, const long long* __restrict__ off, long long P,
    const double* __restrict__ feats, const double* __restrict__ ed, int nbd,
    const double* __restrict__ ec, int nbc, unsigned int* __restrict__ counts) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  int lo = 0, hi = F - 1;  // last frame whose first row is <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const int bd = hist_bin(ed, nbd, feats[2 * i]);
  const int bc = hist_bin(ec, nbc, feats[2 * i + 1]);
  if (bd >= 0 && bc >= 0)
    atomicAdd(&counts[(size_t)lo * nbd * nbc + (size_t)bd * nbc + bc], 1u);
}

// One workgroup per frame: total = sum of the frame's counts (exact in
// 64-bit integers), then hist = count / total (0 / 0 = NaN as in numpy).
__global__ __launch_bounds__(256) void frame_norm_kernel(int B,
                                                          const unsigned int* __restrict__ counts,
                                                          double* __restrict__ hist) {
  __shared__ unsigned long long part[256];
  const int f = blockIdx.x;
  const unsigned int* c = counts + (size_t)f * B;
  unsigned long long s = 0;
  for (int b = threadIdx.x; b < B; b += blockDim.x) s += c[b];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  const double total = (double)part[0];
  double* h = hist + (size_t)f * B;
  for (int b = threadIdx.x; b < B; b += blockDim.x) h[b] = (double)c[b] / total;
}

// Chi-square distances of a 16 x 16 tile of frame pairs (rows i0.., columns
// j0..), j-tile >= i-tile only; bins are staged through LDS 64 at a time.
// Thread (ti, tj) owns pair (i0 + ti, j0 + tj) and writes it and its mirror.
constexpr int kPairTile = 16;
constexpr int kBinChunk = 64;

__global__ __launch_bounds__(256) void chi2_kernel(int F, int B, int ntiles,
                                                    const double* __restrict__ hist,
                                                    double* __restrict__ out) {
  __shared__ double hi_s[kPairTile][kBinChunk + 1];
  __shared__ double hj_s[kPairTile][kBinChunk + 1];
  // blockIdx.x enumerates the upper-triangle tiles (ti <= tj) row by row
  int t = blockIdx.x, ti = 0;
  while (t >= ntiles - ti) {
    t -= ntiles - ti;
    ++ti;
  }
  const int tj = ti + t;
  const int i0 = ti * kPairTile, j0 = tj * kPairTile;
  const int a = threadIdx.x / kPairTile, b = threadIdx.x % kPairTile;
  const int i = i0 + a, j = j0 + b;
  double acc = 0.0;
  for (int c0 = 0; c0 < B; c0 += kBinChunk) {
    const int cn = min(kBinChunk, B - c0);
    for (int k = threadIdx.x; k < kPairTile * kBinChunk; k += blockDim.x) {
      const int r = k / kBinChunk, c = k % kBinChunk;
      const bool okc = c < cn;
      hi_s[r][c] = (okc && i0 + r < F) ? hist[(size_t)(i0 + r) * B + c0 + c] : 0.0;
      hj_s[r][c] = (okc && j0 + r < F) ? hist[(size_t)(j0 + r) * B + c0 + c] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < cn; ++c) {
      const double x = hi_s[a][c], y = hj_s[b][c];
      const double s = x + y;
      const double d = x - y;
      if (s != 0.0) acc += (d * d) / s;
    }
    __syncthreads();
  }
  if (i < F && j < F && j >= i) {
    const double v = (i == j) ? 0.0 : acc / 2.0;
    out[(size_t)i * F + j] = v;
    out[(size_t)j * F + i] = v;
  }
}

__global__ void zero_u32_kernel(unsigned int* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace
}  // namespace cpet

using namespace cpet;

extern "C" {

int cpet_cluster_histograms(cpet_ctx* ctx, int F, const int64_t* offsets, const double* feats,
                            int nb_d, const double* edges_d, int nb_c, const double* edges_c,
                            double* hists, uint32_t flags) {
  if (!ctx || F <= 0 || !offsets || !feats || nb_d <= 0 || nb_c <= 0 || !edges_d ||
      !edges_c || !hists) {
    set_error("cpet_cluster_histograms: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  // offsets are always read on the host (sizes and validation)
  for (int f = 0; f < F; ++f)
    if (offsets[f] > offsets[f + 1] || offsets[0] != 0) {
      set_error("cpet_cluster_histograms: offsets must start at 0 and not decrease");
      return CPET_ERR_ARG;
    }
  for (int k = 0; k < nb_d; ++k)
    if (!(edges_d[k] < edges_d[k + 1])) {
      set_error("cpet_cluster_histograms: distance edges must increase");
      return CPET_ERR_ARG;
    }
  for (int k = 0; k < nb_c; ++k)
    if (!(edges_c[k] < edges_c[k + 1])) {
      set_error("cpet_cluster_histograms: curvature edges must increase");
      return CPET_ERR_ARG;
    }
  const long long P = offsets[F];
  const long long B = (long long)nb_d * nb_c;
  if (B > (1LL << 28)) {
    set_error("cpet_cluster_histograms: %lld bins per histogram is too many", B);
    return CPET_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const void* dfeats = nullptr;
  if (P > 0 && (rc = stage_in(ctx, kSlotCluster0, feats, (size_t)P * 2 * sizeof(double), dev, &dfeats)))
    return rc;
  // small tables: offsets, edges, counts in one buffer
  const size_t o_off = 0;
  const size_t o_ed = round_up((int)((F + 1) * sizeof(long long)), 256);
  const size_t o_ec = o_ed + round_up((int)((nb_d + 1) * sizeof(double)), 256);
  const size_t o_cnt = o_ec + round_up((int)((nb_c + 1) * sizeof(double)), 256);
  void* tab = nullptr;
  if ((rc = ctx_buf(ctx, kSlotCluster1, o_cnt + (size_t)F * B * sizeof(unsigned int), &tab)))
    return rc;
  char* w = (char*)tab;
  CPET_HIP_TRY(hipMemcpyAsync(w + o_off, offsets, (F + 1) * sizeof(long long), hipMemcpyHostToDevice, st));
  CPET_HIP_TRY(hipMemcpyAsync(w + o_ed, edges_d, (nb_d + 1) * sizeof(double), hipMemcpyHostToDevice, st));
  CPET_HIP_TRY(hipMemcpyAsync(w + o_ec, edges_c, (nb_c + 1) * sizeof(double), hipMemcpyHostToDevice, st));
  unsigned int* counts = (unsigned int*)(w + o_cnt);
  const long long nc = (long long)F * B;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, counts, nc);
  if (P > 0)
    hipLaunchKernelGGL(frame_hist_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, F,
                       (const long long*)(w + o_off), P, (const double*)dfeats,
                       (const double*)(w + o_ed), nb_d, (const double*)(w + o_ec), nb_c, counts);
  double* dh = hists;
  if (!dev) {
    void* ob = nullptr;
    if ((rc = ctx_buf(ctx, kSlotCluster2, (size_t)nc * sizeof(double), &ob))) return rc;
    dh = (double*)ob;
  }
  hipLaunchKernelGGL(frame_norm_kernel, dim3(F), dim3(256), 0, st, (int)B, counts, dh);
  CPET_HIP_TRY(hipGetLastError());
  if (!dev) CPET_HIP_TRY(hipMemcpyAsync(hists, dh, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, st));
  if (!dev || !(flags & CPET_F_NO_SYNC)) CPET_HIP_TRY(hipStreamSynchronize(st));
  return CPET_OK;
}

int cpet_chi2_matrix(cpet_ctx* ctx, int F, int B, const double* hists, double* out,
                     uint32_t flags) {
  if (!ctx || F <= 0 || B <= 0 || !hists || !out) {
    set_error("cpet_chi2_matrix: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const void* dh = nullptr;
  if ((rc = stage_in(ctx, kSlotCluster0, hists, (size_t)F * B * sizeof(double), dev, &dh)))
    return rc;
  double* dout = out;
  if (!dev) {
    void* ob = nullptr;
    if ((rc = ctx_buf(ctx, kSlotCluster2, (size_t)F * F * sizeof(double), &ob))) return rc;
    dout = (double*)ob;
  }
  const int nt = (F + kPairTile - 1) / kPairTile;
  const long long tiles = (long long)nt * (nt + 1) / 2;
  if ((rc = prof_mark(ctx, 0))) return rc;
  hipLaunchKernelGGL(chi2_kernel, dim3((unsigned)tiles), dim3(256), 0, st, F, B, nt,
                     (const double*)dh, dout);
  CPET_HIP_TRY(hipGetLastError());
  if ((rc = prof_mark(ctx, 1))) return rc;
  if (!dev) CPET_HIP_TRY(hipMemcpyAsync(out, dout, (size_t)F * F * sizeof(double), hipMemcpyDeviceToHost, st));
  if (!dev || !(flags & CPET_F_NO_SYNC)) {
    CPET_HIP_TRY(hipStreamSynchronize(st));
    if ((rc = prof_collect(ctx, ctx->marks))) return rc;
  }
  return CPET_OK;
}

}  // extern "C"
