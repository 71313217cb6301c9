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
// 64-bit integers), then hist = count / total (0 / 0 = NaN as in numpy), or
// the counts themselves when !normalise.
__global__ __launch_bounds__(256) void frame_norm_kernel(int B,
                                                          const unsigned int* __restrict__ counts,
                                                          double* __restrict__ hist,
                                                          int normalise) {
  if (!normalise) {
    const unsigned int* c = counts + (size_t)blockIdx.x * B;
    double* h = hist + (size_t)blockIdx.x * B;
    for (int b = threadIdx.x; b < B; b += blockDim.x) h[b] = (double)c[b];
    return;
  }
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

// ---- order statistics by radix select -------------------------------------
// make_histograms takes the global min, max and the 25th / 75th percentiles
// (scipy.stats.iqr -> numpy.percentile, linear method) of every frame's
// features (CPET/utils/calculator.py:664-670).  Every one of those is built
// from order statistics: the k-th smallest value of a column.  Each requested
// statistic (a "target") is found by a most-significant-digit radix select
// over 64-bit order-preserving keys, eight 8-bit digits: a pass counts, for
// every target, the digit values of the elements whose higher digits equal
// the target's decided prefix, and a one-workgroup pick kernel then descends
// into the digit bucket that holds the target's remaining rank.  Targets
// whose column and prefix agree share one count (their leader's).  The
// result is exact (it is an element of the column); NaNs sort last, as in
// numpy's sort, and are counted per column so the caller can apply numpy's
// NaN rule; -0.0 is taken as +0.0.
constexpr int kSelMaxTargets = 16;
constexpr int kSelMaxCols = 8;
constexpr int kSelPasses = 8;

struct SelTarget {
  unsigned long long prefix;  // decided high digits (lower digits zero)
  unsigned long long rank;    // rank within the elements sharing the prefix
  int col;
  int pad;
};

__device__ inline unsigned long long order_key(double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  if ((b & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) return ~0ull;  // NaN: last
  if ((b << 1) == 0) b = 0;                                              // -0.0 -> +0.0
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ inline double key_value(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

__device__ inline unsigned long long pass_mask(int pass) {
  return pass == 0 ? 0ull : (~0ull << (64 - 8 * pass));
}

// Target t's leader: the first target with the same column and prefix.
__device__ inline int sel_leader(const unsigned long long* pre, const int* col, int t,
                                 unsigned long long mask) {
  for (int u = 0; u < t; ++u)
    if (col[u] == col[t] && (pre[u] & mask) == (pre[t] & mask)) return u;
  return t;
}

// One counting pass over the flat (n, cols) array.  ghist[t][256] gets the
// digit counts of leader targets; nan_cnt[c] the NaNs of column c (pass 0).
__global__ __launch_bounds__(256) void select_count_kernel(
    long long total, int cols, const double* __restrict__ x, int T,
    const SelTarget* __restrict__ tg, int pass, unsigned int* __restrict__ ghist,
    unsigned long long* __restrict__ nan_cnt) {
  __shared__ unsigned int h[kSelMaxTargets][256];
  __shared__ unsigned long long s_pre[kSelMaxTargets];
  __shared__ int s_col[kSelMaxTargets];
  __shared__ int s_lead[kSelMaxTargets];
  __shared__ unsigned int s_nan[kSelMaxCols];
  const unsigned long long mask = pass_mask(pass);
  const int shift = 56 - 8 * pass;
  for (int k = threadIdx.x; k < T * 256; k += blockDim.x) h[k >> 8][k & 255] = 0u;
  if ((int)threadIdx.x < T) {
    s_pre[threadIdx.x] = tg[threadIdx.x].prefix;
    s_col[threadIdx.x] = tg[threadIdx.x].col;
  }
  if ((int)threadIdx.x < kSelMaxCols) s_nan[threadIdx.x] = 0u;
  __syncthreads();
  if ((int)threadIdx.x < T) s_lead[threadIdx.x] = sel_leader(s_pre, s_col, threadIdx.x, mask);
  __syncthreads();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
    const int c = (int)(j % cols);
    const unsigned long long key = order_key(x[j]);
    if (pass == 0 && key == ~0ull) atomicAdd(&s_nan[c], 1u);  // only NaNs map to ~0
    const unsigned int d = (unsigned int)(key >> shift) & 255u;
    for (int t = 0; t < T; ++t)
      if (s_lead[t] == t && s_col[t] == c && (key & mask) == s_pre[t]) atomicAdd(&h[t][d], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < T * 256; k += blockDim.x) {
    const int t = k >> 8;
    const unsigned int v = h[t][k & 255];
    if (v && s_lead[t] == t) atomicAdd(&ghist[k], v);
  }
  if (pass == 0 && (int)threadIdx.x < cols && s_nan[threadIdx.x])
    atomicAdd(&nan_cnt[threadIdx.x], (unsigned long long)s_nan[threadIdx.x]);
}

// One thread per target: descend into the digit bucket holding the target's
// rank (the leader's counts), and after the last pass write the value.
__global__ void select_pick_kernel(int T, SelTarget* __restrict__ tg, int pass,
                                   const unsigned int* __restrict__ ghist,
                                   double* __restrict__ values) {
  __shared__ unsigned long long s_pre[kSelMaxTargets];
  __shared__ int s_col[kSelMaxTargets];
  const int t = threadIdx.x;
  if (t < T) {
    s_pre[t] = tg[t].prefix;
    s_col[t] = tg[t].col;
  }
  __syncthreads();
  if (t >= T) return;
  const unsigned long long mask = pass_mask(pass);
  const unsigned int* hc = ghist + 256 * sel_leader(s_pre, s_col, t, mask);
  unsigned long long r = tg[t].rank;
  int b = 0;
  for (; b < 255; ++b) {
    if (r < hc[b]) break;
    r -= hc[b];
  }
  const unsigned long long pre = s_pre[t] | ((unsigned long long)b << (56 - 8 * pass));
  tg[t].prefix = pre;
  tg[t].rank = r;
  if (pass == kSelPasses - 1) values[t] = key_value(pre);
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
  hipLaunchKernelGGL(frame_norm_kernel, dim3(F), dim3(256), 0, st, (int)B, counts, dh,
                     (flags & CPET_F_COUNTS) ? 0 : 1);
  CPET_HIP_TRY(hipGetLastError());
  if (!dev) CPET_HIP_TRY(hipMemcpyAsync(hists, dh, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, st));
  if (!dev || !(flags & CPET_F_NO_SYNC)) CPET_HIP_TRY(hipStreamSynchronize(st));
  return CPET_OK;
}

int cpet_order_stats(cpet_ctx* ctx, int64_t n, int cols, const double* x, int n_sel,
                     const int32_t* sel_cols, const int64_t* sel_ranks, double* values,
                     int64_t* nan_counts, uint32_t flags) {
  if (!ctx || n <= 0 || cols <= 0 || cols > kSelMaxCols || !x || n_sel <= 0 ||
      n_sel > kSelMaxTargets || !sel_cols || !sel_ranks || !values) {
    set_error("cpet_order_stats: bad arguments (1 <= cols <= %d, 1 <= n_sel <= %d)",
              kSelMaxCols, kSelMaxTargets);
    return CPET_ERR_ARG;
  }
  if (n > (int64_t)UINT32_MAX) {
    set_error("cpet_order_stats: %lld rows is more than the 32-bit digit counts hold",
              (long long)n);
    return CPET_ERR_ARG;
  }
  SelTarget tgh[kSelMaxTargets];
  for (int k = 0; k < n_sel; ++k) {
    if (sel_cols[k] < 0 || sel_cols[k] >= cols || sel_ranks[k] < 0 || sel_ranks[k] >= n) {
      set_error("cpet_order_stats: selection %d (column %d, rank %lld) is outside the array",
                k, (int)sel_cols[k], (long long)sel_ranks[k]);
      return CPET_ERR_ARG;
    }
    tgh[k] = SelTarget{0ull, (unsigned long long)sel_ranks[k], (int)sel_cols[k], 0};
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const long long total = (long long)n * cols;
  const void* dx = nullptr;
  if ((rc = stage_in(ctx, kSlotCluster0, x, (size_t)total * sizeof(double), dev, &dx))) return rc;
  // tables: targets | per-pass digit counts | NaN counts | values
  const size_t o_tg = 0;
  const size_t o_h = 512;
  const size_t hist_words = (size_t)kSelPasses * n_sel * 256;
  const size_t o_nan = o_h + hist_words * sizeof(unsigned int);
  const size_t o_val = o_nan + kSelMaxCols * sizeof(unsigned long long);
  const size_t bytes = o_val + kSelMaxTargets * sizeof(double);
  static_assert(kSelMaxTargets * sizeof(SelTarget) <= 512, "target table");
  static_assert(kSelMaxCols * sizeof(unsigned long long) % sizeof(unsigned int) == 0, "align");
  void* tab = nullptr;
  if ((rc = ctx_buf(ctx, kSlotCluster1, bytes, &tab))) return rc;
  char* w = (char*)tab;
  SelTarget* dtg = (SelTarget*)(w + o_tg);
  unsigned int* dh = (unsigned int*)(w + o_h);
  unsigned long long* dnan = (unsigned long long*)(w + o_nan);
  double* dval = (double*)(w + o_val);
  CPET_HIP_TRY(hipMemcpyAsync(dtg, tgh, n_sel * sizeof(SelTarget), hipMemcpyHostToDevice, st));
  // the counts and the NaN counters are contiguous u32 words
  const long long zw = (long long)(hist_words + kSelMaxCols * 2);
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((zw + 255) / 256)), dim3(256), 0, st, dh, zw);
  const long long want = (total + 255) / 256;
  const unsigned blocks = (unsigned)(want < 4096 ? want : 4096);
  for (int p = 0; p < kSelPasses; ++p) {
    unsigned int* hp = dh + (size_t)p * n_sel * 256;
    hipLaunchKernelGGL(select_count_kernel, dim3(blocks), dim3(256), 0, st, total, cols,
                       (const double*)dx, n_sel, (const SelTarget*)dtg, p, hp, dnan);
    hipLaunchKernelGGL(select_pick_kernel, dim3(1), dim3(64), 0, st, n_sel, dtg, p,
                       (const unsigned int*)hp, dval);
  }
  CPET_HIP_TRY(hipGetLastError());
  unsigned long long nanh[kSelMaxCols];
  CPET_HIP_TRY(hipMemcpyAsync(values, dval, n_sel * sizeof(double), hipMemcpyDeviceToHost, st));
  CPET_HIP_TRY(hipMemcpyAsync(nanh, dnan, cols * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  CPET_HIP_TRY(hipStreamSynchronize(st));
  if (nan_counts)
    for (int c = 0; c < cols; ++c) nan_counts[c] = (int64_t)nanh[c];
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
