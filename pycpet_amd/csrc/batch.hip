// C-ABI of the multi-frame (batched) entry points: many frames of different
// sizes in one launch each for the volume and the streamline paths.  The
// reference processes frames one at a time from a Python loop
// (CPET/source/CPET.py:96-229); these calls take the whole frame set.

#include <cmath>
#include <vector>

#include "cpet_internal.h"

namespace cpet {
int field_batch_launch(cpet_ctx* ctx, int F, const int32_t* geo3, const float* axes_dev,
                       const float4* ch, const double4* chd, const int32_t* d_chg_off,
                       int ch_total, void* E_out, void* phi_out, int soften, uint32_t flags);
int topo_batch_run(cpet_ctx* ctx, int F, const int32_t* n_starts, const float* starts,
                   const int32_t* n_iter, const float* h, const float* dims, const float4* ch,
                   const int32_t* d_chg_off, const std::vector<int32_t>& chg_off, double* feats,
                   int32_t* steps, int32_t* exited, int nb_d, const double* edges_d, int nb_c,
                   const double* edges_c, uint32_t* hist, uint32_t flags);
}  // namespace cpet

using namespace cpet;

extern "C" {

int cpet_batch_volume(cpet_ctx* ctx, int F, const int32_t* n_charges, const float* x,
                      const float* q, const int32_t* grid_dims, const float* axes, void* E_out,
                      void* phi_out, int soften, uint32_t flags) {
  if (!ctx || F <= 0 || !n_charges || !x || !q || !grid_dims || !axes || (!E_out && !phi_out)) {
    set_error("cpet_batch_volume: bad arguments");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  long long M = 0, naxes = 0;
  for (int f = 0; f < F; ++f) {
    if (grid_dims[3 * f] <= 0 || grid_dims[3 * f + 1] <= 0 || grid_dims[3 * f + 2] <= 0) {
      set_error("frame %d: grid dimensions must be positive", f);
      return CPET_ERR_ARG;
    }
    M += (long long)grid_dims[3 * f] * grid_dims[3 * f + 1] * grid_dims[3 * f + 2];
    naxes += grid_dims[3 * f] + grid_dims[3 * f + 1] + grid_dims[3 * f + 2];
  }
  const size_t osz = (flags & CPET_F_OUT_F64) ? sizeof(double) : sizeof(float);
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const float4* ch = nullptr;
  const double4* chd = nullptr;
  std::vector<int32_t> chg_off;
  const int32_t* d_chg_off = nullptr;
  if ((rc = pack_charges_batch(ctx, 0, F, n_charges, x, q, dev, &ch, &chd, chg_off, &d_chg_off)))
    return rc;
  const void* daxes = nullptr;
  if ((rc = stage_in(ctx, kSlotAxes, axes, (size_t)naxes * sizeof(float), dev, &daxes))) return rc;
  void* dE = E_out;
  void* dphi = phi_out;
  if (!dev) {
    if (E_out && (rc = ctx_buf(ctx, kSlotOut0, (size_t)M * 3 * osz, &dE))) return rc;
    if (phi_out && (rc = ctx_buf(ctx, kSlotOut1, (size_t)M * osz, &dphi))) return rc;
  }
  rc = field_batch_launch(ctx, F, grid_dims, (const float*)daxes, ch, chd, d_chg_off,
                          chg_off.empty() ? 0 : chg_off.back(),
                          E_out ? dE : nullptr, phi_out ? dphi : nullptr, soften, flags);
  if (rc) return rc;
  if (!dev) {
    if (E_out)
      CPET_HIP_TRY(hipMemcpyAsync(E_out, dE, (size_t)M * 3 * osz, hipMemcpyDeviceToHost, ctx->stream));
    if (phi_out)
      CPET_HIP_TRY(hipMemcpyAsync(phi_out, dphi, (size_t)M * osz, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (!dev || !(flags & CPET_F_NO_SYNC)) {
    CPET_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((rc = prof_collect(ctx, ctx->marks))) return rc;
  }
  return CPET_OK;
}

int cpet_batch_topo(cpet_ctx* ctx, int F, const int32_t* n_charges, const float* x,
                    const float* q, const int32_t* n_starts, const float* starts,
                    const int32_t* n_iter, const float* h, const float* dims, double* feats,
                    int32_t* steps, int32_t* exited, int nb_d, const double* edges_d, int nb_c,
                    const double* edges_c, uint32_t* hist, uint32_t flags) {
  if (!ctx || F <= 0 || !n_charges || !x || !q || !n_starts || !starts || !n_iter || !h ||
      !dims || !feats) {
    set_error("cpet_batch_topo: bad arguments");
    return CPET_ERR_ARG;
  }
  if (hist && (nb_d <= 0 || nb_c <= 0 || !edges_d || !edges_c)) {
    set_error("histogram requested without valid bins/edges");
    return CPET_ERR_ARG;
  }
  const bool dev = (flags & CPET_F_DEVICE) != 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int rc = ctx_activate(ctx);
  if (rc) return rc;
  const float4* ch = nullptr;
  std::vector<int32_t> chg_off;
  const int32_t* d_chg_off = nullptr;
  if ((rc = pack_charges_batch(ctx, 1, F, n_charges, x, q, dev, &ch, nullptr, chg_off, &d_chg_off)))
    return rc;
  return topo_batch_run(ctx, F, n_starts, starts, n_iter, h, dims, ch, d_chg_off, chg_off,
                        feats, steps, exited, nb_d, edges_d, nb_c, edges_c, hist, flags);
}

}  // extern "C"
