// Internal declarations shared by the HIP translation units of libcpet_hip.so.
// Nothing here is part of the C-ABI; the public surface is include/cpet_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpet_hip.h"

namespace cpet {

// Coulomb constant in eV*A/e^2, as used by every reference kernel
// (CPET/utils/math_module.c:305,412,462).
constexpr double kCoulomb = 14.3996451;

// Charges are padded to a multiple of kTile with zero charges placed far away
// (their terms underflow to exactly 0), so kernels never handle a remainder.
constexpr int kTile = 64;
constexpr float kPadCoord = 1.0e18f;

// Precision modes of the field kernels (bits 8..10 of the flags word).
enum Mode : int {
  kModeAuto = 0,   // E: FP32 tiled; phi: mixed (default)
  kModeNaive = 1,  // one FP32 running sum over all charges
  kModeKahan = 2,  // Kahan-compensated FP32 inside a tile, FP64 across tiles
  kModeF64 = 3,    // everything in FP64
  kModeTiled = 4,  // FP32 per 64-charge tile, FP64 across tiles
  kModeMixed = 5,  // FP64 d and r^2, FP32 rsq seed + FP64 Newton, FP64 sums
};

// What a field launch computes.
// Streamline field sources (topo_run): charges, or point dipoles with the
// reference's or the physical pair term.
enum SrcKind : int {
  kSrcCharge = 0,
  kSrcDipoleRef = 1,
  kSrcDipolePhys = 2,
  kSrcChargeRk4 = 3,
  kSrcChargeF64 = 4,     // propagate sum in FP64 (CPET_F_PREC_F64)
  kSrcChargeF64Rk4 = 5
};

enum What : int {
  kWhatESoft = 0,     // E with the r^2 >= 1e-6 floor (volume)
  kWhatPhi = 1,       // phi only, no floor (volume_ESP)
  kWhatESoftPhi = 2,  // both, E floored and phi not
  kWhatEHard = 3,     // E without floor (point field)
};

void set_error(const char* fmt, ...);

#define CPET_HIP_TRY(expr)                                                   \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::cpet::set_error("%s failed: %s (%s:%d)", #expr,                      \
                        hipGetErrorString(e_), __FILE__, __LINE__);          \
      return CPET_ERR_HIP;                                                   \
    }                                                                        \
  } while (0)

// A grow-only device buffer owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace cpet

// Context: one device, one stream, a set of scratch buffers.  Calls on one
// context are serialised by its mutex; different contexts run concurrently.
struct cpet_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // stream work is queued on
  hipStream_t own = nullptr;     // the context's own stream
  bool own_stream = false;
  int num_cus = 0;
  int tune[6][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};  // autotuner overrides
  std::mutex mu;
  std::vector<cpet::DevBuf> bufs;  // indexed by Slot
  // host-side pinned staging for small results
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  // opt-in per-phase timing with HIP events on the context's stream
  bool profile = false;
  hipEvent_t ev[8] = {};
  float phase_ms[4] = {0.f, 0.f, 0.f, 0.f};
  int marks = 0;  // events recorded by the last call, not yet collected
  // device counter of the points the last auto-mode field call recomputed
  // in FP64 (null when that call had no check)
  int* fix_count = nullptr;
  // last host table uploaded per key (upload_table): an identical table at
  // the same device address is not copied again, so repeated calls of one
  // shape enqueue kernels only (and can be captured into a HIP graph)
  struct TableCache {
    const char* dst = nullptr;
    std::vector<char> data;
  };
  TableCache tables[16];
  // Fixed scratch arena (the allocation-free entry point): when set,
  // ctx_buf carves every slot from it for the current call instead of
  // allocating, and fails when it is too small.
  char* arena = nullptr;
  size_t arena_bytes = 0;
  size_t arena_used = 0;
  void* arena_slot[32] = {};
  size_t arena_slot_bytes[32] = {};
};

namespace cpet {

enum Slot : int {
  kSlotCharges = 0,  // packed float4 charges
  kSlotCharges64,    // the same charges as double4
  kSlotPts,          // staged points / starts
  kSlotAxes,         // grid axes
  kSlotOut0,         // staged outputs
  kSlotOut1,
  kSlotPartial,      // split partial sums
  kSlotX,            // staged raw x
  kSlotQ,            // staged raw q
  kSlotTopoState,    // final points, steps, flags
  kSlotTopoOrder,    // work order
  kSlotTopoMisc,     // counters, histogram, edges
  kSlotIter,         // staged n_iter
  kSlotOut2,
  kSlotFrame0,       // frame-transform scratch
  kSlotFrame1,
  kSlotFrame2,
  kSlotBatchTables,  // volume batch charge offsets (written by upload_table only)
  kSlotBatchTables2, // topo batch charge offsets (written by upload_table only)
  kSlotBatchWork,    // batch counters, histograms, edges (kernel-written)
  kSlotVolTables,    // batch volume tables (written by upload_table only)
  kSlotTopoTables,   // batch topo tables (written by upload_table only)
  kSlotFix,          // auto-mode check: counter + list of points to recompute
  kSlotCluster0,     // cluster step: staged features / histograms
  kSlotCluster1,     // cluster step: edges, offsets, counts
  kSlotCluster2,     // cluster step: output
  kSlotNear,         // grid kernels: charge-near-grid flags
  kSlotFar,          // streamlines: far-field charge records + per-frame flags
  kSlotCount
};

// Returns a device buffer of at least `bytes` for `slot`, growing it if needed
// (or, for a context with an arena, a region of the arena).
int ctx_buf(cpet_ctx* ctx, int slot, size_t bytes, void** out);
// Forgets the arena regions of the previous call.
void arena_reset(cpet_ctx* ctx);
// cpet_topo after argument validation, with the context locked and active.
int topo_call(cpet_ctx* ctx, int N, int L, const float* starts, const int32_t* n_iter, float h,
              const float* hdims, const float* x, const float* q, double* feats,
              int32_t* steps, int32_t* exited, int nb_d, const double* edges_d, int nb_c,
              const double* edges_c, uint32_t* hist, uint32_t flags);
// Validation shared by cpet_topo and its allocation-free variant; copies
// dims into hdims.
int topo_check_args(int N, int L, const float* starts, const int32_t* n_iter, float h,
                    const float* dims, const float* x, const float* q, const double* feats,
                    int nb_d, const double* edges_d, int nb_c, const double* edges_c,
                    const uint32_t* hist, float* hdims);
// Makes ctx->device current on the calling thread.
int ctx_activate(cpet_ctx* ctx);
// Copies host->device (or passes a device pointer through when dev != 0).
int stage_in(cpet_ctx* ctx, int slot, const void* src, size_t bytes, bool dev,
             const void** out);
// Host table -> device copy on the context's stream, skipped when the same
// bytes were last uploaded to the same address under this key.  Overlapping
// entries of other keys, and the entries inside a scratch buffer that is
// reallocated, are forgotten.  Fails while the stream is being captured.
enum TableKey : int {
  kTabPackX = 0, kTabPackC, kTabPackX2, kTabPackC2,
  kTabVolGeo, kTabVolAx, kTabVolOut, kTabVolBlk,
  kTabTopoHd, kTabTopoSl, kTabTopoAs, kTabTopoBlk,
  kTabCount
};
int upload_table(cpet_ctx* ctx, int key, void* dst, const void* src, size_t bytes);
// Global default context for the reference-compatible entry points.
int default_ctx(cpet_ctx** out);

// Packs x (N,3) and q (N) into padded float4 charges (and, when out64 is
// non-null, a double4 copy); returns the padded count through npad.  x/q may
// be host or device pointers (dev flag).
int pack_charges(cpet_ctx* ctx, int N, const float* x, const float* q,
                 bool dev, const float4** out, const double4** out64, int* npad);

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Packs F frames' charges (x, q concatenated, nch[f] rows each) into one
// padded float4 buffer (and a double4 copy when out64 is non-null); each
// frame starts at chg_off[f] (a multiple of kTile).  chg_off (F+1) is filled
// on the host; *d_chg_off points at its device copy.
// `which` (0 volume, 1 topo) keeps the two batch kinds' offset tables apart,
// so one step may run both and still re-upload nothing.
int pack_charges_batch(cpet_ctx* ctx, int which, int F, const int32_t* nch, const float* x,
                       const float* q, bool dev, const float4** out,
                       const double4** out64, std::vector<int32_t>& chg_off,
                       const int32_t** d_chg_off);

// Phase timing: mark(ctx, i) records event i when profiling is on;
// collect(ctx, n) (after the stream has been synchronised) stores the n-1
// intervals between consecutive marks in ctx->phase_ms.
int prof_mark(cpet_ctx* ctx, int i);
int prof_collect(cpet_ctx* ctx, int n);

// numpy.searchsorted(edges, v, side='right') with the histogramdd rule that
// a value equal to the last edge falls into the last bin.  Returns the bin
// or -1 for an outlier (including NaN).
__device__ __forceinline__ int hist_bin(const double* __restrict__ e, int nb,
                                        double v) {
  int lo = 0, hi = nb + 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (v == e[nb]) lo -= 1;
  return (lo >= 1 && lo <= nb) ? lo - 1 : -1;
}

}  // namespace cpet
