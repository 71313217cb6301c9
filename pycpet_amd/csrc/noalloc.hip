// The allocation-free streamline entry point, cpet_topo_noalloc: cpet_topo on
// device pointers with every scratch buffer (packed charges, far-field
// records, work order, final points, counters, histogram staging) carved
// from a fixed device array, for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It has no counterpart in the reference.
//
// The array is one file-scope __device__ array; each GPU that runs the
// library holds its own copy until the process ends.  Host code never uses
// its name as an address: the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current.
// Calls on one GPU share the array: the library serialises them on the
// host, but calls queued on different streams of one GPU that overlap on
// the device race on it.  Everything that must start at zero (queue
// counters, bucket counts, histograms, flags) is cleared by a fill kernel
// inside each call; nothing else of the previous call's contents is read.

#include <mutex>

#include "cpet_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)CPET_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceArena {
  std::mutex mu;
  bool ready = false;
  char* ptr = nullptr;       // device address of this device's copy of the array
  cpet_ctx* ctx = nullptr;   // arena context (no stream, no buffers of its own)
};

DeviceArena g_arena[kMaxDevices];

// Set-up on the calling thread's current device, once: the device address
// of the array and an arena context.  Called with g_arena[dev].mu held.
int arena_setup(int dev, DeviceArena& a) {
  if (a.ready) return CPET_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    cpet::set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                    hipGetErrorString(e));
    return CPET_ERR_HIP;
  }
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    cpet::set_error("device %d attributes: %s", dev, hipGetErrorString(e));
    return CPET_ERR_HIP;
  }
  cpet_ctx* c = new cpet_ctx();
  c->device = dev;
  c->num_cus = cus;
  c->bufs.resize(cpet::kSlotCount);
  c->arena = (char*)p;
  c->arena_bytes = kScratchBytes;
  a.ptr = (char*)p;
  a.ctx = c;
  a.ready = true;
  return CPET_OK;
}

}  // namespace

extern "C" {

int cpet_topo_noalloc(void* hip_stream, int N, int L, const float* starts, const int32_t* n_iter,
                      float h, const float* dims, const float* x, const float* q, double* feats,
                      int32_t* steps, int32_t* exited, int nb_d, const double* edges_d, int nb_c,
                      const double* edges_c, uint32_t* hist, uint32_t flags) {
  float hdims[3];
  int rc = cpet::topo_check_args(N, L, starts, n_iter, h, dims, x, q, feats, nb_d, edges_d,
                                 nb_c, edges_c, hist, hdims);
  if (rc) return rc;
  if ((flags & CPET_F_PREC_MASK) == CPET_F_PREC_F64 || (flags & CPET_F_TOPO_RK4)) {
    cpet::set_error("cpet_topo_noalloc runs the default precision and the Euler walk only");
    return CPET_ERR_ARG;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    cpet::set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    return CPET_ERR_NODEV;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    cpet::set_error("device %d beyond the %d supported", dev, kMaxDevices);
    return CPET_ERR_ARG;
  }
  DeviceArena& a = g_arena[dev];
  std::lock_guard<std::mutex> lk(a.mu);
  if ((rc = arena_setup(dev, a))) return rc;
  cpet_ctx* c = a.ctx;
  std::lock_guard<std::mutex> lk2(c->mu);
  c->stream = (hipStream_t)hip_stream;
  cpet::arena_reset(c);
  return cpet::topo_call(c, N, L, starts, n_iter, h, hdims, x, q, feats, steps, exited, nb_d,
                         edges_d, nb_c, edges_c, hist, flags | CPET_F_DEVICE);
}

}  // extern "C"
