// VALU issue-rate probe for gfx950: the per-lane cost of v_fma_f32, of
// v_rsq_f32 and of the two interleaved, each over 8 independent chains per
// lane, so that the roofline ceilings of the field kernels can be computed
// from measured rates rather than assumed ones.
//
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         pycpet_amd/csrc/probe/issue_rate.hip -o /tmp/issue_rate && /tmp/issue_rate
//
// Every lane writes one float to out[gid] with gid < blocks * 256 = n.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));     \
      return 1;                                                               \
    }                                                                         \
  } while (0)

template <int OP>
__global__ __launch_bounds__(256) void rate_kernel(float* out, int iters, float seed) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float a[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) a[u] = seed + 0.001f * (gid & 63) + 0.01f * u;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (OP == 0) {
        a[u] = __builtin_fmaf(a[u], 0.999f, 0.0001f);
      } else if (OP == 1) {
        a[u] = __builtin_amdgcn_rsqf(a[u]);
      } else {
        const float r = __builtin_amdgcn_rsqf(a[u]);
        a[u] = __builtin_fmaf(r, 0.5f, a[u]);
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int u = 0; u < 8; ++u) s += a[u];
  out[gid] = s;
}

int main() {
  const int blocks = 256 * 32, threads = 256, iters = 4096;
  const size_t n = (size_t)blocks * threads;
  int dev = 0, cus = 0, khz = 0;
  CHECK(hipGetDevice(&dev));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, dev));
  float* d = nullptr;
  CHECK(hipMalloc(&d, n * sizeof(float)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const char* names[3] = {"v_fma_f32", "v_rsq_f32", "rsq+fma"};
  std::printf("CUs %d, peak clock %.0f MHz; per lane-op, 8 independent chains per lane\n", cus,
              khz / 1e3);
  for (int op = 0; op < 3; ++op) {
    float ms = 0.f;
    for (int rep = 0; rep < 3; ++rep) {
      CHECK(hipEventRecord(e0, 0));
      if (op == 0) hipLaunchKernelGGL(rate_kernel<0>, dim3(blocks), dim3(threads), 0, 0, d, iters, 1.0f);
      if (op == 1) hipLaunchKernelGGL(rate_kernel<1>, dim3(blocks), dim3(threads), 0, 0, d, iters, 1.0f);
      if (op == 2) hipLaunchKernelGGL(rate_kernel<2>, dim3(blocks), dim3(threads), 0, 0, d, iters, 1.0f);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1, 0));
      CHECK(hipEventSynchronize(e1));
      CHECK(hipEventElapsedTime(&ms, e0, e1));
    }
    const double ops = (double)n * iters * 8;
    const double per_s = ops / (ms * 1e-3);
    std::printf("%-10s %8.3f ms  %6.2f T lane-ops/s  %6.2f lane-ops/clk/CU at peak clock  %.3e s/lane-op\n",
                names[op], ms, per_s / 1e12, per_s / (khz * 1e3 * cus), 1.0 / per_s);
  }
  CHECK(hipFree(d));
  return 0;
}
