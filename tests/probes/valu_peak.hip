// FP32/FP64 FMA issue-rate microbenchmark for the VALU roofline of the
// backprojection kernels: long straight-line bodies of independent FMA chains
// (4, 8, 16 per lane), no memory traffic inside the timed loop.  Build and run:
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         tests/probes/valu_peak.hip -o tests/probes/valu_peak && tests/probes/valu_peak
// Each line prints the rate with an FMA counted as 2 FLOPs.
#include <hip/hip_runtime.h>
#include <stdio.h>
#define CHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("err %s\n", hipGetErrorString(e_)); return 1; } } while (0)
template <typename T, int CH>
__global__ __launch_bounds__(256) void k_fma(T* out, int iters) {
  T x[CH];
  const T a = (T)(0.9999 + threadIdx.x * 1e-9), b = (T)(1e-3 + threadIdx.x * 1e-9);
#pragma unroll
  for (int k = 0; k < CH; ++k) x[k] = (T)(threadIdx.x * 1e-6 + k);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int k = 0; k < CH; ++k) x[k] = fma(x[k], a, b);
  }
  T s = 0;
#pragma unroll
  for (int k = 0; k < CH; ++k) s += x[k];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
template <typename T, int CH>
int run(const char* name, T* o, int blocks, int iters) {
  hipEvent_t e0, e1; CHK(hipEventCreate(&e0)); CHK(hipEventCreate(&e1));
  float ms = 0;
  for (int rep = 0; rep < 3; ++rep) {
    CHK(hipEventRecord(e0));
    hipLaunchKernelGGL((k_fma<T, CH>), dim3(blocks), dim3(256), 0, 0, o, iters);
    CHK(hipEventRecord(e1)); CHK(hipEventSynchronize(e1));
    CHK(hipEventElapsedTime(&ms, e0, e1));
  }
  double fmas = (double)blocks * 256 * iters * 8 * CH;
  printf("%s chains=%d: %.3f ms  %.1f TFLOP/s (FMA=2)\n", name, CH, ms, 2 * fmas / ms * 1e-9);
  return 0;
}
int main() {
  const int blocks = 256 * 16, iters = 2000;
  float* of; double* od;
  CHK(hipMalloc(&of, blocks * 256 * sizeof(float)));
  CHK(hipMalloc(&od, blocks * 256 * sizeof(double)));
  run<float, 4>("fp32", of, blocks, iters);
  run<float, 8>("fp32", of, blocks, iters);
  run<float, 16>("fp32", of, blocks, iters);
  run<double, 8>("fp64", od, blocks, iters);
  hipDeviceProp_t p; CHK(hipGetDeviceProperties(&p, 0));
  printf("device %s CUs %d clock %d kHz\n", p.gcnArchName, p.multiProcessorCount, p.clockRate);
  return 0;
}
