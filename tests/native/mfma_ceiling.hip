// mfma_ceiling.hip -- sustained MFMA rate on the device: every wavefront
// issues NA independent v_mfma_*_16x16x4 accumulations per step from
// registers (no memory traffic), for the FP64 and FP32 forms the XC build
// contractions use.  W is the launch bound's minimum waves per SIMD: with
// W = 1 the compiler keeps the accumulators in AccVGPRs (16 FP64
// accumulators: 128 of them), with W = 2 in ArchVGPRs, as the tiled
// contraction kernels do.  Prints one line per (form, accumulators, W,
// workgroups): the rate a contraction kernel could reach if it did nothing
// but MFMAs.  Built by __graft_entry__.build(); run by
// tests/test_gpu_mfma_ceiling.py.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

template <int NA, int W>
__global__ __launch_bounds__(256, W) void k_f64(double *out, int steps, double x) {
  d4 acc[NA];
  for (int i = 0; i < NA; ++i) acc[i] = d4{0.0, 0.0, 0.0, 0.0};
  double a = x + threadIdx.x, b = 2.0 * x;
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    a += 1e-9;
  }
  double t = 0.0;
  for (int i = 0; i < NA; ++i) t += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = t;
}

template <int NA, int W>
__global__ __launch_bounds__(256, W) void k_f32(double *out, int steps, double x) {
  f4 acc[NA];
  for (int i = 0; i < NA; ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
  float a = (float)x + threadIdx.x, b = 2.0f * (float)x;
  for (int s = 0; s < steps; ++s) {
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
    a += 1e-3f;
  }
  float t = 0.f;
  for (int i = 0; i < NA; ++i) t += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = t;
}

template <typename K>
static void run(const char *form, int na, int w, K kernel, int wgs, int steps, double *d) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  kernel<<<wgs, 256>>>(d, 16, 1.0);  // warm-up
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e0));
  kernel<<<wgs, 256>>>(d, steps, 1.0);
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double flops = (double)wgs * 4 * steps * na * 2.0 * 16 * 16 * 4;
  std::printf("%s accumulators %d bound %d workgroups %d: %.3f ms %.1f TFLOP/s\n", form, na, w,
              wgs, ms, flops / (ms * 1e-3) / 1e12);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
}

int main() {
  const int maxwgs = 4096, steps = 8000;
  double *d = nullptr;
  CHECK(hipMalloc(&d, sizeof(double) * maxwgs * 256));
  for (int wgs : {512, 2048, 4096}) {
    run("fp64", 16, 1, k_f64<16, 1>, wgs, steps, d);
    run("fp64", 16, 2, k_f64<16, 2>, wgs, steps, d);
    run("fp64", 8, 1, k_f64<8, 1>, wgs, steps, d);
    run("fp64", 8, 2, k_f64<8, 2>, wgs, steps, d);
    run("fp32", 16, 1, k_f32<16, 1>, wgs, steps, d);
    run("fp32", 16, 2, k_f32<16, 2>, wgs, steps, d);
  }
  CHECK(hipFree(d));
  return 0;
}
