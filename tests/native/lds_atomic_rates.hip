// lds_atomic_rates.hip -- two device rates the XC build kernels are shaped by:
//
// 1. ds_read_b128 of an MFMA operand: lane l reads 16 bytes of row l % 16,
//    4-float group l / 16, in the row layouts the contraction kernels could
//    use (all offsets inside a 32 KB LDS array).  Printed per layout: cycles
//    per wavefront instruction per CU (at 2.4 GHz).
// 2. FP64 atomic adds in the Vxc scatter's shape (each wavefront instruction
//    covers 4 rows x 16 consecutive doubles of a row-major matrix), into a
//    75 MB and a 2.3 MB matrix, plus plain stores of the same shape.
//
// Built by __graft_entry__.build(); run by tests/test_gpu_lds_atomic_rates.py.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f4 __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                 \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

// byte offset of lane's 16-byte read; every layout stays below 2.5 KB
__device__ int layout_offset(int layout, int lane) {
  const int r = lane & 15, g = lane >> 4;
  switch (layout) {
    case 0: return 16 * lane;                                       // linear
    case 1: return 64 * r + 16 * (g ^ ((r >> 1) & 3));              // 64 B rows, slot ^ (row/2 % 4)
    case 2: return 64 * r + 16 * (g ^ (((r >> 1) & 3) ^ ((r >> 3) & 1)));  // 64 B rows, other swizzle
    case 3: return 64 * r + 16 * g;                                 // 64 B rows
    case 4: return 144 * r + 16 * g;                                // rows padded to 144 B
    default: return 80 * r + 16 * g;                                // rows padded to 80 B
  }
}
const char *kLayoutName[] = {"linear", "64B rows, slot ^ (row/2 % 4)",
                             "64B rows, slot ^ (row/2 % 4) ^ (row/8 % 2)", "64B rows",
                             "144B rows", "80B rows"};
constexpr int kLayouts = 6;

__global__ __launch_bounds__(256) void k_lds(float *out, int layout, int iters) {
  __shared__ f4 s[2048];  // 32 KB
  for (int i = threadIdx.x; i < 2048; i += 256) s[i] = f4{(float)i, 1.f, 2.f, 3.f};
  __syncthreads();
  const int off = layout_offset(layout, threadIdx.x & 63) >> 4;  // f4 index < 160
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int it = 0; it < iters; ++it) {
    asm volatile("" ::: "memory");  // the reads stay in the loop
    acc += s[off] + s[off + 256] + s[off + 512] + s[off + 768];
  }
  out[blockIdx.x * 256 + threadIdx.x] = acc[0] + acc[1] + acc[2] + acc[3];
}

__device__ __forceinline__ unsigned mix(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// rows [row0, row0 + 4) x columns [col0, col0 + 16) per wavefront instruction,
// row0 <= nrows - 4, col0 <= ncols - 16
template <bool ATOMIC>
__global__ __launch_bounds__(256) void k_scatter(double *m, int nrows, int ncols, int reps) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int rep = 0; rep < reps; ++rep) {
    const unsigned h = mix(blockIdx.x * 977u + rep * 131u + wave * 7u);
    const int row0 = (int)(h % (unsigned)(nrows - 3));
    const int col0 = (int)((h >> 7) % (unsigned)(ncols / 16)) * 16;
    double *p = m + (long long)(row0 + (lane >> 4)) * ncols + col0 + (lane & 15);
    const double v = 1e-3 * (lane + 1);
    if (ATOMIC) atomicAdd(p, v);
    else *p = v;
  }
}

int main() {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  auto timed = [&](auto launch) {
    launch(true);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e0));
    launch(false);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms;
  };
  {
    const int wgs = 2048, iters = 4096;
    float *d = nullptr;
    CHECK(hipMalloc(&d, sizeof(float) * wgs * 256));
    for (int layout = 0; layout < kLayouts; ++layout) {
      const double ms = timed([&](bool warm) { k_lds<<<wgs, 256>>>(d, layout, warm ? 16 : iters); });
      const double per_cu = (double)wgs * 4 * iters * 4 / 256;  // wavefront reads per CU
      std::printf("lds layout %d (%s): %.3f ms, %.2f cycles per wavefront read per CU\n", layout,
                  kLayoutName[layout], ms, ms * 1e6 / per_cu * 2.4);
    }
    CHECK(hipFree(d));
  }
  {
    const int ncols = 3072, big = 3072, small = 96, wgs = 8192, reps = 64;
    double *m = nullptr;
    CHECK(hipMalloc(&m, sizeof(double) * (size_t)big * ncols));
    CHECK(hipMemset(m, 0, sizeof(double) * (size_t)big * ncols));
    const double n = (double)wgs * 256 * reps;
    for (int rows : {big, small}) {
      const double ma = timed([&](bool warm) { k_scatter<true><<<wgs, 256>>>(m, rows, ncols, warm ? 2 : reps); });
      const double ms = timed([&](bool warm) { k_scatter<false><<<wgs, 256>>>(m, rows, ncols, warm ? 2 : reps); });
      std::printf("scatter rows %d (%.1f MB): atomics %.3g per s, stores %.3g per s\n", rows,
                  rows * (double)ncols * 8 / 1e6, n / (ma * 1e-3), n / (ms * 1e-3));
    }
    CHECK(hipFree(m));
  }
  return 0;
}
