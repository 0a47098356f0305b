// contract.hip -- MFMA contraction kernels of the XC build for gfx950.
//
// Per active leaf (n_b <= 256 points, m_b local AOs padded to 32):
//   k_density_mfma: T = Phi D_loc on the matrix cores (16x16x4 tiles, one
//     wavefront per 32-point block, D_loc columns gathered from the doubled
//     density through the leaf's AO map), then per point
//       rho = sum_nu Phi(nu,p) T(p,nu),  grad rho = 2 sum_nu grad Phi(nu,p) T(p,nu)
//     accumulated in FP64, the functional in FP64, and the Vxc coefficients
//       c0 = w vrho / 2,  c_k = 2 w vsigma d_k rho
//     (density.cpp:32-43, dft_kernels.cpp:8-70, xc_potential_matrix.cpp:19-41).
//   k_vxc_mfma: A = c0 Phi + c . grad Phi is built 32 AOs at a time in LDS,
//     K_loc = Phi^T A on the matrix cores, and K_loc^T is scattered into the
//     global accumulator with FP64 atomics (Vxc = K + K^T afterwards).
//
// MFMA operand layout for 16x16x4 (checked on the device): lane l supplies
// A[l%16][l/16] and B[l/16][l%16].  Accumulator register r of lane l holds
// D[row][l%16] with row = 4*(l/16)+r for the f32 instruction and
// row = l/16 + 4r for the f64 one (Mfma<T>::row).  The contraction index is
// permuted within each group of 16 so that lane group g uses k = 4g+s at step
// s (s = 0..3): the four values a lane needs along a point-fastest row are then
// contiguous (one vector load).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "contract.hpp"
#include "xc_functionals.hpp"

namespace occxc {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  using acc_t = d4;
  using vec_t = d4;
  __device__ static acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return g + 4 * r; }
  // values at rows row(g, 0..3) of a point-fastest array
  __device__ static d4 rows4(const double *base, int g) {
    return d4{base[g], base[g + 4], base[g + 8], base[g + 12]};
  }
};
template <> struct Mfma<float> {
  using acc_t = f4;
  using vec_t = f4;
  __device__ static acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int r) { return 4 * g + r; }
  __device__ static f4 rows4(const float *base, int g) {
    return *reinterpret_cast<const f4 *>(base + 4 * g);
  }
};

template <typename T>
__device__ __forceinline__ typename Mfma<T>::vec_t load4(const T *p) {
  return *reinterpret_cast<const typename Mfma<T>::vec_t *>(p);
}

__device__ __forceinline__ double sum16(double v) {
  v += __shfl_xor(v, 8, 16);
  v += __shfl_xor(v, 4, 16);
  v += __shfl_xor(v, 2, 16);
  v += __shfl_xor(v, 1, 16);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
__device__ double block_sum_256(double v, double *scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  const double r = scratch[0] + scratch[1] + scratch[2] + scratch[3];
  __syncthreads();
  return r;
}
__device__ double block_max_256(double v, double *scratch) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  const double r = fmax(fmax(scratch[0], scratch[1]), fmax(scratch[2], scratch[3]));
  __syncthreads();
  return r;
}

// ------------------------------------------------------------- density --
template <typename T>
__global__ __launch_bounds__(256) void k_density_mfma(BuildArgs a, const T *__restrict__ ao,
                                                      const T *__restrict__ D2t) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  __shared__ double s_q[4][kMaxLeafPts];
  __shared__ double s_red[4];
  const int leaf = blockIdx.x;
  const LeafDev L = a.leaves[leaf];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const T *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  const int nblk = L.np >> 5;
  for (int blk = wave; blk < nblk; blk += 4) {
    const int p0 = blk * 32;
    double part[2][4][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[i][r][q] = 0.0;
    for (int c = 0; c < L.mp; c += 32) {
      acc_t acc[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::zero();
      const long long gn0 = map[c + l16], gn1 = map[c + 16 + l16];
      for (int mc = 0; mc < L.mp; mc += 16) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int mu = mc + 4 * g + s;
          const long long gm = (long long)map[mu] * a.nbf;
          const T *row = phi + (long long)mu * L.np + p0 + l16;
          const T a0 = row[0], a1 = row[16];
          const T b0 = D2t[gm + gn0], b1 = D2t[gm + gn1];
          acc[0][0] = M::mma(a0, b0, acc[0][0]);
          acc[0][1] = M::mma(a0, b1, acc[0][1]);
          acc[1][0] = M::mma(a1, b0, acc[1][0]);
          acc[1][1] = M::mma(a1, b1, acc[1][1]);
        }
      }
      // rho and grad rho partial sums for this lane's points and AO
#pragma unroll
      for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const long long off = (long long)(c + nt * 16 + l16) * L.np + p0 + pt * 16;
          const auto f = M::rows4(phi + off, g);
#pragma unroll
          for (int r = 0; r < 4; ++r) part[pt][r][0] += (double)f[r] * (double)acc[pt][nt][r];
          if (deriv) {
            const auto fx = M::rows4(phi + cs + off, g);
            const auto fy = M::rows4(phi + 2 * cs + off, g);
            const auto fz = M::rows4(phi + 3 * cs + off, g);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double t = 2.0 * (double)acc[pt][nt][r];
              part[pt][r][1] += (double)fx[r] * t;
              part[pt][r][2] += (double)fy[r] * t;
              part[pt][r][3] += (double)fz[r] * t;
            }
          }
        }
    }
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = sum16(part[pt][r][q]);
          if (l16 == 0) s_q[q][p0 + pt * 16 + M::row(g, r)] = v;
        }
  }
  __syncthreads();

  // functional, weights, Vxc coefficients, leaf partials (one thread per point)
  const int p = threadIdx.x;
  const bool live = p < L.np, real = p < L.n;
  const double rho = live ? s_q[0][p] : 0.0;
  const double gx = live ? s_q[1][p] : 0.0, gy = live ? s_q[2][p] : 0.0,
               gz = live ? s_q[3][p] : 0.0;
  const double w = live ? a.pw[L.pt_off + p] : 0.0;
  const double nel = block_sum_256(real ? rho * w : 0.0, s_red);
  const double rmax = block_max_256(real ? rho : -1e300, s_red);
  const bool active = !(rmax < a.dthr);
  double e_sum = 0.0, vr_sum = 0.0, vs_sum = 0.0;
  if (active && live && rho >= kDensThreshold) {
    const double sg = fmax(gx * gx + gy * gy + gz * gz, kSigmaMin);
    for (int f = 0; f < a.nfunc; ++f) {
      double e, vr, vs;
      xc_point(a.func_ids[f], rho, sg, e, vr, vs);
      e_sum += e;
      vr_sum += vr;
      vs_sum += vs;
    }
  }
  const double exc = block_sum_256(active && real ? w * e_sum : 0.0, s_red);
  if (live) {
    const long long q = L.pt_off + p;
    a.coef[q] = active ? 0.5 * w * vr_sum : 0.0;
    a.coef[a.ptot + q] = active ? 2.0 * w * vs_sum * gx : 0.0;
    a.coef[2 * a.ptot + q] = active ? 2.0 * w * vs_sum * gy : 0.0;
    a.coef[3 * a.ptot + q] = active ? 2.0 * w * vs_sum * gz : 0.0;
  }
  if (threadIdx.x == 0) {
    a.leaf_part[2 * leaf] = active ? exc : 0.0;
    a.leaf_part[2 * leaf + 1] = nel;
    a.leaf_active[leaf] = active ? 1 : 0;
  }
}

// ----------------------------------------------------------------- Vxc --
template <typename T>
__global__ __launch_bounds__(256) void k_vxc_mfma(BuildArgs a, const T *__restrict__ ao, int ld) {
  using M = Mfma<T>;
  using acc_t = typename M::acc_t;
  constexpr int CH = 32;
  extern __shared__ double smem[];
  double(*s_c)[kMaxLeafPts] = reinterpret_cast<double(*)[kMaxLeafPts]>(smem);
  T *s_A = reinterpret_cast<T *>(smem + 4 * kMaxLeafPts);
  const int leaf = blockIdx.x;
  if (!a.leaf_active[leaf]) return;
  const LeafDev L = a.leaves[leaf];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const T *phi = ao + L.ao_off;
  const long long cs = (long long)L.mp * L.np;
  const int *map = a.ao_map + L.map_off;
  const bool deriv = a.deriv != 0;
  for (int i = threadIdx.x; i < L.np; i += blockDim.x)
    for (int k = 0; k < 4; ++k) s_c[k][i] = a.coef[k * a.ptot + L.pt_off + i];
  __syncthreads();
  const int ntile = L.mp >> 4;
  for (int c = 0; c < L.mp; c += CH) {
    for (int idx = threadIdx.x; idx < CH * L.np; idx += blockDim.x) {
      const int vl = idx / L.np, p = idx - vl * L.np;
      const long long o = (long long)(c + vl) * L.np + p;
      double v = s_c[0][p] * (double)phi[o];
      if (deriv)
        v += s_c[1][p] * (double)phi[cs + o] + s_c[2][p] * (double)phi[2 * cs + o] +
             s_c[3][p] * (double)phi[3 * cs + o];
      s_A[vl * ld + p] = static_cast<T>(v);
    }
    __syncthreads();
    for (int mt = wave; mt < ntile; mt += 4) {
      acc_t acc[2] = {M::zero(), M::zero()};
      const T *arow = phi + (long long)(mt * 16 + l16) * L.np + 4 * g;
      const T *b0row = s_A + l16 * ld + 4 * g;
      const T *b1row = s_A + (16 + l16) * ld + 4 * g;
      for (int pc = 0; pc < L.np; pc += 16) {
        const auto av = load4(arow + pc);
        const auto b0 = load4(b0row + pc);
        const auto b1 = load4(b1row + pc);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          acc[0] = M::mma(av[s], b0[s], acc[0]);
          acc[1] = M::mma(av[s], b1[s], acc[1]);
        }
      }
      // acc[nt][r] = K_loc(mu = mt*16 + row(g, r), nu = c + nt*16 + l16); store K^T
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const int nu = c + nt * 16 + l16;
        if (nu >= L.m) continue;
        const long long gn = map[nu];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int mu = mt * 16 + M::row(g, r);
          if (mu < L.m) atomicAdd(a.K + gn + (long long)map[mu] * a.nbf, (double)acc[nt][r]);
        }
      }
    }
    __syncthreads();
  }
}

__global__ void k_to_float(const double *__restrict__ x, float *__restrict__ y, long long n) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    y[k] = static_cast<float>(x[k]);
}

void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

static size_t vxc_smem(const ContractPlan &p, size_t elem, int *ld) {
  const int pad = (int)(16 / elem);
  *ld = p.max_np + pad;
  return sizeof(double) * 4 * kMaxLeafPts + elem * 32 * (size_t)(*ld);
}

ContractPlan make_contract_plan(const std::vector<LeafDev> &leaves, int nbf, int deriv,
                                int precision, hipStream_t s) {
  (void)deriv;
  (void)s;
  ContractPlan p;
  p.use_mfma = std::getenv("OCCXC_REFERENCE_KERNELS") == nullptr;
  for (const auto &L : leaves) p.max_np = std::max(p.max_np, L.np);
  if (precision != 0) {
    p.d2f.alloc((size_t)std::max(nbf, 1) * std::max(nbf, 1));
    p.bytes += p.d2f.bytes();
  }
  int ld = 0;
  const size_t sm64 = vxc_smem(p, sizeof(double), &ld), sm32 = vxc_smem(p, sizeof(float), &ld);
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_mfma<double>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm64));
  OCCXC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_vxc_mfma<float>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm32));
  return p;
}

void run_density(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
                 hipStream_t s) {
  if (a.nleaves == 0) return;
  if (precision == 0) {
    k_density_mfma<double><<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), a.D2);
  } else {
    const long long nn = (long long)a.nbf * a.nbf;
    k_to_float<<<(int)std::min<long long>((nn + 255) / 256, 4096), 256, 0, s>>>(
        a.D2, plan.d2f.get(), nn);
    check_launch("k_to_float");
    k_density_mfma<float><<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao),
                                                    plan.d2f.get());
  }
  check_launch("k_density_mfma");
}

void run_vxc(const ContractPlan &plan, const BuildArgs &a, const void *ao, int precision,
             hipStream_t s) {
  if (a.nleaves == 0) return;
  int ld = 0;
  if (precision == 0) {
    const size_t sm = vxc_smem(plan, sizeof(double), &ld);
    k_vxc_mfma<double><<<a.nleaves, 256, sm, s>>>(a, static_cast<const double *>(ao), ld);
  } else {
    const size_t sm = vxc_smem(plan, sizeof(float), &ld);
    k_vxc_mfma<float><<<a.nleaves, 256, sm, s>>>(a, static_cast<const float *>(ao), ld);
  }
  check_launch("k_vxc_mfma");
}

} // namespace occxc
