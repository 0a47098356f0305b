// grad.hip -- XC contribution to the nuclear gradient on gfx950.
//
// Replaces DFT::compute_xc_gradient_impl<d, Restricted> (include/occ/dft/
// dft.h:598-749) and kernels::process_grid_block_gradient<d, Restricted>
// (src/dft/dft_gradient_kernels.cpp:473-552).  The reference forms, per leaf,
// three m x m matrices V_k (accumulate_gradient_contribution_R, :26-90) and
// contracts them with 2 P (P = 2 D):
//   G(k, A) -= sum_{mu on A} sum_nu 2 P(mu,nu) V_k(mu,nu),
//   V_k(mu,nu) = sum_p [d_k phi_mu aow_nu + aow_k,mu phi_nu]          (GGA)
//   aow   = c0 phi + sum_j c_j d_j phi,  aow_k = c0 d_k phi + sum_j c_j d_kj phi
// with the build's coefficients c0 = w vrho / 2 and c_j = 2 w vsigma d_j rho
// (for LDA c_j = 0 and V_k = d_k phi^T (w vrho phi)).  Summing over nu first
// gives the same number with two leaf GEMMs instead of three m x m ones:
//   G(k, A) = -2 sum_{mu on A} sum_p [d_k phi_mu U_mu + aow_k,mu T_mu],
//   T = Phi P (the density's own product), U = Aow P.
// One workgroup per active leaf; each wavefront owns 32-AO chunks of the
// leaf and, for each 32-point block, forms T and U on the FP64 matrix cores
// (A operand Phi and Aow = c0 Phi + c . grad Phi built in registers, B operand
// P gathered through the AO map), then adds the integrand into per-lane sums
// of its AOs.  The per-AO sums go to gpart; a second kernel reduces them per
// atom in a fixed order (no atomics: the gradient repeats bit for bit).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace occxc {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename S> __device__ __forceinline__ d4 load4d(const S *p);
template <> __device__ __forceinline__ d4 load4d<double>(const double *p) {
  return *reinterpret_cast<const d4 *>(p);
}
template <> __device__ __forceinline__ d4 load4d<float>(const float *p) {
  const f4 v = *reinterpret_cast<const f4 *>(p);
  return d4{(double)v[0], (double)v[1], (double)v[2], (double)v[3]};
}

// Same leaf order over the XCDs as the build's contraction kernels.
__device__ __forceinline__ int xcd_leaf(int b, int n) {
  constexpr int kXcd = 8;
  const int x = b % kXcd, i = b / kXcd, q = n / kXcd, r = n % kXcd;
  return x * q + min(x, r) + i;
}

// MFMA layout (v_mfma_f64_16x16x4f64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l holds row l/16 + 4r,
// column l%16.  Point row i of a 16-point tile is fed as point
// 4(i%4) + i/4, so lane group g's accumulator rows are the 4 consecutive
// points 4g..4g+3 (vector loads in the integrand).  The contraction index
// mu = mc + 4g + s (step s = 0..3) is any permutation shared by A and B.
template <typename S, bool GGA>
__global__ __launch_bounds__(256, 2) void k_xc_grad(BuildArgs a, const S *__restrict__ ao,
                                                   const double *__restrict__ ao2,
                                                   double *__restrict__ gpart) {
  __shared__ double s_c[4][kMaxLeafPts];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  double *gp = gpart + 3LL * L.map_off;
  if (!a.leaf_active[leaf]) {
    // skipped by the max-rho rule (dft_gradient_kernels.cpp:488-491)
    for (int i = threadIdx.x; i < 3 * L.mp; i += blockDim.x) gp[i] = 0.0;
    return;
  }
  constexpr int NC = GGA ? 4 : 1;
  for (int i = threadIdx.x; i < L.np; i += blockDim.x)
#pragma unroll
    for (int q = 0; q < NC; ++q) s_c[q][i] = a.coef[q * a.ptot + L.pt_off + i];
  __syncthreads();
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int pa = 4 * (l16 & 3) + (l16 >> 2);
  const long long cs = (long long)L.mp * L.np;
  const S *phi = ao + L.ao_off;
  // gradient of phi: cache components 1..3 (GGA) or the extra block (LDA);
  // Hessian xx,xy,xz,yy,yz,zz: the extra block (GGA)
  const double *x2 = ao2 + L.ao_off / a.ncomp * (GGA ? 6 : 3);
  const double *Dm = a.D2 + MD.moff;
  const int *map = a.ao_map + L.map_off;
  for (int c = 32 * wave; c < L.mp; c += 128) {
    const int gn0 = map[c + l16], gn1 = map[c + 16 + l16];
    double gs[2][3];
#pragma unroll
    for (int j = 0; j < 2; ++j) gs[j][0] = gs[j][1] = gs[j][2] = 0.0;
    for (int p0 = 0; p0 < L.np; p0 += 32) {
      d4 accT[2][2], accU[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) accT[i][j] = accU[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      double ca[2][4];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) ca[i][q] = (GGA || q == 0) ? s_c[q][p0 + 16 * i + pa] : 0.0;
      for (int mc = 0; mc < L.mp; mc += 16) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int mu = mc + 4 * g + s4;
          const long long ro = (long long)mu * L.np + p0 + pa;
          const long long col = (long long)map[mu] * MD.nbf;
          const double f0 = (double)phi[ro], f1 = (double)phi[ro + 16];
          const double d0 = Dm[col + gn0], d1 = Dm[col + gn1];
          accT[0][0] = mma(f0, d0, accT[0][0]);
          accT[0][1] = mma(f0, d1, accT[0][1]);
          accT[1][0] = mma(f1, d0, accT[1][0]);
          accT[1][1] = mma(f1, d1, accT[1][1]);
          if (GGA) {
            const double u0 = ca[0][0] * f0 + ca[0][1] * (double)phi[cs + ro] +
                              ca[0][2] * (double)phi[2 * cs + ro] +
                              ca[0][3] * (double)phi[3 * cs + ro];
            const double u1 = ca[1][0] * f1 + ca[1][1] * (double)phi[cs + ro + 16] +
                              ca[1][2] * (double)phi[2 * cs + ro + 16] +
                              ca[1][3] * (double)phi[3 * cs + ro + 16];
            accU[0][0] = mma(u0, d0, accU[0][0]);
            accU[0][1] = mma(u0, d1, accU[0][1]);
            accU[1][0] = mma(u1, d0, accU[1][0]);
            accU[1][1] = mma(u1, d1, accU[1][1]);
          }
        }
      }
      // integrand at the lane's 4 points of each tile for its AO column
#pragma unroll
      for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const int pb = p0 + pt * 16 + 4 * g;
          const long long off = (long long)(c + nt * 16 + l16) * L.np + pb;
          d4 gx, gy, gz;
          if (GGA) {
            gx = load4d<S>(phi + cs + off);
            gy = load4d<S>(phi + 2 * cs + off);
            gz = load4d<S>(phi + 3 * cs + off);
          } else {
            gx = load4d<double>(x2 + off);
            gy = load4d<double>(x2 + cs + off);
            gz = load4d<double>(x2 + 2 * cs + off);
          }
          if (!GGA) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double t = 2.0 * s_c[0][pb + r] * accT[pt][nt][r];
              gs[nt][0] += gx[r] * t;
              gs[nt][1] += gy[r] * t;
              gs[nt][2] += gz[r] * t;
            }
          } else {
            const d4 hxx = load4d<double>(x2 + off), hxy = load4d<double>(x2 + cs + off),
                     hxz = load4d<double>(x2 + 2 * cs + off),
                     hyy = load4d<double>(x2 + 3 * cs + off),
                     hyz = load4d<double>(x2 + 4 * cs + off),
                     hzz = load4d<double>(x2 + 5 * cs + off);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const double c0 = s_c[0][pb + r], c1 = s_c[1][pb + r], c2 = s_c[2][pb + r],
                           c3 = s_c[3][pb + r];
              const double t = accT[pt][nt][r], u = accU[pt][nt][r];
              gs[nt][0] += gx[r] * u + (c0 * gx[r] + c1 * hxx[r] + c2 * hxy[r] + c3 * hxz[r]) * t;
              gs[nt][1] += gy[r] * u + (c0 * gy[r] + c1 * hxy[r] + c2 * hyy[r] + c3 * hyz[r]) * t;
              gs[nt][2] += gz[r] * u + (c0 * gz[r] + c1 * hxz[r] + c2 * hyz[r] + c3 * hzz[r]) * t;
            }
          }
        }
    }
    // sum over the four lane groups (the points) of each AO column
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double v = gs[nt][k];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (g == 0) gp[3 * (c + nt * 16 + l16) + k] = v;
      }
  }
}

// Meta-GGA tau terms (accumulate_gradient_contribution_R<2>,
// dft_gradient_kernels.cpp:152-172): V_k += sum_j d_kj phi^T ((vtau / 2) d_j phi)
// contracted with P, i.e. per AO mu and direction k
//   sum_p sum_j d_kj phi_mu (vtau / 2) T_j,mu,   T_j = (d_j Phi) P,
// with vtau / 2 = 2 c_tau (c_tau = w vtau / 4, a.coef_tau), into gpart in the
// normalisation of k_xc_grad's partials.  Same leaf / chunk / point-block
// structure as k_xc_grad: the three T_j of a chunk and point block on the
// FP64 matrix cores (A operand d_j phi from the AO cache, B operand P
// gathered through the AO map), the Hessian from ao2 in the integrand.
template <typename S>
__global__ __launch_bounds__(256, 2) void k_xc_grad_tau(BuildArgs a, const S *__restrict__ ao,
                                                       const double *__restrict__ ao2,
                                                       double *__restrict__ gpart) {
  __shared__ double s_t[kMaxLeafPts];
  const int leaf = xcd_leaf(blockIdx.x, gridDim.x);
  const LeafDev L = a.leaves[leaf];
  double *gp = gpart + 3LL * L.map_off;
  if (!a.leaf_active[leaf]) {
    for (int i = threadIdx.x; i < 3 * L.mp; i += blockDim.x) gp[i] = 0.0;
    return;
  }
  for (int i = threadIdx.x; i < L.np; i += blockDim.x)
    s_t[i] = 2.0 * a.coef_tau[L.pt_off + i];
  __syncthreads();
  const MolDev MD = a.mols[L.mol];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int pa = 4 * (l16 & 3) + (l16 >> 2);
  const long long cs = (long long)L.mp * L.np;
  const S *phi = ao + L.ao_off;
  const double *x2 = ao2 + L.ao_off / a.ncomp * 6;
  const double *Dm = a.D2 + MD.moff;
  const int *map = a.ao_map + L.map_off;
  for (int c = 32 * wave; c < L.mp; c += 128) {
    const int gn0 = map[c + l16], gn1 = map[c + 16 + l16];
    double gs[2][3];
#pragma unroll
    for (int j = 0; j < 2; ++j) gs[j][0] = gs[j][1] = gs[j][2] = 0.0;
    for (int p0 = 0; p0 < L.np; p0 += 32) {
      d4 acc[3][2][2];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[d][i][j] = d4{0.0, 0.0, 0.0, 0.0};
      for (int mc = 0; mc < L.mp; mc += 16) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int mu = mc + 4 * g + s4;
          const long long ro = (long long)mu * L.np + p0 + pa;
          const long long col = (long long)map[mu] * MD.nbf;
          const double d0 = Dm[col + gn0], d1 = Dm[col + gn1];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double f0 = (double)phi[(d + 1) * cs + ro];
            const double f1 = (double)phi[(d + 1) * cs + ro + 16];
            acc[d][0][0] = mma(f0, d0, acc[d][0][0]);
            acc[d][0][1] = mma(f0, d1, acc[d][0][1]);
            acc[d][1][0] = mma(f1, d0, acc[d][1][0]);
            acc[d][1][1] = mma(f1, d1, acc[d][1][1]);
          }
        }
      }
#pragma unroll
      for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const int pb = p0 + pt * 16 + 4 * g;
          const long long off = (long long)(c + nt * 16 + l16) * L.np + pb;
          const d4 hxx = load4d<double>(x2 + off), hxy = load4d<double>(x2 + cs + off),
                   hxz = load4d<double>(x2 + 2 * cs + off),
                   hyy = load4d<double>(x2 + 3 * cs + off),
                   hyz = load4d<double>(x2 + 4 * cs + off),
                   hzz = load4d<double>(x2 + 5 * cs + off);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double ct = s_t[pb + r];
            const double tx = ct * acc[0][pt][nt][r], ty = ct * acc[1][pt][nt][r],
                         tz = ct * acc[2][pt][nt][r];
            gs[nt][0] += hxx[r] * tx + hxy[r] * ty + hxz[r] * tz;
            gs[nt][1] += hxy[r] * tx + hyy[r] * ty + hyz[r] * tz;
            gs[nt][2] += hxz[r] * tx + hyz[r] * ty + hzz[r] * tz;
          }
        }
    }
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double v = gs[nt][k];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (g == 0) gp[3 * (c + nt * 16 + l16) + k] = v;
      }
  }
}

// grad[3 A + k] = -2 sum of gpart (plus gpart2, the beta spin of an
// unrestricted gradient, when given) over the entries of atom A, in a fixed
// order (per-thread strided sums, then a fixed tree).
__global__ __launch_bounds__(256) void k_grad_reduce(const double *__restrict__ gpart,
                                                     const double *__restrict__ gpart2,
                                                     const double *__restrict__ gpart3,
                                                     const double *__restrict__ gpart4,
                                                     const int *__restrict__ atom_off,
                                                     const int *__restrict__ atom_ent,
                                                     double *__restrict__ grad) {
  __shared__ double s_r[3][256];
  const int A = blockIdx.x;
  const int e0 = atom_off[A], e1 = atom_off[A + 1];
  double v[3] = {0.0, 0.0, 0.0};
  for (int i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
    const long long e = 3LL * atom_ent[i];
    v[0] += gpart[e];
    v[1] += gpart[e + 1];
    v[2] += gpart[e + 2];
    if (gpart2) {
      v[0] += gpart2[e];
      v[1] += gpart2[e + 1];
      v[2] += gpart2[e + 2];
    }
    if (gpart3) {
      v[0] += gpart3[e] + gpart4[e];
      v[1] += gpart3[e + 1] + gpart4[e + 1];
      v[2] += gpart3[e + 2] + gpart4[e + 2];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) s_r[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 3; ++k) s_r[k][threadIdx.x] += s_r[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) grad[3 * A + threadIdx.x] = -2.0 * s_r[threadIdx.x][0];
}

void check(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

void launch_xc_gradient_partials(const BuildArgs &a, const void *ao, bool ao_fp32,
                                 const double *ao2, double *gpart, hipStream_t s) {
  if (a.nleaves == 0) return;
  const bool gga = a.deriv != 0;
  if (ao_fp32) {
    const float *p = static_cast<const float *>(ao);
    if (gga) k_xc_grad<float, true><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
    else k_xc_grad<float, false><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
  } else {
    const double *p = static_cast<const double *>(ao);
    if (gga) k_xc_grad<double, true><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
    else k_xc_grad<double, false><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
  }
  check("k_xc_grad");
}

void launch_xc_gradient_tau(const BuildArgs &a, const void *ao, bool ao_fp32,
                            const double *ao2, double *gpart, hipStream_t s) {
  if (a.nleaves == 0) return;
  if (ao_fp32)
    k_xc_grad_tau<float><<<a.nleaves, 256, 0, s>>>(a, static_cast<const float *>(ao), ao2, gpart);
  else
    k_xc_grad_tau<double><<<a.nleaves, 256, 0, s>>>(a, static_cast<const double *>(ao), ao2, gpart);
  check("k_xc_grad_tau");
}

void launch_grad_reduce(const double *gpart, const double *gpart2, int natoms,
                        const int *atom_off, const int *atom_ent, double *grad, hipStream_t s,
                        const double *gpart3, const double *gpart4) {
  k_grad_reduce<<<natoms, 256, 0, s>>>(gpart, gpart2, gpart3, gpart4, atom_off, atom_ent, grad);
  check("k_grad_reduce");
}

void launch_xc_gradient(const BuildArgs &a, const void *ao, bool ao_fp32, const double *ao2,
                        double *gpart, int natoms, const int *atom_off, const int *atom_ent,
                        double *grad, hipStream_t s) {
  if (a.nleaves > 0) {
    const bool gga = a.deriv != 0;
    if (ao_fp32) {
      const float *p = static_cast<const float *>(ao);
      if (gga) k_xc_grad<float, true><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
      else k_xc_grad<float, false><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
    } else {
      const double *p = static_cast<const double *>(ao);
      if (gga) k_xc_grad<double, true><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
      else k_xc_grad<double, false><<<a.nleaves, 256, 0, s>>>(a, p, ao2, gpart);
    }
    check("k_xc_grad");
  }
  k_grad_reduce<<<natoms, 256, 0, s>>>(gpart, nullptr, nullptr, nullptr, atom_off, atom_ent,
                                       grad);
  check("k_grad_reduce");
}

} // namespace occxc
