// setup.hip -- the grid hierarchy and shell screening of a context on the
// device: SpatialGridHierarchy (spatial_grid_hierarchy.cpp:28-147) and the
// per-leaf significant-shell lists of compute_K_dft (dft.h:297-320).
//
// Per molecule: bounding box, 21-bit-per-axis Morton codes, a stable radix
// sort of (code, index) (the order std::sort gives the pairs), leaves of 128
// points (the last one 129..256, or the whole grid when it has <= 256
// points), centroid / max-distance spheres, and the screening test
// |origin - centre| - extent < radius for every (leaf, shell).  Every value
// that decides a leaf or a screening outcome is formed with the same
// operations in the same order as the host restatement (sums in point order,
// no fused multiply-adds), so the device layout is bit-identical to the host
// one.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace occxc {

namespace {

constexpr int kLeafTarget = 128;  // SpatialHierarchySettings: target 128, max 256
constexpr int kLeafMax = 256;

void check(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ unsigned long long spread3(unsigned long long v) {
  v &= 0x1fffffULL;
  unsigned long long r = 0;
  for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ULL) << (3 * i);
  return r;
}

// bounding box: per-block min/max, then one block folds them (min and max
// are exact in any order)
__global__ __launch_bounds__(256) void k_bbox_part(const double *__restrict__ xyz, long long n,
                                                   double *__restrict__ part) {
  __shared__ double s[6][256];
  double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    for (int k = 0; k < 3; ++k) {
      const double x = xyz[3 * i + k];
      v[k] = fmin(v[k], x);
      v[3 + k] = fmax(v[3 + k], x);
    }
  for (int k = 0; k < 6; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 6; ++k)
        s[k][threadIdx.x] = k < 3 ? fmin(s[k][threadIdx.x], s[k][threadIdx.x + w])
                                  : fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// box = {min x, min y, min z, extent x, y, z}; an extent below 1e-10 becomes 1
__global__ void k_bbox_final(const double *__restrict__ part, int nb, double *__restrict__ box) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) {
      mn[k] = fmin(mn[k], part[6 * b + k]);
      mx[k] = fmax(mx[k], part[6 * b + 3 + k]);
    }
  for (int k = 0; k < 3; ++k) {
    double e = mx[k] - mn[k];
    if (e < 1e-10) e = 1.0;
    box[k] = mn[k];
    box[3 + k] = e;
  }
}

__global__ __launch_bounds__(256) void k_morton(const double *__restrict__ xyz, long long n,
                                                const double *__restrict__ box,
                                                unsigned long long *__restrict__ code,
                                                int *__restrict__ idx) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  const double scale = (double)((1ULL << 21) - 1);
  unsigned long long q[3];
  for (int k = 0; k < 3; ++k) {
    double t = (xyz[3 * i + k] - box[k]) / box[3 + k];
    t = fmin(fmax(t, 0.0), 1.0);
    q[k] = (unsigned long long)(t * scale);
  }
  code[i] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
  idx[i] = (int)i;
}

__device__ __forceinline__ long long leaf_count(long long n, int nleaves, int li) {
  return li < nleaves - 1 ? kLeafTarget : n - (long long)kLeafTarget * li;
}

// One thread per leaf: centroid summed in point order, then the largest
// distance to it (no fused multiply-adds, as on the host).
__global__ __launch_bounds__(256) void k_leaf_geom(const double *__restrict__ xyz,
                                                   const int *__restrict__ perm, long long n,
                                                   int nleaves, double *__restrict__ center,
                                                   double *__restrict__ radius) {
#pragma clang fp contract(off)
  const int li = blockIdx.x * 256 + threadIdx.x;
  if (li >= nleaves) return;
  const long long off = (long long)kLeafTarget * li, cnt = leaf_count(n, nleaves, li);
  double c[3] = {0.0, 0.0, 0.0};
  for (long long i = 0; i < cnt; ++i) {
    const long long o = perm[off + i];
    for (int k = 0; k < 3; ++k) c[k] += xyz[3 * o + k];
  }
  for (int k = 0; k < 3; ++k) c[k] /= (double)cnt;
  double r = 0.0;
  for (long long i = 0; i < cnt; ++i) {
    const long long o = perm[off + i];
    const double dx = xyz[3 * o] - c[0], dy = xyz[3 * o + 1] - c[1], dz = xyz[3 * o + 2] - c[2];
    r = fmax(r, sqrt(dx * dx + dy * dy + dz * dz));
  }
  for (int k = 0; k < 3; ++k) center[3 * li + k] = c[k];
  radius[li] = r;
}

__device__ __forceinline__ bool significant(const ShellDev &sh, double ext, const double *c,
                                            double r) {
#pragma clang fp contract(off)
  const double dx = sh.x - c[0], dy = sh.y - c[1], dz = sh.z - c[2];
  const double dist = sqrt(dx * dx + dy * dy + dz * dz);
  return dist - ext < r;
}

// One wavefront per leaf: significant shells and their AO count.
__global__ __launch_bounds__(64) void k_screen_count(const double *__restrict__ center,
                                                     const double *__restrict__ radius,
                                                     int nleaves, const ShellDev *__restrict__ sh,
                                                     const double *__restrict__ ext, int nsh,
                                                     int *__restrict__ nsig, int *__restrict__ nao) {
  const int li = blockIdx.x;
  if (li >= nleaves) return;
  const double c[3] = {center[3 * li], center[3 * li + 1], center[3 * li + 2]};
  const double r = radius[li];
  int ns = 0, na = 0;
  for (int s = threadIdx.x; s < nsh; s += 64) {
    const ShellDev S = sh[s];
    if (significant(S, ext[s], c, r)) {
      ++ns;
      na += S.size;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    ns += __shfl_down(ns, o, 64);
    na += __shfl_down(na, o, 64);
  }
  if (threadIdx.x == 0) {
    nsig[li] = ns;
    nao[li] = na;
  }
}

// One wavefront per active leaf: its significant shells in shell order
// (leaf_shells, global shell index sb + s) and their AOs (ao_map, the
// molecule's own AO index bf0 + q), padding AOs mapped to AO 0; then its
// points padded to np with copies of the first point at weight 0.
__global__ __launch_bounds__(64) void k_fill_leaf(const LeafDev *__restrict__ leaves,
                                                  const FillLeaf *__restrict__ fill,
                                                  const double *__restrict__ center,
                                                  const double *__restrict__ radius,
                                                  const ShellDev *__restrict__ shells,
                                                  const double *__restrict__ ext,
                                                  const double *__restrict__ xyz,
                                                  const double *__restrict__ w,
                                                  const int *__restrict__ perm,
                                                  int *__restrict__ leaf_shells,
                                                  int *__restrict__ ao_map, double *__restrict__ px,
                                                  double *__restrict__ py, double *__restrict__ pz,
                                                  double *__restrict__ pw, int *__restrict__ porig) {
  const LeafDev L = leaves[blockIdx.x];
  const FillLeaf F = fill[blockIdx.x];
  const int lane = threadIdx.x;
  const double c[3] = {center[3 * F.src_leaf], center[3 * F.src_leaf + 1],
                       center[3 * F.src_leaf + 2]};
  const double r = radius[F.src_leaf];
  int nsig = 0, nao = 0;
  for (int s0 = 0; s0 < F.nsh; s0 += 64) {
    const int s = s0 + lane;
    ShellDev S{};
    bool sig = false;
    if (s < F.nsh) {
      S = shells[F.sb + s];
      sig = significant(S, ext[F.sb + s], c, r);
    }
    const unsigned long long ball = __ballot(sig);
    const unsigned long long below = lane ? (ball & ((1ULL << lane) - 1ULL)) : 0ULL;
    const int rank = __popcll(below);
    // AO offset: inclusive scan of the sizes of the significant shells
    int sz = sig ? S.size : 0, scan = sz;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(scan, o, 64);
      if (lane >= o) scan += t;
    }
    const int total = __shfl(scan, 63, 64);
    if (sig) {
      leaf_shells[L.sh_off + nsig + rank] = F.sb + s;
      const int a0 = nao + scan - sz;
      for (int q = 0; q < S.size; ++q) ao_map[L.map_off + a0 + q] = S.bf0 + q;
    }
    nsig += __popcll(ball);
    nao += total;
  }
  for (int q = L.m + lane; q < L.mp; q += 64) ao_map[L.map_off + q] = 0;
  for (int p = lane; p < L.np; p += 64) {
    const long long src = perm[F.src_pt + (p < L.n ? p : 0)] + F.mol_pt;
    const long long q = L.pt_off + p;
    px[q] = xyz[3 * src];
    py[q] = xyz[3 * src + 1];
    pz[q] = xyz[3 * src + 2];
    pw[q] = p < L.n ? w[src] : 0.0;
    porig[q] = p < L.n ? perm[F.src_pt + p] : -1;
  }
}

} // namespace

int hierarchy_leaf_count(long long n) {
  if (n <= 0) return 0;
  if (n <= kLeafMax) return 1;
  return (int)((n - kLeafMax + kLeafTarget - 1) / kLeafTarget) + 1;
}

size_t sort_temp_bytes(long long nmax) {
  size_t bytes = 0;
  if (nmax <= 0) return 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (unsigned long long *)nullptr,
                                         (unsigned long long *)nullptr, (int *)nullptr,
                                         (int *)nullptr, (int)nmax, 0, 63) != hipSuccess)
    throw std::runtime_error("hierarchy: radix sort size query failed");
  return bytes;
}

void launch_hierarchy(const double *xyz, long long n, int nleaves, double *part, double *box,
                      unsigned long long *keys_in, unsigned long long *keys_out, int *idx_in,
                      int *perm, void *temp, size_t temp_bytes, double *center, double *radius,
                      hipStream_t s) {
  if (n <= 0) return;
  const int nb = (int)std::min<long long>((n + 255) / 256, 256);
  k_bbox_part<<<nb, 256, 0, s>>>(xyz, n, part);
  k_bbox_final<<<1, 64, 0, s>>>(part, nb, box);
  k_morton<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(xyz, n, box, keys_in, idx_in);
  check("hierarchy: Morton codes");
  size_t tb = temp_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(temp, tb, keys_in, keys_out, idx_in, perm, (int)n, 0, 63,
                                         s) != hipSuccess)
    throw std::runtime_error("hierarchy: radix sort failed");
  k_leaf_geom<<<(nleaves + 255) / 256, 256, 0, s>>>(xyz, perm, n, nleaves, center, radius);
  check("hierarchy: leaf spheres");
}

void launch_screen_count(const double *center, const double *radius, int nleaves,
                         const ShellDev *shells, const double *ext, int nsh, int *nsig, int *nao,
                         hipStream_t s) {
  if (nleaves <= 0) return;
  k_screen_count<<<nleaves, 64, 0, s>>>(center, radius, nleaves, shells, ext, nsh, nsig, nao);
  check("screening: count");
}

void launch_fill_leaves(const LeafDev *leaves, const FillLeaf *fill, int nactive,
                        const double *center, const double *radius, const ShellDev *shells,
                        const double *ext, const double *xyz, const double *w, const int *perm,
                        int *leaf_shells, int *ao_map, double *px, double *py, double *pz,
                        double *pw, int *porig, hipStream_t s) {
  if (nactive <= 0) return;
  k_fill_leaf<<<nactive, 64, 0, s>>>(leaves, fill, center, radius, shells, ext, xyz, w, perm,
                                     leaf_shells, ao_map, px, py, pz, pw, porig);
  check("screening: leaf tables");
}

} // namespace occxc
