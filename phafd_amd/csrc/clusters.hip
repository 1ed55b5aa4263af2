// clusters.hip -- compute grid/clusters (ComputeGridClusters::end_of_step,
// src/compute_grid_clusters.cpp:66-228) on one subdomain holding the box.
//
// The reference marks every grid point beyond the threshold with the id
// i + (j*Nx + k)*Ny ((i, j, k) = (z, y, x)) and sweeps min-propagation over
// the 6 neighbours (periodic) until nothing changes, so a point ends with the
// smallest id of its connected component.  Here the components come from a
// concurrent union-find instead of sweeps (their number grows with the
// cluster's diameter), ordered by the key (id, index):
//   init     parent[p] = p inside a cluster, -1 outside
//   merge    every inside point unites with its +x, +y, +z neighbours; a
//            root is linked under the root of smaller key with atomicCAS, so
//            keys strictly decrease along parent pointers, every find ends,
//            and a root holds the smallest key of its tree
//   flatten  parent[p] = root
//   label    out[p] = id(root) (as a double, -1 outside)
// The root of a component is its member of smallest key, so the label is the
// component's smallest id (ids collide when Nz > Ny; the index breaks ties)
// whatever order the atomics take.
#include "internal.hpp"

namespace phafd {
namespace {

__device__ __forceinline__ int ld(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ int uf_find(int *parent, int p) {
  int q = ld(parent + p);
  while (q != p) {   // keys strictly decrease along parent pointers
    p = q;
    q = ld(parent + p);
  }
  return p;
}

struct Grid3 {
  int nx, ny, nz;
  // the reference's id i + (j*Nx + k)*Ny of flat point p = (i*Ny + j)*Nx + k
  __device__ int id(int p) const {
    const int k = p % nx, j = (p / nx) % ny, i = p / (nx * ny);
    return i + (j * nx + k) * ny;
  }
  __device__ unsigned long long key(int p) const {
    return ((unsigned long long)(unsigned)id(p) << 32) | (unsigned)p;
  }
};

__device__ void uf_union(const Grid3 &g, int *parent, int a, int b) {
  a = uf_find(parent, a);
  b = uf_find(parent, b);
  while (a != b) {
    if (g.key(a) < g.key(b)) { const int t = a; a = b; b = t; }   // link the larger key under the smaller
    const int old = atomicCAS(parent + a, a, b);
    if (old == a) return;
    a = uf_find(parent, old);
    b = uf_find(parent, b);
  }
}

template <class R>
__global__ void k_cl_init(size_t n, const R *__restrict__ phi, double thr, int gt, int *__restrict__ parent) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const double v = phi[p];
    const bool in = gt ? (v > thr) : (v < thr);
    parent[p] = in ? (int)p : -1;
  }
}

__global__ void k_cl_merge(Grid3 g, int *parent) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    if (ld(parent + p) < 0) continue;
    const int x = (int)(p % nx), y = (int)((p / nx) % ny), z = (int)(p / ((size_t)nx * ny));
    const size_t q[3] = {p - x + (x + 1 == nx ? 0 : x + 1),
                         p + (size_t)((y + 1 == ny ? 0 : y + 1) - y) * nx,
                         p + (size_t)((z + 1 == nz ? 0 : z + 1) - z) * nx * ny};
    for (int d = 0; d < 3; d++)
      if (q[d] != p && ld(parent + q[d]) >= 0) uf_union(g, parent, (int)p, (int)q[d]);
  }
}

__global__ void k_cl_flatten(size_t n, int *parent) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
    if (ld(parent + p) >= 0) parent[p] = uf_find(parent, (int)p);
}

__global__ void k_cl_label(Grid3 g, const int *__restrict__ parent, double *__restrict__ out) {
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int r = parent[p];
    out[p] = r < 0 ? -1.0 : (double)g.id(r);
  }
}

int grid_blocks(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 16384); }

}  // namespace

template <class R>
void launch_grid_clusters(int nx, int ny, int nz, const R *phi, double thr, bool gt, int *parent, double *out,
                          hipStream_t s) {
  const size_t n = (size_t)nx * ny * nz;
  const int b = grid_blocks(n);
  const Grid3 g{nx, ny, nz};
  hipLaunchKernelGGL(k_cl_init<R>, dim3(b), dim3(256), 0, s, n, phi, thr, gt ? 1 : 0, parent);
  hipLaunchKernelGGL(k_cl_merge, dim3(b), dim3(256), 0, s, g, parent);
  hipLaunchKernelGGL(k_cl_flatten, dim3(b), dim3(256), 0, s, n, parent);
  hipLaunchKernelGGL(k_cl_label, dim3(b), dim3(256), 0, s, g, parent, out);
  HIPCHK(hipGetLastError());
}
template void launch_grid_clusters<double>(int, int, int, const double *, double, bool, int *, double *,
                                           hipStream_t);
template void launch_grid_clusters<float>(int, int, int, const float *, double, bool, int *, double *, hipStream_t);

}  // namespace phafd
