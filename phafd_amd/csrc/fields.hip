// fields.hip -- pointwise grid kernels of the Model-B step (FP64).
//
// All arithmetic is written in the operation order of the reference so that,
// built with -ffp-contract=off, a kernel given the same inputs as the
// reference loop reproduces it (log/exp aside, which may differ by an ulp).
#include "internal.hpp"
#include "spectral_ops.hpp"

namespace phafd {
namespace {

constexpr int TPB = 256;

// grad multiply, src/fixgrid_conjugate.cpp:80-105, for the three components.
// K layout index = (kx*nz + kz)*ny + ky.
__global__ void k_grad_multiply(const double2 *__restrict__ ph, double2 *__restrict__ gx,
                                double2 *__restrict__ gy, double2 *__restrict__ gz, int hx,
                                int ny, int nz, SpecConst c) {
  const size_t M = (size_t)hx * nz * ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(i % ny);
    const size_t t = i / ny;
    const int kz = (int)(t % nz);
    const int kx = (int)(t / nz);
    const double2 v = ph[i];
    gx[i] = grad_mode(v, c.dqx, (double)kx);
    gz[i] = grad_mode(v, c.dqy_kz, (double)signed_k(kz, nz));   // ft_gradphi[1] -> gradphi[2]
    gy[i] = grad_mode(v, c.dqz_ky, (double)signed_k(ky, ny));   // ft_gradphi[2] -> gradphi[1]
  }
}

// Conjugate::update (src/conjugate.cpp:163-327, one rank) as a pointwise map.
// Bulk kx in [1, Nx/2): every mode updates itself.  Planes kx = 0, Nx/2:
// "primary" modes (the ones single()/line_single() update) write themselves
// and conj() into their mirror (-ky, -kz); mirror threads do nothing; the
// eight self-conjugate modes use real_update; the origin is only rescaled.
__device__ __forceinline__ void update_index(double2 *__restrict__ ph, const double2 *__restrict__ mu,
                                             const double2 *__restrict__ noise, size_t idx, int hx,
                                             int ny, int nz, const UpdateParams &u) {
  const int A = ny, B = nz, H = hx - 1;
  const int i = (int)(idx % ny);   // ky
  const size_t t = idx / ny;
  const int j = (int)(t % nz);     // kz
  const int k = (int)(t / nz);     // kx
  const double2 nz0 = make_double2(0.0, 0.0);
  if (k != 0 && k != H) {
    ph[idx] = mode_update(ph[idx], mu[idx], noise ? noise[idx] : nz0, i, j, k, A, B, u);
    return;
  }
  const bool self = (i == 0 || i == A / 2) && (j == 0 || j == B / 2);
  if (self) {
    if (k == 0 && i == 0 && j == 0) {   // origin_update
      const double2 v = ph[idx];
      ph[idx] = make_double2(v.x * u.normalization, v.y * u.normalization);
    } else {
      ph[idx] = mode_update(ph[idx], mu[idx], noise ? noise[idx] : nz0, i, j, k, A, B, u);
    }
    return;
  }
  const bool primary = (j >= 1 && j < B / 2) || (j == 0 && i >= 1 && i < A / 2) ||
                       (j == B / 2 && i > A / 2);
  if (!primary) return;
  const double2 r = mode_update(ph[idx], mu[idx], noise ? noise[idx] : nz0, i, j, k, A, B, u);
  ph[idx] = r;
  const int im = (A - i) % A, jm = (B - j) % B;
  ph[((size_t)k * nz + jm) * ny + im] = make_double2(r.x, -r.y);
}

__global__ void k_update(double2 *__restrict__ ph, const double2 *__restrict__ mu,
                         const double2 *__restrict__ noise, int hx, int ny, int nz,
                         UpdateParams u) {
  const size_t M = (size_t)hx * nz * ny;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < M;
       idx += (size_t)gridDim.x * blockDim.x)
    update_index(ph, mu, noise, idx, hx, ny, nz, u);
}

// only the planes kx = 0 and kx = Nx/2 (the bulk is fused into the z pass)
__global__ void k_update_edges(double2 *__restrict__ ph, const double2 *__restrict__ mu,
                               const double2 *__restrict__ noise, int hx, int ny, int nz,
                               UpdateParams u) {
  const size_t P = (size_t)nz * ny;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < 2 * P;
       w += (size_t)gridDim.x * blockDim.x) {
    const size_t idx = w < P ? w : (size_t)(hx - 1) * P + (w - P);
    update_index(ph, mu, noise, idx, hx, ny, nz, u);
  }
}

__global__ void k_floryhuggins(size_t n, double *__restrict__ phi, double *__restrict__ mu,
                               double tv, double chi, unsigned long long *counters) {
  __shared__ unsigned cnt[2];
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
  __syncthreads();
  unsigned nl = 0, ng = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double p = phi[i];
    if (p <= 0) {
      p = -p;
      nl++;
      phi[i] = p;
    } else if (p >= 1) {
      p = 1 - (p - 1);
      ng++;
      phi[i] = p;
    }
    mu[i] += tv * (log(p / (1 - p)) + chi * (1 - 2 * p));
  }
  if (nl) atomicAdd(&cnt[0], nl);
  if (ng) atomicAdd(&cnt[1], ng);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (cnt[0]) atomicAdd(&counters[0], (unsigned long long)cnt[0]);
    if (cnt[1]) atomicAdd(&counters[1], (unsigned long long)cnt[1]);
  }
}

__global__ void k_fill_zero(uint4 *p, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}
__global__ void k_fill_zero_bytes(unsigned char *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Flory-Huggins + square-gradient free energy (defined by this build):
// sum of phi ln phi + (1-phi) ln(1-phi) + chi phi (1-phi) and |grad phi|^2,
// block partial sums in FP64 then one atomicAdd per block.
__global__ void k_energy_fh(size_t n, const double *__restrict__ phi, const double *__restrict__ gx,
                            const double *__restrict__ gy, const double *__restrict__ gz,
                            double tv, double chi, double tvg, double *out2) {
  __shared__ double s0[TPB], s1[TPB];
  double a = 0, b = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double p = phi[i];
    if (p <= 0) p = -p;
    else if (p >= 1) p = 1 - (p - 1);
    if (tv != 0.0) a += p * log(p) + (1 - p) * log(1 - p) + chi * p * (1 - p);
    if (gx) b += gx[i] * gx[i] + gy[i] * gy[i] + gz[i] * gz[i];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int st = TPB / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      s0[threadIdx.x] += s0[threadIdx.x + st];
      s1[threadIdx.x] += s1[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out2[0], tv * s0[0]);
    atomicAdd(&out2[1], 0.5 * tvg * s1[0]);
  }
}

int grid_for(size_t n) {
  size_t b = (n + TPB - 1) / TPB;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

void grad_multiply(const FFTPlan &p, const double2 *phihat, double2 *gx, double2 *gy,
                   double2 *gz, SpecConst c, hipStream_t s) {
  probe_begin(K_GRAD_MUL, s);
  hipLaunchKernelGGL(k_grad_multiply, dim3(grid_for(p.M())), dim3(TPB), 0, s, phihat, gx, gy, gz,
                     p.hx, p.ny, p.nz, c);
  probe_end(K_GRAD_MUL, s, 64.0 * p.M());
  HIPCHK(hipGetLastError());
}

void spectral_update(const FFTPlan &p, double2 *phihat, const double2 *muhat,
                     const double2 *noise_k, UpdateParams u, hipStream_t s) {
  probe_begin(K_UPDATE, s);
  hipLaunchKernelGGL(k_update, dim3(grid_for(p.M())), dim3(TPB), 0, s, phihat, muhat, noise_k,
                     p.hx, p.ny, p.nz, u);
  probe_end(K_UPDATE, s, (noise_k ? 64.0 : 48.0) * p.M());
  HIPCHK(hipGetLastError());
}

void spectral_update_edges(const FFTPlan &p, double2 *phihat, const double2 *muhat,
                           const double2 *noise_k, UpdateParams u, hipStream_t s) {
  hipLaunchKernelGGL(k_update_edges, dim3(grid_for(2 * (size_t)p.ny * p.nz)), dim3(TPB), 0, s, phihat,
                     muhat, noise_k, p.hx, p.ny, p.nz, u);
  HIPCHK(hipGetLastError());
}

void floryhuggins(size_t n, double *phi, double *mu, double tv, double chi,
                  unsigned long long *counters, hipStream_t s) {
  probe_begin(K_FH, s);
  hipLaunchKernelGGL(k_floryhuggins, dim3(grid_for(n)), dim3(TPB), 0, s, n, phi, mu, tv, chi,
                     counters);
  probe_end(K_FH, s, 24.0 * n);
  HIPCHK(hipGetLastError());
}

void fill_zero(void *p, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0 && (bytes & 15) == 0) {
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(bytes / 16)), dim3(TPB), 0, s,
                       reinterpret_cast<uint4 *>(p), bytes / 16);
  } else {
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(grid_for(bytes)), dim3(TPB), 0, s,
                       reinterpret_cast<unsigned char *>(p), bytes);
  }
  HIPCHK(hipGetLastError());
}

void reduce_energy_fh(size_t n, const double *phi, const double *gx, const double *gy,
                      const double *gz, double tv, double chi, double tvg, double *out2,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_energy_fh, dim3(grid_for(n) > 1024 ? 1024 : grid_for(n)), dim3(TPB), 0, s,
                     n, phi, gx, gy, gz, tv, chi, tvg, out2);
  HIPCHK(hipGetLastError());
}

}  // namespace phafd
