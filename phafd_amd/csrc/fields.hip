// fields.hip -- pointwise grid kernels of the Model-B step, FP64 and FP32.
//
// All arithmetic is written in the operation order of the reference so that,
// built with -ffp-contract=off, a kernel given the same inputs as the
// reference loop reproduces it (log/exp aside, which may differ by an ulp).
// The FP32 instances round every operand to float; energies and the k = 0
// mode stay in FP64.
#include "internal.hpp"
#include "spectral_ops.hpp"

namespace phafd {
namespace {

constexpr int TPB = 256;

__device__ __forceinline__ float flog(float x) { return logf(x); }
__device__ __forceinline__ double flog(double x) { return log(x); }

// grad multiply, src/fixgrid_conjugate.cpp:80-105, for the three components.
// K layout index = (kx*nz + kz)*ny + ky.
template <class C>
__global__ void k_grad_multiply(const C *__restrict__ ph, C *__restrict__ gx, C *__restrict__ gy,
                                C *__restrict__ gz, int hx, int ny, int nz, SpecConst c, Slab sl) {
  using R = typename C::value_type;
  const size_t M = (size_t)hx * nz * ny;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < M * sl.n;
       w += (size_t)gridDim.x * blockDim.x) {
    const size_t loc = w % M, i = (w / M) * sl.sc + loc;
    const int ky = (int)(loc % ny);
    const size_t t = loc / ny;
    const int kz = (int)(t % nz);
    const int kx = (int)(t / nz);
    const C v = ph[i];
    gx[i] = grad_mode(v, (R)c.dqx, (R)kx);
    gz[i] = grad_mode(v, (R)c.dqy_kz, (R)signed_k(kz, nz));   // ft_gradphi[1] -> gradphi[2]
    gy[i] = grad_mode(v, (R)c.dqz_ky, (R)signed_k(ky, ny));   // ft_gradphi[2] -> gradphi[1]
  }
}

// Conjugate::update (src/conjugate.cpp:163-327, one rank) as a pointwise map.
// Bulk kx in [1, Nx/2): every mode updates itself.  Planes kx = 0, Nx/2:
// "primary" modes (the ones single()/line_single() update) write themselves
// and conj() into their mirror (-ky, -kz); mirror threads do nothing; the
// eight self-conjugate modes use real_update; the origin is only rescaled
// (or, with `origin`, set to the FP64 sum there times the normalisation).
template <bool DEV, class C>
__device__ __forceinline__ void update_index(C *__restrict__ ph, const C *__restrict__ mu,
                                             const C *__restrict__ noise, size_t idx, int hx, int ny,
                                             int nz, const UpdateParams &u, const double *origin) {
  using R = typename C::value_type;
  const int A = ny, B = nz, H = hx - 1;
  const int i = (int)(idx % ny);   // ky
  const size_t t = idx / ny;
  const int j = (int)(t % nz);     // kz
  const int k = (int)(t / nz);     // kx
  if (k != 0 && k != H) {
    ph[idx] = mode_update(ph[idx], mu[idx], mode_noise<DEV>(noise, idx, u, i, j, k, A, B, false), i, j, k, A, B, u);
    return;
  }
  const bool self = (i == 0 || i == A / 2) && (j == 0 || j == B / 2);
  if (self) {
    if (k == 0 && i == 0 && j == 0) {   // origin_update
      if (origin) {
        ph[idx] = cpx<C>((R)(origin[0] * u.normalization), R(0));
      } else {
        const C v = ph[idx];
        ph[idx] = cpx<C>(v.x * (R)u.normalization, v.y * (R)u.normalization);
      }
    } else {   // real_update
      ph[idx] = mode_update(ph[idx], mu[idx], mode_noise<DEV>(noise, idx, u, i, j, k, A, B, true), i, j, k, A, B, u);
    }
    return;
  }
  const bool primary = (j >= 1 && j < B / 2) || (j == 0 && i >= 1 && i < A / 2) ||
                       (j == B / 2 && i > A / 2);
  if (!primary) return;
  const C r = mode_update(ph[idx], mu[idx], mode_noise<DEV>(noise, idx, u, i, j, k, A, B, false), i, j, k, A, B, u);
  ph[idx] = r;
  const int im = (A - i) % A, jm = (B - j) % B;
  ph[((size_t)k * nz + jm) * ny + im] = cpx<C>(r.x, -r.y);
}

template <class C, bool DEV>
__global__ void k_update(C *__restrict__ ph, const C *__restrict__ mu, const C *__restrict__ noise, int hx,
                         int ny, int nz, UpdateParams u, const double *origin, Slab sl) {
  const size_t M = (size_t)hx * nz * ny;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < M * sl.n;
       w += (size_t)gridDim.x * blockDim.x) {
    const size_t r = w / M, off = r * sl.sc;
    update_index<DEV>(ph + off, mu + off, noise ? noise + off : nullptr, w % M, hx, ny, nz, u,
                      origin ? origin + r : nullptr);
  }
}

// only the planes kx = 0 and kx = Nx/2 (the bulk is fused into the z pass)
template <class C, bool DEV>
__global__ void k_update_edges(C *__restrict__ ph, const C *__restrict__ mu, const C *__restrict__ noise,
                               int hx, int ny, int nz, UpdateParams u, const double *origin, size_t sc) {
  ph += blockIdx.y * sc;
  mu += blockIdx.y * sc;
  if (noise) noise += blockIdx.y * sc;
  if (origin) origin += blockIdx.y;
  const size_t P = (size_t)nz * ny;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < 2 * P;
       w += (size_t)gridDim.x * blockDim.x) {
    const size_t idx = w < P ? w : (size_t)(hx - 1) * P + (w - P);
    update_index<DEV>(ph, mu, noise, idx, hx, ny, nz, u, origin);
  }
}

// What the update (update_index) adds as noise at K-layout index idx with
// the device generator: the mode's own draw where it is updated, the
// conjugate of its primary's draw on a mirror, nothing at the origin.
__device__ double2 update_noise_at(const UpdateParams &u, size_t idx, int hx, int ny, int nz) {
  const int A = ny, B = nz, H = hx - 1;
  const int i = (int)(idx % ny);
  const size_t t = idx / ny;
  const int j = (int)(t % nz);
  const int k = (int)(t / nz);
  if (k != 0 && k != H) return device_noise(u, idx, i, j, k, A, B, false);
  const bool self = (i == 0 || i == A / 2) && (j == 0 || j == B / 2);
  if (self) {
    if (k == 0 && i == 0 && j == 0) return make_double2(0.0, 0.0);
    return device_noise(u, idx, i, j, k, A, B, true);
  }
  const bool primary = (j >= 1 && j < B / 2) || (j == 0 && i >= 1 && i < A / 2) || (j == B / 2 && i > A / 2);
  if (primary) return device_noise(u, idx, i, j, k, A, B, false);
  const int im = (A - i) % A, jm = (B - j) % B;
  const double2 v = device_noise(u, ((size_t)k * nz + jm) * ny + im, im, jm, k, A, B, false);
  return make_double2(v.x, -v.y);
}
// out[kz][ky][kx] (canonical layout)
__global__ void k_noise_field(double2 *__restrict__ out, int hx, int ny, int nz, UpdateParams u) {
  const size_t M = (size_t)hx * ny * nz;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < M; w += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(w % hx);
    const size_t t = w / hx;
    const int ky = (int)(t % ny);
    const int kz = (int)(t / ny);
    out[w] = update_noise_at(u, ((size_t)kx * nz + kz) * ny + ky, hx, ny, nz);
  }
}

// FixGridFloryHuggins::post_force, src/fixgrid_floryhuggins.cpp:65-85
template <class R>
__global__ void k_floryhuggins(size_t n, R *__restrict__ phi, R *__restrict__ mu, R tv, R chi,
                               unsigned long long *counters, size_t sr, const double *chi_b) {
  phi += blockIdx.y * sr;
  mu += blockIdx.y * sr;
  counters += 2 * blockIdx.y;
  if (chi_b) chi = (R)chi_b[blockIdx.y];
  __shared__ unsigned cnt[2];
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
  __syncthreads();
  unsigned nl = 0, ng = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    R p = phi[i];
    if (p <= 0) {
      p = -p;
      nl++;
      phi[i] = p;
    } else if (p >= 1) {
      p = 1 - (p - 1);
      ng++;
      phi[i] = p;
    }
    mu[i] += tv * (flog(p / (1 - p)) + chi * (1 - 2 * p));
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

// FP64 reductions in a fixed order, the same bits on every run: each block
// writes its partial (a fixed grid-stride share, combined in a fixed tree),
// then one block folds the partials in index order and adds them to out.
constexpr int RED_BLOCKS = 2048;
template <int W>
__global__ void k_fold_partials(int nb, const double *__restrict__ part, double *out) {
  __shared__ double ws[W][TPB / 64];
#pragma unroll
  for (int k = 0; k < W; k++) {
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += TPB) a += part[(size_t)i * W + k];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) a += __shfl_down(a, d, 64);
    if ((threadIdx.x & 63) == 0) ws[k][threadIdx.x >> 6] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < W; k++) {
      double t = 0;
      for (int w = 0; w < TPB / 64; w++) t += ws[k][w];
      out[k] += t;
    }
  }
}

// Flory-Huggins + square-gradient free energy (defined by this build):
// sum of phi ln phi + (1-phi) ln(1-phi) + chi phi (1-phi) and |grad phi|^2,
// evaluated and summed in FP64 (block partials part[2 * block + 0..1]).
template <class R>
__global__ void k_energy_fh(size_t n, const R *__restrict__ phi, const R *__restrict__ gx,
                            const R *__restrict__ gy, const R *__restrict__ gz, double tv, double chi,
                            double tvg, double *part) {
  __shared__ double s0[TPB], s1[TPB];
  double a = 0, b = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double p = phi[i];
    if (p <= 0) p = -p;
    else if (p >= 1) p = 1 - (p - 1);
    if (tv != 0.0) a += p * log(p) + (1 - p) * log(1 - p) + chi * p * (1 - p);
    if (gx) {
      const double x = gx[i], y = gy[i], z = gz[i];
      b += x * x + y * y + z * z;
    }
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
    part[2 * blockIdx.x] = tv * s0[0];
    part[2 * blockIdx.x + 1] = 0.5 * tvg * s1[0];
  }
}

// part[block] = the block's FP64 sum of its grid-stride share of x
template <class R>
__global__ void k_sum_fp64(size_t n, const R *__restrict__ x, double *part) {
  __shared__ double ws[TPB / 64];
  double a = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a += (double)x[i];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) a += __shfl_down(a, d, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < TPB / 64; w++) t += ws[w];
    part[blockIdx.x] = t;
  }
}

int grid_for(size_t n) {
  size_t b = (n + TPB - 1) / TPB;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

template <class C>
void grad_multiply_t(const FFTPlan &p, const C *phihat, C *gx, C *gy, C *gz, SpecConst c, hipStream_t s, Slab sl) {
  if (sl.n > 1 && sl.sc < p.M()) throw std::invalid_argument("slab stride shorter than a spectrum");
  probe_begin(K_GRAD_MUL, s);
  hipLaunchKernelGGL(k_grad_multiply<C>, dim3(grid_for(p.M() * sl.n)), dim3(TPB), 0, s, phihat, gx, gy, gz, p.hx,
                     p.ny, p.nz, c, sl);
  probe_end(K_GRAD_MUL, s, 4.0 * sizeof(C) * p.M() * sl.n);
  HIPCHK(hipGetLastError());
}
template <class C>
void spectral_update_t(const FFTPlan &p, C *phihat, const C *muhat, const C *noise_k, UpdateParams u,
                       const double *origin, hipStream_t s, Slab sl) {
  if (sl.n > 1 && sl.sc < p.M()) throw std::invalid_argument("slab stride shorter than a spectrum");
  probe_begin(K_UPDATE, s);
  if (u.noise_dev)
    hipLaunchKernelGGL((k_update<C, true>), dim3(grid_for(p.M() * sl.n)), dim3(TPB), 0, s, phihat, muhat, noise_k,
                       p.hx, p.ny, p.nz, u, origin, sl);
  else
    hipLaunchKernelGGL((k_update<C, false>), dim3(grid_for(p.M() * sl.n)), dim3(TPB), 0, s, phihat, muhat, noise_k,
                       p.hx, p.ny, p.nz, u, origin, sl);
  probe_end(K_UPDATE, s, (noise_k ? 4.0 : 3.0) * sizeof(C) * p.M() * sl.n);
  HIPCHK(hipGetLastError());
}
template <class C>
void spectral_update_edges_t(const FFTPlan &p, C *phihat, const C *muhat, const C *noise_k, UpdateParams u,
                             const double *origin, hipStream_t s, Slab sl) {
  if (u.noise_dev)
    hipLaunchKernelGGL((k_update_edges<C, true>), dim3(grid_for(2 * (size_t)p.ny * p.nz), sl.n), dim3(TPB), 0, s,
                       phihat, muhat, noise_k, p.hx, p.ny, p.nz, u, origin, sl.sc);
  else
    hipLaunchKernelGGL((k_update_edges<C, false>), dim3(grid_for(2 * (size_t)p.ny * p.nz), sl.n), dim3(TPB), 0, s,
                       phihat, muhat, noise_k, p.hx, p.ny, p.nz, u, origin, sl.sc);
  HIPCHK(hipGetLastError());
}
template <class R>
void floryhuggins_t(size_t n, R *phi, R *mu, double tv, double chi, unsigned long long *counters,
                    hipStream_t s, Slab sl, const double *chi_b) {
  probe_begin(K_FH, s);
  hipLaunchKernelGGL(k_floryhuggins<R>, dim3(grid_for(n), sl.n), dim3(TPB), 0, s, n, phi, mu, (R)tv, (R)chi,
                     counters, sl.sr, chi_b);
  probe_end(K_FH, s, 3.0 * sizeof(R) * n * sl.n);
  HIPCHK(hipGetLastError());
}
template <class R>
void reduce_energy_fh_t(size_t n, const R *phi, const R *gx, const R *gy, const R *gz, double tv, double chi,
                        double tvg, double *out2, hipStream_t s) {
  const int nb = grid_for(n) > 1024 ? 1024 : grid_for(n);
  double *part = reduce_partials(s, 2 * (size_t)RED_BLOCKS);
  hipLaunchKernelGGL(k_energy_fh<R>, dim3(nb), dim3(TPB), 0, s, n, phi, gx, gy, gz, tv, chi, tvg, part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fold_partials<2>, dim3(1), dim3(TPB), 0, s, nb, (const double *)part, out2);
  HIPCHK(hipGetLastError());
}
template <class R>
void sum_fp64_t(size_t n, const R *x, double *out, hipStream_t s) {
  const int nb = grid_for(n) > RED_BLOCKS ? RED_BLOCKS : grid_for(n);
  double *part = reduce_partials(s, 2 * (size_t)RED_BLOCKS);
  hipLaunchKernelGGL(k_sum_fp64<R>, dim3(nb), dim3(TPB), 0, s, n, x, part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fold_partials<1>, dim3(1), dim3(TPB), 0, s, nb, (const double *)part, out);
  HIPCHK(hipGetLastError());
}

}  // namespace

#define PHAFD_FIELD_DEF(RT, CT)                                                                              \
  void grad_multiply(const FFTPlan &p, const CT *phihat, CT *gx, CT *gy, CT *gz, SpecConst c, hipStream_t s,  \
                     Slab sl) {                                                                              \
    grad_multiply_t(p, phihat, gx, gy, gz, c, s, sl);                                                        \
  }                                                                                                          \
  void spectral_update(const FFTPlan &p, CT *phihat, const CT *muhat, const CT *noise_k, UpdateParams u,    \
                       const double *origin, hipStream_t s, Slab sl) {                                       \
    spectral_update_t(p, phihat, muhat, noise_k, u, origin, s, sl);                                          \
  }                                                                                                          \
  void spectral_update_edges(const FFTPlan &p, CT *phihat, const CT *muhat, const CT *noise_k,              \
                             UpdateParams u, const double *origin, hipStream_t s, Slab sl) {                 \
    spectral_update_edges_t(p, phihat, muhat, noise_k, u, origin, s, sl);                                    \
  }                                                                                                          \
  void floryhuggins(size_t n, RT *phi, RT *mu, double tv, double chi, unsigned long long *counters,         \
                    hipStream_t s, Slab sl, const double *chi_b) {                                           \
    floryhuggins_t(n, phi, mu, tv, chi, counters, s, sl, chi_b);                                             \
  }                                                                                                          \
  void reduce_energy_fh(size_t n, const RT *phi, const RT *gx, const RT *gy, const RT *gz, double tv,       \
                        double chi, double tvg, double *out2, hipStream_t s) {                               \
    reduce_energy_fh_t(n, phi, gx, gy, gz, tv, chi, tvg, out2, s);                                          \
  }                                                                                                          \
  void sum_fp64(size_t n, const RT *x, double *out, hipStream_t s) { sum_fp64_t(n, x, out, s); }
PHAFD_FIELD_DEF(double, double2)
PHAFD_FIELD_DEF(float, float2)
#undef PHAFD_FIELD_DEF

void noise_field(int hx, int ny, int nz, UpdateParams u, double2 *out_canon, hipStream_t s) {
  hipLaunchKernelGGL(k_noise_field, dim3(grid_for((size_t)hx * ny * nz)), dim3(TPB), 0, s, out_canon, hx, ny, nz,
                     u);
  HIPCHK(hipGetLastError());
}

void fill_zero(void *p, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0 && (bytes & 15) == 0) {
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(bytes / 16)), dim3(TPB), 0, s, reinterpret_cast<uint4 *>(p),
                       bytes / 16);
  } else {
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(grid_for(bytes)), dim3(TPB), 0, s,
                       reinterpret_cast<unsigned char *>(p), bytes);
  }
  HIPCHK(hipGetLastError());
}

// ------------------------------------------------- surrogate block check
// The plain update of one mode, phihat <- phihat - c (muhat + s phihat)
// (src/conjugate_volfrac.cpp:89-133; c = mobility q^2 dt, s = temp/volFH
// gamma q^2), split as mu = f2 phi + n(phi) with f2 = temp/volFH (1/m +
// 1/(1-m) - 2 chi) + 2 mean(rho) (m = mean phi), is phihat <- G phihat -
// c nhat with G = 1 - c (f2 + s); over K steps
//   phi_K = G^K phi_0 - c sum_{k<K} G^(K-1-k) n_k.
namespace {
// A run of m consecutive powers G^j (j = 0 .. m-1): S = sum G^j, T = sum
// j G^j, P = G^m.  Two runs concatenate (the second shifted by the first's
// m) as S = S1 + P1 S2, T = T1 + P1 (T2 + m1 S2), P = P1 P2, so the K-term
// sums of a block come from binary doubling in ~2 log2 K steps, with no
// cancellation near G = 1 (unlike (1 - G^K) / (1 - G)).
struct GeoRun {
  double S, T, P, m;
};
__device__ __forceinline__ GeoRun geo_cat(const GeoRun &a, const GeoRun &b) {
  return GeoRun{a.S + a.P * b.S, a.T + a.P * (b.T + a.m * b.S), a.P * b.P, a.m + b.m};
}
__device__ __forceinline__ GeoRun geo_run(double G, int K) {
  GeoRun r{0.0, 0.0, 1.0, 0.0}, b{1.0, 0.0, G, 1.0};
  while (K) {
    if (K & 1) r = geo_cat(r, b);
    K >>= 1;
    if (K) b = geo_cat(b, b);
  }
  return r;
}

struct ModeK {
  double c, G, GK, S, T;   // S = sum_{j<K} G^j, T = sum_{j<K} j G^j
};
__device__ __forceinline__ ModeK mode_k(size_t w, int ny, int nz, const UpdateParams &u, double f2, int K) {
  const int i = (int)(w % ny);                    // ky
  const int j = (int)((w / ny) % nz);             // kz
  const int k = (int)(w / ((size_t)ny * nz));     // kx
  const double qz = (i > ny / 2) ? -u.qky_scale * (ny - i) : u.qky_scale * i;
  const double qy = (j > nz / 2) ? -u.qkz_scale * (nz - j) : u.qkz_scale * j;
  const double qx = u.qx_scale * k;
  const double q2 = qx * qx + qy * qy + qz * qz;
  ModeK r;
  r.c = u.mobility * q2 * u.dt;
  r.G = 1.0 - r.c * (f2 + u.temp / u.volFH * u.gamma * q2);
  const GeoRun g = geo_run(r.G, K);
  r.GK = g.P;
  r.S = g.S;
  r.T = g.T;
  return r;
}
__device__ __forceinline__ double f2_of(const double *sums, double N, double tv, double chi) {
  const double m = sums[0] / N, rbar = sums[1] / N;
  return tv * (1.0 / m + 1.0 / (1.0 - m) - 2.0 * chi) + 2.0 * rbar;
}

// mu of the plain step at phi without the reflection: Flory-Huggins
// (src/fixgrid_floryhuggins.cpp:71-83) plus the grid-atom term 2 (phi -
// phi_t) rho for one atom type (rho = sum of the kernel values e)
__global__ void k_mu_surrogate(size_t n, const double *__restrict__ phi, const float *__restrict__ rho, double tv,
                               double chi, double phit, double *__restrict__ mu) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double p = phi[i];
    mu[i] = tv * (log(p / (1 - p)) + chi * (1 - 2 * p)) + 2 * (p - phit) * (double)rho[i];
  }
}
// K steps of one mode with the linear part exact and n = muhat - f2 phihat
// held at its block-start value (exponential Euler over the block):
//   out = G^K p0 - c (1 + G + ... + G^(K-1)) n0
__global__ void k_etd(const double2 *__restrict__ p0, const double2 *__restrict__ m0, double2 *__restrict__ out,
                      int ny, int nz, size_t M, UpdateParams u, double tv, double chi, const double *__restrict__ sums,
                      double N, int K) {
  const double f2 = f2_of(sums, N, tv, chi);
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < M; w += (size_t)gridDim.x * blockDim.x) {
    const ModeK q = mode_k(w, ny, nz, u, f2, K);
    const double S = q.S;
    const double2 a = p0[w], ma = m0[w];
    double2 r;
    r.x = q.GK * a.x - q.c * S * (ma.x - f2 * a.x);
    r.y = q.GK * a.y - q.c * S * (ma.y - f2 * a.y);
    out[w] = r;
  }
}

// per block: sums over the half spectrum (weights 1 on kx = 0, Nx/2, else 2)
// of |R|^2 and |ph1 - ph0|^2 with R = ph1 - G^K ph0 + c (A n0 + B n1),
// n = muhat - f2 phihat interpolated linearly across the block: A = sum_k
// G^(K-1-k) (1 - k/K) = (T + S) / K, B = sum_k G^(K-1-k) k/K =
// ((K - 1) S - T) / K (j = K-1-k; S, T the sums of ModeK)
__global__ __launch_bounds__(256) void k_block_residual(const double2 *__restrict__ ph0,
                                                        const double2 *__restrict__ ph1,
                                                        const double2 *__restrict__ mu0,
                                                        const double2 *__restrict__ mu1, int nx, int ny, int nz,
                                                        size_t M, UpdateParams u, double tv, double chi,
                                                        const double *__restrict__ sums, double N, int K,
                                                        double *__restrict__ part) {
  __shared__ double r1[256], r2[256];
  const double f2 = f2_of(sums, N, tv, chi);
  double sr = 0, sd = 0;
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < M; w += (size_t)gridDim.x * blockDim.x) {
    const ModeK q = mode_k(w, ny, nz, u, f2, K);
    const double A = (q.T + q.S) / K, B = ((K - 1) * q.S - q.T) / K;
    const double2 a = ph0[w], b = ph1[w], m0 = mu0[w], m1 = mu1[w];
    const double n0x = m0.x - f2 * a.x, n0y = m0.y - f2 * a.y;
    const double n1x = m1.x - f2 * b.x, n1y = m1.y - f2 * b.y;
    const double rx = b.x - q.GK * a.x + q.c * (A * n0x + B * n1x);
    const double ry = b.y - q.GK * a.y + q.c * (A * n0y + B * n1y);
    const double dx = b.x - a.x, dy = b.y - a.y;
    const int kx = (int)(w / ((size_t)ny * nz));
    const double wt = (kx == 0 || 2 * kx == nx) ? 1.0 : 2.0;
    sr += wt * (rx * rx + ry * ry);
    sd += wt * (dx * dx + dy * dy);
  }
  r1[threadIdx.x] = sr;
  r2[threadIdx.x] = sd;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r1[threadIdx.x] += r1[threadIdx.x + o];
      r2[threadIdx.x] += r2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}
}  // namespace

void surrogate_mu(size_t n, const double *phi, const float *rho, double tv, double chi, double phit, double *mu,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_mu_surrogate, dim3(grid_for(n)), dim3(TPB), 0, s, n, phi, rho, tv, chi, phit, mu);
  HIPCHK(hipGetLastError());
}
void etd_block(const FFTPlan &p, const double2 *p0, const double2 *m0, double2 *out, UpdateParams u, double tv,
               double chi, const double *sums, int K, hipStream_t s) {
  hipLaunchKernelGGL(k_etd, dim3(grid_for(p.M())), dim3(TPB), 0, s, p0, m0, out, p.ny, p.nz, p.M(), u, tv, chi, sums,
                     (double)p.N(), K);
  HIPCHK(hipGetLastError());
}
void block_residual_partials(const FFTPlan &p, const double2 *ph0, const double2 *ph1, const double2 *mu0,
                             const double2 *mu1, UpdateParams u, double tv, double chi, const double *sums, int K,
                             double *part, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_block_residual, dim3(nb), dim3(TPB), 0, s, ph0, ph1, mu0, mu1, p.nx, p.ny, p.nz, p.M(), u,
                     tv, chi, sums, (double)p.N(), K, part);
  HIPCHK(hipGetLastError());
}

}  // namespace phafd
