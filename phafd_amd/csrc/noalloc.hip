// noalloc.hip -- grad(phi) without any device allocation.
//
// phafd_hip_gradphi_noalloc runs the kernels of phafd_hip_gradphi (the
// r2c of phi, the i*q multiply, one c2r per component;
// src/fixgrid_conjugate.cpp:62-116) on caller-owned device arrays.  Its
// intermediates -- the half spectrum of phi and one gradient spectrum -- and
// the FFT twiddle tables live in `scratch`, a static device array of
// PHAFD_NOALLOC_SCRATCH_BYTES that each GPU holds until the process ends:
//
//   [0, 64 KiB)        twiddles exp(-2 pi i m/L) at entry L-2 and
//                      exp(+2 pi i m/L) at entry NTAB + L-2, L = 2, 4 .. 1024
//   next 16 M bytes    phi's half spectrum (K layout, M = (Nx/2+1) Ny Nz)
//   next 16 M bytes    the spectrum of one gradient component
//
// Set-up, on the first call on a device with that device current: the
// array's device address from hipGetSymbolAddress (kept per device; host
// code never uses the array's own name, a host-side stand-in) and one copy
// of the twiddles.  After that a call enqueues kernels only: no hipMalloc,
// hipFree, event, stream or library allocation.  Every element a call reads
// from the spectra it has written first, so nothing needs clearing; the
// twiddles are never written again.  Calls share their GPU's array, so
// calls that overlap on one GPU (other streams, other host threads) race.
#include <mutex>
#include <string>
#include <vector>

#include "../../include/phafd_hip.h"
#include "internal.hpp"
#include "spectral_ops.hpp"

using namespace phafd;

namespace {

constexpr size_t SCRATCH_BYTES = PHAFD_NOALLOC_SCRATCH_BYTES;
constexpr size_t TAB_BYTES = 64 << 10;
constexpr int NTAB = 2046;   // sum of L = 2, 4, .., 1024
static_assert(2 * NTAB * sizeof(double2) <= TAB_BYTES, "twiddle region");
static_assert((SCRATCH_BYTES - TAB_BYTES) / (2 * sizeof(double2)) == PHAFD_GRADPHI_NOALLOC_MAX_MODES,
              "the published mode limit is what the spectrum region holds");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAXDEV = 64;
std::mutex g_mu;
unsigned char *g_scr[MAXDEV];   // device address of `scratch` on each device (null: not set up)

// the current device's copy of `scratch`, set up on first use
unsigned char *scratch_of(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scr[dev]) return g_scr[dev];
  void *p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
  std::vector<double2> tab(2 * NTAB);
  for (int L = 2; L <= 1024; L *= 2) twiddle_host(&tab[L - 2], &tab[NTAB + L - 2], L);
  HIPCHK(hipMemcpy(p, tab.data(), sizeof(double2) * tab.size(), hipMemcpyHostToDevice));
  g_scr[dev] = (unsigned char *)p;
  return g_scr[dev];
}

// one component of the gradient multiply of phafd_hip_gradphi (the same
// grad_mode arithmetic per mode): dir 0 -> x (kx), 1 -> y (signed ky),
// 2 -> z (signed kz); K layout index = (kx*nz + kz)*ny + ky
__global__ void k_grad_one(const double2 *__restrict__ ph, double2 *__restrict__ g, int hx, int ny, int nz,
                           int dir, double dq) {
  const size_t M = (size_t)hx * nz * ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M; i += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(i % ny);
    const size_t t = i / ny;
    const int kz = (int)(t % nz);
    const int kx = (int)(t / nz);
    const int n = dir == 0 ? kx : dir == 1 ? signed_k(ky, ny) : signed_k(kz, nz);
    g[i] = grad_mode(ph[i], dq, (double)n);
  }
}

}  // namespace

namespace phafd {
void set_stateless_error(const std::string &msg);
}

extern "C" int phafd_hip_gradphi_noalloc(int nx, int ny, int nz, const double *period, const double *phi,
                                         double *gx, double *gy, double *gz, void *stream) {
  try {
    if (!period || !phi || !gx || !gy || !gz) throw std::invalid_argument("gradphi_noalloc: null pointer");
    for (int d = 0; d < 3; d++)
      if (!(period[d] > 0)) throw std::invalid_argument("gradphi_noalloc: period must be positive");
    FFTPlan plan;
    plan.set_shape(nx, ny, nz);   // powers of two in [4, 1024]
    const size_t M = plan.M(), N = plan.N();
    if (M > (size_t)PHAFD_GRADPHI_NOALLOC_MAX_MODES)
      throw std::invalid_argument("gradphi_noalloc: (Nx/2+1)*Ny*Nz = " + std::to_string(M) +
                                  " half-spectrum modes exceed the scratch limit of " +
                                  std::to_string((long long)PHAFD_GRADPHI_NOALLOC_MAX_MODES) + " (" +
                                  std::to_string(SCRATCH_BYTES >> 20) + " MiB per GPU)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAXDEV) throw std::invalid_argument("gradphi_noalloc: device ordinal beyond 63");
    unsigned char *base = scratch_of(dev);
    double2 *tab = (double2 *)base;
    auto view = [&](DBuf<double2> &f, DBuf<double2> &b, int L) {
      f.wrap(tab + (L - 2), L);
      b.wrap(tab + NTAB + (L - 2), L);
    };
    view(plan.tw_x_f, plan.tw_x_b, nx);
    view(plan.tw_h_f, plan.tw_h_b, nx / 2);
    view(plan.tw_y_f, plan.tw_y_b, ny);
    view(plan.tw_z_f, plan.tw_z_b, nz);
    double2 *ph = (double2 *)(base + TAB_BYTES);
    double2 *gh = ph + M;
    hipStream_t s = (hipStream_t)stream;
    fft_r2c(plan, phi, ph, s);
    // SpecConst of phafd_hip_gradphi: x by 2 pi / period[0], y by
    // 2 pi / period[1], z by 2 pi / period[2] (PI of src/domain.cpp:206-208)
    const double PI_REF = 3.141592;
    const double dq[3] = {2 * PI_REF / period[0], 2 * PI_REF / period[1], 2 * PI_REF / period[2]};
    double *out[3] = {gx, gy, gz};
    const unsigned blocks = (unsigned)std::min<size_t>((M + 255) / 256, 4096);
    for (int d = 0; d < 3; d++) {
      hipLaunchKernelGGL(k_grad_one, dim3(blocks), dim3(256), 0, s, ph, gh, plan.hx, ny, nz, d, dq[d]);
      HIPCHK(hipGetLastError());
      fft_c2r(plan, gh, out[d], (double)N, s);
    }
    return PHAFD_OK;
  } catch (const HipError &e) {
    set_stateless_error(e.what());
    return PHAFD_E_HIP;
  } catch (const std::invalid_argument &e) {
    set_stateless_error(e.what());
    return PHAFD_E_ARG;
  } catch (const std::exception &e) {
    set_stateless_error(e.what());
    return PHAFD_E_RUNTIME;
  }
}
