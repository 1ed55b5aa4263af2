// spectral_ops.hpp -- per-mode arithmetic of the k-space hooks, shared by the
// stand-alone pointwise kernels and the kernels that fuse them into the z pass
// of the FFT.  Written in the reference's operation order; the library is
// built with -ffp-contract=off so no multiply-add is fused behind its back.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace phafd {

__device__ __forceinline__ int signed_k(int i, int n) { return i > n / 2 ? i - n : i; }

// One component of the gradient multiply, src/fixgrid_conjugate.cpp:80-105:
// (phihat * complex(0, dq)) * n with std::complex semantics,
// (a + ib) * (0 + id) = (a*0 - b*d, a*d + b*0), then times the integer wave
// number n as a double.
__device__ __forceinline__ double2 grad_mode(double2 v, double dq, double n) {
  const double re = v.x * 0.0 - v.y * dq;
  const double im = v.x * dq + v.y * 0.0;
  return make_double2(re * n, im * n);
}

// One mode of ConjugateVolFrac::complex_update / real_update,
// src/conjugate_volfrac.cpp:89-133 (reference indices i = ky, j = kz, k = kx;
// A = Ny, B = Nz).
__device__ __forceinline__ double2 mode_update(double2 phi, double2 mu, double2 noise, int i,
                                               int j, int k, int A, int B,
                                               const UpdateParams &u) {
  const double qz = (i > A / 2) ? -u.qky_scale * (A - i) : u.qky_scale * i;   // qzs[i]
  const double qy = (j > B / 2) ? -u.qkz_scale * (B - j) : u.qkz_scale * j;   // qys[j]
  const double qx = u.qx_scale * k;
  const double q2 = qx * qx + qy * qy + qz * qz;
  const double s = u.temp / u.volFH * u.gamma * q2;
  const double c = u.mobility * q2 * u.dt;
  const double tx = mu.x + s * phi.x, tyy = mu.y + s * phi.y;
  const double ux = phi.x - c * tx, uy = phi.y - c * tyy;
  return make_double2(ux * u.normalization + noise.x, uy * u.normalization + noise.y);
}

}  // namespace phafd
