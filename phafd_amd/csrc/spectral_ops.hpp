// spectral_ops.hpp -- per-mode arithmetic of the k-space hooks, shared by the
// stand-alone pointwise kernels and the kernels that fuse them into the z pass
// of the FFT, for FP64 (double2) and FP32 (float2) spectra.  Written in the
// reference's operation order; the library is built with -ffp-contract=off
// so no multiply-add is fused behind its back.  In FP64 every cast below is
// the identity.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace phafd {

__device__ __forceinline__ int signed_k(int i, int n) { return i > n / 2 ? i - n : i; }

template <class C>
__device__ __forceinline__ C cpx(typename C::value_type x, typename C::value_type y) {
  C r;
  r.x = x;
  r.y = y;
  return r;
}

// One component of the gradient multiply, src/fixgrid_conjugate.cpp:80-105:
// (phihat * complex(0, dq)) * n with std::complex semantics,
// (a + ib) * (0 + id) = (a*0 - b*d, a*d + b*0), then times the integer wave
// number n.
template <class C>
__device__ __forceinline__ C grad_mode(C v, typename C::value_type dq, typename C::value_type n) {
  using R = typename C::value_type;
  const R re = v.x * R(0) - v.y * dq;
  const R im = v.x * dq + v.y * R(0);
  return cpx<C>(re * n, im * n);
}

// One mode of ConjugateVolFrac::complex_update / real_update,
// src/conjugate_volfrac.cpp:89-133 (reference indices i = ky, j = kz, k = kx;
// A = Ny, B = Nz).
template <class C>
__device__ __forceinline__ C mode_update(C phi, C mu, C noise, int i, int j, int k, int A, int B,
                                         const UpdateParams &u) {
  using R = typename C::value_type;
  const R qky = (R)u.qky_scale, qkz = (R)u.qkz_scale;
  const R qz = (i > A / 2) ? -qky * (A - i) : qky * i;   // qzs[i]
  const R qy = (j > B / 2) ? -qkz * (B - j) : qkz * j;   // qys[j]
  const R qx = (R)u.qx_scale * k;
  const R q2 = qx * qx + qy * qy + qz * qz;
  const R s = (R)u.temp / (R)u.volFH * (R)u.gamma * q2;
  const R c = (R)u.mobility * q2 * (R)u.dt;
  const R tx = mu.x + s * phi.x, tyy = mu.y + s * phi.y;
  const R ux = phi.x - c * tx, uy = phi.y - c * tyy;
  return cpx<C>(ux * (R)u.normalization + noise.x, uy * (R)u.normalization + noise.y);
}

}  // namespace phafd
