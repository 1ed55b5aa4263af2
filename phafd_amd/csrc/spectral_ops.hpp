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
// A = Ny, B = Nz).  With u.semi the linear gradient term is implicit:
// (phi - c mu) / (1 + c s) in place of phi - c (mu + s phi), the form left
// in a comment at src/conjugate.cpp:356 (opt-in; the reference runs the
// explicit form).
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
  if (u.semi) {
    const R den = R(1) + c * s;
    const R ux = (phi.x - c * mu.x) / den, uy = (phi.y - c * mu.y) / den;
    return cpx<C>(ux * (R)u.normalization + noise.x, uy * (R)u.normalization + noise.y);
  }
  const R tx = mu.x + s * phi.x, tyy = mu.y + s * phi.y;
  const R ux = phi.x - c * tx, uy = phi.y - c * tyy;
  return cpx<C>(ux * (R)u.normalization + noise.x, uy * (R)u.normalization + noise.y);
}

// ---------------------------------------------------------- device noise
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 64x64->128 multiplies with
// the published multipliers and Weyl key increments.  The same function
// as numpy.random.Philox's block function.
__device__ __forceinline__ void philox4x64_10(unsigned long long c[4], unsigned long long k0,
                                              unsigned long long k1) {
  const unsigned long long M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const unsigned long long W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += W0; k1 += W1; }
    const unsigned long long hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const unsigned long long hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const unsigned long long n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}
// the top 53 bits as a double in [-1/2, 1/2) (exact)
__device__ __forceinline__ double unit_centered(unsigned long long w) {
  return (double)(w >> 11) * 0x1.0p-53 - 0.5;
}
// The thermal noise ConjugateVolFrac adds to one updated mode
// (src/conjugate_volfrac.cpp:95-96 complex, :118 real): amplitude
// sqrt(12 T M / L^3) |q| sqrt(dt) on each part of a complex mode and
// sqrt(24 T M / L^3) |q| sqrt(dt) on a self-conjugate (real) mode, times
// uniform draws on [-1/2, 1/2).  The draws come from Philox keyed by
// (u.key0, u.key1) at counter (kidx, u.step, 0, 0), kidx the K-layout mode
// index, instead of the reference's serial mt19937 stream, so every mode is
// drawn independently on the device.  Computed in FP64.
__device__ __forceinline__ double2 device_noise(const UpdateParams &u, size_t kidx, int i, int j, int k, int A,
                                                int B, bool real) {
  const double qz = (i > A / 2) ? -u.qky_scale * (A - i) : u.qky_scale * i;
  const double qy = (j > B / 2) ? -u.qkz_scale * (B - j) : u.qkz_scale * j;
  const double qx = u.qx_scale * k;
  const double q2 = qx * qx + qy * qy + qz * qz;
  unsigned long long w[4] = {(unsigned long long)kidx, u.step, 0ull, 0ull};
  philox4x64_10(w, u.key0, u.key1);
  if (real) return make_double2(u.pref_r * sqrt(q2) * u.sqrtdt * unit_centered(w[0]), 0.0);
  const double a = u.pref_c * sqrt(q2) * u.sqrtdt;
  return make_double2(a * unit_centered(w[0]), a * unit_centered(w[1]));
}
// the noise of one updated mode: the device generator (DEV, a template
// argument so that kernels without it carry none of its code), a host field,
// or none
template <bool DEV, class C>
__device__ __forceinline__ C mode_noise(const C *__restrict__ noise, size_t idx, const UpdateParams &u, int i,
                                        int j, int k, int A, int B, bool real) {
  using R = typename C::value_type;
  if (DEV) {
    const double2 v = device_noise(u, idx, i, j, k, A, B, real);
    return cpx<C>((R)v.x, (R)v.y);
  }
  if (noise) return noise[idx];
  return cpx<C>(R(0), R(0));
}

}  // namespace phafd
