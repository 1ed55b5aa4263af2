// unet.hip -- kernels and host code of the surrogate network (unet.hpp).
//
// Activations are channels-last ([z][y][x][c], x fastest) in bf16.  A 3^3
// convolution is an implicit GEMM on the v_mfma_f32_16x16x32_bf16 units:
// M = 16 consecutive x points of one row, N = 16 output channels, K = 32 =
// four groups of 8 input channels, a group being (tap, 8 channels); periodic
// wrap on every axis.  Each wave computes two rows (32 points) x all output
// channels; weights are packed once per update in the exact lane order of the
// B operand, so a wave reads them with one 16-byte load per lane.  The weight
// gradient is the same GEMM with K = grid points (staged transposed in LDS);
// the data gradient is a forward convolution with the transposed, flipped
// kernel.  Every reduction runs in a fixed order: training is reproducible.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <string>

#include "phafd_hip.h"
#include "unet.hpp"

namespace phafd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int TPB = 256;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (the conversion PyTorch applies), by gfx950's
// v_cvt_pk_bf16_f32 (one instruction per pair)
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ uint32_t f2bf2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}
__device__ __forceinline__ bool bf_pos(uint16_t h) { return !(h & 0x8000u) && h != 0; }

__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

size_t nblk(size_t n) {
  size_t b = (n + TPB - 1) / TPB;
  return b < 1 ? 1 : (b > 65535u * 4 ? 65535u * 4 : b);
}

// 8 bf16 channels <-> float
__device__ __forceinline__ void unpack8(const u32x4 &u, float f[8]) {
#pragma unroll
  for (int e = 0; e < 4; e++) {
    f[2 * e] = __uint_as_float(u[e] << 16);
    f[2 * e + 1] = __uint_as_float(u[e] & 0xffff0000u);
  }
}
__device__ __forceinline__ u32x4 pack8(const float f[8]) {
  u32x4 u;
#pragma unroll
  for (int e = 0; e < 4; e++) u[e] = f2bf2(f[2 * e], f[2 * e + 1]);
  return u;
}

// ------------------------------------------------------------ convolution
// Y[v][co] = act(bias[co] + sum_{tap, ci} W[co][ci][tap] X[v + off(tap)][ci])
// W packed: [chunk][n-tile][lane][8] (B-operand order).  A block of 4 waves
// computes tiles of 16 (x) by 4 * MTR (y) points, MTR rows per wave, for zb
// consecutive z: its input planes (18 x (4 MTR + 2) halo points for a 3^3
// kernel) sit in an LDS ring of three, one new plane loaded per z (periodic
// wrap applied on load), so the K loop reads the A operands from LDS at
// fixed offsets per tap.  A plane is stored group-major ([8-channel
// group][y][x], 16 bytes per entry): the 16 lanes of a row read 256
// contiguous bytes, free of bank conflicts.  Grid: (nx/16) * (ny/(4 MTR)) *
// (nz/zb) blocks.
// K order of a layer's packed weights: chunk c and lane group kg (16 lanes,
// one 8-channel group each) -> the tap (return value, -1 past the last
// group) and the channel group cg.  3^3 kernels: chunks come in threes,
// one per dy (dy = c % 3 - 1), over load phases p = c / 3 of four (dz, dx,
// cg) columns (column 4p + kg), so a wave reads a column's rows once for the
// three dy taps.  1x1 kernels: group g = 4c + kg, cg = g (one tap).
__host__ __device__ inline int conv_k_group(int c, int kg, int CG, int taps, int &cg) {
  if (taps == 27) {
    const int q = 4 * (c / 3) + kg;
    if (q >= 9 * CG) return -1;
    const int comb = q / CG;   // dz * 3 + dx (each 0..2)
    cg = q % CG;
    return (comb / 3) * 9 + (c % 3) * 3 + comb % 3;
  }
  const int g = 4 * c + kg;
  if (g >= taps * CG) return -1;
  cg = g % CG;
  return g / CG;
}
__host__ __device__ constexpr int conv_chunks(int CG, int taps) {
  return taps == 27 ? 3 * ((9 * CG + 3) / 4) : (taps * CG + 3) / 4;
}

template <int CIN, int TAPS, int MTR>
struct ConvTile {
  static constexpr int CG = CIN / 8;
  static constexpr int TY = 4 * MTR;
  static constexpr int HX = TAPS == 27 ? 18 : 16;
  static constexpr int HY = TAPS == 27 ? TY + 2 : TY;
  static constexpr int HZ = TAPS == 27 ? 3 : 1;
  static constexpr int PLANE = HY * HX * CG;   // 16-byte entries per plane
};

// Q (optional): a coarse (half-resolution) input of the same channels added
// on load, X[v] + Q[parent(v)] rounded to bf16 -- the decoder's skip sum,
// bit for bit what k_upadd would have stored.
// Fused epilogues (EPI): EPI_POOL also writes P, the 2x2x2 mean of the
// stored (bf16) outputs, bit for bit what k_pool computes from Y (needs an
// even zb); EPI_HEAD (16 output channels) writes out[v] = head(Y[v]) as
// k_head does, in place of Y (Y may be null).
enum { EPI_NONE = 0, EPI_POOL = 1, EPI_HEAD = 2 };
// the value of the lane's x neighbour (lx ^ 1) by DPP quad_perm [1, 0, 3, 2]
__device__ __forceinline__ float xpair(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
// the head's sum order (shared by k_head): four partials over consecutive
// channel quads, then bias + ((p0 + p1) + (p2 + p3))
__device__ __forceinline__ float head_quad(const float *w, const float a[4]) {
  float p = w[0] * a[0];
  p = fmaf(w[1], a[1], p);
  p = fmaf(w[2], a[2], p);
  return fmaf(w[3], a[3], p);
}

template <int CIN, int COUT, int TAPS, bool RELU, int MTR, bool UP, int EPI>
__global__ __launch_bounds__(256) void k_conv(const u32x4 *__restrict__ X, const u32x4 *__restrict__ Q,
                                              const u32x4 *__restrict__ W, const float *__restrict__ bias,
                                              uint16_t *__restrict__ Y, int nx, int ny, int nz, int zb,
                                              uint16_t *__restrict__ P, const float *__restrict__ hw,
                                              const float *__restrict__ hb, float *__restrict__ hout) {
  static_assert(EPI != EPI_HEAD || COUT == 16, "the fused head reads 16 channels");
  static_assert(EPI != EPI_POOL || MTR % 2 == 0, "the fused pool pairs rows");
  using T = ConvTile<CIN, TAPS, MTR>;
  constexpr int CG = T::CG;
  constexpr int G = TAPS * CG;
  constexpr int NC = conv_chunks(CG, TAPS);
  constexpr int NT = COUT / 16;
  constexpr bool DYR = TAPS == 27;   // rows reused across the dy taps
  constexpr int NP = NC / 3;                     // load phases (DYR)
  constexpr int H = TAPS == 27 ? 1 : 0;   // halo width
  constexpr int PER = (T::PLANE + 255) / 256;
  __shared__ u32x4 tile[T::HZ * T::PLANE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xt = nx >> 4, yt = ny / T::TY;
  int b = blockIdx.x;
  const int bx = b % xt;
  b /= xt;
  const int by = b % yt;
  const int z0 = (b / yt) * zb;
  const int x0 = bx * 16, y0 = by * T::TY;
  const int kg = lane >> 4;
  const int lx = lane & 15;
  // this thread's share of a plane load: offset within a z plane of X (group
  // fastest, coalesced) and the group-major LDS slot; the same every plane
  int goff[PER], loff[PER], qoff[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int t = threadIdx.x + 256 * k;
    const int cg = t % CG, v = t / CG;
    const int hx = v % T::HX, hy = v / T::HX;
    const int xx = wrap1(x0 - H + hx, nx), yy = wrap1(y0 - H + hy, ny);
    goff[k] = t < T::PLANE ? (yy * nx + xx) * CG + cg : -1;
    loff[k] = (cg * T::HY + hy) * T::HX + hx;
    qoff[k] = ((yy >> 1) * (nx >> 1) + (xx >> 1)) * CG + cg;
  }
  const size_t zstride = (size_t)ny * nx * CG;
  const size_t qzstride = (size_t)(ny >> 1) * (nx >> 1) * CG;
  // plane loads in two halves: fetch (global -> registers) and put
  // (registers -> LDS), so the fetch of the next plane is in flight while
  // the current one is computed
  u32x4 px[PER], pq[UP ? PER : 1];
  auto fetch = [&](int z) {
    const int wz = wrap1(z, nz);
    const u32x4 *Xz = X + (size_t)wz * zstride;
#pragma unroll
    for (int k = 0; k < PER; k++)
      if (goff[k] >= 0) px[k] = Xz[goff[k]];
    if (UP) {
      const u32x4 *Qz = Q + (size_t)(wz >> 1) * qzstride;
#pragma unroll
      for (int k = 0; k < (UP ? PER : 0); k++)
        if (goff[k] >= 0) pq[k] = Qz[qoff[k]];
    }
  };
  auto put = [&](u32x4 *dst) {
#pragma unroll
    for (int k = 0; k < PER; k++)
      if (goff[k] >= 0) {
        if (UP) {
          float a[8], q[8];
          unpack8(px[k], a);
          unpack8(pq[UP ? k : 0], q);
#pragma unroll
          for (int e = 0; e < 8; e++) a[e] = a[e] + q[e];
          dst[loff[k]] = pack8(a);
        } else {
          dst[loff[k]] = px[k];
        }
      }
  };
  // this lane's A (points) offset for K chunk c within a ring plane and the
  // plane it reads (dz + 1), packed as (dz + 1) << 20 | offset; -1 past the
  // last group
  auto chunk_code = [&](int c) -> int {
    const int g = 4 * c + kg;
    const int tap = g / CG, cg = g % CG;
    int dx = 0, dy = 0, dzi = 0;
    if (TAPS == 27) {
      dzi = tap / 9;
      dy = (tap / 3) % 3 - 1;
      dx = tap % 3 - 1;
    }
    const int off = (cg * T::HY + wave * MTR + dy + H) * T::HX + (lx + dx + H);
    return g < G ? (dzi << 20) | off : -1;
  };
  // DYR: this lane's column of load phase p -- its ring plane (dz + 1) and
  // the offset of its first row (y0 + wave MTR - 1), packed as above; the
  // phase reads rows 0 .. MTR + 1 from there, and dy tap e (dy = e - 1) of
  // output row r is row r + e
  auto col_code = [&](int p) -> int {
    const int q = 4 * p + kg;
    const int comb = q / CG, cg = q % CG;
    const int off = (cg * T::HY + wave * MTR) * T::HX + (lx + comb % 3);
    return q < 9 * CG ? ((comb / 3) << 20) | off : -1;
  };
  constexpr bool PRE = DYR ? NP <= 9 : NC <= 16;   // codes kept in registers (fully unrolled K loop)
  constexpr int NCODE = PRE ? (DYR ? NP : NC) : 1;
  int code[NCODE];
  if (PRE) {
#pragma unroll
    for (int c = 0; c < NCODE; c++) code[c] = DYR ? col_code(c) : chunk_code(c);
  }
  // the output channels of this lane: 4 * kg + j of each n-tile
  float bb[NT][4];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int j = 0; j < 4; j++) bb[t][j] = bias ? bias[t * 16 + 4 * kg + j] : 0.f;
  if (TAPS == 27) {
    fetch(z0 - 1);
    put(tile);
    fetch(z0);
    put(tile + T::PLANE);
    fetch(z0 + 1);
  } else {
    fetch(z0);
  }
  // the weight loads do not depend on z: kept in registers across the z
  // loop when they fit (16 chunk tiles, 64 VGPRs), else reloaded per z (L1 /
  // L2 hits) through a lane index the compiler cannot hoist
  constexpr bool WHOIST = NC * NT <= 16;
  float pool_part[EPI == EPI_POOL ? MTR / 2 : 1][NT][4];
  for (int zi = 0; zi < zb; zi++) {
    const int z = z0 + zi;
    int wl = lane;
    if (!WHOIST) asm volatile("" : "+v"(wl));
    // plane z + d lives in ring slot (zi + 1 + d) % 3; a 1x1 kernel reads its
    // one plane z from slot 0.  The registers hold plane z + 1 (z for 1x1);
    // the next one is fetched once they are in LDS.
    put(TAPS == 27 ? tile + ((zi + 2) % 3) * T::PLANE : tile);
    __syncthreads();
    if (zi + 1 < zb) fetch(TAPS == 27 ? z + 2 : z + 1);
    const int sb0 = TAPS == 27 ? (zi % 3) * T::PLANE : 0;
    const int sb1 = TAPS == 27 ? ((zi + 1) % 3) * T::PLANE : 0;
    const int sb2 = TAPS == 27 ? ((zi + 2) % 3) * T::PLANE : 0;
    f32x4 acc[MTR][NT];
#pragma unroll
    for (int r = 0; r < MTR; r++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto chunk = [&](int c, int cd) {
      bf16x8 a[MTR];
      if (cd >= 0) {
        const int pz = cd >> 20;
        const int base = (pz == 0 ? sb0 : (pz == 1 ? sb1 : sb2)) + (cd & 0xfffff);
#pragma unroll
        for (int r = 0; r < MTR; r++) a[r] = __builtin_bit_cast(bf16x8, tile[base + r * T::HX]);
      } else {
#pragma unroll
        for (int r = 0; r < MTR; r++) a[r] = __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u});
      }
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const bf16x8 w = __builtin_bit_cast(bf16x8, W[((size_t)c * NT + t) * 64 + wl]);
        // weights as the A operand (rows = output channels), points as B
#pragma unroll
        for (int r = 0; r < MTR; r++)
          acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, a[r], acc[r][t], 0, 0, 0);
      }
    };
    auto phase = [&](int p, int cd) {
      bf16x8 a[MTR + 2];
      if (cd >= 0) {
        const int pz = cd >> 20;
        const int base = (pz == 0 ? sb0 : (pz == 1 ? sb1 : sb2)) + (cd & 0xfffff);
#pragma unroll
        for (int r = 0; r < MTR + 2; r++) a[r] = __builtin_bit_cast(bf16x8, tile[base + r * T::HX]);
      } else {
#pragma unroll
        for (int r = 0; r < MTR + 2; r++) a[r] = __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u});
      }
#pragma unroll
      for (int e = 0; e < 3; e++)
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const bf16x8 w = __builtin_bit_cast(bf16x8, W[((size_t)(3 * p + e) * NT + t) * 64 + wl]);
#pragma unroll
          for (int r = 0; r < MTR; r++)
            acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, a[r + e], acc[r][t], 0, 0, 0);
        }
    };
    if (DYR) {
      if (PRE) {
#pragma unroll
        for (int p = 0; p < NP; p++) phase(p, code[PRE ? p : 0]);
      } else {
#pragma unroll 2
        for (int p = 0; p < NP; p++) phase(p, col_code(p));
      }
    } else if (PRE) {
#pragma unroll
      for (int c = 0; c < NC; c++) chunk(c, code[PRE ? c : 0]);
    } else {
#pragma unroll 2
      for (int c = 0; c < NC; c++) chunk(c, chunk_code(c));
    }
    // D layout: lane holds rows 4 * kg + j (output channels of the n-tile),
    // column lx (the point x0 + lx): 4 consecutive channels, one 8-byte store
    float fe[NT][4];   // EPI_POOL: the even row's stored values
#pragma unroll
    for (int r = 0; r < MTR; r++) {
      const size_t pt = ((size_t)z * ny + (y0 + wave * MTR + r)) * nx + x0 + lx;
#pragma unroll
      for (int t = 0; t < NT; t++) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          v[j] = acc[r][t][j] + bb[t][j];
          if (RELU) v[j] = fmaxf(v[j], 0.f);
        }
        uint2 o;
        o.x = f2bf2(v[0], v[1]);
        o.y = f2bf2(v[2], v[3]);
        if (EPI != EPI_HEAD) *reinterpret_cast<uint2 *>(Y + pt * COUT + t * 16 + 4 * kg) = o;
        float a[4] = {__uint_as_float(o.x << 16), __uint_as_float(o.x & 0xffff0000u), __uint_as_float(o.y << 16),
                      __uint_as_float(o.y & 0xffff0000u)};
        if (EPI == EPI_POOL && !(r & 1)) {
#pragma unroll
          for (int j = 0; j < 4; j++) fe[t][j] = a[j];
        }
        if (EPI == EPI_POOL && (r & 1)) {
          // children in k_pool's order: x fastest, then y, then z; the sum
          // over the even plane is kept in registers until the odd one
          const int rp = r >> 1;
          float m[4];
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const float c0 = fe[t][j], c2 = a[j];
            const float c1 = xpair(c0), c3 = xpair(c2);
            float sum = (zi & 1) ? pool_part[EPI == EPI_POOL ? rp : 0][t][j] : 0.f;
            sum = sum + c0;
            sum = sum + c1;
            sum = sum + c2;
            sum = sum + c3;
            pool_part[EPI == EPI_POOL ? rp : 0][t][j] = sum;
            m[j] = sum * 0.125f;
          }
          if ((zi & 1) && !(lx & 1)) {
            const size_t cp = ((size_t)(z >> 1) * (ny >> 1) + ((y0 + wave * MTR + r) >> 1)) * (nx >> 1) +
                              ((x0 + lx) >> 1);
            uint2 q;
            q.x = f2bf2(m[0], m[1]);
            q.y = f2bf2(m[2], m[3]);
            *reinterpret_cast<uint2 *>(P + cp * COUT + t * 16 + 4 * kg) = q;
          }
        }
        if (EPI == EPI_HEAD) {
          // lanes kg = 0..3 hold the channel quads of point pt: partials,
          // then (p0 + p1) + (p2 + p3) across the lane groups
          float pq = head_quad(hw + 4 * kg, a);
          pq = pq + __shfl_xor(pq, 16);
          pq = pq + __shfl_xor(pq, 32);
          if (kg == 0) hout[pt] = hb[0] + pq;
        }
      }
    }
    __syncthreads();   // the ring slot read here is reloaded next
  }
}

// weight gradient partials: part[(tap * nb + blk) * 4 + wave][ci][co] =
// sum over this wave's points v of X[v + off(tap)][ci] * D[v][co].  Grid
// (nb, TAPS); each block takes groups of 256 points (64 per wave), stages X
// (shifted) and D transposed in LDS, and runs K = 32-point MFMA steps.
template <int CIN, int COUT, int TAPS>
__global__ __launch_bounds__(256) void k_conv_dw(const uint16_t *__restrict__ X, const uint16_t *__restrict__ D,
                                                 float *__restrict__ part, int nx, int ny, int nz) {
  constexpr int MI = (CIN + 15) / 16;
  constexpr int NT = COUT / 16;
  constexpr int VS = 256 + 8;   // LDS row stride in points (16-byte aligned rows)
  __shared__ __align__(16) uint16_t xs[CIN * VS];
  __shared__ __align__(16) uint16_t ds[COUT * VS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tap = blockIdx.y;
  int dx = 0, dy = 0, dz = 0;
  if (TAPS == 27) {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
  }
  const size_t N = (size_t)nx * ny * nz;
  const size_t ngroups = (N + 255) / 256;
  f32x4 acc[MI][NT];
#pragma unroll
  for (int m = 0; m < MI; m++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (size_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const size_t v = grp * 256 + threadIdx.x;
    if (v >= N) {   // past the last point: zero rows
#pragma unroll
      for (int c = 0; c < CIN; c++) xs[c * VS + threadIdx.x] = 0;
#pragma unroll
      for (int c = 0; c < COUT; c++) ds[c * VS + threadIdx.x] = 0;
    } else {
      const int x = (int)(v % nx);
      const size_t t1 = v / nx;
      const int y = (int)(t1 % ny), z = (int)(t1 / ny);
      const size_t vs = ((size_t)wrap1(z + dz, nz) * ny + wrap1(y + dy, ny)) * nx + wrap1(x + dx, nx);
      const u32x4 *xp = reinterpret_cast<const u32x4 *>(X + vs * CIN);
#pragma unroll
      for (int q = 0; q < CIN / 8; q++) {
        const u32x4 w = xp[q];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          xs[(q * 8 + 2 * e) * VS + threadIdx.x] = (uint16_t)(w[e] & 0xffffu);
          xs[(q * 8 + 2 * e + 1) * VS + threadIdx.x] = (uint16_t)(w[e] >> 16);
        }
      }
      const u32x4 *dp = reinterpret_cast<const u32x4 *>(D + v * COUT);
#pragma unroll
      for (int q = 0; q < COUT / 8; q++) {
        const u32x4 w = dp[q];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          ds[(q * 8 + 2 * e) * VS + threadIdx.x] = (uint16_t)(w[e] & 0xffffu);
          ds[(q * 8 + 2 * e + 1) * VS + threadIdx.x] = (uint16_t)(w[e] >> 16);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; s++) {
      const int p0 = wave * 64 + s * 32 + 8 * (lane >> 4);
      bf16x8 a[MI], bm[NT];
#pragma unroll
      for (int m = 0; m < MI; m++) {
        const int ci = m * 16 + (lane & 15);
        if (ci < CIN) a[m] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&xs[ci * VS + p0]));
        else a[m] = __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u});
      }
#pragma unroll
      for (int t = 0; t < NT; t++)
        bm[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(&ds[(t * 16 + (lane & 15)) * VS + p0]));
#pragma unroll
      for (int m = 0; m < MI; m++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], bm[t], acc[m][t], 0, 0, 0);
    }
    __syncthreads();
  }
  float *o = part + (((size_t)tap * gridDim.x + blockIdx.x) * 4 + wave) * (CIN * COUT);
#pragma unroll
  for (int m = 0; m < MI; m++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int ci = m * 16 + 4 * (lane >> 4) + j;
        if (ci < CIN) o[ci * COUT + t * 16 + (lane & 15)] = acc[m][t][j];
      }
}

// G_w[co][ci][tap] += sum over the nb*4 partials (fixed order)
__global__ void k_dw_reduce(const float *__restrict__ part, int nparts, int cin, int cout, int taps,
                            float *__restrict__ gw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = cin * cout * taps;
  if (i >= total) return;
  const int co = i % cout, ci = (i / cout) % cin, tap = i / (cout * cin);
  const float *p = part + (size_t)tap * nparts * (cin * cout) + ci * cout + co;
  float s = 0.f;
  for (int k = 0; k < nparts; k++) s += p[(size_t)k * cin * cout];
  gw[((size_t)co * cin + ci) * taps + tap] += s;
}

// column sums of D[N][C] per block: part[blk][C]
template <int C>
__global__ __launch_bounds__(256) void k_colsum(const uint16_t *__restrict__ D, size_t N, float *__restrict__ part) {
  __shared__ float red[256];
  constexpr int PER = 256 / C;   // points per block iteration
  const int c = threadIdx.x % C, p = threadIdx.x / C;
  float s = 0.f;
  for (size_t v = (size_t)blockIdx.x * PER + p; v < N; v += (size_t)gridDim.x * PER) s += bf2f(D[v * C + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int k = 0; k < PER; k++) t += red[k * C + threadIdx.x];
    part[(size_t)blockIdx.x * C + threadIdx.x] = t;
  }
}
__global__ void k_colsum_reduce(const float *__restrict__ part, int nb, int C, float *__restrict__ g) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int k = 0; k < nb; k++) s += part[(size_t)k * C + c];
  g[c] += s;
}

// ----------------------------------------------------- pooling / upsampling
// fine point of coarse point p, child k (dz, dy, dx bits)
__device__ __forceinline__ size_t child(size_t p, int k, int cx, int cy) {
  const int x = (int)(p % cx);
  const size_t t = p / cx;
  const int y = (int)(t % cy), z = (int)(t / cy);
  const int fx = 2 * x + (k & 1), fy = 2 * y + ((k >> 1) & 1), fz = 2 * z + (k >> 2);
  return ((size_t)fz * (2 * cy) + fy) * (2 * cx) + fx;
}
__device__ __forceinline__ size_t parent(size_t f, int fx_n, int fy_n) {
  const int x = (int)(f % fx_n);
  const size_t t = f / fx_n;
  const int y = (int)(t % fy_n), z = (int)(t / fy_n);
  return ((size_t)(z >> 1) * (fy_n >> 1) + (y >> 1)) * (fx_n >> 1) + (x >> 1);
}

// One thread per (point, group of 8 channels), 16-byte loads and stores;
// C channels = C / 8 groups per point.

// P[p] = mean of the 8 children (coarse dims cx, cy; Nc coarse points)
__global__ void k_pool(const u32x4 *__restrict__ X, u32x4 *__restrict__ P, size_t Nc, int CG, int cx, int cy) {
  const size_t n = Nc * CG;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t p = i / CG;
    const int g = (int)(i % CG);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 8; k++) {
      float f[8];
      unpack8(X[child(p, k, cx, cy) * CG + g], f);
#pragma unroll
      for (int e = 0; e < 8; e++) s[e] += f[e];
    }
#pragma unroll
    for (int e = 0; e < 8; e++) s[e] *= 0.125f;
    P[i] = pack8(s);
  }
}
// dX[f] = dA[f] + dP[parent] / 8 (fine dims fx, fy)
__global__ void k_pool_bwd_add(const u32x4 *__restrict__ dP, const u32x4 *__restrict__ dA, u32x4 *__restrict__ dX,
                               size_t Nf, int CG, int fx, int fy) {
  const size_t n = Nf * CG;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t f = i / CG;
    const int g = (int)(i % CG);
    float a[8], q[8];
    unpack8(dA[i], a);
    unpack8(dP[parent(f, fx, fy) * CG + g], q);
#pragma unroll
    for (int e = 0; e < 8; e++) a[e] = a[e] + q[e] * 0.125f;
    dX[i] = pack8(a);
  }
}
// T[f] = S[f] + Q[parent]
__global__ void k_upadd(const u32x4 *__restrict__ Q, const u32x4 *__restrict__ S, u32x4 *__restrict__ T, size_t Nf,
                        int CG, int fx, int fy) {
  const size_t n = Nf * CG;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t f = i / CG;
    const int g = (int)(i % CG);
    float a[8], q[8];
    unpack8(S[i], a);
    unpack8(Q[parent(f, fx, fy) * CG + g], q);
#pragma unroll
    for (int e = 0; e < 8; e++) a[e] = a[e] + q[e];
    T[i] = pack8(a);
  }
}
// dQ[p] = sum of the 8 children of dT
__global__ void k_downsum(const u32x4 *__restrict__ dT, u32x4 *__restrict__ dQ, size_t Nc, int CG, int cx, int cy) {
  const size_t n = Nc * CG;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t p = i / CG;
    const int g = (int)(i % CG);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 8; k++) {
      float f[8];
      unpack8(dT[child(p, k, cx, cy) * CG + g], f);
#pragma unroll
      for (int e = 0; e < 8; e++) s[e] += f[e];
    }
    dQ[i] = pack8(s);
  }
}
// D = D where A > 0, else 0 (gradient through relu, A its output)
__global__ void k_mask(uint16_t *__restrict__ D, const uint16_t *__restrict__ A, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    if (!bf_pos(A[i])) D[i] = 0;
}

// ------------------------------------------------------------------ head
// out[v] = b + sum_c w[c] A[v][c] (16 channels, FP32)
__global__ void k_head(const u32x4 *__restrict__ A, const float *__restrict__ w, const float *__restrict__ b,
                       float *__restrict__ out, size_t N) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < N; v += (size_t)gridDim.x * blockDim.x) {
    float p[4];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const u32x4 u = A[v * 2 + q];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const float a[4] = {__uint_as_float(u[2 * h] << 16), __uint_as_float(u[2 * h] & 0xffff0000u),
                            __uint_as_float(u[2 * h + 1] << 16), __uint_as_float(u[2 * h + 1] & 0xffff0000u)};
        p[2 * q + h] = head_quad(w + 8 * q + 4 * h, a);
      }
    }
    out[v] = b[0] + ((p[0] + p[1]) + (p[2] + p[3]));
  }
}
// dA[v][c] = (A[v][c] > 0) ? dout[v] w[c] : 0; part[blk][17] = (sum dout A[.][c], sum dout)
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t *__restrict__ A, const float *__restrict__ w,
                                                  const float *__restrict__ dout, uint16_t *__restrict__ dA, size_t N,
                                                  float *__restrict__ part) {
  __shared__ float red[256][17];
  float s[17];
#pragma unroll
  for (int c = 0; c < 17; c++) s[c] = 0.f;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < N; v += (size_t)gridDim.x * blockDim.x) {
    const float d = dout[v];
#pragma unroll
    for (int c = 0; c < 16; c++) {
      const uint16_t a = A[v * 16 + c];
      dA[v * 16 + c] = bf_pos(a) ? f2bf(d * w[c]) : (uint16_t)0;
      s[c] += d * bf2f(a);
    }
    s[16] += d;
  }
#pragma unroll
  for (int c = 0; c < 17; c++) red[threadIdx.x][c] = s[c];
  __syncthreads();
  if (threadIdx.x < 17) {
    float t = 0.f;
    for (int k = 0; k < 256; k++) t += red[k][threadIdx.x];
    part[(size_t)blockIdx.x * 17 + threadIdx.x] = t;
  }
}
__global__ void k_head_reduce(const float *__restrict__ part, int nb, float *__restrict__ gw, float *__restrict__ gb) {
  const int c = threadIdx.x;
  if (c >= 17) return;
  float s = 0.f;
  for (int k = 0; k < nb; k++) s += part[(size_t)k * 17 + c];
  if (c < 16) gw[c] += s;
  else gb[0] += s;
}

// --------------------------------------------------------- loss, optimiser
// dout = scale * 2 (out - y); part[blk] = sum (out - y)^2 (FP64)
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ out, const float *__restrict__ y, size_t N,
                                             float scale, float *__restrict__ dout, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < N; v += (size_t)gridDim.x * blockDim.x) {
    const float d = out[v] - y[v];
    dout[v] = scale * 2.f * d;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float *__restrict__ p, float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       size_t n, float lr_t, float b1, float b2, float eps) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
    g[i] = 0.f;
  }
}

// packed weight operand of a layer: out[((c * NT + t) * 64 + l) * 8 + j] =
// bf16 of W[co][ci][tap] with (tap, cg) = conv_k_group(c, l / 16), ci = cg * 8
// + j, co = 16 t + l % 16; zero past the last group.  transposed: the weights
// of the data gradient (CIN' = cout, COUT' = cin, tap' = taps - 1 - tap).
__global__ void k_pack(const float *__restrict__ W, int cin, int cout, int taps, bool transposed,
                       uint16_t *__restrict__ out) {
  const int CIN = transposed ? cout : cin, COUT = transposed ? cin : cout;
  const int CG = CIN / 8, NC = conv_chunks(CG, taps), NT = COUT / 16;
  const size_t total = (size_t)NC * NT * 64 * 8;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(i % 8);
    const int l = (int)((i / 8) % 64);
    const int t = (int)((i / 512) % NT);
    const int c = (int)(i / (512 * (size_t)NT));
    int cg = 0;
    const int tap = conv_k_group(c, l >> 4, CG, taps, cg);
    float w = 0.f;
    if (tap >= 0) {
      const int ci = cg * 8 + j, co = t * 16 + (l & 15);
      if (!transposed) w = W[((size_t)co * cin + ci) * taps + tap];
      else w = W[((size_t)ci * cin + co) * taps + (taps - 1 - tap)];   // W[co'=ci][ci'=co]
    }
    out[i] = f2bf(w);
  }
}
size_t packed_size(int cin, int cout, int taps) {
  return (size_t)conv_chunks(cin / 8, taps) * (cout / 16) * 64 * 8;
}

// ------------------------------------------------------ features / apply
// channels: (phi - mean) / s0, rho0 / s1, rho1 / s1, (mean - mref) / s3
__global__ void k_features(size_t N, const double *__restrict__ phi, const float *__restrict__ rho0,
                           const float *__restrict__ rho1, const double *__restrict__ phisum, double s0, double s1,
                           double mref, double s3, u32x4 *__restrict__ X0) {
  const double mean = phisum[0] / (double)N;
  const uint32_t c3 = f2bf((float)((mean - mref) / s3));
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < N; v += (size_t)gridDim.x * blockDim.x) {
    const uint32_t c0 = f2bf((float)((phi[v] - mean) / s0));
    const uint32_t c1 = f2bf((float)(rho0[v] / s1));
    const uint32_t c2 = f2bf((float)(rho1[v] / s1));
    X0[v] = u32x4{c0 | (c1 << 16), c2 | (c3 << 16), 0u, 0u};
  }
}
__global__ void k_features_ch(size_t N, const float *__restrict__ ch, u32x4 *__restrict__ X0) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < N; v += (size_t)gridDim.x * blockDim.x) {
    const uint32_t c0 = f2bf(ch[v]), c1 = f2bf(ch[N + v]), c2 = f2bf(ch[2 * N + v]), c3 = f2bf(ch[3 * N + v]);
    X0[v] = u32x4{c0 | (c1 << 16), c2 | (c3 << 16), 0u, 0u};
  }
}
__global__ void k_apply(size_t N, const double *__restrict__ phi, const float *__restrict__ out,
                        const double *__restrict__ outsum, double scale, double *__restrict__ po) {
  const double mean = outsum[0] / (double)N;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < N; v += (size_t)gridDim.x * blockDim.x)
    po[v] = phi[v] + scale * ((double)out[v] - mean);
}

// ------------------------------------------------------------ launchers
// rows per wave: 4 when the rows of a level allow it and the halo tile fits
// comfortably in LDS (64 input channels: 2)
// z slices per block: the largest of 8, 4, 2, 1 dividing nz that leaves at
// least 2048 blocks (8 per CU)
int conv_zb(int tiles_xy, int nz) {
  for (int zb = 8; zb > 1; zb >>= 1)
    if (nz % zb == 0 && (size_t)tiles_xy * (nz / zb) >= 2048) return zb;
  return 1;
}

// rows per wave: 4 when the rows of a level allow it, the halo planes fit
// comfortably in LDS and the accumulators in registers (64 input or output
// channels: 2; a skip sum formed on load: 2, whose registers for the
// coarse input would otherwise cost a wave per SIMD)
// the fused epilogue's outputs (EPI_POOL: P; EPI_HEAD: hw, hb, hout)
struct ConvEpi {
  uint16_t *P = nullptr;
  const float *hw = nullptr, *hb = nullptr;
  float *hout = nullptr;
};

template <int CIN, int COUT, int TAPS, bool RELU, int EPI = EPI_NONE>
void conv(const uint16_t *X, const uint16_t *W, const float *bias, uint16_t *Y, int nx, int ny, int nz,
          hipStream_t s, const uint16_t *Q = nullptr, const ConvEpi &e = ConvEpi()) {
  const bool four = CIN < 64 && COUT < 64 && ny % 16 == 0 && !Q;
  const int ty = four ? 16 : 8;
  const int tiles = (nx / 16) * (ny / ty);
  const int zb = conv_zb(tiles, nz);
  if (EPI == EPI_POOL && zb % 2) {   // z pairs would straddle blocks: separate pool
    conv<CIN, COUT, TAPS, RELU>(X, W, bias, Y, nx, ny, nz, s, Q);
    const size_t Nc = (size_t)nx * ny * nz / 8;
    hipLaunchKernelGGL(k_pool, dim3(nblk(Nc * (COUT / 8))), dim3(TPB), 0, s, reinterpret_cast<const u32x4 *>(Y),
                       reinterpret_cast<u32x4 *>(e.P), Nc, COUT / 8, nx / 2, ny / 2);
    HIPCHK(hipGetLastError());
    return;
  }
  const unsigned blocks = (unsigned)(tiles * (nz / zb));
  const u32x4 *x = reinterpret_cast<const u32x4 *>(X), *q = reinterpret_cast<const u32x4 *>(Q),
              *w = reinterpret_cast<const u32x4 *>(W);
#define PHAFD_CONV_LAUNCH(M, U)                                                                                     \
  hipLaunchKernelGGL((k_conv<CIN, COUT, TAPS, RELU, M, U, EPI>), dim3(blocks), dim3(256), 0, s, x, q, w, bias, Y,  \
                     nx, ny, nz, zb, e.P, e.hw, e.hb, e.hout)
  if (four && Q) PHAFD_CONV_LAUNCH(4, true);
  else if (four) PHAFD_CONV_LAUNCH(4, false);
  else if (Q) PHAFD_CONV_LAUNCH(2, true);
  else PHAFD_CONV_LAUNCH(2, false);
#undef PHAFD_CONV_LAUNCH
  HIPCHK(hipGetLastError());
}

template <int CIN, int COUT, int TAPS>
void conv_dw(const uint16_t *X, const uint16_t *D, float *part, float *gw, int nx, int ny, int nz, hipStream_t s) {
  const size_t N = (size_t)nx * ny * nz;
  const int ngroups = (int)((N + 255) / 256);
  const int nb = ngroups < 256 ? ngroups : 256;
  hipLaunchKernelGGL((k_conv_dw<CIN, COUT, TAPS>), dim3(nb, TAPS), dim3(256), 0, s, X, D, part, nx, ny, nz);
  HIPCHK(hipGetLastError());
  const int total = CIN * COUT * TAPS;
  hipLaunchKernelGGL(k_dw_reduce, dim3((total + 255) / 256), dim3(256), 0, s, part, nb * 4, CIN, COUT, TAPS, gw);
  HIPCHK(hipGetLastError());
}

template <int C>
void colsum(const uint16_t *D, size_t N, float *part, float *g, hipStream_t s) {
  constexpr int PER = 256 / C;
  size_t nb = (N + PER - 1) / PER;
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL((k_colsum<C>), dim3((unsigned)nb), dim3(256), 0, s, D, N, part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_colsum_reduce, dim3(1), dim3(64), 0, s, part, (int)nb, C, g);
  HIPCHK(hipGetLastError());
}

void ew_launch_check() { HIPCHK(hipGetLastError()); }
u32x4 *V(uint16_t *p) { return reinterpret_cast<u32x4 *>(p); }

}  // namespace

bool unet_shape_ok(int nx, int ny, int nz) {
  return nx >= 64 && ny >= 32 && nz >= 4 && nx % 64 == 0 && ny % 32 == 0 && nz % 4 == 0;
}

void UNet::init(int dev, uint64_t seed) {
  device = dev;
  HIPCHK(hipSetDevice(dev));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const int spec[9][5] = {{8, 16, 27, 1, 1},  {16, 16, 27, 1, 1}, {16, 32, 27, 2, 1},
                          {32, 64, 27, 3, 1}, {64, 32, 1, 3, 0},  {32, 32, 27, 2, 1},
                          {32, 16, 1, 2, 0},  {16, 16, 27, 1, 1}, {16, 1, 1, 1, 0}};
  size_t off = 0, poff = 0;
  for (auto &q : spec) {
    UNetLayer l;
    l.cin = q[0]; l.cout = q[1]; l.taps = q[2]; l.level = q[3]; l.relu = q[4] != 0;
    l.w = off;
    off += (size_t)l.cin * l.cout * l.taps;
    l.b = off;
    off += l.cout;
    l.pf = l.pt = 0;
    if (l.cout >= 16) {
      l.pf = poff;
      poff += packed_size(l.cin, l.cout, l.taps);
      l.pt = poff;
      if (l.cin >= 16) poff += packed_size(l.cout, l.cin, l.taps);
    }
    L.push_back(l);
  }
  nparam = off;
  npack = poff;
  P.alloc(nparam); G.alloc(nparam); Mo.alloc(nparam); Vo.alloc(nparam);
  Wp.alloc(npack);
  // He-uniform weights (fan in = cin_used * taps), zero biases; layer 0 uses
  // 4 of its 8 input channels, the head starts at zero (the base unchanged)
  std::vector<float> h(nparam, 0.f);
  std::mt19937_64 gen(seed);
  for (size_t li = 0; li < L.size(); li++) {
    const UNetLayer &l = L[li];
    const int cin_used = li == 0 ? UNET_IN : l.cin;
    const double bound = li == 8 ? 0.0 : std::sqrt(6.0 / (cin_used * l.taps));
    std::uniform_real_distribution<double> U(-bound, bound);
    for (int co = 0; co < l.cout; co++)
      for (int ci = 0; ci < l.cin; ci++)
        for (int t = 0; t < l.taps; t++) {
          const double w = (ci < cin_used && bound > 0) ? U(gen) : 0.0;
          h[l.w + ((size_t)co * l.cin + ci) * l.taps + t] = (float)w;
        }
  }
  HIPCHK(hipMemcpy(P.p, h.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
  fill_zero(G.p, sizeof(float) * nparam, stream);
  fill_zero(Mo.p, sizeof(float) * nparam, stream);
  fill_zero(Vo.p, sizeof(float) * nparam, stream);
  HIPCHK(hipStreamSynchronize(stream));
  rng = seed ^ 0x5851f42d4c957f2dull;
  packed = false;
}

void UNet::pack(hipStream_t s) {
  for (const UNetLayer &l : L) {
    if (l.cout < 16) continue;
    hipLaunchKernelGGL(k_pack, dim3(64), dim3(256), 0, s, P.p + l.w, l.cin, l.cout, l.taps, false, Wp.p + l.pf);
    if (l.cin >= 16)
      hipLaunchKernelGGL(k_pack, dim3(64), dim3(256), 0, s, P.p + l.w, l.cin, l.cout, l.taps, true, Wp.p + l.pt);
  }
  HIPCHK(hipGetLastError());
  packed = true;
}

void UNet::ensure(int nx, int ny, int nz, bool train) {
  if (!unet_shape_ok(nx, ny, nz))
    throw std::invalid_argument("surrogate grid must have nx % 64 == 0, ny % 32 == 0, nz % 4 == 0");
  if (A.nx == nx && A.ny == ny && A.nz == nz && (A.train || !train)) return;
  A = UNetActs();
  A.nx = nx; A.ny = ny; A.nz = nz; A.train = train;
  const size_t N1 = (size_t)nx * ny * nz, N2 = N1 / 8, N3 = N2 / 8;
  A.ae1.alloc(N1 * 16); A.s1.alloc(N1 * 16); A.t1.alloc(N1 * 16); A.a1.alloc(N1 * 16);
  A.p1.alloc(N2 * 16); A.s2.alloc(N2 * 32); A.t2.alloc(N2 * 32); A.a2.alloc(N2 * 32); A.q1.alloc(N2 * 16);
  A.p2.alloc(N3 * 32); A.a3.alloc(N3 * 64); A.q2.alloc(N3 * 32);
  A.out.alloc(N1);
  A.red.alloc(4096);
  if (train) {
    A.g1a.alloc(N1 * 16); A.g1b.alloc(N1 * 16); A.g1c.alloc(N1 * 16);
    A.g2a.alloc(N2 * 32); A.g2b.alloc(N2 * 32); A.g2c.alloc(N2 * 32);
    A.g3a.alloc(N3 * 64); A.g3b.alloc(N3 * 64);
    A.dout.alloc(N1);
    // partials of the weight gradients: taps * 256 blocks * 4 waves * cin * cout
    A.part.alloc((size_t)27 * 256 * 4 * 64 * 64);
  }
}

void UNet::forward(const uint16_t *X0, float *out, hipStream_t s) {
  if (!packed) pack(s);
  const int n1x = A.nx, n1y = A.ny, n1z = A.nz;
  const int n2x = n1x / 2, n2y = n1y / 2, n2z = n1z / 2, n3x = n1x / 4, n3y = n1y / 4, n3z = n1z / 4;
  const size_t N1 = (size_t)n1x * n1y * n1z, N2 = N1 / 8, N3 = N2 / 8;
  const uint16_t *W = Wp.p;
  const float *p = P.p;
  conv<8, 16, 27, true>(X0, W + L[0].pf, p + L[0].b, A.ae1.p, n1x, n1y, n1z, s);
  const double conv_flops = 2.0 * 27 * 16 * 16 * (double)N1;   // 3^3 x 16 -> 16 at full resolution
  // the encoder's convolutions also write the pooled level below
  ConvEpi e1, e2;
  e1.P = A.p1.p;
  e2.P = A.p2.p;
  probe_begin(K_UNET_CONV, s);
  conv<16, 16, 27, true, EPI_POOL>(A.ae1.p, W + L[1].pf, p + L[1].b, A.s1.p, n1x, n1y, n1z, s, nullptr, e1);
  probe_end(K_UNET_CONV, s, conv_flops);
  conv<16, 32, 27, true, EPI_POOL>(A.p1.p, W + L[2].pf, p + L[2].b, A.s2.p, n2x, n2y, n2z, s, nullptr, e2);
  conv<32, 64, 27, true>(A.p2.p, W + L[3].pf, p + L[3].b, A.a3.p, n3x, n3y, n3z, s);
  conv<64, 32, 1, false>(A.a3.p, W + L[4].pf, p + L[4].b, A.q2.p, n3x, n3y, n3z, s);
  // the skip sums t = s + up(q): stored when training (the backward pass
  // reads them), else formed by the next convolution as it loads its planes
  if (A.train)
    hipLaunchKernelGGL(k_upadd, dim3(nblk(N2 * 4)), dim3(TPB), 0, s, V(A.q2.p), V(A.s2.p), V(A.t2.p), N2, 4, n2x,
                       n2y);
  if (A.train) conv<32, 32, 27, true>(A.t2.p, W + L[5].pf, p + L[5].b, A.a2.p, n2x, n2y, n2z, s);
  else conv<32, 32, 27, true>(A.s2.p, W + L[5].pf, p + L[5].b, A.a2.p, n2x, n2y, n2z, s, A.q2.p);
  conv<32, 16, 1, false>(A.a2.p, W + L[6].pf, p + L[6].b, A.q1.p, n2x, n2y, n2z, s);
  if (A.train) {
    hipLaunchKernelGGL(k_upadd, dim3(nblk(N1 * 2)), dim3(TPB), 0, s, V(A.q1.p), V(A.s1.p), V(A.t1.p), N1, 2, n1x,
                       n1y);
    conv<16, 16, 27, true>(A.t1.p, W + L[7].pf, p + L[7].b, A.a1.p, n1x, n1y, n1z, s);
    hipLaunchKernelGGL(k_head, dim3(nblk(N1)), dim3(TPB), 0, s, reinterpret_cast<const u32x4 *>(A.a1.p),
                       p + L[8].w, p + L[8].b, out, N1);
  } else {
    // the head applied in the last convolution's epilogue (a1 not stored)
    ConvEpi eh;
    eh.hw = p + L[8].w;
    eh.hb = p + L[8].b;
    eh.hout = out;
    conv<16, 16, 27, true, EPI_HEAD>(A.s1.p, W + L[7].pf, p + L[7].b, nullptr, n1x, n1y, n1z, s, A.q1.p, eh);
  }
  ew_launch_check();
  (void)n3z;
}

void UNet::backward(const uint16_t *X0, const float *dout, hipStream_t s) {
  if (!A.train) throw std::logic_error("activations were not allocated for training");
  const int n1x = A.nx, n1y = A.ny, n1z = A.nz;
  const int n2x = n1x / 2, n2y = n1y / 2, n2z = n1z / 2, n3x = n1x / 4, n3y = n1y / 4, n3z = n1z / 4;
  const size_t N1 = (size_t)n1x * n1y * n1z, N2 = N1 / 8, N3 = N2 / 8;
  const uint16_t *W = Wp.p;
  const float *p = P.p;
  float *g = G.p, *part = A.part.p;
  // head
  {
    const unsigned nb = 512;
    hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(TPB), 0, s, A.a1.p, p + L[8].w, dout, A.g1a.p, N1, part);
    hipLaunchKernelGGL(k_head_reduce, dim3(1), dim3(64), 0, s, part, (int)nb, g + L[8].w, g + L[8].b);
    ew_launch_check();
  }
  // d1 (input t1, output a1; g1a = its masked output gradient)
  conv_dw<16, 16, 27>(A.t1.p, A.g1a.p, part, g + L[7].w, n1x, n1y, n1z, s);
  colsum<16>(A.g1a.p, N1, part, g + L[7].b, s);
  conv<16, 16, 27, false>(A.g1a.p, W + L[7].pt, nullptr, A.g1b.p, n1x, n1y, n1z, s);   // dT1
  // q1: dQ1 = downsum(dT1)
  hipLaunchKernelGGL(k_downsum, dim3(nblk(N2 * 2)), dim3(TPB), 0, s, V(A.g1b.p), V(A.g2c.p), N2, 2, n2x, n2y);
  ew_launch_check();
  conv_dw<32, 16, 1>(A.a2.p, A.g2c.p, part, g + L[6].w, n2x, n2y, n2z, s);
  colsum<16>(A.g2c.p, N2, part, g + L[6].b, s);
  conv<16, 32, 1, false>(A.g2c.p, W + L[6].pt, nullptr, A.g2a.p, n2x, n2y, n2z, s);   // dA2
  hipLaunchKernelGGL(k_mask, dim3(nblk(N2 * 32)), dim3(TPB), 0, s, A.g2a.p, A.a2.p, N2 * 32);
  ew_launch_check();
  // d2
  conv_dw<32, 32, 27>(A.t2.p, A.g2a.p, part, g + L[5].w, n2x, n2y, n2z, s);
  colsum<32>(A.g2a.p, N2, part, g + L[5].b, s);
  conv<32, 32, 27, false>(A.g2a.p, W + L[5].pt, nullptr, A.g2b.p, n2x, n2y, n2z, s);   // dT2
  // q2: dQ2 = downsum(dT2)
  hipLaunchKernelGGL(k_downsum, dim3(nblk(N3 * 4)), dim3(TPB), 0, s, V(A.g2b.p), V(A.g3b.p), N3, 4, n3x, n3y);
  ew_launch_check();
  conv_dw<64, 32, 1>(A.a3.p, A.g3b.p, part, g + L[4].w, n3x, n3y, n3z, s);
  colsum<32>(A.g3b.p, N3, part, g + L[4].b, s);
  conv<32, 64, 1, false>(A.g3b.p, W + L[4].pt, nullptr, A.g3a.p, n3x, n3y, n3z, s);   // dA3
  hipLaunchKernelGGL(k_mask, dim3(nblk(N3 * 64)), dim3(TPB), 0, s, A.g3a.p, A.a3.p, N3 * 64);
  ew_launch_check();
  // b
  conv_dw<32, 64, 27>(A.p2.p, A.g3a.p, part, g + L[3].w, n3x, n3y, n3z, s);
  colsum<64>(A.g3a.p, N3, part, g + L[3].b, s);
  conv<64, 32, 27, false>(A.g3a.p, W + L[3].pt, nullptr, A.g3b.p, n3x, n3y, n3z, s);   // dP2
  // dS2 = dT2 + pool_bwd(dP2), through relu of e2
  hipLaunchKernelGGL(k_pool_bwd_add, dim3(nblk(N2 * 4)), dim3(TPB), 0, s, V(A.g3b.p), V(A.g2b.p), V(A.g2a.p), N2, 4,
                     n2x, n2y);
  hipLaunchKernelGGL(k_mask, dim3(nblk(N2 * 32)), dim3(TPB), 0, s, A.g2a.p, A.s2.p, N2 * 32);
  ew_launch_check();
  // e2
  conv_dw<16, 32, 27>(A.p1.p, A.g2a.p, part, g + L[2].w, n2x, n2y, n2z, s);
  colsum<32>(A.g2a.p, N2, part, g + L[2].b, s);
  conv<32, 16, 27, false>(A.g2a.p, W + L[2].pt, nullptr, A.g2c.p, n2x, n2y, n2z, s);   // dP1
  // dS1 = dT1 + pool_bwd(dP1), through relu of e1b
  hipLaunchKernelGGL(k_pool_bwd_add, dim3(nblk(N1 * 2)), dim3(TPB), 0, s, V(A.g2c.p), V(A.g1b.p), V(A.g1c.p), N1, 2,
                     n1x, n1y);
  hipLaunchKernelGGL(k_mask, dim3(nblk(N1 * 16)), dim3(TPB), 0, s, A.g1c.p, A.s1.p, N1 * 16);
  ew_launch_check();
  // e1b
  conv_dw<16, 16, 27>(A.ae1.p, A.g1c.p, part, g + L[1].w, n1x, n1y, n1z, s);
  colsum<16>(A.g1c.p, N1, part, g + L[1].b, s);
  conv<16, 16, 27, false>(A.g1c.p, W + L[1].pt, nullptr, A.g1a.p, n1x, n1y, n1z, s);   // dAe1
  hipLaunchKernelGGL(k_mask, dim3(nblk(N1 * 16)), dim3(TPB), 0, s, A.g1a.p, A.ae1.p, N1 * 16);
  ew_launch_check();
  // e1 (no input gradient)
  conv_dw<8, 16, 27>(X0, A.g1a.p, part, g + L[0].w, n1x, n1y, n1z, s);
  colsum<16>(A.g1a.p, N1, part, g + L[0].b, s);
}

double UNet::loss_grad(const float *target, double scale, hipStream_t s) {
  const size_t N1 = (size_t)A.nx * A.ny * A.nz;
  const unsigned nb = 1024;
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(TPB), 0, s, A.out.p, target, N1, (float)scale, A.dout.p, A.red.p);
  HIPCHK(hipGetLastError());
  std::vector<double> h(nb);
  HIPCHK(hipMemcpyAsync(h.data(), A.red.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double t = 0.0;
  for (double x : h) t += x;
  return t / (double)N1;
}

void UNet::adam(double lr, double b1, double b2, double eps, hipStream_t s) {
  adam_t++;
  const double lr_t = lr * std::sqrt(1.0 - std::pow(b2, (double)adam_t)) / (1.0 - std::pow(b1, (double)adam_t));
  hipLaunchKernelGGL(k_adam, dim3(nblk(nparam)), dim3(TPB), 0, s, P.p, G.p, Mo.p, Vo.p, nparam, (float)lr_t,
                     (float)b1, (float)b2, (float)eps);
  HIPCHK(hipGetLastError());
  pack(s);
}

void unet_features(size_t N, const double *phi, const float *rho0, const float *rho1, const double *phisum,
                   const double *norm, uint16_t *X0, hipStream_t s) {
  hipLaunchKernelGGL(k_features, dim3(nblk(N)), dim3(TPB), 0, s, N, phi, rho0, rho1, phisum, norm[0], norm[1],
                     norm[2], norm[3], reinterpret_cast<u32x4 *>(X0));
  HIPCHK(hipGetLastError());
}
void unet_features_from_channels(size_t N, const float *ch4, uint16_t *X0, hipStream_t s) {
  hipLaunchKernelGGL(k_features_ch, dim3(nblk(N)), dim3(TPB), 0, s, N, ch4, reinterpret_cast<u32x4 *>(X0));
  HIPCHK(hipGetLastError());
}
void unet_apply(size_t N, const double *phi, const float *out, const double *outsum, double scale, double *phi_out,
                hipStream_t s) {
  hipLaunchKernelGGL(k_apply, dim3(nblk(N)), dim3(TPB), 0, s, N, phi, out, outsum, scale, phi_out);
  HIPCHK(hipGetLastError());
}

}  // namespace phafd

// ================================================================= C ABI
namespace {
thread_local std::string g_unet_err;

template <class Fn>
int uguard(phafd_unet *u, Fn fn) {
  try {
    if (u) (void)hipSetDevice(u->net.device);
    return fn();
  } catch (const phafd::HipError &e) {
    (u ? u->net.err : g_unet_err) = e.what();
    return PHAFD_E_HIP;
  } catch (const std::invalid_argument &e) {
    (u ? u->net.err : g_unet_err) = e.what();
    return PHAFD_E_ARG;
  } catch (const std::exception &e) {
    (u ? u->net.err : g_unet_err) = e.what();
    return PHAFD_E_RUNTIME;
  }
}

// boundary parameter count: layer 0 takes UNET_IN input channels (8 stored)
size_t boundary_count(const phafd::UNet &n) { return n.nparam - (size_t)16 * (8 - phafd::UNET_IN) * 27; }

// boundary <-> stored parameter vectors
void to_stored(const phafd::UNet &n, const float *in, std::vector<float> &out) {
  out.assign(n.nparam, 0.f);
  const phafd::UNetLayer &l0 = n.L[0];
  size_t k = 0;
  for (int co = 0; co < l0.cout; co++)
    for (int ci = 0; ci < phafd::UNET_IN; ci++)
      for (int t = 0; t < 27; t++) out[l0.w + ((size_t)co * l0.cin + ci) * 27 + t] = in[k++];
  std::memcpy(out.data() + l0.b, in + k, sizeof(float) * (n.nparam - l0.b));
}
void to_boundary(const phafd::UNet &n, const std::vector<float> &st, float *out) {
  const phafd::UNetLayer &l0 = n.L[0];
  size_t k = 0;
  for (int co = 0; co < l0.cout; co++)
    for (int ci = 0; ci < phafd::UNET_IN; ci++)
      for (int t = 0; t < 27; t++) out[k++] = st[l0.w + ((size_t)co * l0.cin + ci) * 27 + t];
  std::memcpy(out + k, st.data() + l0.b, sizeof(float) * (n.nparam - l0.b));
}

uint64_t next_rand(uint64_t &s) {   // splitmix64
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
}  // namespace

extern "C" {

int phafd_unet_create(int device, unsigned long long seed, phafd_unet **out) {
  if (!out) return PHAFD_E_ARG;
  *out = nullptr;
  auto u = std::make_unique<phafd_unet>();
  const int rc = uguard(nullptr, [&] {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw std::invalid_argument("no such HIP device");
    u->net.init(device, seed);
    return 0;
  });
  if (rc != 0) return rc;
  *out = u.release();
  return 0;
}

int phafd_unet_destroy(phafd_unet *u) {
  if (!u) return 0;
  (void)hipSetDevice(u->net.device);
  if (u->net.stream) {
    (void)hipStreamSynchronize(u->net.stream);
    phafd::scan_forget_stream(u->net.stream);
    (void)hipStreamDestroy(u->net.stream);
  }
  delete u;
  return 0;
}

const char *phafd_unet_last_error(const phafd_unet *u) { return u ? u->net.err.c_str() : g_unet_err.c_str(); }

int phafd_unet_nparams(phafd_unet *u, int64_t *n) {
  return uguard(u, [&] {
    *n = (int64_t)boundary_count(u->net);
    return 0;
  });
}

int phafd_unet_get_params(phafd_unet *u, float *host) {
  return uguard(u, [&] {
    std::vector<float> st(u->net.nparam);
    HIPCHK(hipMemcpy(st.data(), u->net.P.p, sizeof(float) * u->net.nparam, hipMemcpyDeviceToHost));
    to_boundary(u->net, st, host);
    return 0;
  });
}

int phafd_unet_set_params(phafd_unet *u, const float *host) {
  return uguard(u, [&] {
    std::vector<float> st;
    to_stored(u->net, host, st);
    phafd::UNet &n = u->net;
    HIPCHK(hipMemcpy(n.P.p, st.data(), sizeof(float) * n.nparam, hipMemcpyHostToDevice));
    phafd::fill_zero(n.G.p, sizeof(float) * n.nparam, n.stream);
    phafd::fill_zero(n.Mo.p, sizeof(float) * n.nparam, n.stream);
    phafd::fill_zero(n.Vo.p, sizeof(float) * n.nparam, n.stream);
    n.adam_t = 0;
    n.pack(n.stream);
    HIPCHK(hipStreamSynchronize(n.stream));
    return 0;
  });
}

int phafd_unet_get_grads(phafd_unet *u, float *host, int zero) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    HIPCHK(hipStreamSynchronize(n.stream));
    std::vector<float> st(n.nparam);
    HIPCHK(hipMemcpy(st.data(), n.G.p, sizeof(float) * n.nparam, hipMemcpyDeviceToHost));
    to_boundary(n, st, host);
    if (zero) {
      phafd::fill_zero(n.G.p, sizeof(float) * n.nparam, n.stream);
      HIPCHK(hipStreamSynchronize(n.stream));
    }
    return 0;
  });
}

int phafd_unet_set_norm(phafd_unet *u, const double *norm5) {
  return uguard(u, [&] {
    for (int k = 0; k < 5; k++) u->net.norm[k] = norm5[k];
    return 0;
  });
}
int phafd_unet_get_norm(phafd_unet *u, double *norm5) {
  return uguard(u, [&] {
    for (int k = 0; k < 5; k++) norm5[k] = u->net.norm[k];
    return 0;
  });
}

int phafd_unet_forward(phafd_unet *u, int nx, int ny, int nz, const float *in4, float *out) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    n.ensure(nx, ny, nz, true);
    const size_t N = (size_t)nx * ny * nz;
    u->hin.alloc(4 * N);
    u->x0.alloc(8 * N);
    HIPCHK(hipMemcpyAsync(u->hin.p, in4, sizeof(float) * 4 * N, hipMemcpyHostToDevice, n.stream));
    phafd::unet_features_from_channels(N, u->hin.p, u->x0.p, n.stream);
    n.forward(u->x0.p, n.A.out.p, n.stream);
    HIPCHK(hipMemcpyAsync(out, n.A.out.p, sizeof(float) * N, hipMemcpyDeviceToHost, n.stream));
    HIPCHK(hipStreamSynchronize(n.stream));
    return 0;
  });
}

int phafd_unet_backward(phafd_unet *u, const float *dout) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    if (!n.A.train || !u->x0.p) throw std::invalid_argument("call phafd_unet_forward first");
    const size_t N = (size_t)n.A.nx * n.A.ny * n.A.nz;
    HIPCHK(hipMemcpyAsync(n.A.dout.p, dout, sizeof(float) * N, hipMemcpyHostToDevice, n.stream));
    n.backward(u->x0.p, n.A.dout.p, n.stream);
    HIPCHK(hipStreamSynchronize(n.stream));
    return 0;
  });
}

int phafd_unet_data_alloc(phafd_unet *u, int nx, int ny, int nz, int nsamples) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    if (nsamples < 1) throw std::invalid_argument("need at least one sample");
    if (!phafd::unet_shape_ok(nx, ny, nz))
      throw std::invalid_argument("surrogate grid must have nx % 64 == 0, ny % 32 == 0, nz % 4 == 0");
    const size_t N = (size_t)nx * ny * nz;
    n.dsX.release();
    n.dsY.release();
    n.dsX.alloc(N * 8 * nsamples);
    n.dsY.alloc(N * nsamples);
    n.ds_n = nsamples;
    n.ds_nx = nx; n.ds_ny = ny; n.ds_nz = nz;
    return 0;
  });
}

int phafd_unet_data_put(phafd_unet *u, int i, const float *in4, const float *target) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    if (i < 0 || i >= n.ds_n) throw std::invalid_argument("sample index out of range");
    const size_t N = (size_t)n.ds_nx * n.ds_ny * n.ds_nz;
    u->hin.alloc(4 * N);
    HIPCHK(hipMemcpyAsync(u->hin.p, in4, sizeof(float) * 4 * N, hipMemcpyHostToDevice, n.stream));
    phafd::unet_features_from_channels(N, u->hin.p, n.dsX.p + (size_t)i * N * 8, n.stream);
    HIPCHK(hipMemcpyAsync(n.dsY.p + (size_t)i * N, target, sizeof(float) * N, hipMemcpyHostToDevice, n.stream));
    HIPCHK(hipStreamSynchronize(n.stream));
    return 0;
  });
}

// nsteps Adam steps over minibatches of `batch` samples drawn from samples
// [0, ntrain); losses[k] = mean square error of step k's minibatch
int phafd_unet_train(phafd_unet *u, int nsteps, int batch, int ntrain, double lr, double *losses) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    if (n.ds_n < 1) throw std::invalid_argument("no training data");
    if (batch < 1 || ntrain < 1 || ntrain > n.ds_n) throw std::invalid_argument("bad batch or training range");
    n.ensure(n.ds_nx, n.ds_ny, n.ds_nz, true);
    const size_t N = (size_t)n.ds_nx * n.ds_ny * n.ds_nz;
    for (int k = 0; k < nsteps; k++) {
      double lsum = 0.0;
      for (int b = 0; b < batch; b++) {
        const int idx = (int)(next_rand(n.rng) % (uint64_t)ntrain);
        const uint16_t *x = n.dsX.p + (size_t)idx * N * 8;
        n.forward(x, n.A.out.p, n.stream);
        lsum += n.loss_grad(n.dsY.p + (size_t)idx * N, 1.0 / ((double)N * batch), n.stream);
        n.backward(x, n.A.dout.p, n.stream);
      }
      n.adam(lr, 0.9, 0.999, 1e-8, n.stream);
      if (losses) losses[k] = lsum / batch;
    }
    HIPCHK(hipStreamSynchronize(n.stream));
    return 0;
  });
}

// mean square error of samples [i0, i0 + count) (no gradients)
int phafd_unet_eval(phafd_unet *u, int i0, int count, double *losses) {
  return uguard(u, [&] {
    phafd::UNet &n = u->net;
    if (i0 < 0 || count < 0 || i0 + count > n.ds_n) throw std::invalid_argument("sample range out of bounds");
    n.ensure(n.ds_nx, n.ds_ny, n.ds_nz, true);
    const size_t N = (size_t)n.ds_nx * n.ds_ny * n.ds_nz;
    for (int k = 0; k < count; k++) {
      const uint16_t *x = n.dsX.p + (size_t)(i0 + k) * N * 8;
      n.forward(x, n.A.out.p, n.stream);
      losses[k] = n.loss_grad(n.dsY.p + (size_t)(i0 + k) * N, 0.0, n.stream);
    }
    return 0;
  });
}

}  // extern "C"
