// fft.hip -- hand-written 3-D real FFT for gfx950 (FP64).
//
// Replaces the FFTW-MPI r2c / c2r plans of the reference (src/grid.cpp:239-280)
// for one GPU holding the whole box.  Three passes per transform, each one
// coalesced read + one coalesced write of the whole array:
//   r2c : x pass  real[Nz][Ny][Nx] -> K[kx][z][y]  (rows packed as Nx/2 complex,
//                                                   post-processed to Nx/2+1 bins)
//         y pass  c2c along contiguous rows of K
//         z pass  c2c along z in tiles of consecutive ky
//   c2r : the same passes in reverse; the x pass forms the packed half-length
//         spectrum using only Re() of the kx = 0 and kx = Nx/2 bins (the
//         half-complex synthesis rule), then inverse FFT and unpack.
// Inside a pass every line is transformed in LDS by a Stockham autosort FFT
// (radix 8/4/2 stages, one register-resident butterfly per work item, barrier
// between read and write phases).  A block holds 4096 complex values (64 KiB)
// and 256 threads, i.e. 16 complex values per thread per stage.
#include <cmath>

#include "internal.hpp"
#include "spectral_ops.hpp"

namespace phafd {
namespace {

constexpr int THREADS = 256;
constexpr int VPT = 16;
constexpr int LDS_ELEMS = THREADS * VPT;   // 4096 complex doubles = 64 KiB

__host__ __device__ constexpr int fft_radix(int L, int s) {
  int r = L, st = 0;
  while (r > 1) {
    int R = (r % 8 == 0 && r != 16) ? 8 : ((r % 4 == 0) ? 4 : 2);
    if (st == s) return R;
    r /= R;
    st++;
  }
  return 1;
}
__host__ __device__ constexpr int fft_nstages(int L) {
  int r = L, st = 0;
  while (r > 1) {
    int R = (r % 8 == 0 && r != 16) ? 8 : ((r % 4 == 0) ? 4 : 2);
    r /= R;
    st++;
  }
  return st;
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 conjg(double2 a) { return make_double2(a.x, -a.y); }
// multiply by SIGN * i
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 a) {
  return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <int SIGN>
__device__ __forceinline__ void dft2(double2 *v) {
  double2 t = v[0];
  v[0] = cadd(t, v[1]);
  v[1] = csub(t, v[1]);
}
template <int SIGN>
__device__ __forceinline__ void dft4(double2 *v) {
  double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  double2 c = cadd(v[1], v[3]), d = mul_si<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}
template <int SIGN>
__device__ __forceinline__ void dft8(double2 *v) {
  double2 e[4] = {v[0], v[2], v[4], v[6]};
  double2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<SIGN>(e);
  dft4<SIGN>(o);
  const double h = 0.70710678118654752440;
  double2 t1 = cmul(o[1], make_double2(h, SIGN * h));
  double2 t2 = mul_si<SIGN>(o[2]);
  double2 t3 = cmul(o[3], make_double2(-h, SIGN * h));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], t1);
  v[5] = csub(e[1], t1);
  v[2] = cadd(e[2], t2);
  v[6] = csub(e[2], t2);
  v[3] = cadd(e[3], t3);
  v[7] = csub(e[3], t3);
}
// 16-point DFT as 4 x 4 (natural-order output)
template <int SIGN>
__device__ __forceinline__ void dft16(double2 *v) {
  const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
  double2 b[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
    double2 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    dft4<SIGN>(t);
    b[n2][0] = t[0];
    if (n2 == 0) {
      b[0][1] = t[1]; b[0][2] = t[2]; b[0][3] = t[3];
    } else if (n2 == 1) {   // W16^{1,2,3}
      b[1][1] = cmul(t[1], make_double2(c1, SIGN * s1));
      b[1][2] = cmul(t[2], make_double2(h, SIGN * h));
      b[1][3] = cmul(t[3], make_double2(s1, SIGN * c1));
    } else if (n2 == 2) {   // W16^{2,4,6}
      b[2][1] = cmul(t[1], make_double2(h, SIGN * h));
      b[2][2] = mul_si<SIGN>(t[2]);
      b[2][3] = cmul(t[3], make_double2(-h, SIGN * h));
    } else {                // W16^{3,6,9}
      b[3][1] = cmul(t[1], make_double2(s1, SIGN * c1));
      b[3][2] = cmul(t[2], make_double2(-h, SIGN * h));
      b[3][3] = cmul(t[3], make_double2(-c1, -SIGN * s1));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    double2 u[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
    dft4<SIGN>(u);
    v[k1] = u[0];
    v[k1 + 4] = u[1];
    v[k1 + 8] = u[2];
    v[k1 + 12] = u[3];
  }
}
template <int R, int SIGN>
__device__ __forceinline__ void dft(double2 *v) {
  if constexpr (R == 2) dft2<SIGN>(v);
  else if constexpr (R == 4) dft4<SIGN>(v);
  else if constexpr (R == 8) dft8<SIGN>(v);
  else dft16<SIGN>(v);
}

// One Stockham stage over `nlines` lines of length L in LDS (line stride LS).
// tw[m] = exp(SIGN * 2 pi i m / L).
template <int L, int R, int NS, int SIGN, int LS>
__device__ __forceinline__ void stage(double2 *buf, int nlines, const double2 *__restrict__ tw) {
  constexpr int NI = (VPT / R) > 0 ? VPT / R : 1;
  constexpr int IPL = L / R;
  const int total = nlines * IPL;
  double2 v[NI][R];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int w = threadIdx.x + it * THREADS;
    if (w < total) {
      const int line = w / IPL, j = w - line * IPL;
      const double2 *b = buf + line * LS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = b[j + r * IPL];
      if constexpr (NS > 1) {
        const int k = j & (NS - 1);
        const int st = k * (L / (NS * R));
#pragma unroll
        for (int r = 1; r < R; ++r) v[it][r] = cmul(v[it][r], tw[st * r]);
      }
      dft<R, SIGN>(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int w = threadIdx.x + it * THREADS;
    if (w < total) {
      const int line = w / IPL, j = w - line * IPL;
      double2 *b = buf + line * LS;
      const int k = j & (NS - 1);
      const int idxD = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) b[idxD + r * NS] = v[it][r];
    }
  }
  __syncthreads();
}

template <int L, int SIGN, int LS, int S = 0, int NS = 1>
__device__ __forceinline__ void fft_lines(double2 *buf, int nlines, const double2 *__restrict__ tw) {
  if constexpr (S < fft_nstages(L)) {
    constexpr int R = fft_radix(L, S);
    stage<L, R, NS, SIGN, LS>(buf, nlines, tw);
    fft_lines<L, SIGN, LS, S + 1, NS * R>(buf, nlines, tw);
  }
}

// ------------------------------------------------------------- x passes
template <int H>
struct XRows {
  static constexpr int ROWS = (LDS_ELEMS / H) > 256 ? 256 : (LDS_ELEMS / H);
};

__device__ __forceinline__ double fh_point(double &p, double m, double tv, double chi,
                                           unsigned &nl, unsigned &ng, bool &wrote) {
  // FixGridFloryHuggins::post_force, src/fixgrid_floryhuggins.cpp:71-83
  if (p <= 0) {
    p = -p;
    nl++;
    wrote = true;
  } else if (p >= 1) {
    p = 1 - (p - 1);
    ng++;
    wrote = true;
  }
  return m + tv * (log(p / (1 - p)) + chi * (1 - 2 * p));
}

template <int H>
__global__ __launch_bounds__(THREADS) void k_r2c_x(const double *__restrict__ in,
                                                   double2 *__restrict__ out, int ny, int nz,
                                                   FHPre pre, const double2 *__restrict__ twh,
                                                   const double2 *__restrict__ twx) {
  constexpr int NX = 2 * H;
  constexpr int ROWS = XRows<H>::ROWS;
  __shared__ double2 buf[ROWS * H];
  __shared__ unsigned cnt[2];
  const int ty = ny < ROWS ? ny : ROWS;
  const int ntile = ny / ty;
  const int z = blockIdx.x / ntile;
  const int y0 = (blockIdx.x - z * ntile) * ty;
  if (pre.on && threadIdx.x < 2) cnt[threadIdx.x] = 0;
  if (pre.on) __syncthreads();
  const int total = ty * H;
  unsigned nl = 0, ng = 0;
  for (int w = threadIdx.x; w < total; w += THREADS) {
    const int r = w / H, n = w - r * H;
    const size_t off = ((size_t)z * ny + y0 + r) * NX + 2 * n;
    double2 v;
    if (!pre.on) {
      v = *reinterpret_cast<const double2 *>(in + off);
    } else {
      double2 p = *reinterpret_cast<const double2 *>(pre.phi + off);
      double2 m = *reinterpret_cast<const double2 *>(pre.mu + off);
      bool wrote = false;
      v.x = fh_point(p.x, m.x, pre.tv, pre.chi, nl, ng, wrote);
      v.y = fh_point(p.y, m.y, pre.tv, pre.chi, nl, ng, wrote);
      if (wrote) *reinterpret_cast<double2 *>(pre.phi + off) = p;
      *reinterpret_cast<double2 *>(pre.mu + off) = v;
    }
    buf[r * H + n] = v;
  }
  if (pre.on) {
    if (nl) atomicAdd(&cnt[0], nl);
    if (ng) atomicAdd(&cnt[1], ng);
  }
  __syncthreads();
  fft_lines<H, -1, H>(buf, ty, twh);
  const int tot2 = ty * (H + 1);
  for (int w = threadIdx.x; w < tot2; w += THREADS) {
    const int k = w / ty, r = w - k * ty;
    const double2 zk = buf[r * H + (k == H ? 0 : k)];
    const double2 zc = conjg(buf[r * H + (k == 0 ? 0 : H - k)]);
    const double2 s = cadd(zk, zc);
    const double2 wd = cmul(twx[k], csub(zk, zc));
    out[((size_t)k * nz + z) * ny + y0 + r] = make_double2(0.5 * (s.x + wd.y), 0.5 * (s.y - wd.x));
  }
  if (pre.on && threadIdx.x == 0) {
    if (cnt[0]) atomicAdd(&pre.counters[0], (unsigned long long)cnt[0]);
    if (cnt[1]) atomicAdd(&pre.counters[1], (unsigned long long)cnt[1]);
  }
}

struct PtrBatch {
  double2 *c[4];
  double *r[4];
  double div[4];   // c2r: divide the output by div[b] when non-zero
};

template <int H>
__global__ __launch_bounds__(THREADS) void k_c2r_x(PtrBatch pb, int ny, int nz,
                                                   const double2 *__restrict__ twh,
                                                   const double2 *__restrict__ twx) {
  constexpr int NX = 2 * H;
  constexpr int ROWS = XRows<H>::ROWS;
  __shared__ double2 buf[ROWS * H];
  __shared__ double2 nyq[ROWS];
  const double2 *__restrict__ in = pb.c[blockIdx.y];
  double *__restrict__ out = pb.r[blockIdx.y];
  const double divisor = pb.div[blockIdx.y];
  const int ty = ny < ROWS ? ny : ROWS;
  const int ntile = ny / ty;
  const int z = blockIdx.x / ntile;
  const int y0 = (blockIdx.x - z * ntile) * ty;
  const int tot = ty * (H + 1);
  for (int w = threadIdx.x; w < tot; w += THREADS) {
    const int k = w / ty, r = w - k * ty;
    const double2 v = in[((size_t)k * nz + z) * ny + y0 + r];
    if (k < H) buf[r * H + k] = v;
    else nyq[r] = v;
  }
  __syncthreads();
  constexpr int HP = H / 2 + 1;
  const int tot2 = ty * HP;
  for (int w = threadIdx.x; w < tot2; w += THREADS) {
    const int r = w / HP, k = w - r * HP;
    double2 *b = buf + r * H;
    if (k == 0) {
      const double a = b[0].x, c = nyq[r].x;   // only Re() of kx = 0 and kx = Nx/2
      b[0] = make_double2(a + c, a - c);
    } else {
      const int k2 = H - k;
      const double2 xk = b[k], xm = b[k2];
      const double2 e1 = cadd(xk, conjg(xm));
      const double2 o1 = cmul(csub(xk, conjg(xm)), twx[k]);
      const double2 z1 = make_double2(e1.x - o1.y, e1.y + o1.x);
      if (k2 != k) {
        const double2 e2 = cadd(xm, conjg(xk));
        const double2 o2 = cmul(csub(xm, conjg(xk)), twx[k2]);
        b[k2] = make_double2(e2.x - o2.y, e2.y + o2.x);
      }
      b[k] = z1;
    }
  }
  __syncthreads();
  fft_lines<H, +1, H>(buf, ty, twh);
  const int tot3 = ty * H;
  for (int w = threadIdx.x; w < tot3; w += THREADS) {
    const int r = w / H, n = w - r * H;
    double2 v = buf[r * H + n];
    if (divisor != 0.0) {
      v.x = v.x / divisor;
      v.y = v.y / divisor;
    }
    *reinterpret_cast<double2 *>(out + ((size_t)z * ny + y0 + r) * NX + 2 * n) = v;
  }
}

// ------------------------------------------------------------- y pass
template <int NY, int SIGN>
__global__ __launch_bounds__(THREADS) void k_c2c_y(PtrBatch pb, int nrows,
                                                   const double2 *__restrict__ tw) {
  constexpr int ROWS = LDS_ELEMS / NY;
  __shared__ double2 buf[ROWS * NY];
  double2 *__restrict__ data = pb.c[blockIdx.y];
  const int row0 = blockIdx.x * ROWS;
  const int nr = (nrows - row0) < ROWS ? (nrows - row0) : ROWS;
  const int total = nr * NY;
  const size_t base = (size_t)row0 * NY;
  for (int w = threadIdx.x; w < total; w += THREADS) buf[w] = data[base + w];
  __syncthreads();
  fft_lines<NY, SIGN, NY>(buf, nr, tw);
  for (int w = threadIdx.x; w < total; w += THREADS) data[base + w] = buf[w];
}

// ------------------------------------------------------------- z pass
template <int NZ>
struct ZTile {
  static constexpr int LINES = LDS_ELEMS / NZ;
  static constexpr int LS = NZ + 1;   // padded line stride (LDS bank spread)
};

template <int NZ, int SIGN>
__global__ __launch_bounds__(THREADS) void k_c2c_z(PtrBatch pb, int ny, int hx,
                                                   const double2 *__restrict__ tw) {
  constexpr int LINES = ZTile<NZ>::LINES;
  constexpr int LS = ZTile<NZ>::LS;
  __shared__ double2 buf[LINES * LS];
  double2 *__restrict__ data = pb.c[blockIdx.y];
  const int ly = ny < LINES ? ny : LINES;
  const int lk = LINES / ly;
  const int nyt = ny / ly;
  const int kxt = blockIdx.x / nyt;
  const int y0 = (blockIdx.x - kxt * nyt) * ly;
  const int kx0 = kxt * lk;
  const int lkv = (hx - kx0) < lk ? (hx - kx0) : lk;
  const int nl = ly * lkv;
  const int total = nl * NZ;
  for (int w = threadIdx.x; w < total; w += THREADS) {
    const int z = w / nl, l = w - z * nl;
    const int kx = kx0 + l / ly, y = y0 + (l % ly);
    buf[l * LS + z] = data[((size_t)kx * NZ + z) * ny + y];
  }
  __syncthreads();
  fft_lines<NZ, SIGN, LS>(buf, nl, tw);
  for (int w = threadIdx.x; w < total; w += THREADS) {
    const int z = w / nl, l = w - z * nl;
    const int kx = kx0 + l / ly, y = y0 + (l % ly);
    data[((size_t)kx * NZ + z) * ny + y] = buf[l * LS + z];
  }
}

// ------------------------------------------------- two-level line FFTs
// A line of length L = R1 * R2 is transformed with one radix-R1 DFT in
// registers, one twiddle, one padded LDS exchange and one radix-R2 DFT:
//
//  form A (input in natural order): item (line l, n2) holds x[R2*n1 + n2]
//    (n1 < R1), does DFT_R1 over n1, multiplies by W_L^(n2*k1), writes row k1
//    of an LDS tile; after a barrier item (l, k1) reads the R2 values of row
//    k1 and does DFT_R2, giving X[k1 + R1*k2] (k2 < R2) in its registers.
//  form B (input in the order form A leaves it): item (l, n1) holds
//    x[n1 + R1*n2] (n2 < R2), does DFT_R2 over n2, multiplies by
//    W_L^(n1*k2), exchanges, and item (l, k2) does DFT_R1 over n1, giving
//    X[k2 + R2*k1] (k1 < R1): natural order again, held strided.
//
// So a forward transform in form A followed by a pointwise map and an inverse
// in form B needs no reordering in between, which is how the z pass fuses the
// gradient multiply and the spectral update.  One LDS round trip per line
// instead of one per Stockham stage.
template <int R1_, int R2_> struct G2 {
  static constexpr int R1 = R1_, R2 = R2_;
  static constexpr int NL = 256 / (R1 > R2 ? R1 : R2);   // lines per block
  static constexpr int RS = R2 + 1;                      // form A tile row (padded)
  static constexpr int LS = R1 * RS + 1;
  static constexpr int RSB = R1 + 1;                     // form B tile row (padded)
  static constexpr int LSB = R2 * RSB + 1;
  static constexpr int LT = LS > LSB ? LS : LSB;         // line stride of the tile
};
template <int L> struct Split;   // c2c passes and the c2r x pass
template <> struct Split<16> { using G = G2<4, 4>; };
template <> struct Split<32> { using G = G2<8, 4>; };
template <> struct Split<64> { using G = G2<8, 8>; };
template <> struct Split<128> { using G = G2<16, 8>; };
template <> struct Split<256> { using G = G2<16, 16>; };
template <int L> struct SplitR2C;   // r2c x pass: every item loads in form A
template <> struct SplitR2C<16> { using G = G2<4, 4>; };
template <> struct SplitR2C<32> { using G = G2<4, 8>; };
template <> struct SplitR2C<64> { using G = G2<8, 8>; };
template <> struct SplitR2C<128> { using G = G2<8, 16>; };
template <> struct SplitR2C<256> { using G = G2<16, 16>; };
constexpr int nl2(int L) {
  return L == 16 ? 64 : L == 32 ? 32 : L == 64 ? 32 : L == 128 ? 16 : L == 256 ? 16 : 0;
}
bool has2(int L) { return nl2(L) > 0; }

template <class G, int SIGN>
__device__ __forceinline__ void stepA1(double2 *v, double2 *xb, int l, int n2, const double2 *__restrict__ tw) {
  dft<G::R1, SIGN>(v);
#pragma unroll
  for (int k1 = 1; k1 < G::R1; k1++) v[k1] = cmul(v[k1], tw[n2 * k1]);
#pragma unroll
  for (int k1 = 0; k1 < G::R1; k1++) xb[l * G::LT + k1 * G::RS + n2] = v[k1];
}
template <class G, int SIGN>
__device__ __forceinline__ void stepA2(double2 *o, const double2 *xb, int l, int k1) {
#pragma unroll
  for (int n2 = 0; n2 < G::R2; n2++) o[n2] = xb[l * G::LT + k1 * G::RS + n2];
  dft<G::R2, SIGN>(o);
}
template <class G, int SIGN>
__device__ __forceinline__ void stepB1(double2 *v, double2 *xb, int l, int n1, const double2 *__restrict__ tw) {
  dft<G::R2, SIGN>(v);
#pragma unroll
  for (int k2 = 1; k2 < G::R2; k2++) v[k2] = cmul(v[k2], tw[n1 * k2]);
#pragma unroll
  for (int k2 = 0; k2 < G::R2; k2++) xb[l * G::LT + k2 * G::RSB + n1] = v[k2];
}
template <class G, int SIGN>
__device__ __forceinline__ void stepB2(double2 *o, const double2 *xb, int l, int k2) {
#pragma unroll
  for (int n1 = 0; n1 < G::R1; n1++) o[n1] = xb[l * G::LT + k2 * G::RSB + n1];
  dft<G::R1, SIGN>(o);
}

// y pass: contiguous rows of length L
template <int L, int SIGN>
__global__ __launch_bounds__(256) void k2_rows(PtrBatch pb, int nrows, const double2 *__restrict__ tw) {
  using G = typename Split<L>::G;
  __shared__ double2 xb[G::NL * G::LT];
  double2 *__restrict__ data = pb.c[blockIdx.y];
  const int row0 = blockIdx.x * G::NL;
  const int tid = threadIdx.x;
  {
    const int n2 = tid % G::R2, l = tid / G::R2;
    if (tid < G::NL * G::R2 && row0 + l < nrows) {
      const double2 *src = data + (size_t)(row0 + l) * L;
      double2 v[G::R1];
#pragma unroll
      for (int n1 = 0; n1 < G::R1; n1++) v[n1] = src[G::R2 * n1 + n2];
      stepA1<G, SIGN>(v, xb, l, n2, tw);
    }
  }
  __syncthreads();
  {
    const int k1 = tid % G::R1, l = tid / G::R1;
    if (tid < G::NL * G::R1 && row0 + l < nrows) {
      double2 o[G::R2];
      stepA2<G, SIGN>(o, xb, l, k1);
      double2 *dst = data + (size_t)(row0 + l) * L;
#pragma unroll
      for (int k2 = 0; k2 < G::R2; k2++) dst[k1 + G::R1 * k2] = o[k2];
    }
  }
}

// z pass: lines along z, NL consecutive ky of one kx per block; the planes
// handled are kx = kx0 + kxstep * (blockIdx.x / (ny / NL))
template <int L, int SIGN>
__global__ __launch_bounds__(256) void k2_cols(PtrBatch pb, int ny, int kx0, int kxstep,
                                               const double2 *__restrict__ tw) {
  using G = typename Split<L>::G;
  __shared__ double2 xb[G::NL * G::LT];
  double2 *__restrict__ data = pb.c[blockIdx.y];
  const int nyt = ny / G::NL;
  const int pl = blockIdx.x / nyt, y0 = (blockIdx.x - pl * nyt) * G::NL;
  const int kx = kx0 + kxstep * pl;
  const size_t base = (size_t)kx * L * ny + y0;
  const int tid = threadIdx.x;
  {
    const int l = tid % G::NL, n2 = tid / G::NL;
    if (tid < G::NL * G::R2) {
      double2 v[G::R1];
#pragma unroll
      for (int n1 = 0; n1 < G::R1; n1++) v[n1] = data[base + (size_t)(G::R2 * n1 + n2) * ny + l];
      stepA1<G, SIGN>(v, xb, l, n2, tw);
    }
  }
  __syncthreads();
  {
    const int l = tid % G::NL, k1 = tid / G::NL;
    if (tid < G::NL * G::R1) {
      double2 o[G::R2];
      stepA2<G, SIGN>(o, xb, l, k1);
#pragma unroll
      for (int k2 = 0; k2 < G::R2; k2++) data[base + (size_t)(k1 + G::R1 * k2) * ny + l] = o[k2];
    }
  }
}

// z pass of phi's r2c fused with the z component of the gradient
// (FixGridConjugate::start_of_step, src/fixgrid_conjugate.cpp:62-116):
// forward z FFT of the y-transformed phi (phihat written in place), times
// (0, dq) * signed kz, inverse z FFT written to `dz`.  Items (l, k1) hold
// kz = k1 + R1*k2 between the two transforms.
template <int L>
__global__ __launch_bounds__(256) void k2_zd(double2 *__restrict__ ph, double2 *__restrict__ dz, int ny,
                                             double dq, const double2 *__restrict__ twf,
                                             const double2 *__restrict__ twb) {
  using G = typename Split<L>::G;
  static_assert(G::NL * G::R1 == 256 && G::NL * G::R2 == 256, "square split expected");
  __shared__ double2 xb[G::NL * G::LT];
  const int nyt = ny / G::NL;
  const int kx = blockIdx.x / nyt, y0 = (blockIdx.x - kx * nyt) * G::NL;
  const size_t base = (size_t)kx * L * ny + y0;
  const int tid = threadIdx.x;
  const int l = tid % G::NL, r = tid / G::NL;
  {
    double2 v[G::R1];
#pragma unroll
    for (int n1 = 0; n1 < G::R1; n1++) v[n1] = ph[base + (size_t)(G::R2 * n1 + r) * ny + l];
    stepA1<G, -1>(v, xb, l, r, twf);
  }
  __syncthreads();
  double2 v[G::R2];
  stepA2<G, -1>(v, xb, l, r);
#pragma unroll
  for (int k2 = 0; k2 < G::R2; k2++) {
    const int kz = r + G::R1 * k2;
    ph[base + (size_t)kz * ny + l] = v[k2];
    v[k2] = grad_mode(v[k2], dq, (double)signed_k(kz, L));
  }
  __syncthreads();
  stepB1<G, +1>(v, xb, l, r, twb);
  __syncthreads();
  double2 o[G::R1];
  stepB2<G, +1>(o, xb, l, r);
#pragma unroll
  for (int k1 = 0; k1 < G::R1; k1++) dz[base + (size_t)(r + G::R2 * k1) * ny + l] = o[k1];
}

// Fused z pass of the conjugate step (Conjugate::update,
// src/conjugate.cpp:163-327, for the bulk planes 1 <= kx < Nx/2 where every
// mode updates itself): forward z FFT of the y-transformed mu (muhat written
// in place), mode update of phihat, inverse z FFT of the result written back
// over phihat (the c2r consumes it, as the reference's c2r destroys its input).
template <int L>
__global__ __launch_bounds__(256) void k2_zupdate(double2 *__restrict__ mu, double2 *__restrict__ ph,
                                                  const double2 *__restrict__ noise, int ny,
                                                  UpdateParams u, const double2 *__restrict__ twf,
                                                  const double2 *__restrict__ twb) {
  using G = typename Split<L>::G;
  static_assert(G::NL * G::R1 == 256 && G::NL * G::R2 == 256, "square split expected");
  __shared__ double2 xb[G::NL * G::LT];
  const int nyt = ny / G::NL;
  const int kx = 1 + (int)(blockIdx.x / nyt), y0 = (blockIdx.x - (kx - 1) * nyt) * G::NL;
  const size_t base = (size_t)kx * L * ny + y0;
  const int tid = threadIdx.x;
  const int l = tid % G::NL, r = tid / G::NL;
  {
    double2 v[G::R1];
#pragma unroll
    for (int n1 = 0; n1 < G::R1; n1++) v[n1] = mu[base + (size_t)(G::R2 * n1 + r) * ny + l];
    stepA1<G, -1>(v, xb, l, r, twf);
  }
  __syncthreads();
  double2 v[G::R2];
  stepA2<G, -1>(v, xb, l, r);
  const int ky = y0 + l;
#pragma unroll
  for (int k2 = 0; k2 < G::R2; k2++) {
    const size_t idx = base + (size_t)(r + G::R1 * k2) * ny + l;
    mu[idx] = v[k2];
    const double2 nz0 = noise ? noise[idx] : make_double2(0.0, 0.0);
    v[k2] = mode_update(ph[idx], v[k2], nz0, ky, r + G::R1 * k2, kx, ny, L, u);
  }
  __syncthreads();
  stepB1<G, +1>(v, xb, l, r, twb);
  __syncthreads();
  double2 o[G::R1];
  stepB2<G, +1>(o, xb, l, r);
#pragma unroll
  for (int k1 = 0; k1 < G::R1; k1++) ph[base + (size_t)(r + G::R2 * k1) * ny + l] = o[k1];
}

// x pass r2c, first half: NL rows of Nx = 2H reals packed as H complex,
// optional Flory-Huggins pre-op on the load, two-level FFT.  Leaves the
// packed transform Z[k] of row l at xb[l*(H+1) + k] (after a barrier).
template <int H>
__device__ __forceinline__ void r2c_head(const double *__restrict__ in, const FHPre &pre, double2 *xb,
                                         int z, int y0, int ny, unsigned &nl, unsigned &ng,
                                         const double2 *__restrict__ twh) {
  using G = typename SplitR2C<H>::G;
  constexpr int NX = 2 * H, ZS = H + 1;
  static_assert(G::NL * ZS <= G::NL * G::LT, "tile reuse");
  const int tid = threadIdx.x;
  {
    const int n2 = tid % G::R2, l = tid / G::R2;
    if (tid < G::NL * G::R2) {
      const size_t rowoff = ((size_t)z * ny + y0 + l) * NX;
      double2 v[G::R1];
      if (!pre.on) {
#pragma unroll
        for (int n1 = 0; n1 < G::R1; n1++)
          v[n1] = *reinterpret_cast<const double2 *>(in + rowoff + 2 * (G::R2 * n1 + n2));
      } else {
        double2 ph[G::R1], m[G::R1];
#pragma unroll
        for (int n1 = 0; n1 < G::R1; n1++) {
          const size_t off = rowoff + 2 * (G::R2 * n1 + n2);
          ph[n1] = *reinterpret_cast<const double2 *>(pre.phi + off);
          m[n1] = *reinterpret_cast<const double2 *>(pre.mu + off);
        }
#pragma unroll
        for (int n1 = 0; n1 < G::R1; n1++) {
          const size_t off = rowoff + 2 * (G::R2 * n1 + n2);
          bool wrote = false;
          double2 q = ph[n1], rr;
          rr.x = fh_point(q.x, m[n1].x, pre.tv, pre.chi, nl, ng, wrote);
          rr.y = fh_point(q.y, m[n1].y, pre.tv, pre.chi, nl, ng, wrote);
          if (wrote) *reinterpret_cast<double2 *>(pre.phi + off) = q;
          *reinterpret_cast<double2 *>(pre.mu + off) = rr;
          v[n1] = rr;
        }
      }
      stepA1<G, -1>(v, xb, l, n2, twh);
    }
  }
  __syncthreads();
  double2 o[G::R2];
  const int k1 = tid % G::R1, l2 = tid / G::R1;
  const bool a2 = tid < G::NL * G::R1;
  if (a2) stepA2<G, -1>(o, xb, l2, k1);
  __syncthreads();
  if (a2) {
#pragma unroll
    for (int k2 = 0; k2 < G::R2; k2++) xb[l2 * ZS + k1 + G::R1 * k2] = o[k2];
  }
  __syncthreads();
}

// r2c post-processing of item w: bin k = w / NL of row l = w % NL
template <int H>
__device__ __forceinline__ double2 r2c_bin(const double2 *xb, int k, int l, const double2 *__restrict__ twx) {
  constexpr int ZS = H + 1;
  const double2 zk = xb[l * ZS + (k == H ? 0 : k)];
  const double2 zc = conjg(xb[l * ZS + (k == 0 ? 0 : H - k)]);
  const double2 sm = cadd(zk, zc);
  const double2 wd = cmul(twx[k], csub(zk, zc));
  return make_double2(0.5 * (sm.x + wd.y), 0.5 * (sm.y - wd.x));
}

// x pass c2r, second half: xb[l*(H+1) + k] holds bins k in [0, H] of NL rows
// (visible to all items); half-complex pre-processing (only Re() of the
// kx = 0 and kx = Nx/2 bins), two-level inverse FFT, unpacked row store.
template <int H>
__device__ __forceinline__ void c2r_tail(double2 *xb, double *__restrict__ out, int z, int y0, int ny,
                                         double divisor, const double2 *__restrict__ twh,
                                         const double2 *__restrict__ twx) {
  using G = typename Split<H>::G;
  constexpr int NX = 2 * H, ZS = H + 1, HP = H / 2 + 1;
  const int tid = threadIdx.x;
  for (int w = tid; w < G::NL * HP; w += 256) {
    const int l = w / HP, k = w - l * HP;
    double2 *b = xb + l * ZS;
    if (k == 0) {
      const double a = b[0].x, c = b[H].x;
      b[0] = make_double2(a + c, a - c);
    } else {
      const int k2 = H - k;
      const double2 xk = b[k], xm = b[k2];
      const double2 e1 = cadd(xk, conjg(xm));
      const double2 o1 = cmul(csub(xk, conjg(xm)), twx[k]);
      const double2 z1 = make_double2(e1.x - o1.y, e1.y + o1.x);
      if (k2 != k) {
        const double2 e2 = cadd(xm, conjg(xk));
        const double2 o2 = cmul(csub(xm, conjg(xk)), twx[k2]);
        b[k2] = make_double2(e2.x - o2.y, e2.y + o2.x);
      }
      b[k] = z1;
    }
  }
  __syncthreads();
  double2 v[G::R1];
  const int n2 = tid % G::R2, l1 = tid / G::R2;
  const bool a1 = tid < G::NL * G::R2;
  if (a1) {
#pragma unroll
    for (int n1 = 0; n1 < G::R1; n1++) v[n1] = xb[l1 * ZS + G::R2 * n1 + n2];
  }
  __syncthreads();
  if (a1) stepA1<G, +1>(v, xb, l1, n2, twh);
  __syncthreads();
  const int k1 = tid % G::R1, l2 = tid / G::R1;
  if (tid < G::NL * G::R1) {
    double2 o[G::R2];
    stepA2<G, +1>(o, xb, l2, k1);
    double2 *row = reinterpret_cast<double2 *>(out + ((size_t)z * ny + y0 + l2) * NX);
#pragma unroll
    for (int k2 = 0; k2 < G::R2; k2++) {
      double2 rr = o[k2];
      if (divisor != 0.0) {
        rr.x = rr.x / divisor;
        rr.y = rr.y / divisor;
      }
      row[k1 + G::R1 * k2] = rr;
    }
  }
}

// x pass r2c into K[kx][z][y] (transposed, NL consecutive y per store run)
template <int H>
__global__ __launch_bounds__(256) void k2_r2c_x(const double *__restrict__ in, double2 *__restrict__ out,
                                                int ny, int nz, FHPre pre, const double2 *__restrict__ twh,
                                                const double2 *__restrict__ twx) {
  using G = typename SplitR2C<H>::G;
  __shared__ double2 xb[G::NL * G::LT];
  __shared__ unsigned cnt[2];
  const int nyt = ny / G::NL;
  const int z = blockIdx.x / nyt, y0 = (blockIdx.x - z * nyt) * G::NL;
  const int tid = threadIdx.x;
  if (pre.on && tid < 2) cnt[tid] = 0;
  if (pre.on) __syncthreads();
  unsigned nl = 0, ng = 0;
  r2c_head<H>(in, pre, xb, z, y0, ny, nl, ng, twh);
  if (pre.on) {
    if (nl) atomicAdd(&cnt[0], nl);
    if (ng) atomicAdd(&cnt[1], ng);
  }
  for (int w = tid; w < G::NL * (H + 1); w += 256) {
    const int k = w / G::NL, l = w - k * G::NL;
    out[((size_t)k * nz + z) * ny + y0 + l] = r2c_bin<H>(xb, k, l, twx);
  }
  if (pre.on) {
    __syncthreads();
    if (tid == 0) {
      if (cnt[0]) atomicAdd(&pre.counters[0], (unsigned long long)cnt[0]);
      if (cnt[1]) atomicAdd(&pre.counters[1], (unsigned long long)cnt[1]);
    }
  }
}

// x pass c2r from K (transposed load, all loads in flight before the LDS
// stores)
template <int H>
__global__ __launch_bounds__(256) void k2_c2r_x(PtrBatch pb, int ny, int nz, const double2 *__restrict__ twh,
                                                const double2 *__restrict__ twx) {
  using G = typename Split<H>::G;
  constexpr int ZS = H + 1, NW = G::NL * (H + 1), NFULL = NW / 256, REM = NW - 256 * NFULL;
  __shared__ double2 xb[G::NL * G::LT];
  const double2 *__restrict__ in = pb.c[blockIdx.y];
  const int nyt = ny / G::NL;
  const int z = blockIdx.x / nyt, y0 = (blockIdx.x - z * nyt) * G::NL;
  const int tid = threadIdx.x;
  double2 t[NFULL], tr = make_double2(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < NFULL; q++) {
    const int w = tid + 256 * q, k = w / G::NL, l = w - k * G::NL;
    t[q] = in[((size_t)k * nz + z) * ny + y0 + l];
  }
  if (REM && tid < REM) {
    const int w = tid + 256 * NFULL, k = w / G::NL, l = w - k * G::NL;
    tr = in[((size_t)k * nz + z) * ny + y0 + l];
  }
#pragma unroll
  for (int q = 0; q < NFULL; q++) {
    const int w = tid + 256 * q, k = w / G::NL, l = w - k * G::NL;
    xb[l * ZS + k] = t[q];
  }
  if (REM && tid < REM) {
    const int w = tid + 256 * NFULL, k = w / G::NL, l = w - k * G::NL;
    xb[l * ZS + k] = tr;
  }
  __syncthreads();
  c2r_tail<H>(xb, pb.r[blockIdx.y], z, y0, ny, pb.div[blockIdx.y], twh, twx);
}

// Spectral x derivative in one pass (the x component of
// FixGridConjugate::start_of_step, src/fixgrid_conjugate.cpp:62-116):
// r2c of the rows of phi (spectrum stored to K for the y and z passes), bins
// times (0, dqx) * kx, c2r, divided by `divisor`.  The multiplier is constant
// along y and z, so the y and z transforms of the reference cancel exactly
// and are left out; only rounding differs.
template <int H>
__global__ __launch_bounds__(256) void k2_r2c_dx(const double *__restrict__ in, double2 *__restrict__ out,
                                                 double *__restrict__ gx, int ny, int nz, double dqx,
                                                 double divisor, const double2 *__restrict__ twhf,
                                                 const double2 *__restrict__ twxf,
                                                 const double2 *__restrict__ twhb,
                                                 const double2 *__restrict__ twxb) {
  using G = typename SplitR2C<H>::G;
  constexpr int ZS = H + 1, NW = G::NL * (H + 1), NQ = (NW + 255) / 256;
  static_assert(G::NL == Split<H>::G::NL && Split<H>::G::LT <= G::LT, "same rows and tile both ways");
  __shared__ double2 xb[G::NL * G::LT];
  const int nyt = ny / G::NL;
  const int z = blockIdx.x / nyt, y0 = (blockIdx.x - z * nyt) * G::NL;
  const int tid = threadIdx.x;
  unsigned nl = 0, ng = 0;
  FHPre none;
  r2c_head<H>(in, none, xb, z, y0, ny, nl, ng, twhf);
  double2 d[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int w = tid + 256 * q;
    if (q + 1 < NQ || w < NW) {
      const int k = w / G::NL, l = w - k * G::NL;
      const double2 X = r2c_bin<H>(xb, k, l, twxf);
      out[((size_t)k * nz + z) * ny + y0 + l] = X;
      d[q] = grad_mode(X, dqx, (double)k);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int w = tid + 256 * q;
    if (q + 1 < NQ || w < NW) {
      const int k = w / G::NL, l = w - k * G::NL;
      xb[l * ZS + k] = d[q];
    }
  }
  __syncthreads();
  c2r_tail<H>(xb, gx, z, y0, ny, divisor, twhb, twxb);
}

// Spectral y derivative of the x-transformed phi, one y line per NL items:
// forward y FFT (stored in place: the y pass of phi's r2c), times
// (0, dq) * signed ky, inverse y FFT written to `dy` (its z transforms
// cancel exactly as in the x derivative).
template <int L>
__global__ __launch_bounds__(256) void k2_rows_d(double2 *__restrict__ data, double2 *__restrict__ dy, int nrows,
                                                 double dq, const double2 *__restrict__ twf,
                                                 const double2 *__restrict__ twb) {
  using G = typename Split<L>::G;
  static_assert(G::NL * G::R1 == 256 && G::NL * G::R2 == 256, "square split expected");
  __shared__ double2 xb[G::NL * G::LT];
  const int row0 = blockIdx.x * G::NL;
  const int tid = threadIdx.x;
  const int r = tid % G::R1, l = tid / G::R1;   // R1 == R2: r is n2, then k1 / n1, then k2
  const bool act = row0 + l < nrows;
  const size_t roff = (size_t)(row0 + l) * L;
  if (act) {
    double2 v[G::R1];
#pragma unroll
    for (int n1 = 0; n1 < G::R1; n1++) v[n1] = data[roff + G::R2 * n1 + r];
    stepA1<G, -1>(v, xb, l, r, twf);
  }
  __syncthreads();
  double2 v[G::R2];
  if (act) {
    stepA2<G, -1>(v, xb, l, r);
#pragma unroll
    for (int k2 = 0; k2 < G::R2; k2++) {
      const int ky = r + G::R1 * k2;
      data[roff + ky] = v[k2];
      v[k2] = grad_mode(v[k2], dq, (double)signed_k(ky, L));
    }
  }
  __syncthreads();
  if (act) stepB1<G, +1>(v, xb, l, r, twb);
  __syncthreads();
  if (act) {
    double2 o[G::R1];
    stepB2<G, +1>(o, xb, l, r);
#pragma unroll
    for (int k1 = 0; k1 < G::R1; k1++) dy[roff + r + G::R2 * k1] = o[k1];
  }
}

// ------------------------------------------------------------ layout
__global__ void k_k_to_canon(const double2 *__restrict__ k, double2 *__restrict__ c, int hx,
                             int ny, int nz) {
  const size_t M = (size_t)hx * ny * nz;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(i % hx);
    const size_t t = i / hx;
    const int ky = (int)(t % ny);
    const int kz = (int)(t / ny);
    c[i] = k[((size_t)kx * nz + kz) * ny + ky];
  }
}
__global__ void k_canon_to_k(const double2 *__restrict__ c, double2 *__restrict__ k, int hx,
                             int ny, int nz) {
  const size_t M = (size_t)hx * ny * nz;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < M;
       i += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(i % hx);
    const size_t t = i / hx;
    const int ky = (int)(t % ny);
    const int kz = (int)(t / ny);
    k[((size_t)kx * nz + kz) * ny + ky] = c[i];
  }
}

// ------------------------------------------------------------ dispatch
#define PHAFD_SIZE_SWITCH(n, MACRO) \
  switch (n) {                      \
    case 2: MACRO(2); break;        \
    case 4: MACRO(4); break;        \
    case 8: MACRO(8); break;        \
    case 16: MACRO(16); break;      \
    case 32: MACRO(32); break;      \
    case 64: MACRO(64); break;      \
    case 128: MACRO(128); break;    \
    case 256: MACRO(256); break;    \
    case 512: MACRO(512); break;    \
    case 1024: MACRO(1024); break;  \
    default: throw std::runtime_error("FFT size must be a power of two in [4, 1024]"); \
  }

void make_table(DBuf<double2> &f, DBuf<double2> &b, int L) {
  std::vector<double2> hf(L), hb(L);
  for (int m = 0; m < L; m++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * m / L;
    hf[m] = make_double2((double)cosl(a), (double)-sinl(a));
    hb[m] = make_double2((double)cosl(a), (double)sinl(a));
  }
  f.alloc(L);
  b.alloc(L);
  HIPCHK(hipMemcpy(f.p, hf.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.p, hb.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
}

template <int H>
int x_blocks(int ny, int nz) {
  constexpr int ROWS = XRows<H>::ROWS;
  const int ty = ny < ROWS ? ny : ROWS;
  return nz * (ny / ty);
}

#define PHAFD_SPLIT_SWITCH(n, MACRO) \
  switch (n) {                       \
    case 16: MACRO(16); break;       \
    case 32: MACRO(32); break;       \
    case 64: MACRO(64); break;       \
    case 128: MACRO(128); break;     \
    case 256: MACRO(256); break;     \
    default: throw std::runtime_error("no two-level split"); \
  }

// fused z kernels need R1 == R2 (every item active in every step)
#define PHAFD_SQUARE_SWITCH(n, MACRO) \
  switch (n) {                        \
    case 16: MACRO(16); break;        \
    case 64: MACRO(64); break;        \
    case 256: MACRO(256); break;      \
    default: throw std::runtime_error("no square two-level split"); \
  }
bool Geo2Square(int L) { return L == 16 || L == 64 || L == 256; }

void run_y(const FFTPlan &p, PtrBatch pb, int batch, int sign, hipStream_t s) {
  const int nrows = p.hx * p.nz;
  const double2 *tw = sign < 0 ? p.tw_y_f.p : p.tw_y_b.p;
  if (has2(p.ny)) {
#define Y2_CASE(NY)                                                                    \
  {                                                                                    \
    dim3 g((nrows + nl2(NY) - 1) / nl2(NY), batch);                                    \
    probe_begin(sign < 0 ? K_Y_FWD : K_Y_INV, s);                                      \
    if (sign < 0) hipLaunchKernelGGL((k2_rows<NY, -1>), g, dim3(256), 0, s, pb, nrows, tw); \
    else hipLaunchKernelGGL((k2_rows<NY, +1>), g, dim3(256), 0, s, pb, nrows, tw);          \
    probe_end(sign < 0 ? K_Y_FWD : K_Y_INV, s, 32.0 * p.M() * batch);                  \
  }
    PHAFD_SPLIT_SWITCH(p.ny, Y2_CASE)
#undef Y2_CASE
    HIPCHK(hipGetLastError());
    return;
  }
#define Y_CASE(NY)                                                                   \
  {                                                                                  \
    constexpr int ROWS = LDS_ELEMS / NY;                                             \
    dim3 g((nrows + ROWS - 1) / ROWS, batch);                                        \
    probe_begin(sign < 0 ? K_Y_FWD : K_Y_INV, s);                                    \
    if (sign < 0) hipLaunchKernelGGL((k_c2c_y<NY, -1>), g, dim3(THREADS), 0, s, pb, nrows, tw); \
    else hipLaunchKernelGGL((k_c2c_y<NY, +1>), g, dim3(THREADS), 0, s, pb, nrows, tw);          \
    probe_end(sign < 0 ? K_Y_FWD : K_Y_INV, s, 32.0 * p.M() * batch);                \
  }
  PHAFD_SIZE_SWITCH(p.ny, Y_CASE)
#undef Y_CASE
  HIPCHK(hipGetLastError());
}

void run_z(const FFTPlan &p, PtrBatch pb, int batch, int sign, hipStream_t s) {
  const double2 *tw = sign < 0 ? p.tw_z_f.p : p.tw_z_b.p;
  if (has2(p.nz) && p.ny % nl2(p.nz) == 0) {
#define Z2_CASE(NZ)                                                                    \
  {                                                                                    \
    dim3 g(p.hx * (p.ny / nl2(NZ)), batch);                                            \
    probe_begin(sign < 0 ? K_Z_FWD : K_Z_INV, s);                                      \
    if (sign < 0) hipLaunchKernelGGL((k2_cols<NZ, -1>), g, dim3(256), 0, s, pb, p.ny, 0, 1, tw); \
    else hipLaunchKernelGGL((k2_cols<NZ, +1>), g, dim3(256), 0, s, pb, p.ny, 0, 1, tw);          \
    probe_end(sign < 0 ? K_Z_FWD : K_Z_INV, s, 32.0 * p.M() * batch);                  \
  }
    PHAFD_SPLIT_SWITCH(p.nz, Z2_CASE)
#undef Z2_CASE
    HIPCHK(hipGetLastError());
    return;
  }
#define Z_CASE(NZ)                                                                   \
  {                                                                                  \
    constexpr int LINES = ZTile<NZ>::LINES;                                          \
    const int ly = p.ny < LINES ? p.ny : LINES;                                      \
    const int lk = LINES / ly;                                                       \
    const int nblk = ((p.hx + lk - 1) / lk) * (p.ny / ly);                           \
    dim3 g(nblk, batch);                                                             \
    probe_begin(sign < 0 ? K_Z_FWD : K_Z_INV, s);                                    \
    if (sign < 0) hipLaunchKernelGGL((k_c2c_z<NZ, -1>), g, dim3(THREADS), 0, s, pb, p.ny, p.hx, tw); \
    else hipLaunchKernelGGL((k_c2c_z<NZ, +1>), g, dim3(THREADS), 0, s, pb, p.ny, p.hx, tw);          \
    probe_end(sign < 0 ? K_Z_FWD : K_Z_INV, s, 32.0 * p.M() * batch);                \
  }
  PHAFD_SIZE_SWITCH(p.nz, Z_CASE)
#undef Z_CASE
  HIPCHK(hipGetLastError());
}

}  // namespace

KernelProbe *g_probe = nullptr;

void probe_begin(int id, hipStream_t s) {
  KernelProbe *p = g_probe;
  if (!p || p->id != id) return;
  if ((int)p->pool.size() < p->used + 2) {
    for (int k = 0; k < 2; k++) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      p->pool.push_back(e);
    }
  }
  HIPCHK(hipEventRecord(p->pool[p->used], s));
}
void probe_end(int id, hipStream_t s, double bytes) {
  KernelProbe *p = g_probe;
  if (!p || p->id != id) return;
  HIPCHK(hipEventRecord(p->pool[p->used + 1], s));
  p->bytes.push_back(bytes);
  p->used += 2;
}
void probe_collect() {
  KernelProbe *p = g_probe;
  if (!p) return;
  for (int k = 0; k + 1 < p->used; k += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, p->pool[k], p->pool[k + 1]));
    p->ms += ms;
    p->total_bytes += p->bytes[k / 2];
    p->launches++;
  }
  p->used = 0;
  p->bytes.clear();
}

void FFTPlan::init(int nx_, int ny_, int nz_) {
  auto ok = [](int n) { return n >= 4 && n <= 1024 && (n & (n - 1)) == 0; };
  if (!ok(nx_) || !ok(ny_) || !ok(nz_))
    throw std::invalid_argument("grid dimensions must be powers of two in [4, 1024]");
  nx = nx_;
  ny = ny_;
  nz = nz_;
  hx = nx / 2 + 1;
  make_table(tw_x_f, tw_x_b, nx);
  make_table(tw_h_f, tw_h_b, nx / 2);
  make_table(tw_y_f, tw_y_b, ny);
  make_table(tw_z_f, tw_z_b, nz);
}

namespace {
void fft_x_r2c(const FFTPlan &p, const double *in, double2 *out, hipStream_t s, const FHPre *pre) {
  FHPre fp;
  if (pre) fp = *pre;
  probe_begin(K_R2C_X, s);
  if (has2(p.nx / 2) && p.ny % nl2(p.nx / 2) == 0) {
#define X2_CASE(HH)                                                                        \
  hipLaunchKernelGGL((k2_r2c_x<HH>), dim3(p.nz * (p.ny / nl2(HH))), dim3(256), 0, s, in, out, p.ny, p.nz, fp, \
                     p.tw_h_f.p, p.tw_x_f.p);
    PHAFD_SPLIT_SWITCH(p.nx / 2, X2_CASE)
#undef X2_CASE
  } else {
#define X_CASE(HH)                                                                            \
  hipLaunchKernelGGL((k_r2c_x<HH>), dim3(x_blocks<HH>(p.ny, p.nz)), dim3(THREADS), 0, s, in, out, \
                     p.ny, p.nz, fp, p.tw_h_f.p, p.tw_x_f.p);
  PHAFD_SIZE_SWITCH(p.nx / 2, X_CASE)
#undef X_CASE
  }
  probe_end(K_R2C_X, s, (fp.on ? 24.0 * p.N() : 8.0 * p.N()) + 16.0 * p.M());
  HIPCHK(hipGetLastError());
}

void fft_x_c2r(const FFTPlan &p, PtrBatch pb, int batch, hipStream_t s) {
  probe_begin(K_C2R_X, s);
  if (has2(p.nx / 2) && p.ny % nl2(p.nx / 2) == 0) {
#define X2_CASE(HH)                                                                          \
  hipLaunchKernelGGL((k2_c2r_x<HH>), dim3(p.nz * (p.ny / nl2(HH)), batch), dim3(256), 0, s, pb, p.ny, p.nz, \
                     p.tw_h_b.p, p.tw_x_b.p);
    PHAFD_SPLIT_SWITCH(p.nx / 2, X2_CASE)
#undef X2_CASE
  } else {
#define X_CASE(HH)                                                                              \
  hipLaunchKernelGGL((k_c2r_x<HH>), dim3(x_blocks<HH>(p.ny, p.nz), batch), dim3(THREADS), 0, s, pb, \
                     p.ny, p.nz, p.tw_h_b.p, p.tw_x_b.p);
  PHAFD_SIZE_SWITCH(p.nx / 2, X_CASE)
#undef X_CASE
  }
  probe_end(K_C2R_X, s, (16.0 * p.M() + 8.0 * p.N()) * batch);
  HIPCHK(hipGetLastError());
}
}  // namespace

void fft_r2c(const FFTPlan &p, const double *in, double2 *out, hipStream_t s, const FHPre *pre) {
  fft_x_r2c(p, in, out, s, pre);
  PtrBatch pb{};
  pb.c[0] = out;
  run_y(p, pb, 1, -1, s);
  run_z(p, pb, 1, -1, s);
}

void fft_c2r_batch(const FFTPlan &p, double2 *const *in, double *const *out, int batch,
                   double divisor, hipStream_t s) {
  if (batch < 1 || batch > 4) throw std::invalid_argument("fft batch must be 1..4");
  PtrBatch pb{};
  for (int b = 0; b < batch; b++) {
    pb.c[b] = in[b];
    pb.r[b] = out[b];
    pb.div[b] = divisor;
  }
  run_z(p, pb, batch, +1, s);
  run_y(p, pb, batch, +1, s);
  fft_x_c2r(p, pb, batch, s);
}

void fft_c2r(const FFTPlan &p, double2 *in, double *out, double divisor, hipStream_t s) {
  fft_c2r_batch(p, &in, &out, 1, divisor, s);
}

bool fft_fused_z_ok(const FFTPlan &p) {
  return has2(p.nz) && Geo2Square(p.nz) && p.ny % nl2(p.nz) == 0 && p.hx >= 3;
}

void fft_r2c_xy(const FFTPlan &p, const double *in, double2 *out, hipStream_t s, const FHPre *pre) {
  fft_x_r2c(p, in, out, s, pre);
  PtrBatch pb{};
  pb.c[0] = out;
  run_y(p, pb, 1, -1, s);
}

bool fft_sep_grad_ok(const FFTPlan &p) {
  const int H = p.nx / 2;
  return fft_fused_z_ok(p) && Geo2Square(p.ny) && has2(H) && p.ny % nl2(H) == 0;
}

void fft_grad_sep(const FFTPlan &p, const double *phi, double2 *ph, double2 *gy_k, double2 *gz_k,
                  double *gx, double *gy, double *gz, SpecConst sc, hipStream_t s) {
  if (!fft_sep_grad_ok(p)) throw std::logic_error("separable gradient not available for this grid");
  const double N = (double)p.nx * p.ny * p.nz;
  probe_begin(K_R2C_X, s);
#define DX_CASE(HH)                                                                                     \
  hipLaunchKernelGGL((k2_r2c_dx<HH>), dim3(p.nz * (p.ny / nl2(HH))), dim3(256), 0, s, phi, ph, gx, p.ny, p.nz, \
                     sc.dqx, (double)p.nx, p.tw_h_f.p, p.tw_x_f.p, p.tw_h_b.p, p.tw_x_b.p);
  PHAFD_SPLIT_SWITCH(p.nx / 2, DX_CASE)
#undef DX_CASE
  probe_end(K_R2C_X, s, 16.0 * p.N() + 16.0 * p.M());
  HIPCHK(hipGetLastError());
  const int nrows = p.hx * p.nz;
  probe_begin(K_Y_FWD, s);
#define DY_CASE(NY)                                                                                    \
  hipLaunchKernelGGL((k2_rows_d<NY>), dim3((nrows + nl2(NY) - 1) / nl2(NY)), dim3(256), 0, s, ph, gy_k, nrows, \
                     sc.dqz_ky, p.tw_y_f.p, p.tw_y_b.p);
  PHAFD_SQUARE_SWITCH(p.ny, DY_CASE)
#undef DY_CASE
  probe_end(K_Y_FWD, s, 48.0 * p.M());
  HIPCHK(hipGetLastError());
  probe_begin(K_Z_GRAD, s);
#define DZ_CASE(NZ)                                                                                    \
  hipLaunchKernelGGL((k2_zd<NZ>), dim3(p.hx * (p.ny / nl2(NZ))), dim3(256), 0, s, ph, gz_k, p.ny, sc.dqy_kz, \
                     p.tw_z_f.p, p.tw_z_b.p);
  PHAFD_SQUARE_SWITCH(p.nz, DZ_CASE)
#undef DZ_CASE
  probe_end(K_Z_GRAD, s, 48.0 * p.M());
  HIPCHK(hipGetLastError());
  PtrBatch pz{};
  pz.c[0] = gz_k;
  run_y(p, pz, 1, +1, s);
  PtrBatch pb{};
  pb.c[0] = gy_k;
  pb.r[0] = gy;
  pb.div[0] = (double)p.nx * p.ny;
  pb.c[1] = gz_k;
  pb.r[1] = gz;
  pb.div[1] = N;
  fft_x_c2r(p, pb, 2, s);
}

void fft_z_update(const FFTPlan &p, double2 *mu, double2 *ph, const double2 *noise, UpdateParams u,
                  hipStream_t s) {
  if (!fft_fused_z_ok(p)) throw std::logic_error("fused z pass not available for this grid");
  const size_t mbulk = (size_t)(p.hx - 2) * p.nz * p.ny;
  probe_begin(K_Z_UPDATE, s);
#define ZU_CASE(NZ)                                                                                  \
  hipLaunchKernelGGL((k2_zupdate<NZ>), dim3((p.hx - 2) * (p.ny / nl2(NZ))), dim3(256), 0, s, mu, ph, noise, \
                     p.ny, u, p.tw_z_f.p, p.tw_z_b.p);
  PHAFD_SQUARE_SWITCH(p.nz, ZU_CASE)
#undef ZU_CASE
  probe_end(K_Z_UPDATE, s, (noise ? 80.0 : 64.0) * mbulk);
  HIPCHK(hipGetLastError());
  // planes kx = 0 and kx = Nx/2: z pass of mu, mirror-aware update, inverse z pass
  PtrBatch pm{}, pp{};
  pm.c[0] = mu;
  pp.c[0] = ph;
  const int H = p.hx - 1;
#define ZE_CASE(NZ)                                                                                     \
  {                                                                                                     \
    dim3 g(2 * (p.ny / nl2(NZ)));                                                                       \
    hipLaunchKernelGGL((k2_cols<NZ, -1>), g, dim3(256), 0, s, pm, p.ny, 0, H, p.tw_z_f.p);              \
    HIPCHK(hipGetLastError());                                                                          \
    spectral_update_edges(p, ph, mu, noise, u, s);                                                      \
    hipLaunchKernelGGL((k2_cols<NZ, +1>), g, dim3(256), 0, s, pp, p.ny, 0, H, p.tw_z_b.p);              \
  }
  PHAFD_SQUARE_SWITCH(p.nz, ZE_CASE)
#undef ZE_CASE
  HIPCHK(hipGetLastError());
}

void fft_c2r_yx_batch(const FFTPlan &p, double2 *const *in, double *const *out, int batch,
                      double divisor, hipStream_t s) {
  if (batch < 1 || batch > 4) throw std::invalid_argument("fft batch must be 1..4");
  PtrBatch pb{};
  for (int b = 0; b < batch; b++) {
    pb.c[b] = in[b];
    pb.r[b] = out[b];
    pb.div[b] = divisor;
  }
  run_y(p, pb, batch, +1, s);
  fft_x_c2r(p, pb, batch, s);
}

void spec_k_to_canon(const FFTPlan &p, const double2 *k, double2 *canon, hipStream_t s) {
  hipLaunchKernelGGL(k_k_to_canon, dim3(1024), dim3(256), 0, s, k, canon, p.hx, p.ny, p.nz);
  HIPCHK(hipGetLastError());
}
void spec_canon_to_k(const FFTPlan &p, const double2 *canon, double2 *k, hipStream_t s) {
  hipLaunchKernelGGL(k_canon_to_k, dim3(1024), dim3(256), 0, s, canon, k, p.hx, p.ny, p.nz);
  HIPCHK(hipGetLastError());
}

}  // namespace phafd
