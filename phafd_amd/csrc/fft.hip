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
template <int R, int SIGN>
__device__ __forceinline__ void dft(double2 *v) {
  if constexpr (R == 2) dft2<SIGN>(v);
  else if constexpr (R == 4) dft4<SIGN>(v);
  else dft8<SIGN>(v);
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
};

template <int H>
__global__ __launch_bounds__(THREADS) void k_c2r_x(PtrBatch pb, int ny, int nz, double divisor,
                                                   const double2 *__restrict__ twh,
                                                   const double2 *__restrict__ twx) {
  constexpr int NX = 2 * H;
  constexpr int ROWS = XRows<H>::ROWS;
  __shared__ double2 buf[ROWS * H];
  __shared__ double2 nyq[ROWS];
  const double2 *__restrict__ in = pb.c[blockIdx.y];
  double *__restrict__ out = pb.r[blockIdx.y];
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

void run_y(const FFTPlan &p, PtrBatch pb, int batch, int sign, hipStream_t s) {
  const int nrows = p.hx * p.nz;
  const double2 *tw = sign < 0 ? p.tw_y_f.p : p.tw_y_b.p;
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

void fft_r2c(const FFTPlan &p, const double *in, double2 *out, hipStream_t s, const FHPre *pre) {
  FHPre fp;
  if (pre) fp = *pre;
  probe_begin(K_R2C_X, s);
#define X_CASE(HH)                                                                            \
  hipLaunchKernelGGL((k_r2c_x<HH>), dim3(x_blocks<HH>(p.ny, p.nz)), dim3(THREADS), 0, s, in, out, \
                     p.ny, p.nz, fp, p.tw_h_f.p, p.tw_x_f.p);
  PHAFD_SIZE_SWITCH(p.nx / 2, X_CASE)
#undef X_CASE
  probe_end(K_R2C_X, s, (fp.on ? 24.0 * p.N() : 8.0 * p.N()) + 16.0 * p.M());
  HIPCHK(hipGetLastError());
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
  }
  run_z(p, pb, batch, +1, s);
  run_y(p, pb, batch, +1, s);
  probe_begin(K_C2R_X, s);
#define X_CASE(HH)                                                                              \
  hipLaunchKernelGGL((k_c2r_x<HH>), dim3(x_blocks<HH>(p.ny, p.nz), batch), dim3(THREADS), 0, s, pb, \
                     p.ny, p.nz, divisor, p.tw_h_b.p, p.tw_x_b.p);
  PHAFD_SIZE_SWITCH(p.nx / 2, X_CASE)
#undef X_CASE
  probe_end(K_C2R_X, s, (16.0 * p.M() + 8.0 * p.N()) * batch);
  HIPCHK(hipGetLastError());
}

void fft_c2r(const FFTPlan &p, double2 *in, double *out, double divisor, hipStream_t s) {
  fft_c2r_batch(p, &in, &out, 1, divisor, s);
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
