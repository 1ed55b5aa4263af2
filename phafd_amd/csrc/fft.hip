// fft.hip -- hand-written 3-D real FFT for gfx950, FP64 and FP32.
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
// Lines of 16..256 points use two-level register FFTs (one LDS exchange per
// line); other sizes use LDS Stockham stages (radix 8/4/2).  The kernels and
// their dispatch live in fft_impl.inc, compiled here once per precision.
#include <cmath>

#include "internal.hpp"
#include "spectral_ops.hpp"

namespace phafd {

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

// exp(-+2 pi i m / L) for m < L, evaluated in long double and rounded
template <class CV>
void twiddle_host(CV *f, CV *b, int L) {
  for (int m = 0; m < L; m++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * m / L;
    f[m].x = (decltype(f[m].x))cosl(a);
    f[m].y = (decltype(f[m].y))-sinl(a);
    b[m].x = (decltype(b[m].x))cosl(a);
    b[m].y = (decltype(b[m].y))sinl(a);
  }
}
template void twiddle_host<double2>(double2 *, double2 *, int);

namespace {
template <class CV>
void make_table(DBuf<CV> &f, DBuf<CV> &b, int L) {
  std::vector<CV> hf(L), hb(L);
  twiddle_host(hf.data(), hb.data(), L);
  f.alloc(L);
  b.alloc(L);
  HIPCHK(hipMemcpy(f.p, hf.data(), sizeof(CV) * L, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.p, hb.data(), sizeof(CV) * L, hipMemcpyHostToDevice));
}
}  // namespace

void FFTPlan::set_shape(int nx_, int ny_, int nz_) {
  auto ok = [](int n) { return n >= 4 && n <= 1024 && (n & (n - 1)) == 0; };
  if (!ok(nx_) || !ok(ny_) || !ok(nz_))
    throw std::invalid_argument("grid dimensions must be powers of two in [4, 1024]");
  nx = nx_;
  ny = ny_;
  nz = nz_;
  hx = nx / 2 + 1;
}

void FFTPlan::init(int nx_, int ny_, int nz_, bool fp32) {
  set_shape(nx_, ny_, nz_);
  make_table(tw_x_f, tw_x_b, nx);
  make_table(tw_h_f, tw_h_b, nx / 2);
  make_table(tw_y_f, tw_y_b, ny);
  make_table(tw_z_f, tw_z_b, nz);
  if (fp32) {
    make_table(tw32_x_f, tw32_x_b, nx);
    make_table(tw32_h_f, tw32_h_b, nx / 2);
    make_table(tw32_y_f, tw32_y_b, ny);
    make_table(tw32_z_f, tw32_z_b, nz);
  }
}

namespace f64 {
using R = double;
using C = double2;
__host__ __device__ __forceinline__ C mk2(R x, R y) { return make_double2(x, y); }
#define FMA(a, b, c) fma(a, b, c)
#define LOG(a) log(a)
#define TWP(name) p.tw_##name.p
#include "fft_impl.inc"
#undef TWP
#undef LOG
#undef FMA
}  // namespace f64

namespace f32 {
using R = float;
using C = float2;
__host__ __device__ __forceinline__ C mk2(R x, R y) { return make_float2(x, y); }
#define FMA(a, b, c) fmaf(a, b, c)
#define LOG(a) logf(a)
#define TWP(name) p.tw32_##name.p
#include "fft_impl.inc"
#undef TWP
#undef LOG
#undef FMA
}  // namespace f32

// ------------------------------------------------------------ public API
#define PHAFD_FFT_API(NS, RT, CT)                                                                          \
  void fft_r2c(const FFTPlan &p, const RT *in, CT *out, hipStream_t s, const FHPreT<RT> *pre, Slab sl) {    \
    NS::fft_r2c_p(p, in, out, s, pre, sl);                                                                   \
  }                                                                                                          \
  void fft_c2r_slab(const FFTPlan &p, CT *in, RT *out, Slab sl, double divisor, hipStream_t s) {            \
    NS::fft_c2r_slab_p(p, in, out, sl, (RT)divisor, s);                                                      \
  }                                                                                                          \
  void fft_c2r_yx_slab(const FFTPlan &p, CT *in, RT *out, Slab sl, double divisor, hipStream_t s) {         \
    NS::fft_c2r_yx_slab_p(p, in, out, sl, (RT)divisor, s);                                                   \
  }                                                                                                          \
  void fft_c2r(const FFTPlan &p, CT *in, RT *out, double divisor, hipStream_t s) {                          \
    NS::fft_c2r_p(p, in, out, (RT)divisor, s);                                                                 \
  }                                                                                                          \
  void fft_c2r_batch(const FFTPlan &p, CT *const *in, RT *const *out, int batch, double divisor,            \
                     hipStream_t s) {                                                                        \
    NS::fft_c2r_batch_p(p, in, out, batch, (RT)divisor, s);                                                    \
  }                                                                                                          \
  void fft_r2c_xy(const FFTPlan &p, const RT *in, CT *out, hipStream_t s, const FHPreT<RT> *pre, Slab sl) { \
    NS::fft_r2c_xy_p(p, in, out, s, pre, sl);                                                                \
  }                                                                                                          \
  void fft_grad_sep(const FFTPlan &p, const RT *phi, CT *ph, CT *gy_k, CT *gz_k, RT *gx, RT *gy, RT *gz,   \
                    SpecConst sc, hipStream_t s, Slab sl) {                                                  \
    NS::fft_grad_sep_p(p, phi, ph, gy_k, gz_k, gx, gy, gz, sc, s, sl);                                       \
  }                                                                                                          \
  void fft_z_update(const FFTPlan &p, CT *mu, CT *ph, const CT *noise, UpdateParams u, const double *origin, \
                    hipStream_t s, Slab sl, bool store_mu) {                                                 \
    NS::fft_z_update_p(p, mu, ph, noise, u, origin, s, sl, store_mu);                                        \
  }                                                                                                          \
  void fft_c2r_yx_batch(const FFTPlan &p, CT *const *in, RT *const *out, int batch, double divisor,         \
                        hipStream_t s) {                                                                     \
    NS::fft_c2r_yx_batch_p(p, in, out, batch, (RT)divisor, s);                                                 \
  }                                                                                                          \
  void fft_z_update_next(const FFTPlan &p, CT *mu, CT *ph, CT *za, CT *zg, const CT *noise, UpdateParams u,  \
                         const double *origin, SpecConst sc, hipStream_t s, bool store_mu) {                \
    NS::fft_z_update_next_p(p, mu, ph, za, zg, noise, u, origin, sc, s, store_mu);                           \
  }                                                                                                          \
  void fft_c2r_next(const FFTPlan &p, CT *za, CT *gy_k, CT *zg, RT *phi, RT *gx, RT *gy, RT *gz, SpecConst sc, \
                    hipStream_t s) {                                                                         \
    NS::fft_c2r_next_p(p, za, gy_k, zg, phi, gx, gy, gz, sc, s);                                             \
  }                                                                                                          \
  void spec_k_to_canon(const FFTPlan &p, const CT *k, CT *canon, hipStream_t s) {                           \
    NS::spec_k_to_canon_p(p, k, canon, s);                                                                     \
  }                                                                                                          \
  void spec_canon_to_k(const FFTPlan &p, const CT *canon, CT *k, hipStream_t s) {                           \
    NS::spec_canon_to_k_p(p, canon, k, s);                                                                     \
  }
PHAFD_FFT_API(f64, double, double2)
PHAFD_FFT_API(f32, float, float2)
#undef PHAFD_FFT_API

bool fft_fused_z_ok(const FFTPlan &p) { return f64::fft_fused_z_ok_p(p); }
bool fft_sep_grad_ok(const FFTPlan &p) { return f64::fft_sep_grad_ok_p(p); }

}  // namespace phafd
