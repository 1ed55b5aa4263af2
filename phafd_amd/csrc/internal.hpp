// internal.hpp -- shared declarations of the gfx950 kernels behind phafd_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace phafd {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::phafd::HipError(std::string(#expr) + ": " + hipGetErrorString(e_) +   \
                              " (" __FILE__ ":" + std::to_string(__LINE__) + ")");   \
  } while (0)

// ----------------------------------------------------------------- memory
template <class T>
struct DBuf {   // device buffer owned by the context (or a view into another's)
  T *p = nullptr;
  size_t n = 0;
  bool own = true;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n), own(o.own) { o.p = nullptr; o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; own = o.own; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p && own) (void)hipFree(p);
    p = nullptr;
    n = 0;
    own = true;
  }
  // non-owning view of `count` elements at q (alloc() within it is a no-op)
  void wrap(T *q, size_t count) {
    release();
    p = q;
    n = count;
    own = false;
  }
  // grow-only, contents not preserved.  A buffer that grows again (list and
  // term buffers sized by each rebuild's counts) takes 25% headroom, so
  // counts that creep up do not free and reallocate (hipFree synchronises
  // the device) at every rebuild.
  void alloc(size_t count) {
    if (count <= n && p) return;
    const bool regrow = p != nullptr && own;
    release();
    size_t c = count ? count : 1;
    if (regrow) c += c / 4;
    HIPCHK(hipMalloc(&p, c * sizeof(T)));
    n = c;
  }
  T *get() const { return p; }
};

// ------------------------------------------------------------ kernel probe
// Optional HIP-event timing of every launch of one kernel (by id) on the
// stream it runs on; per launch the caller states the kernel's algorithmic
// bytes (logical inputs + outputs, each read or written once), or for the
// surrogate's MFMA convolution (K_UNET_CONV) its algorithmic flops.
enum KernelId {
  K_R2C_X = 0, K_Y_FWD, K_Y_INV, K_Z_FWD, K_Z_INV, K_C2R_X, K_GRAD_MUL, K_UPDATE, K_FH,
  K_GA_DEPOSIT, K_GA_FORCE, K_GRID_COUNT, K_GRID_FILL, K_HALF_COUNT, K_HALF_FILL, K_LJ,
  K_Z_GRAD, K_Z_UPDATE, K_UNET_CONV, K_C2R_XD, K_NKERNELS
};
struct KernelProbe {
  int id = -1;
  std::vector<hipEvent_t> pool;
  std::vector<double> bytes;
  int used = 0;
  double ms = 0, total_bytes = 0;
  int64_t launches = 0;
};
extern KernelProbe *g_probe;
void probe_begin(int id, hipStream_t s);
void probe_end(int id, hipStream_t s, double bytes);
void probe_collect();   // after the stream has been synchronised

// ------------------------------------------------------------------- FFT
// K layout of half spectra on the device: [Nx/2+1][Nz][Ny] (ky fastest), so
// that the y pass reads contiguous rows and the z pass reads contiguous runs
// of ky.  Canonical host layout is [Nz][Ny][Nx/2+1].  Every transform exists
// in FP64 (double / double2) and FP32 (float / float2), as overloads.
// exp(-2 pi i m / L) into f[m], exp(+2 pi i m / L) into b[m], m < L
// (long double evaluation rounded to the element type): the FFT twiddles
template <class CV>
void twiddle_host(CV *f, CV *b, int L);

struct FFTPlan {
  int nx = 0, ny = 0, nz = 0, hx = 0;
  DBuf<double2> tw_x_f, tw_x_b;   // length nx   (post/pre processing of the real pass)
  DBuf<double2> tw_h_f, tw_h_b;   // length nx/2 (complex FFT of the packed row)
  DBuf<double2> tw_y_f, tw_y_b;   // length ny
  DBuf<double2> tw_z_f, tw_z_b;   // length nz
  DBuf<float2> tw32_x_f, tw32_x_b, tw32_h_f, tw32_h_b, tw32_y_f, tw32_y_b, tw32_z_f, tw32_z_b;
  void init(int nx, int ny, int nz, bool fp32 = false);
  // the sizes alone (validated); the tables are left to the caller (wrap)
  void set_shape(int nx, int ny, int nz);
  size_t M() const { return (size_t)hx * nz * ny; }
  size_t N() const { return (size_t)nx * ny * nz; }
};

// A slab of n systems of one grid: real arrays sr, spectra sc elements apart
// (replica r at base + r * stride); a single system is a slab of one.
struct Slab {
  int n = 1;
  size_t sr = 0, sc = 0;
};

// pre-op fused into the load of the r2c x pass
template <class R>
struct FHPreT {
  bool on = false;
  R *phi = nullptr;           // read, reflected values written back
  R *mu = nullptr;            // read, mu + FH written back
  double tv = 1, chi = 1;     // temp/volFH, chi
  unsigned long long *counters = nullptr;   // [nless, ngreat] (per replica in a slab)
  const double *chi_b = nullptr;            // per-replica chi in a slab, or nullptr
};
using FHPre = FHPreT<double>;

struct SpecConst;
struct UpdateParams;
#define PHAFD_FFT_DECL(RT, CT)                                                                         \
  void fft_r2c(const FFTPlan &p, const RT *in, CT *out, hipStream_t s, const FHPreT<RT> *pre = nullptr, \
               Slab sl = Slab());                                                                        \
  void fft_c2r_slab(const FFTPlan &p, CT *in, RT *out, Slab sl, double divisor, hipStream_t s);          \
  void fft_c2r_yx_slab(const FFTPlan &p, CT *in, RT *out, Slab sl, double divisor, hipStream_t s);       \
  void fft_c2r(const FFTPlan &p, CT *in, RT *out, double divisor, hipStream_t s);                        \
  void fft_c2r_batch(const FFTPlan &p, CT *const *in, RT *const *out, int batch, double divisor,         \
                     hipStream_t s);                                                                     \
  void fft_r2c_xy(const FFTPlan &p, const RT *in, CT *out, hipStream_t s, const FHPreT<RT> *pre = nullptr, \
                  Slab sl = Slab());                                                                     \
  void fft_grad_sep(const FFTPlan &p, const RT *phi, CT *ph, CT *gy_k, CT *gz_k, RT *gx, RT *gy, RT *gz, \
                    SpecConst sc, hipStream_t s, Slab sl = Slab());                                      \
  void fft_z_update(const FFTPlan &p, CT *mu, CT *ph, const CT *noise, UpdateParams u,                   \
                    const double *origin, hipStream_t s, Slab sl = Slab(), bool store_mu = true);        \
  void fft_c2r_yx_batch(const FFTPlan &p, CT *const *in, RT *const *out, int batch, double divisor,      \
                        hipStream_t s);                                                                  \
  void fft_z_update_next(const FFTPlan &p, CT *mu, CT *ph, CT *za, CT *zg, const CT *noise, UpdateParams u,  \
                         const double *origin, SpecConst sc, hipStream_t s, bool store_mu);                \
  void fft_c2r_next(const FFTPlan &p, CT *za, CT *gy_k, CT *zg, RT *phi, RT *gx, RT *gy, RT *gz, SpecConst sc, \
                    hipStream_t s);                                                                         \
  void spec_k_to_canon(const FFTPlan &p, const CT *k, CT *canon, hipStream_t s);                         \
  void spec_canon_to_k(const FFTPlan &p, const CT *canon, CT *k, hipStream_t s);
// fft_r2c / fft_c2r: whole transforms (c2r destroys `in`; divisor != 0
// divides the output).  fft_c2r_batch: up to 4 spectra per launch.
// fft_r2c_xy: the x and y passes of an r2c.  fft_grad_sep: the gradient of
// phi by separable spectral derivatives (grids with fft_sep_grad_ok): gx from
// one x-line kernel, gy from the y pass, gz from the z pass; ph receives
// phihat; gy_k, gz_k are scratch spectra; equal in exact arithmetic to r2c,
// grad multiply, three c2r / N.  fft_z_update: z pass of mu's r2c, the
// spectral update of phihat and the z pass of its c2r in one kernel (grids
// with fft_fused_z_ok); finish with fft_c2r_yx_batch.  origin != nullptr
// (FP32 form): the k = 0 mode is taken from the FP64 sum there.  store_mu
// false: muhat of the planes 1 <= kx < Nx/2 is left as the y pass wrote it
// (for steps whose muhat nobody reads; the update itself is unchanged).
// fft_z_update_next / fft_c2r_next (grids with fft_sep_grad_ok, one
// system): the update of a step that another step of the run follows, and
// the c2r of phi together with the next step's gradient: phihat keeps N
// times the updated spectrum (the next step's r2c of phi, up to rounding),
// za and zg hold the z passes of phi and of its z gradient, gy_k is scratch;
// phi and grad phi come out as the next step's start_of_step would make them.
PHAFD_FFT_DECL(double, double2)
PHAFD_FFT_DECL(float, float2)
#undef PHAFD_FFT_DECL
bool fft_fused_z_ok(const FFTPlan &p);
bool fft_sep_grad_ok(const FFTPlan &p);

// ------------------------------------------------------------ field kernels
struct SpecConst {
  double dqx, dqy_kz, dqz_ky;   // 2*3.141592/Lx, /Lz, /Ly (src/domain.cpp:206-208)
};
struct UpdateParams {
  double mobility, temp, volFH, gamma, dt;
  double normalization, qx_scale, qky_scale, qkz_scale;
  // semi != 0: the linear gradient term taken implicitly (opt-in scheme)
  int semi;
  // noise_dev != 0: thermal noise generated per mode on the device from a
  // counter-based generator (Philox4x64-10 keyed by (key0, key1), counter
  // (K-layout mode index, step, 0, 0)); pref_c, pref_r: the complex and real
  // prefactors sqrt(12 T M / L^3), sqrt(24 T M / L^3); sqrtdt = sqrt(dt)
  int noise_dev;
  unsigned long long key0, key1, step;
  double pref_c, pref_r, sqrtdt;
};
#define PHAFD_FIELD_DECL(RT, CT)                                                                     \
  void grad_multiply(const FFTPlan &p, const CT *phihat, CT *gx, CT *gy, CT *gz, SpecConst c,        \
                     hipStream_t s, Slab sl = Slab());                                               \
  void spectral_update(const FFTPlan &p, CT *phihat, const CT *muhat, const CT *noise_k,             \
                       UpdateParams u, const double *origin, hipStream_t s, Slab sl = Slab());       \
  void spectral_update_edges(const FFTPlan &p, CT *phihat, const CT *muhat, const CT *noise_k,       \
                             UpdateParams u, const double *origin, hipStream_t s, Slab sl = Slab()); \
  void floryhuggins(size_t n, RT *phi, RT *mu, double tv, double chi, unsigned long long *counters,  \
                    hipStream_t s, Slab sl = Slab(), const double *chi_b = nullptr);                 \
  void reduce_energy_fh(size_t n, const RT *phi, const RT *gx, const RT *gy, const RT *gz, double tv, \
                        double chi, double tvg, double *out2, hipStream_t s);                        \
  void sum_fp64(size_t n, const RT *x, double *out, hipStream_t s);
// grad_multiply: the three i*q multiplies of the gradient.  spectral_update:
// Conjugate::update as a pointwise map (origin != nullptr: the k = 0 mode is
// set from the FP64 sum there times the normalisation); _edges: only the
// planes kx = 0 and kx = Nx/2.  reduce_energy_fh: FH + square-gradient
// energies summed in FP64 into out2[0..1].  sum_fp64: out[0] += sum of x in
// FP64.
PHAFD_FIELD_DECL(double, double2)
PHAFD_FIELD_DECL(float, float2)
#undef PHAFD_FIELD_DECL
void fill_zero(void *p, size_t bytes, hipStream_t s);
// the device noise of one update (u.noise_dev set) in the canonical layout
// [Nz][Ny][Nx/2+1], as the update of that step adds it
void noise_field(int hx, int ny, int nz, UpdateParams u, double2 *out_canon, hipStream_t s);
// surrogate blocks (the plain update linearised about the mean: mu = f2 phi
// + n, f2 = temp/volFH (1/m + 1/(1-m) - 2 chi) + 2 mean(rho); sums = FP64
// {sum phi0, sum rho} on the device).  surrogate_mu: mu of the plain step at
// phi (no reflection, one atom type); etd_block: K steps of the spectrum p0
// (mu spectrum m0) with the linear part exact, n = muhat - f2 phihat held;
// block_residual_partials: per-block sums {|R|^2, |ph1 - ph0|^2} of the
// K-step residual with n interpolated linearly across the block.
void surrogate_mu(size_t n, const double *phi, const float *rho, double tv, double chi, double phit, double *mu,
                  hipStream_t s);
void etd_block(const FFTPlan &p, const double2 *p0, const double2 *m0, double2 *out, UpdateParams u, double tv,
               double chi, const double *sums, int K, hipStream_t s);
void block_residual_partials(const FFTPlan &p, const double2 *ph0, const double2 *ph1, const double2 *mu0,
                             const double2 *mu1, UpdateParams u, double tv, double chi, const double *sums, int K,
                             double *part, int nb, hipStream_t s);

// --------------------------------------------------------------- scan utils
// exclusive scan of n int32 values into int64 offsets; out has n+1 entries,
// out[n] = total
void scan_exclusive_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s);
void scan_exclusive_i32(const int *in, int *out, size_t n, hipStream_t s);   // out[n] = total
void scan_forget_stream(hipStream_t s);   // frees the scan workspace of a stream about to be destroyed
// per-stream space for n double partial sums of a reduction on stream s
// (grown after draining s; kept until scan_forget_stream)
double *reduce_partials(hipStream_t s, size_t n);

// compute grid/clusters (clusters.hip): out[p] = the smallest reference id
// i + (j*Nx + k)*Ny of p's periodic 6-connected component of points with
// phi < thr (gt: phi > thr), -1 outside; parent holds Nx*Ny*Nz ints
template <class R>
void launch_grid_clusters(int nx, int ny, int nz, const R *phi, double thr, bool gt, int *parent, double *out,
                          hipStream_t s);

// ----------------------------------------------------------- atom kernels
struct BinGeom {
  int nbinx, nbiny, nbinz, mbinx, mbiny, mbinz, mbinxlo, mbinylo, mbinzlo, mbins;
  double binsizex, binsizey, binsizez, bininvx, bininvy, bininvz;
  double boxlo[3], boxhi[3];
};

}  // namespace phafd
