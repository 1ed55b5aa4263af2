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
struct DBuf {   // device buffer owned by the context
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {   // grow-only, contents not preserved
    if (count <= n && p) return;
    release();
    size_t c = count ? count : 1;
    HIPCHK(hipMalloc(&p, c * sizeof(T)));
    n = c;
  }
  T *get() const { return p; }
};

// ------------------------------------------------------------ kernel probe
// Optional HIP-event timing of every launch of one kernel (by id) on the
// stream it runs on; per launch the caller states the kernel's algorithmic
// bytes (logical inputs + outputs, each read or written once).
enum KernelId {
  K_R2C_X = 0, K_Y_FWD, K_Y_INV, K_Z_FWD, K_Z_INV, K_C2R_X, K_GRAD_MUL, K_UPDATE, K_FH,
  K_GA_DEPOSIT, K_GA_FORCE, K_GRID_COUNT, K_GRID_FILL, K_HALF_COUNT, K_HALF_FILL, K_LJ,
  K_Z_GRAD, K_Z_UPDATE, K_NKERNELS
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
// of ky.  Canonical host layout is [Nz][Ny][Nx/2+1].
struct FFTPlan {
  int nx = 0, ny = 0, nz = 0, hx = 0;
  DBuf<double2> tw_x_f, tw_x_b;   // length nx   (post/pre processing of the real pass)
  DBuf<double2> tw_h_f, tw_h_b;   // length nx/2 (complex FFT of the packed row)
  DBuf<double2> tw_y_f, tw_y_b;   // length ny
  DBuf<double2> tw_z_f, tw_z_b;   // length nz
  void init(int nx, int ny, int nz);
  size_t M() const { return (size_t)hx * nz * ny; }
  size_t N() const { return (size_t)nx * ny * nz; }
};

// pre-op fused into the load of the r2c x pass
struct FHPre {
  bool on = false;
  double *phi = nullptr;      // read, reflected values written back
  double *mu = nullptr;       // read, mu + FH written back
  double tv = 1, chi = 1;     // temp/volFH, chi
  unsigned long long *counters = nullptr;   // [nless, ngreat]
};

void fft_r2c(const FFTPlan &p, const double *in, double2 *out, hipStream_t s,
             const FHPre *pre = nullptr);
// c2r destroys `in`; `divisor` != 0 divides the output (gradphi /= N)
void fft_c2r(const FFTPlan &p, double2 *in, double *out, double divisor, hipStream_t s);
// batched c2r over `batch` spectra (one launch per pass)
void fft_c2r_batch(const FFTPlan &p, double2 *const *in, double *const *out, int batch,
                   double divisor, hipStream_t s);

// Split transforms for the fused z pass.  fft_r2c_xy does the x and y passes
// of an r2c; fft_z_update finishes mu's transform in place, applies the
// spectral update to phihat and leaves it z-inverted (finish with
// fft_c2r_yx_batch).
struct SpecConst;
struct UpdateParams;
bool fft_fused_z_ok(const FFTPlan &p);
void fft_r2c_xy(const FFTPlan &p, const double *in, double2 *out, hipStream_t s,
                const FHPre *pre = nullptr);
// Gradient of phi by separable spectral derivatives (grids with
// fft_sep_grad_ok): gx from one x-line kernel, gy from the y pass, gz from the
// z pass; ph receives phihat (the full r2c of phi); gy_k, gz_k are scratch
// spectra.  Equal in exact arithmetic to r2c, grad multiply, three c2r / N.
bool fft_sep_grad_ok(const FFTPlan &p);
void fft_grad_sep(const FFTPlan &p, const double *phi, double2 *ph, double2 *gy_k, double2 *gz_k,
                  double *gx, double *gy, double *gz, SpecConst sc, hipStream_t s);
void fft_z_update(const FFTPlan &p, double2 *mu, double2 *ph, const double2 *noise, UpdateParams u,
                  hipStream_t s);
void fft_c2r_yx_batch(const FFTPlan &p, double2 *const *in, double *const *out, int batch,
                      double divisor, hipStream_t s);

// K layout <-> canonical layout (host-visible order) conversions on device
void spec_k_to_canon(const FFTPlan &p, const double2 *k, double2 *canon, hipStream_t s);
void spec_canon_to_k(const FFTPlan &p, const double2 *canon, double2 *k, hipStream_t s);

// ------------------------------------------------------------ field kernels
struct SpecConst {
  double dqx, dqy_kz, dqz_ky;   // 2*3.141592/Lx, /Lz, /Ly (src/domain.cpp:206-208)
};
void grad_multiply(const FFTPlan &p, const double2 *phihat, double2 *gx, double2 *gy,
                   double2 *gz, SpecConst c, hipStream_t s);
struct UpdateParams {
  double mobility, temp, volFH, gamma, dt;
  double normalization, qx_scale, qky_scale, qkz_scale;
};
void spectral_update(const FFTPlan &p, double2 *phihat, const double2 *muhat,
                     const double2 *noise_k, UpdateParams u, hipStream_t s);
void spectral_update_edges(const FFTPlan &p, double2 *phihat, const double2 *muhat,
                           const double2 *noise_k, UpdateParams u, hipStream_t s);
void floryhuggins(size_t n, double *phi, double *mu, double tv, double chi,
                  unsigned long long *counters, hipStream_t s);
void fill_zero(void *p, size_t bytes, hipStream_t s);
// nan check of phi[0] + rebuild decision reduced into flags[0..1]
void reduce_energy_fh(size_t n, const double *phi, const double *gx, const double *gy,
                      const double *gz, double tv, double chi, double tvg, double *out2,
                      hipStream_t s);

// --------------------------------------------------------------- scan utils
// exclusive scan of n int32 values into int64 offsets; out has n+1 entries,
// out[n] = total
void scan_exclusive_i32_to_i64(const int *in, int64_t *out, size_t n, hipStream_t s);
void scan_exclusive_i32(const int *in, int *out, size_t n, hipStream_t s);   // out[n] = total

// ----------------------------------------------------------- atom kernels
struct BinGeom {
  int nbinx, nbiny, nbinz, mbinx, mbiny, mbinz, mbinxlo, mbinylo, mbinzlo, mbins;
  double binsizex, binsizey, binsizez, bininvx, bininvy, bininvz;
  double boxlo[3], boxhi[3];
};

}  // namespace phafd
