/*
 * oracle.cpp -- CPU FP64 restatement of PHAFD's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Single-threaded, compiled with
 * -O3 -ffp-contract=off so that the order of floating-point operations is the
 * one written here, which follows the reference line by line where it
 * matters for bit-exact neighbour lists.  Every function cites the reference
 * file:line it restates (paths relative to the reference source tree).
 *
 * Parity pins (what checks each part; the tests live under tests/):
 *   FFT r2c / c2r ........ numpy rfftn / irfftn golden vectors incl. non-Hermitian
 *                          c2r inputs, plus brute-force DFT at 4^3 / 8x4x2
 *   gradphi .............. closed form for single sine modes (3.141592 wavenumber)
 *   spectral update ...... closed-form single-mode recurrence, exact mass
 *                          conservation, Hermitian symmetry of the k=0 planes
 *   Flory-Huggins ........ closed form per point incl. both reflection branches
 *   noise draw order ..... C++ standard KAT of std::mt19937 (10000th draw) and
 *                          a per-mode draw-count/ordering test
 *   bins / stencils ...... brute-force bin membership, hand-enumerated stencils
 *   grid & half lists .... O(N*M) brute-force pair sets on small boxes
 *   ghosts ............... hand-enumerated images near faces (z quirk incl.)
 *   gridatom pairs ....... single atom on a tiny grid (closed-form deposit,
 *                          zero force for constant phi)
 *   lj/cut ............... two-body analytic force, Newton's third law
 *   chain (build-defined). finite-difference gradient of its own energy
 *
 * Parity unpinned (nothing in the reference tree or in mathematics fixes it):
 *   - the semiflexible chain dynamics: the reference delegates them to the
 *     external beadrodpmer library (no source, no version, no test in the
 *     tree); the chain fix here is a build-defined harmonic bond + bending
 *     model, pinned only by its own energy's finite-difference gradient;
 *   - FFTW's c2r output on non-Hermitian input at the kx = 0 and kx = Nx/2
 *     bins: the standard half-complex synthesis (only Re() of those bins
 *     enters) is assumed, as numpy's irfftn does;
 *   - the fftw_arr container layout (padding, reverseFlat): inferred from its
 *     call sites, src/npair_grid_bin.cpp:71-116 and
 *     src/pair_gridatom_gaussian.cpp:71-77.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <omp.h>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

typedef std::complex<double> cplx;

namespace {

const double PI_REF = 3.141592;   // src/domain.cpp:206-208 truncates pi
const int IMGMASK = 1023, IMGMAX = 512, IMGBITS = 10, IMG2BITS = 20; // src/domain.cpp:10-13
const int OWNED = 0, LOCAL = 1, GHOST = 2;                           // src/atom.hpp:18-20
const double BIG = 1e20;                                             // src/comm_brick.cpp:8
const double SMALL = 1.0e-6;                                         // src/nbin_standard.cpp:25
const int ONEATOM = 2000;                                            // src/neighbor.cpp:25

bool is_pow2(int n) { return n >= 1 && (n & (n - 1)) == 0; }

/* ------------------------------------------------------------------ FFT --
 * Iterative radix-2 complex FFT (Cooley-Tukey, bit-reversed input order).
 * Twiddles are computed in long double.  Unnormalised, sign -1 forward.   */
struct FFT1 {
  int n = 0;
  std::vector<int> rev;
  std::vector<cplx> wf, wb;
  explicit FFT1(int n_) : n(n_) {
    if (!is_pow2(n)) throw std::runtime_error("oracle FFT needs power-of-two sizes");
    rev.resize(n);
    int lg = 0;
    while ((1 << lg) < n) lg++;
    for (int i = 0; i < n; i++) {
      int r = 0;
      for (int b = 0; b < lg; b++) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
      rev[i] = r;
    }
    wf.resize(n / 2 + 1);
    wb.resize(n / 2 + 1);
    for (int k = 0; k <= n / 2; k++) {
      long double a = 2.0L * 3.14159265358979323846264338327950288L * k / n;
      wf[k] = cplx((double)cosl(a), (double)-sinl(a));
      wb[k] = cplx((double)cosl(a), (double)sinl(a));
    }
  }
  void run(cplx *a, int sign) const {
    for (int i = 0; i < n; i++) if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    const std::vector<cplx> &w = sign < 0 ? wf : wb;
    for (int len = 2; len <= n; len <<= 1) {
      int step = n / len;
      for (int i = 0; i < n; i += len)
        for (int j = 0; j < len / 2; j++) {
          cplx u = a[i + j];
          cplx t = a[i + j + len / 2] * w[j * step];
          a[i + j] = u + t;
          a[i + j + len / 2] = u - t;
        }
    }
  }
};

/* 3-D r2c of real[Nz][Ny][Nx] into spec[Nz][Ny][Nx/2+1] (unnormalised).
 * Restates fftw_mpi_plan_dft_r2c_3d (src/grid.cpp:239-242) up to layout.   */
void fft3_r2c(int nx, int ny, int nz, const double *in, cplx *out) {
  FFT1 fx(nx), fy(ny), fz(nz);
  const int hx = nx / 2 + 1;
  const int L = std::max(nx, std::max(ny, nz));
#pragma omp parallel
  {
    std::vector<cplx> line(L);
#pragma omp for schedule(static)
    for (int zy = 0; zy < nz * ny; zy++) {
      const double *r = in + (size_t)zy * nx;
      for (int x = 0; x < nx; x++) line[x] = cplx(r[x], 0.0);
      fx.run(line.data(), -1);
      cplx *o = out + (size_t)zy * hx;
      for (int k = 0; k < hx; k++) o[k] = line[k];
    }
#pragma omp for schedule(static)
    for (int zk = 0; zk < nz * hx; zk++) {
      const int z = zk / hx, k = zk % hx;
      for (int y = 0; y < ny; y++) line[y] = out[((size_t)z * ny + y) * hx + k];
      fy.run(line.data(), -1);
      for (int y = 0; y < ny; y++) out[((size_t)z * ny + y) * hx + k] = line[y];
    }
#pragma omp for schedule(static)
    for (int yk = 0; yk < ny * hx; yk++) {
      const int y = yk / hx, k = yk % hx;
      for (int z = 0; z < nz; z++) line[z] = out[((size_t)z * ny + y) * hx + k];
      fz.run(line.data(), -1);
      for (int z = 0; z < nz; z++) out[((size_t)z * ny + y) * hx + k] = line[z];
    }
  }
}

/* 3-D c2r (unnormalised) with the half-complex synthesis rule: inverse c2c
 * along z and y over every kx, then a 1-D c2r along x that uses only the real
 * part of the kx=0 and kx=Nx/2 bins.  Restates fftw_mpi_plan_dft_c2r_3d
 * (src/grid.cpp:244-280) for Hermitian input; for the non-Hermitian grad
 * spectra of src/fixgrid_conjugate.cpp:80-105 this is the stated convention
 * (numpy irfftn follows the same rule and pins it). The input is preserved. */
void fft3_c2r(int nx, int ny, int nz, const cplx *in, double *out) {
  FFT1 fx(nx), fy(ny), fz(nz);
  const int hx = nx / 2 + 1;
  std::vector<cplx> w(in, in + (size_t)nz * ny * hx);
  const int L = std::max(nx, std::max(ny, nz));
#pragma omp parallel
  {
    std::vector<cplx> line(L);
#pragma omp for schedule(static)
    for (int yk = 0; yk < ny * hx; yk++) {
      const int y = yk / hx, k = yk % hx;
      for (int z = 0; z < nz; z++) line[z] = w[((size_t)z * ny + y) * hx + k];
      fz.run(line.data(), +1);
      for (int z = 0; z < nz; z++) w[((size_t)z * ny + y) * hx + k] = line[z];
    }
#pragma omp for schedule(static)
    for (int zk = 0; zk < nz * hx; zk++) {
      const int z = zk / hx, k = zk % hx;
      for (int y = 0; y < ny; y++) line[y] = w[((size_t)z * ny + y) * hx + k];
      fy.run(line.data(), +1);
      for (int y = 0; y < ny; y++) w[((size_t)z * ny + y) * hx + k] = line[y];
    }
#pragma omp for schedule(static)
    for (int zy = 0; zy < nz * ny; zy++) {
      const cplx *c = w.data() + (size_t)zy * hx;
      line[0] = cplx(c[0].real(), 0.0);
      for (int k = 1; k < nx / 2; k++) {
        line[k] = c[k];
        line[nx - k] = std::conj(c[k]);
      }
      if (nx > 1) line[nx / 2] = cplx(c[nx / 2].real(), 0.0);
      fx.run(line.data(), +1);
      double *r = out + (size_t)zy * nx;
      for (int x = 0; x < nx; x++) r[x] = line[x].real();
    }
  }
}

/* signed wavenumber index: i > N/2 -> i-N (Nyquist stays positive),
 * src/fixgrid_conjugate.cpp:82-93, src/conjugate.cpp:146-157               */
inline int signed_k(int i, int n) { return i > n / 2 ? i - n : i; }

/* grad multiply + three c2r + divide, src/fixgrid_conjugate.cpp:62-116.
 * Canonical spectrum index (kz,ky,kx); the reference's transposed (i,j,k)
 * has i = ky, j = kz.  dqx = 2pi/Lx, "dqy" = 2pi/Lz multiplies the kz index,
 * "dqz" = 2pi/Ly multiplies the ky index (src/domain.cpp:206-208); the c2r of
 * ft_gradphi[1] lands in gradphi[2] and vice versa (src/grid.cpp:269-280),
 * so real-space gx, gy, gz are d/dx, d/dy, d/dz.                           */
void gradphi_impl(int nx, int ny, int nz, const double period[3], const double *phi,
                  double *gx, double *gy, double *gz, std::vector<cplx> *phihat_out) {
  const int hx = nx / 2 + 1;
  const size_t M = (size_t)nz * ny * hx;
  std::vector<cplx> ph(M), g0(M), g1(M), g2(M);
  fft3_r2c(nx, ny, nz, phi, ph.data());
  const cplx idqx(0, 2 * PI_REF / period[0]);
  const cplx idqy(0, 2 * PI_REF / period[2]);
  const cplx idqz(0, 2 * PI_REF / period[1]);
#pragma omp parallel for schedule(static)
  for (int ky = 0; ky < ny; ky++) {
    double l = (double)signed_k(ky, ny);
    for (int kz = 0; kz < nz; kz++) {
      double m = (double)signed_k(kz, nz);
      for (int k = 0; k < hx; k++) {
        double n = k;
        size_t idx = ((size_t)kz * ny + ky) * hx + k;
        g0[idx] = ph[idx] * idqx * n;
        g1[idx] = ph[idx] * idqy * m;
        g2[idx] = ph[idx] * idqz * l;
      }
    }
  }
  fft3_c2r(nx, ny, nz, g0.data(), gx);
  fft3_c2r(nx, ny, nz, g1.data(), gz);   // swapped target, src/grid.cpp:270-273
  fft3_c2r(nx, ny, nz, g2.data(), gy);   // swapped target, src/grid.cpp:276-279
  double factor = (double)nx * ny * nz;  // src/fixgrid_conjugate.cpp:111
  const size_t N = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; i++) { gx[i] /= factor; gy[i] /= factor; gz[i] /= factor; }
  if (phihat_out) phihat_out->swap(ph);
}

/* FixGridFloryHuggins::post_force, src/fixgrid_floryhuggins.cpp:65-85 */
void floryhuggins_impl(int64_t n, double *phi, double *mu, double temp, double volFH,
                       double chi, int64_t &nless, int64_t &ngreat) {
  int64_t nl = 0, ng = 0;
#pragma omp parallel for schedule(static) reduction(+ : nl, ng)
  for (int64_t i = 0; i < n; i++) {
    if (phi[i] <= 0) {
      phi[i] = -phi[i];
      nl += 1;
    } else if (phi[i] >= 1) {
      phi[i] = 1 - (phi[i] - 1);
      ng += 1;
    }
    mu[i] += temp / volFH * (log(phi[i] / (1 - phi[i])) + chi * (1 - 2 * phi[i]));
  }
  nless += nl;
  ngreat += ng;
}

/* Conjugate::update single-rank traversal, src/conjugate.cpp:163-327, with
 * ConjugateVolFrac's per-mode formula, src/conjugate_volfrac.cpp:72-144.
 * Noise modes: off (new flag: the reference has no switch, finding F2),
 * mt19937 (the reference's draws, in its traversal order), or a per-mode
 * host field in the canonical spectrum layout.                             */
struct SpectralParams {
  int nx, ny, nz;
  double period[3];
  double mobility, temp, volFH, gamma, dt;
  int noise_mode;
  int semi;              // the opt-in semi-implicit form (src/conjugate.cpp:356, a comment there)
  const double *noise;   // canonical layout (re,im) when noise_mode == FIELD
  std::mt19937 *gen;     // when noise_mode == MT19937
  std::vector<double> *record;   // optional: record draws per mode (noise field generation)
};

struct SpectralUpdater {
  SpectralParams p;
  int hx, A, B;          // A = Ny (ky count), B = Nz (kz count)
  std::vector<double> qzs, qys;   // qzs over ky, qys over kz (src/conjugate.cpp:130-161)
  double normalization, complexprefactor, realprefactor, sqrtdt;
  std::uniform_real_distribution<double> real_dist{-0.5, 0.5};
  cplx *phi;
  const cplx *mu;

  explicit SpectralUpdater(const SpectralParams &pp) : p(pp) {
    hx = p.nx / 2 + 1;
    A = p.ny;
    B = p.nz;
    const double dz = 2 * PI_REF / p.period[1];  // Domain::dqz()
    const double dy = 2 * PI_REF / p.period[2];  // Domain::dqy()
    qzs.resize(A);
    qys.resize(B);
    for (int i = 0; i < A; i++)
      qzs[i] = (i > A / 2) ? -dz * (A - i) : dz * i;
    for (int j = 0; j < B; j++)
      qys[j] = (j > B / 2) ? -dy * (B - j) : dy * j;
    // ConjugateVolFrac::reset_dt, src/conjugate_volfrac.cpp:72-86
    double invLcubed = 1.0 / (p.period[0] * p.period[1] * p.period[2]);
    normalization = 1.0 / ((double)p.nx * p.nz * p.ny);
    complexprefactor = sqrt(12 * p.temp * p.mobility * invLcubed);
    realprefactor = sqrt(24 * p.temp * p.mobility * invLcubed);
    sqrtdt = sqrt(p.dt);
  }
  inline size_t idx(int i, int j, int k) const { return ((size_t)j * A + i) * hx + k; }
  inline double q2_of(int i, int j, int k, double &qx) const {
    double qz = qzs[i], qy = qys[j];
    qx = (2 * PI_REF / p.period[0]) * k;
    return qx * qx + qy * qy + qz * qz;
  }
  inline double draw() {
    double u = real_dist(*p.gen);
    return u;
  }
  // the bracket of src/conjugate_volfrac.cpp:106-109 (explicit Euler), or
  // with p.semi the linear gradient term implicit: (phi - c mu) / (1 + c s),
  // c = M q^2 dt, s = T/v gamma q^2, the denominator of the formula left in
  // a comment at src/conjugate.cpp:356
  cplx step_mode(cplx ph, cplx m, double q2) const {
    if (p.semi) {
      const double c = p.mobility * q2 * p.dt;
      const double sg = p.temp / p.volFH * p.gamma * q2;
      return (ph - c * m) / (1.0 + c * sg);
    }
    return ph - p.mobility * q2 * p.dt * (m + p.temp / p.volFH * p.gamma * q2 * ph);
  }
  // src/conjugate_volfrac.cpp:89-113
  void complex_update(int i, int j, int k) {
    double qx, q2 = q2_of(i, j, k, qx);
    cplx noise(0.0, 0.0);
    size_t id = idx(i, j, k);
    if (p.noise_mode == ORC_NOISE_MT19937) {
      double u1 = draw(), u2 = draw();
      noise.real(complexprefactor * sqrt(q2) * sqrtdt * u1);
      noise.imag(complexprefactor * sqrt(q2) * sqrtdt * u2);
      if (p.record) { (*p.record)[2 * id] = noise.real(); (*p.record)[2 * id + 1] = noise.imag(); }
    } else if (p.noise_mode == ORC_NOISE_FIELD) {
      noise = cplx(p.noise[2 * id], p.noise[2 * id + 1]);
    }
    phi[id] = step_mode(phi[id], mu[id], q2) * normalization + noise;
  }
  // src/conjugate_volfrac.cpp:116-133
  void real_update(int i, int j, int k) {
    double qx, q2 = q2_of(i, j, k, qx);
    cplx noise(0.0, 0.0);
    size_t id = idx(i, j, k);
    if (p.noise_mode == ORC_NOISE_MT19937) {
      noise = realprefactor * sqrt(q2) * sqrtdt * draw();
      if (p.record) { (*p.record)[2 * id] = noise.real(); (*p.record)[2 * id + 1] = 0.0; }
    } else if (p.noise_mode == ORC_NOISE_FIELD) {
      noise = cplx(p.noise[2 * id], p.noise[2 * id + 1]);
    }
    phi[id] = step_mode(phi[id], mu[id], q2) * normalization + noise;
  }
  // src/conjugate_volfrac.cpp:135-141
  void origin_update() { phi[idx(0, 0, 0)] *= normalization; }

  // src/conjugate.cpp:255-302
  void single(int nx) {
    int nz = 0;
    for (int j = 1; j < B / 2; j++) {
      complex_update(nz, j, nx);
      phi[idx(nz, B - j, nx)] = std::conj(phi[idx(nz, j, nx)]);
    }
    for (int i = 1; i < A / 2; i++)
      for (int j = 1; j < B / 2; j++) {
        complex_update(i, j, nx);
        phi[idx(A - i, B - j, nx)] = std::conj(phi[idx(i, j, nx)]);
      }
    for (int i = A / 2 + 1; i < A; i++)
      for (int j = 1; j < B / 2; j++) {
        complex_update(i, j, nx);
        phi[idx(A - i, B - j, nx)] = std::conj(phi[idx(i, j, nx)]);
      }
    nz = A / 2;
    for (int j = 1; j < B / 2; j++) {
      complex_update(nz, j, nx);
      phi[idx(nz, B - j, nx)] = std::conj(phi[idx(nz, j, nx)]);
    }
  }
  // src/conjugate.cpp:304-327
  void line_single(int nx) {
    int ny = 0;
    for (int i = 1; i < A / 2; i++) {
      complex_update(i, ny, nx);
      phi[idx(A - i, ny, nx)] = std::conj(phi[idx(i, ny, nx)]);
    }
    ny = B / 2;
    for (int i = A / 2 + 1; i < A; i++) {
      complex_update(i, ny, nx);
      phi[idx(A - i, ny, nx)] = std::conj(phi[idx(i, ny, nx)]);
    }
  }
  // src/conjugate.cpp:163-247 (nprocs == 1 branch)
  void update(cplx *ph, const cplx *m) {
    phi = ph;
    mu = m;
    const int localNx = hx;
    for (int nz = 0; nz < A; nz++)
      for (int ny = 0; ny < B; ny++)
        for (int nx = 1; nx < localNx - 1; nx++) complex_update(nz, ny, nx);
    for (int nx = 0; nx < localNx; nx += localNx - 1) {
      single(nx);
      line_single(nx);
      if (localNx - 1 == 0) break;
    }
    int nx = localNx - 1;
    real_update(0, B / 2, 0);
    real_update(0, B / 2, nx);
    real_update(0, 0, nx);
    origin_update();
    int nz = A / 2;
    real_update(nz, 0, 0);
    real_update(nz, B / 2, 0);
    real_update(nz, B / 2, nx);
    real_update(nz, 0, nx);
  }
};

/* utility::make_unique_seed for rank 0, src/utility.cpp:213-235 */
int make_unique_seed(int baseseed) {
  std::mt19937 gen;
  std::uniform_int_distribution<int> integer_dist;
  gen.seed(baseseed);
  return integer_dist(gen);
}

/* ============================================================ system === */
struct PairStyle {
  int style;           // ORC_PAIR_*
  double maxcut;
  // gridatom
  std::vector<double> eps, phit, cutsq;
  // lj/cut, src/pair_lj_cut.cpp:101-117
  double lj1 = 0, lj2 = 0, ljcutsq = 0;
  // harmonic/cut, src/pair_harmonic_cut.cpp:111-124
  double hk = 0, hcut = 0;
};

struct FixDef {
  enum { CONJ, FH, GRADPHI, DRAG, CHAIN, SEMIFLEX } kind;
  // conjugate
  int seed = 0, noise_mode = 0;
  double mobility = 1, temp = 1, volFH = 1, gamma = 1, chi = 1;
  std::mt19937 gen;
  // gradphi
  int once = 0;
  // drag / chain
  std::vector<int> start, end;
  double drag = 0, r0 = 1, kbond = 0, kbend = 0;
  // atom/semiflexible/*: one polymer per group chunk
  int tether = 0;
  double bondlength = 1, zpar = 1, zperp = 1, bending = 0, brtemp = 0;
  std::vector<int> pseed;            // per-polymer seeds
  std::vector<double> R0, FB, T;     // per atom: start-of-step positions, Brownian force, rod tensions
  int64_t brstep = 0;                // first_step calls
};

struct Swap {
  int dim;             // 0,1,2
  double lo, hi;
  double shift;        // applied to coordinate `dim` (x/y/z swaps)
  std::vector<int> sendlist;
  std::vector<double> pbcz;   // per-entry z shift for the copy swap
  std::vector<int> labelz;    // per-entry labels for the copy swap
  int firstrecv = 0;
};

}  // namespace

struct orc_sys {
  std::string err;
  int nx, ny, nz, hx;
  double boxlo[3], boxhi[3], period[3], sublo[3], subhi[3];
  int zghost_mode;
  size_t N, M;
  std::vector<double> phi, mu, grad[3];
  std::vector<cplx> ftphi, ftmu;
  std::vector<double> noisefield;   // consumed by one update
  int semi = 0;
  // atoms (src/atom.hpp)
  int nowned = 0, ngathered = 0, nlocal = 0, nghost = 0, ntypes = 0;
  std::vector<double> xs, uxs, Fs;
  std::vector<int> tags, types, images, labels, moltags, owner;
  // comm
  double cutghost = 0;
  std::vector<Swap> swaps;
  // neighbor
  bool neigh_set = false;
  double skin = 0, cutneigh = 0, triggersq = 0;
  int ago = 0, ncalls = 0;
  std::vector<double> xholds;
  int nbinx, nbiny, nbinz, mbinx, mbiny, mbinz, mbinxlo, mbinylo, mbinzlo, mbins;
  double binsizex, binsizey, binsizez, bininvx, bininvy, bininvz;
  std::vector<int> binhead, bins, atom2bin;
  std::set<int> local_atom_indices;
  std::vector<int> stencil_full, stencil_half;
  bool need_full = false, need_half = false;
  // lists (CSR)
  std::vector<int> g_ilist, g_num, g_off, g_neigh;
  std::vector<int> h_ilist, h_num, h_off, h_neigh;
  bool list_overflow = false;
  // styles
  std::vector<PairStyle> pairs;
  std::vector<FixDef> fixes;
  double dt = 0;
  int64_t timestep = 0, firststep = 0;
  int64_t nless = 0, ngreat = 0;
  bool setup_done = false;
  bool nan_seen = false;

  double dx() const { return period[0] / nx; }   // src/grid.cpp:187-201
  double dy() const { return period[1] / ny; }
  double dz() const { return period[2] / nz; }
};

namespace {

/* ----------------------------------------------------------- domain --- */
// Domain::unmap / map, src/domain.cpp:166-193
void unmap(const orc_sys *s, const double *x, int image, double *u) {
  int xbox = (image & IMGMASK) - IMGMAX;
  int ybox = (image >> IMGBITS & IMGMASK) - IMGMAX;
  int zbox = (image >> IMG2BITS) - IMGMAX;
  u[0] = x[0] + xbox * s->period[0];
  u[1] = x[1] + ybox * s->period[1];
  u[2] = x[2] + zbox * s->period[2];
}
void map(const orc_sys *s, const double *u, int image, double *x) {
  int xbox = (image & IMGMASK) - IMGMAX;
  int ybox = (image >> IMGBITS & IMGMASK) - IMGMAX;
  int zbox = (image >> IMG2BITS) - IMGMAX;
  x[0] = u[0] - xbox * s->period[0];
  x[1] = u[1] - ybox * s->period[1];
  x[2] = u[2] - zbox * s->period[2];
}

// Domain::pbc, src/domain.cpp:102-163
void pbc(orc_sys *s) {
  for (int i = 0; i < s->nowned; i++) {
    double *x = &s->xs[3 * i];
    int &img = s->images[i];
    int idim, otherdims;
    if (x[0] < s->boxlo[0]) {
      x[0] += s->period[0];
      idim = img & IMGMASK; otherdims = img ^ idim; idim--; idim &= IMGMASK; img = otherdims | idim;
    }
    if (x[0] >= s->boxhi[0]) {
      x[0] -= s->period[0];
      x[0] = (x[0] > s->boxlo[0] ? x[0] : s->boxlo[0]);
      idim = img & IMGMASK; otherdims = img ^ idim; idim++; idim &= IMGMASK; img = otherdims | idim;
    }
    if (x[1] < s->boxlo[1]) {
      x[1] += s->period[1];
      idim = (img >> IMGBITS) & IMGMASK; otherdims = img ^ (idim << IMGBITS); idim--; idim &= IMGMASK;
      img = otherdims | (idim << IMGBITS);
    }
    if (x[1] >= s->boxhi[1]) {
      x[1] -= s->period[1];
      x[1] = (x[1] > s->boxlo[1] ? x[1] : s->boxlo[1]);
      idim = (img >> IMGBITS) & IMGMASK; otherdims = img ^ (idim << IMGBITS); idim++; idim &= IMGMASK;
      img = otherdims | (idim << IMGBITS);
    }
    if (x[2] < s->boxlo[2]) {
      x[2] += s->period[2];
      idim = img >> IMG2BITS; otherdims = img ^ (idim << IMG2BITS); idim--; idim &= IMGMASK;
      img = otherdims | (idim << IMG2BITS);
    }
    if (x[2] >= s->boxhi[2]) {
      x[2] -= s->period[2];
      x[2] = (x[2] > s->boxlo[2] ? x[2] : s->boxlo[2]);
      idim = img >> IMG2BITS; otherdims = img ^ (idim << IMG2BITS); idim++; idim &= IMGMASK;
      img = otherdims | (idim << IMG2BITS);
    }
    unmap(s, x, img, &s->uxs[3 * i]);
  }
}

/* ------------------------------------------------------------- comm --- */
// CommBrick::setup, src/comm_brick.cpp:96-187 (nprocs == 1).  Swap list:
// [0] copy swap (z), then (periodic mode only) z lo / z hi image swaps, then
// y lo, y hi, x lo, x hi.  In reference mode the z image swaps are absent,
// which is what one rank gives (finding F3).
void comm_setup(orc_sys *s) {
  s->cutghost = s->cutneigh;
  s->swaps.clear();
  Swap z; z.dim = 2; z.shift = 0;
  int zstart = (int)s->sublo[2];   // int truncation, src/comm_brick.cpp:110-111
  int zend = (int)s->subhi[2];
  z.lo = zend - s->cutghost;       // slablo, src/comm_brick.cpp:138-139
  z.hi = zstart + s->cutghost;     // slabhi, src/comm_brick.cpp:142-143
  s->swaps.push_back(z);
  std::vector<int> dims;
  if (s->zghost_mode == ORC_ZGHOST_PERIODIC) dims.push_back(2);
  dims.push_back(1);
  dims.push_back(0);
  for (int dim : dims) {
    for (int ineed = 0; ineed < 2; ineed++) {
      Swap w; w.dim = dim;
      if (ineed % 2 == 0) {
        w.lo = -BIG; w.hi = s->boxlo[dim] + s->cutghost; w.shift = s->period[dim];
      } else {
        w.lo = s->boxhi[dim] - s->cutghost; w.hi = BIG; w.shift = -s->period[dim];
      }
      s->swaps.push_back(w);
    }
  }
}

void grow_gathered(orc_sys *s, int last) {
  s->xs.resize(3 * (size_t)last);
  s->Fs.resize(3 * (size_t)last);
  s->tags.resize(last);
  s->types.resize(last);
  s->labels.resize(last);
  s->owner.resize(last);
}

// CommBrick::borders, src/comm_brick.cpp:263-426 (sendproc == me branches)
void borders(orc_sys *s) {
  s->nlocal = s->ngathered = s->nghost = 0;
  // copy swap: src/comm_brick.cpp:299-316
  {
    Swap &w = s->swaps[0];
    w.sendlist.clear(); w.pbcz.clear(); w.labelz.clear();
    int zstart = (int)s->sublo[2], zend = (int)s->subhi[2];
    const double zprd = s->period[2];
    int nlocalsend = 0;
    for (int i = 0; i < s->nowned; i++) {
      double zc = s->xs[3 * i + 2];
      if (zc >= zstart && zc <= zend) {
        w.sendlist.push_back(i); w.pbcz.push_back(0); w.labelz.push_back(LOCAL); nlocalsend++;
      } else if (zc >= w.lo) {
        w.sendlist.push_back(i); w.pbcz.push_back(-zprd); w.labelz.push_back(GHOST);
      } else if (zc <= w.hi) {
        w.sendlist.push_back(i); w.pbcz.push_back(zprd); w.labelz.push_back(GHOST);
      }
    }
    int first = s->nowned + s->ngathered;
    int n = (int)w.sendlist.size();
    grow_gathered(s, first + n);
    for (int q = 0; q < n; q++) {   // Atom::pack_border_in_z + unpack_border
      int j = w.sendlist[q];
      s->xs[3 * (first + q) + 0] = s->xs[3 * j + 0];
      s->xs[3 * (first + q) + 1] = s->xs[3 * j + 1];
      s->xs[3 * (first + q) + 2] = s->xs[3 * j + 2] + w.pbcz[q];
      s->tags[first + q] = s->tags[j];
      s->types[first + q] = s->types[j];
      s->labels[first + q] = w.labelz[q];
      s->owner[first + q] = j;
    }
    w.firstrecv = first;
    s->ngathered += n;
    s->nlocal += nlocalsend;
  }
  // image swaps, src/comm_brick.cpp:337-474 (x/y branch, plus z in periodic mode)
  size_t iswap = 1;
  while (iswap < s->swaps.size()) {
    int dim = s->swaps[iswap].dim;
    int nfirst = s->nowned, nlast = s->nowned + s->ngathered;
    for (int ineed = 0; ineed < 2; ineed++, iswap++) {
      Swap &w = s->swaps[iswap];
      w.sendlist.clear();
      for (int i = nfirst; i < nlast; i++) {
        double c = s->xs[3 * i + dim];
        if (c >= w.lo && c <= w.hi) w.sendlist.push_back(i);
      }
      int first = s->nowned + s->ngathered;
      int n = (int)w.sendlist.size();
      grow_gathered(s, first + n);
      double pbcv[3] = {0, 0, 0};
      pbcv[dim] = w.shift;
      for (int q = 0; q < n; q++) {   // Atom::pack_border
        int j = w.sendlist[q];
        s->xs[3 * (first + q) + 0] = s->xs[3 * j + 0] + pbcv[0];
        s->xs[3 * (first + q) + 1] = s->xs[3 * j + 1] + pbcv[1];
        s->xs[3 * (first + q) + 2] = s->xs[3 * j + 2] + pbcv[2];
        s->tags[first + q] = s->tags[j];
        s->types[first + q] = s->types[j];
        s->labels[first + q] = GHOST;
        s->owner[first + q] = s->owner[j];
      }
      w.firstrecv = first;
      s->ngathered += n;
    }
  }
  s->nghost = s->ngathered - s->nlocal;
}

// CommBrick::forward_comm, src/comm_brick.cpp:190-224 (self swaps)
void forward_comm(orc_sys *s) {
  for (size_t iswap = 0; iswap < s->swaps.size(); iswap++) {
    Swap &w = s->swaps[iswap];
    int n = (int)w.sendlist.size();
    for (int q = 0; q < n; q++) {
      int j = w.sendlist[q];
      double *o = &s->xs[3 * (size_t)(w.firstrecv + q)];
      if (iswap == 0) {   // Atom::pack_comm_in_z
        o[0] = s->xs[3 * j + 0];
        o[1] = s->xs[3 * j + 1];
        o[2] = s->xs[3 * j + 2] + w.pbcz[q];
      } else {            // Atom::pack_comm
        double pbcv[3] = {0, 0, 0};
        pbcv[w.dim] = w.shift;
        o[0] = s->xs[3 * j + 0] + pbcv[0];
        o[1] = s->xs[3 * j + 1] + pbcv[1];
        o[2] = s->xs[3 * j + 2] + pbcv[2];
      }
    }
  }
}

// CommBrick::reverse_comm, src/comm_brick.cpp:228-258 + Atom::unpack_reverse
void reverse_comm(orc_sys *s) {
  for (int iswap = (int)s->swaps.size() - 1; iswap >= 0; iswap--) {
    Swap &w = s->swaps[iswap];
    int n = (int)w.sendlist.size();
    for (int q = 0; q < n; q++) {
      int j = w.sendlist[q];
      const double *b = &s->Fs[3 * (size_t)(w.firstrecv + q)];
      s->Fs[3 * j + 0] += b[0];
      s->Fs[3 * j + 1] += b[1];
      s->Fs[3 * j + 2] += b[2];
    }
  }
}

/* --------------------------------------------------------- neighbor --- */
// NBin::coord2bin, src/nbin.cpp:61-94 (and NPair::coord2bin, src/npair.cpp:125-158)
int coord2bin(const orc_sys *s, double x, double y, double z) {
  int ix, iy, iz;
  if (x >= s->boxhi[0]) ix = (int)((x - s->boxhi[0]) * s->bininvx) + s->nbinx;
  else if (x >= s->boxlo[0]) { ix = (int)((x - s->boxlo[0]) * s->bininvx); ix = std::min(ix, s->nbinx - 1); }
  else ix = (int)((x - s->boxlo[0]) * s->bininvx) - 1;
  if (y >= s->boxhi[1]) iy = (int)((y - s->boxhi[1]) * s->bininvy) + s->nbiny;
  else if (y >= s->boxlo[1]) { iy = (int)((y - s->boxlo[1]) * s->bininvy); iy = std::min(iy, s->nbiny - 1); }
  else iy = (int)((y - s->boxlo[1]) * s->bininvy) - 1;
  if (z >= s->boxhi[2]) iz = (int)((z - s->boxhi[2]) * s->bininvz) + s->nbinz;
  else if (z >= s->boxlo[2]) { iz = (int)((z - s->boxlo[2]) * s->bininvz); iz = std::min(iz, s->nbinz - 1); }
  else iz = (int)((z - s->boxlo[2]) * s->bininvz) - 1;
  return (iz - s->mbinzlo) * s->mbiny * s->mbinx + (iy - s->mbinylo) * s->mbinx + (ix - s->mbinxlo);
}

// NBinStandard::setup_bins, src/nbin_standard.cpp:53-175
void setup_bins(orc_sys *s) {
  double cutghost = s->cutneigh;
  double bsubboxlo[3], bsubboxhi[3], bbox[3];
  for (int d = 0; d < 3; d++) {
    bsubboxlo[d] = s->sublo[d] - cutghost;
    bsubboxhi[d] = s->subhi[d] + cutghost;
    bbox[d] = s->boxhi[d] - s->boxlo[d];
  }
  double binsize_optimal = 0.5 * s->cutneigh;
  if (binsize_optimal == 0.0) binsize_optimal = bbox[0];
  double binsizeinv = 1.0 / binsize_optimal;
  s->nbinx = (int)(bbox[0] * binsizeinv);
  s->nbiny = (int)(bbox[1] * binsizeinv);
  s->nbinz = (int)(bbox[2] * binsizeinv);
  if (s->nbinx == 0) s->nbinx = 1;
  if (s->nbiny == 0) s->nbiny = 1;
  if (s->nbinz == 0) s->nbinz = 1;
  s->binsizex = bbox[0] / s->nbinx;
  s->binsizey = bbox[1] / s->nbiny;
  s->binsizez = bbox[2] / s->nbinz;
  s->bininvx = 1.0 / s->binsizex;
  s->bininvy = 1.0 / s->binsizey;
  s->bininvz = 1.0 / s->binsizez;
  if (binsize_optimal * s->bininvx > 100 || binsize_optimal * s->bininvy > 100 ||
      binsize_optimal * s->bininvz > 100)
    throw std::runtime_error("Cannot use neighbor bins - box size << cutoff");
  int mbinxhi, mbinyhi, mbinzhi;
  double coord;
  coord = bsubboxlo[0] - SMALL * bbox[0];
  s->mbinxlo = (int)((coord - s->boxlo[0]) * s->bininvx);
  if (coord < s->boxlo[0]) s->mbinxlo = s->mbinxlo - 1;
  coord = bsubboxhi[0] + SMALL * bbox[0];
  mbinxhi = (int)((coord - s->boxlo[0]) * s->bininvx);
  coord = bsubboxlo[1] - SMALL * bbox[1];
  s->mbinylo = (int)((coord - s->boxlo[1]) * s->bininvy);
  if (coord < s->boxlo[1]) s->mbinylo = s->mbinylo - 1;
  coord = bsubboxhi[1] + SMALL * bbox[1];
  mbinyhi = (int)((coord - s->boxlo[1]) * s->bininvy);
  coord = bsubboxlo[2] - SMALL * bbox[2];
  s->mbinzlo = (int)((coord - s->boxlo[2]) * s->bininvz);
  if (coord < s->boxlo[2]) s->mbinzlo = s->mbinzlo - 1;
  coord = bsubboxhi[2] + SMALL * bbox[2];
  mbinzhi = (int)((coord - s->boxlo[2]) * s->bininvz);
  s->mbinxlo -= 1; mbinxhi += 1; s->mbinx = mbinxhi - s->mbinxlo + 1;
  s->mbinylo -= 1; mbinyhi += 1; s->mbiny = mbinyhi - s->mbinylo + 1;
  s->mbinzlo -= 1; mbinzhi += 1; s->mbinz = mbinzhi - s->mbinzlo + 1;
  s->mbins = s->mbinx * s->mbiny * s->mbinz + 1;
  s->binhead.assign(s->mbins, -1);
}

// NStencil::create_setup + bin_distance + full/half create,
// src/nstencil.cpp:101-143, src/nstencil_full_bin_3d.cpp:26-37,
// src/nstencil_half_bin_3d.cpp:26-38
double bin_distance(const orc_sys *s, int i, int j, int k) {
  double delx, dely, delz;
  if (i > 0) delx = (i - 1) * s->binsizex; else if (i == 0) delx = 0.0; else delx = (i + 1) * s->binsizex;
  if (j > 0) dely = (j - 1) * s->binsizey; else if (j == 0) dely = 0.0; else dely = (j + 1) * s->binsizey;
  if (k > 0) delz = (k - 1) * s->binsizez; else if (k == 0) delz = 0.0; else delz = (k + 1) * s->binsizez;
  return (delx * delx + dely * dely + delz * delz);
}
void create_stencils(orc_sys *s) {
  double cutneighsq = s->cutneigh * s->cutneigh;
  int sx = (int)(s->cutneigh * s->bininvx); if (sx * s->binsizex < s->cutneigh) sx++;
  int sy = (int)(s->cutneigh * s->bininvy); if (sy * s->binsizey < s->cutneigh) sy++;
  int sz = (int)(s->cutneigh * s->bininvz); if (sz * s->binsizez < s->cutneigh) sz++;
  s->stencil_full.clear();
  for (int k = -sz; k <= sz; k++)
    for (int j = -sy; j <= sy; j++)
      for (int i = -sx; i <= sx; i++)
        if (bin_distance(s, i, j, k) < cutneighsq)
          s->stencil_full.push_back(k * s->mbiny * s->mbinx + j * s->mbinx + i);
  s->stencil_half.clear();
  for (int k = 0; k <= sz; k++)
    for (int j = -sy; j <= sy; j++)
      for (int i = -sx; i <= sx; i++)
        if (k > 0 || j > 0 || (j == 0 && i > 0))
          if (bin_distance(s, i, j, k) < cutneighsq)
            s->stencil_half.push_back(k * s->mbiny * s->mbinx + j * s->mbinx + i);
}

// NBinStandard::bin_atoms, src/nbin_standard.cpp:181-229
void bin_atoms(orc_sys *s) {
  std::fill(s->binhead.begin(), s->binhead.end(), -1);
  s->local_atom_indices.clear();
  int nstart = s->nowned, nall = s->nowned + s->ngathered;
  s->atom2bin.assign(nall - nstart, 0);
  s->bins.assign(nall - nstart, 0);
  for (int i = nall - 1; i >= nstart; i--) {
    if (s->labels[i] == GHOST) {
      int ibin = coord2bin(s, s->xs[3 * i], s->xs[3 * i + 1], s->xs[3 * i + 2]);
      s->atom2bin[i - nstart] = ibin;
      s->bins[i - nstart] = s->binhead[ibin];
      s->binhead[ibin] = i;
    } else {
      s->local_atom_indices.insert(i);
    }
  }
  for (auto rit = s->local_atom_indices.rbegin(); rit != s->local_atom_indices.rend(); ++rit) {
    int ibin = coord2bin(s, s->xs[3 * *rit], s->xs[3 * *rit + 1], s->xs[3 * *rit + 2]);
    s->atom2bin[*rit - nstart] = ibin;
    s->bins[*rit - nstart] = s->binhead[ibin];
    s->binhead[ibin] = *rit;
  }
}

// NPairGridBin::build, src/npair_grid_bin.cpp:22-127 (CSR instead of MyPage)
void build_gridlist(orc_sys *s) {
  s->g_ilist.clear(); s->g_num.clear(); s->g_off.clear(); s->g_neigh.clear();
  const double cutneighsq = s->cutneigh * s->cutneigh;
  const int nstart = s->nowned;
  const int nst = (int)s->stencil_full.size();
  int gridindex = 0;
  const int local0start = 0;
  for (int k = 0; k < s->nz; k++) {
    double ztmp = (k + local0start) * s->dz() + s->boxlo[2];
    for (int j = 0; j < s->ny; j++) {
      double ytmp = j * s->dy() + s->boxlo[1];
      for (int i = 0; i < s->nx; i++) {
        int n = 0;
        size_t off = s->g_neigh.size();
        double xtmp = i * s->dx() + s->boxlo[0];
        int ibin = coord2bin(s, xtmp, ytmp, ztmp);
        for (int m = 0; m < nst; m++)
          for (int ai = s->binhead[ibin + s->stencil_full[m]]; ai >= 0; ai = s->bins[ai - nstart]) {
            double delx = xtmp - s->xs[3 * ai];
            double dely = ytmp - s->xs[3 * ai + 1];
            double delz = ztmp - s->xs[3 * ai + 2];
            double rsq = delx * delx + dely * dely + delz * delz;
            if (rsq <= cutneighsq) { s->g_neigh.push_back(ai); n++; }
          }
        if (n > 0) {
          s->g_ilist.push_back(gridindex++);
          s->g_off.push_back((int)off);
          s->g_num.push_back(n);
        } else {
          gridindex++;
        }
        if (n > ONEATOM) s->list_overflow = true;
      }
    }
  }
}

// NPairHalfBinNewton::build, src/npair_half_bin_newton.cpp:38-124
void build_halflist(orc_sys *s) {
  s->h_ilist.clear(); s->h_num.clear(); s->h_off.clear(); s->h_neigh.clear();
  const double cutneighsq = s->cutneigh * s->cutneigh;
  const int nstart = s->nowned;
  const int nst = (int)s->stencil_half.size();
  for (int i : s->local_atom_indices) {
    int n = 0;
    size_t off = s->h_neigh.size();
    double xtmp = s->xs[3 * i], ytmp = s->xs[3 * i + 1], ztmp = s->xs[3 * i + 2];
    for (int j = s->bins[i - nstart]; j >= 0; j = s->bins[j - nstart]) {
      if (!s->local_atom_indices.count(j)) {
        if (s->xs[3 * j + 2] < ztmp) continue;
        if (s->xs[3 * j + 2] == ztmp) {
          if (s->xs[3 * j + 1] < ytmp) continue;
          if (s->xs[3 * j + 1] == ytmp && s->xs[3 * j] < xtmp) continue;
        }
      }
      double delx = xtmp - s->xs[3 * j];
      double dely = ytmp - s->xs[3 * j + 1];
      double delz = ztmp - s->xs[3 * j + 2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq <= cutneighsq) { s->h_neigh.push_back(j); n++; }
    }
    int ibin = s->atom2bin[i - nstart];
    for (int k = 0; k < nst; k++)
      for (int j = s->binhead[ibin + s->stencil_half[k]]; j >= 0; j = s->bins[j - nstart]) {
        double delx = xtmp - s->xs[3 * j];
        double dely = ytmp - s->xs[3 * j + 1];
        double delz = ztmp - s->xs[3 * j + 2];
        double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq <= cutneighsq) { s->h_neigh.push_back(j); n++; }
      }
    s->h_ilist.push_back(i);
    s->h_off.push_back((int)off);
    s->h_num.push_back(n);
    if (n > ONEATOM) s->list_overflow = true;
  }
}

// Neighbor::build, src/neighbor.cpp:99-125
void neighbor_build(orc_sys *s) {
  s->ago = 0;
  s->ncalls++;
  s->xholds.assign(s->xs.begin(), s->xs.begin() + 3 * (size_t)s->nowned);
  bin_atoms(s);
  if (s->need_full) build_gridlist(s);
  if (s->need_half) build_halflist(s);
}

// Neighbor::decide + check_distance, src/neighbor.cpp:70-97 (delay 0, every 1)
bool neighbor_decide(orc_sys *s) {
  s->ago++;
  int flag = 0;
  for (int i = 0; i < s->nowned; i++) {
    double d0 = s->xs[3 * i] - s->xholds[3 * i];
    double d1 = s->xs[3 * i + 1] - s->xholds[3 * i + 1];
    double d2 = s->xs[3 * i + 2] - s->xholds[3 * i + 2];
    double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    if (rsq > s->triggersq) flag = 1;
  }
  return flag != 0;
}

/* ------------------------------------------------------------ pairs --- */
// PairGridAtomGaussian::compute (src/pair_gridatom_gaussian.cpp:45-120),
// PairGridAtomLJish::compute (src/pair_gridatom_lj_ish.cpp:44-181) and
// PairGridAtomLJishLinear::compute (src/pair_gridatom_lj_ish_linear.cpp:46-179)
void pair_gridatom(orc_sys *s, const PairStyle &p, double *energy) {
  const double Ox = s->sublo[0], Oy = s->sublo[1], Oz = s->sublo[2];
  const double dv = s->dx() * s->dy() * s->dz();
  for (size_t ii = 0; ii < s->g_ilist.size(); ii++) {
    int gridindex = s->g_ilist[ii];
    int i = gridindex / (s->ny * s->nx);          // reverseFlat: (z, y, x)
    int j = (gridindex / s->nx) % s->ny;
    int k = gridindex % s->nx;
    double xtmp = k * s->dx() + Ox;
    double ytmp = j * s->dy() + Oy;
    double ztmp = i * s->dz() + Oz;
    const int *jlist = &s->g_neigh[s->g_off[ii]];
    int jnum = s->g_num[ii];
    for (int ajj = 0; ajj < jnum; ajj++) {
      int aj = jlist[ajj];
      int ajtype = s->types[aj];
      double delx = xtmp - s->xs[3 * aj];
      double dely = ytmp - s->xs[3 * aj + 1];
      double delz = ztmp - s->xs[3 * aj + 2];
      double rsq = delx * delx + dely * dely + delz * delz;
      bool in = (p.style == ORC_PAIR_GAUSSIAN) ? (rsq <= p.cutsq[ajtype]) : (rsq < p.cutsq[ajtype]);
      if (!in) continue;
      rsq /= p.cutsq[ajtype];
      double expfac;
      if (p.style == ORC_PAIR_LJISH_LINEAR) {   // mu += e, F -= grad phi e dv; no energy
        expfac = -1 * exp(-1. / (1 - rsq) + 1) * p.eps[ajtype];
        s->mu[gridindex] += expfac;
        expfac *= dv;
        s->Fs[3 * aj + 0] -= s->grad[0][gridindex] * expfac;
        s->Fs[3 * aj + 1] -= s->grad[1][gridindex] * expfac;
        s->Fs[3 * aj + 2] -= s->grad[2][gridindex] * expfac;
        continue;
      }
      if (p.style == ORC_PAIR_GAUSSIAN)
        expfac = (exp(-rsq / 2.0) - exp(-0.5)) * exp(0.5) * p.eps[ajtype];
      else
        expfac = exp(-1. / (1 - rsq) + 1) * p.eps[ajtype];
      double delphi = (s->phi[gridindex] - p.phit[ajtype]);
      s->mu[gridindex] += 2 * delphi * expfac;
      if (energy) *energy += delphi * delphi * expfac * dv;
      expfac *= 2 * delphi * dv;
      s->Fs[3 * aj + 0] -= s->grad[0][gridindex] * expfac;
      s->Fs[3 * aj + 1] -= s->grad[1][gridindex] * expfac;
      s->Fs[3 * aj + 2] -= s->grad[2][gridindex] * expfac;
    }
  }
}

// PairLJCut::compute, src/pair_lj_cut.cpp:46-119
void pair_ljcut(orc_sys *s, const PairStyle &p, double *energy) {
  for (size_t ii = 0; ii < s->h_ilist.size(); ii++) {
    int i = s->h_ilist[ii];
    double xtmp = s->xs[3 * i], ytmp = s->xs[3 * i + 1], ztmp = s->xs[3 * i + 2];
    const int *jlist = &s->h_neigh[s->h_off[ii]];
    int jnum = s->h_num[ii];
    for (int jj = 0; jj < jnum; jj++) {
      int j = jlist[jj];
      double delx = xtmp - s->xs[3 * j];
      double dely = ytmp - s->xs[3 * j + 1];
      double delz = ztmp - s->xs[3 * j + 2];
      double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq < p.ljcutsq) {
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        double forcelj = r6inv * (p.lj1 * r6inv - p.lj2);
        double fpair = forcelj * r2inv;
        s->Fs[3 * i + 0] += delx * fpair;
        s->Fs[3 * i + 1] += dely * fpair;
        s->Fs[3 * i + 2] += delz * fpair;
        s->Fs[3 * j + 0] -= delx * fpair;
        s->Fs[3 * j + 1] -= dely * fpair;
        s->Fs[3 * j + 2] -= delz * fpair;
        if (energy) *energy += r6inv * (p.lj1 / 12.0 * r6inv - p.lj2 / 6.0);
      }
    }
  }
}

// PairHarmonicCut::compute, src/pair_harmonic_cut.cpp:45-108: a repulsive
// spring k (cut - r) along the pair axis for r < cut; coincident atoms are
// pushed apart along (0.001, 0.001, 0.001); no energy is tallied.
void pair_harmonic(orc_sys *s, const PairStyle &p) {
  for (size_t ii = 0; ii < s->h_ilist.size(); ii++) {
    int i = s->h_ilist[ii];
    double xtmp = s->xs[3 * i], ytmp = s->xs[3 * i + 1], ztmp = s->xs[3 * i + 2];
    const int *jlist = &s->h_neigh[s->h_off[ii]];
    int jnum = s->h_num[ii];
    for (int jj = 0; jj < jnum; jj++) {
      int j = jlist[jj];
      double delx = xtmp - s->xs[3 * j];
      double dely = ytmp - s->xs[3 * j + 1];
      double delz = ztmp - s->xs[3 * j + 2];
      double rij = sqrt(delx * delx + dely * dely + delz * delz);
      if (!(rij < p.hcut)) continue;
      if (rij == 0) {
        delx = dely = delz = 0.001;
        rij = sqrt(delx * delx + dely * dely + delz * delz);
      }
      double fpair = -1 * p.hk * (rij - p.hcut) / rij;
      s->Fs[3 * i + 0] += delx * fpair;
      s->Fs[3 * i + 1] += dely * fpair;
      s->Fs[3 * i + 2] += delz * fpair;
      s->Fs[3 * j + 0] -= delx * fpair;
      s->Fs[3 * j + 1] -= dely * fpair;
      s->Fs[3 * j + 2] -= delz * fpair;
    }
  }
}

/* ------------------------------------------------------- chain fix ---
 * Build-defined semiflexible chain (the reference's bead-rod dynamics live
 * in the absent beadrodpmer library, src/fixatom_semiflexible.cpp:144-191).
 * Harmonic bonds U = k/2 (r - r0)^2 and bending U = kb (1 - t_a . t_b)
 * between consecutive unit bond vectors, evaluated on unwrapped coordinates
 * of each chain chunk.  Per-bead force is gathered in a fixed order:
 * bond(b-1,b), bond(b,b+1), angle centred at b-1, b, b+1.                  */
inline void bond_force(const double *ua, const double *ub, double k, double r0, double *fa_on_b) {
  // force on b from bond (a,b)
  double d0 = ub[0] - ua[0], d1 = ub[1] - ua[1], d2 = ub[2] - ua[2];
  double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double c = -k * (r - r0) / r;
  fa_on_b[0] = c * d0; fa_on_b[1] = c * d1; fa_on_b[2] = c * d2;
}
// gradient pieces of cos(theta) for the angle (a, b, c): d1 = b - a, d2 = c - b
inline void angle_terms(const double *ua, const double *ubb, const double *uc, double ga[3], double gb[3]) {
  double d1[3] = {ubb[0] - ua[0], ubb[1] - ua[1], ubb[2] - ua[2]};
  double d2[3] = {uc[0] - ubb[0], uc[1] - ubb[1], uc[2] - ubb[2]};
  double r1s = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
  double r2s = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  double r1 = sqrt(r1s), r2 = sqrt(r2s);
  double dot = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
  double inv12 = 1.0 / (r1 * r2);
  double cs = dot * inv12;
  for (int c = 0; c < 3; c++) {
    ga[c] = d2[c] * inv12 - cs * d1[c] / r1s;   // d cos / d d1
    gb[c] = d1[c] * inv12 - cs * d2[c] / r2s;   // d cos / d d2
  }
}
void chain_forces(const double *u, int nb, double k, double r0, double kb, double *f) {
  for (int b = 0; b < nb; b++) {
    double acc[3] = {0, 0, 0}, t[3], ga[3], gb[3];
    if (b >= 1) { bond_force(&u[3 * (b - 1)], &u[3 * b], k, r0, t); for (int c = 0; c < 3; c++) acc[c] += t[c]; }
    if (b + 1 < nb) { bond_force(&u[3 * (b + 1)], &u[3 * b], k, r0, t); for (int c = 0; c < 3; c++) acc[c] += t[c]; }
    // angle centred at b-1: atoms (b-2, b-1, b), b is the third atom -> +kb*gb
    if (b >= 2) { angle_terms(&u[3 * (b - 2)], &u[3 * (b - 1)], &u[3 * b], ga, gb); for (int c = 0; c < 3; c++) acc[c] += kb * gb[c]; }
    // angle centred at b: atoms (b-1, b, b+1), b is the middle -> kb*(ga - gb)
    if (b >= 1 && b + 1 < nb) { angle_terms(&u[3 * (b - 1)], &u[3 * b], &u[3 * (b + 1)], ga, gb); for (int c = 0; c < 3; c++) acc[c] += kb * (ga[c] - gb[c]); }
    // angle centred at b+1: atoms (b, b+1, b+2), b is the first -> -kb*ga
    if (b + 2 < nb) { angle_terms(&u[3 * b], &u[3 * (b + 1)], &u[3 * (b + 2)], ga, gb); for (int c = 0; c < 3; c++) acc[c] += -kb * ga[c]; }
    f[3 * b + 0] = acc[0]; f[3 * b + 1] = acc[1]; f[3 * b + 2] = acc[2];
  }
}
double chain_energy(const double *u, int nb, double k, double r0, double kb) {
  double e = 0;
  for (int b = 0; b + 1 < nb; b++) {
    double d0 = u[3 * b + 3] - u[3 * b], d1 = u[3 * b + 4] - u[3 * b + 1], d2 = u[3 * b + 5] - u[3 * b + 2];
    double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    e += 0.5 * k * (r - r0) * (r - r0);
  }
  for (int b = 1; b + 1 < nb; b++) {
    const double *a = &u[3 * (b - 1)], *m = &u[3 * b], *c = &u[3 * (b + 1)];
    double d1[3] = {m[0] - a[0], m[1] - a[1], m[2] - a[2]};
    double d2[3] = {c[0] - m[0], c[1] - m[1], c[2] - m[2]};
    double r1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
    double r2 = sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
    e += kb * (1 - (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) / (r1 * r2));
  }
  return e;
}

void chain_halfstep(orc_sys *s, const FixDef &f) {
  for (size_t c = 0; c < f.start.size(); c++) {
    int st = f.start[c], nb = f.end[c] - f.start[c];
    std::vector<double> fc(3 * (size_t)nb);
    chain_forces(&s->uxs[3 * (size_t)st], nb, f.kbond, f.r0, f.kbend, fc.data());
    for (int b = 0; b < nb; b++) {
      int j = st + b;
      for (int q = 0; q < 3; q++) {
        double tot = s->Fs[3 * j + q] + fc[3 * b + q];
        s->uxs[3 * j + q] += f.drag * tot * s->dt * 0.5;
      }
      map(s, &s->uxs[3 * j], s->images[j], &s->xs[3 * j]);
    }
  }
}

// FixAtomDrag::initial_integrate / final_integrate, src/fixatom_drag.cpp:43-66
void drag_halfstep(orc_sys *s, const FixDef &f) {
  for (size_t c = 0; c < f.start.size(); c++)
    for (int j = f.start[c]; j < f.end[c]; j++)
      for (int q = 0; q < 3; q++) s->xs[3 * j + q] += f.drag * s->Fs[3 * j + q] * s->dt * 0.5;
}

/* --------------------------------------------------- bead-rod chains ---
 * fix atom/semiflexible/{notether,singletether,doubletether}
 * (src/fixatom_semiflexible.cpp:27-191).  The reference hands the dynamics
 * to beadrodpmer (absent here, no version pinned), so this restates a
 * published bead-rod Brownian dynamics instead (Montesi, Morse & Pasquali,
 * J. Chem. Phys. 122, 084903 (2005)):
 *   - rigid rods of length b between consecutive beads, tension T_j;
 *   - bending U = kb sum_j (1 - u_j.u_{j+1});
 *   - the metric pseudo-potential (kT/2) ln det G, G_jj = w_j + w_{j+1},
 *     G_{j,j+1} = -w_{j+1} u_j.u_{j+1}, under which the rigid chain samples
 *     independent bond orientations weighted by exp(-U/kT) (Fixman);
 *   - bead mobility H_i = w_i [t t / zeta_para + (I - t t) / zeta_perp]
 *     about the bead tangent t (w_i = 0 for a tethered bead: bead 0 for
 *     singletether, beads 0 and N-1 for doubletether);
 *   - Brownian force sqrt(2 kT / dt) [sqrt(zeta_para) t t + sqrt(zeta_perp)
 *     (I - t t)] xi at the step's start, xi of unit variance;
 *   - first_step: the tensions that keep every rod length constant to first
 *     order, then the half step R* = R0 + dt/2 V(R0);
 *   - second_step: R1 = R0 + dt V(R*), every mobility, direction and force
 *     taken at R*, with the tensions solved by Newton iteration on
 *     |R1_{j+1} - R1_j|^2 = b^2 (tridiagonal Jacobian), at most itermax
 *     iterations (src/fixatom_semiflexible.cpp:160-187).
 * xi: uniform deviates (w >> 11) 2^-53 - 1/2 scaled by sqrt(12), words 0-2
 * of Philox4x64-10(counter {bead in polymer, step, 0, 0}, key {polymer seed,
 * 0x4245414452}).  Parity with beadrodpmer itself: unpinned.              */
constexpr uint64_t BR_KEY1 = 0x4245414452ull;
constexpr double BR_TOL = 1e-12;   // |R1_{j+1} - R1_j|^2 - b^2 within BR_TOL b^2

// Philox4x64-10 (Salmon, Moraes, Dror & Shaw, SC'11): ten rounds of the
// 64-bit multiply-xor round with Weyl key increments
void philox4x64_10_host(uint64_t c[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  for (int r = 0; r < 10; r++) {
    const unsigned __int128 p0 = (unsigned __int128)M0 * c[0];
    const unsigned __int128 p1 = (unsigned __int128)M1 * c[2];
    const uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    const uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += W0;
    k1 += W1;
  }
}

struct BRParams {
  double b, zpar, zperp, kb, temp, dt;
  int tether;
};

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// chain geometry at positions R (3N): unit bonds u (3m), inverse lengths il,
// bead tangents t (3N)
struct BRGeom {
  int N = 0, m = 0;
  std::vector<double> u, il, t, w;
  void compute(const double *R, int n, int tether) {
    N = n;
    m = n - 1;
    u.assign(3 * (size_t)m, 0.0);
    il.assign(m, 0.0);
    t.assign(3 * (size_t)N, 0.0);
    w.assign(N, 1.0);
    if (tether >= 1) w[0] = 0.0;
    if (tether >= 2) w[N - 1] = 0.0;
    for (int j = 0; j < m; j++) {
      double d[3];
      for (int q = 0; q < 3; q++) d[q] = R[3 * (j + 1) + q] - R[3 * j + q];
      const double l = sqrt(dot3(d, d));
      il[j] = 1.0 / l;
      for (int q = 0; q < 3; q++) u[3 * j + q] = d[q] * il[j];
    }
    for (int i = 0; i < N; i++) {
      double *ti = &t[3 * i];
      if (m == 0) continue;   // a lone bead: isotropic, zeta_perp
      if (i == 0 || i == N - 1) {
        const double *uu = &u[3 * (i == 0 ? 0 : m - 1)];
        for (int q = 0; q < 3; q++) ti[q] = uu[q];
        continue;
      }
      double sv[3];
      for (int q = 0; q < 3; q++) sv[q] = u[3 * (i - 1) + q] + u[3 * i + q];
      const double n2 = sqrt(dot3(sv, sv));
      if (n2 > 1e-12) for (int q = 0; q < 3; q++) ti[q] = sv[q] / n2;
      else for (int q = 0; q < 3; q++) ti[q] = u[3 * i + q];
    }
  }
  // out = H_i v
  void mob(const BRParams &p, int i, const double *v, double *out) const {
    if (w[i] == 0.0) { out[0] = out[1] = out[2] = 0.0; return; }
    const double *ti = &t[3 * i];
    const double tv = dot3(ti, v);
    const double a = 1.0 / p.zperp, c = 1.0 / p.zpar - 1.0 / p.zperp;
    for (int q = 0; q < 3; q++) out[q] = v[q] * a + c * tv * ti[q];
  }
};

// F (3N) += bending and metric pseudo-potential forces.  The metric is the
// rod-tension matrix A = G H G^T (G: rows grad |R_{j+1} - R_j|), i.e.
// A_jj = u_j.(H_j + H_{j+1}) u_j, A_{j,j+1} = -u_j.H_{j+1} u_{j+1}; with the
// bead tangents of BRGeom every entry is a function of the bond cosines
// c_j = u_j.u_{j+1} only ((t.u)^2 = (1 + c)/2 at an interior bead, 1 at an
// end), so U_ps = (kT/2) ln det A acts through the cosines like the bending
// energy: F = sum_j (kb - (kT/2) d ln det A / d c_j) grad c_j.
void br_bend_ps(const BRParams &p, const BRGeom &g, double *F) {
  const int m = g.m, N = g.N;
  if (m < 2) return;
  const double ip = 1.0 / p.zperp, kh = 1.0 / p.zpar - 1.0 / p.zperp;
  std::vector<double> c(m - 1), a(m), off(m, 0.0), al(m), be(m);
  for (int j = 0; j + 1 < m; j++) c[j] = dot3(&g.u[3 * j], &g.u[3 * (j + 1)]);
  for (int j = 0; j < m; j++) {
    const double aL = j == 0 ? 1.0 : 0.5 * (1.0 + c[j - 1]);       // (t_j.u_j)^2
    const double aR = j + 1 == N - 1 ? 1.0 : 0.5 * (1.0 + c[j]);   // (t_{j+1}.u_j)^2
    a[j] = (g.w[j] + g.w[j + 1]) * ip + kh * (g.w[j] * aL + g.w[j + 1] * aR);
    if (j + 1 < m) off[j] = -g.w[j + 1] * (c[j] * ip + kh * 0.5 * (1.0 + c[j]));
  }
  auto safe = [](double x) { return fabs(x) > 1e-300; };
  al[0] = a[0];
  for (int j = 1; j < m; j++) al[j] = safe(al[j - 1]) ? a[j] - off[j - 1] * off[j - 1] / al[j - 1] : a[j];
  be[m - 1] = a[m - 1];
  for (int j = m - 2; j >= 0; j--) be[j] = safe(be[j + 1]) ? a[j] - off[j] * off[j] / be[j + 1] : a[j];
  // diagonal of A^-1: 1 / (al_j + be_j - a_j)
  std::vector<double> dinv(m);
  for (int j = 0; j < m; j++) {
    const double den = al[j] + be[j] - a[j];
    dinv[j] = safe(den) ? 1.0 / den : 0.0;
  }
  for (int j = 0; j + 1 < m; j++) {
    double e = p.kb;
    if (p.temp != 0.0) {
      // (A^-1)_{j,j+1} = -off_j / be_{j+1} (A^-1)_jj
      const double oinv = safe(be[j + 1]) ? -off[j] / be[j + 1] * dinv[j] : 0.0;
      const double w1 = g.w[j + 1];
      const double dlog = 2.0 * oinv * (-w1 * (ip + 0.5 * kh)) + 0.5 * kh * w1 * (dinv[j] + dinv[j + 1]);
      e -= 0.5 * p.temp * dlog;
    }
    const double *u0 = &g.u[3 * j], *u1 = &g.u[3 * (j + 1)];
    for (int q = 0; q < 3; q++) {
      const double pj = g.il[j] * (u1[q] - c[j] * u0[q]);
      const double qj = g.il[j + 1] * (u0[q] - c[j] * u1[q]);
      F[3 * j + q] -= e * pj;
      F[3 * (j + 1) + q] += e * (pj - qj);
      F[3 * (j + 2) + q] += e * qj;
    }
  }
}

// Thomas algorithm: lo_j x_{j-1} + di_j x_j + up_j x_{j+1} = r_j (x in r);
// a vanishing pivot leaves its row's unknown 0
void br_thomas(int m, const double *lo, const double *di, const double *up, double *r) {
  std::vector<double> cp(m);
  for (int j = 0; j < m; j++) {
    const double den = di[j] - (j > 0 ? lo[j] * cp[j - 1] : 0.0);
    if (fabs(den) > 1e-300) {
      cp[j] = up[j] / den;
      r[j] = (r[j] - (j > 0 ? lo[j] * r[j - 1] : 0.0)) / den;
    } else {
      cp[j] = 0.0;
      r[j] = 0.0;
    }
  }
  for (int j = m - 2; j >= 0; j--) r[j] -= cp[j] * r[j + 1];
}

// tensions T (m) that keep u_j.(V_{j+1} - V_j) = 0 for the forces F
void br_tension_linear(const BRParams &p, const BRGeom &g, const double *F, double *T) {
  const int m = g.m;
  std::vector<double> lo(m, 0.0), di(m), up(m, 0.0);
  for (int j = 0; j < m; j++) {
    const double *uj = &g.u[3 * j];
    double h0[3], h1[3], f0[3], f1[3];
    g.mob(p, j, uj, h0);
    g.mob(p, j + 1, uj, h1);
    di[j] = -(dot3(uj, h1) + dot3(uj, h0));
    if (j > 0) {
      double hl[3];
      g.mob(p, j, &g.u[3 * (j - 1)], hl);
      lo[j] = dot3(uj, hl);
    }
    if (j + 1 < m) {
      double hu[3];
      g.mob(p, j + 1, &g.u[3 * (j + 1)], hu);
      up[j] = dot3(uj, hu);
    }
    g.mob(p, j, &F[3 * j], f0);
    g.mob(p, j + 1, &F[3 * (j + 1)], f1);
    double df[3] = {f1[0] - f0[0], f1[1] - f0[1], f1[2] - f0[2]};
    T[j] = -dot3(uj, df);
  }
  br_thomas(m, lo.data(), di.data(), up.data(), T);
}

// total force (without tension) of the bead-rod step: external F + Brownian
// FB + bending and pseudo-potential at the geometry g
std::vector<double> br_forces(const BRParams &p, const BRGeom &g, const double *Fext, const double *FB) {
  std::vector<double> F(3 * (size_t)g.N);
  for (size_t k = 0; k < F.size(); k++) F[k] = Fext[k] + FB[k];
  br_bend_ps(p, g, F.data());
  return F;
}

// Brownian forces FB (3N) at the tangents of g
void br_noise(const BRParams &p, const BRGeom &g, uint64_t seed, int64_t step, double *FB) {
  const double amp = p.temp > 0 ? sqrt(2.0 * p.temp / p.dt) : 0.0;
  const double s12 = sqrt(12.0), sp = sqrt(p.zpar), sq = sqrt(p.zperp);
  for (int i = 0; i < g.N; i++) {
    double xi[3] = {0.0, 0.0, 0.0};
    if (amp != 0.0) {
      uint64_t c[4] = {(uint64_t)i, (uint64_t)step, 0, 0};
      philox4x64_10_host(c, seed, BR_KEY1);
      for (int q = 0; q < 3; q++) xi[q] = s12 * ((double)(c[q] >> 11) * 0x1.0p-53 - 0.5);
    }
    const double *t = &g.t[3 * i];
    const double tx = dot3(t, xi);
    for (int q = 0; q < 3; q++) FB[3 * i + q] = amp * (sp * tx * t[q] + sq * (xi[q] - tx * t[q]));
  }
}

// first_step of one polymer of N beads: R (3N) holds R0 on entry and R* on
// return; R0, FB (3N) and T (N-1) are kept for second_step
void br_first_step(const BRParams &p, uint64_t seed, int64_t step, int N, double *R, const double *Fext, double *R0,
                   double *FB, double *T) {
  std::copy(R, R + 3 * (size_t)N, R0);
  BRGeom g;
  g.compute(R0, N, p.tether);
  br_noise(p, g, seed, step, FB);
  std::vector<double> F = br_forces(p, g, Fext, FB);
  if (g.m > 0) br_tension_linear(p, g, F.data(), T);
  for (int i = 0; i < N; i++) {
    double tot[3], v[3];
    for (int q = 0; q < 3; q++) {
      tot[q] = F[3 * i + q];
      if (i < g.m) tot[q] += T[i] * g.u[3 * i + q];
      if (i > 0) tot[q] -= T[i - 1] * g.u[3 * (i - 1) + q];
    }
    g.mob(p, i, tot, v);
    for (int q = 0; q < 3; q++) R[3 * i + q] = R0[3 * i + q] + 0.5 * p.dt * v[q];
  }
}

// second_step of one polymer: R holds R* on entry and R1 on return; T (the
// first step's tensions on entry) is the Newton start.  Returns the Newton
// iterations taken, itermax + 1 when the rods were not brought to length.
int br_second_step(const BRParams &p, int N, double *R, const double *Fext, const double *R0, const double *FB,
                   double *T, int itermax) {
  BRGeom g;
  g.compute(R, N, p.tether);
  std::vector<double> F = br_forces(p, g, Fext, FB);
  const int m = g.m;
  std::vector<double> P(3 * (size_t)N), hA(3 * (size_t)m), hB(3 * (size_t)m), D(3 * (size_t)m);
  std::vector<double> r(m), lo(m), di(m), up(m);
  for (int i = 0; i < N; i++) {
    double v[3];
    g.mob(p, i, &F[3 * i], v);
    for (int q = 0; q < 3; q++) P[3 * i + q] = R0[3 * i + q] + p.dt * v[q];
  }
  for (int j = 0; j < m; j++) {
    g.mob(p, j, &g.u[3 * j], &hA[3 * j]);
    g.mob(p, j + 1, &g.u[3 * j], &hB[3 * j]);
  }
  int it = 0;
  const double tol = BR_TOL * p.b * p.b;
  for (;; it++) {
    double gmax = 0.0;
    for (int j = 0; j < m; j++) {
      for (int q = 0; q < 3; q++) {
        double d = P[3 * (j + 1) + q] - P[3 * j + q];
        double tq = -T[j] * hB[3 * j + q] - T[j] * hA[3 * j + q];
        if (j + 1 < m) tq += T[j + 1] * hA[3 * (j + 1) + q];
        if (j > 0) tq += T[j - 1] * hB[3 * (j - 1) + q];
        D[3 * j + q] = d + p.dt * tq;
      }
      r[j] = dot3(&D[3 * j], &D[3 * j]) - p.b * p.b;
      gmax = std::max(gmax, fabs(r[j]));
    }
    if (gmax <= tol) break;
    if (it == itermax) {
      it = itermax + 1;
      break;
    }
    for (int j = 0; j < m; j++) {
      const double *Dj = &D[3 * j];
      double s2[3] = {hA[3 * j] + hB[3 * j], hA[3 * j + 1] + hB[3 * j + 1], hA[3 * j + 2] + hB[3 * j + 2]};
      di[j] = -2.0 * p.dt * dot3(Dj, s2);
      lo[j] = j > 0 ? 2.0 * p.dt * dot3(Dj, &hB[3 * (j - 1)]) : 0.0;
      up[j] = j + 1 < m ? 2.0 * p.dt * dot3(Dj, &hA[3 * (j + 1)]) : 0.0;
    }
    br_thomas(m, lo.data(), di.data(), up.data(), r.data());
    for (int j = 0; j < m; j++) T[j] -= r[j];
  }
  for (int i = 0; i < N; i++)
    for (int q = 0; q < 3; q++) {
      double v = P[3 * i + q];
      if (i < m) v += p.dt * T[i] * hA[3 * i + q];
      if (i > 0) v -= p.dt * T[i - 1] * hB[3 * (i - 1) + q];
      R[3 * i + q] = v;
    }
  return it;
}

// bending energy kb sum (1 - u_j.u_{j+1}) of one polymer
double br_energy(const double *R, int N, double kb) {
  BRGeom g;
  g.compute(R, N, 0);
  double e = 0.0;
  for (int j = 0; j + 1 < g.m; j++) e += kb * (1.0 - dot3(&g.u[3 * j], &g.u[3 * (j + 1)]));
  return e;
}

/* ----------------------------------------------------------- phases --- */
void phase_start(orc_sys *s) {
  for (auto &f : s->fixes) {
    if (f.kind == FixDef::CONJ ||
        (f.kind == FixDef::GRADPHI && (!f.once || s->timestep == s->firststep))) {
      std::vector<cplx> ph;
      gradphi_impl(s->nx, s->ny, s->nz, s->period, s->phi.data(), s->grad[0].data(),
                   s->grad[1].data(), s->grad[2].data(), f.kind == FixDef::CONJ ? &ph : nullptr);
      if (f.kind == FixDef::CONJ) s->ftphi.swap(ph);
    }
  }
}

void phase_neigh(orc_sys *s) {
  if (!s->neigh_set) return;
  if (neighbor_decide(s)) {
    pbc(s);
    borders(s);
    neighbor_build(s);
  } else {
    forward_comm(s);
  }
}

void phase_pair(orc_sys *s) {
  // Integrate::run, src/integrate.cpp:130-146: the second pass sees the same
  // inputs as the first (finding F4), so one pass after mu = 0 gives the
  // state the reference holds after its second pass.
  std::fill(s->mu.begin(), s->mu.end(), 0.0);
  std::fill(s->Fs.begin(), s->Fs.end(), 0.0);
  for (auto &p : s->pairs) {
    if (p.style == ORC_PAIR_LJCUT) pair_ljcut(s, p, nullptr);
    else if (p.style == ORC_PAIR_HARMONIC) pair_harmonic(s, p);
    else pair_gridatom(s, p, nullptr);
  }
  reverse_comm(s);
}

BRParams br_params(const orc_sys *s, const FixDef &f) {
  return BRParams{f.bondlength, f.zpar, f.zperp, f.bending, f.brtemp, s->dt, f.tether};
}

// FixAtomSemiFlexible::initial_integrate, src/fixatom_semiflexible.cpp:144-156
void semiflex_first(orc_sys *s, FixDef &f) {
  const BRParams p = br_params(s, f);
  f.brstep++;
  for (size_t c = 0; c < f.start.size(); c++) {
    const int st = f.start[c], nb = f.end[c] - f.start[c];
    if (nb <= 0) continue;
    const size_t o = 3 * (size_t)st;
    br_first_step(p, (uint64_t)(uint32_t)f.pseed[c], f.brstep, nb, &s->uxs[o], &s->Fs[o], &f.R0[o], &f.FB[o],
                  &f.T[st]);
    for (int j = st; j < st + nb; j++) map(s, &s->uxs[3 * j], s->images[j], &s->xs[3 * j]);
  }
}

// FixAtomSemiFlexible::final_integrate, src/fixatom_semiflexible.cpp:158-191
void semiflex_second(orc_sys *s, FixDef &f) {
  const BRParams p = br_params(s, f);
  const int itermax = 50;
  int errflag = 0;
  for (size_t c = 0; c < f.start.size(); c++) {
    const int st = f.start[c], nb = f.end[c] - f.start[c];
    if (nb <= 0) continue;
    const size_t o = 3 * (size_t)st;
    const int it = br_second_step(p, nb, &s->uxs[o], &s->Fs[o], &f.R0[o], &f.FB[o], &f.T[st], itermax);
    for (int j = st; j < st + nb; j++) map(s, &s->uxs[3 * j], s->images[j], &s->xs[3 * j]);
    if (it > itermax) errflag = 1;
  }
  if (errflag) throw std::runtime_error("Failed to maintain polymer bond-length constraints.");
}

// FixAtomSemiFlexible::check_bonds, src/fixatom_semiflexible.cpp:114-140
void semiflex_check_bonds(const orc_sys *s, const FixDef &f) {
  for (size_t c = 0; c < f.start.size(); c++)
    for (int mu = f.start[c] + 1; mu < f.end[c]; mu++) {
      const double *a = &s->uxs[3 * (size_t)mu], *b = &s->uxs[3 * (size_t)(mu - 1)];
      const double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
      if (fabs(sqrt(dot3(d, d)) - f.bondlength) > 1e-3)
        throw std::runtime_error("Adjacent atom displacements do not match the FixAtomSemiFlexible "
                                 "bondlength specified.");
    }
}

void phase_integrate(orc_sys *s) {
  for (auto &f : s->fixes) {
    if (f.kind == FixDef::DRAG) drag_halfstep(s, f);
    else if (f.kind == FixDef::CHAIN) chain_halfstep(s, f);
    else if (f.kind == FixDef::SEMIFLEX) semiflex_first(s, f);
  }
}

void phase_postforce(orc_sys *s) {
  for (auto &f : s->fixes)
    if (f.kind == FixDef::FH)
      floryhuggins_impl((int64_t)s->N, s->phi.data(), s->mu.data(), f.temp, f.volFH, f.chi,
                        s->nless, s->ngreat);
}

void phase_prefinal(orc_sys *s) {
  for (auto &f : s->fixes)
    if (f.kind == FixDef::CONJ) {
      s->ftmu.resize(s->M);
      fft3_r2c(s->nx, s->ny, s->nz, s->mu.data(), s->ftmu.data());
    }
}

void phase_final(orc_sys *s) {
  for (auto &f : s->fixes) {
    if (f.kind == FixDef::CONJ) {
      SpectralParams p{};
      p.nx = s->nx; p.ny = s->ny; p.nz = s->nz;
      for (int d = 0; d < 3; d++) p.period[d] = s->period[d];
      p.mobility = f.mobility; p.temp = f.temp; p.volFH = f.volFH; p.gamma = f.gamma; p.dt = s->dt;
      p.noise_mode = f.noise_mode;
      p.semi = s->semi;
      p.noise = s->noisefield.empty() ? nullptr : s->noisefield.data();
      if (p.noise_mode == ORC_NOISE_FIELD && !p.noise) p.noise_mode = ORC_NOISE_OFF;
      p.gen = &f.gen;
      SpectralUpdater up(p);
      up.update(s->ftphi.data(), s->ftmu.data());
      if (f.noise_mode == ORC_NOISE_FIELD) s->noisefield.clear();   // one field per update
    } else if (f.kind == FixDef::DRAG) {
      drag_halfstep(s, f);
    } else if (f.kind == FixDef::CHAIN) {
      chain_halfstep(s, f);
    } else if (f.kind == FixDef::SEMIFLEX) {
      semiflex_second(s, f);
    }
  }
}

void phase_postfinal(orc_sys *s) {
  for (auto &f : s->fixes)
    if (f.kind == FixDef::CONJ) fft3_c2r(s->nx, s->ny, s->nz, s->ftphi.data(), s->phi.data());
}

void run_phase(orc_sys *s, int ph) {
  switch (ph) {
    case ORC_PH_START: phase_start(s); break;
    case ORC_PH_NEIGH: phase_neigh(s); break;
    case ORC_PH_PAIR: phase_pair(s); break;
    case ORC_PH_INITIAL: phase_integrate(s); break;
    case ORC_PH_POSTFORCE: phase_postforce(s); break;
    case ORC_PH_PREFINAL: phase_prefinal(s); break;
    case ORC_PH_FINAL: phase_final(s); break;
    case ORC_PH_POSTFINAL: phase_postfinal(s); break;
    default: throw std::runtime_error("bad phase");
  }
}

template <class F>
int guard(orc_sys *s, F f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    if (s) s->err = e.what();
    return -1;
  }
}

}  // namespace

/* ================================================================ C API */
extern "C" {

int orc_fft_r2c(int nx, int ny, int nz, const double *in, double *out) {
  return guard(nullptr, [&] { fft3_r2c(nx, ny, nz, in, reinterpret_cast<cplx *>(out)); });
}
int orc_fft_c2r(int nx, int ny, int nz, const double *in, double *out) {
  return guard(nullptr, [&] { fft3_c2r(nx, ny, nz, reinterpret_cast<const cplx *>(in), out); });
}
int orc_gradphi(int nx, int ny, int nz, const double period[3], const double *phi, double *gx,
                double *gy, double *gz) {
  return guard(nullptr, [&] { gradphi_impl(nx, ny, nz, period, phi, gx, gy, gz, nullptr); });
}
int orc_floryhuggins(int64_t n, double *phi, double *mu, double temp, double volFH, double chi,
                     int64_t *nless, int64_t *ngreat) {
  int64_t a = 0, b = 0;
  floryhuggins_impl(n, phi, mu, temp, volFH, chi, a, b);
  if (nless) *nless = a;
  if (ngreat) *ngreat = b;
  return 0;
}
int orc_spectral_update(int nx, int ny, int nz, const double period[3], double *phihat,
                        const double *muhat, double mobility, double temp, double volFH,
                        double gamma, double dt, int noise_mode, int seed, const double *noise) {
  return guard(nullptr, [&] {
    std::mt19937 gen;
    gen.seed(seed);
    SpectralParams p{};
    p.nx = nx; p.ny = ny; p.nz = nz;
    for (int d = 0; d < 3; d++) p.period[d] = period[d];
    p.mobility = mobility; p.temp = temp; p.volFH = volFH; p.gamma = gamma; p.dt = dt;
    p.noise_mode = noise_mode; p.noise = noise; p.gen = &gen;
    SpectralUpdater up(p);
    up.update(reinterpret_cast<cplx *>(phihat), reinterpret_cast<const cplx *>(muhat));
  });
}
int orc_noise_field(int nx, int ny, int nz, const double period[3], int seed, double mobility,
                    double temp, double dt, double *noise) {
  return guard(nullptr, [&] {
    const size_t M = (size_t)nz * ny * (nx / 2 + 1);
    std::vector<cplx> ph(M, cplx(0, 0)), mu(M, cplx(0, 0));
    std::vector<double> rec(2 * M, 0.0);
    std::mt19937 gen;
    gen.seed(seed);
    SpectralParams p{};
    p.nx = nx; p.ny = ny; p.nz = nz;
    for (int d = 0; d < 3; d++) p.period[d] = period[d];
    p.mobility = mobility; p.temp = temp; p.volFH = 1; p.gamma = 1; p.dt = dt;
    p.noise_mode = ORC_NOISE_MT19937; p.gen = &gen; p.record = &rec;
    SpectralUpdater up(p);
    up.update(ph.data(), mu.data());
    std::memcpy(noise, rec.data(), sizeof(double) * 2 * M);
  });
}
int orc_mt19937_uniform(int seed, int n, double *out) {
  std::mt19937 gen;
  gen.seed(seed);
  std::uniform_real_distribution<double> d(-0.5, 0.5);
  for (int i = 0; i < n; i++) out[i] = d(gen);
  return 0;
}
int orc_make_unique_seed(int seed) { return make_unique_seed(seed); }
int orc_set_threads(int n) {
  if (n < 1) return -1;
  omp_set_num_threads(n);
  return 0;
}
int orc_threads() { return omp_get_max_threads(); }
int orc_mt19937_raw(int seed, int n, uint32_t *out) {
  std::mt19937 gen;
  gen.seed(seed);
  for (int i = 0; i < n; i++) out[i] = (uint32_t)gen();
  return 0;
}

orc_sys *orc_create(int nx, int ny, int nz, const double boxlo[3], const double period[3],
                    int zghost_mode) {
  if (!is_pow2(nx) || !is_pow2(ny) || !is_pow2(nz) || nx < 2 || ny < 2 || nz < 2) return nullptr;
  orc_sys *s = new orc_sys();
  s->nx = nx; s->ny = ny; s->nz = nz; s->hx = nx / 2 + 1;
  for (int d = 0; d < 3; d++) {
    s->boxlo[d] = boxlo[d];
    s->period[d] = period[d];
    s->boxhi[d] = boxlo[d] + period[d];   // src/domain.cpp:71-72
  }
  // Domain::set_subbox, src/domain.cpp:78-99 (local0start = 0)
  double dzz = s->period[2] / s->nz;
  s->sublo[0] = s->boxlo[0]; s->sublo[1] = s->boxlo[1]; s->sublo[2] = dzz * 0 + s->boxlo[2];
  s->subhi[0] = s->boxhi[0]; s->subhi[1] = s->boxhi[1]; s->subhi[2] = dzz * (0 + s->nz) + s->boxlo[2];
  s->zghost_mode = zghost_mode;
  s->N = (size_t)nx * ny * nz;
  s->M = (size_t)nz * ny * s->hx;
  s->phi.assign(s->N, 0.0);
  s->mu.assign(s->N, 0.0);
  for (int d = 0; d < 3; d++) s->grad[d].assign(s->N, 0.0);
  s->ftphi.assign(s->M, cplx(0, 0));
  s->ftmu.assign(s->M, cplx(0, 0));
  return s;
}
void orc_destroy(orc_sys *s) { delete s; }
const char *orc_last_error(orc_sys *s) { return s ? s->err.c_str() : ""; }

int orc_set_field(orc_sys *s, int which, const double *data) {
  std::vector<double> *f = which == ORC_PHI ? &s->phi : which == ORC_MU ? &s->mu
                           : (which >= 2 && which <= 4) ? &s->grad[which - 2] : nullptr;
  if (!f) return -1;
  std::memcpy(f->data(), data, sizeof(double) * s->N);
  return 0;
}
int orc_get_field(orc_sys *s, int which, double *data) {
  const std::vector<double> *f = which == ORC_PHI ? &s->phi : which == ORC_MU ? &s->mu
                                 : (which >= 2 && which <= 4) ? &s->grad[which - 2] : nullptr;
  if (!f) return -1;
  std::memcpy(data, f->data(), sizeof(double) * s->N);
  return 0;
}

// Atom::setup + ReadAtoms + Atom::add_atom + check_tags_and_types,
// src/atom.cpp:26-110, 140-252
int orc_set_atoms(orc_sys *s, int n, const double *xyz, const int *type, const int *tag,
                  const int *mol) {
  return guard(s, [&] {
    s->nowned = n;
    s->ngathered = s->nlocal = s->nghost = 0;
    s->xs.assign(xyz, xyz + 3 * (size_t)n);
    s->uxs.assign(3 * (size_t)n, 0.0);
    s->Fs.assign(3 * (size_t)n, 0.0);
    s->types.assign(type, type + n);
    s->tags.resize(n);
    s->moltags.assign(n, 0);
    s->labels.assign(n, OWNED);
    s->owner.resize(n);
    int img = (IMGMAX << IMG2BITS) | (IMGMAX << IMGBITS) | IMGMAX;   // Domain::set_image
    s->images.assign(n, img);
    int tmax = -1;
    for (int i = 0; i < n; i++) {
      s->tags[i] = tag ? tag[i] : i;
      if (mol) s->moltags[i] = mol[i];
      s->owner[i] = i;
      tmax = std::max(tmax, type[i]);
    }
    std::set<int> tset(s->tags.begin(), s->tags.end());
    if ((int)tset.size() != n) throw std::runtime_error("duplicate IDs found on processor(s).");
    s->ntypes = tmax + 1;
  });
}
int orc_get_atoms(orc_sys *s, double *xyz, double *uxyz, int *image) {
  if (xyz) std::memcpy(xyz, s->xs.data(), sizeof(double) * 3 * s->nowned);
  if (uxyz) std::memcpy(uxyz, s->uxs.data(), sizeof(double) * 3 * s->nowned);
  if (image) std::memcpy(image, s->images.data(), sizeof(int) * s->nowned);
  return 0;
}
int orc_get_forces(orc_sys *s, double *f) {
  std::memcpy(f, s->Fs.data(), sizeof(double) * 3 * s->nowned);
  return 0;
}

// Pair settings (src/pair_gridatom_gaussian.cpp:123-140,
// src/pair_gridatom_lj_ish.cpp:184-204, src/pair_gridatom_lj_ish_linear.cpp:180-197,
// src/pair_lj_cut.cpp:101-117, src/pair_harmonic_cut.cpp:111-124)
int orc_add_pair(orc_sys *s, int style, double a, double b, double c) {
  if (s->ntypes <= 0) { s->err = "Cannot create pair style before atoms are made."; return -1; }
  if (style < ORC_PAIR_GAUSSIAN || style > ORC_PAIR_HARMONIC) { s->err = "Unrecognised pair style."; return -1; }
  PairStyle p;
  p.style = style;
  if (style == ORC_PAIR_HARMONIC) {   // a = cut, b = springK
    p.maxcut = a;
    p.hcut = a;
    p.hk = b;
    s->need_half = true;
  } else if (style == ORC_PAIR_LJCUT) {
    p.maxcut = a;
    p.lj1 = 48 * b * pow(c, 12.0);
    p.lj2 = 24 * b * pow(c, 6.0);
    p.ljcutsq = a * a;
    s->need_half = true;
  } else {
    p.maxcut = a;
    p.eps.assign(s->ntypes, 1.0);
    p.phit.assign(s->ntypes, 0.1);
    p.cutsq.assign(s->ntypes, a * a);
    s->need_full = true;
  }
  s->pairs.push_back(p);
  return (int)s->pairs.size() - 1;
}
// pairCoeffs (src/pair_gridatom_gaussian.cpp:143-181)
int orc_pair_coeff(orc_sys *s, int ipair, int type, double eps, double phit, double nucwidth) {
  if (ipair < 0 || ipair >= (int)s->pairs.size()) return -1;
  PairStyle &p = s->pairs[ipair];
  if (type < 0 || type >= (int)p.eps.size()) { s->err = "atom type out of range."; return -1; }
  if (!std::isnan(eps)) p.eps[type] = eps;
  if (!std::isnan(phit)) p.phit[type] = phit;
  if (!std::isnan(nucwidth)) {
    if (nucwidth > p.maxcut) { s->err = "nucwidth cannot be larger than maximum cut."; return -1; }
    p.cutsq[type] = nucwidth * nucwidth;
  }
  return 0;
}
// FixGridConjugate::init + ConjugateVolFrac::readCoeffs,
// src/fixgrid_conjugate.cpp:21-45, src/conjugate_volfrac.cpp:29-69
int orc_add_fix_conjugate(orc_sys *s, int seed, double mobility, double temp, double volFH,
                          double gamma, int noise_mode) {
  FixDef f;
  f.kind = FixDef::CONJ;
  f.seed = make_unique_seed(seed);
  f.mobility = mobility; f.temp = temp; f.volFH = volFH; f.gamma = gamma;
  f.noise_mode = noise_mode;
  f.gen.seed(f.seed);
  s->fixes.push_back(f);
  return 0;
}
int orc_set_semi_implicit(orc_sys *s, int on) {
  s->semi = on ? 1 : 0;
  return 0;
}
int orc_set_noise(orc_sys *s, const double *noise) {
  if (!noise) { s->noisefield.clear(); return 0; }
  s->noisefield.assign(noise, noise + 2 * s->M);
  return 0;
}
int orc_add_fix_floryhuggins(orc_sys *s, double temp, double volFH, double chi) {
  FixDef f;
  f.kind = FixDef::FH;
  f.temp = temp; f.volFH = volFH; f.chi = chi;
  s->fixes.push_back(f);
  return 0;
}
int orc_add_fix_gradphi(orc_sys *s, int once) {
  FixDef f;
  f.kind = FixDef::GRADPHI;
  f.once = once;
  s->fixes.push_back(f);
  return 0;
}
int orc_add_fix_drag(orc_sys *s, int nchunk, const int *start, const int *end, double drag) {
  FixDef f;
  f.kind = FixDef::DRAG;
  f.start.assign(start, start + nchunk);
  f.end.assign(end, end + nchunk);
  f.drag = drag;
  s->fixes.push_back(f);
  return 0;
}
int orc_add_fix_chain(orc_sys *s, int nchunk, const int *start, const int *end, double r0,
                      double kbond, double kbend, double drag) {
  FixDef f;
  f.kind = FixDef::CHAIN;
  f.start.assign(start, start + nchunk);
  f.end.assign(end, end + nchunk);
  f.r0 = r0; f.kbond = kbond; f.kbend = kbend; f.drag = drag;
  s->fixes.push_back(f);
  return 0;
}
// fix atom/semiflexible/{notether,singletether,doubletether} <group> <seed>
// bondlength b zeta_para zpar zeta_perp zperp bending kb temp T over group
// chunks (one polymer each); per-polymer seeds as the reference draws them
// (src/fixatom_semiflexible.cpp:44-80): uniform_int_distribution<int> on
// mt19937(make_unique_seed(seed))
int orc_add_fix_semiflexible(orc_sys *s, int tether, int nchunk, const int *start, const int *end, int seed,
                             double bondlength, double zpar, double zperp, double bending, double temp) {
  return guard(s, [&] {
    if (tether < 0 || tether > 2) throw std::invalid_argument("tether must be 0, 1 or 2");
    if (!(bondlength > 0) || !(zpar > 0) || !(zperp > 0) || temp < 0)
      throw std::invalid_argument("bondlength, zeta_para, zeta_perp must be positive, temp non-negative");
    FixDef f;
    f.kind = FixDef::SEMIFLEX;
    f.tether = tether;
    f.start.assign(start, start + nchunk);
    f.end.assign(end, end + nchunk);
    f.bondlength = bondlength; f.zpar = zpar; f.zperp = zperp; f.bending = bending; f.brtemp = temp;
    std::mt19937 gen;
    std::uniform_int_distribution<int> integer_dist;
    gen.seed(make_unique_seed(seed));
    for (int c = 0; c < nchunk; c++) f.pseed.push_back(integer_dist(gen));
    f.R0.assign(3 * (size_t)s->nowned, 0.0);
    f.FB.assign(3 * (size_t)s->nowned, 0.0);
    f.T.assign((size_t)s->nowned, 0.0);
    s->fixes.push_back(f);
  });
}
int orc_semiflexible_seeds(orc_sys *s, int ifix, int *out) {
  return guard(s, [&] {
    int k = 0;
    for (auto &f : s->fixes)
      if (f.kind == FixDef::SEMIFLEX && k++ == ifix) {
        std::copy(f.pseed.begin(), f.pseed.end(), out);
        return;
      }
    throw std::invalid_argument("no such semiflexible fix");
  });
}
// one bead-rod timestep (first_step, then second_step with the same
// external forces) of one polymer; *iters = the corrector's iterations
int orc_beadrod_step(int tether, int N, double *R, const double *Fext, uint64_t seed, int64_t step, double b,
                     double zpar, double zperp, double bending, double temp, double dt, int itermax, int *iters) {
  return guard(nullptr, [&] {
    if (N < 1 || tether < 0 || tether > 2) throw std::invalid_argument("bad polymer");
    const BRParams p{b, zpar, zperp, bending, temp, dt, tether};
    std::vector<double> R0(3 * (size_t)N), FB(3 * (size_t)N), T((size_t)N, 0.0);
    br_first_step(p, seed, step, N, R, Fext, R0.data(), FB.data(), T.data());
    *iters = br_second_step(p, N, R, Fext, R0.data(), FB.data(), T.data(), itermax);
  });
}
int orc_philox4x64_10(const uint64_t *ctr, const uint64_t *key, uint64_t *out) {
  uint64_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox4x64_10_host(c, key[0], key[1]);
  std::copy(c, c + 4, out);
  return 0;
}
// Neighbor::setup, src/neighbor.cpp:39-68, then CommBrick::setup
int orc_neighbor(orc_sys *s, double skin) {
  return guard(s, [&] {
    s->skin = skin;
    double cutoff = 0.0;
    for (auto &p : s->pairs) cutoff = std::max(cutoff, p.maxcut);
    s->triggersq = 0.25 * skin * skin;
    if (cutoff > 0.0) s->cutneigh = skin + cutoff;
    setup_bins(s);
    create_stencils(s);
    comm_setup(s);
    s->neigh_set = true;
  });
}
// Integrate::setup, src/integrate.cpp:33-88
int orc_setup(orc_sys *s, double dt) {
  return guard(s, [&] {
    s->dt = dt;
    if (s->neigh_set) {
      pbc(s);
      borders(s);
      neighbor_build(s);
    } else {
      for (int i = 0; i < s->nowned; i++) unmap(s, &s->xs[3 * i], s->images[i], &s->uxs[3 * i]);
    }
    std::fill(s->Fs.begin(), s->Fs.end(), 0.0);
    for (auto &f : s->fixes)
      if (f.kind == FixDef::SEMIFLEX) semiflex_check_bonds(s, f);
    s->setup_done = true;
  });
}
// Integrate::run, src/integrate.cpp:90-219
int orc_run(orc_sys *s, int nsteps) {
  return guard(s, [&] {
    for (int i = 0; i < nsteps; i++) {
      s->timestep++;
      for (int ph = ORC_PH_START; ph <= ORC_PH_POSTFINAL; ph++) run_phase(s, ph);
      if (std::isnan(s->phi[0])) { s->nan_seen = true; throw std::runtime_error("NAN encountered in phi."); }
    }
  });
}
int orc_phase(orc_sys *s, int phase) {
  return guard(s, [&] {
    if (phase == ORC_PH_START) s->timestep++;
    run_phase(s, phase);
  });
}
// ComputePair::end_of_step (src/compute_pair.cpp:43-85): the forces due to
// pair style ipair alone, from a fresh compute on the current gathered
// positions and fields plus reverse_comm; the owned forces are then put
// back.  A grid-atom style's compute adds its deposit to mu again, as the
// reference's pair compute does (mu is not zeroed first).
int orc_compute_pair(orc_sys *s, int ipair, double *f) {
  return guard(s, [&] {
    if (ipair < 0 || ipair >= (int)s->pairs.size())
      throw std::invalid_argument("Cannot compute pair style, it hasn't been created!");
    const size_t no = 3 * (size_t)s->nowned;
    std::vector<double> keep(s->Fs.begin(), s->Fs.begin() + no);
    std::fill(s->Fs.begin(), s->Fs.end(), 0.0);
    const PairStyle &p = s->pairs[ipair];
    if (p.style == ORC_PAIR_LJCUT) pair_ljcut(s, p, nullptr);
    else if (p.style == ORC_PAIR_HARMONIC) pair_harmonic(s, p);
    else pair_gridatom(s, p, nullptr);
    reverse_comm(s);
    std::memcpy(f, s->Fs.data(), sizeof(double) * no);
    std::copy(keep.begin(), keep.end(), s->Fs.begin());
  });
}
// ComputeComplex::in_fourier (src/compute_complex.cpp:54-91), called between
// ORC_PH_PREFINAL and ORC_PH_FINAL: ft_phi (the start-of-step r2c of phi,
// which the reference transforms again at pre_final_integrate from the same
// phi) or ft_chempot, in the reference's ft array order: its arrays are
// built (Nx, Nz, Ny) (src/grid.cpp:214-224), so (i, j, k) = (ky, kz, kx)
int orc_compute_complex(orc_sys *s, int which, int quant, double *out) {
  return guard(s, [&] {
    const std::vector<cplx> &a = which == 0 ? s->ftphi : s->ftmu;
    if (which < 0 || which > 1 || a.size() != s->M) throw std::runtime_error("spectrum not available");
    const int hx = s->nx / 2 + 1;
    const double prefac = 1. / (s->dx() * s->dy() * s->dz());
    size_t count = 0;
    for (int i = 0; i < s->ny; i++)
      for (int j = 0; j < s->nz; j++)
        for (int k = 0; k < hx; k++) {
          const cplx z = a[((size_t)j * s->ny + i) * hx + k];
          if (quant == 2) out[count++] = std::abs(z) * prefac;
          else if (quant == 3) out[count++] = std::norm(z) * prefac * prefac;
          else if (quant == 0) out[count++] = z.real() * prefac;
          else out[count++] = z.imag() * prefac;
        }
  });
}
// ComputeGridClusters::end_of_step (src/compute_grid_clusters.cpp:66-228)
// with one rank: the reference takes its z neighbours from lefts / rights,
// the end planes of the neighbouring ranks; with the box on one rank those
// are this rank's own planes Nz-1 and 0, so z is periodic like x and y.
// Points beyond the threshold start with id i + (j*Nx + k)*Ny and the
// in-place min sweeps run until one changes nothing.
int orc_compute_grid_clusters(orc_sys *s, double threshold, int condition, double *out) {
  return guard(s, [&] {
    const int Nz = s->nz, Ny = s->ny, Nx = s->nx;
    std::vector<int> a((size_t)Nx * Ny * Nz);
    auto at = [&](int i, int j, int k) -> int & { return a[(size_t)k + ((size_t)i * Ny + j) * Nx]; };
    for (int i = 0; i < Nz; i++)
      for (int j = 0; j < Ny; j++)
        for (int k = 0; k < Nx; k++) {
          const double v = s->phi[(size_t)k + ((size_t)i * Ny + j) * Nx];
          const bool in = condition == 0 ? v < threshold : v > threshold;
          at(i, j, k) = in ? i + (j * Nx + k) * Ny : -1;
        }
    bool done = false;
    while (!done) {
      done = true;
      for (int i = 0; i < Nz; i++)
        for (int j = 0; j < Ny; j++)
          for (int k = 0; k < Nx; k++) {
            if (at(i, j, k) == -1) continue;
            int *nb[6] = {&at(i == 0 ? Nz - 1 : i - 1, j, k), &at(i == Nz - 1 ? 0 : i + 1, j, k),
                          &at(i, j == 0 ? Ny - 1 : j - 1, k), &at(i, j == Ny - 1 ? 0 : j + 1, k),
                          &at(i, j, k == 0 ? Nx - 1 : k - 1), &at(i, j, k == Nx - 1 ? 0 : k + 1)};
            for (int *c : nb) {
              if (*c == -1 || *c == at(i, j, k)) continue;
              at(i, j, k) = *c = std::min(at(i, j, k), *c);
              done = false;
            }
          }
    }
    for (size_t p = 0; p < a.size(); p++) out[p] = a[p];
  });
}
// Integrate::run_until_touching (src/integrate.cpp:221-415): full steps;
// after each, |x[nowned-1] - x[0]| (held positions) against touch_point,
// with the reference's since_last_cyc bookkeeping; returns 1 when it stopped
int orc_run_until_touching(orc_sys *s, int nsteps, double touch_point, int continue_running, int delay_step,
                           int cap, int64_t *occurred, int *n_occurred, int64_t *ended, int *n_ended,
                           int *steps_run) {
  int rc = 0, nocc = 0, nend = 0, done = 0;
  int since = delay_step;
  bool breakflag = false;
  int g = guard(s, [&] {
    if (s->nowned < 1) throw std::runtime_error("run_until_touching needs atoms");
    for (int it = 0; it < nsteps; it++) {
      s->timestep++;
      for (int ph = ORC_PH_START; ph <= ORC_PH_POSTFINAL; ph++) run_phase(s, ph);
      if (std::isnan(s->phi[0])) throw std::runtime_error("NAN encountered in phi.");
      done = it + 1;
      const double *a = &s->xs[0], *b = &s->xs[3 * (size_t)(s->nowned - 1)];
      const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) < touch_point) {
        if (since >= delay_step) { if (nocc < cap && occurred) occurred[nocc] = s->timestep; nocc++; }
        since = 0;
        breakflag = true;
      } else {
        if (breakflag && since == delay_step) { if (nend < cap && ended) ended[nend] = s->timestep - delay_step; nend++; }
        since += 1;
      }
      if (!continue_running && breakflag) { rc = 1; break; }
    }
  });
  if (n_occurred) *n_occurred = nocc;
  if (n_ended) *n_ended = nend;
  if (steps_run) *steps_run = done;
  return g < 0 ? g : rc;
}
int orc_counts(orc_sys *s, int64_t *o) {
  o[0] = s->nowned; o[1] = s->ngathered; o[2] = s->nlocal; o[3] = s->nghost;
  o[4] = s->ncalls; o[5] = s->nless; o[6] = s->ngreat; o[7] = s->list_overflow ? 1 : 0;
  return 0;
}
static int copy_list(const std::vector<int> &il, const std::vector<int> &num,
                     const std::vector<int> &off, const std::vector<int> &nb, int64_t cap,
                     int *ilist, int *numneigh, int *neigh, int64_t *inum, int64_t *nentries) {
  int64_t tot = 0;
  for (int v : num) tot += v;
  *inum = (int64_t)il.size();
  *nentries = tot;
  if (!ilist) return 0;
  if ((int64_t)il.size() > cap || tot > cap) return -1;
  int64_t w = 0;
  for (size_t i = 0; i < il.size(); i++) {
    ilist[i] = il[i];
    numneigh[i] = num[i];
    for (int q = 0; q < num[i]; q++) neigh[w++] = nb[off[i] + q];
  }
  return 0;
}
int orc_gridlist(orc_sys *s, int64_t cap, int *ilist, int *numneigh, int *neigh, int64_t *inum,
                 int64_t *nentries) {
  return copy_list(s->g_ilist, s->g_num, s->g_off, s->g_neigh, cap, ilist, numneigh, neigh, inum,
                   nentries);
}
int orc_halflist(orc_sys *s, int64_t cap, int *ilist, int *numneigh, int *neigh, int64_t *inum,
                 int64_t *nentries) {
  return copy_list(s->h_ilist, s->h_num, s->h_off, s->h_neigh, cap, ilist, numneigh, neigh, inum,
                   nentries);
}
int orc_gathered(orc_sys *s, double *xyz, int *tag, int *label, int *owner) {
  int n = s->ngathered, f = s->nowned;
  if (xyz) std::memcpy(xyz, &s->xs[3 * (size_t)f], sizeof(double) * 3 * n);
  if (tag) std::memcpy(tag, &s->tags[f], sizeof(int) * n);
  if (label) std::memcpy(label, &s->labels[f], sizeof(int) * n);
  if (owner) std::memcpy(owner, &s->owner[f], sizeof(int) * n);
  return n;
}
int orc_bins(orc_sys *s, int *info, int *binhead, int *bins, int *atom2bin) {
  if (info) {
    info[0] = s->nbinx; info[1] = s->nbiny; info[2] = s->nbinz;
    info[3] = s->mbinx; info[4] = s->mbiny; info[5] = s->mbinz;
    info[6] = s->mbinxlo; info[7] = s->mbinylo; info[8] = s->mbinzlo; info[9] = s->mbins;
    info[10] = (int)s->stencil_full.size(); info[11] = (int)s->stencil_half.size();
  }
  if (binhead) std::memcpy(binhead, s->binhead.data(), sizeof(int) * s->binhead.size());
  if (bins) std::memcpy(bins, s->bins.data(), sizeof(int) * s->bins.size());
  if (atom2bin) std::memcpy(atom2bin, s->atom2bin.data(), sizeof(int) * s->atom2bin.size());
  return 0;
}
// Energies defined by this build (the reference computes none, finding F11):
// out[0] Flory-Huggins + square-gradient free energy, out[1] grid-atom
// coupling, out[2] lj/cut, out[3] chain bond + bending.
int orc_energy(orc_sys *s, double *out) {
  return guard(s, [&] {
    const double dv = s->dx() * s->dy() * s->dz();
    double efh = 0;
    for (auto &f : s->fixes) {
      if (f.kind == FixDef::FH) {
        double e = 0;
        for (size_t i = 0; i < s->N; i++) {
          double p = s->phi[i];
          if (p <= 0) p = -p; else if (p >= 1) p = 1 - (p - 1);
          e += p * log(p) + (1 - p) * log(1 - p) + f.chi * p * (1 - p);
        }
        efh += f.temp / f.volFH * e * dv;
      }
      if (f.kind == FixDef::CONJ) {
        double e = 0;
        for (size_t i = 0; i < s->N; i++)
          e += s->grad[0][i] * s->grad[0][i] + s->grad[1][i] * s->grad[1][i] + s->grad[2][i] * s->grad[2][i];
        efh += 0.5 * f.temp / f.volFH * f.gamma * e * dv;
      }
    }
    double ega = 0, elj = 0;
    std::vector<double> mu_save = s->mu, f_save = s->Fs;
    for (auto &p : s->pairs) {
      if (p.style == ORC_PAIR_LJCUT) pair_ljcut(s, p, &elj);
      else if (p.style == ORC_PAIR_HARMONIC) continue;
      else pair_gridatom(s, p, &ega);
    }
    s->mu.swap(mu_save);
    s->Fs.swap(f_save);
    double ech = 0;
    for (auto &f : s->fixes)
      if (f.kind == FixDef::CHAIN)
        for (size_t c = 0; c < f.start.size(); c++)
          ech += chain_energy(&s->uxs[3 * (size_t)f.start[c]], f.end[c] - f.start[c], f.kbond, f.r0, f.kbend);
      else if (f.kind == FixDef::SEMIFLEX)
        for (size_t c = 0; c < f.start.size(); c++)
          if (f.end[c] > f.start[c])
            ech += br_energy(&s->uxs[3 * (size_t)f.start[c]], f.end[c] - f.start[c], f.bending);
    out[0] = efh; out[1] = ega; out[2] = elj; out[3] = ech;
  });
}

}  // extern "C"
