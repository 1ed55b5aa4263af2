/*
 * oracle.h -- C API of the CPU FP64 oracle for PHAFD's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ (OpenMP over FFT
 * lines and pointwise loops only) restatement of the reference's CPU algorithm (one MPI rank holding the whole
 * box).  It is the checker that the HIP path is compared against: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load it.
 * Nothing in the product links or calls it.
 *
 * Layouts used at this boundary (identical to the product C-ABI):
 *   real fields   : double[Nz][Ny][Nx], x fastest (flat = (z*Ny+y)*Nx+x, the
 *                   order of fftw_arr reverseFlat used by the reference's
 *                   src/npair_grid_bin.cpp:71-116)
 *   spectra       : complex double[Nz][Ny][Nx/2+1] interleaved (re,im), kx
 *                   fastest (numpy rfftn layout)
 *   atoms         : double[n][3] (x,y,z per atom, like Eigen::Matrix3Xd
 *                   column-major storage in src/atom.hpp:38)
 */
#ifndef PHAFD_ORACLE_H
#define PHAFD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

enum { ORC_PHI = 0, ORC_MU = 1, ORC_GRADX = 2, ORC_GRADY = 3, ORC_GRADZ = 4 };
enum { ORC_ZGHOST_REFERENCE = 0, ORC_ZGHOST_PERIODIC = 1 };
enum { ORC_NOISE_OFF = 0, ORC_NOISE_MT19937 = 1, ORC_NOISE_FIELD = 2 };
enum { ORC_PAIR_GAUSSIAN = 0, ORC_PAIR_LJISH = 1, ORC_PAIR_LJCUT = 2, ORC_PAIR_LJISH_LINEAR = 3,
       ORC_PAIR_HARMONIC = 4 };

/* ---- stateless kernels (unit-level parity) ---- */
int orc_fft_r2c(int nx, int ny, int nz, const double *in, double *out);
int orc_fft_c2r(int nx, int ny, int nz, const double *in, double *out);
int orc_gradphi(int nx, int ny, int nz, const double period[3],
                const double *phi, double *gx, double *gy, double *gz);
int orc_floryhuggins(int64_t n, double *phi, double *mu, double temp,
                     double volFH, double chi, int64_t *nless, int64_t *ngreat);
int orc_spectral_update(int nx, int ny, int nz, const double period[3],
                        double *phihat, const double *muhat, double mobility,
                        double temp, double volFH, double gamma, double dt,
                        int noise_mode, int seed, const double *noise);
int orc_noise_field(int nx, int ny, int nz, const double period[3], int seed,
                    double mobility, double temp, double dt, double *noise);
int orc_mt19937_uniform(int seed, int n, double *out);
int orc_make_unique_seed(int seed);
/* OpenMP threads of the FFT lines and pointwise loops (results do not
 * depend on the count: every line and point is computed alone) */
/* bead-rod semiflexible polymers (tether 0 none, 1 bead 0, 2 both ends) */
int orc_add_fix_semiflexible(orc_sys *s, int tether, int nchunk, const int *start, const int *end, int seed,
                             double bondlength, double zpar, double zperp, double bending, double temp);
int orc_semiflexible_seeds(orc_sys *s, int ifix, int *out);
int orc_philox4x64_10(const uint64_t *ctr, const uint64_t *key, uint64_t *out);
int orc_beadrod_step(int tether, int N, double *R, const double *Fext, uint64_t seed, int64_t step, double b,
                     double zpar, double zperp, double bending, double temp, double dt, int itermax, int *iters);
int orc_set_threads(int n);
int orc_threads(void);
int orc_mt19937_raw(int seed, int n, uint32_t *out);

/* ---- whole-system state machine (mirrors Integrate::setup / run) ---- */
orc_sys *orc_create(int nx, int ny, int nz, const double boxlo[3],
                    const double period[3], int zghost_mode);
void orc_destroy(orc_sys *);
const char *orc_last_error(orc_sys *);
int orc_set_field(orc_sys *, int which, const double *data);
int orc_get_field(orc_sys *, int which, double *data);
int orc_set_atoms(orc_sys *, int n, const double *xyz, const int *type,
                  const int *tag, const int *mol);
int orc_get_atoms(orc_sys *, double *xyz, double *uxyz, int *image);
int orc_get_forces(orc_sys *, double *f);
int orc_add_pair(orc_sys *, int style, double a, double b, double c);
int orc_pair_coeff(orc_sys *, int ipair, int type, double eps, double phit,
                   double nucwidth);
int orc_add_fix_conjugate(orc_sys *, int seed, double mobility, double temp,
                          double volFH, double gamma, int noise_mode);
int orc_set_noise(orc_sys *, const double *noise);
int orc_set_semi_implicit(orc_sys *, int on);
int orc_add_fix_floryhuggins(orc_sys *, double temp, double volFH, double chi);
int orc_add_fix_gradphi(orc_sys *, int once);
int orc_add_fix_drag(orc_sys *, int nchunk, const int *start, const int *end,
                     double drag);
int orc_add_fix_chain(orc_sys *, int nchunk, const int *start, const int *end,
                      double r0, double kbond, double kbend, double drag);
int orc_neighbor(orc_sys *, double skin);
int orc_setup(orc_sys *, double dt);
int orc_run(orc_sys *, int nsteps);
int orc_counts(orc_sys *, int64_t *out8);
int orc_gridlist(orc_sys *, int64_t cap, int *ilist, int *numneigh, int *neigh,
                 int64_t *inum, int64_t *nentries);
int orc_halflist(orc_sys *, int64_t cap, int *ilist, int *numneigh, int *neigh,
                 int64_t *inum, int64_t *nentries);
int orc_gathered(orc_sys *, double *xyz, int *tag, int *label, int *owner);
int orc_bins(orc_sys *, int *info, int *binhead, int *bins, int *atom2bin);
int orc_energy(orc_sys *, double *out4);

/* ---- individual hook phases (for per-hook parity) ---- */
enum {
  ORC_PH_START = 0,    /* fixes start_of_step (gradphi)            */
  ORC_PH_NEIGH = 1,    /* decide -> pbc+borders+build | forward     */
  ORC_PH_PAIR = 2,     /* mu=0, F=0, pairs compute, reverse_comm    */
  ORC_PH_INITIAL = 3,  /* fixes initial_integrate                   */
  ORC_PH_POSTFORCE = 4,/* fixes post_force (Flory-Huggins)          */
  ORC_PH_PREFINAL = 5, /* r2c mu                                    */
  ORC_PH_FINAL = 6,    /* spectral update, integrator second half   */
  ORC_PH_POSTFINAL = 7 /* c2r phi                                   */
};
int orc_phase(orc_sys *, int phase);
/* ComputeComplex::in_fourier between ORC_PH_PREFINAL and ORC_PH_FINAL: which
 * 0 ft_phi, 1 ft_chempot; quant 0 real, 1 imag, 2 modulus, 3 norm; out in
 * the reference's ft order [ky][kz][kx] */
int orc_compute_complex(orc_sys *, int which, int quant, double *out);
/* ComputeGridClusters::end_of_step, one rank (periodic z): condition 0 lt,
 * 1 gt; out[Nz][Ny][Nx] cluster ids, -1 outside */
int orc_compute_grid_clusters(orc_sys *, double threshold, int condition, double *out);
/* Integrate::run_until_touching; arguments as phafd_hip_run_until_touching */
int orc_run_until_touching(orc_sys *, int nsteps, double touch_point, int continue_running, int delay_step,
                           int cap, int64_t *occurred, int *n_occurred, int64_t *ended, int *n_ended,
                           int *steps_run);
/* ComputePair::end_of_step: owned forces of pair style ipair alone */
int orc_compute_pair(orc_sys *, int ipair, double *f);

/* ---- dump writers (oracle_vtk.cpp): the reference's std::ostream output
 * of src/dump.cpp for one rank; arguments as phafd_vtk_write_* ---- */
const char *orc_vtk_last_error(void);
int orc_vtk_write_grid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                       int nattr, const char *const *names, const double *const *arrays);
int orc_vtk_write_pgrid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                        int nattr, const char *const *names, const char *source);
int orc_vtk_write_atoms(const char *path, int n, int nattr, const char *const *names, const void *const *arrays,
                        int precision, int fmt);
int orc_vtk_write_patoms(const char *path, int nattr, const char *const *names, const char *source);
int orc_vtk_collection(const char *path, int part, double time, const char *file);
int orc_lammps_write_step(const char *path, int64_t timestep, int n, const double *boxlo, const double *boxhi,
                          int nattr, const char *const *names, const void *const *arrays, int precision, int fmt);

#ifdef __cplusplus
}
#endif
#endif
