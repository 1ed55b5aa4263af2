/*
 * phafd_hip.h -- C ABI of the MI355X (gfx950) implementation of PHAFD's
 * per-timestep hot path: Model-B pseudo-spectral concentration update,
 * grid-atom coupling with its binned neighbour lists, bead-bead LJ and
 * semiflexible-chain forces, for one GPU holding the whole periodic box.
 *
 * Plain C: pointers, sizes, int status codes.  No exception, no C++ or torch
 * type crosses this boundary.  Every entry point returns 0 (or a
 * non-negative id) on success and a negative PHAFD_E* code on failure; the
 * message is available from phafd_hip_last_error().
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the reference source tree).  The reference runs these as C++ virtual
 * hooks on one MPI rank (src/fix.hpp:16-29, src/pair.hpp:36-42); here one
 * context owns all device state and runs the same hook sequence on one HIP
 * stream.
 *
 * Host layouts at this boundary:
 *   real field : double[Nz][Ny][Nx], x fastest (flat = (z*Ny+y)*Nx+x)
 *   spectrum   : double[Nz][Ny][Nx/2+1][2] (re, im), kx fastest
 *   atoms      : double[n][3] positions / forces, int[n] types, tags, ...
 * Threading: one context per host thread; calls on a context are not
 * thread-safe.  get_* calls synchronise the context stream.
 */
#ifndef PHAFD_HIP_H
#define PHAFD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phafd_hip_ctx phafd_hip_ctx;

/* error codes */
enum {
  PHAFD_OK = 0,
  PHAFD_E_ARG = -1,      /* invalid argument / state (std::invalid_argument)  */
  PHAFD_E_HIP = -2,      /* HIP runtime error                                 */
  PHAFD_E_NAN = -3,      /* "NAN encountered in phi." (src/integrate.cpp:190-201) */
  PHAFD_E_RUNTIME = -4,  /* runtime error (std::runtime_error)                */
  PHAFD_E_OVERFLOW = -5  /* neighbour row longer than oneatom = 2000; the
                            reference only warns (src/npair_grid_bin.cpp:119-121) */
};

/* fields (src/grid.hpp:28-36) */
enum { PHAFD_PHI = 0, PHAFD_MU = 1, PHAFD_GRADX = 2, PHAFD_GRADY = 3, PHAFD_GRADZ = 4 };
enum { PHAFD_PHIHAT = 0, PHAFD_MUHAT = 1 };
/* ghost images along z on one GPU (single-rank CommBrick keeps none) */
enum { PHAFD_ZGHOST_REFERENCE = 0, PHAFD_ZGHOST_PERIODIC = 1 };
/* thermal noise of grid/conjugate/volfrac: off, generated on the device
 * (PHAFD_NOISE_DEVICE, see phafd_hip_noise_field), or one host field per
 * update (PHAFD_NOISE_FIELD, see phafd_hip_set_noise) */
enum { PHAFD_NOISE_OFF = 0, PHAFD_NOISE_DEVICE = 1, PHAFD_NOISE_FIELD = 2 };
/* grid-atom pair styles */
/* grid-atom pair styles of phafd_hip_pair_gridatom (lj/cut has its own call) */
enum { PHAFD_PAIR_GAUSSIAN = 0, PHAFD_PAIR_LJISH = 1, PHAFD_PAIR_LJISH_LINEAR = 3 };
enum { PHAFD_FP64 = 0, PHAFD_FP32 = 1 };
/* hook phases of one timestep (src/integrate.cpp:105-203) */
enum {
  PHAFD_PH_START = 0, PHAFD_PH_NEIGH = 1, PHAFD_PH_PAIR = 2, PHAFD_PH_INITIAL = 3,
  PHAFD_PH_POSTFORCE = 4, PHAFD_PH_PREFINAL = 5, PHAFD_PH_FINAL = 6, PHAFD_PH_POSTFINAL = 7
};

typedef struct {
  int nx, ny, nz;          /* grid_style Nx Ny Nz (src/grid.cpp:88-113); powers of two, 4..1024 */
  double boxlo[3];         /* domain boxorigin (src/domain.cpp:50-56) */
  double period[3];        /* domain boxdims   (src/domain.cpp:44-49) */
  int zghost;              /* PHAFD_ZGHOST_* */
  int precision;           /* PHAFD_FP64 (the drop-in) or PHAFD_FP32 */
  int device;              /* HIP device ordinal */
} phafd_hip_desc;

/* ---- lifecycle: PHAFD::create + Domain::set_box + Grid::create
 *      (src/phafd.cpp:108-121, src/domain.cpp:30-99, src/grid.cpp:88-113,205-286) */
int phafd_hip_create(const phafd_hip_desc *desc, phafd_hip_ctx **out);
int phafd_hip_destroy(phafd_hip_ctx *ctx);
const char *phafd_hip_last_error(const phafd_hip_ctx *ctx);

/* ---- fields: Grid public arrays phi, chempot, gradphi (src/grid.hpp:28-36) */
int phafd_hip_set_field(phafd_hip_ctx *ctx, int which, const double *host);
int phafd_hip_get_field(phafd_hip_ctx *ctx, int which, double *host);
/* the same with rows ld doubles apart (ld >= Nx): row (z, y) of Nx values at
 * host + (z*Ny + y)*ld.  The reference's real fftw arrays pad each x row to
 * 2*(Nx/2+1) doubles ("real fftw arrays aren't contiguous", src/dump.cpp:
 * 1181-1187, src/read_dump.cpp:344-350), so Grid::phi->data() binds with
 * ld = 2*(Nx/2+1); padding entries are neither read nor written. */
int phafd_hip_set_field_ld(phafd_hip_ctx *ctx, int which, const double *host, int64_t ld);
int phafd_hip_get_field_ld(phafd_hip_ctx *ctx, int which, double *host, int64_t ld);
/* ft_phi / ft_chempot (src/grid.hpp:34-35), returned in the canonical layout */
int phafd_hip_get_spectrum(phafd_hip_ctx *ctx, int which, double *host);

/* ---- atoms: Atom::setup + ReadAtoms + check_tags_and_types
 *      (src/atom.cpp:26-110,140-252, src/read_atoms.cpp:58-97) */
int phafd_hip_set_atoms(phafd_hip_ctx *ctx, int n, const double *xyz, const int *type,
                        const int *tag, const int *mol);
/* Atom::images of the owned atoms as read by read_dump (src/read_dump.cpp:
 * 240-244), in Domain's packed image bits; after phafd_hip_set_atoms, which
 * sets every image to zero (src/atom.cpp:245), and before phafd_hip_setup */
int phafd_hip_set_images(phafd_hip_ctx *ctx, const int *image);
/* Atom::xs, uxs, images (src/atom.hpp:38-47); any pointer may be NULL */
int phafd_hip_get_atoms(phafd_hip_ctx *ctx, double *xyz, double *uxyz, int *image);
/* Atom::Fs of owned atoms after reverse_comm (src/comm_brick.cpp:228-258) */
int phafd_hip_get_forces(phafd_hip_ctx *ctx, double *f);

/* compute <id> pair <pairstyle>: ComputePair::end_of_step (src/compute_pair.cpp:
 * 43-85).  f[nowned][3] receives the forces due to pair style ipair (the index
 * its creation returned) alone, computed afresh on the current gathered
 * positions and fields, as at the end of a step; the owned forces are left as
 * they were.  Like the reference's pair compute, a grid-atom style adds its
 * chempot deposit to mu once more (mu is zeroed again by the next step). */
int phafd_hip_compute_pair(phafd_hip_ctx *ctx, int ipair, double *f);

/* ---- Fourier-space computes.
 * phafd_hip_fourier_capture(on): the last step of each later phafd_hip_run
 * keeps ft_phi and ft_chempot as they are at the in_fourier hook
 * (src/integrate.cpp:156-160: after their r2c, before the update).
 * compute <id> complex ft_phi|ft_chempot real|imag|modulus|norm:
 * ComputeComplex::in_fourier (src/compute_complex.cpp:54-91) of the kept
 * spectrum, scaled by 1/(dx dy dz) (squared for norm); out holds
 * Ny*Nz*(Nx/2+1) values in the reference's ft array order [ky][kz][kx].
 * fix <id> grid/ave <every> <repeat> <freq> c_<id>: FixGridAve
 * (src/fixgrid_ave.cpp:92-215) of such a compute, accumulated on the device
 * in one of 8 slots: first starts the window from zero, last divides by
 * repeat; out (may be NULL) receives the slot after the call. */
enum { PHAFD_CPX_REAL = 0, PHAFD_CPX_IMAG = 1, PHAFD_CPX_MODULUS = 2, PHAFD_CPX_NORM = 3 };
int phafd_hip_fourier_capture(phafd_hip_ctx *ctx, int on);
int phafd_hip_compute_complex(phafd_hip_ctx *ctx, int which, int quant, double *out);
int phafd_hip_grid_ave(phafd_hip_ctx *ctx, int slot, int which, int quant, int first, int last, int repeat,
                       double *out);

/* compute <id> grid/clusters phi <threshold> lt|gt: ComputeGridClusters::
 * end_of_step (src/compute_grid_clusters.cpp:66-228) on one subdomain, the
 * box periodic in x, y and z: out[Nz][Ny][Nx] = the smallest id
 * i + (j*Nx + k)*Ny ((i, j, k) = (z, y, x)) of the point's 6-connected
 * cluster of points with phi < threshold (lt) or > threshold (gt), -1
 * elsewhere */
enum { PHAFD_CLUSTER_LT = 0, PHAFD_CLUSTER_GT = 1 };
int phafd_hip_compute_grid_clusters(phafd_hip_ctx *ctx, double threshold, int condition, double *out);

/* run_until_touching <nsteps> <touch_point> [continue] [delay]:
 * Integrate::run_until_touching (src/integrate.cpp:221-415).  After every
 * step the distance between the last and the first owned atom (positions as
 * held, no minimum image) is compared with touch_point.  occurred[] receives
 * the steps the reference reports as "cyclisation occured" (a touch at least
 * delay_step steps after the previous one), ended[] the steps of "cyclisation
 * interval ended"; at most cap of each are stored, the counts are complete.
 * Without continue_running the run stops after the first touching step and
 * the call returns 1 (0 when the steps ran out); steps_run counts the steps
 * done.  The reference's extra computes and dumps at the stop are the
 * caller's. */
int phafd_hip_run_until_touching(phafd_hip_ctx *ctx, int nsteps, double touch_point, int continue_running,
                                 int delay_step, int cap, int64_t *occurred, int *n_occurred, int64_t *ended,
                                 int *n_ended, int *steps_run);

/* ---- dumps and restart (formats and attribute names of include/phafd_host.h).
 * dump <name> grid <file> ... attributes n phi chempot gradphi_x ...:
 * Dump::write_grid_timestep (src/dump.cpp:681-780) of the current fields.
 * dump <name> atom/owned <file> ... attributes n x ux ix F ID molID type
 * [decimalplaces p] [fmt lf|e]: Dump::write_atom_timestep (src/dump.cpp:
 * 360-418).  The arrays are snapshotted in device memory on the step stream,
 * so with async != 0 the call returns at once and later steps may run while
 * the copy to the host and the file write proceed (a grid and an atom dump
 * may be in flight together); the next dump of the same kind, a read_dump
 * or phafd_hip_dump_wait() waits for it and reports its error. */
int phafd_hip_dump_grid(phafd_hip_ctx *ctx, const char *path, int nattr, const char *const *names, int async);
int phafd_hip_dump_atoms(phafd_hip_ctx *ctx, const char *path, int nattr, const char *const *names, int precision,
                         int fmt, int async);
int phafd_hip_dump_wait(phafd_hip_ctx *ctx);
/* grid_populate read_dump grid <file> attributes n phi ... [no_padding]:
 * ReadDump::process_grid_attributes (src/read_dump.cpp:90-185) into the
 * context's fields (phi must be among the attributes) */
int phafd_hip_read_dump_grid(phafd_hip_ctx *ctx, const char *path, int nattr, const char *const *names,
                             int no_padding);

/* ---- pair styles, in script order (src/input.cpp:99-153).
 * pair gridatom/gaussian|gridatom/LJish|gridatom/LJish/linear <cut>:
 * Pair::settings (src/pair_gridatom_gaussian.cpp:123-140,
 * src/pair_gridatom_lj_ish.cpp:184-204, src/pair_gridatom_lj_ish_linear.cpp:180-197).
 * gridatom/LJish/linear deposits mu += e and pulls atoms with grad phi e dv,
 * e = -eps exp(1 - 1/(1 - r^2/w^2)) (src/pair_gridatom_lj_ish_linear.cpp:46-179);
 * its phi coefficient is unused.  Returns the pair id (>= 0). */
int phafd_hip_pair_gridatom(phafd_hip_ctx *ctx, int style, double cut);
/* pairCoeffs <type> epsilon e phi p nucwidth w (src/pair_gridatom_gaussian.cpp:143-181);
 * pass NaN to leave a value at its default */
int phafd_hip_pair_coeff(phafd_hip_ctx *ctx, int pair, int type, double epsilon, double phi,
                         double nucwidth);
/* pair lj/cut <cut> <epsilon> <sigma> (src/pair_lj_cut.cpp:101-117) */
int phafd_hip_pair_ljcut(phafd_hip_ctx *ctx, double cut, double epsilon, double sigma);
/* pair harmonic/cut <cut> <springK> (src/pair_harmonic_cut.cpp:111-124): bead pairs
 * closer than cut are pushed apart with fpair = -springK (r - cut) / r
 * (src/pair_harmonic_cut.cpp:45-108), over the same half list as lj/cut;
 * no energy.  Returns the pair id (>= 0). */
int phafd_hip_pair_harmoniccut(phafd_hip_ctx *ctx, double cut, double springK);

/* ---- fixes, in script order (src/input.cpp:175-230) */
/* fix grid/conjugate/volfrac <name> <seed> mobility M temp T volFH v gamma g
 * (src/fixgrid_conjugate.cpp:21-45, src/conjugate_volfrac.cpp:29-69).
 * noise_mode: PHAFD_NOISE_OFF (new switch: the reference always adds noise),
 * PHAFD_NOISE_DEVICE (fresh noise at every update, drawn on the device:
 * phafd_hip_noise_field(..., make_unique_seed(seed), step, ...) is exactly
 * what the update of timestep `step` adds) or PHAFD_NOISE_FIELD (one host
 * field per update, see phafd_hip_set_noise). */
int phafd_hip_fix_conjugate(phafd_hip_ctx *ctx, int seed, double mobility, double temp,
                            double volFH, double gamma, int noise_mode);
/* PHAFD_NOISE_FIELD: queue a host noise field xi (canonical spectrum layout:
 * the value ConjugateVolFrac adds per mode, src/conjugate_volfrac.cpp:
 * 95-96,118, e.g. from phafd_noise_next in the reference's mt19937 order).
 * Each spectral update consumes one queued field, oldest first, so a
 * phafd_hip_run(n) needs n fields queued (else PHAFD_E_ARG, nothing run).
 * NULL clears the queue. */
int phafd_hip_set_noise(phafd_hip_ctx *ctx, const double *xi);
/* number of queued, unconsumed noise fields */
int phafd_hip_noise_queued(phafd_hip_ctx *ctx, int64_t *n);
/* fix grid/floryhuggins <name> temp T volFH v chi c (src/fixgrid_floryhuggins.cpp:22-53) */
int phafd_hip_fix_floryhuggins(phafd_hip_ctx *ctx, double temp, double volFH, double chi);
/* fix grid/gradphi <name> once|every (src/fixgrid_gradphi.cpp:21-33) */
int phafd_hip_fix_gradphi(phafd_hip_ctx *ctx, int once);
/* fix atom/drag <name> <group> <drag> over group chunks [start,end)
 * (src/fixatom_drag.cpp:19-66, src/group.cpp:31-207) */
int phafd_hip_fix_drag(phafd_hip_ctx *ctx, int nchunk, const int *start, const int *end,
                       double drag);
/* fix atom/semiflexible/harmonic: replaces the bead-rod dynamics of
 * atom/semiflexible/{notether,singletether,doubletether}
 * (src/fixatom_semiflexible.cpp:27-191, which delegate to
 * the absent beadrodpmer library) with harmonic bonds k/2 (r-r0)^2 and
 * bending kb (1 - t_a.t_b) on unwrapped chain coordinates, integrated
 * overdamped with the same two half steps as atom/drag. */
int phafd_hip_fix_chain(phafd_hip_ctx *ctx, int nchunk, const int *start, const int *end,
                        double r0, double kbond, double kbend, double drag);
/* fix atom/semiflexible/{notether,singletether,doubletether} <group> <seed>
 * bondlength b zeta_para zpar zeta_perp zperp bending kb temp T
 * (src/fixatom_semiflexible.cpp:27-191; tether 0, 1, 2), one polymer per
 * group chunk, per-polymer seeds drawn as the reference draws them
 * (uniform_int_distribution<int> on mt19937(make_unique_seed(seed))).  The
 * reference delegates the dynamics to beadrodpmer, which is not available;
 * this is the published bead-rod Brownian dynamics of Montesi, Morse &
 * Pasquali (J. Chem. Phys. 122, 084903, 2005): rigid rods, bending
 * kb (1 - u_j.u_{j+1}), anisotropic bead friction about the bead tangent,
 * metric pseudo-potential (kT/2) ln det(G H G^T), midpoint predictor
 * (initial_integrate) and Newton corrector of the rod tensions
 * (final_integrate, at most 50 iterations; past that the step returns
 * PHAFD_E_RUNTIME "Failed to maintain polymer bond-length constraints.").
 * Tethered beads (bead 0; bead 0 and N-1) do not move.  Brownian draws:
 * Philox4x64-10 keyed by the polymer seed (uniform, unit variance).
 * phafd_hip_setup checks every bond against b (|delta| > 1e-3 is the
 * reference's error).  At most 700 beads per polymer; plain runs only (not
 * batched members or the surrogate). */
int phafd_hip_fix_semiflexible(phafd_hip_ctx *ctx, int tether, int nchunk, const int *start, const int *end,
                               int seed, double bondlength, double zeta_para, double zeta_perp, double bending,
                               double temp);

/* ---- neighbor <skin>: Neighbor::setup + CommBrick::setup
 *      (src/neighbor.cpp:39-68, src/comm_brick.cpp:96-187) */
int phafd_hip_neighbor(phafd_hip_ctx *ctx, double skin);

/* ---- run: Integrate::setup / Integrate::run (src/integrate.cpp:33-219) */
int phafd_hip_setup(phafd_hip_ctx *ctx, double dt);
int phafd_hip_run(phafd_hip_ctx *ctx, int nsteps);
/* one hook phase of a timestep (PHAFD_PH_*); PHAFD_PH_START advances the step */
int phafd_hip_phase(phafd_hip_ctx *ctx, int phase);

/* ---- inspection (for parity checks and diagnostics) */
/* {nowned, ngathered, nlocal, nghost, neighbor builds, num_less_zero,
 *  num_great_zero, list overflow} (src/atom.hpp:27-31, src/neighbor.cpp:105,
 *  src/fixgrid_floryhuggins.cpp:74-80) */
int phafd_hip_counts(phafd_hip_ctx *ctx, int64_t *out8);
/* NPairGridBin list as ilist / numneigh / flattened neighbours
 * (src/npair_grid_bin.cpp:22-127); call with ilist == NULL to query sizes.
 * Building it prints the reference's "Neighbor list overflow" warning to
 * stderr when a row holds more than 2000 atoms (src/npair_grid_bin.cpp:
 * 118-121), as the half list's build does for its rows */
int phafd_hip_gridlist(phafd_hip_ctx *ctx, int64_t cap, int *ilist, int *numneigh,
                       int *neigh, int64_t *inum, int64_t *nentries);
/* NPairHalfBinNewton list (src/npair_half_bin_newton.cpp:38-124) */
int phafd_hip_halflist(phafd_hip_ctx *ctx, int64_t cap, int *ilist, int *numneigh,
                       int *neigh, int64_t *inum, int64_t *nentries);
/* gathered copies / ghost images built by CommBrick::borders
 * (src/comm_brick.cpp:263-426); returns ngathered */
int phafd_hip_gathered(phafd_hip_ctx *ctx, double *xyz, int *tag, int *label, int *owner);
/* NBinStandard geometry and linked lists binhead / bins / atom2bin
 * (src/nbin_standard.cpp:53-229); info[12] = nbin xyz, mbin xyz, mbinlo xyz,
 * mbins, full stencil size, half stencil size */
int phafd_hip_bins(phafd_hip_ctx *ctx, int *info, int *binhead, int *bins, int *atom2bin);
/* energies defined by this implementation (the reference computes none):
 * {Flory-Huggins + square-gradient, grid-atom coupling, lj/cut, chain} */
int phafd_hip_energy(phafd_hip_ctx *ctx, double *out4);
/* per-hook GPU time in ms accumulated over run() calls (hipEvent timers):
 * {gradphi, neighbor, pair, integrate, floryhuggins+r2c mu, update, c2r phi, total};
 * reset when reset != 0 */
int phafd_hip_timers(phafd_hip_ctx *ctx, double *ms8, int reset);
/* run-time options (defaults in brackets):
 *  PHAFD_OPT_EAGER_GRIDLIST [0]: store the NPairGridBin list at every build
 *    and walk it for the mu deposit.  By default the step walks the same bins
 *    in the same order without storing the list (bit-identical mu) and the
 *    list is materialised from the last build only when phafd_hip_gridlist()
 *    asks for it.
 *  PHAFD_OPT_FUSE_FH [1]: apply Flory-Huggins inside the load of the r2c of
 *    mu when the two hooks are adjacent (same arithmetic per point).
 *  PHAFD_OPT_FUSE_SPECTRAL [1]: fold the k-space gradient multiply and the
 *    spectral update into the z pass of the FFT (forward z, pointwise map,
 *    inverse z in one kernel) where the grid allows it (Nz in {16, 32, 64,
 *    128, 256, 512}).
 *    Same per-mode arithmetic; the inverse z transform factors differently,
 *    so results agree with the unfused path to rounding, not bitwise.  With
 *    it on, phafd_hip_run() leaves muhat complete and phihat consumed by the
 *    c2r, as without it.
 *  PHAFD_OPT_EAGER_HALFLIST [0]: build the NPairHalfBinNewton list with the
 *    full neighbour cut at every build and walk it for lj/cut.  By default the
 *    step walks a pruned copy (pairs closer than lj cut + skin at the build, in
 *    the same relative order: bit-identical forces) and the full list is
 *    materialised from the last build when phafd_hip_halflist() asks.
 *  PHAFD_OPT_ATOM_WINDOWS [1]: phafd_hip_run_surrogate runs the atom steps
 *    of a block in speculative windows (one host read-back per window; the
 *    steps after a rebuild decision are skipped on the device and redone
 *    after the rebuild): the same steps, lists and sums as 0, which reads
 *    the decision back after every step.
 *  PHAFD_OPT_SEMI_IMPLICIT [0]: the k-space update takes the linear
 *    gradient term implicitly, phihat <- (phihat - c muhat) / (1 + c s) / N
 *    + xi with c = M q^2 dt, s = T/v gamma q^2 (the semi-implicit form kept
 *    in a comment at src/conjugate.cpp:356); 0 is the reference's explicit
 *    update (src/conjugate_volfrac.cpp:89-133), the drop-in.
 *  PHAFD_OPT_NEXT_GRADIENT [1]: inside phafd_hip_run, a step that another
 *    step follows also makes the next step's grad phi from its updated
 *    spectrum (the next r2c of phi is N times that spectrum in exact
 *    arithmetic, src/fixgrid_conjugate.cpp:62-154), so the next
 *    start_of_step transforms nothing; results agree with 0 to rounding.
 *    The last step of a run (and every step of run_until_touching) leaves
 *    the fields as the reference does. */
enum { PHAFD_OPT_EAGER_GRIDLIST = 0, PHAFD_OPT_FUSE_FH = 1, PHAFD_OPT_FUSE_SPECTRAL = 2,
       PHAFD_OPT_EAGER_HALFLIST = 3, PHAFD_OPT_ATOM_WINDOWS = 4, PHAFD_OPT_SEMI_IMPLICIT = 5,
       PHAFD_OPT_NEXT_GRADIENT = 6 };
int phafd_hip_set_option(phafd_hip_ctx *ctx, int key, double value);
/* HIP-event timing of every launch of one kernel during run():
 * kernel_id >= 0 (re)starts the probe for that kernel (ids in the order
 * r2c x, y fwd, y inv, z fwd, z inv, c2r x, grad multiply, update, FH,
 * grid-atom deposit, grid-atom force, grid list count, grid list fill,
 * half list count, half list fill, lj/cut, fused z gradient, fused z
 * update, the surrogate's full-resolution 16->16 convolution, the c2r x
 * pass of phi with its x gradient); kernel_id < 0
 * stops it.  out3 = {total ms, launches, algorithmic bytes (flops for the
 * convolution)} accumulated so far; also records during
 * phafd_hip_run_surrogate. */
int phafd_hip_probe(phafd_hip_ctx *ctx, int kernel_id, double *out3);

/* ---- batched replicas (no counterpart in the reference, which runs one
 * system per MPI job).  n independent systems of one grid advanced together:
 * each member is a context configured through the calls above (fields,
 * atoms, pairs, fixes, neighbor); members need one grid/conjugate/volfrac fix
 * (shared mobility, temp, volFH, gamma; noise off) and one grid/floryhuggins
 * fix (chi may differ per member).  The grid side of a step runs as one
 * launch per pass over all members, captured in hipGraphs; the atom side per
 * member on a few streams.  Each member's results equal, bit for bit, those
 * of the same context run alone with PHAFD_OPT_NEXT_GRADIENT 0 (the batch
 * takes every step's gradient from phi).  setup() also autotunes the members per
 * batched launch and the number of member streams (choices that leave the
 * arithmetic unchanged). */
typedef struct phafd_hip_batch phafd_hip_batch;
int phafd_hip_batch_create(const phafd_hip_desc *desc, int nmembers, phafd_hip_batch **out);
/* member i (owned by the batch; do not destroy it) */
int phafd_hip_batch_member(phafd_hip_batch *batch, int i, phafd_hip_ctx **member);
int phafd_hip_batch_setup(phafd_hip_batch *batch, double dt);
int phafd_hip_batch_run(phafd_hip_batch *batch, int nsteps);
/* on = 0 launches the batched passes directly instead of replaying the graphs */
int phafd_hip_batch_set_graphs(phafd_hip_batch *batch, int on);
/* {members, members per launch, member streams, graphs on, gradient-block ms
 *  for the four launch sizes tried, pair-pass ms for the four stream counts} */
int phafd_hip_batch_info(phafd_hip_batch *batch, double *out12);
/* The members' pair phase (grid-atom force walk, mu sums, lj/cut, each
 * ghost-force fold level, the harmonic chains' two half steps) as one
 * launch per pass over every member (items: blocks of 64-256 atoms of one
 * member, taken round-robin by a persistent grid; both sizes autotuned)
 * instead of per-member
 * launches on the member streams.  mode -1 (default): the autotuner of
 * phafd_hip_batch_setup times both on the set-up state and keeps the faster;
 * 0: per-member; 1: one launch per pass.  The arithmetic is the same either
 * way (bit-identical members).  Members with other pair styles than one
 * grid-atom style and lj/cut run per-member. */
int phafd_hip_batch_set_fused(phafd_hip_batch *batch, int mode);
/* out14 = {one launch per pass in use (0/1), its set-up time in ms (best
 * shape), the best per-member time in ms, the tuned shape: atoms per item,
 * persistent grid cap, the set-up times in ms of the six shapes tried
 * (256 x {1024, 2048, 4096}, 128 x {2048, 4096}, 64 x 4096), then the
 * whole-step choice of mode -1: decided (0/1), mean step ms per-member, mean
 * step ms one launch per pass}.  In mode -1 the pair-phase timing of set-up
 * picks the mode of the first steps; steps 16-47 then run the modes in
 * blocks of 8 ordered A B B A (the same numbers either way; a steady drift
 * of the step time cancels), every step but a block's first timed whole,
 * and the faster mode stays. */
int phafd_hip_batch_fused_info(phafd_hip_batch *batch, double *out14);
int phafd_hip_batch_destroy(phafd_hip_batch *batch);
const char *phafd_hip_batch_last_error(const phafd_hip_batch *batch);

/* ---- learned surrogate (no counterpart in the reference).  A three-level
 * 3D convolutional U-Net (16/32/64 channels, 3^3 kernels, periodic; BF16
 * weights and activations, FP32 accumulation on the MFMA units) correcting a
 * spectral block integrator over K steps; trained on the GPU by the
 * library's own forward and backward kernels.  Grids: nx % 64 == 0,
 * ny % 32 == 0, nz % 4 == 0.  Network input: four channels per grid point,
 * [4][Nz][Ny][Nx] floats: (phi0 - mean phi0) / norm[0], rho0 / norm[1],
 * rho1 / norm[1], (mean phi0 - norm[2]) / norm[3]; rho0, rho1 being the bead
 * density (phafd_hip_density) at the block's start and end; output:
 * (phi after K steps - base) / norm[4], base from phafd_hip_surrogate_base.
 * Parameters: per layer the PyTorch Conv3d weight [cout][cin][k][k][k] then
 * the bias, layers e1 (4->16, k 3), e1b (16->16), e2 (16->32, at 1/2
 * resolution), b (32->64, 1/4), q2 (64->32, k 1), d2 (32->32, 1/2; input
 * nearest-upsampled q2 plus e2's output), q1 (32->16, k 1), d1 (16->16,
 * full; input upsampled q1 plus e1b's output), head (16->1, k 1); relu after
 * every k 3 layer; 2^3 average pooling down. */
typedef struct phafd_unet phafd_unet;
int phafd_unet_create(int device, unsigned long long seed, phafd_unet **out);
int phafd_unet_destroy(phafd_unet *net);
const char *phafd_unet_last_error(const phafd_unet *net);
int phafd_unet_nparams(phafd_unet *net, int64_t *n);
int phafd_unet_get_params(phafd_unet *net, float *params);
int phafd_unet_set_params(phafd_unet *net, const float *params);   /* also resets the optimiser */
/* norm5 = {phi scale, rho scale, mean reference, mean scale, output scale} */
int phafd_unet_set_norm(phafd_unet *net, const double *norm5);
int phafd_unet_get_norm(phafd_unet *net, double *norm5);
/* out[Nz][Ny][Nx] of one input; the input is kept for phafd_unet_backward */
int phafd_unet_forward(phafd_unet *net, int nx, int ny, int nz, const float *in4, float *out);
/* adds the gradient of sum(dout * out) with respect to the parameters to
 * the gradient buffer (read, and optionally zeroed, by phafd_unet_get_grads) */
int phafd_unet_backward(phafd_unet *net, const float *dout);
int phafd_unet_get_grads(phafd_unet *net, float *grads, int zero);
/* training set resident on the device: nsamples inputs (in4 layout) and
 * targets ([Nz][Ny][Nx]) */
int phafd_unet_data_alloc(phafd_unet *net, int nx, int ny, int nz, int nsamples);
int phafd_unet_data_put(phafd_unet *net, int i, const float *in4, const float *target);
/* nsteps Adam steps (beta 0.9, 0.999) on minibatches of `batch` samples
 * drawn from samples [0, ntrain); losses[k]: minibatch mean square error */
int phafd_unet_train(phafd_unet *net, int nsteps, int batch, int ntrain, double lr, double *losses);
int phafd_unet_eval(phafd_unet *net, int i0, int count, double *losses);

/* bead density rho[Nz][Ny][Nx] of the context's atoms: the sum over atoms of
 * the grid-atom kernel value e within its range (PairGridAtomGaussian /
 * LJish, the factor of 2 (phi - phi_t) in the mu term,
 * src/pair_gridatom_gaussian.cpp:95-101) */
int phafd_hip_density(phafd_hip_ctx *ctx, float *rho);
/* The plain update (src/conjugate_volfrac.cpp:89-133) per Fourier mode is
 * phi <- G phi - c n with c = M q^2 dt, G = 1 - c (f'' + T/v gamma q^2),
 * n = mu - f'' phi, mu = Flory-Huggins + 2 (phi - phi_t) rho,
 * f'' = T/v (1/m + 1/(1-m) - 2 chi) + 2 mean(rho0), m = mean(phi0); over K
 * steps phi_K = G^K phi_0 - c sum_k G^(K-1-k) n_k.
 * base[Nz][Ny][Nx]: that sum with n held at its block-start value n0
 * (density rho0): base = G^K phi0 - c (1 + G + ... + G^(K-1)) n0
 * (exponential Euler over the block). */
int phafd_hip_surrogate_base(phafd_hip_ctx *ctx, const double *phi0, const float *rho0, int K, double *base);
/* the check of a surrogate block: out2 = {|R| / |phi1 - phi0|, mass drift},
 * R = phi1 - G^K phi0 + c (A n0 + B n1) in Fourier space (n0 with rho0, n1
 * at phi1 with rho1, A = sum_k G^(K-1-k) (1 - k/K), B = sum_k G^(K-1-k) k/K):
 * the K-step equation with n linear across the block */
int phafd_hip_block_residual(phafd_hip_ctx *ctx, const double *phi0, const double *phi1, const float *rho0,
                             const float *rho1, int K, double *out2);
/* nsteps (a multiple of K) in blocks of K: the atoms take K steps (lj/cut and
 * chains every step, the grid-atom force of the block-start field held),
 * then phi after K steps is predicted as the base plus the network's output
 * times norm[4] (mean removed).  Accepted when the block residual <= tol and
 * the mass drift <= 1e-12; otherwise the atoms are put back and the block
 * runs as K plain steps.  stats8 = {blocks, accepted, fallbacks, max
 * residual, last residual, max |mass drift|, ms in the network path, ms in
 * total}.  FP64 contexts with one conjugate and one Flory-Huggins fix, noise
 * off, at most one grid-atom style with one phi_t. */
int phafd_hip_run_surrogate(phafd_hip_ctx *ctx, phafd_unet *net, int nsteps, int K, double tol, double *stats8);

/* ---- stateless kernels (unit-level parity; allocate internally) */
/* fftw_mpi_plan_dft_r2c_3d / _c2r_3d (src/grid.cpp:239-280), unnormalised;
 * c2r uses only Re() of the kx = 0 and kx = Nx/2 bins */
int phafd_hip_fft_r2c(int nx, int ny, int nz, const double *in, double *out);
int phafd_hip_fft_c2r(int nx, int ny, int nz, const double *in, double *out);
/* FixGridConjugate::start_of_step grad(phi) (src/fixgrid_conjugate.cpp:62-116) */
int phafd_hip_gradphi(int nx, int ny, int nz, const double *period, const double *phi,
                      double *gx, double *gy, double *gz);
/* FixGridFloryHuggins::post_force (src/fixgrid_floryhuggins.cpp:65-85) */
int phafd_hip_floryhuggins(int64_t n, double *phi, double *mu, double temp, double volFH,
                           double chi, int64_t *nless, int64_t *ngreat);
/* The thermal noise the update of timestep `step` adds in
 * PHAFD_NOISE_DEVICE mode, xi[Nz][Ny][Nx/2+1][2] (canonical layout), for
 * ConjugateVolFrac's amplitudes (src/conjugate_volfrac.cpp:72-133) and its
 * Hermitian rules (src/conjugate.cpp:163-327): a mode updated with
 * complex_update gets sqrt(12 T M / L^3) |q| sqrt(dt) (U1 + i U2), a
 * self-conjugate mode sqrt(24 T M / L^3) |q| sqrt(dt) U1, the mirror of an
 * updated mode on the kx = 0, Nx/2 planes the conjugate of its noise, the
 * origin 0.  U1, U2 = the top 53 bits of words 0 and 1 of
 * Philox4x64-10(counter {m, step, 0, 0}, key {key0, 0x5048414644}) as
 * doubles in [-1/2, 1/2), m = (kx*Nz + kz)*Ny + ky of the updated mode;
 * key0 = make_unique_seed(seed) of the fix. */
int phafd_hip_noise_field(int nx, int ny, int nz, const double *period, double mobility, double temp,
                          double dt, uint64_t key0, int64_t step, double *xi);
/* Conjugate::update + ConjugateVolFrac (src/conjugate.cpp:163-327,
 * src/conjugate_volfrac.cpp:72-144) on a canonical-layout spectrum */
int phafd_hip_spectral_update(int nx, int ny, int nz, const double *period, double *phihat,
                              const double *muhat, double mobility, double temp, double volFH,
                              double gamma, double dt, const double *noise);

/* ---- no device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The gradient of
 * phafd_hip_gradphi, computed by the same kernels, on caller-owned DEVICE
 * arrays phi, gx, gy, gz (double[Nz][Ny][Nx]), enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) and not waited for.  The
 * intermediates (two half spectra) and the FFT twiddles live in a static
 * device array of PHAFD_NOALLOC_SCRATCH_BYTES that each GPU holds until the
 * process ends; grids with (Nx/2+1)*Ny*Nz above
 * PHAFD_GRADPHI_NOALLOC_MAX_MODES (128^3 fits, 256x128x128 does not) are
 * rejected with PHAFD_E_ARG.  The first call on a GPU sets that GPU up
 * (symbol lookup, a 64 KiB twiddle copy); later calls only launch kernels.
 * All calls on one GPU share its array: calls that overlap on one GPU (other
 * streams or host threads) race.  Errors: phafd_hip_last_error(NULL). */
#define PHAFD_NOALLOC_SCRATCH_BYTES (40u * 1024u * 1024u)
#define PHAFD_GRADPHI_NOALLOC_MAX_MODES 1308672
int phafd_hip_gradphi_noalloc(int nx, int ny, int nz, const double *period, const double *phi,
                              double *gx, double *gy, double *gz, void *stream);

#ifdef __cplusplus
}
#endif
#endif
