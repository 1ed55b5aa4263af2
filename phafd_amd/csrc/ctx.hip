// ctx.hip -- the phafd_hip context: device state of one whole-box subdomain,
// the per-timestep hook sequence of Integrate::run (src/integrate.cpp:105-203)
// and the C ABI declared in include/phafd_hip.h.
#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <thread>

#include "../../include/phafd_hip.h"
#include <cstdio>
#include <random>

#include "../../include/phafd_host.h"
#include "atoms.hpp"
#include "internal.hpp"
#include "unet.hpp"

using namespace phafd;

namespace {

thread_local std::string g_create_err;
const double PI_REF = 3.141592;   // src/domain.cpp:206-208
const double BIG = 1e20;          // src/comm_brick.cpp:8
const double SMALL = 1.0e-6;      // src/nbin_standard.cpp:25

struct PairDef {
  int style;   // 0 gaussian, 1 LJish, 3 LJish/linear (grid-atom); 2 lj/cut, 4 harmonic/cut (bead pairs)
  double maxcut = 0;
  std::vector<double> eps, phit, cutsq;
  DBuf<double> d_eps, d_phit, d_cutsq;
  double lj1 = 0, lj2 = 0, ljcutsq = 0;
  double hk = 0, hcut = 0;   // harmonic/cut spring constant and cut
};
// lj/cut and harmonic/cut act between beads over the half list; the other
// styles are grid-atom couplings
bool bead_pair(const PairDef &p) { return p.style == 2 || p.style == 4; }

struct FixDef {
  enum Kind { CONJ, FH, GRADPHI, DRAG, CHAIN, SEMIFLEX } kind;
  double mobility = 1, temp = 1, volFH = 1, gamma = 1, chi = 1;
  int seed = 0, noise_mode = 0, once = 0;
  unsigned long long noise_key = 0;   // device noise: make_unique_seed(seed)
  std::vector<int> start, end;
  double drag = 0, r0 = 1, kbond = 0, kbend = 0;
  int nmem = 0, nchain = 0, maxlen = 0;
  DBuf<int> members, cstart, cend;
  DBuf<int> chs, che;   // chains (group chunks) as atom ranges [chs, che)
  // atom/semiflexible/*: bead-rod polymers, one per group chunk
  int tether = 0;
  double bondlength = 1, zpar = 1, zperp = 1, bending = 0, brtemp = 0;
  DBuf<int> d_seed;              // per non-empty chain: the polymer's seed
  DBuf<double4> br_R0, br_FB;    // per owned atom: step-start position, Brownian force
  DBuf<double> br_T;             // per owned atom: its bond's tension (first step -> second step)
  unsigned long long brstep = 0; // first_step calls
};

struct SwapDev {
  int dim = 2;
  double lo = 0, hi = 0, shift = 0;
  int first = 0;   // relative gathered index of the first received entry
  int count = 0;
  DBuf<int> send;  // sources: owned index (copy swap) or relative gathered index
  // second image swap of a dimension: no atom can be sent by both swaps of
  // the dimension (lo of this swap above hi of the first), so their reverse
  // folds touch different targets and may share one launch
  bool disjoint_pair = false;
};

// a dump in flight: the fields or atom arrays snapshotted into HBM on the
// step stream, copied to pinned host memory on the dump stream, written to
// disk by the writer thread (while the steps go on when asynchronous)
// Integrate::run_until_touching state (src/integrate.cpp:221-415)
struct Touch {
  double touch_point = 0;
  bool cont = false;
  int delay = 0, since = 0;
  bool breakflag = false, stopped = false;
  int steps = 0;
  std::vector<int64_t> occurred, ended;
};

struct DumpJob {
  DBuf<double> stage;
  char *host = nullptr;
  size_t host_bytes = 0;
  hipStream_t cs = nullptr;
  hipEvent_t snap = nullptr, done = nullptr;
  std::thread th;
  std::string err;
};

}  // namespace

template <class R> struct Vec2;
template <> struct Vec2<double> { using T = double2; };
template <> struct Vec2<float> { using T = float2; };

// fields and spectra of one precision
template <class R>
struct FieldSet {
  using C = typename Vec2<R>::T;
  DBuf<R> phi, mu, grad[3];
  DBuf<C> phihat, muhat, ghat[3], canon_tmp;
  // host noise fields queued by phafd_hip_set_noise, K layout; each spectral
  // update consumes the one at noise_head
  std::vector<DBuf<C>> noise_q;
  void alloc(size_t N, size_t M) {
    phi.alloc(N); mu.alloc(N);
    for (int d = 0; d < 3; d++) { grad[d].alloc(N); ghat[d].alloc(M); }
    phihat.alloc(M); muhat.alloc(M);
  }
};

struct phafd_hip_ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  int nx = 0, ny = 0, nz = 0;
  double boxlo[3], boxhi[3], period[3], sublo[3], subhi[3];
  int zghost = 0;
  FFTPlan plan;
  size_t N = 0, M = 0;
  bool own_stream = true;    // false for a batch member (the batch's stream)
  bool fp32 = false;         // PHAFD_FP32: fields, spectra and grid-atom terms in FP32
  FieldSet<double> fd;
  FieldSet<float> fs;
  DBuf<double> mass64;       // FP32 form: sum of phi in FP64, the k = 0 mode
  size_t noise_head = 0;      // next unconsumed entry of the noise queue
  bool semi_implicit = false; // PHAFD_OPT_SEMI_IMPLICIT
  DBuf<unsigned long long> counters;   // [nless, ngreat]
  DBuf<int> flags;                     // [0] rebuild, [1] overflow
  DBuf<double> esum;                   // energy accumulators
  int *h_flags = nullptr;              // pinned, 32 ints: [2..3] phi(0,0,0), [4] scalar reads, [6..7] window,
                                       // [8..15] the step flags, [16..17] phi(0,0,0) of odd pipelined steps
                                       // stop step and flags, [8..15] flags
  // owned atoms
  int nowned = 0, ntypes = 0;
  bool atoms_set = false;
  std::vector<int> h_tag, h_type, h_mol;
  DBuf<double4> x, ux, F, xhold;
  DBuf<int> image, type;
  // gathered
  int ngathered = 0, nlocal = 0, nghost = 0, gcap = 0;
  DBuf<double4> gx, gshift, gF;
  DBuf<int> gowner, glabel;
  DBuf<int> tmp_flag, tmp_pos, tmp_cat, tmp_part;
  std::vector<SwapDev> swaps;
  double cutghost = 0;
  // neighbour
  bool neigh_set = false, need_full = false, need_half = false;
  double skin = 0, cutneigh = 0, triggersq = 0;
  int64_t ncalls = 0;
  BinGeom bg{};
  std::vector<int> stencil_full, stencil_half, stencil_lj;
  DBuf<int> d_stencil_full, d_stencil_half, d_stencil_lj, d_sfx, d_sfy, d_sfz;
  // The step's lj/cut list keeps only the half-list pairs closer than
  // (lj cut + skin) at the build (prune_sq), walking only the half-stencil bins
  // that can hold them: every pair that can come inside the lj cut before the
  // next build is kept, in the reference's relative order, so the forces are
  // the same bit for bit.  The full reference list is materialised from the
  // build snapshot when asked for (half_full) or at every build (eager).
  double prune_sq = 0;
  bool half_full = false, eager_halflist = false;
  DBuf<double4> gx_build;             // gathered positions at the last build
  bool grid_built = false;            // grid list materialised for the last build
  bool eager_gridlist = false;        // build (and use) the stored grid list in the step
  DBuf<int> a2b, bincount, binstart, bincursor, binatoms;
  DBuf<int> g_cnt, g_neigh;
  DBuf<int64_t> g_off;
  // The grid-atom pairs of the last build: every (atom, grid point) within
  // (largest grid-atom cut + skin/2) at the build, per atom in bin order
  // (ga_aoff / ga_ps.x, ascending grid index) and grouped by grid point in
  // the reference's list order, ascending q (ga_poff; ga_ps.y: the
  // atom-major entry's slot); ga_mterm: each slot's mu term of the current
  // pass (NaN out of the cut)
  DBuf<int> ga_acnt, ga_aoff, ga_amq, ga_rank, ga_pcnt, ga_poff, ga_key, ga_ent;
  DBuf<int2> ga_ps;
  DBuf<double> ga_mterm;
  DBuf<float> ga_mterm32;
  int64_t ga_nent = 0, ga_calls = -1;   // entries; the build (ncalls) they belong to
  int64_t g_nent = 0;
  DBuf<int> h_ilist, h_rowof, h_cnt, h_off, h_neigh;
  int h_nrows = 0, h_nent = 0;
  DBuf<int> t_cnt, t_off, t_cur, t_pair, t_i;
  bool lists_valid = false;
  int pending_decision = -1;   // rebuild decision prefetched at the end of the last step
  bool store_muhat = true;     // the fused z update writes muhat (false inside a run but its last step)
  // in_fourier: the last step of a run copies ft_phi (before the update) and
  // ft_chempot (K layout, FP64) for compute complex / fix grid/ave
  bool fourier_capture = false;
  bool cap_valid = false;
  DBuf<double2> cap_hat[2];
  DBuf<double> cc_out, ave[8];
  DBuf<int> cl_parent;   // compute grid/clusters union-find
  struct Touch *touch = nullptr;    // run_until_touching in progress
  double *h_ends = nullptr;         // pinned: x of the first and last owned atom (2 double4)
  DBuf<double> cl_out;
  bool atom_windows = true;    // the surrogate's atom steps in speculative windows
  // styles
  std::vector<std::unique_ptr<PairDef>> pairs;
  std::vector<std::unique_ptr<FixDef>> fixes;
  double dt = 0;
  int64_t timestep = 0, firststep = 0;
  bool setup_done = false;
  // timers
  hipEvent_t ev[9] = {};
  hipEvent_t ev_pre[2] = {};   // a start_of_step enqueued ahead of the last read-back
  hipEvent_t evp[2][10] = {};  // pipelined steps: hook events 0-8 and the flags' read-back, per step parity
  double tacc[8] = {0};
  bool fuse_fh = true;
  bool fuse_spectral = true;
  // PHAFD_OPT_NEXT_GRADIENT: a step that another step of the same run
  // follows leaves phi, grad phi and phihat as the next start_of_step makes
  // them (grad_ready), and that start_of_step then has nothing to transform
  bool next_grad = true, grad_ready = false;
  // surrogate blocks: bead densities at both ends, network input and
  // output, the predicted phi, mu, the base, FP64 partial sums, frozen
  // grid-atom forces, the atoms' block-start state
  DBuf<float> sg_rho, sg_rho1, sg_out;
  DBuf<unsigned long long> sg_acc;   // fixed-point density accumulator
  DBuf<uint16_t> sg_x0;
  DBuf<double> sg_phi, sg_mu0, sg_mu1, sg_base, sg_red, sg_sum;
  DBuf<double2> sg_ph0, sg_mh0;   // spectra of phi and mu at the block start
  DBuf<int> sg_win;               // speculative atom-step windows: [0] stop step, [1] OR of decide flags
  DBuf<double4> sg_fga, sg_x, sg_ux;
  DBuf<int> sg_img;
  DumpJob dump[2];   // [0] grid, [1] atoms: the two kinds write concurrently
  DBuf<double4> cp_keep;   // owned forces kept across a compute pair

  double dx() const { return period[0] / nx; }   // src/grid.cpp:187-201
  double dy() const { return period[1] / ny; }
  double dz() const { return period[2] / nz; }
  Box box() const {
    Box b;
    for (int d = 0; d < 3; d++) { b.lo[d] = boxlo[d]; b.hi[d] = boxhi[d]; b.period[d] = period[d]; }
    return b;
  }
  GatheredView gview() { return GatheredView{gx.p, gshift.p, gowner.p, glabel.p}; }
  template <class R> FieldSet<R> &fields();
  template <class R> GridViewT<R> gridview() {
    GridViewT<R> v;
    v.nx = nx; v.ny = ny; v.nz = nz;
    v.dx = dx(); v.dy = dy(); v.dz = dz();
    v.dv = dx() * dy() * dz();
    for (int d = 0; d < 3; d++) v.lo[d] = sublo[d];
    FieldSet<R> &f = fields<R>();
    v.phi = f.phi.p; v.mu = f.mu.p; v.gx = f.grad[0].p; v.gy = f.grad[1].p; v.gz = f.grad[2].p;
    return v;
  }
  size_t real_bytes() const { return fp32 ? sizeof(float) : sizeof(double); }
};
template <> FieldSet<double> &phafd_hip_ctx::fields<double>() { return fd; }
template <> FieldSet<float> &phafd_hip_ctx::fields<float>() { return fs; }

namespace {

template <class Fn>
int guard(phafd_hip_ctx *c, Fn fn) {
  try {
    if (c) (void)hipSetDevice(c->device);
    return fn();
  } catch (const HipError &e) {
    if (c) c->err = e.what();
    return PHAFD_E_HIP;
  } catch (const std::invalid_argument &e) {
    if (c) c->err = e.what();
    return PHAFD_E_ARG;
  } catch (const std::exception &e) {
    if (c) c->err = e.what();
    std::string w = e.what();
    if (w.find("NAN") != std::string::npos) return PHAFD_E_NAN;
    return PHAFD_E_RUNTIME;
  }
}

void sync(phafd_hip_ctx *c) { HIPCHK(hipStreamSynchronize(c->stream)); }

// two device ints read back with one synchronise
void read_int2(phafd_hip_ctx *c, const int *a, const int *b, int &va, int &vb) {
  HIPCHK(hipMemcpyAsync(c->h_flags + 4, a, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->h_flags + 5, b, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  va = c->h_flags[4];
  vb = c->h_flags[5];
}

int read_int(phafd_hip_ctx *c, const int *dptr) {
  HIPCHK(hipMemcpyAsync(c->h_flags + 4, dptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  return c->h_flags[4];
}

// ---------------------------------------------------------------- comm
// CommBrick::setup (src/comm_brick.cpp:96-187, nprocs == 1): the copy swap,
// then (periodic mode only) two z image swaps, then y and x image swaps.
void comm_setup(phafd_hip_ctx *c) {
  c->cutghost = c->cutneigh;
  c->swaps.clear();
  c->swaps.resize(1);
  SwapDev &z = c->swaps[0];
  z.dim = 2;
  const int zstart = (int)c->sublo[2], zend = (int)c->subhi[2];
  z.lo = zend - c->cutghost;
  z.hi = zstart + c->cutghost;
  std::vector<int> dims;
  if (c->zghost == PHAFD_ZGHOST_PERIODIC) dims.push_back(2);
  dims.push_back(1);
  dims.push_back(0);
  for (int dim : dims)
    for (int ineed = 0; ineed < 2; ineed++) {
      c->swaps.emplace_back();
      SwapDev &w = c->swaps.back();
      w.dim = dim;
      if (ineed == 0) { w.lo = -BIG; w.hi = c->boxlo[dim] + c->cutghost; w.shift = c->period[dim]; }
      else {
        w.lo = c->boxhi[dim] - c->cutghost;
        w.hi = BIG;
        w.shift = -c->period[dim];
        w.disjoint_pair = w.lo > c->swaps[c->swaps.size() - 2].hi;
      }
    }
}

void ensure_gathered(phafd_hip_ctx *c, int need) {
  if (need <= c->gcap) return;
  int cap = std::max(need + need / 4 + 1024, 2 * c->nowned + 1024);
  // grow preserving the first ngathered entries
  DBuf<double4> nx_, ns_;
  DBuf<int> no_, nl_;
  nx_.alloc(cap); ns_.alloc(cap); no_.alloc(cap); nl_.alloc(cap);
  if (c->ngathered > 0) {
    HIPCHK(hipMemcpyAsync(nx_.p, c->gx.p, sizeof(double4) * c->ngathered, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(ns_.p, c->gshift.p, sizeof(double4) * c->ngathered, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(no_.p, c->gowner.p, sizeof(int) * c->ngathered, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(nl_.p, c->glabel.p, sizeof(int) * c->ngathered, hipMemcpyDeviceToDevice, c->stream));
    sync(c);
  }
  std::swap(c->gx.p, nx_.p); std::swap(c->gx.n, nx_.n);
  std::swap(c->gshift.p, ns_.p); std::swap(c->gshift.n, ns_.n);
  std::swap(c->gowner.p, no_.p); std::swap(c->gowner.n, no_.n);
  std::swap(c->glabel.p, nl_.p); std::swap(c->glabel.n, nl_.n);
  c->gF.alloc(cap);
  c->gcap = cap;
}

// ordered compaction of flags[0..n): positions into tmp_pos, returns count
int compact_count(phafd_hip_ctx *c, int n) {
  scan_exclusive_i32(c->tmp_flag.p, c->tmp_pos.p, (size_t)n, c->stream);
  return read_int(c, c->tmp_pos.p + n);
}

// CommBrick::borders, src/comm_brick.cpp:263-426 (sendproc == me branches)
void borders(phafd_hip_ctx *c) {
  const int n = c->nowned;
  c->tmp_flag.alloc((size_t)std::max(n, 1) * 4 + 16);
  c->tmp_pos.alloc((size_t)std::max(n, 1) * 4 + 16);
  c->tmp_cat.alloc((size_t)std::max(n, 1) + 1);
  c->ngathered = c->nlocal = c->nghost = 0;
  ensure_gathered(c, n + 16);
  // copy swap
  SwapDev &z = c->swaps[0];
  const int zstart = (int)c->sublo[2], zend = (int)c->subhi[2];
  // nlocal (the LOCAL entries of the copy swap): per-block counts of the
  // classification, read back with the swap's count and summed here
  const int nbc = n > 0 ? copy_classify_blocks(n) : 0;
  c->tmp_part.alloc((size_t)nbc + 1);
  launch_copy_classify(n, c->x.p, (double)zstart, (double)zend, z.lo, z.hi, c->tmp_cat.p, c->tmp_flag.p,
                       c->tmp_part.p, c->stream);
  scan_exclusive_i32(c->tmp_flag.p, c->tmp_pos.p, (size_t)n, c->stream);
  std::vector<int> part((size_t)nbc + 1);
  HIPCHK(hipMemcpyAsync(part.data(), c->tmp_pos.p + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  if (nbc) HIPCHK(hipMemcpyAsync(part.data() + 1, c->tmp_part.p, sizeof(int) * nbc, hipMemcpyDeviceToHost, c->stream));
  sync(c);
  const int n0 = part[0];
  int nloc = 0;
  for (int b = 1; b <= nbc; b++) nloc += part[b];
  z.send.alloc(n0 + 1);
  launch_copy_scatter(n, c->x.p, c->type.p, c->tmp_cat.p, c->tmp_pos.p, c->period[2], z.send.p, c->gview(), c->stream);
  z.first = 0;
  z.count = n0;
  c->ngathered = n0;
  c->nlocal = nloc;
  // image swaps, a dimension at a time: both swaps of a dimension scan the
  // same entries [0, nscan), so their flags and counts come back together
  size_t iswap = 1;
  while (iswap < c->swaps.size()) {
    const int nscan = c->ngathered;
    const size_t stride = (size_t)nscan + 16;
    c->tmp_flag.alloc(2 * stride);
    c->tmp_pos.alloc(2 * stride);
    SwapDev *w[2] = {&c->swaps[iswap], &c->swaps[iswap + 1]};
    int cnt[2] = {0, 0};
    for (int k = 0; k < 2; k++) {
      launch_img_flag(nscan, c->gx.p, w[k]->dim, w[k]->lo, w[k]->hi, c->tmp_flag.p + k * stride, c->stream);
      scan_exclusive_i32(c->tmp_flag.p + k * stride, c->tmp_pos.p + k * stride, (size_t)nscan, c->stream);
    }
    read_int2(c, c->tmp_pos.p + nscan, c->tmp_pos.p + stride + nscan, cnt[0], cnt[1]);
    ensure_gathered(c, c->ngathered + cnt[0] + cnt[1] + 16);
    for (int k = 0; k < 2; k++) {
      w[k]->send.alloc(cnt[k] + 1);
      w[k]->first = c->ngathered;
      w[k]->count = cnt[k];
      launch_img_scatter(nscan, c->tmp_flag.p + k * stride, c->tmp_pos.p + k * stride, w[k]->dim, w[k]->shift,
                         w[k]->first, w[k]->send.p, c->gview(), c->stream);
      c->ngathered += cnt[k];
    }
    iswap += 2;
  }
  c->nghost = c->ngathered - c->nlocal;
}

void forward_comm(phafd_hip_ctx *c, StepGate gate = StepGate()) {
  launch_forward(c->ngathered, c->x.p, c->type.p, c->gview(), c->stream, gate);
}

// reverse_comm with the owned forces starting from init (instead of their
// current values): when the copy swap sends every owned atom once, the copy
// and the last fold are one pass; the sums are those of copy + reverse_comm
void reverse_comm_from(phafd_hip_ctx *c, const double4 *init, StepGate gate = StepGate());

void reverse_comm(phafd_hip_ctx *c) {
  for (int iswap = (int)c->swaps.size() - 1; iswap >= 1; iswap--) {
    SwapDev &w = c->swaps[iswap];
    if (w.disjoint_pair) {   // the dimension's two swaps in one launch (same sums)
      SwapDev &v = c->swaps[iswap - 1];
      launch_reverse2(w.count, w.send.p, w.first, v.count, v.send.p, v.first, c->gF.p, c->stream);
      iswap--;
      continue;
    }
    launch_reverse(w.count, w.send.p, w.first, c->gF.p, c->gF.p, c->stream);
  }
  SwapDev &z = c->swaps[0];
  launch_reverse(z.count, z.send.p, z.first, c->F.p, c->gF.p, c->stream);
}

bool reverse_from_one_pass(const phafd_hip_ctx *c) { return c->swaps[0].count == c->nowned; }
void reverse_comm_from(phafd_hip_ctx *c, const double4 *init, StepGate gate) {
  SwapDev &z = c->swaps[0];
  if (!reverse_from_one_pass(c)) {
    if (gate.stop) throw std::logic_error("gated fold needs the one-pass copy swap");
    launch_copy4(c->nowned, init, c->F.p, c->stream);
    reverse_comm(c);
    return;
  }
  for (int iswap = (int)c->swaps.size() - 1; iswap >= 1; iswap--) {
    SwapDev &w = c->swaps[iswap];
    if (w.disjoint_pair) {
      SwapDev &v = c->swaps[iswap - 1];
      launch_reverse2(w.count, w.send.p, w.first, v.count, v.send.p, v.first, c->gF.p, c->stream, gate);
      iswap--;
      continue;
    }
    launch_reverse(w.count, w.send.p, w.first, c->gF.p, c->gF.p, c->stream, gate);
  }
  launch_reverse_init(z.count, z.send.p, z.first, c->F.p, init, c->gF.p, c->stream, gate);
}

// ------------------------------------------------------------ neighbour
// NBinStandard::setup_bins, src/nbin_standard.cpp:53-175
void setup_bins(phafd_hip_ctx *c) {
  BinGeom &b = c->bg;
  double cutghost = c->cutneigh;
  double bsubboxlo[3], bsubboxhi[3], bbox[3];
  for (int d = 0; d < 3; d++) {
    bsubboxlo[d] = c->sublo[d] - cutghost;
    bsubboxhi[d] = c->subhi[d] + cutghost;
    bbox[d] = c->boxhi[d] - c->boxlo[d];
    b.boxlo[d] = c->boxlo[d];
    b.boxhi[d] = c->boxhi[d];
  }
  double binsize_optimal = 0.5 * c->cutneigh;
  if (binsize_optimal == 0.0) binsize_optimal = bbox[0];
  double binsizeinv = 1.0 / binsize_optimal;
  b.nbinx = (int)(bbox[0] * binsizeinv);
  b.nbiny = (int)(bbox[1] * binsizeinv);
  b.nbinz = (int)(bbox[2] * binsizeinv);
  if (b.nbinx == 0) b.nbinx = 1;
  if (b.nbiny == 0) b.nbiny = 1;
  if (b.nbinz == 0) b.nbinz = 1;
  b.binsizex = bbox[0] / b.nbinx;
  b.binsizey = bbox[1] / b.nbiny;
  b.binsizez = bbox[2] / b.nbinz;
  b.bininvx = 1.0 / b.binsizex;
  b.bininvy = 1.0 / b.binsizey;
  b.bininvz = 1.0 / b.binsizez;
  if (binsize_optimal * b.bininvx > 100 || binsize_optimal * b.bininvy > 100 ||
      binsize_optimal * b.bininvz > 100)
    throw std::runtime_error("Cannot use neighbor bins - box size << cutoff");
  int hi[3];
  int *lo[3] = {&b.mbinxlo, &b.mbinylo, &b.mbinzlo};
  const double inv[3] = {b.bininvx, b.bininvy, b.bininvz};
  for (int d = 0; d < 3; d++) {
    double coord = bsubboxlo[d] - SMALL * bbox[d];
    *lo[d] = (int)((coord - c->boxlo[d]) * inv[d]);
    if (coord < c->boxlo[d]) *lo[d] = *lo[d] - 1;
    coord = bsubboxhi[d] + SMALL * bbox[d];
    hi[d] = (int)((coord - c->boxlo[d]) * inv[d]);
    *lo[d] -= 1;
    hi[d] += 1;
  }
  b.mbinx = hi[0] - b.mbinxlo + 1;
  b.mbiny = hi[1] - b.mbinylo + 1;
  b.mbinz = hi[2] - b.mbinzlo + 1;
  b.mbins = b.mbinx * b.mbiny * b.mbinz + 1;
}

// NStencil create_setup / bin_distance / full + half create,
// src/nstencil.cpp:101-143, src/nstencil_full_bin_3d.cpp:26-37,
// src/nstencil_half_bin_3d.cpp:26-38
void create_stencils(phafd_hip_ctx *c) {
  const BinGeom &b = c->bg;
  const double cutneighsq = c->cutneigh * c->cutneigh;
  auto bin_distance = [&](int i, int j, int k) {
    double delx, dely, delz;
    if (i > 0) delx = (i - 1) * b.binsizex; else if (i == 0) delx = 0.0; else delx = (i + 1) * b.binsizex;
    if (j > 0) dely = (j - 1) * b.binsizey; else if (j == 0) dely = 0.0; else dely = (j + 1) * b.binsizey;
    if (k > 0) delz = (k - 1) * b.binsizez; else if (k == 0) delz = 0.0; else delz = (k + 1) * b.binsizez;
    return (delx * delx + dely * dely + delz * delz);
  };
  int sx = (int)(c->cutneigh * b.bininvx); if (sx * b.binsizex < c->cutneigh) sx++;
  int sy = (int)(c->cutneigh * b.bininvy); if (sy * b.binsizey < c->cutneigh) sy++;
  int sz = (int)(c->cutneigh * b.bininvz); if (sz * b.binsizez < c->cutneigh) sz++;
  c->stencil_full.clear();
  std::vector<int> fx, fy, fz;
  for (int k = -sz; k <= sz; k++)
    for (int j = -sy; j <= sy; j++)
      for (int i = -sx; i <= sx; i++)
        if (bin_distance(i, j, k) < cutneighsq) {
          c->stencil_full.push_back(k * b.mbiny * b.mbinx + j * b.mbinx + i);
          fx.push_back(i); fy.push_back(j); fz.push_back(k);
        }
  c->d_sfx.alloc(fx.size() + 1); c->d_sfy.alloc(fx.size() + 1); c->d_sfz.alloc(fx.size() + 1);
  if (!fx.empty()) {
    HIPCHK(hipMemcpy(c->d_sfx.p, fx.data(), sizeof(int) * fx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_sfy.p, fy.data(), sizeof(int) * fy.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_sfz.p, fz.data(), sizeof(int) * fz.size(), hipMemcpyHostToDevice));
  }
  c->stencil_half.clear();
  for (int k = 0; k <= sz; k++)
    for (int j = -sy; j <= sy; j++)
      for (int i = -sx; i <= sx; i++)
        if (k > 0 || j > 0 || (j == 0 && i > 0))
          if (bin_distance(i, j, k) < cutneighsq)
            c->stencil_half.push_back(k * b.mbiny * b.mbinx + j * b.mbinx + i);
  double ljcut = 0.0;
  for (auto &p : c->pairs) {
    if (p->style == 2) ljcut = std::max(ljcut, std::sqrt(p->ljcutsq));
    if (p->style == 4) ljcut = std::max(ljcut, p->hcut);
  }
  const double prune = ljcut + c->skin;
  c->prune_sq = prune < c->cutneigh ? prune * prune : cutneighsq;
  c->stencil_lj.clear();
  for (int k = 0; k <= sz; k++)
    for (int j = -sy; j <= sy; j++)
      for (int i = -sx; i <= sx; i++)
        if (k > 0 || j > 0 || (j == 0 && i > 0))
          if (bin_distance(i, j, k) < cutneighsq && bin_distance(i, j, k) <= c->prune_sq)
            c->stencil_lj.push_back(k * b.mbiny * b.mbinx + j * b.mbinx + i);
  c->d_stencil_lj.alloc(c->stencil_lj.size() + 1);
  if (!c->stencil_lj.empty())
    HIPCHK(hipMemcpy(c->d_stencil_lj.p, c->stencil_lj.data(), sizeof(int) * c->stencil_lj.size(), hipMemcpyHostToDevice));
  c->d_stencil_full.alloc(c->stencil_full.size() + 1);
  c->d_stencil_half.alloc(c->stencil_half.size() + 1);
  if (!c->stencil_full.empty())
    HIPCHK(hipMemcpy(c->d_stencil_full.p, c->stencil_full.data(), sizeof(int) * c->stencil_full.size(), hipMemcpyHostToDevice));
  if (!c->stencil_half.empty())
    HIPCHK(hipMemcpy(c->d_stencil_half.p, c->stencil_half.data(), sizeof(int) * c->stencil_half.size(), hipMemcpyHostToDevice));
}

// NBinStandard::bin_atoms as a counting sort, src/nbin_standard.cpp:181-229
void bin_atoms(phafd_hip_ctx *c) {
  const int ng = c->ngathered, nb = c->bg.mbins;
  c->a2b.alloc((size_t)ng + 1);
  c->binatoms.alloc((size_t)ng + 1);
  c->bincount.alloc((size_t)nb + 1);
  c->bincursor.alloc((size_t)nb + 1);
  c->binstart.alloc((size_t)nb + 2);
  launch_fill_i32(nb, c->bincount.p, 0, c->stream);
  launch_fill_i32(nb, c->bincursor.p, 0, c->stream);
  launch_fill_i32(1, c->flags.p + 2, 0, c->stream);
  launch_atom2bin(ng, c->gx.p, c->bg, c->a2b.p, c->bincount.p, c->flags.p + 2, c->stream);
  if (read_int(c, c->flags.p + 2))
    throw std::runtime_error("Non-numeric positions - simulation unstable (atom outside the bin grid)");
  scan_exclusive_i32(c->bincount.p, c->binstart.p, (size_t)nb, c->stream);
  launch_bin_place(ng, c->a2b.p, c->binstart.p, c->bincursor.p, c->binatoms.p, c->stream);
  launch_bin_sort(nb, c->binstart.p, c->binatoms.p, c->glabel.p, c->stream);
}

// flags[1], read at a build's count synchronise: bit 0 a row of more than
// Neighbor::oneatom (2000) entries, which the reference reports on stderr
// and goes on (src/npair_grid_bin.cpp:118-121, src/npair_half_bin_newton.cpp:
// 117-120); bit 1 a fill that found more entries than its count (the writes
// past the row's scanned segment were dropped), an error
void list_flags(phafd_hip_ctx *c, int ovf) {
  if (!ovf) return;
  launch_fill_i32(1, c->flags.p + 1, 0, c->stream);
  if (ovf & 2) throw std::runtime_error("neighbor list fill exceeded the entries counted for its row");
  std::fprintf(stderr, "Neighbor list overflow, terrible things are going to happen.\n");
}

// NPairGridBin::build (src/npair_grid_bin.cpp:22-127)
void build_gridlist(phafd_hip_ctx *c) {
  GridView gv = c->gridview<double>();
  const double4 *gxb = c->gx_build.p;
  const double cutneighsq = c->cutneigh * c->cutneigh;
  c->g_cnt.alloc(c->N + 1);
  c->g_off.alloc(c->N + 1);
  probe_begin(K_GRID_COUNT, c->stream);
  launch_grid_list(false, gv, c->bg, c->d_stencil_full.p, (int)c->stencil_full.size(), c->binstart.p,
                   c->binatoms.p, gxb, cutneighsq, c->nowned, c->g_cnt.p, (const int64_t *)nullptr, nullptr,
                   c->flags.p + 1, c->stream);
  probe_end(K_GRID_COUNT, c->stream, 4.0 * c->N);
  scan_exclusive_i32_to_i64(c->g_cnt.p, c->g_off.p, c->N, c->stream);
  int64_t tot = 0;
  int ovf = 0;
  HIPCHK(hipMemcpyAsync(&tot, c->g_off.p + c->N, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&ovf, c->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  list_flags(c, ovf);
  c->g_nent = tot;
  c->g_neigh.alloc((size_t)tot + 1);
  probe_begin(K_GRID_FILL, c->stream);
  launch_grid_list(true, gv, c->bg, c->d_stencil_full.p, (int)c->stencil_full.size(), c->binstart.p,
                   c->binatoms.p, gxb, cutneighsq, c->nowned, c->g_cnt.p, c->g_off.p, c->g_neigh.p,
                   c->flags.p + 1, c->stream);
  c->grid_built = true;
  probe_end(K_GRID_FILL, c->stream, 8.0 * c->N + 4.0 * tot);
}

// NPairHalfBinNewton::build (src/npair_half_bin_newton.cpp:38-124) plus the
// transposed slots used by the deterministic force gather
// full = the reference's list (cut = cutneigh), else the pruned lj/cut list;
// both from the bins and positions of the last build
void build_halflist(phafd_hip_ctx *c, bool full) {
  const int ng = c->ngathered;
  const double cutneighsq = full ? c->cutneigh * c->cutneigh : c->prune_sq;
  const DBuf<int> &st = full ? c->d_stencil_half : c->d_stencil_lj;
  const int nst = (int)(full ? c->stencil_half.size() : c->stencil_lj.size());
  const double4 *pos = c->gx_build.p;
  c->tmp_flag.alloc((size_t)ng + 16);
  c->tmp_pos.alloc((size_t)ng + 16);
  launch_local_flag(ng, c->glabel.p, c->tmp_flag.p, c->stream);
  scan_exclusive_i32(c->tmp_flag.p, c->tmp_pos.p, (size_t)ng, c->stream);
  const int nrows = c->nlocal;   // the LOCAL entries (label 1), counted at the borders
  c->h_nrows = nrows;
  c->h_ilist.alloc((size_t)nrows + 1);
  c->h_rowof.alloc((size_t)ng + 1);
  launch_local_scatter(ng, c->tmp_flag.p, c->tmp_pos.p, c->h_ilist.p, c->h_rowof.p, c->stream);
  c->h_cnt.alloc((size_t)nrows + 1);
  c->h_off.alloc((size_t)nrows + 2);
  launch_half_list(false, nrows, c->h_ilist.p, c->a2b.p, c->binstart.p, c->binatoms.p, c->glabel.p,
                   st.p, nst, pos, cutneighsq, c->nowned, c->h_cnt.p, nullptr, nullptr, c->flags.p + 1, c->stream);
  scan_exclusive_i32(c->h_cnt.p, c->h_off.p, (size_t)nrows, c->stream);
  int nent = 0, ovf = 0;
  HIPCHK(hipMemcpyAsync(&nent, c->h_off.p + nrows, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&ovf, c->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  list_flags(c, ovf);
  c->h_nent = nent;
  c->h_neigh.alloc((size_t)nent + 1);
  launch_half_list(true, nrows, c->h_ilist.p, c->a2b.p, c->binstart.p, c->binatoms.p, c->glabel.p,
                   st.p, nst, pos, cutneighsq, c->nowned, c->h_cnt.p, c->h_off.p, c->h_neigh.p, c->flags.p + 1,
                   c->stream);
  c->t_cnt.alloc((size_t)ng + 1);
  c->t_cur.alloc((size_t)ng + 1);
  c->t_off.alloc((size_t)ng + 2);
  c->t_pair.alloc((size_t)nent + 1);
  c->t_i.alloc((size_t)nent + 1);
  launch_fill_i32(ng, c->t_cnt.p, 0, c->stream);
  launch_fill_i32(ng, c->t_cur.p, 0, c->stream);
  launch_t_count(nent, c->h_neigh.p, c->nowned, c->t_cnt.p, c->stream);
  scan_exclusive_i32(c->t_cnt.p, c->t_off.p, (size_t)ng, c->stream);
  launch_t_place(nrows, c->h_ilist.p, c->h_off.p, c->h_neigh.p, c->nowned, c->t_off.p, c->t_cur.p,
                 c->t_pair.p, c->t_i.p, c->stream);
  launch_t_sort(ng, c->t_off.p, c->t_pair.p, c->t_i.p, c->stream);
  c->half_full = full;
}

// Neighbor::build, src/neighbor.cpp:99-125
// the largest grid-atom cut + skin/2, with a margin far above rounding
double ga_reach_of(const phafd_hip_ctx *c) {
  double r = 0;
  for (auto &p : c->pairs) {
    if (bead_pair(*p)) continue;
    for (double cs : p->cutsq) r = std::max(r, std::sqrt(cs));
  }
  return r > 0 ? (r + 0.5 * c->skin) * (1 + 1e-9) + 1e-9 : 0.0;
}
// the grid-atom pairs of the build: the atom-major walk (count, scan,
// fill), then the grouping by grid point (count with ranks, scan, place,
// order by q)
void build_ga_pairs(phafd_hip_ctx *c) {
  const double reach = ga_reach_of(c);
  if (reach <= 0) return;
  const int ng = c->ngathered;
  GridView gv = c->gridview<double>();
  c->ga_acnt.alloc((size_t)ng + 1);
  c->ga_aoff.alloc((size_t)ng + 2);
  launch_ga_am_walk(ng, gv, c->gx_build.p, c->binatoms.p, reach * reach, c->ga_acnt.p, nullptr, nullptr, nullptr,
                    c->stream);
  scan_exclusive_i32(c->ga_acnt.p, c->ga_aoff.p, (size_t)ng, c->stream);
  const int nent = read_int(c, c->ga_aoff.p + ng);
  if (nent < 0) throw std::runtime_error("grid-atom pairs of a build exceed 2^31");
  c->ga_nent = nent;
  c->ga_ps.alloc((size_t)nent + 1);
  c->ga_amq.alloc((size_t)nent + 1);
  launch_ga_am_walk(ng, gv, c->gx_build.p, c->binatoms.p, reach * reach, nullptr, c->ga_aoff.p, c->ga_ps.p,
                    c->ga_amq.p, c->stream);
  c->ga_pcnt.alloc(c->N + 1);
  c->ga_poff.alloc(c->N + 2);
  c->ga_rank.alloc((size_t)nent + 1);
  c->ga_key.alloc((size_t)nent + 1);
  c->ga_ent.alloc((size_t)nent + 1);
  fill_zero(c->ga_pcnt.p, sizeof(int) * c->N, c->stream);
  launch_ga_pm_count(nent, c->ga_ps.p, c->ga_pcnt.p, c->ga_rank.p, c->stream);
  scan_exclusive_i32(c->ga_pcnt.p, c->ga_poff.p, c->N, c->stream);
  launch_ga_pm_place(nent, c->ga_ps.p, c->ga_rank.p, c->ga_amq.p, c->ga_poff.p, c->ga_key.p, c->ga_ent.p, c->N,
                     c->stream);
  if (c->fp32) c->ga_mterm32.alloc((size_t)nent + 1);
  else c->ga_mterm.alloc((size_t)nent + 1);
  c->ga_calls = c->ncalls;
}

void neighbor_build(phafd_hip_ctx *c) {
  c->ncalls++;
  launch_copy4(c->nowned, c->x.p, c->xhold.p, c->stream);
  bin_atoms(c);
  c->gx_build.alloc((size_t)c->ngathered + 1);
  launch_copy4(c->ngathered, c->gx.p, c->gx_build.p, c->stream);
  c->grid_built = false;
  if (c->need_full && c->eager_gridlist) build_gridlist(c);
  if (c->need_full) build_ga_pairs(c);
  if (c->need_half) build_halflist(c, c->eager_halflist);
  c->lists_valid = true;
}

// Device flags: [0] and [4] the rebuild decision (Neighbor::decide) made at
// the end of a step, slot 4 * (timestep & 1); [1] list overflow; [2] atom
// outside the bins; [5] an immediate decide.
// A step's decide clears the slot of the step after it, so no fill runs per
// step; a run loop clears its first step's slot once.
int decide_slot(const phafd_hip_ctx *c) { return 4 * (int)(c->timestep & 1); }
void clear_decide_slot(phafd_hip_ctx *c, int64_t step) {
  launch_fill_i32(1, c->flags.p + 4 * (int)(step & 1), 0, c->stream);
}
void enqueue_decide(phafd_hip_ctx *c) {
  const int sl = decide_slot(c);
  launch_decide(c->nowned, c->x.p, c->xhold.p, c->triggersq, c->flags.p + sl, c->flags.p + (4 - sl), c->stream);
}
// the step's flags to h_flags[8..15], read after the step's synchronise
void enqueue_flags_readback(phafd_hip_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->h_flags + 8, c->flags.p, 8 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
}
// the bead-rod fix's failure flag: flags[3] (read back with each step's
// flags); FixAtomSemiFlexible::final_integrate's error (src/fixatom_semiflexible.cpp:176-188)
void semiflex_failed(phafd_hip_ctx *c) {
  std::fprintf(stderr, "bad timestep from no_tether of polymer on processor 0 at timestep %lld\n",
               (long long)c->timestep);
  throw std::runtime_error("Failed to maintain polymer bond-length constraints.");
}
bool has_semiflex(const phafd_hip_ctx *c) {
  for (auto &f : c->fixes)
    if (f->kind == FixDef::SEMIFLEX) return true;
  return false;
}

// the checks after step `step`'s synchronise (src/integrate.cpp:190-201)
void check_step_flags(phafd_hip_ctx *c, int64_t step) {
  if (c->h_flags[11]) semiflex_failed(c);
  if (c->neigh_set) {
    const int f = c->h_flags[8 + 4 * (int)(step & 1)];
    if (f & 2) throw std::runtime_error("Non-numeric positions - simulation unstable");
    c->pending_decision = f & 1;
  }
}

bool neighbor_decide(phafd_hip_ctx *c) {
  if (c->pending_decision >= 0) {
    bool r = c->pending_decision != 0;
    c->pending_decision = -1;
    return r;
  }
  launch_fill_i32(1, c->flags.p + 5, 0, c->stream);
  launch_decide(c->nowned, c->x.p, c->xhold.p, c->triggersq, c->flags.p + 5, nullptr, c->stream);
  const int f = read_int(c, c->flags.p + 5);
  if (f & 2) throw std::runtime_error("Non-numeric positions - simulation unstable");
  return (f & 1) != 0;
}

void upload_pair(PairDef &p, double skin) {
  int nt = (int)p.eps.size();
  p.d_eps.alloc(nt); p.d_phit.alloc(nt); p.d_cutsq.alloc(nt);
  HIPCHK(hipMemcpy(p.d_eps.p, p.eps.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_phit.p, p.phit.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.d_cutsq.p, p.cutsq.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
}

GAParams ga_params(const PairDef &p) {
  GAParams g;
  g.style = p.style;
  g.ntypes = (int)p.eps.size();
  if (g.ntypes > GA_MAXTYPES) throw std::invalid_argument("grid-atom styles take at most 64 atom types");
  g.eps = p.d_eps.p; g.phit = p.d_phit.p; g.cutsq = p.d_cutsq.p;
  g.em05 = exp(-0.5);
  g.e05 = exp(0.5);
  return g;
}

// ---------------------------------------------------------- grid hooks
SpecConst spec_const(const phafd_hip_ctx *c) {
  SpecConst k;
  k.dqx = 2 * PI_REF / c->period[0];      // Domain::dqx
  k.dqy_kz = 2 * PI_REF / c->period[2];   // Domain::dqy
  k.dqz_ky = 2 * PI_REF / c->period[1];   // Domain::dqz
  return k;
}

// FixGridConjugate::start_of_step / FixGridGradPhi::start_of_step,
// src/fixgrid_conjugate.cpp:62-116: r2c, grad multiply, three c2r, / N
bool spectral_fused(const phafd_hip_ctx *c) { return c->fuse_spectral && fft_fused_z_ok(c->plan); }

// FP32 form: the k = 0 mode of the spectral update comes from the FP64 sum
// of phi taken here (mass), so the mean of phi does not drift with FP32
// rounding of the transform.
template <class R>
void compute_gradphi(phafd_hip_ctx *c, typename Vec2<R>::T *phihat_dst, bool mass) {
  using C = typename Vec2<R>::T;
  FieldSet<R> &F = c->fields<R>();
  C *ins[3] = {F.ghat[0].p, F.ghat[1].p, F.ghat[2].p};
  R *outs[3] = {F.grad[0].p, F.grad[1].p, F.grad[2].p};
  const double factor = (double)c->nx * c->ny * c->nz;
  if (mass && c->fp32) {
    fill_zero(c->mass64.p, sizeof(double), c->stream);
    sum_fp64(c->N, F.phi.p, c->mass64.p, c->stream);
  }
  if (c->fuse_spectral && fft_sep_grad_ok(c->plan)) {
    fft_grad_sep(c->plan, F.phi.p, phihat_dst, ins[1], ins[2], outs[0], outs[1], outs[2], spec_const(c),
                 c->stream);
    return;
  }
  fft_r2c(c->plan, F.phi.p, phihat_dst, c->stream);
  grad_multiply(c->plan, phihat_dst, F.ghat[0].p, F.ghat[1].p, F.ghat[2].p, spec_const(c), c->stream);
  fft_c2r_batch(c->plan, ins, outs, 3, factor, c->stream);
}
const double *origin_ptr(const phafd_hip_ctx *c) { return c->fp32 ? c->mass64.p : nullptr; }

UpdateParams update_params(const phafd_hip_ctx *c, const FixDef &f) {
  UpdateParams u{};
  u.mobility = f.mobility; u.temp = f.temp; u.volFH = f.volFH; u.gamma = f.gamma; u.dt = c->dt;
  u.normalization = 1.0 / ((double)c->nx * c->nz * c->ny);   // src/conjugate_volfrac.cpp:78
  u.qx_scale = 2 * PI_REF / c->period[0];
  u.qky_scale = 2 * PI_REF / c->period[1];
  u.qkz_scale = 2 * PI_REF / c->period[2];
  u.semi = c->semi_implicit ? 1 : 0;
  if (f.noise_mode == PHAFD_NOISE_DEVICE) {
    // ConjugateVolFrac::reset_dt prefactors, src/conjugate_volfrac.cpp:72-86
    const double invL3 = 1.0 / (c->period[0] * c->period[1] * c->period[2]);
    u.noise_dev = 1;
    u.key0 = f.noise_key;
    u.key1 = 0x5048414644ull;   // fixed second key word of this generator
    u.step = (unsigned long long)c->timestep;
    u.pref_c = std::sqrt(12 * f.temp * f.mobility * invL3);
    u.pref_r = std::sqrt(24 * f.temp * f.mobility * invL3);
    u.sqrtdt = std::sqrt(c->dt);
  }
  return u;
}

// The host noise field of this update in PHAFD_NOISE_FIELD mode: the oldest
// unconsumed phafd_hip_set_noise field, consumed here (the reference draws
// fresh noise at every update, src/conjugate_volfrac.cpp:95-96,118).
template <class R>
const typename Vec2<R>::T *take_noise(phafd_hip_ctx *c, const FixDef &f) {
  if (f.noise_mode != PHAFD_NOISE_FIELD) return nullptr;
  FieldSet<R> &F = c->fields<R>();
  if (c->noise_head >= F.noise_q.size())
    throw std::invalid_argument("grid/conjugate/volfrac in noise field mode: every update needs its own "
                                "phafd_hip_set_noise field and none is queued");
  return F.noise_q[c->noise_head++].p;
}
size_t noise_queued(const phafd_hip_ctx *c) {
  const size_t n = c->fp32 ? c->fs.noise_q.size() : c->fd.noise_q.size();
  return n - c->noise_head;
}
// a run of nsteps with a noise-field conjugate fix needs nsteps fields queued
void require_noise_fields(const phafd_hip_ctx *c, int64_t nsteps) {
  for (auto &f : c->fixes)
    if (f->kind == FixDef::CONJ && f->noise_mode == PHAFD_NOISE_FIELD && (int64_t)noise_queued(c) < nsteps)
      throw std::invalid_argument("grid/conjugate/volfrac in noise field mode: a run of " + std::to_string(nsteps) +
                                  " steps needs as many phafd_hip_set_noise fields, " +
                                  std::to_string(noise_queued(c)) + " queued");
}

// ---------------------------------------------------------------- phases
template <class R>
void phase_start_t(phafd_hip_ctx *c) {
  FieldSet<R> &F = c->fields<R>();
  const bool ready = c->grad_ready;
  c->grad_ready = false;
  for (auto &fp : c->fixes) {
    FixDef &f = *fp;
    if (f.kind == FixDef::CONJ && ready) {
      if (c->fp32) {   // the k = 0 mode of the FP32 update: the FP64 sum of phi
        fill_zero(c->mass64.p, sizeof(double), c->stream);
        sum_fp64(c->N, F.phi.p, c->mass64.p, c->stream);
      }
    } else if (f.kind == FixDef::CONJ) compute_gradphi<R>(c, F.phihat.p, true);
    else if (f.kind == FixDef::GRADPHI && (!f.once || c->timestep == c->firststep))
      compute_gradphi<R>(c, F.muhat.p, false);
  }
}
void phase_start(phafd_hip_ctx *c) {
  if (c->fp32) phase_start_t<float>(c);
  else phase_start_t<double>(c);
}

void phase_neigh(phafd_hip_ctx *c) {
  if (!c->neigh_set) return;
  if (neighbor_decide(c)) {
    launch_pbc(c->nowned, c->x.p, c->ux.p, c->image.p, c->box(), c->stream);
    borders(c);
    neighbor_build(c);
  } else {
    forward_comm(c);
  }
}

// One grid-atom pair style (PairGridAtomGaussian / PairGridAtomLJish
// compute) over the pairs of the last build: the atom-side walk (forces in
// ascending grid index, each kept pair's mu term), then the grid-side sum
// of the terms in the reference's list order (or, with the eager option,
// the walk of the stored NPairGridBin list; the two are bit-identical).
// first: mu has not been zeroed this step; the pass writes every point.
template <class R> DBuf<R> &ga_mterm(phafd_hip_ctx *c);
template <> DBuf<double> &ga_mterm<double>(phafd_hip_ctx *c) { return c->ga_mterm; }
template <> DBuf<float> &ga_mterm<float>(phafd_hip_ctx *c) { return c->ga_mterm32; }
template <class R>
void grid_atom_pass(phafd_hip_ctx *c, PairDef &p, double *energy, bool first) {
  FieldSet<R> &Fs = c->fields<R>();
  const GridViewT<R> gvr = c->gridview<R>();
  GAParams ga = ga_params(p);
  if (c->ga_calls != c->ncalls) build_ga_pairs(c);
  DBuf<R> &mt = ga_mterm<R>(c);
  mt.alloc((size_t)c->ga_nent + 1);
  // algorithmic bytes: phi and its gradient read once (4 fields), each
  // gathered atom's position read and force written
  probe_begin(K_GA_FORCE, c->stream);
  launch_ga_force(c->ngathered, gvr, ga, c->gx.p, c->binatoms.p, c->ga_aoff.p, c->ga_ps.p, mt.p, c->gF.p, energy,
                  c->stream);
  probe_end(K_GA_FORCE, c->stream, 4.0 * sizeof(R) * c->N + 64.0 * c->ngathered);
  probe_begin(K_GA_DEPOSIT, c->stream);
  if (!(c->eager_gridlist && c->grid_built)) {
    launch_ga_list_mu(c->N, Fs.mu.p, c->ga_poff.p, mt.p, first, c->stream);
  } else {
    if (c->fp32) throw std::invalid_argument("the stored grid list walk is FP64 only");
    if (first) fill_zero(Fs.mu.p, sizeof(R) * c->N, c->stream);
    launch_ga_deposit(c->gridview<double>(), ga, c->g_cnt.p, c->g_off.p, c->g_neigh.p, c->gx.p, c->nowned, nullptr,
                      c->stream);
  }
  probe_end(K_GA_DEPOSIT, c->stream, 2.0 * sizeof(R) * c->N + (first ? 0.0 : sizeof(R) * c->N));
}

// Integrate::run pair block, src/integrate.cpp:130-146 (one pass, finding F4)
void phase_pair(phafd_hip_ctx *c, double *energy_ga = nullptr, double *energy_lj = nullptr) {
  // mu = 0 before the pair computes: done by the first grid-atom pass, which
  // writes every grid point, when there is one
  bool mu_pending = true;
  const bool any_ga = c->neigh_set && c->lists_valid &&
                      std::any_of(c->pairs.begin(), c->pairs.end(), [](const std::unique_ptr<PairDef> &p) {
                        return !bead_pair(*p);
                      });
  if (!any_ga) {
    fill_zero(c->fp32 ? (void *)c->fs.mu.p : (void *)c->fd.mu.p, c->real_bytes() * c->N, c->stream);
    mu_pending = false;
  }
  launch_fill_d4(c->ngathered, c->gF.p, c->stream);
  launch_fill_d4(c->nowned, c->F.p, c->stream);
  if (!c->neigh_set || !c->lists_valid) return;
  for (auto &pp : c->pairs) {
    PairDef &p = *pp;
    if (p.style == 2) {
      LJParams lj{p.lj1, p.lj2, p.ljcutsq};
      probe_begin(K_LJ, c->stream);
      launch_lj_force(c->ngathered, lj, c->gx.p, c->h_rowof.p, c->h_off.p, c->h_neigh.p, c->t_off.p,
                      c->t_pair.p, c->t_i.p, c->nowned, c->gF.p, energy_lj, c->stream);
      probe_end(K_LJ, c->stream, 64.0 * c->ngathered + 4.0 * c->h_nent);
    } else if (p.style == 4) {
      launch_harmonic_force(c->ngathered, HarmParams{p.hk, p.hcut}, c->gx.p, c->h_rowof.p, c->h_off.p, c->h_neigh.p,
                            c->t_off.p, c->t_pair.p, c->t_i.p, c->nowned, c->gF.p, c->stream);
    } else {
      if (c->fp32) grid_atom_pass<float>(c, p, energy_ga, mu_pending);
      else grid_atom_pass<double>(c, p, energy_ga, mu_pending);
      mu_pending = false;
    }
  }
  reverse_comm(c);
}

// second: the final_integrate half (FixAtomDrag and the harmonic chains do
// the same in both halves; the bead-rod fix runs first_step, then second_step)
void integrate_half(phafd_hip_ctx *c, FixDef &f, bool second = false) {
  if (f.kind == FixDef::SEMIFLEX) {
    BRParamsDev p{f.bondlength, f.zpar, f.zperp, f.bending, f.brtemp, c->dt, f.tether, 50, f.brstep};
    if (!second) {
      p.step = ++f.brstep;
      launch_br_first(f.nchain, f.maxlen, f.chs.p, f.che.p, f.d_seed.p, c->F.p, c->ux.p, c->x.p, c->image.p,
                      f.br_R0.p, f.br_FB.p, f.br_T.p, p, c->box(), c->stream);
    } else {
      launch_br_second(f.nchain, f.maxlen, f.chs.p, f.che.p, c->F.p, c->ux.p, c->x.p, c->image.p, f.br_R0.p,
                       f.br_FB.p, f.br_T.p, c->flags.p + 3, p, c->box(), c->stream);
    }
    return;
  }
  if (f.kind == FixDef::DRAG) {
    launch_drag(f.nmem, f.members.p, c->x.p, c->F.p, f.drag, c->dt, c->stream);
  } else if (f.kind == FixDef::CHAIN) {
    ChainParams cp{f.r0, f.kbond, f.kbend, f.drag};
    launch_chain_step(f.nchain, f.maxlen, f.chs.p, f.che.p, c->F.p, c->ux.p, c->x.p, c->image.p, cp, c->dt,
                      c->box(), c->stream);
  }
}

void phase_initial(phafd_hip_ctx *c) {
  for (auto &f : c->fixes) integrate_half(c, *f);
}

// The chain fix when it is the only atom integrator (and its chains fit two
// LDS buffers of 32 KB): then both half steps of a timestep run as one
// launch at initial_integrate.  F is computed once per step by the pair pass
// and the hooks between initial_integrate and final_integrate work on the
// fields only, so nothing reads or moves the atoms in between and the
// positions are those of two separate half steps.
FixDef *fused_chain_fix(phafd_hip_ctx *c) {
  FixDef *ch = nullptr;
  int n = 0;
  for (auto &f : c->fixes) {
    if (f->kind == FixDef::CONJ || f->kind == FixDef::FH) continue;
    n++;
    if (f->kind == FixDef::CHAIN) ch = f.get();
  }
  return (n == 1 && ch && ch->maxlen <= 1024) ? ch : nullptr;
}
void integrate_both_halves(phafd_hip_ctx *c, FixDef &f, StepGate gate = StepGate()) {
  ChainParams cp{f.r0, f.kbond, f.kbend, f.drag};
  launch_chain_step(f.nchain, f.maxlen, f.chs.p, f.che.p, c->F.p, c->ux.p, c->x.p, c->image.p, cp, c->dt, c->box(),
                    c->stream, 2, gate);
}

template <class R>
void phase_postforce_t(phafd_hip_ctx *c) {
  FieldSet<R> &F = c->fields<R>();
  for (auto &fp : c->fixes)
    if (fp->kind == FixDef::FH)
      floryhuggins(c->N, F.phi.p, F.mu.p, fp->temp / fp->volFH, fp->chi, c->counters.p, c->stream);
}

template <class R>
void phase_prefinal_t(phafd_hip_ctx *c, const FHPreT<R> *pre = nullptr) {
  FieldSet<R> &F = c->fields<R>();
  for (auto &fp : c->fixes)
    if (fp->kind == FixDef::CONJ) fft_r2c(c->plan, F.mu.p, F.muhat.p, c->stream, pre);
}

template <class R>
void phase_final_t(phafd_hip_ctx *c, bool atoms = true) {
  FieldSet<R> &F = c->fields<R>();
  for (auto &fp : c->fixes) {
    FixDef &f = *fp;
    if (f.kind == FixDef::CONJ) {
      const auto *nz = take_noise<R>(c, f);
      spectral_update(c->plan, F.phihat.p, F.muhat.p, nz, update_params(c, f), origin_ptr(c), c->stream);
    } else if (atoms) {
      integrate_half(c, f, true);
    }
  }
}

template <class R>
void phase_postfinal_t(phafd_hip_ctx *c) {
  FieldSet<R> &F = c->fields<R>();
  for (auto &fp : c->fixes)
    if (fp->kind == FixDef::CONJ) fft_c2r(c->plan, F.phihat.p, F.phi.p, 0.0, c->stream);
}

// The same three hooks with the z pass of r2c(mu), the update and the z pass
// of c2r(phihat) folded into one kernel; used by run_steps, where the hooks
// always run back to back (one conjugate fix).
template <class R>
void phase_prefinal_split(phafd_hip_ctx *c, const FHPreT<R> *pre) {
  FieldSet<R> &F = c->fields<R>();
  fft_r2c_xy(c->plan, F.mu.p, F.muhat.p, c->stream, pre);
}
template <class R>
void phase_final_split(phafd_hip_ctx *c, bool atoms = true, bool next = false) {
  FieldSet<R> &F = c->fields<R>();
  for (auto &fp : c->fixes) {
    FixDef &f = *fp;
    if (f.kind == FixDef::CONJ && next) {
      const auto *nz = take_noise<R>(c, f);
      fft_z_update_next(c->plan, F.muhat.p, F.phihat.p, F.ghat[0].p, F.ghat[2].p, nz, update_params(c, f),
                        origin_ptr(c), spec_const(c), c->stream, c->store_muhat);
    } else if (f.kind == FixDef::CONJ) {
      const auto *nz = take_noise<R>(c, f);
      fft_z_update(c->plan, F.muhat.p, F.phihat.p, nz, update_params(c, f), origin_ptr(c), c->stream, Slab(),
                   c->store_muhat);
    } else if (atoms) {
      integrate_half(c, f, true);
    }
  }
}
template <class R>
void phase_postfinal_split(phafd_hip_ctx *c, bool next = false) {
  FieldSet<R> &F = c->fields<R>();
  if (next) {
    fft_c2r_next(c->plan, F.ghat[0].p, F.ghat[1].p, F.ghat[2].p, F.phi.p, F.grad[0].p, F.grad[1].p, F.grad[2].p,
                 spec_const(c), c->stream);
    c->grad_ready = true;
    return;
  }
  auto *in = F.phihat.p;
  R *out = F.phi.p;
  fft_c2r_yx_batch(c->plan, &in, &out, 1, 0.0, c->stream);
}

void run_phase(phafd_hip_ctx *c, int ph) {
  switch (ph) {
    case PHAFD_PH_START: phase_start(c); break;
    case PHAFD_PH_NEIGH: phase_neigh(c); break;
    case PHAFD_PH_PAIR: phase_pair(c); break;
    case PHAFD_PH_INITIAL: phase_initial(c); break;
    case PHAFD_PH_POSTFORCE:
      if (c->fp32) phase_postforce_t<float>(c);
      else phase_postforce_t<double>(c);
      break;
    case PHAFD_PH_PREFINAL:
      if (c->fp32) phase_prefinal_t<float>(c);
      else phase_prefinal_t<double>(c);
      break;
    case PHAFD_PH_FINAL:
      if (c->fp32) phase_final_t<float>(c);
      else phase_final_t<double>(c);
      if (has_semiflex(c) && read_int(c, c->flags.p + 3)) semiflex_failed(c);
      break;
    case PHAFD_PH_POSTFINAL:
      if (c->fp32) phase_postfinal_t<float>(c);
      else phase_postfinal_t<double>(c);
      break;
    default: throw std::invalid_argument("bad phase");
  }
}

void require_setup(phafd_hip_ctx *c) {
  if (!c->setup_done) throw std::invalid_argument("call phafd_hip_setup before running");
}

// FH fused into the r2c of mu when the two hooks are adjacent and alone
template <class R>
bool can_fuse_fh(phafd_hip_ctx *c, FHPreT<R> &pre) {
  if (!c->fuse_fh) return false;
  int nconj = 0, nfh = 0;
  const FixDef *fh = nullptr;
  for (auto &f : c->fixes) {
    if (f->kind == FixDef::CONJ) nconj++;
    if (f->kind == FixDef::FH) { nfh++; fh = f.get(); }
  }
  if (nconj != 1 || nfh != 1) return false;
  pre.on = true;
  pre.phi = c->fields<R>().phi.p;
  pre.mu = c->fields<R>().mu.p;
  pre.tv = fh->temp / fh->volFH;
  pre.chi = fh->chi;
  pre.counters = c->counters.p;
  return true;
}

// the check of Integrate::run_until_touching after a step
// (src/integrate.cpp:355-385): the distance between the last and the first
// owned atom (positions as held, no minimum image) against touch_point;
// true when the run stops here
bool touch_check(phafd_hip_ctx *c, int64_t step, int done) {
  Touch &t = *c->touch;
  const double *a = c->h_ends, *b = c->h_ends + 4;
  const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
  t.steps = done;
  if (dist < t.touch_point) {
    if (t.since >= t.delay) t.occurred.push_back(step);   // "cyclisation occured at step"
    t.since = 0;
    t.breakflag = true;
  } else {
    if (t.breakflag && t.since == t.delay) t.ended.push_back(step - t.delay);   // "interval ended at step"
    t.since += 1;
  }
  if (!t.cont && t.breakflag) {
    t.stopped = true;
    return true;
  }
  return false;
}

__global__ void k_widen_c(size_t n, const float2 *__restrict__ a, double2 *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = make_double2(a[i].x, a[i].y);
}
void capture_spectrum(phafd_hip_ctx *c, int k, const double2 *src) {
  c->cap_hat[k].alloc(c->M);
  HIPCHK(hipMemcpyAsync(c->cap_hat[k].p, src, sizeof(double2) * c->M, hipMemcpyDeviceToDevice, c->stream));
}
void capture_spectrum(phafd_hip_ctx *c, int k, const float2 *src) {
  c->cap_hat[k].alloc(c->M);
  const int blocks = (int)std::min<size_t>((c->M + 255) / 256, 8192);
  hipLaunchKernelGGL(k_widen_c, dim3(blocks), dim3(256), 0, c->stream, c->M, src, c->cap_hat[k].p);
  HIPCHK(hipGetLastError());
}

// see run_steps_t: the steps with the flags read back mid-step, step k + 1
// enqueued before step k ends
template <class R>
void run_steps_pipelined(phafd_hip_ctx *c, int nsteps, bool fused, bool zsplit, bool next_ok, FHPreT<R> &pre,
                         FixDef *fc) {
  auto phi0_slot = [&](int set) { return c->h_flags + (set ? 16 : 2); };
  // the timers and the NaN check of a completed step
  auto finish = [&](int set, int64_t step) {
    hipEvent_t *E = c->evp[set];
    for (int k = 0; k < 7; k++) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, E[k], E[k + 1]));
      c->tacc[k] += ms;
    }
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, E[0], E[8]));
    c->tacc[7] += tot;
    R phi0;
    std::memcpy(&phi0, phi0_slot(set), sizeof(R));
    if (std::isnan(phi0)) {
      sync(c);   // a later step in flight ends first
      c->timestep = step;
      c->grad_ready = false;
      c->pending_decision = -1;
      throw std::runtime_error("NAN encountered in phi.");
    }
  };
  c->grad_ready = false;
  int64_t prev_step = -1;
  for (int it = 0; it < nsteps; it++) {
    const int set = it & 1;
    hipEvent_t *E = c->evp[set];
    c->timestep++;
    const int64_t step = c->timestep;
    const bool last = it == nsteps - 1;
    c->store_muhat = last;
    HIPCHK(hipEventRecord(E[0], c->stream));
    phase_start(c);
    HIPCHK(hipEventRecord(E[1], c->stream));
    phase_neigh(c);
    HIPCHK(hipEventRecord(E[2], c->stream));
    phase_pair(c);
    HIPCHK(hipEventRecord(E[3], c->stream));
    if (fc) integrate_both_halves(c, *fc);
    else phase_initial(c);
    HIPCHK(hipEventRecord(E[4], c->stream));
    // next step's rebuild decision (Neighbor::decide): the atoms are final
    if (c->neigh_set) enqueue_decide(c);
    enqueue_flags_readback(c);
    HIPCHK(hipEventRecord(E[9], c->stream));
    if (!fused) phase_postforce_t<R>(c);
    if (zsplit) phase_prefinal_split<R>(c, fused ? &pre : nullptr);
    else phase_prefinal_t<R>(c, fused ? &pre : nullptr);
    HIPCHK(hipEventRecord(E[5], c->stream));
    const bool capture = c->fourier_capture && last;
    if (capture) capture_spectrum(c, 0, c->fields<R>().phihat.p);   // in_fourier: ft_phi before the update
    const bool next = next_ok && !last;
    if (zsplit) phase_final_split<R>(c, !fc, next);
    else phase_final_t<R>(c, !fc);
    HIPCHK(hipEventRecord(E[6], c->stream));
    if (zsplit) phase_postfinal_split<R>(c, next);
    else phase_postfinal_t<R>(c);
    if (capture) {   // ft_chempot: the last step of a run stores the complete muhat
      capture_spectrum(c, 1, c->fields<R>().muhat.p);
      c->cap_valid = true;
    }
    HIPCHK(hipEventRecord(E[7], c->stream));
    // NaN check of phi(0,0,0) (src/integrate.cpp:190-201)
    HIPCHK(hipMemcpyAsync(phi0_slot(set), c->fields<R>().phi.p, sizeof(R), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(E[8], c->stream));
    // this step's flags (mid-step); the previous step is complete by now
    HIPCHK(hipEventSynchronize(E[9]));
    if (prev_step >= 0) finish(set ^ 1, prev_step);
    try {
      check_step_flags(c, step);
    } catch (...) {
      // the reference checks phi at the end of the step, before the next
      // step's neighbour decision: a NaN in phi takes precedence
      sync(c);
      c->grad_ready = false;
      finish(set, step);
      throw;
    }
    prev_step = step;
  }
  sync(c);
  probe_collect();
  if (prev_step >= 0) finish((nsteps - 1) & 1, prev_step);
}

template <class R>
void run_steps_t(phafd_hip_ctx *c, int nsteps) {
  FHPreT<R> pre;
  const bool fused = can_fuse_fh(c, pre);
  int nconj = 0;
  for (auto &f : c->fixes) nconj += f->kind == FixDef::CONJ;
  const bool zsplit = nconj == 1 && spectral_fused(c);
  int ngrad = 0;
  for (auto &f : c->fixes) ngrad += f->kind == FixDef::GRADPHI;
  // a run that may stop after any step (run_until_touching) keeps every
  // step's fields as the reference leaves them
  const bool next_ok = zsplit && c->next_grad && !c->touch && ngrad == 0 && fft_sep_grad_ok(c->plan);
  c->grad_ready = false;
  FixDef *fc = fused_chain_fix(c);
  if (c->neigh_set) clear_decide_slot(c, c->timestep + 1);
  // Pipelined steps: when every atom move of a step happens at
  // initial_integrate (the fused chain fix, or no atom integrator), the next
  // step's rebuild decision is known right after it.  The decision and the
  // flags are then read back mid-step, and the host enqueues step k + 1
  // while the device still runs step k's grid passes; the NaN check of step
  // k's phi(0,0,0) happens once step k + 1 is enqueued (after PHAFD_E_NAN the
  // fields hold one step more than the step reported).
  bool atom_final = false;
  for (auto &f : c->fixes)
    if (f->kind != FixDef::CONJ && f->kind != FixDef::FH && f->kind != FixDef::GRADPHI && f.get() != fc)
      atom_final = true;
  if (!c->touch && !atom_final) {   // an active kernel probe collects its events after the run
    run_steps_pipelined<R>(c, nsteps, fused, zsplit, next_ok, pre, fc);
    return;
  }
  bool started = false;   // this step's start_of_step went in before the last read-back
  for (int it = 0; it < nsteps; it++) {
    if (!started) c->timestep++;
    const int64_t step = c->timestep;
    // muhat is observable only after the run: its last step stores it
    c->store_muhat = it == nsteps - 1;
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    if (!started) phase_start(c);
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    phase_neigh(c);
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    phase_pair(c);
    HIPCHK(hipEventRecord(c->ev[3], c->stream));
    if (fc) integrate_both_halves(c, *fc);
    else phase_initial(c);
    HIPCHK(hipEventRecord(c->ev[4], c->stream));
    if (!fused) phase_postforce_t<R>(c);
    if (zsplit) phase_prefinal_split<R>(c, fused ? &pre : nullptr);
    else phase_prefinal_t<R>(c, fused ? &pre : nullptr);
    HIPCHK(hipEventRecord(c->ev[5], c->stream));
    const bool capture = c->fourier_capture && it == nsteps - 1;
    if (capture) capture_spectrum(c, 0, c->fields<R>().phihat.p);   // in_fourier: ft_phi before the update
    const bool next = next_ok && it + 1 < nsteps;
    if (zsplit) phase_final_split<R>(c, !fc, next);
    else phase_final_t<R>(c, !fc);
    HIPCHK(hipEventRecord(c->ev[6], c->stream));
    if (zsplit) phase_postfinal_split<R>(c, next);
    else phase_postfinal_t<R>(c);
    if (capture) {   // ft_chempot: the last step of a run stores the complete muhat
      capture_spectrum(c, 1, c->fields<R>().muhat.p);
      c->cap_valid = true;
    }
    HIPCHK(hipEventRecord(c->ev[7], c->stream));
    // next step's rebuild decision (Neighbor::decide) + NaN check of
    // phi(0,0,0) (src/integrate.cpp:190-201), read back together
    if (c->neigh_set) enqueue_decide(c);
    enqueue_flags_readback(c);
    HIPCHK(hipMemcpyAsync(c->h_flags + 2, c->fields<R>().phi.p, sizeof(R), hipMemcpyDeviceToHost, c->stream));
    if (c->touch) {   // the chain ends' positions after the step, read back with the flags
      HIPCHK(hipMemcpyAsync(c->h_ends, c->x.p, sizeof(double4), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(c->h_ends + 4, c->x.p + (c->nowned - 1), sizeof(double4), hipMemcpyDeviceToHost,
                            c->stream));
    }
    HIPCHK(hipEventRecord(c->ev[8], c->stream));
    // The next step's start_of_step (grad phi: the fields only, not the
    // rebuild decision) goes in before the host waits for the read-back, so
    // the device works through the host's turnaround.
    started = it + 1 < nsteps && !c->touch;   // a touching run may stop after any step
    if (started) {
      c->timestep++;
      HIPCHK(hipEventRecord(c->ev_pre[0], c->stream));
      phase_start(c);
      HIPCHK(hipEventRecord(c->ev_pre[1], c->stream));
    }
    sync(c);
    probe_collect();
    for (int k = 0; k < 7; k++) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
      c->tacc[k] += ms;
    }
    float tot = 0;
    HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[8]));
    c->tacc[7] += tot;
    if (started) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->ev_pre[0], c->ev_pre[1]));
      c->tacc[0] += ms;
      c->tacc[7] += ms;
    }
    R phi0;
    std::memcpy(&phi0, c->h_flags + 2, sizeof(R));
    if (std::isnan(phi0) && started) c->timestep = step;   // the next step did not happen
    if (std::isnan(phi0)) {
      c->grad_ready = false;
      throw std::runtime_error("NAN encountered in phi.");
    }
    try {
      check_step_flags(c, step);
    } catch (...) {
      if (started) c->timestep = step;
      c->grad_ready = false;
      throw;
    }
    if (c->touch && touch_check(c, step, it + 1)) break;
  }
}

void run_steps(phafd_hip_ctx *c, int nsteps) {
  require_setup(c);
  if (c->fp32) run_steps_t<float>(c, nsteps);
  else run_steps_t<double>(c, nsteps);
}

// flatten group chunks (src/group.cpp:31-207) into member lists
void build_members(phafd_hip_ctx *c, FixDef &f) {
  std::vector<int> mem, cs, ce, hs, he;
  std::vector<char> seen(c->nowned, 0);
  f.maxlen = 0;
  for (size_t k = 0; k < f.start.size(); k++) {
    if (f.start[k] < 0 || f.end[k] > c->nowned || f.end[k] < f.start[k])
      throw std::invalid_argument("fix group chunk out of range");
    if (f.end[k] == f.start[k]) continue;
    for (int j = f.start[k]; j < f.end[k]; j++) {
      if (seen[j] && (f.kind == FixDef::CHAIN || f.kind == FixDef::SEMIFLEX))
        throw std::invalid_argument("chain chunks overlap");
      seen[j] = 1;
      mem.push_back(j);
      cs.push_back(f.start[k]);
      ce.push_back(f.end[k]);
    }
    hs.push_back(f.start[k]);
    he.push_back(f.end[k]);
    f.maxlen = std::max(f.maxlen, f.end[k] - f.start[k]);
  }
  if (f.kind == FixDef::CHAIN && f.maxlen > 2048)
    throw std::invalid_argument("chains longer than 2048 beads are not supported");
  if (f.kind == FixDef::SEMIFLEX && f.maxlen > BR_MAXLEN)
    throw std::invalid_argument("bead-rod chains longer than 700 beads are not supported");
  f.nmem = (int)mem.size();
  f.nchain = (int)hs.size();
  f.members.alloc(mem.size() + 1);
  f.cstart.alloc(mem.size() + 1);
  f.cend.alloc(mem.size() + 1);
  f.chs.alloc(hs.size() + 1);
  f.che.alloc(he.size() + 1);
  if (!mem.empty()) {
    HIPCHK(hipMemcpy(f.members.p, mem.data(), sizeof(int) * mem.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.cstart.p, cs.data(), sizeof(int) * cs.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.cend.p, ce.data(), sizeof(int) * ce.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.chs.p, hs.data(), sizeof(int) * hs.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.che.p, he.data(), sizeof(int) * he.size(), hipMemcpyHostToDevice));
  }
}

template <class R>
R *field_ptr(phafd_hip_ctx *c, int which) {
  FieldSet<R> &F = c->fields<R>();
  switch (which) {
    case PHAFD_PHI: return F.phi.p;
    case PHAFD_MU: return F.mu.p;
    case PHAFD_GRADX: return F.grad[0].p;
    case PHAFD_GRADY: return F.grad[1].p;
    case PHAFD_GRADZ: return F.grad[2].p;
  }
  throw std::invalid_argument("unknown field");
}

// the host side always holds FP64; the FP32 form rounds on the way in
void field_in(phafd_hip_ctx *c, int which, const double *h) {
  if (!c->fp32) {
    HIPCHK(hipMemcpyAsync(field_ptr<double>(c, which), h, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
  } else {
    std::vector<float> t(c->N);
    for (size_t i = 0; i < c->N; i++) t[i] = (float)h[i];
    HIPCHK(hipMemcpyAsync(field_ptr<float>(c, which), t.data(), sizeof(float) * c->N, hipMemcpyHostToDevice,
                          c->stream));
    sync(c);
  }
  sync(c);
}
void field_out(phafd_hip_ctx *c, int which, double *h) {
  if (!c->fp32) {
    HIPCHK(hipMemcpyAsync(h, field_ptr<double>(c, which), sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    sync(c);
  } else {
    std::vector<float> t(c->N);
    HIPCHK(hipMemcpyAsync(t.data(), field_ptr<float>(c, which), sizeof(float) * c->N, hipMemcpyDeviceToHost,
                          c->stream));
    sync(c);
    for (size_t i = 0; i < c->N; i++) h[i] = t[i];
  }
}

std::vector<double4> to_d4(const double *xyz, int n) {
  std::vector<double4> v(n);
  for (int i = 0; i < n; i++) v[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
  return v;
}
void from_d4(const std::vector<double4> &v, double *xyz) {
  for (size_t i = 0; i < v.size(); i++) {
    xyz[3 * i] = v[i].x;
    xyz[3 * i + 1] = v[i].y;
    xyz[3 * i + 2] = v[i].z;
  }
}
template <class T>
std::vector<T> download(phafd_hip_ctx *c, const T *d, size_t n) {
  std::vector<T> h(n);
  if (n) {
    HIPCHK(hipMemcpyAsync(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
  }
  return h;
}

}  // namespace

namespace {

// where a batch member's state lives: the batch's stream, its slab of fields
// (member `index` at index * N reals / index * M spectra) and its FH counters
struct MemberSlot {
  hipStream_t stream;
  FieldSet<double> *slab;
  int index;
  unsigned long long *counters;
};

// set up a context for one grid (slot: as member `index` of a batch)
void init_ctx(phafd_hip_ctx *c, const phafd_hip_desc *desc, const MemberSlot *slot) {
    if (desc->precision != PHAFD_FP64 && desc->precision != PHAFD_FP32)
      throw std::invalid_argument("precision must be PHAFD_FP64 or PHAFD_FP32");
    c->fp32 = desc->precision == PHAFD_FP32;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (desc->device < 0 || desc->device >= ndev) throw std::invalid_argument("no such HIP device");
    HIPCHK(hipSetDevice(desc->device));
    c->plan.init(desc->nx, desc->ny, desc->nz, desc->precision == PHAFD_FP32);
    c->nx = desc->nx; c->ny = desc->ny; c->nz = desc->nz;
    for (int d = 0; d < 3; d++) {
      if (!(desc->period[d] > 0)) throw std::invalid_argument("need to set box dimensions.");
      c->boxlo[d] = desc->boxlo[d];
      c->period[d] = desc->period[d];
      c->boxhi[d] = c->boxlo[d] + c->period[d];   // src/domain.cpp:71-72
    }
    // Domain::set_subbox, src/domain.cpp:78-99 (local0start = 0)
    const double dzz = c->period[2] / c->nz;
    c->sublo[0] = c->boxlo[0]; c->sublo[1] = c->boxlo[1]; c->sublo[2] = dzz * 0 + c->boxlo[2];
    c->subhi[0] = c->boxhi[0]; c->subhi[1] = c->boxhi[1]; c->subhi[2] = dzz * (0 + c->nz) + c->boxlo[2];
    c->zghost = desc->zghost;
    c->N = c->plan.N();
    c->M = c->plan.M();
    if (slot) {   // member of a batch: its stream, fields in the batch slab, counters in the batch's array
      if (c->fp32) throw std::invalid_argument("batched replicas are FP64");
      c->stream = slot->stream;
      c->own_stream = false;
      FieldSet<double> &F = c->fd;
      const size_t r = (size_t)slot->index;
      F.phi.wrap(slot->slab->phi.p + r * c->N, c->N);
      F.mu.wrap(slot->slab->mu.p + r * c->N, c->N);
      for (int d = 0; d < 3; d++) {
        F.grad[d].wrap(slot->slab->grad[d].p + r * c->N, c->N);
        F.ghat[d].wrap(slot->slab->ghat[d].p + r * c->M, c->M);
      }
      F.phihat.wrap(slot->slab->phihat.p + r * c->M, c->M);
      F.muhat.wrap(slot->slab->muhat.p + r * c->M, c->M);
      c->counters.wrap(slot->counters + 2 * r, 2);
    } else {
      HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      if (c->fp32) c->fs.alloc(c->N, c->M);
      else c->fd.alloc(c->N, c->M);
      c->counters.alloc(2);
    }
    c->mass64.alloc(1);
    c->flags.alloc(8);
    c->esum.alloc(8);
    HIPCHK(hipHostMalloc((void **)&c->h_flags, 128, hipHostMallocDefault));
    for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
    for (auto &e : c->ev_pre) HIPCHK(hipEventCreate(&e));
    for (auto &set : c->evp)
      for (auto &e : set) HIPCHK(hipEventCreate(&e));
    for (int w = PHAFD_PHI; w <= PHAFD_GRADZ; w++)
      fill_zero(c->fp32 ? (void *)field_ptr<float>(c, w) : (void *)field_ptr<double>(c, w),
                c->real_bytes() * c->N, c->stream);
    fill_zero(c->counters.p, 16, c->stream);
    fill_zero(c->flags.p, 16, c->stream);
  sync(c);
}

}  // namespace

// =================================================================== C ABI
namespace phafd {
// the message phafd_hip_last_error(NULL) reports on this thread
void set_stateless_error(const std::string &msg) { g_create_err = msg; }
}  // namespace phafd

extern "C" {

int phafd_hip_create(const phafd_hip_desc *desc, phafd_hip_ctx **out) {
  if (!desc || !out) return PHAFD_E_ARG;
  *out = nullptr;
  auto c = std::make_unique<phafd_hip_ctx>();
  c->device = desc->device;
  int rc = guard(c.get(), [&] {
    init_ctx(c.get(), desc, nullptr);
    return 0;
  });
  if (rc != 0) {
    g_create_err = c->err;   // reported by phafd_hip_last_error(NULL)
    return rc;
  }
  *out = c.release();
  return 0;
}

int phafd_hip_destroy(phafd_hip_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
  for (auto &e : c->ev_pre) if (e) (void)hipEventDestroy(e);
  for (auto &set : c->evp)
    for (auto &e : set) if (e) (void)hipEventDestroy(e);
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  if (c->h_ends) (void)hipHostFree(c->h_ends);
  for (auto &d : c->dump) {
    if (d.th.joinable()) d.th.join();
    if (d.cs) (void)hipStreamSynchronize(d.cs);
    if (d.host) (void)hipHostFree(d.host);
    if (d.snap) (void)hipEventDestroy(d.snap);
    if (d.done) (void)hipEventDestroy(d.done);
    if (d.cs) (void)hipStreamDestroy(d.cs);
  }
  if (c->stream && c->own_stream) {
    scan_forget_stream(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
  return 0;
}

const char *phafd_hip_last_error(const phafd_hip_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int phafd_hip_set_field(phafd_hip_ctx *c, int which, const double *h) {
  return guard(c, [&] {
    field_in(c, which, h);
    return 0;
  });
}
int phafd_hip_get_field(phafd_hip_ctx *c, int which, double *h) {
  return guard(c, [&] {
    field_out(c, which, h);
    return 0;
  });
}
// rows ld doubles apart on the host (the reference's padded fftw rows)
int phafd_hip_set_field_ld(phafd_hip_ctx *c, int which, const double *h, int64_t ld) {
  return guard(c, [&] {
    if (!h || ld < c->nx) throw std::invalid_argument("set_field_ld: need a buffer and ld >= Nx");
    const size_t rows = (size_t)c->ny * c->nz;
    std::vector<double> t(c->N);
    for (size_t r = 0; r < rows; r++) std::memcpy(&t[r * c->nx], h + r * (size_t)ld, sizeof(double) * c->nx);
    field_in(c, which, t.data());
    return 0;
  });
}
int phafd_hip_get_field_ld(phafd_hip_ctx *c, int which, double *h, int64_t ld) {
  return guard(c, [&] {
    if (!h || ld < c->nx) throw std::invalid_argument("get_field_ld: need a buffer and ld >= Nx");
    const size_t rows = (size_t)c->ny * c->nz;
    std::vector<double> t(c->N);
    field_out(c, which, t.data());
    for (size_t r = 0; r < rows; r++) std::memcpy(h + r * (size_t)ld, &t[r * c->nx], sizeof(double) * c->nx);
    return 0;
  });
}
int phafd_hip_get_spectrum(phafd_hip_ctx *c, int which, double *h) {
  return guard(c, [&] {
    if (which != PHAFD_PHIHAT && which != PHAFD_MUHAT) throw std::invalid_argument("unknown spectrum");
    if (!c->fp32) {
      FieldSet<double> &F = c->fd;
      F.canon_tmp.alloc(c->M);
      spec_k_to_canon(c->plan, which == PHAFD_PHIHAT ? F.phihat.p : F.muhat.p, F.canon_tmp.p, c->stream);
      HIPCHK(hipMemcpyAsync(h, F.canon_tmp.p, sizeof(double2) * c->M, hipMemcpyDeviceToHost, c->stream));
      sync(c);
    } else {
      FieldSet<float> &F = c->fs;
      F.canon_tmp.alloc(c->M);
      spec_k_to_canon(c->plan, which == PHAFD_PHIHAT ? F.phihat.p : F.muhat.p, F.canon_tmp.p, c->stream);
      std::vector<float2> t(c->M);
      HIPCHK(hipMemcpyAsync(t.data(), F.canon_tmp.p, sizeof(float2) * c->M, hipMemcpyDeviceToHost, c->stream));
      sync(c);
      for (size_t i = 0; i < c->M; i++) { h[2 * i] = t[i].x; h[2 * i + 1] = t[i].y; }
    }
    return 0;
  });
}

int phafd_hip_set_atoms(phafd_hip_ctx *c, int n, const double *xyz, const int *type, const int *tag,
                        const int *mol) {
  return guard(c, [&] {
    if (n < 0 || (n > 0 && (!xyz || !type))) throw std::invalid_argument("bad atom arrays");
    c->nowned = n;
    c->h_type.assign(type, type + n);
    c->h_tag.resize(n);
    c->h_mol.assign(n, 0);
    int tmax = -1;
    for (int i = 0; i < n; i++) {
      c->h_tag[i] = tag ? tag[i] : i;
      if (mol) c->h_mol[i] = mol[i];
      if (type[i] < 0) throw std::invalid_argument("negative atom type");
      tmax = std::max(tmax, type[i]);
    }
    std::set<int> tset(c->h_tag.begin(), c->h_tag.end());
    if ((int)tset.size() != n) throw std::runtime_error("duplicate IDs found on processor(s).");
    c->ntypes = tmax + 1;   // src/atom.cpp:190-203
    c->x.alloc(n + 1); c->ux.alloc(n + 1); c->F.alloc(n + 1); c->xhold.alloc(n + 1);
    c->image.alloc(n + 1); c->type.alloc(n + 1);
    if (n) {
      std::vector<double4> v = to_d4(xyz, n);
      HIPCHK(hipMemcpy(c->x.p, v.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->xhold.p, v.data(), sizeof(double4) * n, hipMemcpyHostToDevice));
      std::vector<int> img(n, (512 << 20) | (512 << 10) | 512);   // Domain::set_image
      HIPCHK(hipMemcpy(c->image.p, img.data(), sizeof(int) * n, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->type.p, type, sizeof(int) * n, hipMemcpyHostToDevice));
      launch_fill_d4(n, c->F.p, c->stream);
      launch_fill_d4(n, c->ux.p, c->stream);
      sync(c);
    }
    c->gcap = 0;
    c->ngathered = 0;
    c->atoms_set = true;
    c->setup_done = false;
    c->lists_valid = false;
    c->pending_decision = -1;
    return 0;
  });
}
int phafd_hip_set_images(phafd_hip_ctx *c, const int *image) {
  return guard(c, [&] {
    if (!c->atoms_set) throw std::runtime_error("set the atoms before their images");
    if (c->nowned && !image) throw std::invalid_argument("null image array");
    if (c->nowned) HIPCHK(hipMemcpy(c->image.p, image, sizeof(int) * c->nowned, hipMemcpyHostToDevice));
    c->setup_done = false;
    c->lists_valid = false;
    return 0;
  });
}
int phafd_hip_get_atoms(phafd_hip_ctx *c, double *xyz, double *uxyz, int *image) {
  return guard(c, [&] {
    if (xyz) from_d4(download(c, c->x.p, c->nowned), xyz);
    if (uxyz) from_d4(download(c, c->ux.p, c->nowned), uxyz);
    if (image && c->nowned) {
      auto v = download(c, c->image.p, c->nowned);
      std::memcpy(image, v.data(), sizeof(int) * c->nowned);
    }
    return 0;
  });
}
int phafd_hip_get_forces(phafd_hip_ctx *c, double *f) {
  return guard(c, [&] {
    from_d4(download(c, c->F.p, c->nowned), f);
    return 0;
  });
}

}  // extern "C"

namespace {

__global__ void k_widen(size_t n, const float *__restrict__ a, double *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// grid attributes of the dump and read_dump commands (src/dump.cpp:969-983,
// src/read_dump.cpp:170-185)
int grid_attribute(const std::string &w) {
  if (w == "phi") return PHAFD_PHI;
  if (w == "chempot") return PHAFD_MU;
  if (w == "gradphi_x") return PHAFD_GRADX;
  if (w == "gradphi_y") return PHAFD_GRADY;
  if (w == "gradphi_z") return PHAFD_GRADZ;
  throw std::runtime_error("Dump error: Attribute does not exist.");
}

std::vector<std::string> attribute_list(int nattr, const char *const *names, const char *what) {
  if (nattr <= 0 || !names) throw std::invalid_argument(std::string("Need attributes to ") + what + " dump files.");
  std::vector<std::string> v;
  for (int k = 0; k < nattr; k++) {
    if (!names[k]) throw std::invalid_argument("null attribute name");
    v.emplace_back(names[k]);
  }
  return v;
}

// the previous dump of the job's kind finished; its error, if any, is this
// call's
void dump_wait_job(DumpJob &d) {
  if (d.th.joinable()) d.th.join();
  if (!d.err.empty()) {
    std::string e;
    e.swap(d.err);
    throw std::runtime_error(e);
  }
}
void dump_wait(phafd_hip_ctx *c) {
  std::string err;
  for (auto &d : c->dump) {
    try {
      dump_wait_job(d);
    } catch (const std::exception &e) {
      if (err.empty()) err = e.what();
    }
  }
  if (!err.empty()) throw std::runtime_error(err);
}

void dump_prepare(DumpJob &d, size_t bytes) {
  if (!d.cs) {
    HIPCHK(hipStreamCreateWithFlags(&d.cs, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&d.snap, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&d.done, hipEventDisableTiming));
  }
  if (d.host_bytes < bytes) {
    if (d.host) HIPCHK(hipHostFree(d.host));
    d.host = nullptr;
    d.host_bytes = 0;
    HIPCHK(hipHostMalloc((void **)&d.host, bytes, hipHostMallocDefault));
    d.host_bytes = bytes;
  }
  d.stage.alloc((bytes + sizeof(double) - 1) / sizeof(double));
}

// stage -> pinned host on the dump stream, after what the step stream has
// enqueued so far (the snapshot); then the write, on the writer thread or here
template <class W>
void dump_launch(phafd_hip_ctx *c, DumpJob &d, size_t bytes, bool async, W &&write) {
  HIPCHK(hipEventRecord(d.snap, c->stream));
  HIPCHK(hipStreamWaitEvent(d.cs, d.snap, 0));
  HIPCHK(hipMemcpyAsync(d.host, d.stage.p, bytes, hipMemcpyDeviceToHost, d.cs));
  HIPCHK(hipEventRecord(d.done, d.cs));
  auto job = [&d, w = std::forward<W>(write)]() mutable {
    try {
      HIPCHK(hipEventSynchronize(d.done));
      w(static_cast<const char *>(d.host));
    } catch (const std::exception &e) {
      d.err = e.what();
    }
  };
  if (async) {
    d.th = std::thread(std::move(job));
  } else {
    job();
    dump_wait_job(d);
  }
}

}  // namespace

extern "C" {

int phafd_hip_dump_grid(phafd_hip_ctx *c, const char *path, int nattr, const char *const *names, int async) {
  return guard(c, [&] {
    DumpJob &job = c->dump[0];
    dump_wait_job(job);
    if (!path) throw std::invalid_argument("null file name");
    auto at = attribute_list(nattr, names, "write");
    std::vector<int> ids;
    for (auto &w : at) ids.push_back(grid_attribute(w));
    const size_t N = c->N, bytes = (size_t)nattr * N * sizeof(double);
    // the format's limits, checked before anything is enqueued (the writer
    // checks them again): a 32-bit byte length and int offsets
    if (N * sizeof(double) > (size_t)UINT32_MAX) throw std::invalid_argument("grid too large for a 32-bit VTK byte length");
    if ((size_t)(nattr - 1) * (N * sizeof(double) + sizeof(unsigned)) > (size_t)INT_MAX)
      throw std::invalid_argument("appended offsets exceed the format's int range");
    dump_prepare(job, bytes);
    for (int k = 0; k < nattr; k++) {
      double *dst = job.stage.p + (size_t)k * N;
      if (!c->fp32) {
        HIPCHK(hipMemcpyAsync(dst, field_ptr<double>(c, ids[k]), N * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
      } else {
        const int blocks = (int)std::min<size_t>((N + 255) / 256, 8192);
        hipLaunchKernelGGL(k_widen, dim3(blocks), dim3(256), 0, c->stream, N, field_ptr<float>(c, ids[k]), dst);
        HIPCHK(hipGetLastError());
      }
    }
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const std::array<double, 3> origin{c->boxlo[0], c->boxlo[1], c->boxlo[2]};
    const std::array<double, 3> spacing{c->dx(), c->dy(), c->dz()};
    dump_launch(c, job, bytes, async != 0, [=, file = std::string(path)](const char *host) {
      std::vector<const char *> nm;
      std::vector<const double *> arr;
      for (int k = 0; k < nattr; k++) {
        nm.push_back(at[k].c_str());
        arr.push_back(reinterpret_cast<const double *>(host) + (size_t)k * N);
      }
      if (phafd_vtk_write_grid(file.c_str(), nx, ny, nz, origin.data(), spacing.data(), nattr, nm.data(),
                               arr.data()) < 0)
        throw std::runtime_error(phafd_host_last_error());
    });
    return 0;
  });
}

int phafd_hip_dump_atoms(phafd_hip_ctx *c, const char *path, int nattr, const char *const *names, int precision,
                         int fmt, int async) {
  return guard(c, [&] {
    DumpJob &job = c->dump[1];
    dump_wait_job(job);
    if (!path) throw std::invalid_argument("null file name");
    if (fmt != PHAFD_VTK_FMT_LF && fmt != PHAFD_VTK_FMT_E)
      throw std::invalid_argument("Dump fmt must be either lf or e.");
    auto at = attribute_list(nattr, names, "write");
    bool has_x = false;
    for (auto &w : at) {
      if (w.rfind("c_", 0) == 0)
        throw std::runtime_error("compute attributes are written with phafd_vtk_write_atoms from the compute's array");
      if (w != "x" && w != "F" && w != "ux" && w != "ix" && w != "ID" && w != "molID" && w != "type" && w != "label")
        throw std::runtime_error("Invalid per atom quantity in dump file.");
      has_x |= (w == "x");
    }
    if (!has_x) throw std::runtime_error("Atom dump must contain positions.");
    // device arrays in the stage: x, ux, F (double4 each), image (int)
    const size_t n = (size_t)c->nowned, v4 = n * sizeof(double4);
    const size_t bytes = 3 * v4 + n * sizeof(int) + 1;
    dump_prepare(job, bytes);
    char *st = reinterpret_cast<char *>(job.stage.p);
    if (n) {
      HIPCHK(hipMemcpyAsync(st, c->x.p, v4, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(st + v4, c->ux.p, v4, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(st + 2 * v4, c->F.p, v4, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(st + 3 * v4, c->image.p, n * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    }
    dump_launch(c, job, bytes, async != 0,
                [=, file = std::string(path), tag = c->h_tag, mol = c->h_mol, type = c->h_type](const char *host) {
                  auto xyz = [&](int which) {
                    std::vector<double> v(3 * n);
                    const double4 *s = reinterpret_cast<const double4 *>(host + which * v4);
                    for (size_t i = 0; i < n; i++) {
                      v[3 * i] = s[i].x; v[3 * i + 1] = s[i].y; v[3 * i + 2] = s[i].z;
                    }
                    return v;
                  };
                  std::vector<std::vector<double>> vec;
                  std::vector<int> img(n), label(n, 0);   // every owned atom is OWNED (0)
                  if (n) std::memcpy(img.data(), host + 3 * v4, n * sizeof(int));
                  std::vector<const char *> nm;
                  std::vector<const void *> arr;
                  vec.reserve(nattr);
                  for (auto &w : at) {
                    nm.push_back(w.c_str());
                    if (w == "x") { vec.push_back(xyz(0)); arr.push_back(vec.back().data()); }
                    else if (w == "ux") { vec.push_back(xyz(1)); arr.push_back(vec.back().data()); }
                    else if (w == "F") { vec.push_back(xyz(2)); arr.push_back(vec.back().data()); }
                    else if (w == "ix") arr.push_back(img.data());
                    else if (w == "ID") arr.push_back(tag.data());
                    else if (w == "molID") arr.push_back(mol.data());
                    else if (w == "type") arr.push_back(type.data());
                    else arr.push_back(label.data());
                  }
                  if (phafd_vtk_write_atoms(file.c_str(), (int)n, nattr, nm.data(), arr.data(), precision, fmt) < 0)
                    throw std::runtime_error(phafd_host_last_error());
                });
    return 0;
  });
}

int phafd_hip_compute_pair(phafd_hip_ctx *c, int ipair, double *f) {
  return guard(c, [&] {
    if (ipair < 0 || ipair >= (int)c->pairs.size())
      throw std::invalid_argument("Cannot compute pair style, it hasn't been created!");
    if (!c->setup_done) throw std::runtime_error("compute pair needs a set-up system");
    if (!f) throw std::invalid_argument("null force array");
    PairDef &p = *c->pairs[ipair];
    const int n = c->nowned;
    c->cp_keep.alloc(n + 1);
    launch_copy4(n, c->F.p, c->cp_keep.p, c->stream);
    launch_fill_d4(c->ngathered, c->gF.p, c->stream);
    launch_fill_d4(n, c->F.p, c->stream);
    if (c->neigh_set && c->lists_valid) {
      if (p.style == 2) {
        LJParams lj{p.lj1, p.lj2, p.ljcutsq};
        launch_lj_force(c->ngathered, lj, c->gx.p, c->h_rowof.p, c->h_off.p, c->h_neigh.p, c->t_off.p,
                        c->t_pair.p, c->t_i.p, n, c->gF.p, nullptr, c->stream);
      } else if (p.style == 4) {
        launch_harmonic_force(c->ngathered, HarmParams{p.hk, p.hcut}, c->gx.p, c->h_rowof.p, c->h_off.p,
                              c->h_neigh.p, c->t_off.p, c->t_pair.p, c->t_i.p, n, c->gF.p, c->stream);
      } else if (c->fp32) {
        grid_atom_pass<float>(c, p, nullptr, false);   // adds its deposit to mu
      } else {
        grid_atom_pass<double>(c, p, nullptr, false);
      }
    }
    reverse_comm(c);
    from_d4(download(c, c->F.p, n), f);
    launch_copy4(n, c->cp_keep.p, c->F.p, c->stream);
    sync(c);
    return 0;
  });
}

}  // extern "C"

namespace {

// ComputeComplex::in_fourier (src/compute_complex.cpp:54-91) on a captured K
// layout spectrum [kx][kz][ky]: out in the reference's ft array order
// [ky][kz][kx] (its arrays are built (Nx, Nz, Ny), src/grid.cpp:214-224).
// One block per (32 ky x 32 kx) tile of one kz plane, transposed in LDS so
// that both the reads (ky runs) and the writes (kx runs) are coalesced.
__global__ void __launch_bounds__(256) k_complex_out(const double2 *__restrict__ hat, int hx, int ny, int nz,
                                                     int quant, double prefac, double *__restrict__ out) {
  __shared__ double2 tile[32][33];
  const int ky0 = blockIdx.x * 32, kx0 = blockIdx.y * 32, kz = blockIdx.z;
  const int tx = threadIdx.x, ty = threadIdx.y;   // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int kx = kx0 + r, ky = ky0 + tx;
    if (kx < hx && ky < ny) tile[r][tx] = hat[((size_t)kx * nz + kz) * ny + ky];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int ky = ky0 + r, kx = kx0 + tx;
    if (kx < hx && ky < ny) {
      const double2 z = tile[tx][r];
      double v;
      if (quant == PHAFD_CPX_MODULUS) v = hypot(z.x, z.y) * prefac;
      else if (quant == PHAFD_CPX_NORM) v = (z.x * z.x + z.y * z.y) * prefac * prefac;
      else if (quant == PHAFD_CPX_REAL) v = z.x * prefac;
      else v = z.y * prefac;
      out[((size_t)ky * nz + kz) * hx + kx] = v;
    }
  }
}

// FixGridAve::end_of_step (src/fixgrid_ave.cpp:170-215): the first sample of
// a window starts from zero, the last divides by repeat
__global__ void k_grid_ave(size_t n, const double *__restrict__ x, double *__restrict__ acc, int first, int last,
                           double repeat) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double a = first ? 0.0 : acc[i];
    a += x[i];
    if (last) a /= repeat;
    acc[i] = a;
  }
}

void complex_to(phafd_hip_ctx *c, int which, int quant, double *dst) {
  if (which != PHAFD_PHIHAT && which != PHAFD_MUHAT) throw std::invalid_argument("Invalid array to compute in compute modulus.");
  if (quant < PHAFD_CPX_REAL || quant > PHAFD_CPX_NORM) throw std::invalid_argument("Invalid argument in compute complex.");
  if (!c->cap_valid) throw std::runtime_error("no captured spectra: enable phafd_hip_fourier_capture before the run");
  if (c->nz > 65535) throw std::invalid_argument("compute complex: Nz above the launch's grid limit");
  const int hx = c->nx / 2 + 1;
  const double prefac = 1. / (c->dx() * c->dy() * c->dz());
  dim3 grid((c->ny + 31) / 32, (hx + 31) / 32, c->nz);
  hipLaunchKernelGGL(k_complex_out, grid, dim3(32, 8), 0, c->stream, c->cap_hat[which].p, hx, c->ny, c->nz, quant,
                     prefac, dst);
  HIPCHK(hipGetLastError());
}

}  // namespace

extern "C" {

int phafd_hip_fourier_capture(phafd_hip_ctx *c, int on) {
  return guard(c, [&] {
    c->fourier_capture = on != 0;
    return 0;
  });
}

int phafd_hip_compute_complex(phafd_hip_ctx *c, int which, int quant, double *out) {
  return guard(c, [&] {
    if (!out) throw std::invalid_argument("null output");
    c->cc_out.alloc(c->M);
    complex_to(c, which, quant, c->cc_out.p);
    HIPCHK(hipMemcpyAsync(out, c->cc_out.p, sizeof(double) * c->M, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return 0;
  });
}

int phafd_hip_grid_ave(phafd_hip_ctx *c, int slot, int which, int quant, int first, int last, int repeat,
                       double *out) {
  return guard(c, [&] {
    if (slot < 0 || slot >= 8) throw std::invalid_argument("grid/ave slot out of range (0..7)");
    if (repeat <= 0) throw std::invalid_argument("Every, repeat, and freq must all be >= 0.");
    c->cc_out.alloc(c->M);
    c->ave[slot].alloc(c->M);
    complex_to(c, which, quant, c->cc_out.p);
    const int blocks = (int)std::min<size_t>((c->M + 255) / 256, 8192);
    hipLaunchKernelGGL(k_grid_ave, dim3(blocks), dim3(256), 0, c->stream, c->M, c->cc_out.p, c->ave[slot].p,
                       first != 0, last != 0, (double)repeat);
    HIPCHK(hipGetLastError());
    if (out) HIPCHK(hipMemcpyAsync(out, c->ave[slot].p, sizeof(double) * c->M, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return 0;
  });
}

int phafd_hip_compute_grid_clusters(phafd_hip_ctx *c, double threshold, int condition, double *out) {
  return guard(c, [&] {
    if (condition != PHAFD_CLUSTER_LT && condition != PHAFD_CLUSTER_GT)
      throw std::invalid_argument("Must specify lt or gt for compute grid cluster.");
    if (!out) throw std::invalid_argument("null output");
    if (c->N > (size_t)INT_MAX) throw std::invalid_argument("grid too large for int cluster ids");
    c->cl_parent.alloc(c->N);
    c->cl_out.alloc(c->N);
    const bool gt = condition == PHAFD_CLUSTER_GT;
    if (c->fp32)
      launch_grid_clusters<float>(c->nx, c->ny, c->nz, c->fs.phi.p, threshold, gt, c->cl_parent.p, c->cl_out.p,
                                  c->stream);
    else
      launch_grid_clusters<double>(c->nx, c->ny, c->nz, c->fd.phi.p, threshold, gt, c->cl_parent.p, c->cl_out.p,
                                   c->stream);
    HIPCHK(hipMemcpyAsync(out, c->cl_out.p, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return 0;
  });
}

int phafd_hip_dump_wait(phafd_hip_ctx *c) {
  return guard(c, [&] {
    dump_wait(c);
    return 0;
  });
}

int phafd_hip_read_dump_grid(phafd_hip_ctx *c, const char *path, int nattr, const char *const *names,
                             int no_padding) {
  return guard(c, [&] {
    dump_wait(c);
    auto at = attribute_list(nattr, names, "read");
    std::vector<int> ids;
    for (auto &w : at) {
      if (w.rfind("c_", 0) == 0 || w.rfind("f_", 0) == 0) throw std::runtime_error("Cannot read computes from dump");
      ids.push_back(grid_attribute(w));
    }
    std::vector<std::vector<double>> buf(nattr, std::vector<double>(c->N));
    std::vector<double *> out;
    for (auto &b : buf) out.push_back(b.data());
    if (phafd_vtk_read_grid(path, c->nx, c->ny, c->nz, nattr, names, out.data(), no_padding) < 0)
      throw std::runtime_error(phafd_host_last_error());
    for (int k = 0; k < nattr; k++) field_in(c, ids[k], buf[k].data());
    return 0;
  });
}

int phafd_hip_pair_gridatom(phafd_hip_ctx *c, int style, double cut) {
  return guard(c, [&] {
    if (c->ntypes <= 0) throw std::runtime_error("Cannot create pair style before atoms are made.");
    if (style != PHAFD_PAIR_GAUSSIAN && style != PHAFD_PAIR_LJISH && style != PHAFD_PAIR_LJISH_LINEAR)
      throw std::invalid_argument("Unrecognised pair style.");
    auto p = std::make_unique<PairDef>();
    p->style = style;
    p->maxcut = cut;
    p->eps.assign(c->ntypes, 1.0);      // defaults, src/pair_gridatom_gaussian.cpp:134-139
    p->phit.assign(c->ntypes, 0.1);
    p->cutsq.assign(c->ntypes, cut * cut);
    c->pairs.push_back(std::move(p));
    c->need_full = true;
    return (int)c->pairs.size() - 1;
  });
}
int phafd_hip_pair_coeff(phafd_hip_ctx *c, int ipair, int type, double eps, double phit, double nucwidth) {
  return guard(c, [&] {
    if (ipair < 0 || ipair >= (int)c->pairs.size()) throw std::invalid_argument("no such pair");
    PairDef &p = *c->pairs[ipair];
    if (bead_pair(p)) return 0;   // PairLJCut / PairHarmonicCut::coeff are empty (src/pair_lj_cut.cpp:119-125,
                                  // src/pair_harmonic_cut.cpp:127-130)
    if (type < 0 || type >= (int)p.eps.size()) throw std::runtime_error("Pair grid/atom atom type out of range.");
    if (!std::isnan(eps)) p.eps[type] = eps;
    if (!std::isnan(phit)) p.phit[type] = phit;
    if (!std::isnan(nucwidth)) {
      if (nucwidth > p.maxcut) throw std::runtime_error("Pair grid/atom nucwidth cannot be larger than maximum cut.");
      p.cutsq[type] = nucwidth * nucwidth;
    }
    return 0;
  });
}
int phafd_hip_pair_ljcut(phafd_hip_ctx *c, double cut, double eps, double sigma) {
  return guard(c, [&] {
    if (c->ntypes <= 0) throw std::runtime_error("Cannot create pair style before atoms are made.");
    auto p = std::make_unique<PairDef>();
    p->style = 2;
    p->maxcut = cut;
    p->lj1 = 48 * eps * pow(sigma, 12.0);
    p->lj2 = 24 * eps * pow(sigma, 6.0);
    p->ljcutsq = cut * cut;
    c->pairs.push_back(std::move(p));
    c->need_half = true;
    return (int)c->pairs.size() - 1;
  });
}

int phafd_hip_pair_harmoniccut(phafd_hip_ctx *c, double cut, double springK) {
  return guard(c, [&] {
    if (c->ntypes <= 0) throw std::runtime_error("Cannot create pair style before atoms are made.");
    auto p = std::make_unique<PairDef>();
    p->style = 4;
    p->maxcut = cut;
    p->hcut = cut;
    p->hk = springK;
    c->pairs.push_back(std::move(p));
    c->need_half = true;
    return (int)c->pairs.size() - 1;
  });
}

int phafd_hip_fix_conjugate(phafd_hip_ctx *c, int seed, double mobility, double temp, double volFH,
                            double gamma, int noise_mode) {
  return guard(c, [&] {
    if (noise_mode != PHAFD_NOISE_OFF && noise_mode != PHAFD_NOISE_FIELD && noise_mode != PHAFD_NOISE_DEVICE)
      throw std::invalid_argument("noise mode must be off, field or device");
    auto f = std::make_unique<FixDef>();
    f->kind = FixDef::CONJ;
    f->seed = seed;
    f->noise_key = (unsigned long long)(unsigned)phafd_make_unique_seed(seed);
    f->mobility = mobility; f->temp = temp; f->volFH = volFH; f->gamma = gamma;
    f->noise_mode = noise_mode;
    c->fixes.push_back(std::move(f));
    return 0;
  });
}
}  // extern "C"
namespace {
// queue one host field; a consumed entry's device buffer is reused
template <class R>
void noise_push(phafd_hip_ctx *c, const double *xi) {
  using C = typename Vec2<R>::T;
  FieldSet<R> &F = c->fields<R>();
  DBuf<C> buf;
  if (c->noise_head > 0) {
    buf = std::move(F.noise_q.front());
    F.noise_q.erase(F.noise_q.begin());
    c->noise_head--;
  }
  buf.alloc(c->M);
  F.canon_tmp.alloc(c->M);
  std::vector<C> t(c->M);
  for (size_t i = 0; i < c->M; i++) {
    t[i].x = (R)xi[2 * i];
    t[i].y = (R)xi[2 * i + 1];
  }
  HIPCHK(hipMemcpyAsync(F.canon_tmp.p, t.data(), sizeof(C) * c->M, hipMemcpyHostToDevice, c->stream));
  spec_canon_to_k(c->plan, F.canon_tmp.p, buf.p, c->stream);
  sync(c);
  F.noise_q.push_back(std::move(buf));
}
}  // namespace
extern "C" {
int phafd_hip_set_noise(phafd_hip_ctx *c, const double *xi) {
  return guard(c, [&] {
    sync(c);   // no update still reads a queued field
    if (!xi) {
      c->fd.noise_q.clear();
      c->fs.noise_q.clear();
      c->noise_head = 0;
      return 0;
    }
    if (c->fp32) noise_push<float>(c, xi);
    else noise_push<double>(c, xi);
    return 0;
  });
}
int phafd_hip_noise_queued(phafd_hip_ctx *c, int64_t *n) {
  return guard(c, [&] {
    if (n) *n = (int64_t)noise_queued(c);
    return 0;
  });
}
int phafd_hip_fix_floryhuggins(phafd_hip_ctx *c, double temp, double volFH, double chi) {
  return guard(c, [&] {
    auto f = std::make_unique<FixDef>();
    f->kind = FixDef::FH;
    f->temp = temp; f->volFH = volFH; f->chi = chi;
    c->fixes.push_back(std::move(f));
    return 0;
  });
}
int phafd_hip_fix_gradphi(phafd_hip_ctx *c, int once) {
  return guard(c, [&] {
    auto f = std::make_unique<FixDef>();
    f->kind = FixDef::GRADPHI;
    f->once = once;
    c->fixes.push_back(std::move(f));
    return 0;
  });
}
int phafd_hip_fix_drag(phafd_hip_ctx *c, int nchunk, const int *start, const int *end, double drag) {
  return guard(c, [&] {
    if (c->ntypes <= 0) throw std::runtime_error("Fix requires atoms to be created.");
    auto f = std::make_unique<FixDef>();
    f->kind = FixDef::DRAG;
    f->start.assign(start, start + nchunk);
    f->end.assign(end, end + nchunk);
    f->drag = drag;
    build_members(c, *f);
    c->fixes.push_back(std::move(f));
    return 0;
  });
}
int phafd_hip_fix_chain(phafd_hip_ctx *c, int nchunk, const int *start, const int *end, double r0,
                        double kbond, double kbend, double drag) {
  return guard(c, [&] {
    if (c->ntypes <= 0) throw std::runtime_error("Fix requires atoms to be created.");
    auto f = std::make_unique<FixDef>();
    f->kind = FixDef::CHAIN;
    f->start.assign(start, start + nchunk);
    f->end.assign(end, end + nchunk);
    f->r0 = r0; f->kbond = kbond; f->kbend = kbend; f->drag = drag;
    build_members(c, *f);
    c->fixes.push_back(std::move(f));
    return 0;
  });
}

int phafd_hip_fix_semiflexible(phafd_hip_ctx *c, int tether, int nchunk, const int *start, const int *end, int seed,
                               double bondlength, double zeta_para, double zeta_perp, double bending, double temp) {
  return guard(c, [&] {
    if (c->ntypes <= 0) throw std::runtime_error("Fix requires atoms to be created.");
    if (tether < 0 || tether > 2) throw std::invalid_argument("tether must be 0 (notether), 1 (singletether) or 2");
    if (!(bondlength > 0) || !(zeta_para > 0) || !(zeta_perp > 0) || !(temp >= 0))
      throw std::invalid_argument("bondlength, zeta_para, zeta_perp must be positive, temp non-negative");
    if (nchunk < 0 || (nchunk > 0 && (!start || !end))) throw std::invalid_argument("bad group chunks");
    auto f = std::make_unique<FixDef>();
    f->kind = FixDef::SEMIFLEX;
    f->tether = tether;
    f->start.assign(start, start + nchunk);
    f->end.assign(end, end + nchunk);
    f->bondlength = bondlength; f->zpar = zeta_para; f->zperp = zeta_perp; f->bending = bending; f->brtemp = temp;
    build_members(c, *f);
    // one seed per chunk, drawn as src/fixatom_semiflexible.cpp:44-80 does;
    // the empty chunks' draws are skipped with them
    std::mt19937 gen;
    std::uniform_int_distribution<int> integer_dist;
    gen.seed(phafd_make_unique_seed(seed));
    std::vector<int> seeds;
    for (int k = 0; k < nchunk; k++) {
      const int v = integer_dist(gen);
      if (end[k] > start[k]) seeds.push_back(v);
    }
    f->d_seed.alloc(seeds.size() + 1);
    if (!seeds.empty())
      HIPCHK(hipMemcpy(f->d_seed.p, seeds.data(), sizeof(int) * seeds.size(), hipMemcpyHostToDevice));
    f->br_R0.alloc(c->nowned + 1);
    f->br_FB.alloc(c->nowned + 1);
    f->br_T.alloc(c->nowned + 1);
    c->fixes.push_back(std::move(f));
    return 0;
  });
}

int phafd_hip_neighbor(phafd_hip_ctx *c, double skin) {
  return guard(c, [&] {
    if (c->pairs.empty() && c->ntypes > 0) throw std::runtime_error("Must call neighbor after pairs are given.");
    c->skin = skin;
    double cutoff = 0.0;
    for (auto &p : c->pairs) cutoff = std::max(cutoff, p->maxcut);
    c->triggersq = 0.25 * skin * skin;
    if (cutoff > 0.0) c->cutneigh = skin + cutoff;
    setup_bins(c);
    create_stencils(c);
    comm_setup(c);
    for (auto &p : c->pairs) if (!bead_pair(*p)) upload_pair(*p, c->skin);
    c->neigh_set = true;
    return 0;
  });
}

int phafd_hip_setup(phafd_hip_ctx *c, double dt) {
  return guard(c, [&] {
    if (dt <= 0.0) throw std::runtime_error("Cannot run simulation without setting dt > 0.");
    c->dt = dt;
    for (auto &p : c->pairs) if (!bead_pair(*p)) upload_pair(*p, c->skin);
    if (c->neigh_set && c->nowned > 0) {
      launch_pbc(c->nowned, c->x.p, c->ux.p, c->image.p, c->box(), c->stream);
      borders(c);
      neighbor_build(c);
    } else if (c->nowned > 0) {
      launch_unmap(c->nowned, c->x.p, c->ux.p, c->image.p, c->box(), c->stream);
    }
    if (c->nowned > 0) launch_fill_d4(c->nowned, c->F.p, c->stream);
    if (c->ngathered > 0) launch_fill_d4(c->ngathered, c->gF.p, c->stream);
    launch_fill_i32(1, c->flags.p + 3, 0, c->stream);
    c->pending_decision = -1;
    sync(c);
    // FixAtomSemiFlexible::setup -> check_bonds (src/fixatom_semiflexible.cpp:84-140)
    if (has_semiflex(c)) {
      std::vector<double4> u(c->nowned);
      if (c->nowned) HIPCHK(hipMemcpy(u.data(), c->ux.p, sizeof(double4) * c->nowned, hipMemcpyDeviceToHost));
      for (auto &fp : c->fixes) {
        if (fp->kind != FixDef::SEMIFLEX) continue;
        for (size_t k = 0; k < fp->start.size(); k++)
          for (int mu = fp->start[k] + 1; mu < fp->end[k]; mu++) {
            const double d0 = u[mu].x - u[mu - 1].x, d1 = u[mu].y - u[mu - 1].y, d2 = u[mu].z - u[mu - 1].z;
            if (std::fabs(std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) - fp->bondlength) > 1e-3)
              throw std::runtime_error("Adjacent atom displacements do not match the FixAtomSemiFlexible "
                                       "bondlength specified.");
          }
      }
    }
    c->setup_done = true;
    return 0;
  });
}

int phafd_hip_run(phafd_hip_ctx *c, int nsteps) {
  return guard(c, [&] {
    require_noise_fields(c, nsteps);
    run_steps(c, nsteps);
    return 0;
  });
}

int phafd_hip_run_until_touching(phafd_hip_ctx *c, int nsteps, double touch_point, int continue_running,
                                 int delay_step, int cap, int64_t *occurred, int *n_occurred, int64_t *ended,
                                 int *n_ended, int *steps_run) {
  return guard(c, [&] {
    require_setup(c);
    if (c->nowned < 1) throw std::runtime_error("run_until_touching needs atoms");
    require_noise_fields(c, nsteps);
    if (!c->h_ends) HIPCHK(hipHostMalloc((void **)&c->h_ends, 8 * sizeof(double), hipHostMallocDefault));
    Touch t;
    t.touch_point = touch_point;
    t.cont = continue_running != 0;
    t.delay = delay_step;
    t.since = delay_step;   // since_last_cyc starts at delay_step
    c->touch = &t;
    try {
      run_steps(c, nsteps);
    } catch (...) {
      c->touch = nullptr;
      throw;
    }
    c->touch = nullptr;
    if (n_occurred) *n_occurred = (int)t.occurred.size();
    if (n_ended) *n_ended = (int)t.ended.size();
    if (steps_run) *steps_run = t.steps;
    for (int i = 0; i < (int)t.occurred.size() && i < cap && occurred; i++) occurred[i] = t.occurred[i];
    for (int i = 0; i < (int)t.ended.size() && i < cap && ended; i++) ended[i] = t.ended[i];
    return t.stopped ? 1 : 0;
  });
}

int phafd_hip_phase(phafd_hip_ctx *c, int phase) {
  return guard(c, [&] {
    require_setup(c);
    if (phase == PHAFD_PH_START) c->timestep++;
    c->pending_decision = -1;
    run_phase(c, phase);
    sync(c);
    return 0;
  });
}

int phafd_hip_counts(phafd_hip_ctx *c, int64_t *o) {
  return guard(c, [&] {
    unsigned long long cnt[2] = {0, 0};
    int fl[4] = {0};
    HIPCHK(hipMemcpyAsync(cnt, c->counters.p, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fl, c->flags.p, 16, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    o[0] = c->nowned; o[1] = c->ngathered; o[2] = c->nlocal; o[3] = c->nghost;
    o[4] = c->ncalls; o[5] = (int64_t)cnt[0]; o[6] = (int64_t)cnt[1]; o[7] = fl[1];
    return 0;
  });
}

int phafd_hip_gridlist(phafd_hip_ctx *c, int64_t cap, int *ilist, int *numneigh, int *neigh,
                       int64_t *inum, int64_t *nentries) {
  return guard(c, [&] {
    if (!c->lists_valid || !c->need_full) { *inum = 0; *nentries = 0; return 0; }
    if (!c->grid_built) build_gridlist(c);   // from the bins and positions of the last build
    auto cnt = download(c, c->g_cnt.p, c->N);
    int64_t ni = 0;
    for (int v : cnt) ni += v > 0;
    *inum = ni;
    *nentries = c->g_nent;
    if (!ilist) return 0;
    if (ni > cap || c->g_nent > cap) throw std::invalid_argument("capacity too small");
    auto nb = download(c, c->g_neigh.p, (size_t)c->g_nent);
    int64_t w = 0;
    for (size_t p = 0; p < c->N; p++)
      if (cnt[p] > 0) { ilist[w] = (int)p; numneigh[w] = cnt[p]; w++; }
    std::memcpy(neigh, nb.data(), sizeof(int) * nb.size());
    return 0;
  });
}

int phafd_hip_halflist(phafd_hip_ctx *c, int64_t cap, int *ilist, int *numneigh, int *neigh,
                       int64_t *inum, int64_t *nentries) {
  return guard(c, [&] {
    if (!c->lists_valid || !c->need_half) { *inum = 0; *nentries = 0; return 0; }
    if (!c->half_full) build_halflist(c, true);   // from the bins and positions of the last build
    *inum = c->h_nrows;
    *nentries = c->h_nent;
    if (!ilist) return 0;
    if (c->h_nrows > cap || c->h_nent > cap) throw std::invalid_argument("capacity too small");
    auto il = download(c, c->h_ilist.p, c->h_nrows);
    auto cn = download(c, c->h_cnt.p, c->h_nrows);
    auto nb = download(c, c->h_neigh.p, c->h_nent);
    for (int r = 0; r < c->h_nrows; r++) { ilist[r] = il[r] + c->nowned; numneigh[r] = cn[r]; }
    if (!nb.empty()) std::memcpy(neigh, nb.data(), sizeof(int) * nb.size());
    return 0;
  });
}

int phafd_hip_gathered(phafd_hip_ctx *c, double *xyz, int *tag, int *label, int *owner) {
  return guard(c, [&] {
    const int ng = c->ngathered;
    if (xyz) from_d4(download(c, c->gx.p, ng), xyz);
    auto own = download(c, c->gowner.p, ng);
    if (owner && ng) std::memcpy(owner, own.data(), sizeof(int) * ng);
    if (tag) for (int i = 0; i < ng; i++) tag[i] = c->h_tag[own[i]];
    if (label && ng) {
      auto lab = download(c, c->glabel.p, ng);
      std::memcpy(label, lab.data(), sizeof(int) * ng);
    }
    return ng;
  });
}

int phafd_hip_bins(phafd_hip_ctx *c, int *info, int *binhead, int *bins, int *atom2bin) {
  return guard(c, [&] {
    const BinGeom &b = c->bg;
    if (info) {
      int v[12] = {b.nbinx, b.nbiny, b.nbinz, b.mbinx, b.mbiny, b.mbinz, b.mbinxlo, b.mbinylo, b.mbinzlo,
                   b.mbins, (int)c->stencil_full.size(), (int)c->stencil_half.size()};
      std::memcpy(info, v, sizeof(v));
    }
    if (!c->lists_valid) return 0;
    const int ng = c->ngathered, nb = b.mbins;
    auto bs = download(c, c->binstart.p, (size_t)nb + 1);
    auto ba = download(c, c->binatoms.p, (size_t)ng);
    if (binhead) for (int k = 0; k < nb; k++) binhead[k] = bs[k] < bs[k + 1] ? ba[bs[k]] + c->nowned : -1;
    if (bins) {
      for (int k = 0; k < nb; k++)
        for (int q = bs[k]; q < bs[k + 1]; q++) bins[ba[q]] = (q + 1 < bs[k + 1]) ? ba[q + 1] + c->nowned : -1;
    }
    if (atom2bin && ng) {
      auto a = download(c, c->a2b.p, (size_t)ng);
      std::memcpy(atom2bin, a.data(), sizeof(int) * ng);
    }
    return 0;
  });
}

int phafd_hip_energy(phafd_hip_ctx *c, double *out) {
  return guard(c, [&] {
    fill_zero(c->esum.p, 8 * sizeof(double), c->stream);
    for (auto &f : c->fixes) {
      if (f->kind == FixDef::FH) {
        // FH part only here; gradient part added by the CONJ fix below
        if (c->fp32)
          reduce_energy_fh(c->N, c->fs.phi.p, (const float *)nullptr, nullptr, nullptr, f->temp / f->volFH, f->chi,
                           0.0, c->esum.p, c->stream);
        else
          reduce_energy_fh(c->N, c->fd.phi.p, (const double *)nullptr, nullptr, nullptr, f->temp / f->volFH,
                           f->chi, 0.0, c->esum.p, c->stream);
      } else if (f->kind == FixDef::CONJ) {
        // square-gradient term: pass phi with chi = 0 would add FH terms, so
        // accumulate only the gradient sum into esum[1] then move it
        if (c->fp32)
          reduce_energy_fh(c->N, c->fs.phi.p, c->fs.grad[0].p, c->fs.grad[1].p, c->fs.grad[2].p, 0.0, 0.0,
                           f->temp / f->volFH * f->gamma, c->esum.p + 2, c->stream);
        else
          reduce_energy_fh(c->N, c->fd.phi.p, c->fd.grad[0].p, c->fd.grad[1].p, c->fd.grad[2].p, 0.0, 0.0,
                           f->temp / f->volFH * f->gamma, c->esum.p + 2, c->stream);
      } else if (f->kind == FixDef::CHAIN) {
        ChainParams cp{f->r0, f->kbond, f->kbend, f->drag};
        launch_chain_energy(f->nmem, f->members.p, f->cstart.p, f->cend.p, c->ux.p, cp, c->esum.p + 6, c->stream);
      } else if (f->kind == FixDef::SEMIFLEX) {   // rigid rods: the bending energy kb (1 - cos) alone
        ChainParams cp{f->bondlength, 0.0, f->bending, 0.0};
        launch_chain_energy(f->nmem, f->members.p, f->cstart.p, f->cend.p, c->ux.p, cp, c->esum.p + 6, c->stream);
      }
    }
    // pair energies: recompute the pair pass into scratch copies of mu / gF
    if (c->neigh_set && c->lists_valid) {
      DBuf<double> mu_save;   // holds either precision (8 bytes per point is enough)
      DBuf<double4> gf_save, f_save;
      mu_save.alloc(c->N);
      void *mu_p = c->fp32 ? (void *)c->fs.mu.p : (void *)c->fd.mu.p;
      gf_save.alloc((size_t)c->ngathered + 1);
      f_save.alloc((size_t)c->nowned + 1);
      HIPCHK(hipMemcpyAsync(mu_save.p, mu_p, c->real_bytes() * c->N, hipMemcpyDeviceToDevice, c->stream));
      if (c->ngathered) HIPCHK(hipMemcpyAsync(gf_save.p, c->gF.p, sizeof(double4) * c->ngathered, hipMemcpyDeviceToDevice, c->stream));
      if (c->nowned) HIPCHK(hipMemcpyAsync(f_save.p, c->F.p, sizeof(double4) * c->nowned, hipMemcpyDeviceToDevice, c->stream));
      phase_pair(c, c->esum.p + 4, c->esum.p + 5);
      HIPCHK(hipMemcpyAsync(mu_p, mu_save.p, c->real_bytes() * c->N, hipMemcpyDeviceToDevice, c->stream));
      if (c->ngathered) HIPCHK(hipMemcpyAsync(c->gF.p, gf_save.p, sizeof(double4) * c->ngathered, hipMemcpyDeviceToDevice, c->stream));
      if (c->nowned) HIPCHK(hipMemcpyAsync(c->F.p, f_save.p, sizeof(double4) * c->nowned, hipMemcpyDeviceToDevice, c->stream));
      sync(c);
    }
    double e[8];
    HIPCHK(hipMemcpyAsync(e, c->esum.p, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    const double dv = c->dx() * c->dy() * c->dz();
    out[0] = (e[0] + e[3]) * dv;   // FH (esum[0]) + gradient (esum[3])
    out[1] = e[4];
    out[2] = e[5];
    out[3] = e[6];
    return 0;
  });
}

int phafd_hip_probe(phafd_hip_ctx *c, int kernel_id, double *out3) {
  return guard(c, [&] {
    static KernelProbe probe;
    if (kernel_id >= 0) {
      if (kernel_id >= K_NKERNELS) throw std::invalid_argument("unknown kernel id");
      probe.id = kernel_id;
      probe.ms = 0; probe.total_bytes = 0; probe.launches = 0; probe.used = 0; probe.bytes.clear();
      g_probe = &probe;
    } else {
      g_probe = nullptr;
    }
    if (out3) { out3[0] = probe.ms; out3[1] = (double)probe.launches; out3[2] = probe.total_bytes; }
    return 0;
  });
}

int phafd_hip_set_option(phafd_hip_ctx *c, int key, double value) {
  return guard(c, [&] {
    switch (key) {
      case PHAFD_OPT_EAGER_GRIDLIST: c->eager_gridlist = value != 0; break;
      case PHAFD_OPT_FUSE_FH: c->fuse_fh = value != 0; break;
      case PHAFD_OPT_FUSE_SPECTRAL: c->fuse_spectral = value != 0; break;
      case PHAFD_OPT_EAGER_HALFLIST: c->eager_halflist = value != 0; break;
      case PHAFD_OPT_ATOM_WINDOWS: c->atom_windows = value != 0; break;
      case PHAFD_OPT_SEMI_IMPLICIT: c->semi_implicit = value != 0; break;
      case PHAFD_OPT_NEXT_GRADIENT: c->next_grad = value != 0; break;
      default: throw std::invalid_argument("unknown option");
    }
    return 0;
  });
}

int phafd_hip_timers(phafd_hip_ctx *c, double *ms, int reset) {
  if (!c) return PHAFD_E_ARG;
  if (ms) for (int k = 0; k < 8; k++) ms[k] = c->tacc[k];
  if (reset) for (double &t : c->tacc) t = 0;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------- stateless kernels
namespace {
struct Scratch {
  FFTPlan plan;
  hipStream_t s = nullptr;
  Scratch(int nx, int ny, int nz) {
    plan.init(nx, ny, nz);
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~Scratch() {
    if (!s) return;
    scan_forget_stream(s);
    (void)hipStreamDestroy(s);
  }
};
template <class Fn>
int sguard(Fn fn) {
  try {
    return fn();
  } catch (const HipError &) {
    return PHAFD_E_HIP;
  } catch (const std::invalid_argument &) {
    return PHAFD_E_ARG;
  } catch (const std::exception &) {
    return PHAFD_E_RUNTIME;
  }
}
}  // namespace

extern "C" {

int phafd_hip_noise_field(int nx, int ny, int nz, const double *period, double mobility, double temp, double dt,
                          uint64_t key0, int64_t step, double *xi) {
  return sguard([&] {
    if (!period || !xi || nx < 2 || ny < 2 || nz < 2 || (nx & 1)) throw std::invalid_argument("noise_field: bad args");
    const int hx = nx / 2 + 1;
    const size_t M = (size_t)hx * ny * nz;
    UpdateParams u{};
    u.dt = dt;
    u.qx_scale = 2 * PI_REF / period[0];
    u.qky_scale = 2 * PI_REF / period[1];
    u.qkz_scale = 2 * PI_REF / period[2];
    const double invL3 = 1.0 / (period[0] * period[1] * period[2]);
    u.noise_dev = 1;
    u.key0 = key0;
    u.key1 = 0x5048414644ull;
    u.step = (unsigned long long)step;
    u.pref_c = std::sqrt(12 * temp * mobility * invL3);
    u.pref_r = std::sqrt(24 * temp * mobility * invL3);
    u.sqrtdt = std::sqrt(dt);
    Scratch sc(nx, ny, nz);
    DBuf<double2> out;
    out.alloc(M);
    noise_field(hx, ny, nz, u, out.p, sc.s);
    HIPCHK(hipStreamSynchronize(sc.s));
    HIPCHK(hipMemcpy(xi, out.p, sizeof(double2) * M, hipMemcpyDeviceToHost));
    return 0;
  });
}

int phafd_hip_fft_r2c(int nx, int ny, int nz, const double *in, double *out) {
  return sguard([&] {
    Scratch sc(nx, ny, nz);
    DBuf<double> r;
    DBuf<double2> k, cn;
    r.alloc(sc.plan.N()); k.alloc(sc.plan.M()); cn.alloc(sc.plan.M());
    HIPCHK(hipMemcpy(r.p, in, sizeof(double) * sc.plan.N(), hipMemcpyHostToDevice));
    fft_r2c(sc.plan, r.p, k.p, sc.s);
    spec_k_to_canon(sc.plan, k.p, cn.p, sc.s);
    HIPCHK(hipStreamSynchronize(sc.s));
    HIPCHK(hipMemcpy(out, cn.p, sizeof(double2) * sc.plan.M(), hipMemcpyDeviceToHost));
    return 0;
  });
}
int phafd_hip_fft_c2r(int nx, int ny, int nz, const double *in, double *out) {
  return sguard([&] {
    Scratch sc(nx, ny, nz);
    DBuf<double> r;
    DBuf<double2> k, cn;
    r.alloc(sc.plan.N()); k.alloc(sc.plan.M()); cn.alloc(sc.plan.M());
    HIPCHK(hipMemcpy(cn.p, in, sizeof(double2) * sc.plan.M(), hipMemcpyHostToDevice));
    spec_canon_to_k(sc.plan, cn.p, k.p, sc.s);
    fft_c2r(sc.plan, k.p, r.p, 0.0, sc.s);
    HIPCHK(hipStreamSynchronize(sc.s));
    HIPCHK(hipMemcpy(out, r.p, sizeof(double) * sc.plan.N(), hipMemcpyDeviceToHost));
    return 0;
  });
}
int phafd_hip_gradphi(int nx, int ny, int nz, const double *period, const double *phi, double *gx,
                      double *gy, double *gz) {
  return sguard([&] {
    Scratch sc(nx, ny, nz);
    const size_t N = sc.plan.N(), M = sc.plan.M();
    DBuf<double> r, g[3];
    DBuf<double2> ph, gh[3];
    r.alloc(N); ph.alloc(M);
    for (int d = 0; d < 3; d++) { g[d].alloc(N); gh[d].alloc(M); }
    HIPCHK(hipMemcpy(r.p, phi, sizeof(double) * N, hipMemcpyHostToDevice));
    fft_r2c(sc.plan, r.p, ph.p, sc.s);
    SpecConst k{2 * PI_REF / period[0], 2 * PI_REF / period[2], 2 * PI_REF / period[1]};
    grad_multiply(sc.plan, ph.p, gh[0].p, gh[1].p, gh[2].p, k, sc.s);
    double2 *ins[3] = {gh[0].p, gh[1].p, gh[2].p};
    double *outs[3] = {g[0].p, g[1].p, g[2].p};
    fft_c2r_batch(sc.plan, ins, outs, 3, (double)nx * ny * nz, sc.s);
    HIPCHK(hipStreamSynchronize(sc.s));
    HIPCHK(hipMemcpy(gx, g[0].p, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gy, g[1].p, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gz, g[2].p, sizeof(double) * N, hipMemcpyDeviceToHost));
    return 0;
  });
}
int phafd_hip_floryhuggins(int64_t n, double *phi, double *mu, double temp, double volFH, double chi,
                           int64_t *nless, int64_t *ngreat) {
  return sguard([&] {
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DBuf<double> p, m;
    DBuf<unsigned long long> cn;
    p.alloc(n); m.alloc(n); cn.alloc(2);
    HIPCHK(hipMemcpy(p.p, phi, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.p, mu, sizeof(double) * n, hipMemcpyHostToDevice));
    fill_zero(cn.p, 16, s);
    floryhuggins((size_t)n, p.p, m.p, temp / volFH, chi, cn.p, s);
    HIPCHK(hipStreamSynchronize(s));
    unsigned long long h[2];
    HIPCHK(hipMemcpy(phi, p.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(mu, m.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(h, cn.p, 16, hipMemcpyDeviceToHost));
    HIPCHK(hipStreamDestroy(s));
    if (nless) *nless = (int64_t)h[0];
    if (ngreat) *ngreat = (int64_t)h[1];
    return 0;
  });
}
int phafd_hip_spectral_update(int nx, int ny, int nz, const double *period, double *phihat,
                              const double *muhat, double mobility, double temp, double volFH,
                              double gamma, double dt, const double *noise) {
  return sguard([&] {
    Scratch sc(nx, ny, nz);
    const size_t M = sc.plan.M();
    DBuf<double2> cn, pk, mk, nk;
    cn.alloc(M); pk.alloc(M); mk.alloc(M);
    HIPCHK(hipMemcpy(cn.p, phihat, sizeof(double2) * M, hipMemcpyHostToDevice));
    spec_canon_to_k(sc.plan, cn.p, pk.p, sc.s);
    HIPCHK(hipStreamSynchronize(sc.s));
    HIPCHK(hipMemcpy(cn.p, muhat, sizeof(double2) * M, hipMemcpyHostToDevice));
    spec_canon_to_k(sc.plan, cn.p, mk.p, sc.s);
    if (noise) {
      HIPCHK(hipStreamSynchronize(sc.s));
      nk.alloc(M);
      HIPCHK(hipMemcpy(cn.p, noise, sizeof(double2) * M, hipMemcpyHostToDevice));
      spec_canon_to_k(sc.plan, cn.p, nk.p, sc.s);
    }
    UpdateParams u{};
    u.mobility = mobility; u.temp = temp; u.volFH = volFH; u.gamma = gamma; u.dt = dt;
    u.normalization = 1.0 / ((double)nx * nz * ny);
    u.qx_scale = 2 * PI_REF / period[0];
    u.qky_scale = 2 * PI_REF / period[1];
    u.qkz_scale = 2 * PI_REF / period[2];
    spectral_update(sc.plan, pk.p, mk.p, noise ? nk.p : nullptr, u, nullptr, sc.s);
    spec_k_to_canon(sc.plan, pk.p, cn.p, sc.s);
    HIPCHK(hipStreamSynchronize(sc.s));
    HIPCHK(hipMemcpy(phihat, cn.p, sizeof(double2) * M, hipMemcpyDeviceToHost));
    return 0;
  });
}

}  // extern "C"

// ================================================================ batches
// Independent replicas of one grid advanced together.  The members are
// ordinary contexts (configured through the calls above) whose fields live in
// one slab; the grid side of a step runs as one launch per pass over all
// members (captured in two hipGraphs), the atom side per member on a few
// member streams.  Every member gives bit for bit what it gives run alone:
// the batched passes run the same kernels on each member's slice.
struct phafd_hip_batch {
  std::string err;
  int device = 0, n = 0;
  size_t N = 0, M = 0;
  hipStream_t stream = nullptr;                  // grid side
  std::vector<hipStream_t> mstreams;             // atom side
  FieldSet<double> slab;
  DBuf<unsigned long long> counters;             // 2 per member
  DBuf<double> chi_b;                            // chi per member
  std::vector<std::unique_ptr<phafd_hip_ctx>> members;
  hipEvent_t ev_grad = nullptr;
  std::vector<hipEvent_t> ev_m;
  hipGraph_t g_grad = nullptr, g_spec = nullptr;
  hipGraphExec_t x_grad = nullptr, x_spec = nullptr;
  bool setup_done = false, use_graphs = true, sep = false, zsplit = false;
  int nstreams = 1, chunk = 1;
  double tune_grad_ms[4] = {0, 0, 0, 0}, tune_pair_ms[4] = {0, 0, 0, 0};
  // the batched pair phase: every member in one launch per pass
  int fused_mode = -1;                 // -1: picked by the autotuner, 0: per-member launches, 1: one launch per pass
  bool fused = false;                  // what runs
  double tune_fused_ms = 0;            // the pair phase in one launch per pass, at setup (best shape)
  MultiShape shape;                    // its launch shape (autotuned)
  // mode -1: steps [BATCH_TRIAL_START, + 4 * BATCH_TRIAL) run the two modes
  // in blocks A B B A (real steps: the modes compute the same numbers),
  // every step but a block's first timed whole; the faster mode is kept
  int64_t steps_run = 0;
  int trial = 0;                       // trial steps done
  bool decided = false;
  double trial_ms[2] = {0, 0};         // mean step time: per-member, one launch per pass
  double tune_shape_ms[6] = {0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev_n;        // member neighbour stage done
  hipEvent_t ev_pair = nullptr;        // batched pair phase done
  unsigned char *h_desc = nullptr;     // pinned staging of the descriptors
  size_t h_desc_bytes = 0;
  DBuf<unsigned char> d_desc;
  UpdateParams u{};
  SpecConst sc{};
  double tv = 1;
};

namespace {

template <class Fn>
int bguard(phafd_hip_batch *b, Fn fn) {
  try {
    if (b) (void)hipSetDevice(b->device);
    return fn();
  } catch (const HipError &e) {
    if (b) b->err = e.what();
    return PHAFD_E_HIP;
  } catch (const std::invalid_argument &e) {
    if (b) b->err = e.what();
    return PHAFD_E_ARG;
  } catch (const std::exception &e) {
    if (b) b->err = e.what();
    std::string w = e.what();
    if (w.find("NAN") != std::string::npos) return PHAFD_E_NAN;
    return PHAFD_E_RUNTIME;
  }
}

// gradient of every member (FixGridConjugate::start_of_step), chunk members
// per launch
void batch_grad(phafd_hip_batch *b, int chunk) {
  const FFTPlan &p = b->members[0]->plan;
  FieldSet<double> &S = b->slab;
  for (int r0 = 0; r0 < b->n; r0 += chunk) {
    Slab sl;
    sl.n = std::min(chunk, b->n - r0);
    sl.sr = b->N;
    sl.sc = b->M;
    const size_t oN = (size_t)r0 * b->N, oM = (size_t)r0 * b->M;
    if (b->sep) {
      fft_grad_sep(p, S.phi.p + oN, S.phihat.p + oM, S.ghat[1].p + oM, S.ghat[2].p + oM, S.grad[0].p + oN,
                   S.grad[1].p + oN, S.grad[2].p + oN, b->sc, b->stream, sl);
    } else {
      fft_r2c(p, S.phi.p + oN, S.phihat.p + oM, b->stream, nullptr, sl);
      grad_multiply(p, S.phihat.p + oM, S.ghat[0].p + oM, S.ghat[1].p + oM, S.ghat[2].p + oM, b->sc, b->stream, sl);
      for (int d = 0; d < 3; d++)
        fft_c2r_slab(p, S.ghat[d].p + oM, S.grad[d].p + oN, sl, (double)b->N, b->stream);
    }
  }
}

// Flory-Huggins fused into r2c(mu), the spectral update and c2r(phihat) of
// every member (post_force .. post_final_integrate)
void batch_spec(phafd_hip_batch *b, int chunk) {
  const FFTPlan &p = b->members[0]->plan;
  FieldSet<double> &S = b->slab;
  for (int r0 = 0; r0 < b->n; r0 += chunk) {
    Slab sl;
    sl.n = std::min(chunk, b->n - r0);
    sl.sr = b->N;
    sl.sc = b->M;
    const size_t oN = (size_t)r0 * b->N, oM = (size_t)r0 * b->M;
    FHPre pre;
    pre.on = true;
    pre.phi = S.phi.p + oN;
    pre.mu = S.mu.p + oN;
    pre.tv = b->tv;
    pre.counters = b->counters.p + 2 * r0;
    pre.chi_b = b->chi_b.p + r0;
    if (b->zsplit) {
      fft_r2c_xy(p, S.mu.p + oN, S.muhat.p + oM, b->stream, &pre, sl);
      fft_z_update(p, S.muhat.p + oM, S.phihat.p + oM, nullptr, b->u, nullptr, b->stream, sl);
      fft_c2r_yx_slab(p, S.phihat.p + oM, S.phi.p + oN, sl, 0.0, b->stream);
    } else {
      fft_r2c(p, S.mu.p + oN, S.muhat.p + oM, b->stream, &pre, sl);
      spectral_update(p, S.phihat.p + oM, S.muhat.p + oM, nullptr, b->u, nullptr, b->stream, sl);
      fft_c2r_slab(p, S.phihat.p + oM, S.phi.p + oN, sl, 0.0, b->stream);
    }
  }
}

void batch_sync(phafd_hip_batch *b) {
  HIPCHK(hipStreamSynchronize(b->stream));
  for (auto s : b->mstreams) HIPCHK(hipStreamSynchronize(s));
}

void batch_assign_streams(phafd_hip_batch *b, int k) {
  batch_sync(b);
  for (int i = 0; i < b->n; i++) b->members[i]->stream = b->mstreams[i % k];
  b->nstreams = k;
}

// time `fn` (enqueued work) over reps, after one warm call
template <class Fn>
double time_ms(phafd_hip_batch *b, int reps, Fn fn) {
  fn();
  batch_sync(b);
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; r++) fn();
  batch_sync(b);
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
}

void destroy_graphs(phafd_hip_batch *b) {
  if (b->x_grad) (void)hipGraphExecDestroy(b->x_grad);
  if (b->x_spec) (void)hipGraphExecDestroy(b->x_spec);
  if (b->g_grad) (void)hipGraphDestroy(b->g_grad);
  if (b->g_spec) (void)hipGraphDestroy(b->g_spec);
  b->x_grad = b->x_spec = nullptr;
  b->g_grad = b->g_spec = nullptr;
}

template <class Fn>
void capture(phafd_hip_batch *b, hipGraph_t *g, hipGraphExec_t *x, Fn fn) {
  KernelProbe *saved = g_probe;   // no event records inside a graph
  g_probe = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  try {
    fn();
  } catch (...) {
    hipGraph_t tmp;
    (void)hipStreamEndCapture(b->stream, &tmp);
    g_probe = saved;
    throw;
  }
  HIPCHK(hipStreamEndCapture(b->stream, g));
  g_probe = saved;
  HIPCHK(hipGraphInstantiate(x, *g, nullptr, nullptr, 0));
}

// The pair phase of every member in one launch per pass (grid-atom force
// walk, mu sums, lj/cut, each fold level, the chains' two half steps):
// descriptors of every member packed into one pinned blob, copied with one
// transfer.  The per-member arithmetic is that of phase_pair +
// integrate_both_halves.  Returns false (nothing launched) when a member
// does not fit: other pair styles, no grid-atom style, folds of another
// shape than member 0's.
struct FusedPlan {
  size_t off_ga, off_lj, off_fold, off_ch, off_int, bytes;
  int nlev = 0, items_ga = 0, nch = 0, maxlen = 0;
  bool lj = false, chains = false, onepass = false;
  std::vector<int> ints;   // boff_ga (nm+1), fold boffs (nlev+1)*(nm+1), coff (nm+1)
};

bool fused_eligible(phafd_hip_batch *b) {
  phafd_hip_ctx *c0 = b->members[0].get();
  for (auto &m : b->members) {
    phafd_hip_ctx *c = m.get();
    if (!c->neigh_set || c->eager_gridlist || c->fp32) return false;
    int nga = 0, nlj = 0;
    for (auto &p : c->pairs) {
      if (p->style == 2) nlj++;
      else if (!bead_pair(*p)) nga++;
      else return false;
    }
    if (nga != 1 || nlj > 1 || nlj != (int)std::count_if(c0->pairs.begin(), c0->pairs.end(),
                                                          [](const std::unique_ptr<PairDef> &p) {
                                                            return p->style == 2;
                                                          }))
      return false;
    if (c->swaps.size() != c0->swaps.size()) return false;
    for (size_t k = 0; k < c->swaps.size(); k++)
      if (c->swaps[k].disjoint_pair != c0->swaps[k].disjoint_pair) return false;
    if ((fused_chain_fix(c) != nullptr) != (fused_chain_fix(c0) != nullptr)) return false;
  }
  return true;
}

// the members' lists and pair arrays for this step (member streams)
void fused_prepare(phafd_hip_ctx *c) {
  if (c->ga_calls != c->ncalls) build_ga_pairs(c);
}

void batch_pair_fused(phafd_hip_batch *b, bool with_chains) {
  const int nm = b->n;
  const int T = b->shape.threads;
  phafd_hip_ctx *c0 = b->members[0].get();
  FusedPlan P;
  P.lj = std::any_of(c0->pairs.begin(), c0->pairs.end(), [](const std::unique_ptr<PairDef> &p) {
    return p->style == 2;
  });
  P.chains = with_chains && fused_chain_fix(c0) != nullptr;
  // fold levels (the order of reverse_comm): pairs of disjoint swaps, single
  // swaps, then the copy swap into F
  std::vector<std::pair<int, int>> levels;   // (swap a, swap b or -1)
  for (int iswap = (int)c0->swaps.size() - 1; iswap >= 1; iswap--) {
    if (c0->swaps[iswap].disjoint_pair) {
      levels.push_back({iswap, iswap - 1});
      iswap--;
    } else {
      levels.push_back({iswap, -1});
    }
  }
  levels.push_back({0, -1});
  P.nlev = (int)levels.size();
  P.onepass = true;
  for (auto &m : b->members) P.onepass = P.onepass && reverse_from_one_pass(m.get());
  std::vector<GAMemberDev> ga(nm);
  std::vector<LJMemberDev> lj(nm);
  std::vector<FoldDev> fold((size_t)(P.nlev + 1) * nm);   // levels, then the fill of F
  std::vector<ChainMemberDev> ch(nm);
  const int stride = nm + 1;
  P.ints.assign((size_t)stride * (P.nlev + 4), 0);
  int *boff_ga = P.ints.data(), *boff_fold = boff_ga + stride, *coff = boff_fold + (size_t)stride * (P.nlev + 1);
  for (int i = 0; i < nm; i++) {
    phafd_hip_ctx *c = b->members[i].get();
    PairDef *pga = nullptr, *plj = nullptr;
    for (auto &p : c->pairs) (p->style == 2 ? plj : pga) = p.get();
    GAMemberDev &g = ga[i];
    g.gv = c->gridview<double>();
    g.ga = ga_params(*pga);
    g.gx = c->gx.p; g.order = c->binatoms.p; g.aoff = c->ga_aoff.p; g.ps = c->ga_ps.p;
    g.mterm = c->ga_mterm.p; g.gF = c->gF.p; g.ng = c->ngathered;
    g.mu = c->fd.mu.p; g.poff = c->ga_poff.p;
    boff_ga[i + 1] = boff_ga[i] + (c->ngathered + T - 1) / T;
    if (P.lj) {
      LJMemberDev &l = lj[i];
      l.gx = c->gx.p; l.rowof = c->h_rowof.p; l.hoff = c->h_off.p; l.hneigh = c->h_neigh.p;
      l.toff = c->t_off.p; l.tpair = c->t_pair.p; l.ti = c->t_i.p; l.nowned = c->nowned; l.ng = c->ngathered;
      l.gF = c->gF.p; l.lj = LJParams{plj->lj1, plj->lj2, plj->ljcutsq};
    }
    for (int L = 0; L < P.nlev; L++) {
      FoldDev &f = fold[(size_t)L * nm + i];
      const SwapDev &wa = c->swaps[levels[L].first];
      f.ca = wa.count; f.fa = wa.first; f.sa = wa.send.p;
      f.cb = 0; f.fb = 0; f.sb = nullptr;
      if (levels[L].second >= 0) {
        const SwapDev &wb = c->swaps[levels[L].second];
        f.cb = wb.count; f.fb = wb.first; f.sb = wb.send.p;
      }
      f.target = levels[L].first == 0 ? c->F.p : c->gF.p;
      f.src = c->gF.p;
      int *bo = boff_fold + (size_t)L * stride;
      bo[i + 1] = bo[i] + (f.ca + f.cb + T - 1) / T;
    }
    FoldDev &fz = fold[(size_t)P.nlev * nm + i];   // the fill of F
    fz = FoldDev{c->nowned, 0, 0, 0, nullptr, nullptr, c->F.p, nullptr};
    int *bz = boff_fold + (size_t)P.nlev * stride;
    bz[i + 1] = bz[i] + (c->nowned + T - 1) / T;
    if (P.chains) {
      FixDef &f = *fused_chain_fix(c);
      ch[i] = ChainMemberDev{f.chs.p, f.che.p, c->F.p, c->ux.p, c->x.p, c->image.p,
                             ChainParams{f.r0, f.kbond, f.kbend, f.drag}, c->box()};
      coff[i + 1] = coff[i] + f.nchain;
      P.maxlen = std::max(P.maxlen, f.maxlen);
    } else {
      coff[i + 1] = coff[i];
    }
  }
  P.items_ga = boff_ga[nm];
  P.nch = coff[nm];
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  P.off_ga = 0;
  P.off_lj = al(P.off_ga + sizeof(GAMemberDev) * nm);
  P.off_fold = al(P.off_lj + sizeof(LJMemberDev) * nm);
  P.off_ch = al(P.off_fold + sizeof(FoldDev) * fold.size());
  P.off_int = al(P.off_ch + sizeof(ChainMemberDev) * nm);
  P.bytes = al(P.off_int + sizeof(int) * P.ints.size());
  if (P.bytes > b->h_desc_bytes) {
    if (b->h_desc) HIPCHK(hipHostFree(b->h_desc));
    b->h_desc = nullptr;
    HIPCHK(hipHostMalloc((void **)&b->h_desc, P.bytes, hipHostMallocDefault));
    b->h_desc_bytes = P.bytes;
  }
  b->d_desc.alloc(P.bytes);
  unsigned char *h = b->h_desc;
  std::memcpy(h + P.off_ga, ga.data(), sizeof(GAMemberDev) * nm);
  std::memcpy(h + P.off_lj, lj.data(), sizeof(LJMemberDev) * nm);
  std::memcpy(h + P.off_fold, fold.data(), sizeof(FoldDev) * fold.size());
  std::memcpy(h + P.off_ch, ch.data(), sizeof(ChainMemberDev) * nm);
  std::memcpy(h + P.off_int, P.ints.data(), sizeof(int) * P.ints.size());
  unsigned char *d = b->d_desc.p;
  HIPCHK(hipMemcpyAsync(d, h, P.bytes, hipMemcpyHostToDevice, b->stream));
  const GAMemberDev *dga = (const GAMemberDev *)(d + P.off_ga);
  const LJMemberDev *dlj = (const LJMemberDev *)(d + P.off_lj);
  const FoldDev *dfold = (const FoldDev *)(d + P.off_fold);
  const ChainMemberDev *dch = (const ChainMemberDev *)(d + P.off_ch);
  const int *di = (const int *)(d + P.off_int);
  const int *dboff_ga = di, *dboff_fold = di + stride, *dcoff = dboff_fold + (size_t)stride * (P.nlev + 1);
  double ng_all = 0;
  for (auto &g : ga) ng_all += g.ng;
  probe_begin(K_GA_FORCE, b->stream);
  launch_ga_force_multi(nm, dga, dboff_ga, P.items_ga, c0->nx, c0->ny, b->shape, b->stream);
  probe_end(K_GA_FORCE, b->stream, 4.0 * sizeof(double) * b->N * nm + 64.0 * ng_all);
  launch_ga_list_mu_multi(nm, b->N, dga, b->stream);
  if (P.lj) launch_lj_multi(nm, dlj, dboff_ga, P.items_ga, b->shape, b->stream);
  for (int L = 0; L < P.nlev; L++) {
    const bool copy_swap = levels[L].first == 0;
    if (copy_swap && !P.onepass)
      launch_fill_multi(nm, dfold + (size_t)P.nlev * nm, dboff_fold + (size_t)P.nlev * stride,
                        boff_fold[(size_t)P.nlev * stride + nm], b->shape, b->stream);
    launch_fold_multi(nm, dfold + (size_t)L * nm, dboff_fold + (size_t)L * stride, boff_fold[(size_t)L * stride + nm],
                      copy_swap && P.onepass, b->shape, b->stream);
  }
  if (P.chains) launch_chain_multi(nm, dch, dcoff, P.nch, P.maxlen, c0->dt, b->stream);
}

// the online choice of the pair-phase mode: after the set-up state has
// given way to the run's (neighbour rebuilds under way), four blocks of
// BATCH_TRIAL steps in the order A B B A
constexpr int BATCH_TRIAL_START = 16, BATCH_TRIAL = 8;

void batch_steps(phafd_hip_batch *b, int nsteps) {
  if (!b->setup_done) throw std::invalid_argument("call phafd_hip_batch_setup before running");
  for (auto &m : b->members)
    if (m->neigh_set) clear_decide_slot(m.get(), m->timestep + 1);
  for (int it = 0; it < nsteps; it++) {
    const bool eligible = fused_eligible(b);
    const bool trial = b->fused_mode < 0 && !b->decided && eligible && b->steps_run >= BATCH_TRIAL_START;
    if (trial) {   // blocks in the order one launch per pass, per-member, per-member, one launch per pass
      const int blk = b->trial / BATCH_TRIAL;   // (a drift of the step time over the trial cancels)
      b->fused = blk == 0 || blk == 3;
    }
    const auto t_step = std::chrono::steady_clock::now();
    for (auto &m : b->members) m->timestep++;
    if (b->use_graphs) HIPCHK(hipGraphLaunch(b->x_grad, b->stream));
    else batch_grad(b, b->chunk);
    HIPCHK(hipEventRecord(b->ev_grad, b->stream));
    const bool fused = b->fused && eligible;
    // a member's neighbour stage (decide, or pbc + ghosts + bins + lists,
    // whose sizes the host reads back) needs positions only: it runs beside
    // the gradient block, and the pair phase waits for the gradient
    for (int i = 0; i < b->n; i++) {
      phafd_hip_ctx *c = b->members[i].get();
      phase_neigh(c);
      HIPCHK(hipStreamWaitEvent(c->stream, b->ev_grad, 0));
      if (fused) {
        fused_prepare(c);
      } else {
        phase_pair(c);
        if (FixDef *fc = fused_chain_fix(c)) integrate_both_halves(c, *fc);
        else phase_initial(c);
      }
      HIPCHK(hipEventRecord(b->ev_m[i], c->stream));
    }
    for (int i = 0; i < b->n; i++) HIPCHK(hipStreamWaitEvent(b->stream, b->ev_m[i], 0));
    if (fused) {
      batch_pair_fused(b, true);
      HIPCHK(hipEventRecord(b->ev_pair, b->stream));
      for (int i = 0; i < b->n; i++) {
        phafd_hip_ctx *c = b->members[i].get();
        HIPCHK(hipStreamWaitEvent(c->stream, b->ev_pair, 0));
        if (!fused_chain_fix(c)) phase_initial(c);
      }
    }
    if (b->use_graphs) HIPCHK(hipGraphLaunch(b->x_spec, b->stream));
    else batch_spec(b, b->chunk);
    for (int i = 0; i < b->n; i++) {
      phafd_hip_ctx *c = b->members[i].get();
      if (!fused_chain_fix(c))
        for (auto &f : c->fixes)
          if (f->kind != FixDef::CONJ && f->kind != FixDef::FH) integrate_half(c, *f, true);
      if (c->neigh_set) enqueue_decide(c);   // next step's rebuild decision (Neighbor::decide)
      enqueue_flags_readback(c);
      HIPCHK(hipMemcpyAsync(c->h_flags + 2, c->fd.phi.p, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    }
    batch_sync(b);
    b->steps_run++;
    if (trial) {
      const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_step).count();
      if (b->trial % BATCH_TRIAL != 0) b->trial_ms[fused ? 1 : 0] += ms / (2 * (BATCH_TRIAL - 1));
      if (++b->trial == 4 * BATCH_TRIAL) {
        b->decided = true;
        b->fused = b->trial_ms[1] < b->trial_ms[0];
      }
    }
    probe_collect();
    for (int i = 0; i < b->n; i++) {   // the checks of src/integrate.cpp:190-201, per member
      phafd_hip_ctx *c = b->members[i].get();
      double phi0;
      std::memcpy(&phi0, c->h_flags + 2, sizeof(double));
      if (std::isnan(phi0)) throw std::runtime_error("NAN encountered in phi (member " + std::to_string(i) + ").");
      check_step_flags(c, c->timestep);
    }
  }
}

}  // namespace

extern "C" {

int phafd_hip_batch_create(const phafd_hip_desc *desc, int nmembers, phafd_hip_batch **out) {
  if (!desc || !out) return PHAFD_E_ARG;
  *out = nullptr;
  auto b = std::make_unique<phafd_hip_batch>();
  b->device = desc->device;
  int rc = bguard(b.get(), [&] {
    if (nmembers < 1) throw std::invalid_argument("a batch needs at least one member");
    if (desc->precision != PHAFD_FP64) throw std::invalid_argument("batched replicas are FP64");
    HIPCHK(hipSetDevice(desc->device));
    b->n = nmembers;
    b->N = (size_t)desc->nx * desc->ny * desc->nz;
    b->M = (size_t)(desc->nx / 2 + 1) * desc->ny * desc->nz;
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->slab.alloc(b->N * b->n, b->M * b->n);
    b->counters.alloc(2 * (size_t)b->n);
    fill_zero(b->counters.p, 16 * (size_t)b->n, b->stream);
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->n; i++) {
      auto c = std::make_unique<phafd_hip_ctx>();
      c->device = desc->device;
      MemberSlot slot{b->stream, &b->slab, i, b->counters.p};
      init_ctx(c.get(), desc, &slot);
      b->members.push_back(std::move(c));
    }
    HIPCHK(hipEventCreateWithFlags(&b->ev_grad, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->ev_pair, hipEventDisableTiming));
    b->ev_m.resize(b->n);
    for (auto &e : b->ev_m) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return 0;
  });
  if (rc != 0) {
    g_create_err = b->err;
    return rc;
  }
  *out = b.release();
  return 0;
}

int phafd_hip_batch_member(phafd_hip_batch *b, int i, phafd_hip_ctx **member) {
  return bguard(b, [&] {
    if (!member || i < 0 || i >= b->n) throw std::invalid_argument("no such member");
    *member = b->members[i].get();
    return 0;
  });
}

int phafd_hip_batch_setup(phafd_hip_batch *b, double dt) {
  return bguard(b, [&] {
    std::vector<double> chi(b->n);
    for (int i = 0; i < b->n; i++) {
      phafd_hip_ctx *c = b->members[i].get();
      int nconj = 0, nfh = 0;
      const FixDef *cj = nullptr, *fh = nullptr;
      for (auto &f : c->fixes) {
        if (f->kind == FixDef::CONJ) { nconj++; cj = f.get(); }
        if (f->kind == FixDef::FH) { nfh++; fh = f.get(); }
        if (f->kind == FixDef::GRADPHI) throw std::invalid_argument("batch members take no grid/gradphi fix");
        if (f->kind == FixDef::SEMIFLEX)
          throw std::invalid_argument("batch members take no atom/semiflexible fix (plain runs only)");
      }
      if (nconj != 1 || nfh != 1)
        throw std::invalid_argument("each batch member needs one grid/conjugate/volfrac and one grid/floryhuggins fix");
      if (cj->noise_mode != PHAFD_NOISE_OFF) throw std::invalid_argument("batch members run with noise off");
      if (!c->fuse_fh || !c->fuse_spectral) throw std::invalid_argument("batch members keep the default options");
      const int rcs = phafd_hip_setup(c, dt);
      if (rcs != PHAFD_OK) throw std::runtime_error("member " + std::to_string(i) + ": " + c->err);
      const UpdateParams u = update_params(c, *cj);
      const double tv = fh->temp / fh->volFH;
      if (i == 0) {
        b->u = u;
        b->tv = tv;
        b->sc = spec_const(c);
      } else if (std::memcmp(&u, &b->u, sizeof(u)) != 0 || tv != b->tv) {
        throw std::invalid_argument("batch members share mobility, temp, volFH and gamma (chi may differ)");
      }
      chi[i] = fh->chi;
    }
    b->chi_b.alloc(b->n);
    HIPCHK(hipMemcpy(b->chi_b.p, chi.data(), sizeof(double) * b->n, hipMemcpyHostToDevice));
    phafd_hip_ctx *c0 = b->members[0].get();
    b->sep = c0->fuse_spectral && fft_sep_grad_ok(c0->plan);
    b->zsplit = spectral_fused(c0);
    // autotune what does not change the arithmetic: members per batched
    // launch (timed on the gradient block, which rewrites the same state) and
    // member streams for the atom side (timed on the pair pass, likewise)
    const int chunks[4] = {b->n, (b->n + 1) / 2, (b->n + 3) / 4, std::min(b->n, 8)};
    double best = 1e30;
    for (int k = 0; k < 4; k++) {
      b->tune_grad_ms[k] = time_ms(b, 3, [&] { batch_grad(b, chunks[k]); });
      if (b->tune_grad_ms[k] < best) { best = b->tune_grad_ms[k]; b->chunk = chunks[k]; }
    }
    const int maxs = std::min(b->n, 8);
    while ((int)b->mstreams.size() < maxs) {
      hipStream_t s;
      HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      b->mstreams.push_back(s);
    }
    const int ks[4] = {1, std::min(2, maxs), std::min(4, maxs), maxs};
    best = 1e30;
    int bestk = 1;
    for (int k = 0; k < 4; k++) {
      batch_assign_streams(b, ks[k]);
      b->tune_pair_ms[k] = time_ms(b, 2, [&] {
        for (auto &m : b->members) phase_pair(m.get());
      });
      if (b->tune_pair_ms[k] < best) { best = b->tune_pair_ms[k]; bestk = ks[k]; }
    }
    batch_assign_streams(b, bestk);
    // the pair phase as one launch per pass for every member, timed on the
    // same state (the chains' step left out of both timings)
    b->tune_fused_ms = 0;
    b->fused = false;
    MultiShape best_shape;
    if (fused_eligible(b)) {
      for (auto &m : b->members) fused_prepare(m.get());
      batch_sync(b);
      // the launch shape: atoms per item (threads) x persistent grid cap
      const MultiShape shapes[6] = {{256, 1024}, {256, 2048}, {256, 4096}, {128, 2048}, {128, 4096}, {64, 4096}};
      double bs = 1e30;
      for (int k = 0; k < 6; k++) {
        b->shape = shapes[k];
        b->tune_shape_ms[k] = time_ms(b, 2, [&] { batch_pair_fused(b, false); });
        if (b->tune_shape_ms[k] < bs) { bs = b->tune_shape_ms[k]; best_shape = shapes[k]; }
      }
      b->shape = best_shape;
      b->tune_fused_ms = bs;
      b->fused = b->fused_mode == 1 || (b->fused_mode < 0 && b->tune_fused_ms < best);
      b->trial = 0;
      b->steps_run = 0;
      b->decided = false;
      b->trial_ms[0] = b->trial_ms[1] = 0;
    }
    batch_grad(b, b->chunk);   // the state the first step expects (start_of_step recomputes it anyway)
    batch_sync(b);
    destroy_graphs(b);
    capture(b, &b->g_grad, &b->x_grad, [&] { batch_grad(b, b->chunk); });
    capture(b, &b->g_spec, &b->x_spec, [&] { batch_spec(b, b->chunk); });
    b->setup_done = true;
    return 0;
  });
}

int phafd_hip_batch_run(phafd_hip_batch *b, int nsteps) {
  return bguard(b, [&] {
    batch_steps(b, nsteps);
    return 0;
  });
}

int phafd_hip_batch_set_graphs(phafd_hip_batch *b, int on) {
  return bguard(b, [&] {
    b->use_graphs = on != 0;
    return 0;
  });
}

int phafd_hip_batch_set_fused(phafd_hip_batch *b, int mode) {
  return bguard(b, [&] {
    if (mode < -1 || mode > 1) throw std::invalid_argument("mode must be -1 (autotuned), 0 or 1");
    b->fused_mode = mode;
    if (b->setup_done && mode >= 0) b->fused = mode == 1 && fused_eligible(b);
    return 0;
  });
}

int phafd_hip_batch_fused_info(phafd_hip_batch *b, double *out14) {
  return bguard(b, [&] {
    out14[0] = b->fused ? 1 : 0;
    out14[1] = b->tune_fused_ms;
    double best = 1e30;
    for (int k = 0; k < 4; k++)
      if (b->tune_pair_ms[k] > 0) best = std::min(best, b->tune_pair_ms[k]);
    out14[2] = best < 1e30 ? best : 0.0;
    out14[3] = b->shape.threads;
    out14[4] = b->shape.cap;
    for (int k = 0; k < 6; k++) out14[5 + k] = b->tune_shape_ms[k];
    out14[11] = b->decided ? 1 : 0;
    out14[12] = b->trial_ms[0];
    out14[13] = b->trial_ms[1];
    return 0;
  });
}

int phafd_hip_batch_info(phafd_hip_batch *b, double *out12) {
  return bguard(b, [&] {
    out12[0] = b->n;
    out12[1] = b->chunk;
    out12[2] = b->nstreams;
    out12[3] = b->use_graphs ? 1 : 0;
    for (int k = 0; k < 4; k++) out12[4 + k] = b->tune_grad_ms[k];
    for (int k = 0; k < 4; k++) out12[8 + k] = b->tune_pair_ms[k];
    return 0;
  });
}

int phafd_hip_batch_destroy(phafd_hip_batch *b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto s : b->mstreams) (void)hipStreamSynchronize(s);
  destroy_graphs(b);
  for (auto &m : b->members) (void)phafd_hip_destroy(m.release());
  if (b->ev_grad) (void)hipEventDestroy(b->ev_grad);
  if (b->ev_pair) (void)hipEventDestroy(b->ev_pair);
  if (b->h_desc) (void)hipHostFree(b->h_desc);
  for (auto e : b->ev_m) if (e) (void)hipEventDestroy(e);
  for (auto s : b->mstreams) {
    scan_forget_stream(s);
    (void)hipStreamDestroy(s);
  }
  if (b->stream) {
    scan_forget_stream(b->stream);
    (void)hipStreamDestroy(b->stream);
  }
  delete b;
  return 0;
}

const char *phafd_hip_batch_last_error(const phafd_hip_batch *b) { return b ? b->err.c_str() : g_create_err.c_str(); }

}  // extern "C"

// ============================================================== surrogate
// Blocks of K steps with the field advanced by the network.  The atoms go
// first: K steps with lj/cut and the chains every step and the grid-atom
// force of the block-start field held.  The bead density is deposited at
// both ends of the block.  phi after K steps is then a base -- the plain
// update with its linear part exact per Fourier mode and the rest of mu held
// at its block-start value (exponential Euler) -- plus the network's
// correction from phi0 and the two densities (mean removed, so the mass is
// kept).  The check of the plain method's equation decides:
// accepted, phi is replaced; rejected, the atoms are put back and the block
// runs as K plain steps.
namespace {

struct SurrogateCfg {
  PairDef *ga = nullptr;
  const FixDef *cj = nullptr, *fh = nullptr;
  double phit = 0;
};
constexpr int SG_NB = 512;   // blocks of the residual reduction

SurrogateCfg surrogate_cfg(phafd_hip_ctx *c) {
  require_setup(c);
  if (c->fp32) throw std::invalid_argument("the surrogate runs with the FP64 form");
  SurrogateCfg g;
  int nconj = 0, nfh = 0, nga = 0;
  for (auto &f : c->fixes) {
    if (f->kind == FixDef::CONJ) { nconj++; g.cj = f.get(); }
    if (f->kind == FixDef::FH) { nfh++; g.fh = f.get(); }
    if (f->kind == FixDef::GRADPHI) throw std::invalid_argument("the surrogate takes no grid/gradphi fix");
    if (f->kind == FixDef::SEMIFLEX)
      throw std::invalid_argument("the surrogate takes no atom/semiflexible fix (plain runs only)");
  }
  if (nconj != 1 || nfh != 1)
    throw std::invalid_argument("the surrogate needs one grid/conjugate/volfrac and one grid/floryhuggins fix");
  if (g.cj->noise_mode != PHAFD_NOISE_OFF) throw std::invalid_argument("the surrogate runs with noise off");
  for (auto &p : c->pairs)
    if (!bead_pair(*p)) { nga++; g.ga = p.get(); }
  if (nga > 1) throw std::invalid_argument("the surrogate takes at most one grid-atom pair style");
  if (g.ga && g.ga->style == PHAFD_PAIR_LJISH_LINEAR)
    throw std::invalid_argument("the surrogate's mu model needs a (phi - phi_t) coupling, not gridatom/LJish/linear");
  if (g.ga) {
    g.phit = g.ga->phit.empty() ? 0.0 : g.ga->phit[0];
    for (double v : g.ga->phit)
      if (v != g.phit) throw std::invalid_argument("the surrogate needs one phi_t for all atom types");
  }
  const size_t N = c->N;
  c->sg_rho.alloc(N); c->sg_rho1.alloc(N); c->sg_out.alloc(N); c->sg_x0.alloc(8 * N); c->sg_acc.alloc(N);
  c->sg_phi.alloc(N); c->sg_mu0.alloc(N); c->sg_mu1.alloc(N); c->sg_base.alloc(N);
  c->sg_ph0.alloc(c->M); c->sg_mh0.alloc(c->M); c->sg_win.alloc(2);
  c->sg_red.alloc(2 * SG_NB + 2);
  c->sg_sum.alloc(4);
  const size_t na = (size_t)std::max(c->nowned, 1);
  c->sg_fga.alloc(na); c->sg_x.alloc(na); c->sg_ux.alloc(na); c->sg_img.alloc(na);
  return g;
}

// gathered positions current without consuming the prefetched rebuild decision
void refresh_gathered(phafd_hip_ctx *c) {
  if (!c->neigh_set) return;
  if (c->pending_decision == 0) forward_comm(c);
  else phase_neigh(c);
}

// 2^32 fixed-point units per max|eps| (a point would need ~2e9 beads at
// full kernel weight to overflow; resolution ~2e-10 max|eps|)
void deposit_density(phafd_hip_ctx *c, const SurrogateCfg &g, float *rho) {
  if (!(g.ga && c->neigh_set && c->lists_valid)) {
    fill_zero(rho, sizeof(float) * c->N, c->stream);
    return;
  }
  double emax = 0;
  for (double e : g.ga->eps) emax = std::max(emax, std::fabs(e));
  if (!(emax > 0)) emax = 1;
  fill_zero(c->sg_acc.p, sizeof(unsigned long long) * c->N, c->stream);
  launch_ga_density(c->ngathered, c->gridview<double>(), ga_params(*g.ga), c->gx.p, 4294967296.0 / emax,
                    c->sg_acc.p, c->N, rho, c->stream);
}

// sg_sum[0..1] = FP64 {sum phi0, sum rho0}
void block_sums(phafd_hip_ctx *c, const double *phi0, const float *rho0) {
  fill_zero(c->sg_sum.p, 4 * sizeof(double), c->stream);
  sum_fp64(c->N, phi0, c->sg_sum.p, c->stream);
  sum_fp64(c->N, rho0, c->sg_sum.p + 1, c->stream);
}

// base = phi0 after K steps by exponential Euler with the block-start
// density rho0.  Leaves sg_ph0 = spectrum of phi0, sg_mh0 = spectrum of
// mu0 = mu(phi0, rho0) (in sg_mu0) and sg_sum set; uses ghat[0].  carried:
// sg_ph0, sg_mh0 and sg_mu0 already hold those of this phi0 and rho0 (the
// previous block's check computed them for its accepted phi1 and rho1).
void surrogate_base(phafd_hip_ctx *c, const SurrogateCfg &g, const double *phi0, const float *rho0, double *base,
                    int K, bool carried = false) {
  FieldSet<double> &F = c->fd;
  const double tv = g.fh->temp / g.fh->volFH;
  const UpdateParams u = update_params(c, *g.cj);
  block_sums(c, phi0, rho0);
  if (!carried) {
    surrogate_mu(c->N, phi0, rho0, tv, g.fh->chi, g.phit, c->sg_mu0.p, c->stream);
    fft_r2c(c->plan, phi0, c->sg_ph0.p, c->stream);
    fft_r2c(c->plan, c->sg_mu0.p, c->sg_mh0.p, c->stream);
  }
  etd_block(c->plan, c->sg_ph0.p, c->sg_mh0.p, F.ghat[0].p, u, tv, g.fh->chi, c->sg_sum.p, K, c->stream);
  fft_c2r(c->plan, F.ghat[0].p, base, (double)c->N, c->stream);
}

// {residual, relative mass drift} of phi1 as phi0 advanced by K steps with
// densities rho0, rho1; needs sg_ph0, sg_mh0 and sg_sum from surrogate_base
// on the same phi0, rho0; leaves mu1 in sg_mu1 and the spectra of phi1 and
// mu1 in ghat[0], ghat[1]
void block_check(phafd_hip_ctx *c, const SurrogateCfg &g, const double *phi0, const double *phi1,
                 const float *rho1, int K, double out[2]) {
  FieldSet<double> &F = c->fd;
  const double tv = g.fh->temp / g.fh->volFH;
  surrogate_mu(c->N, phi1, rho1, tv, g.fh->chi, g.phit, c->sg_mu1.p, c->stream);
  fft_r2c(c->plan, phi1, F.ghat[0].p, c->stream);
  fft_r2c(c->plan, c->sg_mu1.p, F.ghat[1].p, c->stream);
  block_residual_partials(c->plan, c->sg_ph0.p, F.ghat[0].p, c->sg_mh0.p, F.ghat[1].p, update_params(c, *g.cj), tv,
                          g.fh->chi, c->sg_sum.p, K, c->sg_red.p, SG_NB, c->stream);
  double *ms = c->sg_red.p + 2 * SG_NB;
  fill_zero(ms, 2 * sizeof(double), c->stream);
  sum_fp64(c->N, phi0, ms, c->stream);
  sum_fp64(c->N, phi1, ms + 1, c->stream);
  std::vector<double> h(2 * SG_NB + 2);
  HIPCHK(hipMemcpyAsync(h.data(), c->sg_red.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  double sr = 0, sd = 0;
  for (int b = 0; b < SG_NB; b++) {
    sr += h[2 * b];
    sd += h[2 * b + 1];
  }
  out[0] = sd > 0 ? std::sqrt(sr / sd) : (sr > 0 ? INFINITY : 0.0);
  out[1] = (h[2 * SG_NB + 1] - h[2 * SG_NB]) / h[2 * SG_NB];
}

// K atom steps (neighbour, lj/cut, chains every step) with the grid-atom
// force of the current field computed once and held.
//
// Where every kernel of a step can be gated (the chains are the only
// integrator, the copy swap folds in one pass, lj/cut is computed), steps
// run in speculative windows of W: the first step takes the host's rebuild
// decision, the later ones run as if no rebuild came, and each step's
// decide records its step in a device word when a rebuild is due for the
// next; the kernels of the steps after it then return at once.  One
// read-back per window then resumes at the step after the decision, with the
// rebuild: the same steps, lists and sums as deciding after every step.
void atom_steps_frozen_field(phafd_hip_ctx *c, const SurrogateCfg &g, int K) {
  phase_start_t<double>(c);   // grad phi of the block-start field
  if (c->neigh_set) clear_decide_slot(c, c->timestep + 1);
  FixDef *fc = fused_chain_fix(c);
  bool any_lj = false;
  for (auto &pp : c->pairs) any_lj = any_lj || bead_pair(*pp);
  bool slot_dirty = false;   // a window ran since the last single step
  int since_build = 0, interval = 8;
  int s = 0;
  while (s < K) {
    const bool can_window = c->atom_windows && s > 0 && c->neigh_set && c->lists_valid && fc && any_lj && reverse_from_one_pass(c) &&
                            c->pending_decision == 0;
    const int W = can_window ? std::min(K - s, std::max(2, std::min(16, interval - since_build))) : 1;
    if (W > 1) launch_fill_i32(2, c->sg_win.p, 0, c->stream);
    else if (slot_dirty && c->neigh_set) {
      clear_decide_slot(c, c->timestep + 1);
      slot_dirty = false;
    }
    const int64_t calls0 = c->ncalls;
    for (int j = 1; j <= W; j++) {
      const StepGate gate = W > 1 ? StepGate{c->sg_win.p, j} : StepGate{};
      c->timestep++;
      if (j == 1) phase_neigh(c);
      else forward_comm(c, gate);
      const bool lists = c->neigh_set && c->lists_valid;
      if (s == 0) {
        launch_fill_d4(c->ngathered, c->gF.p, c->stream);
        launch_fill_d4(c->nowned, c->F.p, c->stream);
        if (g.ga && lists) {
          grid_atom_pass<double>(c, *g.ga, nullptr, true);
          reverse_comm(c);
        }
        launch_copy4(c->nowned, c->F.p, c->sg_fga.p, c->stream);
      }
      // F starts as the held grid-atom force; lj/cut, the only pair force of
      // these steps, writes gF from zero and the fold adds it: the same sums
      // as zeroing both, folding and adding the held force after
      bool any = false;
      if (lists) {
        for (auto &pp : c->pairs) {
          if (!bead_pair(*pp)) continue;
          if (pp->style == 2) {
            LJParams lj{pp->lj1, pp->lj2, pp->ljcutsq};
            launch_lj_force(c->ngathered, lj, c->gx.p, c->h_rowof.p, c->h_off.p, c->h_neigh.p, c->t_off.p,
                            c->t_pair.p, c->t_i.p, c->nowned, c->gF.p, nullptr, c->stream, !any, gate);
          } else {
            launch_harmonic_force(c->ngathered, HarmParams{pp->hk, pp->hcut}, c->gx.p, c->h_rowof.p, c->h_off.p,
                                  c->h_neigh.p, c->t_off.p, c->t_pair.p, c->t_i.p, c->nowned, c->gF.p, c->stream,
                                  !any, gate);
          }
          any = true;
        }
      }
      if (any) reverse_comm_from(c, c->sg_fga.p, gate);
      else launch_copy4(c->nowned, c->sg_fga.p, c->F.p, c->stream);
      if (fc) {
        integrate_both_halves(c, *fc, gate);
      } else {
        phase_initial(c);
        for (auto &f : c->fixes)
          if (f->kind != FixDef::CONJ && f->kind != FixDef::FH) integrate_half(c, *f, true);
      }
      // next step's rebuild decision (Neighbor::decide)
      if (W > 1) launch_decide(c->nowned, c->x.p, c->xhold.p, c->triggersq, c->sg_win.p + 1, nullptr, c->stream, gate);
      else if (c->neigh_set) enqueue_decide(c);
    }
    enqueue_flags_readback(c);
    if (W > 1)
      HIPCHK(hipMemcpyAsync(c->h_flags + 6, c->sg_win.p, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    if (c->ncalls != calls0) {   // rebuilt at this window's first step
      interval = std::max(2, since_build);
      since_build = 0;
    }
    if (W == 1) {
      check_step_flags(c, c->timestep);
      since_build += 1;
      s += 1;
      continue;
    }
      const int stop = c->h_flags[6], f = c->h_flags[7];
    if (f & 2) throw std::runtime_error("Non-numeric positions - simulation unstable");
    const int done = stop > 0 ? stop : W;
    c->timestep -= W - done;
    c->pending_decision = stop > 0 ? 1 : 0;
    since_build += done;
    s += done;
    slot_dirty = true;
  }
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// stats8: {blocks, accepted, fallbacks, max residual, last residual, max
// |mass drift|, ms in the network path (densities, base, network, check),
// ms in total}
void surrogate_steps(phafd_hip_ctx *c, UNet &net, int nsteps, int K, double tol, double *stats) {
  SurrogateCfg g = surrogate_cfg(c);
  if (K < 1 || nsteps < 0 || nsteps % K) throw std::invalid_argument("nsteps must be a multiple of K >= 1");
  if (net.device != c->device) throw std::invalid_argument("the network lives on another device");
  if (!unet_shape_ok(c->nx, c->ny, c->nz))
    throw std::invalid_argument("surrogate grid must have nx % 64 == 0, ny % 32 == 0, nz % 4 == 0");
  net.ensure(c->nx, c->ny, c->nz, false);
  double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const auto t_all = std::chrono::steady_clock::now();
  const size_t N = c->N;
  // After an accepted block the next one starts from exactly the checked
  // phi1 and, with no list build in between, the same bead density: the
  // check's mu1 and spectra are that block's mu0 and spectra.
  bool carry = false;
  int64_t carry_calls = -1;
  for (int blk = 0; blk < nsteps / K; blk++) {
    auto t0 = std::chrono::steady_clock::now();
    refresh_gathered(c);
    if (carry && c->ncalls != carry_calls) carry = false;
    if (!carry) deposit_density(c, g, c->sg_rho.p);
    // the atoms' block-start state, for a fallback
    const int64_t calls0 = c->ncalls, step0 = c->timestep;
    const int pending0 = c->pending_decision;
    launch_copy4(c->nowned, c->x.p, c->sg_x.p, c->stream);
    launch_copy4(c->nowned, c->ux.p, c->sg_ux.p, c->stream);
    HIPCHK(hipMemcpyAsync(c->sg_img.p, c->image.p, sizeof(int) * c->nowned, hipMemcpyDeviceToDevice, c->stream));
    st[6] += ms_since(t0);
    atom_steps_frozen_field(c, g, K);
    t0 = std::chrono::steady_clock::now();
    refresh_gathered(c);
    deposit_density(c, g, c->sg_rho1.p);
    const int64_t calls1 = c->ncalls;
    const double *phi0 = c->fd.phi.p;
    surrogate_base(c, g, phi0, c->sg_rho.p, c->sg_base.p, K, carry);
    unet_features(N, phi0, c->sg_rho.p, c->sg_rho1.p, c->sg_sum.p, net.norm, c->sg_x0.p, c->stream);
    net.forward(c->sg_x0.p, c->sg_out.p, c->stream);
    fill_zero(c->sg_sum.p + 2, sizeof(double), c->stream);
    sum_fp64(N, c->sg_out.p, c->sg_sum.p + 2, c->stream);
    unet_apply(N, c->sg_base.p, c->sg_out.p, c->sg_sum.p + 2, net.norm[4], c->sg_phi.p, c->stream);
    double chk[2];
    block_check(c, g, phi0, c->sg_phi.p, c->sg_rho1.p, K, chk);   // ends with a synchronise
    probe_collect();
    st[6] += ms_since(t0);
    st[0] += 1;
    st[3] = std::max(st[3], chk[0]);
    st[4] = chk[0];
    st[5] = std::max(st[5], std::fabs(chk[1]));
    if (chk[0] <= tol && std::fabs(chk[1]) <= 1e-12) {
      HIPCHK(hipMemcpyAsync(c->fd.phi.p, c->sg_phi.p, sizeof(double) * N, hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->h_flags + 2, c->fd.phi.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      sync(c);
      double p0;
      std::memcpy(&p0, c->h_flags + 2, sizeof(double));
      if (std::isnan(p0)) throw std::runtime_error("NAN encountered in phi.");
      st[1] += 1;
      FieldSet<double> &F = c->fd;
      carry = F.ghat[0].own && F.ghat[1].own;
      if (carry) {
        std::swap(c->sg_ph0, F.ghat[0]);
        std::swap(c->sg_mh0, F.ghat[1]);
        std::swap(c->sg_mu0, c->sg_mu1);
        std::swap(c->sg_rho, c->sg_rho1);
        carry_calls = calls1;
      }
    } else {
      carry = false;
      // put the atoms back; lists rebuilt in between are rebuilt again at
      // the restored positions
      launch_copy4(c->nowned, c->sg_x.p, c->x.p, c->stream);
      launch_copy4(c->nowned, c->sg_ux.p, c->ux.p, c->stream);
      HIPCHK(hipMemcpyAsync(c->image.p, c->sg_img.p, sizeof(int) * c->nowned, hipMemcpyDeviceToDevice, c->stream));
      c->timestep = step0;
      c->pending_decision = c->ncalls != calls0 ? 1 : pending0;
      run_steps_t<double>(c, K);
      st[2] += 1;
    }
  }
  st[7] = ms_since(t_all);
  if (stats) std::memcpy(stats, st, sizeof(st));
}

}  // namespace

extern "C" {

int phafd_hip_density(phafd_hip_ctx *c, float *rho) {
  return guard(c, [&] {
    const SurrogateCfg g = surrogate_cfg(c);
    refresh_gathered(c);
    deposit_density(c, g, c->sg_rho.p);
    HIPCHK(hipMemcpyAsync(rho, c->sg_rho.p, sizeof(float) * c->N, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return 0;
  });
}

int phafd_hip_surrogate_base(phafd_hip_ctx *c, const double *phi0, const float *rho0, int K, double *base) {
  return guard(c, [&] {
    const SurrogateCfg g = surrogate_cfg(c);
    if (!phi0 || !rho0 || !base || K < 1) throw std::invalid_argument("bad arguments");
    HIPCHK(hipMemcpyAsync(c->sg_base.p, phi0, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sg_rho.p, rho0, sizeof(float) * c->N, hipMemcpyHostToDevice, c->stream));
    DBuf<double> out;
    out.alloc(c->N);
    surrogate_base(c, g, c->sg_base.p, c->sg_rho.p, out.p, K);
    HIPCHK(hipMemcpyAsync(base, out.p, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return 0;
  });
}

int phafd_hip_block_residual(phafd_hip_ctx *c, const double *phi0, const double *phi1, const float *rho0,
                             const float *rho1, int K, double *out2) {
  return guard(c, [&] {
    const SurrogateCfg g = surrogate_cfg(c);
    if (!phi0 || !phi1 || !rho0 || !rho1 || !out2 || K < 1) throw std::invalid_argument("bad arguments");
    DBuf<double> a, b, scratch;
    a.alloc(c->N); b.alloc(c->N); scratch.alloc(c->N);
    HIPCHK(hipMemcpyAsync(a.p, phi0, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.p, phi1, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sg_rho.p, rho0, sizeof(float) * c->N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sg_rho1.p, rho1, sizeof(float) * c->N, hipMemcpyHostToDevice, c->stream));
    surrogate_base(c, g, a.p, c->sg_rho.p, scratch.p, K);   // spectra of phi0, mu0 and the sums
    block_check(c, g, a.p, b.p, c->sg_rho1.p, K, out2);
    return 0;
  });
}

int phafd_hip_run_surrogate(phafd_hip_ctx *c, phafd_unet *net, int nsteps, int K, double tol, double *stats8) {
  return guard(c, [&] {
    if (!net) throw std::invalid_argument("no network");
    surrogate_steps(c, net->net, nsteps, K, tol, stats8);
    return 0;
  });
}

}  // extern "C"
