// atoms.hpp -- kernel parameter blocks and launchers of atoms.hip.
#pragma once
#include "internal.hpp"

namespace phafd {

struct Box {
  double lo[3], hi[3], period[3];
};

// gathered atoms (LOCAL copies and GHOST images), indexed relative to nowned
struct GatheredView {
  double4 *x;       // x, y, z, type
  double4 *shift;   // periodic shift applied to the owner's position
  int *owner;       // owned index
  int *label;       // 1 LOCAL, 2 GHOST
};

// grid geometry (always FP64) and the fields of one precision
template <class R>
struct GridViewT {
  int nx, ny, nz;
  double dx, dy, dz, dv;
  double lo[3];     // sublo (grid point (0,0,0))
  const R *phi;
  R *mu;
  const R *gx, *gy, *gz;
};
using GridView = GridViewT<double>;

constexpr int GA_LJISH_LINEAR = 3;   // gridatom/LJish/linear (the public PHAFD_PAIR_LJISH_LINEAR)
constexpr int GA_MAXTYPES = 64;      // atom types of a grid-atom style
struct GAParams {
  int style;        // 0 gaussian, 1 LJish, 3 LJish/linear
  int ntypes;       // <= GA_MAXTYPES
  const double *eps, *phit, *cutsq;
  double em05, e05;   // exp(-0.5), exp(0.5) evaluated on the host
};

// Speculative windows of atom steps: stop (device) holds the 1-based step
// of the window whose rebuild decision fired (0: none); the kernels of a
// later step of the window return at once.  step 0 / stop null: not gated.
struct StepGate {
  int *stop = nullptr;
  int step = 0;
};

struct LJParams {
  double lj1, lj2, cutsq;
};
struct HarmParams {   // pair harmonic/cut <cut> <k> (src/pair_harmonic_cut.cpp:111-124)
  double k, cut;
};

struct ChainParams {
  double r0, kbond, kbend, drag;
};

void launch_pbc(int n, double4 *x, double4 *ux, int *image, const Box &b, hipStream_t s);
void launch_unmap(int n, const double4 *x, double4 *ux, const int *image, const Box &b, hipStream_t s);
// Neighbor::decide: flag |= 1 if an atom moved more than sqrt(triggersq)
// since the build, |= 2 for a non-finite position; clear (may be null) is
// zeroed by the same launch
void launch_decide(int n, const double4 *x, const double4 *xh, double triggersq, int *flag, int *clear,
                   hipStream_t s, StepGate gate = StepGate());
void launch_copy4(int n, const double4 *a, double4 *b, hipStream_t s);
// part[b] receives block b's count of LOCAL atoms (copy_classify_blocks(n)
// blocks)
int copy_classify_blocks(int n);
void launch_copy_classify(int n, const double4 *x, double zstart, double zend, double lo, double hi,
                          int *cat, int *flag, int *part, hipStream_t s);
void launch_copy_scatter(int n, const double4 *x, const int *type, const int *cat, const int *pos,
                         double zprd, int *send, GatheredView g, hipStream_t s);
void launch_img_flag(int nscan, const double4 *gx, int dim, double lo, double hi, int *flag, hipStream_t s);
void launch_img_scatter(int nscan, const int *flag, const int *pos, int dim, double shift, int first,
                        int *send, GatheredView g, hipStream_t s);
void launch_forward(int ng, const double4 *x, const int *type, GatheredView g, hipStream_t s,
                    StepGate gate = StepGate());
void launch_reverse(int cnt, const int *send, int first, double4 *target, const double4 *gF, hipStream_t s,
                    StepGate gate = StepGate());
void launch_reverse_init(int cnt, const int *send, int first, double4 *target, const double4 *init,
                         const double4 *gF, hipStream_t s, StepGate gate = StepGate());
void launch_reverse2(int ca, const int *sa, int fa, int cb, const int *sb, int fb, double4 *gF, hipStream_t s,
                     StepGate gate = StepGate());
void launch_atom2bin(int ng, const double4 *gx, const BinGeom &b, int *a2b, int *bincount, int *bad,
                     hipStream_t s);
void launch_bin_place(int ng, const int *a2b, const int *binstart, int *cursor, int *binatoms, hipStream_t s);
void launch_bin_sort(int nbins, const int *binstart, int *binatoms, const int *label, hipStream_t s);
void launch_grid_list(bool fill, const GridView &gv, const BinGeom &b, const int *stencil, int nst,
                      const int *binstart, const int *binatoms, const double4 *gx, double cutneighsq,
                      int nowned, int *cnt, const int64_t *off, int *neigh, int *overflow, hipStream_t s);
// the same with 32-bit offsets (the pruned grid-atom list)
void launch_grid_list(bool fill, const GridView &gv, const BinGeom &b, const int *stencil, int nst,
                      const int *binstart, const int *binatoms, const double4 *gx, double cutneighsq,
                      int nowned, int *cnt, const int *off, int *neigh, int *overflow, hipStream_t s);
void launch_local_flag(int ng, const int *label, int *flag, hipStream_t s);
void launch_local_scatter(int ng, const int *flag, const int *pos, int *ilist, int *rowof, hipStream_t s);
void launch_half_list(bool fill, int nrows, const int *ilist, const int *a2b, const int *binstart,
                      const int *binatoms, const int *label, const int *stencil, int nst, const double4 *gx,
                      double cutneighsq, int nowned, int *cnt, const int *off, int *neigh, int *overflow,
                      hipStream_t s);
void launch_t_count(int nent, const int *neigh, int nowned, int *tcnt, hipStream_t s);
void launch_t_place(int nrows, const int *ilist, const int *off, const int *neigh, int nowned,
                    const int *toff, int *tcur, int *tpair, int *ti, hipStream_t s);
void launch_t_sort(int ng, const int *toff, int *tpair, int *ti, hipStream_t s);
void launch_ga_deposit(const GridView &gv, const GAParams &ga, const int *cnt, const int64_t *off,
                       const int *neigh, const double4 *gx, int nowned, double *energy, hipStream_t s);
// bead density (surrogate input): rho = sum of the kernel value e of every
// in-cut pair, accumulated in acc[N] (zeroed by the caller) as integers
// e * scale and converted to FP32 (order-independent, so bit-reproducible)
void launch_ga_density(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, double scale,
                       unsigned long long *acc, size_t N, float *rho, hipStream_t s);
// The grid-atom pairs of a build (from the build positions gxb; order =
// the bin-sorted atom array): launch_ga_am_walk with off == null counts
// each atom's grid points within sqrt(reachsq) into cnt[q], with off (the
// scanned counts) writes them to ps[].x in ascending grid index, with their
// atom's q in amq (the am list); launch_ga_pm_count / _place group the am
// entries by grid point in ascending q (pcnt zeroed by the caller, poff its
// scan; key, ent scratch): ps[entry].y = its slot in the point's segment.
void launch_ga_am_walk(int ng, const GridView &gv, const double4 *gxb, const int *order, double reachsq, int *cnt,
                       const int *off, int2 *ps, int *amq, hipStream_t s);
void launch_ga_pm_count(int nent, const int2 *ps, int *pcnt, int *rank, hipStream_t s);
void launch_ga_pm_place(int nent, int2 *ps, const int *rank, const int *amq, const int *poff, int *key, int *ent,
                        size_t N, hipStream_t s);
// One grid-atom pair style (PairGridAtomGaussian / LJish / LJish/linear
// compute) over the pairs of the last build.  launch_ga_force: gF(a) -= the
// atom's terms in ascending grid index (FP32 form: Kahan-summed FP32 terms
// subtracted once); mterm[ps[entry].y] = the pair's mu term, NaN out of the
// cut; energy (may be null) += the coupling energy.
// launch_ga_list_mu: mu(g) (first: from zero, else added to) += the terms
// of its segment in order (ascending q, the reference's order).
#define PHAFD_GA_DECL(RT)                                                                                       \
  void launch_ga_force(int ng, const GridViewT<RT> &gv, const GAParams &ga, const double4 *gx, const int *order, \
                       const int *aoff, const int2 *ps, RT *mterm, double4 *gF, double *energy, hipStream_t s);  \
  void launch_ga_list_mu(size_t N, RT *mu, const int *poff, const RT *mterm, bool first, hipStream_t s);
PHAFD_GA_DECL(double)
PHAFD_GA_DECL(float)
#undef PHAFD_GA_DECL
// lj/cut forces of every gathered atom added to gF (zero_init: written from
// zero, the fill of gF left out when lj/cut is its only writer)
void launch_lj_force(int ng, const LJParams &lj, const double4 *gx, const int *rowof, const int *hoff,
                     const int *hneigh, const int *toff, const int *tpair, const int *ti, int nowned,
                     double4 *gF, double *energy, hipStream_t s, bool zero_init = false,
                     StepGate gate = StepGate());
// harmonic/cut forces over the same list and in the same order (no energy)
void launch_harmonic_force(int ng, const HarmParams &h, const double4 *gx, const int *rowof, const int *hoff,
                           const int *hneigh, const int *toff, const int *tpair, const int *ti, int nowned,
                           double4 *gF, hipStream_t s, bool zero_init = false, StepGate gate = StepGate());
void launch_drag(int n, const int *members, double4 *x, const double4 *F, double drag, double dt,
                 hipStream_t s);
// halves (1 or 2) half steps of the chains (contiguous atom ranges
// [chs[c], che[c]), at most maxlen beads, halves * maxlen * 32 bytes of LDS
// per block); 2: both halves of a timestep, with the same F
void launch_chain_step(int nchains, int maxlen, const int *chs, const int *che, const double4 *F, double4 *ux,
                       double4 *x, const int *image, const ChainParams &cp, double dt, const Box &b, hipStream_t s,
                       int halves = 1, StepGate gate = StepGate());
void launch_chain_energy(int n, const int *members, const int *cstart, const int *cend, const double4 *ux,
                         const ChainParams &cp, double *energy, hipStream_t s);
// The batched pair phase: every member of a batch in one launch per pass.
// Items are blocks of 256 atoms (or entries) of one member; boff[m] the
// first item of member m, boff[nm] their total (device arrays).
struct GAMemberDev {
  GridViewT<double> gv;
  GAParams ga;
  const double4 *gx;
  const int *order, *aoff;
  const int2 *ps;
  double *mterm;
  double4 *gF;
  int ng;
  double *mu;          // list_mu target (written from zero)
  const int *poff;
};
struct LJMemberDev {
  const double4 *gx;
  const int *rowof, *hoff, *hneigh, *toff, *tpair, *ti;
  int nowned, ng;
  double4 *gF;
  LJParams lj;
};
// one fold level of one member: swap a (ca entries) and, for a dimension's
// two disjoint swaps, swap b (cb entries; 0 otherwise)
struct FoldDev {
  int ca, fa, cb, fb;
  const int *sa, *sb;
  double4 *target;
  const double4 *src;
};
struct ChainMemberDev {
  const int *chs, *che;
  const double4 *F;
  double4 *ux, *x;
  const int *image;
  ChainParams cp;
  Box bx;
};
// launch shape of the item kernels: threads per block = atoms (entries)
// per item, and the cap on the persistent grid; the item offsets boff are
// counted in items of `threads`
struct MultiShape {
  int threads = 256, cap = 2048;
};
// gF of every gathered atom written (from zero: the fill folded in)
void launch_ga_force_multi(int nm, const GAMemberDev *d, const int *boff, int nitems, int nx, int ny,
                           const MultiShape &sh, hipStream_t s);
void launch_ga_list_mu_multi(int nm, size_t N, const GAMemberDev *d, hipStream_t s);
void launch_lj_multi(int nm, const LJMemberDev *d, const int *boff, int nitems, const MultiShape &sh,
                     hipStream_t s);
void launch_fold_multi(int nm, const FoldDev *d, const int *boff, int nitems, bool zero, const MultiShape &sh,
                       hipStream_t s);
// target[0 .. ca) = 0 of every member
void launch_fill_multi(int nm, const FoldDev *d, const int *boff, int nitems, const MultiShape &sh, hipStream_t s);
// both half steps of the harmonic chains of every member (coff: chain offsets)
void launch_chain_multi(int nm, const ChainMemberDev *d, const int *coff, int nchains, int maxlen, double dt,
                        hipStream_t s);

// atom/semiflexible bead-rod polymers (one chain per block, in LDS)
constexpr int BR_MAXLEN = 700;
struct BRParamsDev {
  double b, zpar, zperp, kb, temp, dt;
  int tether, itermax;              // tether 0 none, 1 bead 0, 2 both ends
  unsigned long long step;          // first_step calls (noise counter)
};
// first_step: ux (R0 -> R*), x = map(ux); R0, FB (Brownian force) and T
// (first-order tensions) kept per atom for second_step
void launch_br_first(int nchains, int maxlen, const int *chs, const int *che, const int *seeds, const double4 *F,
                     double4 *ux, double4 *x, const int *image, double4 *R0, double4 *FB, double *T,
                     const BRParamsDev &p, const Box &b, hipStream_t s);
// second_step: ux (R* -> R1), x = map(ux); fail |= 1 past itermax iterations
void launch_br_second(int nchains, int maxlen, const int *chs, const int *che, const double4 *F, double4 *ux,
                      double4 *x, const int *image, const double4 *R0, const double4 *FB, double *T, int *fail,
                      const BRParamsDev &p, const Box &b, hipStream_t s);
void launch_fill_i32(int n, int *p, int v, hipStream_t s);
void launch_fill_d4(int n, double4 *p, hipStream_t s);
void launch_add4(int n, const double4 *a, double4 *b, hipStream_t s);   // b.xyz += a.xyz

}  // namespace phafd
