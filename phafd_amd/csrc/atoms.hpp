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
struct GAParams {
  int style;        // 0 gaussian, 1 LJish, 3 LJish/linear
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
// grid-atom pass; order = bin-sorted atom array, npairs = device pointer to
// the number of pairs (the last entry of the scanned counts), cap = capacity
// of the pair arrays
void launch_ga_count(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, const int *order,
                     const double *cutsq, int *acnt, hipStream_t s);
// bead density (surrogate input): rho = sum of the kernel value e of every
// in-cut pair, accumulated in acc[N] (zeroed by the caller) as integers
// e * scale and converted to FP32 (order-independent, so bit-reproducible)
void launch_ga_density(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, double scale,
                       unsigned long long *acc, size_t N, float *rho, hipStream_t s);
void launch_ga_enum(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, const int *order,
                    const int *aoff, int64_t cap, int *pg, int *pq, int *overflow, hipStream_t s);
// launch_ga_enum / _pair / _fold: the atom side of the pass (pairs of each
// atom in walk order at aoff[q]; per pair the mu record and the force term;
// the force terms folded into gF in order); the FP32 form (float fields,
// float4 terms) folds forces and sums mu with Kahan compensation, energies
// stay FP64.  launch_ga_place puts
// each term's key into its grid point's segment, then its value at its rank
// in key order (sv); launch_ga_sum adds each segment in order (first: mu is
// written, summed from 0, instead of added to)
#define PHAFD_GA_DECL(RT, V4)                                                                                   \
  void launch_ga_pair(int64_t cap, const int *npairs, const GridViewT<RT> &gv, const GAParams &ga,             \
                      const double4 *gx, const int *order, const int *pg, const int *pq, V4 *ft, RT *pv,      \
                      int *ppos, int *gcnt, bool emit, double *energy, hipStream_t s);                        \
  void launch_ga_fold(int ng, const int *order, const int *aoff, const V4 *ft, double4 *gF, hipStream_t s);   \
  void launch_ga_place(int64_t cap, const int *npairs, const int *pg, const int *ppos, const int *pq,         \
                       const RT *pv, const int *goff, int *skey, RT *sv, hipStream_t s);                      \
  void launch_ga_sum(size_t N, const int *goff, const int *skey, const RT *sv, RT *mu, bool first,           \
                     hipStream_t s);
PHAFD_GA_DECL(double, double4)
PHAFD_GA_DECL(float, float4)
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
void launch_fill_i32(int n, int *p, int v, hipStream_t s);
void launch_fill_d4(int n, double4 *p, hipStream_t s);
void launch_add4(int n, const double4 *a, double4 *b, hipStream_t s);   // b.xyz += a.xyz

}  // namespace phafd
