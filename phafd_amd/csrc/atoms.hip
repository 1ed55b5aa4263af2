// atoms.hip -- atom-side kernels of the per-timestep hot path (FP64):
// periodic wrap, ghost images, binning, neighbour lists, pair forces,
// force fold-back and the overdamped integrators.
//
// Every kernel that decides list membership or position arithmetic uses the
// reference's exact expressions (built with -ffp-contract=off), so lists and
// bins are bit-identical to the reference's on the same positions.
#include "atoms.hpp"
#include "spectral_ops.hpp"

namespace phafd {
namespace {

constexpr int TPB = 256;
constexpr int IMGMASK = 1023, IMGMAX = 512, IMGBITS = 10, IMG2BITS = 20;   // src/domain.cpp:10-13

int nblk(size_t n) {
  size_t b = (n + TPB - 1) / TPB;
  if (b < 1) b = 1;
  if (b > 65535 * 4) b = 65535 * 4;
  return (int)b;
}
#define GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)(n); i += (size_t)gridDim.x * blockDim.x)

// a kernel of a speculative window's step that follows the step whose
// rebuild decision fired (StepGate)
__device__ __forceinline__ bool gate_off(const StepGate &g) {
  if (!g.stop) return false;
  const int t = *g.stop;
  return t > 0 && g.step > t;
}

// ------------------------------------------------------------ domain
// Domain::pbc + unmap, src/domain.cpp:102-178
__global__ void k_pbc(int n, double4 *x, double4 *ux, int *image, Box b) {
  GRID_STRIDE(i, n) {
    double4 p = x[i];
    int img = image[i];
    int idim, otherdims;
    if (p.x < b.lo[0]) {
      p.x += b.period[0];
      idim = img & IMGMASK; otherdims = img ^ idim; idim--; idim &= IMGMASK; img = otherdims | idim;
    }
    if (p.x >= b.hi[0]) {
      p.x -= b.period[0];
      p.x = (p.x > b.lo[0] ? p.x : b.lo[0]);
      idim = img & IMGMASK; otherdims = img ^ idim; idim++; idim &= IMGMASK; img = otherdims | idim;
    }
    if (p.y < b.lo[1]) {
      p.y += b.period[1];
      idim = (img >> IMGBITS) & IMGMASK; otherdims = img ^ (idim << IMGBITS); idim--; idim &= IMGMASK;
      img = otherdims | (idim << IMGBITS);
    }
    if (p.y >= b.hi[1]) {
      p.y -= b.period[1];
      p.y = (p.y > b.lo[1] ? p.y : b.lo[1]);
      idim = (img >> IMGBITS) & IMGMASK; otherdims = img ^ (idim << IMGBITS); idim++; idim &= IMGMASK;
      img = otherdims | (idim << IMGBITS);
    }
    if (p.z < b.lo[2]) {
      p.z += b.period[2];
      idim = img >> IMG2BITS; otherdims = img ^ (idim << IMG2BITS); idim--; idim &= IMGMASK;
      img = otherdims | (idim << IMG2BITS);
    }
    if (p.z >= b.hi[2]) {
      p.z -= b.period[2];
      p.z = (p.z > b.lo[2] ? p.z : b.lo[2]);
      idim = img >> IMG2BITS; otherdims = img ^ (idim << IMG2BITS); idim++; idim &= IMGMASK;
      img = otherdims | (idim << IMG2BITS);
    }
    x[i] = p;
    image[i] = img;
    const int xb = (img & IMGMASK) - IMGMAX, yb = (img >> IMGBITS & IMGMASK) - IMGMAX,
              zb = (img >> IMG2BITS) - IMGMAX;
    ux[i] = make_double4(p.x + xb * b.period[0], p.y + yb * b.period[1], p.z + zb * b.period[2], 0.0);
  }
}

__global__ void k_unmap(int n, const double4 *x, double4 *ux, const int *image, Box b) {
  GRID_STRIDE(i, n) {
    const double4 p = x[i];
    const int img = image[i];
    const int xb = (img & IMGMASK) - IMGMAX, yb = (img >> IMGBITS & IMGMASK) - IMGMAX,
              zb = (img >> IMG2BITS) - IMGMAX;
    ux[i] = make_double4(p.x + xb * b.period[0], p.y + yb * b.period[1], p.z + zb * b.period[2], 0.0);
  }
}

// Neighbor::check_distance, src/neighbor.cpp:79-97
// clear != nullptr: block 0 also zeroes *clear (the flag of the step after)
// gate.stop non-null: a speculative window's step; a decision records its
// step in *gate.stop (the first one wins)
__global__ void k_decide(int n, const double4 *x, const double4 *xh, double triggersq, int *flag, int *clear,
                         StepGate gate) {
  if (gate_off(gate)) return;
  if (clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0;
  int f = 0;
  GRID_STRIDE(i, n) {
    const double4 a = x[i], h = xh[i];
    const double d0 = a.x - h.x, d1 = a.y - h.y, d2 = a.z - h.z;
    const double rsq = d0 * d0 + d1 * d1 + d2 * d2;
    if (rsq > triggersq) f |= 1;
    if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z)) f |= 2;   // reported as an error
  }
  if (f) {
    atomicOr(flag, f);
    if (gate.stop) atomicCAS(gate.stop, 0, gate.step);
  }
}

__global__ void k_copy4(int n, const double4 *a, double4 *b) {
  GRID_STRIDE(i, n) b[i] = a[i];
}

// ------------------------------------------------------------- borders
// copy swap classification, src/comm_brick.cpp:299-316 (one rank):
// 1 = LOCAL copy, 2 = GHOST at z - Lz, 3 = GHOST at z + Lz, 0 = not sent
// part[block] = the block's count of LOCAL (category 1) atoms
__global__ void k_copy_classify(int n, const double4 *x, double zstart, double zend, double lo,
                                double hi, int *cat, int *flag, int *part) {
  __shared__ int wsum[TPB / 64];
  int local = 0;
  GRID_STRIDE(i, n) {
    const double z = x[i].z;
    int c = 0;
    if (z >= zstart && z <= zend) c = 1;
    else if (z >= lo) c = 2;
    else if (z <= hi) c = 3;
    cat[i] = c;
    flag[i] = c != 0;
    local += c == 1;
  }
  // every lane of the block reaches this point
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < TPB / 64; w++) t += wsum[w];
    part[blockIdx.x] = t;
  }
}
__global__ void k_copy_scatter(int n, const double4 *x, const int *type, const int *cat,
                               const int *pos, double zprd, int *send, GatheredView g) {
  GRID_STRIDE(i, n) {
    const int c = cat[i];
    if (!c) continue;
    const int q = pos[i];
    const double pz = c == 1 ? 0.0 : (c == 2 ? -zprd : zprd);
    const double4 p = x[i];
    send[q] = (int)i;
    g.x[q] = make_double4(p.x, p.y, p.z + pz, (double)type[i]);
    g.shift[q] = make_double4(0.0, 0.0, pz, 0.0);
    g.owner[q] = (int)i;
    g.label[q] = c == 1 ? 1 : 2;
  }
}
// image swap selection, src/comm_brick.cpp:420-428 (inclusive slab test)
__global__ void k_img_flag(int nscan, const double4 *gx, int dim, double lo, double hi, int *flag) {
  GRID_STRIDE(i, nscan) {
    const double4 p = gx[i];
    const double c = dim == 0 ? p.x : (dim == 1 ? p.y : p.z);
    flag[i] = (c >= lo && c <= hi) ? 1 : 0;
  }
}
// Atom::pack_border + unpack_border, src/atom.cpp:327-376
__global__ void k_img_scatter(int nscan, const int *flag, const int *pos, int dim, double shift,
                              int first, int *send, GatheredView g) {
  GRID_STRIDE(i, nscan) {
    if (!flag[i]) continue;
    const int q = pos[i];
    const int d = first + q;
    double4 p = g.x[i];
    double4 s = g.shift[i];
    if (dim == 0) { p.x = p.x + shift; s.x = shift; }
    else if (dim == 1) { p.y = p.y + shift; s.y = shift; }
    else { p.z = p.z + shift; s.z = shift; }
    send[q] = (int)i;
    g.x[d] = p;
    g.shift[d] = s;
    g.owner[d] = g.owner[i];
    g.label[d] = 2;
  }
}

// CommBrick::forward_comm, src/comm_brick.cpp:190-224: every gathered entry
// is its owner's position plus one periodic shift per coordinate (the other
// swaps add 0.0, which is exact), so one pass reproduces the swap sequence.
// Local copies (label 1) carry a zero shift, so only ghosts read theirs; the
// type in .w is the owner's (as the border swaps set it).
__global__ void k_forward(int ng, const double4 *x, const int *type, GatheredView g, StepGate gate) {
  if (gate_off(gate)) return;
  GRID_STRIDE(i, ng) {
    const int o = g.owner[i];
    const double4 p = x[o];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (g.label[i] != 1) {
      const double4 s = g.shift[i];
      sx = s.x;
      sy = s.y;
      sz = s.z;
    }
    g.x[i] = make_double4(p.x + sx, p.y + sy, p.z + sz, (double)type[o]);
  }
}

// CommBrick::reverse_comm for one swap, src/comm_brick.cpp:228-258 +
// Atom::unpack_reverse, src/atom.cpp:255-269 (targets within a swap are distinct)
__global__ void k_reverse(int cnt, const int *send, int first, double4 *target, const double4 *gF, StepGate gate) {
  if (gate_off(gate)) return;
  GRID_STRIDE(q, cnt) {
    const int j = send[q];
    const double4 b = gF[first + q];
    double4 t = target[j];
    t.x += b.x;
    t.y += b.y;
    t.z += b.z;
    target[j] = t;
  }
}
// the copy swap's fold when every owned atom is in its send list once:
// target[send[q]] = init[send[q]] + gF[first + q] (init: the owned atoms'
// force before the fold), the copy and the fold in one pass
__global__ void k_reverse_init(int cnt, const int *send, int first, double4 *target, const double4 *init,
                               const double4 *gF, StepGate gate) {
  if (gate_off(gate)) return;
  GRID_STRIDE(q, cnt) {
    const int j = send[q];
    const double4 b = gF[first + q];
    double4 t = init[j];
    t.x += b.x;
    t.y += b.y;
    t.z += b.z;
    target[j] = t;
  }
}

// the two image swaps of one dimension when no atom is in both send lists:
// items [0, ca) fold swap a, [ca, ca + cb) swap b; every target is written
// by one item, so the sums are those of two sequential launches
__global__ void k_reverse2(int ca, const int *sa, int fa, int cb, const int *sb, int fb, double4 *gF,
                           StepGate gate) {
  if (gate_off(gate)) return;
  GRID_STRIDE(q, ca + cb) {
    const bool ina = q < ca;
    const int r = ina ? q : q - ca;
    const int j = ina ? sa[r] : sb[r];
    const double4 b = gF[(ina ? fa : fb) + r];
    double4 t = gF[j];
    t.x += b.x;
    t.y += b.y;
    t.z += b.z;
    gF[j] = t;
  }
}

// ---------------------------------------------------------------- bins
// NBin::coord2bin, src/nbin.cpp:61-94
__device__ __forceinline__ int coord2bin(const BinGeom &b, double x, double y, double z) {
  int ix, iy, iz;
  if (x >= b.boxhi[0]) ix = (int)((x - b.boxhi[0]) * b.bininvx) + b.nbinx;
  else if (x >= b.boxlo[0]) { ix = (int)((x - b.boxlo[0]) * b.bininvx); ix = ix < b.nbinx - 1 ? ix : b.nbinx - 1; }
  else ix = (int)((x - b.boxlo[0]) * b.bininvx) - 1;
  if (y >= b.boxhi[1]) iy = (int)((y - b.boxhi[1]) * b.bininvy) + b.nbiny;
  else if (y >= b.boxlo[1]) { iy = (int)((y - b.boxlo[1]) * b.bininvy); iy = iy < b.nbiny - 1 ? iy : b.nbiny - 1; }
  else iy = (int)((y - b.boxlo[1]) * b.bininvy) - 1;
  if (z >= b.boxhi[2]) iz = (int)((z - b.boxhi[2]) * b.bininvz) + b.nbinz;
  else if (z >= b.boxlo[2]) { iz = (int)((z - b.boxlo[2]) * b.bininvz); iz = iz < b.nbinz - 1 ? iz : b.nbinz - 1; }
  else iz = (int)((z - b.boxlo[2]) * b.bininvz) - 1;
  return (iz - b.mbinzlo) * b.mbiny * b.mbinx + (iy - b.mbinylo) * b.mbinx + (ix - b.mbinxlo);
}

// bins outside [0, mbins-1) (non-finite or escaped positions) are clamped to
// the spare last bin and flagged, so no kernel ever indexes out of range
__global__ void k_atom2bin(int ng, const double4 *gx, BinGeom b, int *a2b, int *bincount, int *bad) {
  GRID_STRIDE(i, ng) {
    const double4 p = gx[i];
    int bin = 0;
    const bool finite = isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
    if (finite) bin = coord2bin(b, p.x, p.y, p.z);
    if (!finite || bin < 0 || bin >= b.mbins - 1) {
      bin = b.mbins - 1;
      atomicOr(bad, 1);
    }
    a2b[i] = bin;
    atomicAdd(&bincount[bin], 1);
  }
}
__global__ void k_bin_place(int ng, const int *a2b, const int *binstart, int *cursor, int *binatoms) {
  GRID_STRIDE(i, ng) {
    const int bin = a2b[i];
    const int slot = binstart[bin] + atomicAdd(&cursor[bin], 1);
    binatoms[slot] = (int)i;
  }
}
// NBinStandard::bin_atoms order (src/nbin_standard.cpp:181-229): each bin
// lists LOCAL atoms in ascending index, then GHOST atoms in ascending index.
__global__ void k_bin_sort(int nbins, const int *binstart, int *binatoms, const int *label) {
  GRID_STRIDE(b, nbins) {
    const int s = binstart[b], e = binstart[b + 1];
    for (int i = s + 1; i < e; i++) {
      const int v = binatoms[i];
      const long long kv = ((long long)(label[v] == 2) << 32) | (unsigned)v;
      int j = i - 1;
      while (j >= s) {
        const int u = binatoms[j];
        const long long ku = ((long long)(label[u] == 2) << 32) | (unsigned)u;
        if (ku <= kv) break;
        binatoms[j + 1] = u;
        j--;
      }
      binatoms[j + 1] = v;
    }
  }
}

// ------------------------------------------------------------ grid list
// NPairGridBin::build, src/npair_grid_bin.cpp:74-125: for every grid point in
// (z, y, x) order, walk the full stencil in order and each bin in its list
// order; keep atoms with rsq <= cutneighsq.  Two passes: count, then fill.
template <bool FILL, class OffT>
__global__ void k_grid_list(GridView gv, BinGeom b, const int *stencil, int nst,
                            const int *binstart, const int *binatoms, const double4 *gx,
                            double cutneighsq, int nowned, int *cnt, const OffT *off,
                            int *neigh, int *overflow) {
  const size_t N = (size_t)gv.nx * gv.ny * gv.nz;
  GRID_STRIDE(p, N) {
    const int i = (int)(p % gv.nx);
    const int j = (int)((p / gv.nx) % gv.ny);
    const int k = (int)(p / ((size_t)gv.nx * gv.ny));
    const double ztmp = (k + 0) * gv.dz + gv.lo[2];
    const double ytmp = j * gv.dy + gv.lo[1];
    const double xtmp = i * gv.dx + gv.lo[0];
    const int ibin = coord2bin(b, xtmp, ytmp, ztmp);
    int n = 0;
    OffT w = FILL ? off[p] : 0;
    const OffT wend = FILL ? off[p + 1] : 0;   // the row's scanned segment
    for (int m = 0; m < nst; m++) {
      const int bb = ibin + stencil[m];
      const int s = binstart[bb], e = binstart[bb + 1];
      for (int q = s; q < e; q++) {
        const int a = binatoms[q];
        const double4 x = gx[a];
        const double delx = xtmp - x.x;
        const double dely = ytmp - x.y;
        const double delz = ztmp - x.z;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq <= cutneighsq) {
          if (FILL) {
            if (w < wend) neigh[w++] = nowned + a;
            else atomicOr(overflow, 2);
          }
          n++;
        }
      }
    }
    if (!FILL) {
      cnt[p] = n;
      if (n > 2000) atomicOr(overflow, 1);
    }
  }
}

// ------------------------------------------------------------ half list
// local (non-GHOST) gathered atoms in ascending order: the std::set
// local_atom_indices of src/nbin_standard.cpp:236-246
__global__ void k_local_flag(int ng, const int *label, int *flag) {
  GRID_STRIDE(i, ng) flag[i] = label[i] != 2;
}
__global__ void k_local_scatter(int ng, const int *flag, const int *pos, int *ilist, int *rowof) {
  GRID_STRIDE(i, ng) {
    if (flag[i]) {
      ilist[pos[i]] = (int)i;
      rowof[i] = pos[i];
    } else {
      rowof[i] = -1;
    }
  }
}
// NPairHalfBinNewton::build, src/npair_half_bin_newton.cpp:63-120
template <bool FILL>
__global__ void k_half_list(int nrows, const int *ilist, const int *a2b, const int *binstart,
                            const int *binatoms, const int *label, const int *stencil, int nst,
                            const double4 *gx, double cutneighsq, int nowned, int *cnt,
                            const int *off, int *neigh, int *overflow) {
  GRID_STRIDE(r, nrows) {
    const int i = ilist[r];
    const double4 xi = gx[i];
    const double xtmp = xi.x, ytmp = xi.y, ztmp = xi.z;
    int n = 0;
    int w = FILL ? off[r] : 0;
    const int wend = FILL ? off[r + 1] : 0;   // the row's scanned segment
    const int ibin = a2b[i];
    // rest of i's own bin (entries after i in the bin's list order)
    const int s0 = binstart[ibin], e0 = binstart[ibin + 1];
    int q = s0;
    while (q < e0 && binatoms[q] != i) q++;
    for (q = q + 1; q < e0; q++) {
      const int j = binatoms[q];
      const double4 xj = gx[j];
      if (label[j] == 2) {
        if (xj.z < ztmp) continue;
        if (xj.z == ztmp) {
          if (xj.y < ytmp) continue;
          if (xj.y == ytmp && xj.x < xtmp) continue;
        }
      }
      const double delx = xtmp - xj.x, dely = ytmp - xj.y, delz = ztmp - xj.z;
      const double rsq = delx * delx + dely * dely + delz * delz;
      if (rsq <= cutneighsq) {
        if (FILL) {
          if (w < wend) neigh[w++] = nowned + j;
          else atomicOr(overflow, 2);
        }
        n++;
      }
    }
    for (int k = 0; k < nst; k++) {
      const int bb = ibin + stencil[k];
      const int s = binstart[bb], e = binstart[bb + 1];
      for (int qq = s; qq < e; qq++) {
        const int j = binatoms[qq];
        const double4 xj = gx[j];
        const double delx = xtmp - xj.x, dely = ytmp - xj.y, delz = ztmp - xj.z;
        const double rsq = delx * delx + dely * dely + delz * delz;
        if (rsq <= cutneighsq) {
          if (FILL) {
          if (w < wend) neigh[w++] = nowned + j;
          else atomicOr(overflow, 2);
        }
          n++;
        }
      }
    }
    if (!FILL) {
      cnt[r] = n;
      if (n > 2000) atomicOr(overflow, 1);
    }
  }
}
// transposed half list: for every atom j, the pair slots e (ascending) that
// name it as a neighbour, with the row atom i
__global__ void k_t_count(int nent, const int *neigh, int nowned, int *tcnt) {
  GRID_STRIDE(e, nent) atomicAdd(&tcnt[neigh[e] - nowned], 1);
}
__global__ void k_t_place(int nrows, const int *ilist, const int *off, const int *neigh,
                          int nowned, const int *toff, int *tcur, int *tpair, int *ti) {
  GRID_STRIDE(r, nrows) {
    for (int e = off[r]; e < off[r + 1]; e++) {
      const int j = neigh[e] - nowned;
      const int slot = toff[j] + atomicAdd(&tcur[j], 1);
      tpair[slot] = e;
      ti[slot] = ilist[r];
    }
  }
}
__global__ void k_t_sort(int ng, const int *toff, int *tpair, int *ti) {
  GRID_STRIDE(g, ng) {
    const int s = toff[g], e = toff[g + 1];
    for (int i = s + 1; i < e; i++) {
      const int v = tpair[i], vi = ti[i];
      int j = i - 1;
      while (j >= s && tpair[j] > v) {
        tpair[j + 1] = tpair[j];
        ti[j + 1] = ti[j];
        j--;
      }
      tpair[j + 1] = v;
      ti[j + 1] = vi;
    }
  }
}

// --------------------------------------------------------- pair kernels
// PairGridAtomGaussian / PairGridAtomLJish compute, the mu side
// (src/pair_gridatom_gaussian.cpp:69-118, src/pair_gridatom_lj_ish.cpp:72-178):
// one thread per grid point walks its list in order and accumulates mu.
__global__ void k_ga_deposit(GridView gv, GAParams ga, const int *cnt, const int64_t *off,
                             const int *neigh, const double4 *gx, int nowned, double *energy) {
  const size_t N = (size_t)gv.nx * gv.ny * gv.nz;
  double esum = 0.0;
  GRID_STRIDE(p, N) {
    const int c = cnt[p];
    if (c == 0) continue;
    const int k = (int)(p % gv.nx);
    const int j = (int)((p / gv.nx) % gv.ny);
    const int i = (int)(p / ((size_t)gv.nx * gv.ny));
    const double xtmp = k * gv.dx + gv.lo[0];
    const double ytmp = j * gv.dy + gv.lo[1];
    const double ztmp = i * gv.dz + gv.lo[2];
    const double ph = gv.phi[p];
    double m = gv.mu[p];
    const int64_t s = off[p];
    for (int q = 0; q < c; q++) {
      const int a = neigh[s + q] - nowned;
      const double4 x = gx[a];
      const int t = (int)x.w;
      const double delx = xtmp - x.x, dely = ytmp - x.y, delz = ztmp - x.z;
      double rsq = delx * delx + dely * dely + delz * delz;
      const double cs = ga.cutsq[t];
      const bool in = ga.style == 0 ? (rsq <= cs) : (rsq < cs);
      if (!in) continue;
      rsq /= cs;
      double expfac;
      if (ga.style == 0) expfac = (exp(-rsq / 2.0) - ga.em05) * ga.e05 * ga.eps[t];
      else expfac = exp(-1. / (1 - rsq) + 1) * ga.eps[t];
      if (ga.style == GA_LJISH_LINEAR) {   // mu += -e (no phi factor, no energy)
        m += -expfac;
        continue;
      }
      const double delphi = (ph - ga.phit[t]);
      m += 2 * delphi * expfac;
      if (energy) esum += delphi * delphi * expfac * gv.dv;
    }
    gv.mu[p] = m;
  }
  if (energy) atomicAdd(energy, esum);
}

__device__ __forceinline__ int bin1(double c, double lo, double hi, double inv, int nbin) {
  if (c >= hi) return (int)((c - hi) * inv) + nbin;
  if (c >= lo) { int i = (int)((c - lo) * inv); return i < nbin - 1 ? i : nbin - 1; }
  return (int)((c - lo) * inv) - 1;
}

// walk box of an atom: the grid points within sqrt(cs) of it, clipped to the
// grid (a hair wider than the sphere; the exact test decides)
struct GABox {
  int i0, i1, j0, j1, k0, k1;
};
template <class GV>
__device__ __forceinline__ GABox ga_box(const GV &gv, const double4 &x, double cs) {
  // products with the inverse spacings, widened by a margin far above their
  // rounding: the box only bounds the walk, the exact test decides
  const double w = sqrt(cs);
  const double mg = 1e-9;
  const double iz = 1.0 / gv.dz, iy = 1.0 / gv.dy, ix = 1.0 / gv.dx;
  GABox b;
  b.i0 = (int)ceil((x.z - w - gv.lo[2]) * iz - mg); b.i1 = (int)floor((x.z + w - gv.lo[2]) * iz + mg);
  b.j0 = (int)ceil((x.y - w - gv.lo[1]) * iy - mg); b.j1 = (int)floor((x.y + w - gv.lo[1]) * iy + mg);
  b.k0 = (int)ceil((x.x - w - gv.lo[0]) * ix - mg); b.k1 = (int)floor((x.x + w - gv.lo[0]) * ix + mg);
  b.i0 = b.i0 < 0 ? 0 : b.i0; b.j0 = b.j0 < 0 ? 0 : b.j0; b.k0 = b.k0 < 0 ? 0 : b.k0;
  b.i1 = b.i1 > gv.nz - 1 ? gv.nz - 1 : b.i1;
  b.j1 = b.j1 > gv.ny - 1 ? gv.ny - 1 : b.j1;
  b.k1 = b.k1 > gv.nx - 1 ? gv.nx - 1 : b.k1;
  return b;
}
__device__ __forceinline__ bool ga_in(const GAParams &ga, double rsq, double cs) {
  return ga.style == 0 ? (rsq <= cs) : (rsq < cs);
}
// x range of one (z, y) row of the box that can be in range: the sphere's
// chord widened by one grid point each side (the exact test stays in the
// loop, so the widening only has to cover the rounding of this estimate)
template <class GV>
__device__ __forceinline__ bool ga_row(const GV &gv, const GABox &b, double xx, double cs, double dyz,
                                       int &k0, int &k1) {
  const double rem = cs - dyz;
  if (rem < -1e-9 * cs) return false;
  const double w = sqrt(rem > 0 ? rem : 0.0);
  const int a0 = (int)floor((xx - w - gv.lo[0]) / gv.dx) - 1;
  const int a1 = (int)ceil((xx + w - gv.lo[0]) / gv.dx) + 1;
  k0 = a0 > b.k0 ? a0 : b.k0;
  k1 = a1 < b.k1 ? a1 : b.k1;
  return k0 <= k1;
}

// Grid-atom pass (PairGridAtomGaussian / PairGridAtomLJish::compute,
// src/pair_gridatom_gaussian.cpp:69-118, src/pair_gridatom_lj_ish.cpp:72-178).
// The reference walks its grid list: for every grid point the listed atoms in
// list order, adding to mu(g) and to F(aj) per in-cut pair.  F(aj) thus
// receives its terms in ascending grid index, and mu(g) its terms in list
// order: the stencil bins in (dz, dy, dx) lexicographic order
// (src/nstencil_full_bin_3d.cpp:30-36), each bin in its list order
// (src/nbin_standard.cpp:181-229).  Bins are numbered z-major, so that is
// ascending position q in the bin-sorted atom array.  Every atom within the
// pair cut of a point is in its list: the list cut is cut + skin and no atom
// moves more than skin/2 between builds.  So here, at each build, the pairs
// within (cut + skin/2) are listed twice: per atom in ascending grid index
// and per grid point in ascending q.  Each step:
//   k_ga_force: one thread per atom walks its pairs, keeps those in the cut
//     now, subtracts their force terms in order and stores each one's mu term;
//   k_ga_list_mu: one thread per grid point adds its pairs' mu terms in q
//     order.
// The exponentials are evaluated once per pair, by the atom side, with every
// lane busy; the sums are those of the reference.
template <class R> struct Vec4;
template <> struct Vec4<double> { using T = double4; };
template <> struct Vec4<float> { using T = float4; };
__device__ __forceinline__ float fexp(float x) { return expf(x); }
__device__ __forceinline__ double fexp(double x) { return exp(x); }

// Kahan-compensated running sum (FP32 form)
struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};
// running sum of a point's mu terms: FP64 plain, in order; FP32 Kahan
template <class R>
struct MuAcc {
  R m;
  __device__ __forceinline__ void add(R v) { m += v; }
  __device__ __forceinline__ R get() const { return m; }
};
template <>
struct MuAcc<float> {
  float m, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = m + y;
    c = (t - m) - y;
    m = t;
  }
  __device__ __forceinline__ float get() const { return m; }
};
// running force of one atom: FP64 subtracts each term from F in order; FP32
// sums the terms with Kahan compensation and subtracts the sum once
template <class R>
struct ForceAcc {
  double4 F;
  __device__ __forceinline__ void sub(double fx, double fy, double fz) { F.x -= fx; F.y -= fy; F.z -= fz; }
  __device__ __forceinline__ double4 get() const { return F; }
};
template <>
struct ForceAcc<float> {
  double4 F;
  Kahan kx, ky, kz;
  __device__ __forceinline__ void sub(float fx, float fy, float fz) { kx.add(fx); ky.add(fy); kz.add(fz); }
  __device__ __forceinline__ double4 get() const {
    double4 r = F;
    r.x -= (double)kx.s;
    r.y -= (double)ky.s;
    r.z -= (double)kz.s;
    return r;
  }
};

// the kernel value e of one in-cut pair (rsq already divided by cutsq): the
// expressions of the reference's loops; FP32 form: rsq in FP64 from the FP64
// positions, e in FP32
template <class R>
__device__ __forceinline__ R ga_expfac(const GAParams &ga, double rsq_over_cs, int t) {
  const R r = (R)rsq_over_cs;
  if (ga.style == 0) return (fexp(-r / R(2)) - (R)ga.em05) * (R)ga.e05 * (R)ga.eps[t];
  return fexp(R(-1) / (1 - r) + 1) * (R)ga.eps[t];
}

// Blocks are dispatched to the 8 XCDs round robin (block b to XCD b % 8),
// each XCD with its own L2.  This maps them so that each XCD takes one
// contiguous range of logical blocks (spatially neighbouring work, sharing
// field and atom lines in its L2); a bijection of [0, G).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) {
  const unsigned x = b % 8u, per = G / 8u, rem = G % 8u;
  const unsigned first = x < rem ? x * (per + 1u) : rem * (per + 1u) + (x - rem) * per;
  return first + b / 8u;
}

// ---- the pair lists of a build.  am: for each atom (bin order q) the grid
// points within reach (largest pair cut + skin/2) of its build position, in
// ascending grid index; pm: the same pairs grouped by grid point, in
// ascending q within a point (the reference's list order).  Every pair in
// the pair cut at any step before the next build is among them (no atom
// moves more than skin/2 in between).
__global__ void k_ga_am_walk(int ng, GridView gv, const double4 *__restrict__ gxb, const int *__restrict__ order,
                             double reachsq, int *__restrict__ cnt, const int *__restrict__ off,
                             int2 *__restrict__ ps, int *__restrict__ amq) {
  GRID_STRIDE(q, ng) {
    const double4 x = gxb[order[q]];
    int n = 0;
    if (isfinite(x.x) && isfinite(x.y) && isfinite(x.z)) {
      const GABox b = ga_box(gv, x, reachsq);
      int w = off ? off[q] : 0;
      for (int i = b.i0; i <= b.i1; i++) {
        const double delz = (i * gv.dz + gv.lo[2]) - x.z;
        for (int j = b.j0; j <= b.j1; j++) {
          const double dely = (j * gv.dy + gv.lo[1]) - x.y;
          for (int k = b.k0; k <= b.k1; k++) {
            const double delx = (k * gv.dx + gv.lo[0]) - x.x;
            if (delx * delx + dely * dely + delz * delz <= reachsq) {
              if (off) {
                ps[w].x = (int)(((size_t)i * gv.ny + j) * gv.nx + k);
                amq[w] = (int)q;
                w++;
              }
              n++;
            }
          }
        }
      }
    }
    if (!off) cnt[q] = n;
  }
}
// per pair: its rank among its point's pairs (arbitrary), and the point count
__global__ void k_ga_pm_count(int nent, const int2 *__restrict__ ps, int *__restrict__ pcnt, int *__restrict__ rank) {
  GRID_STRIDE(e, nent) rank[e] = atomicAdd(&pcnt[ps[e].x], 1);
}
// per pair: its place in its point's segment, carrying (q, entry)
__global__ void k_ga_pm_place(int nent, const int2 *__restrict__ ps, const int *__restrict__ rank,
                              const int *__restrict__ amq, const int *__restrict__ poff, int *__restrict__ key,
                              int *__restrict__ ent) {
  GRID_STRIDE(e, nent) {
    const int slot = poff[ps[e].x] + rank[e];
    key[slot] = amq[e];
    ent[slot] = e;
  }
}
// each point's segment into ascending q, then ps[entry].y = its slot.
// GM_U points per lane, 64 apart; segments of up to 8 are sorted in
// registers (their loads issued together), longer ones in memory.
constexpr int GM_U = 8;
__global__ __launch_bounds__(256) void k_ga_pm_sort(size_t N, const int *__restrict__ poff, int *__restrict__ key,
                                                    int *__restrict__ ent, int2 *__restrict__ ps) {
  const size_t wbase = ((size_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * (64 * GM_U);
  const int lane = threadIdx.x & 63;
  int s[GM_U], n[GM_U];
#pragma unroll
  for (int u = 0; u < GM_U; u++) {
    const size_t p = wbase + lane + 64 * u;
    s[u] = p < N ? poff[p] : 0;
    n[u] = p < N ? poff[p + 1] - s[u] : 0;
  }
#pragma unroll
  for (int u = 0; u < GM_U; u++) {
    const int a = s[u], m = n[u];
    if (m == 0) continue;
    if (m == 1) {
      ps[ent[a]].y = a;
    } else if (m == 2) {
      const int k0 = key[a], k1 = key[a + 1], e0 = ent[a], e1 = ent[a + 1];
      const bool sw = k0 > k1;
      ps[sw ? e1 : e0].y = a;
      ps[sw ? e0 : e1].y = a + 1;
    } else if (m <= 8) {
      int kk[8], ee[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        kk[i] = i < m ? key[a + i] : 0x7fffffff;
        ee[i] = i < m ? ent[a + i] : 0;
      }
      // insertion sort over the fixed 8 (keys of a segment differ)
#pragma unroll
      for (int i = 1; i < 8; i++) {
#pragma unroll
        for (int j = i; j > 0; j--) {
          const bool sw = kk[j - 1] > kk[j];
          const int tk = sw ? kk[j - 1] : kk[j], te = sw ? ee[j - 1] : ee[j];
          kk[j - 1] = sw ? kk[j] : kk[j - 1];
          ee[j - 1] = sw ? ee[j] : ee[j - 1];
          kk[j] = tk;
          ee[j] = te;
        }
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (i < m) {
          key[a + i] = kk[i];
          ent[a + i] = ee[i];
          ps[ee[i]].y = a + i;
        }
      }
    } else {
      for (int i = a + 1; i < a + m; i++) {
        const int kv = key[i], ev = ent[i];
        int j = i - 1;
        while (j >= a && key[j] > kv) {
          key[j + 1] = key[j];
          ent[j + 1] = ent[j];
          j--;
        }
        key[j + 1] = kv;
        ent[j + 1] = ev;
      }
      for (int i = a; i < a + m; i++) ps[ent[i]].y = i;
    }
  }
}

// The step's grid-atom pass, atom side: one thread per atom (bin order q)
// walks its build pairs in ascending grid index, keeps those in the pair
// cut now (the reference's test), and subtracts their force terms in that
// order.  The mu term of each pair it keeps goes to its grid point's slot,
// mterm[ps[entry].y] (NaN for a pair out of the cut).  Entries are taken
// GA_B at a time: the tests and field loads first, then the terms in
// order, so that up to GA_B pairs' loads are in flight.  Grid sides are
// powers of two (lx = log2 nx, lxy = log2 (nx ny)).
constexpr int GA_B = 8;
// the force walk of gathered atom number q of the bin-sorted order (ZI: gF
// starts from zero instead of its stored value, and atoms without pairs
// get zero written -- the fill and the walk in one)
template <class R, bool ZI>
__device__ __forceinline__ void ga_force_atom(size_t q, const GridViewT<R> &gv, const GAParams &ga,
                                              const double4 *__restrict__ gx, const int *__restrict__ order,
                                              const int *__restrict__ aoff, const int2 *__restrict__ ps,
                                              R *__restrict__ mterm, double4 *__restrict__ gF, double &esum,
                                              bool want_energy, int lx, int lxy) {
  const int mx = gv.nx - 1, my = gv.ny - 1;
  const int s = aoff[q], e = aoff[q + 1];
  const int a = order[q];
  if (s == e) {
    if (ZI) gF[a] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  const double4 x = gx[a];
  const double4 F0 = ZI ? make_double4(0.0, 0.0, 0.0, 0.0) : gF[a];
  const int t = (int)x.w;
  const double cs = ga.cutsq[t];
  const R phit = (R)ga.phit[t];
  ForceAcc<R> acc;
  acc.F = F0;
  bool any = false;
  // the pair entries of the next batch load while this batch's grid
  // values load and its terms are formed
  int2 nx[GA_B];
#pragma unroll
  for (int u = 0; u < GA_B; u++) nx[u] = s + u < e ? ps[s + u] : make_int2(0, 0);
  for (int c0 = s; c0 < e; c0 += GA_B) {
    int pp[GA_B], so[GA_B];
#pragma unroll
    for (int u = 0; u < GA_B; u++) {
      pp[u] = nx[u].x;
      so[u] = nx[u].y;
    }
#pragma unroll
    for (int u = 0; u < GA_B; u++) nx[u] = c0 + GA_B + u < e ? ps[c0 + GA_B + u] : make_int2(0, 0);
    bool in[GA_B];
    double rs[GA_B];
    R ph[GA_B], fx[GA_B], fy[GA_B], fz[GA_B];
#pragma unroll
    for (int u = 0; u < GA_B; u++) {
      const int p = pp[u];
      const int k = p & mx, j = (p >> lx) & my, i = p >> lxy;
      const double delz = (i * gv.dz + gv.lo[2]) - x.z;
      const double dely = (j * gv.dy + gv.lo[1]) - x.y;
      const double delx = (k * gv.dx + gv.lo[0]) - x.x;
      const double rsq = delx * delx + dely * dely + delz * delz;
      in[u] = c0 + u < e && ga_in(ga, rsq, cs);   // NaN positions: false
      rs[u] = rsq;
      ph[u] = fx[u] = fy[u] = fz[u] = R(0);
      if (in[u]) {
        ph[u] = gv.phi[p];
        fx[u] = gv.gx[p];
        fy[u] = gv.gy[p];
        fz[u] = gv.gz[p];
      } else if (c0 + u < e) {
        mterm[so[u]] = __builtin_nan("");
      }
    }
#pragma unroll
    for (int u = 0; u < GA_B; u++) {
      if (!in[u]) continue;
      R expfac = ga_expfac<R>(ga, rs[u] / cs, t);
      if (ga.style == GA_LJISH_LINEAR) {
        // PairGridAtomLJishLinear::compute (src/pair_gridatom_lj_ish_linear.cpp:
        // 93-160): e = -1 * bump * eps, mu += e, force weight e dv, no energy
        expfac = -expfac;
        mterm[so[u]] = expfac;
        expfac *= (R)gv.dv;
      } else {
        const R delphi = (ph[u] - phit);
        mterm[so[u]] = 2 * delphi * expfac;
        if (want_energy) esum += (double)delphi * (double)delphi * (double)expfac * gv.dv;
        expfac *= 2 * delphi * (R)gv.dv;
      }
      acc.sub(fx[u] * expfac, fy[u] * expfac, fz[u] * expfac);
      any = true;
    }
  }
  if (any || ZI) gF[a] = acc.get();
}

template <class R>
__global__ __launch_bounds__(256) void k_ga_force(int ng, GridViewT<R> gv, GAParams ga,
                                                  const double4 *__restrict__ gx, const int *__restrict__ order,
                                                  const int *__restrict__ aoff, const int2 *__restrict__ ps,
                                                  R *__restrict__ mterm,
                                                  double4 *__restrict__ gF, double *energy, int lx, int lxy) {
  double esum = 0.0;
  const size_t bl = xcd_block(blockIdx.x, gridDim.x);
  for (size_t q = bl * blockDim.x + threadIdx.x; q < (size_t)ng; q += (size_t)gridDim.x * blockDim.x)
    ga_force_atom<R, false>(q, gv, ga, gx, order, aoff, ps, mterm, gF, esum, energy != nullptr, lx, lxy);
  if (energy && esum != 0.0) atomicAdd(energy, esum);
}

// the member whose item range [boff[m], boff[m + 1]) holds item
__device__ __forceinline__ int member_of(const int *__restrict__ boff, int nm, int item) {
  int lo = 0, hi = nm;   // boff[lo] <= item < boff[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (boff[mid] <= item) lo = mid;
    else hi = mid;
  }
  return lo;
}

// every member's force walk in one launch: item = a block of 256 gathered
// atoms of one member, items taken round-robin by a grid of at most a few
// blocks per CU (gF written from zero: the fill folded in)
__global__ __launch_bounds__(256) void k_ga_force_multi(const GAMemberDev *__restrict__ d,
                                                        const int *__restrict__ boff, int nm, int lx, int lxy) {
  const int nitems = boff[nm];
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int m = member_of(boff, nm, item);
    const GAMemberDev &g = d[m];
    const size_t q = (size_t)(item - boff[m]) * blockDim.x + threadIdx.x;
    double esum = 0.0;
    if (q < (size_t)g.ng)
      ga_force_atom<double, true>(q, g.gv, g.ga, g.gx, g.order, g.aoff, g.ps, g.mterm, g.gF, esum, false, lx, lxy);
  }
}

// The step's grid-atom pass, grid side: mu(g) (first: from zero, else added
// to) += the mu terms of its pairs, held in its segment in the reference's
// list order, skipping the pairs out of the cut (NaN).  A wave takes
// 64 * GM_U consecutive points, GM_U per lane 64 apart, and issues the loads
// of its points (offsets, then up to four terms each) before using them.
template <class R>
__device__ __forceinline__ void ga_list_mu_block(size_t blk, size_t N, R *__restrict__ mu,
                                                 const int *__restrict__ poff, const R *__restrict__ mterm,
                                                 int first) {
  const size_t wbase = (blk * (blockDim.x / 64) + threadIdx.x / 64) * (64 * GM_U);
  const int lane = threadIdx.x & 63;
  int s[GM_U], n[GM_U];
  R m[GM_U];
#pragma unroll
  for (int u = 0; u < GM_U; u++) {
    const size_t p = wbase + lane + 64 * u;
    s[u] = p <= N ? poff[p] : 0;   // poff[N] is the total
    m[u] = (first || p >= N) ? R(0) : mu[p];
  }
  // a point's end is the next point's start: the next lane's, or for lane
  // 63 the next row's lane 0 (one more load after the last row)
  const size_t plast = wbase + 64 * GM_U;
  const int tail = (lane == 63 && plast <= N) ? poff[plast] : 0;
#pragma unroll
  for (int u = 0; u < GM_U; u++) {
    const int nxt = __shfl(s[u], (lane + 1) & 63, 64);
    const int row0 = u + 1 < GM_U ? __shfl(s[u + 1 < GM_U ? u + 1 : u], 0, 64) : tail;
    const size_t p = wbase + lane + 64 * u;
    n[u] = p < N ? (lane == 63 ? row0 : nxt) - s[u] : 0;
  }
  R v0[GM_U];
#pragma unroll
  for (int u = 0; u < GM_U; u++) v0[u] = n[u] > 0 ? mterm[s[u]] : R(0);
#pragma unroll
  for (int u = 0; u < GM_U; u++) {
    const size_t p = wbase + lane + 64 * u;
    if (p >= N) continue;
    MuAcc<R> acc;
    acc.m = m[u];
    if (n[u] > 0 && !isnan(v0[u])) acc.add(v0[u]);
    for (int c = s[u] + 1; c < s[u] + n[u]; c++) {
      const R x = mterm[c];
      if (!isnan(x)) acc.add(x);
    }
    mu[p] = acc.get();
  }
}
template <class R>
__global__ __launch_bounds__(256) void k_ga_list_mu(size_t N, R *__restrict__ mu, const int *__restrict__ poff,
                                                    const R *__restrict__ mterm, int first) {
  ga_list_mu_block<R>(blockIdx.x, N, mu, poff, mterm, first);
}
// every member's mu sums in one launch (blockIdx.y: the member; mu from zero)
__global__ __launch_bounds__(256) void k_ga_list_mu_multi(const GAMemberDev *__restrict__ d, size_t N) {
  const GAMemberDev &g = d[blockIdx.y];
  ga_list_mu_block<double>(blockIdx.x, N, g.mu, g.poff, g.mterm, 1);
}

// PairLJCut::compute, src/pair_lj_cut.cpp:46-119, gathered per atom: atom k
// receives +d*fpair from its own row and -d*fpair from every row that lists
// it; the reference adds these in ascending pair-slot order, which the merge
// of the (sorted) transposed slots with the own row reproduces.
__device__ __forceinline__ bool lj_pair(const double4 &xi, const double4 &xj, const LJParams &lj,
                                        double &fx, double &fy, double &fz, double &e) {
  const double delx = xi.x - xj.x, dely = xi.y - xj.y, delz = xi.z - xj.z;
  const double rsq = delx * delx + dely * dely + delz * delz;
  if (!(rsq < lj.cutsq)) return false;
  const double r2inv = 1.0 / rsq;
  const double r6inv = r2inv * r2inv * r2inv;
  const double forcelj = r6inv * (lj.lj1 * r6inv - lj.lj2);
  const double fpair = forcelj * r2inv;
  fx = delx * fpair;
  fy = dely * fpair;
  fz = delz * fpair;
  e = r6inv * (lj.lj1 / 12.0 * r6inv - lj.lj2 / 6.0);
  return true;
}
// PairHarmonicCut::compute, src/pair_harmonic_cut.cpp:45-108: fpair =
// -k (r - cut) / r inside the cut; coincident atoms are pushed apart along
// (0.001, 0.001, 0.001) as the reference does; no energy
__device__ __forceinline__ bool harm_pair(const double4 &xi, const double4 &xj, const HarmParams &h, double &fx,
                                          double &fy, double &fz, double &e) {
  double delx = xi.x - xj.x, dely = xi.y - xj.y, delz = xi.z - xj.z;
  double rij = sqrt(delx * delx + dely * dely + delz * delz);
  if (!(rij < h.cut)) return false;
  if (rij == 0) {
    delx = dely = delz = 0.001;
    rij = sqrt(delx * delx + dely * dely + delz * delz);
  }
  const double fpair = -1 * h.k * (rij - h.cut) / rij;
  fx = delx * fpair;
  fy = dely * fpair;
  fz = delz * fpair;
  e = 0.0;
  return true;
}
struct LJLaw {
  LJParams p;
  __device__ bool operator()(const double4 &a, const double4 &b, double &fx, double &fy, double &fz,
                             double &e) const {
    return lj_pair(a, b, p, fx, fy, fz, e);
  }
};
struct HarmLaw {
  HarmParams p;
  __device__ bool operator()(const double4 &a, const double4 &b, double &fx, double &fy, double &fz,
                             double &e) const {
    return harm_pair(a, b, p, fx, fy, fz, e);
  }
};
// one bead-bead pair style over the pruned half list, gathered per atom in
// the reference's slot order (the law: lj/cut or harmonic/cut)
template <class Law>
__device__ __forceinline__ void pair_force_atom(int a, const Law &law, const double4 *gx, const int *rowof,
                                                const int *hoff, const int *hneigh, const int *toff,
                                                const int *tpair, const int *ti, int nowned, double4 *gF,
                                                bool want_energy, double &esum, int zero_init) {
  const double4 xa = gx[a];
  double4 F = zero_init ? make_double4(0.0, 0.0, 0.0, 0.0) : gF[a];
  const int row = rowof[a];
  const int rs = row >= 0 ? hoff[row] : 0x7fffffff;
  const int re = row >= 0 ? hoff[row + 1] : 0x7fffffff;
  int q = toff[a];
  const int q1 = toff[a + 1];
  double fx, fy, fz, e;
  for (; q < q1 && tpair[q] < rs; q++) {
    if (law(gx[ti[q]], xa, fx, fy, fz, e)) { F.x -= fx; F.y -= fy; F.z -= fz; }
  }
  if (row >= 0) {
    for (int s = rs; s < re; s++) {
      if (law(xa, gx[hneigh[s] - nowned], fx, fy, fz, e)) {
        F.x += fx; F.y += fy; F.z += fz;
        if (want_energy) esum += e;
      }
    }
  }
  for (; q < q1; q++) {
    if (law(gx[ti[q]], xa, fx, fy, fz, e)) { F.x -= fx; F.y -= fy; F.z -= fz; }
  }
  gF[a] = F;
}
template <class Law>
__global__ void k_pair_force(int ng, Law law, const double4 *gx, const int *rowof,
                             const int *hoff, const int *hneigh, const int *toff, const int *tpair,
                             const int *ti, int nowned, double4 *gF, double *energy, int zero_init,
                             StepGate gate) {
  if (gate_off(gate)) return;
  double esum = 0.0;
  GRID_STRIDE(a, ng) {
    pair_force_atom(a, law, gx, rowof, hoff, hneigh, toff, tpair, ti, nowned, gF, energy != nullptr, esum,
                    zero_init);
  }
  if (energy) atomicAdd(energy, esum);
}
// every member's lj/cut in one launch (items: blocks of 256 gathered atoms)
__global__ __launch_bounds__(256) void k_lj_multi(const LJMemberDev *__restrict__ d, const int *__restrict__ boff,
                                                  int nm) {
  const int nitems = boff[nm];
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int m = member_of(boff, nm, item);
    const LJMemberDev &g = d[m];
    const int a = (item - boff[m]) * blockDim.x + threadIdx.x;
    double esum = 0.0;
    if (a < g.ng)
      pair_force_atom(a, LJLaw{g.lj}, g.gx, g.rowof, g.hoff, g.hneigh, g.toff, g.tpair, g.ti, g.nowned, g.gF,
                      false, esum, 0);
  }
}
// one fold level of every member (a swap, or a dimension's two disjoint
// swaps): target[send[r]] (+)= src[first + r]; zero: the target starts from
// zero (the fill folded in, one-pass copy swap)
__global__ __launch_bounds__(256) void k_fold_multi(const FoldDev *__restrict__ d, const int *__restrict__ boff,
                                                    int nm, int zero) {
  const int nitems = boff[nm];
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int m = member_of(boff, nm, item);
    const FoldDev &f = d[m];
    const int r = (item - boff[m]) * blockDim.x + threadIdx.x;
    if (r >= f.ca + f.cb) continue;
    const bool ina = r < f.ca;
    const int rr = ina ? r : r - f.ca;
    const int j = ina ? f.sa[rr] : f.sb[rr];
    const double4 b = f.src[(ina ? f.fa : f.fb) + rr];
    double4 t = zero ? make_double4(0.0, 0.0, 0.0, 0.0) : f.target[j];
    t.x += b.x;
    t.y += b.y;
    t.z += b.z;
    f.target[j] = t;
  }
}
// the owned forces of every member set to zero (before a fold that adds)
__global__ __launch_bounds__(256) void k_fill_multi(const FoldDev *__restrict__ d, const int *__restrict__ boff,
                                                    int nm) {
  const int nitems = boff[nm];
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    const int m = member_of(boff, nm, item);
    const int r = (item - boff[m]) * blockDim.x + threadIdx.x;
    if (r < d[m].ca) d[m].target[r] = make_double4(0.0, 0.0, 0.0, 0.0);
  }
}

// ---------------------------------------------------------- integrators
// FixAtomDrag half step, src/fixatom_drag.cpp:43-66
__global__ void k_drag(int n, const int *members, double4 *x, const double4 *F, double drag,
                       double dt) {
  GRID_STRIDE(q, n) {
    const int j = members[q];
    double4 p = x[j];
    const double4 f = F[j];
    p.x += drag * f.x * dt * 0.5;
    p.y += drag * f.y * dt * 0.5;
    p.z += drag * f.z * dt * 0.5;
    x[j] = p;
  }
}

// chain forces (build-defined semiflexible style), gathered per bead in a
// fixed order: bond(b-1,b), bond(b,b+1), angles centred at b-1, b, b+1
__device__ __forceinline__ void bond_f(const double4 &ua, const double4 &ub, double k, double r0,
                                       double *o) {
  const double d0 = ub.x - ua.x, d1 = ub.y - ua.y, d2 = ub.z - ua.z;
  const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double c = -k * (r - r0) / r;
  o[0] = c * d0; o[1] = c * d1; o[2] = c * d2;
}
__device__ __forceinline__ void angle_t(const double4 &ua, const double4 &ub, const double4 &uc,
                                        double *ga, double *gb) {
  const double d1[3] = {ub.x - ua.x, ub.y - ua.y, ub.z - ua.z};
  const double d2[3] = {uc.x - ub.x, uc.y - ub.y, uc.z - ub.z};
  const double r1s = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
  const double r2s = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
  const double r1 = sqrt(r1s), r2 = sqrt(r2s);
  const double dot = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
  const double inv12 = 1.0 / (r1 * r2);
  const double cs = dot * inv12;
  for (int c = 0; c < 3; c++) {
    ga[c] = d2[c] * inv12 - cs * d1[c] / r1s;
    gb[c] = d1[c] * inv12 - cs * d2[c] / r2s;
  }
}
// bonded force on bead b of a chain of nb beads (unwrapped positions u)
__device__ __forceinline__ void chain_bead_force(const double4 *u, int b, int nb, const ChainParams &cp,
                                                 double acc[3]) {
  double t[3], ga[3], gb[3];
  acc[0] = acc[1] = acc[2] = 0;
  if (b >= 1) { bond_f(u[b - 1], u[b], cp.kbond, cp.r0, t); for (int c = 0; c < 3; c++) acc[c] += t[c]; }
  if (b + 1 < nb) { bond_f(u[b + 1], u[b], cp.kbond, cp.r0, t); for (int c = 0; c < 3; c++) acc[c] += t[c]; }
  if (b >= 2) { angle_t(u[b - 2], u[b - 1], u[b], ga, gb); for (int c = 0; c < 3; c++) acc[c] += cp.kbend * gb[c]; }
  if (b >= 1 && b + 1 < nb) { angle_t(u[b - 1], u[b], u[b + 1], ga, gb); for (int c = 0; c < 3; c++) acc[c] += cp.kbend * (ga[c] - gb[c]); }
  if (b + 2 < nb) { angle_t(u[b], u[b + 1], u[b + 2], ga, gb); for (int c = 0; c < 3; c++) acc[c] += -cp.kbend * ga[c]; }
}

// One half step of the chains, one chain per block: the chain's unwrapped
// positions staged in LDS (all beads read before any moves), every bead's
// bonded force from them, then x += drag (F + Fchain) dt / 2 on the
// unwrapped position and Domain::map (src/domain.cpp:181-193) back to x.
// Chains are contiguous atom ranges [chs[c], che[c]).
// halves = 2: both half steps of a timestep in one launch (F is the same for
// both; the first half's positions stay in LDS, in the second buffer at
// us + stride, and only the second half is written out)
__device__ __forceinline__ void chain_step_body(int s, int nb, const double4 *F, double4 *ux, double4 *x,
                                                const int *image, const ChainParams &cp, double dt, const Box &bx,
                                                int halves, int stride, double4 *us) {
  double4 *cur = us, *nxt = us + stride;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) cur[b] = ux[s + b];
  __syncthreads();
  for (int h = 0; h < halves; h++) {
    const bool last = h + 1 == halves;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      double acc[3];
      chain_bead_force(cur, b, nb, cp, acc);
      const int j = s + b;
      const double4 f = F[j];
      double4 u = cur[b];
      const double tx = f.x + acc[0], ty = f.y + acc[1], tz = f.z + acc[2];
      u.x += cp.drag * tx * dt * 0.5;
      u.y += cp.drag * ty * dt * 0.5;
      u.z += cp.drag * tz * dt * 0.5;
      if (!last) {
        nxt[b] = u;
        continue;
      }
      ux[j] = u;
      const int img = image[j];
      const int xb = (img & IMGMASK) - IMGMAX, yb = (img >> IMGBITS & IMGMASK) - IMGMAX,
                zb = (img >> IMG2BITS) - IMGMAX;
      // .w of x and ux is 0 for every owned atom (set at upload, kept by
      // every kernel that moves atoms), so x is written without reading it
      x[j] = make_double4(u.x - xb * bx.period[0], u.y - yb * bx.period[1], u.z - zb * bx.period[2], u.w);
    }
    if (!last) {
      __syncthreads();
      double4 *t = cur;
      cur = nxt;
      nxt = t;
    }
  }
}
__global__ void k_chain_step(const int *chs, const int *che, const double4 *F, double4 *ux, double4 *x,
                             const int *image, ChainParams cp, double dt, Box bx, int halves, int stride,
                             StepGate gate) {
  if (gate_off(gate)) return;   // uniform over the block: *gate.stop is fixed during this launch
  extern __shared__ double4 us[];
  const int s = chs[blockIdx.x], nb = che[blockIdx.x] - s;
  chain_step_body(s, nb, F, ux, x, image, cp, dt, bx, halves, stride, us);
}
// both half steps of every member's chains in one launch: block = one chain
// (coff: the members' chain offsets)
__global__ void k_chain_multi(const ChainMemberDev *__restrict__ d, const int *__restrict__ coff, int nm, double dt,
                              int stride) {
  extern __shared__ double4 us[];
  const int m = member_of(coff, nm, blockIdx.x);
  const ChainMemberDev &g = d[m];
  const int c = blockIdx.x - coff[m];
  const int s = g.chs[c], nb = g.che[c] - s;
  chain_step_body(s, nb, g.F, g.ux, g.x, g.image, g.cp, dt, g.bx, 2, stride, us);
}
__global__ void k_chain_energy(int n, const int *members, const int *cstart, const int *cend,
                               const double4 *ux, ChainParams cp, double *energy) {
  double e = 0;
  GRID_STRIDE(q, n) {
    const int j = members[q];
    const int s = cstart[q], eidx = cend[q];
    const int b = j - s, nb = eidx - s;
    const double4 *u = ux + s;
    if (b + 1 < nb) {
      const double d0 = u[b + 1].x - u[b].x, d1 = u[b + 1].y - u[b].y, d2 = u[b + 1].z - u[b].z;
      const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      e += 0.5 * cp.kbond * (r - cp.r0) * (r - cp.r0);
    }
    if (b >= 1 && b + 1 < nb) {
      const double4 a = u[b - 1], m = u[b], c = u[b + 1];
      const double d1[3] = {m.x - a.x, m.y - a.y, m.z - a.z};
      const double d2[3] = {c.x - m.x, c.y - m.y, c.z - m.z};
      const double r1 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
      const double r2 = sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
      e += cp.kbend * (1 - (d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2]) / (r1 * r2));
    }
  }
  atomicAdd(energy, e);
}

// ------------------------------------------------------- bead-rod chains
// fix atom/semiflexible/{notether,singletether,doubletether}
// (src/fixatom_semiflexible.cpp:144-191): the bead-rod Brownian dynamics of
// Montesi, Morse & Pasquali (J. Chem. Phys. 122, 084903, 2005) -- rigid rods,
// bending kb (1 - u_j.u_{j+1}), mobility w_i [t t / zeta_para + (I - t t) /
// zeta_perp] about the bead tangent, the metric pseudo-potential
// (kT/2) ln det A of the rod-tension matrix A = G H G^T, Brownian force
// sqrt(2 kT / dt) zeta^(1/2) xi at the step start, midpoint predictor with
// first-order tensions, corrector with Newton-iterated tensions (at most
// itermax).  One 64-thread block per chain; the chain lives in LDS (29
// doubles per bead), serial tridiagonal sweeps on lane 0.

constexpr int BR_PER_BEAD = 29;
constexpr double BR_TOL = 1e-12;   // |R_{j+1} - R_j|^2 - b^2 within BR_TOL b^2
constexpr unsigned long long BR_KEY1 = 0x4245414452ull;

struct BRLds {
  double *Rx, *Ry, *Rz, *Fx, *Fy, *Fz, *tx, *ty, *tz;           // per bead
  double *ux, *uy, *uz, *il, *cs, *T;                            // per bond
  double *hAx, *hAy, *hAz, *hBx, *hBy, *hBz, *Dx, *Dy, *Dz;      // per bond
  double *r, *lo, *di, *up, *cp;                                 // per bond (tridiagonal)
  __device__ void carve(double *sh, int L) {
    double **a[BR_PER_BEAD] = {&Rx, &Ry, &Rz, &Fx, &Fy, &Fz, &tx, &ty, &tz, &ux, &uy, &uz, &il, &cs, &T,
                               &hAx, &hAy, &hAz, &hBx, &hBy, &hBz, &Dx, &Dy, &Dz, &r, &lo, &di, &up, &cp};
    for (int k = 0; k < BR_PER_BEAD; k++) *a[k] = sh + (size_t)k * L;
  }
};

__device__ __forceinline__ double br_w(const BRParamsDev &p, int i, int N) {
  return ((p.tether >= 1 && i == 0) || (p.tether >= 2 && i == N - 1)) ? 0.0 : 1.0;
}
// out = H_i v
__device__ __forceinline__ void br_mob(const BRParamsDev &p, const BRLds &S, int i, int N, double vx, double vy,
                                       double vz, double &ox, double &oy, double &oz) {
  if (br_w(p, i, N) == 0.0) { ox = oy = oz = 0.0; return; }
  const double t0 = S.tx[i], t1 = S.ty[i], t2 = S.tz[i];
  const double tv = t0 * vx + t1 * vy + t2 * vz;
  const double a = 1.0 / p.zperp, c = 1.0 / p.zpar - 1.0 / p.zperp;
  ox = vx * a + c * tv * t0;
  oy = vy * a + c * tv * t1;
  oz = vz * a + c * tv * t2;
}

// bonds, tangents and cosines of the positions in S.R
__device__ void br_geom(BRLds &S, int N) {
  const int m = N - 1, tid = threadIdx.x;
  for (int j = tid; j < m; j += blockDim.x) {
    const double d0 = S.Rx[j + 1] - S.Rx[j], d1 = S.Ry[j + 1] - S.Ry[j], d2 = S.Rz[j + 1] - S.Rz[j];
    const double l = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double il = 1.0 / l;
    S.il[j] = il;
    S.ux[j] = d0 * il;
    S.uy[j] = d1 * il;
    S.uz[j] = d2 * il;
  }
  __syncthreads();
  for (int i = tid; i < N; i += blockDim.x) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (m > 0) {
      if (i == 0 || i == N - 1) {
        const int j = i == 0 ? 0 : m - 1;
        t0 = S.ux[j]; t1 = S.uy[j]; t2 = S.uz[j];
      } else {
        const double s0 = S.ux[i - 1] + S.ux[i], s1 = S.uy[i - 1] + S.uy[i], s2 = S.uz[i - 1] + S.uz[i];
        const double n = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
        if (n > 1e-12) { t0 = s0 / n; t1 = s1 / n; t2 = s2 / n; }
        else { t0 = S.ux[i]; t1 = S.uy[i]; t2 = S.uz[i]; }
      }
    }
    S.tx[i] = t0; S.ty[i] = t1; S.tz[i] = t2;
  }
  for (int j = tid; j + 1 < m; j += blockDim.x)
    S.cs[j] = S.ux[j] * S.ux[j + 1] + S.uy[j] * S.uy[j + 1] + S.uz[j] * S.uz[j + 1];
  __syncthreads();
}

// lane 0: lo_j x_{j-1} + di_j x_j + up_j x_{j+1} = r_j, x into r (Thomas);
// a vanishing pivot leaves its row's unknown 0
__device__ void br_thomas(BRLds &S, int m) {
  if (threadIdx.x == 0) {
    double cprev = 0.0, rprev = 0.0;
    for (int j = 0; j < m; j++) {
      const double den = S.di[j] - (j > 0 ? S.lo[j] * cprev : 0.0);
      double cj, rj;
      if (fabs(den) > 1e-300) {
        cj = S.up[j] / den;
        rj = (S.r[j] - (j > 0 ? S.lo[j] * rprev : 0.0)) / den;
      } else {
        cj = 0.0;
        rj = 0.0;
      }
      S.cp[j] = cj;
      S.r[j] = rj;
      cprev = cj;
      rprev = rj;
    }
    for (int j = m - 2; j >= 0; j--) {
      rprev = S.r[j] - S.cp[j] * rprev;
      S.r[j] = rprev;
    }
  }
  __syncthreads();
}

// S.F += bending and metric pseudo-potential forces (the tension matrix A
// as a function of the cosines: A_jj = (w_j + w_{j+1}) / zeta_perp + kh
// (w_j (t_j.u_j)^2 + w_{j+1} (t_{j+1}.u_j)^2), A_{j,j+1} = -w_{j+1} (c_j /
// zeta_perp + kh (1 + c_j) / 2), kh = 1/zeta_para - 1/zeta_perp; the forward
// and backward pivots give diag and first off-diagonal of A^-1).  The
// scratch arrays of the Newton iteration hold a, off, pivots and weights.
__device__ void br_bend_ps(const BRParamsDev &p, BRLds &S, int N) {
  const int m = N - 1, tid = threadIdx.x;
  if (m < 2) return;
  double *a = S.Dx, *off = S.Dy, *al = S.Dz, *be = S.lo, *dinv = S.di, *e = S.up;
  const double ip = 1.0 / p.zperp, kh = 1.0 / p.zpar - 1.0 / p.zperp;
  for (int j = tid; j < m; j += blockDim.x) {
    const double aL = j == 0 ? 1.0 : 0.5 * (1.0 + S.cs[j - 1]);
    const double aR = j + 1 == N - 1 ? 1.0 : 0.5 * (1.0 + S.cs[j]);
    const double w0 = br_w(p, j, N), w1 = br_w(p, j + 1, N);
    a[j] = (w0 + w1) * ip + kh * (w0 * aL + w1 * aR);
    off[j] = j + 1 < m ? -w1 * (S.cs[j] * ip + kh * 0.5 * (1.0 + S.cs[j])) : 0.0;
  }
  __syncthreads();
  // forward pivots on lane 0, backward pivots on lane 1
  if (tid == 0) {
    double v = a[0];
    al[0] = v;
    for (int j = 1; j < m; j++) {
      v = fabs(v) > 1e-300 ? a[j] - off[j - 1] * off[j - 1] / v : a[j];
      al[j] = v;
    }
  } else if (tid == 1) {
    double v = a[m - 1];
    be[m - 1] = v;
    for (int j = m - 2; j >= 0; j--) {
      v = fabs(v) > 1e-300 ? a[j] - off[j] * off[j] / v : a[j];
      be[j] = v;
    }
  }
  __syncthreads();
  for (int j = tid; j < m; j += blockDim.x) {
    const double den = al[j] + be[j] - a[j];
    dinv[j] = fabs(den) > 1e-300 ? 1.0 / den : 0.0;
  }
  __syncthreads();
  for (int j = tid; j + 1 < m; j += blockDim.x) {
    double ej = p.kb;
    if (p.temp != 0.0) {
      const double oinv = fabs(be[j + 1]) > 1e-300 ? -off[j] / be[j + 1] * dinv[j] : 0.0;
      const double w1 = br_w(p, j + 1, N);
      const double dlog = 2.0 * oinv * (-w1 * (ip + 0.5 * kh)) + 0.5 * kh * w1 * (dinv[j] + dinv[j + 1]);
      ej -= 0.5 * p.temp * dlog;
    }
    e[j] = ej;
  }
  __syncthreads();
  // bead i gathers, in the order of the bond-pair loop: pair i-2 (+e q),
  // pair i-1 (+e (p - q)), pair i (-e p)
  for (int i = tid; i < N; i += blockDim.x) {
    double f[3] = {S.Fx[i], S.Fy[i], S.Fz[i]};
    for (int j = i - 2; j <= i; j++) {
      if (j < 0 || j + 1 >= m) continue;
      const double u0[3] = {S.ux[j], S.uy[j], S.uz[j]}, u1[3] = {S.ux[j + 1], S.uy[j + 1], S.uz[j + 1]};
      const double c = S.cs[j];
      for (int q = 0; q < 3; q++) {
        const double pj = S.il[j] * (u1[q] - c * u0[q]);
        const double qj = S.il[j + 1] * (u0[q] - c * u1[q]);
        if (j == i - 2) f[q] += e[j] * qj;
        else if (j == i - 1) f[q] += e[j] * (pj - qj);
        else f[q] -= e[j] * pj;
      }
    }
    S.Fx[i] = f[0]; S.Fy[i] = f[1]; S.Fz[i] = f[2];
  }
  __syncthreads();
}

__device__ __forceinline__ double4 br_map(double x, double y, double z, int img, const Box &bx) {
  const int xb = (img & IMGMASK) - IMGMAX, yb = (img >> IMGBITS & IMGMASK) - IMGMAX, zb = (img >> IMG2BITS) - IMGMAX;
  return make_double4(x - xb * bx.period[0], y - yb * bx.period[1], z - zb * bx.period[2], 0.0);
}

// FixAtomSemiFlexible::initial_integrate -> first_step
__global__ __launch_bounds__(64) void k_br_first(const int *chs, const int *che, const int *seeds, const double4 *F,
                                                 double4 *ux, double4 *x, const int *image, double4 *R0,
                                                 double4 *FB, double *Tg, BRParamsDev p, Box bx, int L) {
  extern __shared__ double sh[];
  BRLds S;
  S.carve(sh, L);
  const int s = chs[blockIdx.x], N = che[blockIdx.x] - s, m = N - 1, tid = threadIdx.x;
  for (int i = tid; i < N; i += blockDim.x) {
    const double4 u = ux[s + i];
    S.Rx[i] = u.x; S.Ry[i] = u.y; S.Rz[i] = u.z;
    R0[s + i] = u;
  }
  __syncthreads();
  br_geom(S, N);
  const double amp = p.temp > 0 ? sqrt(2.0 * p.temp / p.dt) : 0.0;
  const double s12 = sqrt(12.0), sp = sqrt(p.zpar), sq = sqrt(p.zperp);
  const unsigned long long key0 = (unsigned long long)(unsigned)seeds[blockIdx.x];
  for (int i = tid; i < N; i += blockDim.x) {
    double xi[3] = {0.0, 0.0, 0.0};
    if (amp != 0.0) {
      unsigned long long w[4] = {(unsigned long long)i, p.step, 0ull, 0ull};
      philox4x64_10(w, key0, BR_KEY1);
      for (int q = 0; q < 3; q++) xi[q] = s12 * unit_centered(w[q]);
    }
    const double t[3] = {S.tx[i], S.ty[i], S.tz[i]};
    const double tx = t[0] * xi[0] + t[1] * xi[1] + t[2] * xi[2];
    double fb[3];
    for (int q = 0; q < 3; q++) fb[q] = amp * (sp * tx * t[q] + sq * (xi[q] - tx * t[q]));
    FB[s + i] = make_double4(fb[0], fb[1], fb[2], 0.0);
    const double4 f = F[s + i];
    S.Fx[i] = f.x + fb[0];
    S.Fy[i] = f.y + fb[1];
    S.Fz[i] = f.z + fb[2];
  }
  __syncthreads();
  br_bend_ps(p, S, N);
  // first-order tensions: u_j.(V_{j+1} - V_j) = 0
  for (int j = tid; j < m; j += blockDim.x) {
    const double u0 = S.ux[j], u1 = S.uy[j], u2 = S.uz[j];
    double h0x, h0y, h0z, h1x, h1y, h1z;
    br_mob(p, S, j, N, u0, u1, u2, h0x, h0y, h0z);
    br_mob(p, S, j + 1, N, u0, u1, u2, h1x, h1y, h1z);
    S.di[j] = -((u0 * h1x + u1 * h1y + u2 * h1z) + (u0 * h0x + u1 * h0y + u2 * h0z));
    double lo = 0.0, up = 0.0;
    if (j > 0) {
      double hx, hy, hz;
      br_mob(p, S, j, N, S.ux[j - 1], S.uy[j - 1], S.uz[j - 1], hx, hy, hz);
      lo = u0 * hx + u1 * hy + u2 * hz;
    }
    if (j + 1 < m) {
      double hx, hy, hz;
      br_mob(p, S, j + 1, N, S.ux[j + 1], S.uy[j + 1], S.uz[j + 1], hx, hy, hz);
      up = u0 * hx + u1 * hy + u2 * hz;
    }
    S.lo[j] = lo;
    S.up[j] = up;
    double f0x, f0y, f0z, f1x, f1y, f1z;
    br_mob(p, S, j, N, S.Fx[j], S.Fy[j], S.Fz[j], f0x, f0y, f0z);
    br_mob(p, S, j + 1, N, S.Fx[j + 1], S.Fy[j + 1], S.Fz[j + 1], f1x, f1y, f1z);
    const double dfx = f1x - f0x, dfy = f1y - f0y, dfz = f1z - f0z;
    S.r[j] = -(u0 * dfx + u1 * dfy + u2 * dfz);
  }
  __syncthreads();
  br_thomas(S, m);
  for (int i = tid; i < N; i += blockDim.x) {
    double tot[3] = {S.Fx[i], S.Fy[i], S.Fz[i]};
    const double uu[3] = {i < m ? S.ux[i] : 0.0, i < m ? S.uy[i] : 0.0, i < m ? S.uz[i] : 0.0};
    const double ul[3] = {i > 0 ? S.ux[i - 1] : 0.0, i > 0 ? S.uy[i - 1] : 0.0, i > 0 ? S.uz[i - 1] : 0.0};
    for (int q = 0; q < 3; q++) {
      if (i < m) tot[q] += S.r[i] * uu[q];
      if (i > 0) tot[q] -= S.r[i - 1] * ul[q];
    }
    double vx, vy, vz;
    br_mob(p, S, i, N, tot[0], tot[1], tot[2], vx, vy, vz);
    const double4 r0 = R0[s + i];
    const double nx = r0.x + 0.5 * p.dt * vx, ny = r0.y + 0.5 * p.dt * vy, nz = r0.z + 0.5 * p.dt * vz;
    ux[s + i] = make_double4(nx, ny, nz, r0.w);
    x[s + i] = br_map(nx, ny, nz, image[s + i], bx);
    if (i < m) Tg[s + i] = S.r[i];
  }
}

// FixAtomSemiFlexible::final_integrate -> second_step(..., itermax): R1 = R0
// + dt V(R*), tensions by Newton from the first step's; fail |= 1 when the
// rods were not brought to length within itermax iterations
__global__ __launch_bounds__(64) void k_br_second(const int *chs, const int *che, const double4 *F, double4 *ux,
                                                  double4 *x, const int *image, const double4 *R0,
                                                  const double4 *FB, double *Tg, int *fail, BRParamsDev p, Box bx,
                                                  int L) {
  extern __shared__ double sh[];
  BRLds S;
  S.carve(sh, L);
  const int s = chs[blockIdx.x], N = che[blockIdx.x] - s, m = N - 1, tid = threadIdx.x;
  for (int i = tid; i < N; i += blockDim.x) {
    const double4 u = ux[s + i];
    S.Rx[i] = u.x; S.Ry[i] = u.y; S.Rz[i] = u.z;
  }
  __syncthreads();
  br_geom(S, N);
  for (int i = tid; i < N; i += blockDim.x) {
    const double4 f = F[s + i], fb = FB[s + i];
    S.Fx[i] = f.x + fb.x;
    S.Fy[i] = f.y + fb.y;
    S.Fz[i] = f.z + fb.z;
  }
  __syncthreads();
  br_bend_ps(p, S, N);
  // P = R0 + dt H F (into R), hA_j = H_j u_j, hB_j = H_{j+1} u_j, T from the first step
  for (int i = tid; i < N; i += blockDim.x) {
    double vx, vy, vz;
    br_mob(p, S, i, N, S.Fx[i], S.Fy[i], S.Fz[i], vx, vy, vz);
    const double4 r0 = R0[s + i];
    S.Rx[i] = r0.x + p.dt * vx;
    S.Ry[i] = r0.y + p.dt * vy;
    S.Rz[i] = r0.z + p.dt * vz;
  }
  for (int j = tid; j < m; j += blockDim.x) {
    br_mob(p, S, j, N, S.ux[j], S.uy[j], S.uz[j], S.hAx[j], S.hAy[j], S.hAz[j]);
    br_mob(p, S, j + 1, N, S.ux[j], S.uy[j], S.uz[j], S.hBx[j], S.hBy[j], S.hBz[j]);
    S.T[j] = Tg[s + j];
  }
  __syncthreads();
  const double tol = BR_TOL * p.b * p.b, dt = p.dt;
  int it = 0;
  for (;; it++) {
    double gmax = 0.0;
    for (int j = tid; j < m; j += blockDim.x) {
      double tq0 = -S.T[j] * S.hBx[j] - S.T[j] * S.hAx[j];
      double tq1 = -S.T[j] * S.hBy[j] - S.T[j] * S.hAy[j];
      double tq2 = -S.T[j] * S.hBz[j] - S.T[j] * S.hAz[j];
      if (j + 1 < m) {
        tq0 += S.T[j + 1] * S.hAx[j + 1];
        tq1 += S.T[j + 1] * S.hAy[j + 1];
        tq2 += S.T[j + 1] * S.hAz[j + 1];
      }
      if (j > 0) {
        tq0 += S.T[j - 1] * S.hBx[j - 1];
        tq1 += S.T[j - 1] * S.hBy[j - 1];
        tq2 += S.T[j - 1] * S.hBz[j - 1];
      }
      const double d0 = (S.Rx[j + 1] - S.Rx[j]) + dt * tq0;
      const double d1 = (S.Ry[j + 1] - S.Ry[j]) + dt * tq1;
      const double d2 = (S.Rz[j + 1] - S.Rz[j]) + dt * tq2;
      S.Dx[j] = d0; S.Dy[j] = d1; S.Dz[j] = d2;
      const double g = (d0 * d0 + d1 * d1 + d2 * d2) - p.b * p.b;
      S.r[j] = g;
      gmax = fmax(gmax, fabs(g));
    }
    for (int o = 32; o > 0; o >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, o, 64));
    if (gmax <= tol) break;                      // uniform: every lane holds the block maximum
    if (it == p.itermax) { it = p.itermax + 1; break; }
    __syncthreads();
    for (int j = tid; j < m; j += blockDim.x) {
      const double d0 = S.Dx[j], d1 = S.Dy[j], d2 = S.Dz[j];
      S.di[j] = -2.0 * dt * (d0 * (S.hAx[j] + S.hBx[j]) + d1 * (S.hAy[j] + S.hBy[j]) + d2 * (S.hAz[j] + S.hBz[j]));
      S.lo[j] = j > 0 ? 2.0 * dt * (d0 * S.hBx[j - 1] + d1 * S.hBy[j - 1] + d2 * S.hBz[j - 1]) : 0.0;
      S.up[j] = j + 1 < m ? 2.0 * dt * (d0 * S.hAx[j + 1] + d1 * S.hAy[j + 1] + d2 * S.hAz[j + 1]) : 0.0;
    }
    __syncthreads();
    br_thomas(S, m);
    for (int j = tid; j < m; j += blockDim.x) S.T[j] -= S.r[j];
    __syncthreads();
  }
  __syncthreads();
  for (int i = tid; i < N; i += blockDim.x) {
    double v[3] = {S.Rx[i], S.Ry[i], S.Rz[i]};
    if (i < m) { v[0] += dt * S.T[i] * S.hAx[i]; v[1] += dt * S.T[i] * S.hAy[i]; v[2] += dt * S.T[i] * S.hAz[i]; }
    if (i > 0) {
      v[0] -= dt * S.T[i - 1] * S.hBx[i - 1];
      v[1] -= dt * S.T[i - 1] * S.hBy[i - 1];
      v[2] -= dt * S.T[i - 1] * S.hBz[i - 1];
    }
    const double w = ux[s + i].w;
    ux[s + i] = make_double4(v[0], v[1], v[2], w);
    x[s + i] = br_map(v[0], v[1], v[2], image[s + i], bx);
  }
  if (tid == 0 && it > p.itermax) atomicOr(fail, 1);
}

__global__ void k_fill_i32(int n, int *p, int v) { GRID_STRIDE(i, n) p[i] = v; }
__global__ void k_fill_d4(int n, double4 *p) { GRID_STRIDE(i, n) p[i] = make_double4(0, 0, 0, 0); }

}  // namespace

// ================================================================ launchers
void launch_pbc(int n, double4 *x, double4 *ux, int *image, const Box &b, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pbc, dim3(nblk(n)), dim3(TPB), 0, s, n, x, ux, image, b);
  HIPCHK(hipGetLastError());
}
void launch_unmap(int n, const double4 *x, double4 *ux, const int *image, const Box &b, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_unmap, dim3(nblk(n)), dim3(TPB), 0, s, n, x, ux, image, b);
  HIPCHK(hipGetLastError());
}
void launch_decide(int n, const double4 *x, const double4 *xh, double triggersq, int *flag, int *clear,
                   hipStream_t s, StepGate gate) {
  if (n <= 0) {
    if (clear) launch_fill_i32(1, clear, 0, s);
    return;
  }
  hipLaunchKernelGGL(k_decide, dim3(nblk(n)), dim3(TPB), 0, s, n, x, xh, triggersq, flag, clear, gate);
  HIPCHK(hipGetLastError());
}
void launch_copy4(int n, const double4 *a, double4 *b, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy4, dim3(nblk(n)), dim3(TPB), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}
int copy_classify_blocks(int n) { return nblk(n); }
void launch_copy_classify(int n, const double4 *x, double zstart, double zend, double lo, double hi,
                          int *cat, int *flag, int *part, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy_classify, dim3(nblk(n)), dim3(TPB), 0, s, n, x, zstart, zend, lo, hi, cat, flag, part);
  HIPCHK(hipGetLastError());
}
void launch_copy_scatter(int n, const double4 *x, const int *type, const int *cat, const int *pos,
                         double zprd, int *send, GatheredView g, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy_scatter, dim3(nblk(n)), dim3(TPB), 0, s, n, x, type, cat, pos, zprd, send, g);
  HIPCHK(hipGetLastError());
}
void launch_img_flag(int nscan, const double4 *gx, int dim, double lo, double hi, int *flag, hipStream_t s) {
  if (nscan <= 0) return;
  hipLaunchKernelGGL(k_img_flag, dim3(nblk(nscan)), dim3(TPB), 0, s, nscan, gx, dim, lo, hi, flag);
  HIPCHK(hipGetLastError());
}
void launch_img_scatter(int nscan, const int *flag, const int *pos, int dim, double shift, int first,
                        int *send, GatheredView g, hipStream_t s) {
  if (nscan <= 0) return;
  hipLaunchKernelGGL(k_img_scatter, dim3(nblk(nscan)), dim3(TPB), 0, s, nscan, flag, pos, dim, shift, first, send, g);
  HIPCHK(hipGetLastError());
}
void launch_forward(int ng, const double4 *x, const int *type, GatheredView g, hipStream_t s, StepGate gate) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_forward, dim3(nblk(ng)), dim3(TPB), 0, s, ng, x, type, g, gate);
  HIPCHK(hipGetLastError());
}
void launch_reverse(int cnt, const int *send, int first, double4 *target, const double4 *gF, hipStream_t s,
                    StepGate gate) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(k_reverse, dim3(nblk(cnt)), dim3(TPB), 0, s, cnt, send, first, target, gF, gate);
  HIPCHK(hipGetLastError());
}
void launch_reverse_init(int cnt, const int *send, int first, double4 *target, const double4 *init,
                         const double4 *gF, hipStream_t s, StepGate gate) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(k_reverse_init, dim3(nblk(cnt)), dim3(TPB), 0, s, cnt, send, first, target, init, gF, gate);
  HIPCHK(hipGetLastError());
}
void launch_reverse2(int ca, const int *sa, int fa, int cb, const int *sb, int fb, double4 *gF, hipStream_t s,
                     StepGate gate) {
  if (ca + cb <= 0) return;
  hipLaunchKernelGGL(k_reverse2, dim3(nblk(ca + cb)), dim3(TPB), 0, s, ca, sa, fa, cb, sb, fb, gF, gate);
  HIPCHK(hipGetLastError());
}
void launch_atom2bin(int ng, const double4 *gx, const BinGeom &b, int *a2b, int *bincount, int *bad,
                     hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_atom2bin, dim3(nblk(ng)), dim3(TPB), 0, s, ng, gx, b, a2b, bincount, bad);
  HIPCHK(hipGetLastError());
}
void launch_bin_place(int ng, const int *a2b, const int *binstart, int *cursor, int *binatoms, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_bin_place, dim3(nblk(ng)), dim3(TPB), 0, s, ng, a2b, binstart, cursor, binatoms);
  HIPCHK(hipGetLastError());
}
void launch_bin_sort(int nbins, const int *binstart, int *binatoms, const int *label, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_sort, dim3(nblk(nbins)), dim3(TPB), 0, s, nbins, binstart, binatoms, label);
  HIPCHK(hipGetLastError());
}
template <class OffT>
void launch_grid_list_t(bool fill, const GridView &gv, const BinGeom &b, const int *stencil, int nst,
                        const int *binstart, const int *binatoms, const double4 *gx, double cutneighsq, int nowned,
                        int *cnt, const OffT *off, int *neigh, int *overflow, hipStream_t s) {
  const size_t N = (size_t)gv.nx * gv.ny * gv.nz;
  if (fill)
    hipLaunchKernelGGL((k_grid_list<true, OffT>), dim3(nblk(N)), dim3(TPB), 0, s, gv, b, stencil, nst, binstart,
                       binatoms, gx, cutneighsq, nowned, cnt, off, neigh, overflow);
  else
    hipLaunchKernelGGL((k_grid_list<false, OffT>), dim3(nblk(N)), dim3(TPB), 0, s, gv, b, stencil, nst, binstart,
                       binatoms, gx, cutneighsq, nowned, cnt, off, neigh, overflow);
  HIPCHK(hipGetLastError());
}
void launch_grid_list(bool fill, const GridView &gv, const BinGeom &b, const int *stencil, int nst,
                      const int *binstart, const int *binatoms, const double4 *gx, double cutneighsq,
                      int nowned, int *cnt, const int64_t *off, int *neigh, int *overflow, hipStream_t s) {
  launch_grid_list_t(fill, gv, b, stencil, nst, binstart, binatoms, gx, cutneighsq, nowned, cnt, off, neigh, overflow,
                     s);
}
void launch_grid_list(bool fill, const GridView &gv, const BinGeom &b, const int *stencil, int nst,
                      const int *binstart, const int *binatoms, const double4 *gx, double cutneighsq,
                      int nowned, int *cnt, const int *off, int *neigh, int *overflow, hipStream_t s) {
  launch_grid_list_t(fill, gv, b, stencil, nst, binstart, binatoms, gx, cutneighsq, nowned, cnt, off, neigh, overflow,
                     s);
}
void launch_local_flag(int ng, const int *label, int *flag, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_local_flag, dim3(nblk(ng)), dim3(TPB), 0, s, ng, label, flag);
  HIPCHK(hipGetLastError());
}
void launch_local_scatter(int ng, const int *flag, const int *pos, int *ilist, int *rowof, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_local_scatter, dim3(nblk(ng)), dim3(TPB), 0, s, ng, flag, pos, ilist, rowof);
  HIPCHK(hipGetLastError());
}
void launch_half_list(bool fill, int nrows, const int *ilist, const int *a2b, const int *binstart,
                      const int *binatoms, const int *label, const int *stencil, int nst, const double4 *gx,
                      double cutneighsq, int nowned, int *cnt, const int *off, int *neigh, int *overflow,
                      hipStream_t s) {
  if (nrows <= 0) return;
  if (fill)
    hipLaunchKernelGGL((k_half_list<true>), dim3(nblk(nrows)), dim3(TPB), 0, s, nrows, ilist, a2b, binstart,
                       binatoms, label, stencil, nst, gx, cutneighsq, nowned, cnt, off, neigh, overflow);
  else
    hipLaunchKernelGGL((k_half_list<false>), dim3(nblk(nrows)), dim3(TPB), 0, s, nrows, ilist, a2b, binstart,
                       binatoms, label, stencil, nst, gx, cutneighsq, nowned, cnt, off, neigh, overflow);
  HIPCHK(hipGetLastError());
}
void launch_t_count(int nent, const int *neigh, int nowned, int *tcnt, hipStream_t s) {
  if (nent <= 0) return;
  hipLaunchKernelGGL(k_t_count, dim3(nblk(nent)), dim3(TPB), 0, s, nent, neigh, nowned, tcnt);
  HIPCHK(hipGetLastError());
}
void launch_t_place(int nrows, const int *ilist, const int *off, const int *neigh, int nowned,
                    const int *toff, int *tcur, int *tpair, int *ti, hipStream_t s) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_t_place, dim3(nblk(nrows)), dim3(TPB), 0, s, nrows, ilist, off, neigh, nowned, toff,
                     tcur, tpair, ti);
  HIPCHK(hipGetLastError());
}
void launch_t_sort(int ng, const int *toff, int *tpair, int *ti, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_t_sort, dim3(nblk(ng)), dim3(TPB), 0, s, ng, toff, tpair, ti);
  HIPCHK(hipGetLastError());
}
void launch_ga_deposit(const GridView &gv, const GAParams &ga, const int *cnt, const int64_t *off,
                       const int *neigh, const double4 *gx, int nowned, double *energy, hipStream_t s) {
  const size_t N = (size_t)gv.nx * gv.ny * gv.nz;
  hipLaunchKernelGGL(k_ga_deposit, dim3(nblk(N)), dim3(TPB), 0, s, gv, ga, cnt, off, neigh, gx, nowned, energy);
  HIPCHK(hipGetLastError());
}
// Bead density: rho(g) = sum over gathered atoms of the grid-atom kernel
// value e at g (the factor of 2 (phi - phi_t) in the mu term), for points
// within the pair cut (an input feature of the surrogate).  Each term is
// rounded to a 64-bit fixed-point integer (e * scale) and added with integer
// atomics, which are associative: the sum, and so rho, does not depend on
// the order of the adds.
__global__ void k_ga_density(int ng, GridView gv, GAParams ga, const double4 *gx, double scale,
                             unsigned long long *acc) {
  GRID_STRIDE(q, ng) {
    const double4 x = gx[q];
    if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z))) continue;
    const int t = (int)x.w;
    const double cs = ga.cutsq[t];
    const GABox b = ga_box(gv, x, cs);
    for (int i = b.i0; i <= b.i1; i++) {
      const double delz = (i * gv.dz + gv.lo[2]) - x.z;
      for (int j = b.j0; j <= b.j1; j++) {
        const double dely = (j * gv.dy + gv.lo[1]) - x.y;
        int rk0, rk1;
        if (!ga_row(gv, b, x.x, cs, dely * dely + delz * delz, rk0, rk1)) continue;
        for (int k = rk0; k <= rk1; k++) {
          const double delx = (k * gv.dx + gv.lo[0]) - x.x;
          const double rsq = delx * delx + dely * dely + delz * delz;
          if (!ga_in(ga, rsq, cs)) continue;
          const double r = rsq / cs;
          double e;
          if (ga.style == 0) e = (exp(-r / 2.0) - ga.em05) * ga.e05 * ga.eps[t];
          else e = exp(-1. / (1 - r) + 1) * ga.eps[t];
          atomicAdd(&acc[((size_t)i * gv.ny + j) * gv.nx + k], (unsigned long long)llrint(e * scale));
        }
      }
    }
  }
}

__global__ void k_fixed_to_float(size_t n, const unsigned long long *acc, double inv_scale, float *rho) {
  GRID_STRIDE(i, n) rho[i] = (float)((double)(long long)acc[i] * inv_scale);
}

__global__ void k_add4(int n, const double4 *a, double4 *b) {
  GRID_STRIDE(i, n) {
    const double4 u = a[i];
    double4 v = b[i];
    v.x += u.x;
    v.y += u.y;
    v.z += u.z;
    b[i] = v;
  }
}
void launch_add4(int n, const double4 *a, double4 *b, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add4, dim3(nblk(n)), dim3(TPB), 0, s, n, a, b);
  HIPCHK(hipGetLastError());
}

void launch_ga_density(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, double scale,
                       unsigned long long *acc, size_t N, float *rho, hipStream_t s) {
  if (ng > 0) {
    hipLaunchKernelGGL(k_ga_density, dim3(nblk(ng)), dim3(TPB), 0, s, ng, gv, ga, gx, scale, acc);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_fixed_to_float, dim3(nblk(N)), dim3(TPB), 0, s, N, acc, 1.0 / scale, rho);
  HIPCHK(hipGetLastError());
}
void launch_ga_am_walk(int ng, const GridView &gv, const double4 *gxb, const int *order, double reachsq, int *cnt,
                       const int *off, int2 *ps, int *amq, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_ga_am_walk, dim3(nblk(ng)), dim3(TPB), 0, s, ng, gv, gxb, order, reachsq, cnt, off, ps, amq);
  HIPCHK(hipGetLastError());
}
void launch_ga_pm_count(int nent, const int2 *ps, int *pcnt, int *rank, hipStream_t s) {
  if (nent <= 0) return;
  hipLaunchKernelGGL(k_ga_pm_count, dim3(nblk(nent)), dim3(TPB), 0, s, nent, ps, pcnt, rank);
  HIPCHK(hipGetLastError());
}
void launch_ga_pm_place(int nent, int2 *ps, const int *rank, const int *amq, const int *poff, int *key, int *ent,
                        size_t N, hipStream_t s) {
  if (nent <= 0) return;
  hipLaunchKernelGGL(k_ga_pm_place, dim3(nblk(nent)), dim3(TPB), 0, s, nent, ps, rank, amq, poff, key, ent);
  HIPCHK(hipGetLastError());
  const size_t per_block = 4 * 64 * GM_U;
  hipLaunchKernelGGL(k_ga_pm_sort, dim3((unsigned)((N + per_block - 1) / per_block)), dim3(256), 0, s, N, poff, key,
                     ent, ps);
  HIPCHK(hipGetLastError());
}
static int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  if ((1 << l) != v) throw std::invalid_argument("grid sides must be powers of two");
  return l;
}
#define PHAFD_GA_DEF(RT)                                                                                        \
  void launch_ga_force(int ng, const GridViewT<RT> &gv, const GAParams &ga, const double4 *gx, const int *order, \
                       const int *aoff, const int2 *ps, RT *mterm, double4 *gF, double *energy, hipStream_t s) { \
    if (ng <= 0) return;                                                                                      \
    const int lx = ilog2(gv.nx), lxy = lx + ilog2(gv.ny);                                                     \
    hipLaunchKernelGGL(k_ga_force<RT>, dim3(nblk(ng)), dim3(TPB), 0, s, ng, gv, ga, gx, order, aoff, ps,      \
                       mterm, gF, energy, lx, lxy);                                                           \
    HIPCHK(hipGetLastError());                                                                                \
  }                                                                                                           \
  void launch_ga_list_mu(size_t N, RT *mu, const int *poff, const RT *mterm, bool first, hipStream_t s) {     \
    const size_t per_block = 4 * 64 * GM_U;                                                                   \
    hipLaunchKernelGGL(k_ga_list_mu<RT>, dim3((unsigned)((N + per_block - 1) / per_block)), dim3(256), 0, s, N, \
                       mu, poff, mterm, first ? 1 : 0);                                                       \
    HIPCHK(hipGetLastError());                                                                                \
  }
PHAFD_GA_DEF(double)
PHAFD_GA_DEF(float)
#undef PHAFD_GA_DEF
void launch_lj_force(int ng, const LJParams &lj, const double4 *gx, const int *rowof, const int *hoff,
                     const int *hneigh, const int *toff, const int *tpair, const int *ti, int nowned,
                     double4 *gF, double *energy, hipStream_t s, bool zero_init, StepGate gate) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_pair_force<LJLaw>, dim3(nblk(ng)), dim3(TPB), 0, s, ng, LJLaw{lj}, gx, rowof, hoff, hneigh,
                     toff, tpair, ti, nowned, gF, energy, zero_init ? 1 : 0, gate);
  HIPCHK(hipGetLastError());
}
void launch_harmonic_force(int ng, const HarmParams &h, const double4 *gx, const int *rowof, const int *hoff,
                           const int *hneigh, const int *toff, const int *tpair, const int *ti, int nowned,
                           double4 *gF, hipStream_t s, bool zero_init, StepGate gate) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_pair_force<HarmLaw>, dim3(nblk(ng)), dim3(TPB), 0, s, ng, HarmLaw{h}, gx, rowof, hoff,
                     hneigh, toff, tpair, ti, nowned, gF, (double *)nullptr, zero_init ? 1 : 0, gate);
  HIPCHK(hipGetLastError());
}
void launch_drag(int n, const int *members, double4 *x, const double4 *F, double drag, double dt,
                 hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_drag, dim3(nblk(n)), dim3(TPB), 0, s, n, members, x, F, drag, dt);
  HIPCHK(hipGetLastError());
}
void launch_chain_step(int nchains, int maxlen, const int *chs, const int *che, const double4 *F, double4 *ux,
                       double4 *x, const int *image, const ChainParams &cp, double dt, const Box &b, hipStream_t s,
                       int halves, StepGate gate) {
  if (nchains <= 0 || maxlen <= 0) return;
  if (halves != 1 && halves != 2) throw std::invalid_argument("chain step: halves must be 1 or 2");
  const int threads = maxlen <= 64 ? 64 : (maxlen <= 128 ? 128 : 256);
  hipLaunchKernelGGL(k_chain_step, dim3(nchains), dim3(threads), sizeof(double4) * maxlen * halves, s, chs, che, F,
                     ux, x, image, cp, dt, b, halves, maxlen, gate);
  HIPCHK(hipGetLastError());
}
void launch_chain_energy(int n, const int *members, const int *cstart, const int *cend, const double4 *ux,
                         const ChainParams &cp, double *energy, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_chain_energy, dim3(nblk(n)), dim3(TPB), 0, s, n, members, cstart, cend, ux, cp, energy);
  HIPCHK(hipGetLastError());
}
size_t br_lds_bytes(int maxlen) { return sizeof(double) * BR_PER_BEAD * (size_t)maxlen; }
void br_lds_allow(const void *k, size_t bytes) {
  if (bytes > 65536) HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}
void launch_br_first(int nchains, int maxlen, const int *chs, const int *che, const int *seeds, const double4 *F,
                     double4 *ux, double4 *x, const int *image, double4 *R0, double4 *FB, double *T,
                     const BRParamsDev &p, const Box &b, hipStream_t s) {
  if (nchains <= 0 || maxlen <= 0) return;
  if (maxlen > BR_MAXLEN) throw std::invalid_argument("bead-rod chains longer than 700 beads are not supported");
  const size_t lds = br_lds_bytes(maxlen);
  br_lds_allow((const void *)k_br_first, lds);
  hipLaunchKernelGGL(k_br_first, dim3(nchains), dim3(64), lds, s, chs, che, seeds, F, ux, x, image, R0, FB, T, p, b,
                     maxlen);
  HIPCHK(hipGetLastError());
}
void launch_br_second(int nchains, int maxlen, const int *chs, const int *che, const double4 *F, double4 *ux,
                      double4 *x, const int *image, const double4 *R0, const double4 *FB, double *T, int *fail,
                      const BRParamsDev &p, const Box &b, hipStream_t s) {
  if (nchains <= 0 || maxlen <= 0) return;
  if (maxlen > BR_MAXLEN) throw std::invalid_argument("bead-rod chains longer than 700 beads are not supported");
  const size_t lds = br_lds_bytes(maxlen);
  br_lds_allow((const void *)k_br_second, lds);
  hipLaunchKernelGGL(k_br_second, dim3(nchains), dim3(64), lds, s, chs, che, F, ux, x, image, R0, FB, T, fail, p,
                     b, maxlen);
  HIPCHK(hipGetLastError());
}
// the batched pair phase: grids of at most 8 blocks per CU take the items
// round-robin
unsigned multi_grid(int nitems, int cap) { return (unsigned)std::max(1, std::min(nitems, cap)); }
void check_shape(const MultiShape &sh) {
  if ((sh.threads != 64 && sh.threads != 128 && sh.threads != 256) || sh.cap < 1)
    throw std::invalid_argument("batched pair phase: items of 64, 128 or 256 atoms, a positive grid cap");
}
void launch_ga_force_multi(int nm, const GAMemberDev *d, const int *boff, int nitems, int nx, int ny,
                           const MultiShape &sh, hipStream_t s) {
  if (nm <= 0 || nitems <= 0) return;
  check_shape(sh);
  const int lx = ilog2(nx), lxy = lx + ilog2(ny);
  hipLaunchKernelGGL(k_ga_force_multi, dim3(multi_grid(nitems, sh.cap)), dim3(sh.threads), 0, s, d, boff, nm, lx,
                     lxy);
  HIPCHK(hipGetLastError());
}
void launch_ga_list_mu_multi(int nm, size_t N, const GAMemberDev *d, hipStream_t s) {
  if (nm <= 0) return;
  const size_t per_block = 4 * 64 * GM_U;
  hipLaunchKernelGGL(k_ga_list_mu_multi, dim3((unsigned)((N + per_block - 1) / per_block), nm), dim3(256), 0, s, d,
                     N);
  HIPCHK(hipGetLastError());
}
void launch_lj_multi(int nm, const LJMemberDev *d, const int *boff, int nitems, const MultiShape &sh,
                     hipStream_t s) {
  if (nm <= 0 || nitems <= 0) return;
  check_shape(sh);
  hipLaunchKernelGGL(k_lj_multi, dim3(multi_grid(nitems, sh.cap)), dim3(sh.threads), 0, s, d, boff, nm);
  HIPCHK(hipGetLastError());
}
void launch_fold_multi(int nm, const FoldDev *d, const int *boff, int nitems, bool zero, const MultiShape &sh,
                       hipStream_t s) {
  if (nm <= 0 || nitems <= 0) return;
  check_shape(sh);
  hipLaunchKernelGGL(k_fold_multi, dim3(multi_grid(nitems, sh.cap)), dim3(sh.threads), 0, s, d, boff, nm,
                     zero ? 1 : 0);
  HIPCHK(hipGetLastError());
}
void launch_fill_multi(int nm, const FoldDev *d, const int *boff, int nitems, const MultiShape &sh, hipStream_t s) {
  if (nm <= 0 || nitems <= 0) return;
  check_shape(sh);
  hipLaunchKernelGGL(k_fill_multi, dim3(multi_grid(nitems, sh.cap)), dim3(sh.threads), 0, s, d, boff, nm);
  HIPCHK(hipGetLastError());
}
void launch_chain_multi(int nm, const ChainMemberDev *d, const int *coff, int nchains, int maxlen, double dt,
                        hipStream_t s) {
  if (nm <= 0 || nchains <= 0 || maxlen <= 0) return;
  if (maxlen > 1024) throw std::invalid_argument("fused chain steps take chains of at most 1024 beads");
  const int threads = maxlen <= 64 ? 64 : (maxlen <= 128 ? 128 : 256);
  hipLaunchKernelGGL(k_chain_multi, dim3(nchains), dim3(threads), sizeof(double4) * maxlen * 2, s, d, coff, nm, dt,
                     maxlen);
  HIPCHK(hipGetLastError());
}
void launch_fill_i32(int n, int *p, int v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n)), dim3(TPB), 0, s, n, p, v);
  HIPCHK(hipGetLastError());
}
void launch_fill_d4(int n, double4 *p, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_d4, dim3(nblk(n)), dim3(TPB), 0, s, n, p);
  HIPCHK(hipGetLastError());
}

}  // namespace phafd
