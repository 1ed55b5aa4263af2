// atoms.hip -- atom-side kernels of the per-timestep hot path (FP64):
// periodic wrap, ghost images, binning, neighbour lists, pair forces,
// force fold-back and the overdamped integrators.
//
// Every kernel that decides list membership or position arithmetic uses the
// reference's exact expressions (built with -ffp-contract=off), so lists and
// bins are bit-identical to the reference's on the same positions.
#include "atoms.hpp"

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
template <bool FILL>
__global__ void k_grid_list(GridView gv, BinGeom b, const int *stencil, int nst,
                            const int *binstart, const int *binatoms, const double4 *gx,
                            double cutneighsq, int nowned, int *cnt, const int64_t *off,
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
    int64_t w = FILL ? off[p] : 0;
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
          if (FILL) neigh[w++] = nowned + a;
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
        if (FILL) neigh[w++] = nowned + j;
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
          if (FILL) neigh[w++] = nowned + j;
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
__device__ __forceinline__ GABox ga_box(const GridView &gv, const double4 &x, double cs) {
  const double w = sqrt(cs);
  const double mg = 1e-9;
  GABox b;
  b.i0 = (int)ceil((x.z - w - gv.lo[2]) / gv.dz - mg); b.i1 = (int)floor((x.z + w - gv.lo[2]) / gv.dz + mg);
  b.j0 = (int)ceil((x.y - w - gv.lo[1]) / gv.dy - mg); b.j1 = (int)floor((x.y + w - gv.lo[1]) / gv.dy + mg);
  b.k0 = (int)ceil((x.x - w - gv.lo[0]) / gv.dx - mg); b.k1 = (int)floor((x.x + w - gv.lo[0]) / gv.dx + mg);
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
__device__ __forceinline__ bool ga_row(const GridView &gv, const GABox &b, double xx, double cs, double dyz,
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
// The reference adds to F(aj) once per listed grid point, in ascending grid
// index, and to mu(g) once per listed atom, in list order.  Every grid point
// with rsq <= cutsq(type) is in the atom's list (the list cut is cutsq + skin
// and no atom has moved more than skin/2 since the build), so walking the
// in-box grid points around the atom in (z, y, x) order finds the same pairs
// in the force's order.  A grid point's list walks the stencil bins in
// (dz, dy, dx) lexicographic order (src/nstencil_full_bin_3d.cpp:30-36) and
// each bin in its list order (src/nbin_standard.cpp:181-229); bins are
// numbered z-major, so that order is ascending position q in the bin-sorted
// atom array: the key mu terms are summed by.
//
// Pipeline: per atom (bin order) count then enumerate its in-range points;
// per pair the expensive part (exp, field gathers, the mu term, the force
// term, the pair's rank among its grid point's terms); per atom the force
// fold in pair order; per pair the placement by grid point; per grid point
// the sum in key order.  `order` is the bin-sorted atom array.
__global__ void k_ga_count(int ng, GridView gv, GAParams ga, const double4 *gx, const int *order,
                           const double *cutsq, int *acnt) {
  GRID_STRIDE(q, ng) {
    const double4 x = gx[order[q]];
    int n = 0;
    if (isfinite(x.x) && isfinite(x.y) && isfinite(x.z)) {
      const double cs = cutsq[(int)x.w];
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
            n += ga_in(ga, rsq, cs) ? 1 : 0;
          }
        }
      }
    }
    acnt[q] = n;
  }
}

// in-range grid point and atom position of every pair, atom by atom
__global__ void k_ga_enum(int ng, GridView gv, GAParams ga, const double4 *gx, const int *order, const int *aoff,
                          int64_t cap, int *pg, int *pq, int *overflow) {
  GRID_STRIDE(q, ng) {
    const double4 x = gx[order[q]];
    if (!isfinite(x.x) || !isfinite(x.y) || !isfinite(x.z)) continue;
    const double cs = ga.cutsq[(int)x.w];
    const GABox b = ga_box(gv, x, cs);
    int w = aoff[q];
    if (aoff[q + 1] > cap) {
      atomicOr(overflow, 1);
      continue;
    }
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
          pg[w] = (int)(((size_t)i * gv.ny + j) * gv.nx + k);
          pq[w] = (int)q;
          w++;
        }
      }
    }
  }
}

__device__ __forceinline__ float fexp(float x) { return expf(x); }
__device__ __forceinline__ double fexp(double x) { return exp(x); }
template <class R> struct Vec4;
template <> struct Vec4<double> { using T = double4; };
template <> struct Vec4<float> { using T = float4; };

// one pair: the loop body of the reference's compute, with the same
// arithmetic (rsq rebuilt from the grid indices exactly as the walk made it).
// FP32 form: rsq in FP64 from the FP64 positions, everything after it in
// FP32; the energy term in FP64.
template <class R>
__global__ void k_ga_pair(const int *npairs, GridViewT<R> gv, GAParams ga, const double4 *gx, const int *order,
                          const int *pg, const int *pq, typename Vec4<R>::T *ft, R *pv, int *ppos, int *gcnt,
                          bool emit, double *energy) {
  const int np = *npairs;
  double esum = 0.0;
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < np; w += gridDim.x * blockDim.x) {
    const int p = pg[w];
    const double4 x = gx[order[pq[w]]];
    const int t = (int)x.w;
    const double cs = ga.cutsq[t];
    const int k = p % gv.nx, j = (p / gv.nx) % gv.ny, i = p / (gv.nx * gv.ny);
    const double ztmp = i * gv.dz + gv.lo[2];
    const double delz = ztmp - x.z;
    const double ytmp = j * gv.dy + gv.lo[1];
    const double dely = ytmp - x.y;
    const double xtmp = k * gv.dx + gv.lo[0];
    const double delx = xtmp - x.x;
    double rsq = delx * delx + dely * dely + delz * delz;
    rsq /= cs;
    const R r = (R)rsq;
    R expfac;
    if (ga.style == 0) expfac = (fexp(-r / R(2)) - (R)ga.em05) * (R)ga.e05 * (R)ga.eps[t];
    else expfac = fexp(R(-1) / (1 - r) + 1) * (R)ga.eps[t];
    if (ga.style == GA_LJISH_LINEAR) {
      // PairGridAtomLJishLinear::compute (src/pair_gridatom_lj_ish_linear.cpp:
      // 93-160): e = -1 * bump * eps (the sign flip is exact); mu += e, the
      // force weight is e dv, and no energy
      expfac = -expfac;
      if (emit) {
        pv[w] = expfac;
        ppos[w] = atomicAdd(&gcnt[p], 1);
      }
      expfac *= (R)gv.dv;
    } else {
      const R delphi = (gv.phi[p] - (R)ga.phit[t]);
      if (emit) {   // mu term, summed later in list order
        pv[w] = 2 * delphi * expfac;
        ppos[w] = atomicAdd(&gcnt[p], 1);
      }
      if (energy) esum += (double)delphi * (double)delphi * (double)expfac * gv.dv;
      expfac *= 2 * delphi * (R)gv.dv;
    }
    typename Vec4<R>::T f;
    f.x = gv.gx[p] * expfac;
    f.y = gv.gy[p] * expfac;
    f.z = gv.gz[p] * expfac;
    f.w = R(0);
    ft[w] = f;
  }
  if (energy && esum != 0.0) atomicAdd(energy, esum);
}

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

// F(a) -= its pair terms in ascending grid index, as the reference adds them
// (FP64); FP32 form: the terms summed with Kahan compensation, then taken
// from F in FP64
__global__ void k_ga_fold(int ng, const int *order, const int *aoff, const double4 *ft, double4 *gF) {
  GRID_STRIDE(q, ng) {
    const int a = order[q];
    const int s = aoff[q], e = aoff[q + 1];
    if (s == e) continue;
    double4 F = gF[a];
    for (int w = s; w < e; w++) {
      const double4 f = ft[w];
      F.x -= f.x;
      F.y -= f.y;
      F.z -= f.z;
    }
    gF[a] = F;
  }
}
__global__ void k_ga_fold(int ng, const int *order, const int *aoff, const float4 *ft, double4 *gF) {
  GRID_STRIDE(q, ng) {
    const int a = order[q];
    const int s = aoff[q], e = aoff[q + 1];
    if (s == e) continue;
    Kahan kx, ky, kz;
    for (int w = s; w < e; w++) {
      const float4 f = ft[w];
      kx.add(f.x);
      ky.add(f.y);
      kz.add(f.z);
    }
    double4 F = gF[a];
    F.x -= (double)kx.s;
    F.y -= (double)ky.s;
    F.z -= (double)kz.s;
    gF[a] = F;
  }
}

// every term's key to its grid point's segment (in arbitrary order: the rank
// came from an atomic)
__global__ void k_ga_place(const int *npairs, const int *pg, const int *ppos, const int *pq, const int *goff,
                           int *skey) {
  const int np = *npairs;
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < np; w += gridDim.x * blockDim.x)
    skey[goff[pg[w]] + ppos[w]] = pq[w];
}

// every term's value to its place in key (list) order inside its segment:
// its rank is the number of smaller keys there (keys of a segment differ)
template <class R>
__global__ void k_ga_order(const int *npairs, const int *pg, const int *pq, const R *pv, const int *goff,
                           const int *skey, R *sv) {
  const int np = *npairs;
  for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < np; w += gridDim.x * blockDim.x) {
    const int p = pg[w], key = pq[w];
    const int s = goff[p], e = goff[p + 1];
    int r = 0;
    for (int j = s; j < e; j++) r += skey[j] < key ? 1 : 0;
    sv[s + r] = pv[w];
  }
}

// mu(g) (+)= its terms, already in key order, four grid points per item;
// FP32 form: Kahan-compensated
template <class R>
struct MuAcc {   // FP64: plain running sum in order
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
// Four grid points per thread.  Their terms are contiguous in sv ([o.x,
// o4)); they are loaded GS at a time (independent loads in flight) and
// added to their point's sum in index order, the same order as one term at
// a time.
template <class R>
__global__ void k_ga_sum(size_t N, const int *__restrict__ goff, const R *__restrict__ sv, R *__restrict__ mu,
                         bool first) {
  constexpr int GS = 8;
  const size_t N4 = N / 4;
  for (size_t u = blockIdx.x * (size_t)blockDim.x + threadIdx.x; u < N4; u += (size_t)gridDim.x * blockDim.x) {
    const size_t p0 = 4 * u;
    const int4 o = *reinterpret_cast<const int4 *>(goff + p0);
    const int o4 = goff[p0 + 4];
    if (!first && o.x == o4) continue;
    MuAcc<R> a0, a1, a2, a3;
    a0.m = first ? R(0) : mu[p0];
    a1.m = first ? R(0) : mu[p0 + 1];
    a2.m = first ? R(0) : mu[p0 + 2];
    a3.m = first ? R(0) : mu[p0 + 3];
    for (int j = o.x; j < o4; j += GS) {
      R t[GS];
#pragma unroll
      for (int k = 0; k < GS; k++) t[k] = j + k < o4 ? sv[j + k] : R(0);
#pragma unroll
      for (int k = 0; k < GS; k++) {
        const int jj = j + k;
        if (jj >= o4) break;
        if (jj < o.y) a0.add(t[k]);
        else if (jj < o.z) a1.add(t[k]);
        else if (jj < o.w) a2.add(t[k]);
        else a3.add(t[k]);
      }
    }
    mu[p0] = a0.get();
    mu[p0 + 1] = a1.get();
    mu[p0 + 2] = a2.get();
    mu[p0 + 3] = a3.get();
  }
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
__global__ void k_pair_force(int ng, Law law, const double4 *gx, const int *rowof,
                             const int *hoff, const int *hneigh, const int *toff, const int *tpair,
                             const int *ti, int nowned, double4 *gF, double *energy, int zero_init,
                             StepGate gate) {
  if (gate_off(gate)) return;
  double esum = 0.0;
  GRID_STRIDE(a, ng) {
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
          if (energy) esum += e;
        }
      }
    }
    for (; q < q1; q++) {
      if (law(gx[ti[q]], xa, fx, fy, fz, e)) { F.x -= fx; F.y -= fy; F.z -= fz; }
    }
    gF[a] = F;
  }
  if (energy) atomicAdd(energy, esum);
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
__global__ void k_chain_step(const int *chs, const int *che, const double4 *F, double4 *ux, double4 *x,
                             const int *image, ChainParams cp, double dt, Box bx, int halves, int stride,
                             StepGate gate) {
  if (gate_off(gate)) return;   // uniform over the block: *gate.stop is fixed during this launch
  extern __shared__ double4 us[];
  const int s = chs[blockIdx.x], nb = che[blockIdx.x] - s;
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
void launch_grid_list(bool fill, const GridView &gv, const BinGeom &b, const int *stencil, int nst,
                      const int *binstart, const int *binatoms, const double4 *gx, double cutneighsq,
                      int nowned, int *cnt, const int64_t *off, int *neigh, int *overflow, hipStream_t s) {
  const size_t N = (size_t)gv.nx * gv.ny * gv.nz;
  if (fill)
    hipLaunchKernelGGL((k_grid_list<true>), dim3(nblk(N)), dim3(TPB), 0, s, gv, b, stencil, nst, binstart,
                       binatoms, gx, cutneighsq, nowned, cnt, off, neigh, overflow);
  else
    hipLaunchKernelGGL((k_grid_list<false>), dim3(nblk(N)), dim3(TPB), 0, s, gv, b, stencil, nst, binstart,
                       binatoms, gx, cutneighsq, nowned, cnt, off, neigh, overflow);
  HIPCHK(hipGetLastError());
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
void launch_ga_count(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, const int *order,
                     const double *cutsq, int *acnt, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_ga_count, dim3(nblk(ng)), dim3(TPB), 0, s, ng, gv, ga, gx, order, cutsq, acnt);
  HIPCHK(hipGetLastError());
}
void launch_ga_enum(int ng, const GridView &gv, const GAParams &ga, const double4 *gx, const int *order,
                    const int *aoff, int64_t cap, int *pg, int *pq, int *overflow, hipStream_t s) {
  if (ng <= 0) return;
  hipLaunchKernelGGL(k_ga_enum, dim3(nblk(ng)), dim3(TPB), 0, s, ng, gv, ga, gx, order, aoff, cap, pg, pq,
                     overflow);
  HIPCHK(hipGetLastError());
}
#define PHAFD_GA_DEF(RT, V4)                                                                                    \
  void launch_ga_pair(int64_t cap, const int *npairs, const GridViewT<RT> &gv, const GAParams &ga,             \
                      const double4 *gx, const int *order, const int *pg, const int *pq, V4 *ft, RT *pv,      \
                      int *ppos, int *gcnt, bool emit, double *energy, hipStream_t s) {                       \
    hipLaunchKernelGGL(k_ga_pair<RT>, dim3(nblk((size_t)cap)), dim3(TPB), 0, s, npairs, gv, ga, gx, order, pg, pq, \
                       ft, pv, ppos, gcnt, emit, energy);                                                    \
    HIPCHK(hipGetLastError());                                                                                \
  }                                                                                                           \
  void launch_ga_fold(int ng, const int *order, const int *aoff, const V4 *ft, double4 *gF, hipStream_t s) {  \
    if (ng <= 0) return;                                                                                      \
    const V4 *f = ft;                                                                                         \
    hipLaunchKernelGGL(k_ga_fold, dim3(nblk(ng)), dim3(TPB), 0, s, ng, order, aoff, f, gF);                   \
    HIPCHK(hipGetLastError());                                                                                \
  }                                                                                                           \
  void launch_ga_place(int64_t cap, const int *npairs, const int *pg, const int *ppos, const int *pq,         \
                       const RT *pv, const int *goff, int *skey, RT *sv, hipStream_t s) {                     \
    hipLaunchKernelGGL(k_ga_place, dim3(nblk((size_t)cap)), dim3(TPB), 0, s, npairs, pg, ppos, pq, goff, skey); \
    HIPCHK(hipGetLastError());                                                                                \
    hipLaunchKernelGGL(k_ga_order<RT>, dim3(nblk((size_t)cap)), dim3(TPB), 0, s, npairs, pg, pq, pv, goff, skey, sv); \
    HIPCHK(hipGetLastError());                                                                                \
  }                                                                                                           \
  void launch_ga_sum(size_t N, const int *goff, const int *skey, const RT *sv, RT *mu, bool first,           \
                     hipStream_t s) {                                                                         \
    (void)skey;                                                                                               \
    if (N % 4) throw std::invalid_argument("grid size must be a multiple of 4");                              \
    hipLaunchKernelGGL(k_ga_sum<RT>, dim3(nblk(N / 4)), dim3(TPB), 0, s, N, goff, sv, mu, first);            \
    HIPCHK(hipGetLastError());                                                                                \
  }
PHAFD_GA_DEF(double, double4)
PHAFD_GA_DEF(float, float4)
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
