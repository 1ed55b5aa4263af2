// ic.cpp -- host-side input generators of the drop-in (plain C++, no GPU):
//   * grid_populate constant/flat_interface/sphere (src/grid.cpp:115-434)
//   * the per-mode thermal noise of grid/conjugate/volfrac in the reference's
//     std::mt19937 draw order (src/conjugate.cpp:163-327,
//     src/conjugate_volfrac.cpp:72-133), for PHAFD_NOISE_FIELD
//   * a synthetic random-walk chain generator for benchmark workloads
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_map>
#include <vector>

#include "phafd_host.h"

namespace {

// utility::make_unique_seed for rank 0, src/utility.cpp:213-235
int unique_seed(int base) {
  std::mt19937 gen;
  std::uniform_int_distribution<int> d;
  gen.seed(base);
  return d(gen);
}

}  // namespace

struct phafd_noise_gen {
  std::mt19937 gen;
  std::uniform_real_distribution<double> dist{-0.5, 0.5};
};

extern "C" {

int phafd_make_unique_seed(int seed) { return unique_seed(seed); }

// Grid::noisy_constant, src/grid.cpp:289-307 (seed through make_unique_seed,
// src/grid.cpp:124-128); out is [Nz][Ny][Nx]
int phafd_ic_noisy_constant(int64_t n, double average, double variance, int seed, double *out) {
  if (n < 0 || (!out && n)) return -1;
  std::uniform_real_distribution<double> real_dist(-0.5, 0.5);
  std::mt19937 gen;
  gen.seed(unique_seed(seed));
  for (int64_t i = 0; i < n; i++) out[i] = average + variance * real_dist(gen);
  return 0;
}

// Grid::flat_interface, src/grid.cpp:310-337
int phafd_ic_flat_interface(int nx, int ny, int nz, double boxlo_x, double period_x, double average,
                            double hi, double lo, double variance, int seed, double *out) {
  std::uniform_real_distribution<double> real_dist(-0.5, 0.5);
  std::mt19937 gen;
  const double Lbox = (average - lo) / (hi - lo) * period_x;
  const double Ox = boxlo_x;
  const double dx = period_x / nx;
  gen.seed(unique_seed(seed));
  size_t w = 0;
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nx; k++) {
        double x = k * dx + Ox;
        if (x - Ox < Lbox) out[w++] = hi + variance * real_dist(gen);
        else out[w++] = lo + variance * real_dist(gen);
      }
  return 0;
}

// Grid::sphere, src/grid.cpp:339-438 (width given as in the script, i.e.
// squared; the reference takes its square root at src/grid.cpp:158)
int phafd_ic_sphere(int nx, int ny, int nz, const double *boxlo, const double *period, double radius,
                    double width_sq, double hi, double av, double *out) {
  const double width = sqrt(width_sq);
  for (int d = 0; d < 3; d++)
    if (radius >= period[d] / 2.) return -1;
  const double dx = period[0] / nx, dy = period[1] / ny, dz = period[2] / nz;
  auto shape = [&](double r) { return tanh((-r + radius) / width); };
  double integral = 0;
  for (int i = 0; i < nz; i++) {
    double z = i * dz + boxlo[2];
    for (int j = 0; j < ny; j++) {
      double y = j * dy + boxlo[1];
      for (int k = 0; k < nx; k++) {
        double x = k * dx + boxlo[0];
        integral += shape(sqrt(x * x + y * y + z * z));
      }
    }
  }
  integral *= dx * dy * dz;
  const double volume = period[0] * period[1] * period[2];
  const double ifrac = integral / volume;
  const double lo = (2 * av - hi * (1 + ifrac)) / (1 - ifrac);
  if (lo < 0 || lo > hi) return -2;
  size_t w = 0;
  for (int i = 0; i < nz; i++) {
    double z = i * dz + boxlo[2];
    for (int j = 0; j < ny; j++) {
      double y = j * dy + boxlo[1];
      for (int k = 0; k < nx; k++) {
        double x = k * dx + boxlo[0];
        out[w++] = 0.5 * (hi - lo) * shape(sqrt(x * x + y * y + z * z)) + 0.5 * (hi + lo);
      }
    }
  }
  return 0;
}

// --------------------------------------------------------------- noise
phafd_noise_gen *phafd_noise_create(int seed) {
  auto *g = new phafd_noise_gen();
  g->gen.seed(unique_seed(seed));   // src/fixgrid_conjugate.cpp:29-31, conjugate_volfrac.cpp:65
  return g;
}
void phafd_noise_destroy(phafd_noise_gen *g) { delete g; }

// One step of noise values xi(k) in the canonical spectrum layout
// [Nz][Ny][Nx/2+1] (re, im), drawn in the traversal order of
// Conjugate::update for one rank: bulk (ky, kz, kx in [1, Nx/2)), then the
// kx = 0 and kx = Nx/2 planes (single then line_single), then the seven
// real_update modes.  Mirror modes and the origin get no draw (zero here).
int phafd_noise_next(phafd_noise_gen *g, int nx, int ny, int nz, const double *period,
                     double mobility, double temp, double dt, double *xi) {
  if (!g || !xi) return -1;
  const int hx = nx / 2 + 1, A = ny, B = nz;
  const double PI_REF = 3.141592;
  const double dqx = 2 * PI_REF / period[0];
  const double dzq = 2 * PI_REF / period[1];   // Domain::dqz over the ky index
  const double dyq = 2 * PI_REF / period[2];   // Domain::dqy over the kz index
  const double invLcubed = 1.0 / (period[0] * period[1] * period[2]);
  const double cpre = sqrt(12 * temp * mobility * invLcubed);
  const double rpre = sqrt(24 * temp * mobility * invLcubed);
  const double sqrtdt = sqrt(dt);
  std::memset(xi, 0, sizeof(double) * 2 * (size_t)hx * ny * nz);
  auto q2 = [&](int i, int j, int k) {
    double qz = (i > A / 2) ? -dzq * (A - i) : dzq * i;
    double qy = (j > B / 2) ? -dyq * (B - j) : dyq * j;
    double qx = dqx * k;
    return qx * qx + qy * qy + qz * qz;
  };
  auto at = [&](int i, int j, int k) { return 2 * (((size_t)j * A + i) * hx + k); };
  auto cplx = [&](int i, int j, int k) {
    double q = q2(i, j, k);
    double u1 = g->dist(g->gen), u2 = g->dist(g->gen);
    size_t o = at(i, j, k);
    xi[o] = cpre * sqrt(q) * sqrtdt * u1;
    xi[o + 1] = cpre * sqrt(q) * sqrtdt * u2;
  };
  auto real = [&](int i, int j, int k) {
    double q = q2(i, j, k);
    size_t o = at(i, j, k);
    xi[o] = rpre * sqrt(q) * sqrtdt * g->dist(g->gen);
    xi[o + 1] = 0.0;
  };
  for (int i = 0; i < A; i++)
    for (int j = 0; j < B; j++)
      for (int k = 1; k < hx - 1; k++) cplx(i, j, k);
  for (int k = 0; k < hx; k += hx - 1) {
    for (int j = 1; j < B / 2; j++) cplx(0, j, k);
    for (int i = 1; i < A / 2; i++)
      for (int j = 1; j < B / 2; j++) cplx(i, j, k);
    for (int i = A / 2 + 1; i < A; i++)
      for (int j = 1; j < B / 2; j++) cplx(i, j, k);
    for (int j = 1; j < B / 2; j++) cplx(A / 2, j, k);
    for (int i = 1; i < A / 2; i++) cplx(i, 0, k);
    for (int i = A / 2 + 1; i < A; i++) cplx(i, B / 2, k);
  }
  const int H = hx - 1;
  real(0, B / 2, 0);
  real(0, B / 2, H);
  real(0, 0, H);
  real(A / 2, 0, 0);
  real(A / 2, B / 2, 0);
  real(A / 2, B / 2, H);
  real(A / 2, 0, H);
  return 0;
}

// ------------------------------------------------------ chain generator
// Synthetic polymer workload: nchains random walks of nbeads with fixed bond
// length, each bead placed at a uniformly random direction from the previous
// one and rejected when closer than minsep to any already placed bead; a
// chain that cannot place a bead within 100 tries is regrown from a new
// random start, so no two beads end up closer than minsep.  Positions are unwrapped walks started at a
// uniform point of the box; wrap them with the periodic box as the reference
// does in Domain::pbc.  Deterministic for a given seed (std::mt19937_64).
int phafd_ic_random_chains(int nchains, int nbeads, double bond, double minsep,
                           const double *boxlo, const double *period, uint64_t seed,
                           double *xyz, int *mol) {
  if (nchains < 0 || nbeads < 1) return -1;
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  const double cell = minsep > 0 ? minsep : 1.0;
  int nc[3];
  for (int d = 0; d < 3; d++) {
    nc[d] = (int)(period[d] / cell);
    if (nc[d] < 1) nc[d] = 1;
  }
  std::unordered_map<int64_t, std::vector<int>> grid;
  auto wrap = [&](double v, int d) {
    double t = fmod(v - boxlo[d], period[d]);
    if (t < 0) t += period[d];
    return t;
  };
  auto key = [&](const double *p) {
    int64_t c[3];
    for (int d = 0; d < 3; d++) {
      c[d] = (int64_t)(wrap(p[d], d) / period[d] * nc[d]);
      if (c[d] >= nc[d]) c[d] = nc[d] - 1;
    }
    return (c[2] * nc[1] + c[1]) * nc[0] + c[0];
  };
  auto too_close = [&](const double *p) {
    if (minsep <= 0) return false;
    int64_t c0[3];
    for (int d = 0; d < 3; d++) {
      c0[d] = (int64_t)(wrap(p[d], d) / period[d] * nc[d]);
      if (c0[d] >= nc[d]) c0[d] = nc[d] - 1;
    }
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int64_t c[3] = {(c0[0] + dx + nc[0]) % nc[0], (c0[1] + dy + nc[1]) % nc[1], (c0[2] + dz + nc[2]) % nc[2]};
          auto it = grid.find((c[2] * nc[1] + c[1]) * nc[0] + c[0]);
          if (it == grid.end()) continue;
          for (int j : it->second) {
            double r2 = 0;
            for (int d = 0; d < 3; d++) {
              double dd = wrap(p[d], d) - wrap(xyz[3 * (size_t)j + d], d);
              dd -= period[d] * std::round(dd / period[d]);
              r2 += dd * dd;
            }
            if (r2 < minsep * minsep) return true;
          }
        }
    return false;
  };
  int idx = 0;
  for (int c = 0; c < nchains; c++) {
    const int first = idx;
    int regrow = 0;
    for (int b = 0; b < nbeads; b++, idx++) {
      double p[3];
      bool placed = false;
      for (int tries = 0; tries < 100; tries++) {
        if (b == 0) {
          for (int d = 0; d < 3; d++) p[d] = boxlo[d] + u01(gen) * period[d];
        } else {
          double z = 2 * u01(gen) - 1, ph = 2 * M_PI * u01(gen), s = sqrt(1 - z * z);
          const double *q = &xyz[3 * (size_t)(idx - 1)];
          p[0] = q[0] + bond * s * cos(ph);
          p[1] = q[1] + bond * s * sin(ph);
          p[2] = q[2] + bond * z;
        }
        if (!too_close(p)) { placed = true; break; }
      }
      if (!placed) {
        // remove this chain's beads from the cell map and start it again
        for (int j = first; j < idx; j++) {
          auto &v = grid[key(&xyz[3 * (size_t)j])];
          v.erase(std::remove(v.begin(), v.end(), j), v.end());
        }
        if (++regrow > 1000) return -2;
        idx = first - 1;
        b = -1;
        continue;
      }
      for (int d = 0; d < 3; d++) xyz[3 * (size_t)idx + d] = p[d];
      if (mol) mol[idx] = c;
      grid[key(p)].push_back(idx);
    }
  }
  return 0;
}

}  // extern "C"
