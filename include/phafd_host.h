/*
 * phafd_host.h -- host-side (CPU) helpers of the drop-in that feed the C ABI
 * of phafd_hip.h: initial conditions, reference-order thermal noise, and a
 * synthetic chain workload generator.  No GPU work happens here.
 */
#ifndef PHAFD_HOST_H
#define PHAFD_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* utility::make_unique_seed for rank 0 (src/utility.cpp:213-235) */
int phafd_make_unique_seed(int seed);
/* grid_populate constant concentration <avg> <var> <seed> -> Grid::noisy_constant
 * (src/grid.cpp:120-130, 289-307) */
int phafd_ic_noisy_constant(int64_t n, double average, double variance, int seed, double *out);
/* grid_populate flat_interface concentration ... (src/grid.cpp:136-151, 310-337) */
int phafd_ic_flat_interface(int nx, int ny, int nz, double boxlo_x, double period_x, double average,
                            double hi, double lo, double variance, int seed, double *out);
/* grid_populate sphere concentration ... (src/grid.cpp:153-166, 339-438) */
int phafd_ic_sphere(int nx, int ny, int nz, const double *boxlo, const double *period, double radius,
                    double width_sq, double hi, double av, double *out);

/* ConjugateVolFrac thermal noise in the reference's draw order
 * (src/conjugate.cpp:163-327, src/conjugate_volfrac.cpp:72-133): one call per
 * timestep yields the per-mode field to pass to phafd_hip_set_noise */
typedef struct phafd_noise_gen phafd_noise_gen;
phafd_noise_gen *phafd_noise_create(int seed);
void phafd_noise_destroy(phafd_noise_gen *g);
int phafd_noise_next(phafd_noise_gen *g, int nx, int ny, int nz, const double *period,
                     double mobility, double temp, double dt, double *xi);

/* synthetic workload: nchains random-walk chains of nbeads (bond length
 * `bond`, beads closer than `minsep` rejected); xyz[n][3], mol[n] */
int phafd_ic_random_chains(int nchains, int nbeads, double bond, double minsep,
                           const double *boxlo, const double *period, uint64_t seed,
                           double *xyz, int *mol);

#ifdef __cplusplus
}
#endif
#endif
