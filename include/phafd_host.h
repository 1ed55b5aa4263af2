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

/* ---- dump files (the reference's VTK formats, one subdomain) ----------
 * Attribute names are the reference's: grid phi, chempot, gradphi_x,
 * gradphi_y, gradphi_z; atoms x, F, ux (three doubles per atom) and ix, ID,
 * molID, type, label (one int per atom).  On error the functions return a
 * negative value and phafd_host_last_error() gives the message. */
enum { PHAFD_VTK_FMT_LF = 0, PHAFD_VTK_FMT_E = 1 };   /* dump ... fmt lf|e */
enum { PHAFD_VTK_PVD_HEADER = 0, PHAFD_VTK_PVD_DATASET = 1, PHAFD_VTK_PVD_FOOTER = 2 };
const char *phafd_host_last_error(void);
/* Dump::write_grid_timestep (src/dump.cpp:681-780): ImageData with the
 * arrays ([Nz][Ny][Nx] doubles each) appended raw after 32-bit byte lengths;
 * one rank, so the extent is 0..Nz-1 and each length is Nx*Ny*Nz doubles */
int phafd_vtk_write_grid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                         int nattr, const char *const *names, const double *const *arrays);
/* Dump::write_grid_timestep_pvti (src/dump.cpp:782-855) */
int phafd_vtk_write_pgrid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                          int nattr, const char *const *names, const char *source);
/* Dump::write_atom_timestep, atom/owned (src/dump.cpp:348-418, 1014-1133);
 * precision <= 0: the stream default; fmt PHAFD_VTK_FMT_* */
int phafd_vtk_write_atoms(const char *path, int n, int nattr, const char *const *names, const void *const *arrays,
                          int precision, int fmt);
/* Dump::write_atom_timestep_pvtp (src/dump.cpp:422-483) */
int phafd_vtk_write_patoms(const char *path, int nattr, const char *const *names, const char *source);
/* dump <name> lammpsatom <file> <every> attributes n ...: Dump::write_lammps_dump
 * (src/dump.cpp:485-679), one timestep appended to path (create or truncate
 * the file when the dump is set up); arrays as phafd_vtk_write_atoms,
 * boxlo/boxhi the box bounds */
int phafd_lammps_write_step(const char *path, int64_t timestep, int n, const double *boxlo, const double *boxhi,
                            int nattr, const char *const *names, const void *const *arrays, int precision, int fmt);
/* Dump::write_collection_header / _middle / _footer (src/dump.cpp:235-313):
 * part PHAFD_VTK_PVD_*; time and file are used by the data-set line */
int phafd_vtk_collection(const char *path, int part, double time, const char *file);
/* ReadDump grid (src/read_dump.cpp:90-185, 307-360): out[k] (Nx*Ny*Nz
 * doubles, may be null to skip) receives requested attribute k; phi must be
 * requested */
int phafd_vtk_read_grid(const char *path, int nx, int ny, int nz, int nattr, const char *const *names,
                        double *const *out, int no_padding);
/* ReadDump atoms (src/read_dump.cpp:188-305): out[k] receives requested
 * attribute k for natoms atoms; x, type, ID (and molID, ix when molecule)
 * must be requested; returns natoms */
int phafd_vtk_read_atoms(const char *path, int natoms, int nattr, const char *const *names, int molecule,
                         void *const *out);

#ifdef __cplusplus
}
#endif
#endif
