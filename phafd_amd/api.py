"""ctypes binding of libphafd_hip.so (include/phafd_hip.h, include/phafd_host.h).

`System` mirrors the reference's script-level surface (domain / grid_style /
atom_populate / pair + pairCoeffs / fix / neighbor / run, src/input.cpp:48-299)
on one MI355X.  The library must have been built (phafd_amd/build.py); there
is no CPU fallback: a missing library raises immediately.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphafd_hip.so")

PHI, MU, GRADX, GRADY, GRADZ = 0, 1, 2, 3, 4
PHIHAT, MUHAT = 0, 1
ZGHOST_REFERENCE, ZGHOST_PERIODIC = 0, 1
NOISE_OFF, NOISE_DEVICE, NOISE_FIELD = 0, 1, 2
PAIR_GAUSSIAN, PAIR_LJISH, PAIR_LJISH_LINEAR = 0, 1, 3
FP64, FP32 = 0, 1
(OPT_EAGER_GRIDLIST, OPT_FUSE_FH, OPT_FUSE_SPECTRAL, OPT_EAGER_HALFLIST, OPT_ATOM_WINDOWS,
 OPT_SEMI_IMPLICIT, OPT_NEXT_GRADIENT) = 0, 1, 2, 3, 4, 5, 6
PH_START, PH_NEIGH, PH_PAIR, PH_INITIAL, PH_POSTFORCE, PH_PREFINAL, PH_FINAL, PH_POSTFINAL = range(8)
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "NaN in phi", -4: "runtime error", -5: "list overflow"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)
_fp = ctypes.POINTER(ctypes.c_float)
_lib = None


class PhafdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("boxlo", ctypes.c_double * 3), ("period", ctypes.c_double * 3),
                ("zghost", ctypes.c_int), ("precision", ctypes.c_int), ("device", ctypes.c_int)]


_cp, _pp = ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)
_vpp = ctypes.POINTER(ctypes.c_void_p)
_GEO = [_cp] + [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_int, _pp]


def lib():
    """Load libphafd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with phafd_amd/build.py "
                           "(__graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    V = ctypes.c_void_p
    sig = {
        "phafd_hip_create": ([ctypes.POINTER(Desc), ctypes.POINTER(V)], ctypes.c_int),
        "phafd_hip_destroy": ([V], ctypes.c_int),
        "phafd_hip_last_error": ([V], ctypes.c_char_p),
        "phafd_hip_set_field": ([V, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_get_field": ([V, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_get_spectrum": ([V, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_set_field_ld": ([V, ctypes.c_int, _dp, ctypes.c_int64], ctypes.c_int),
        "phafd_hip_get_field_ld": ([V, ctypes.c_int, _dp, ctypes.c_int64], ctypes.c_int),
        "phafd_hip_noise_queued": ([V, _lp], ctypes.c_int),
        "phafd_hip_noise_field": ([ctypes.c_int] * 3 + [_dp] + [ctypes.c_double] * 3 + [ctypes.c_uint64,
                                  ctypes.c_int64, _dp], ctypes.c_int),
        "phafd_hip_set_atoms": ([V, ctypes.c_int, _dp, _ip, _ip, _ip], ctypes.c_int),
        "phafd_hip_get_atoms": ([V, _dp, _dp, _ip], ctypes.c_int),
        "phafd_hip_set_images": ([V, _ip], ctypes.c_int),
        "phafd_hip_get_forces": ([V, _dp], ctypes.c_int),
        "phafd_hip_pair_gridatom": ([V, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "phafd_hip_pair_coeff": ([V, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 3, ctypes.c_int),
        "phafd_hip_pair_ljcut": ([V] + [ctypes.c_double] * 3, ctypes.c_int),
        "phafd_hip_pair_harmoniccut": ([V] + [ctypes.c_double] * 2, ctypes.c_int),
        "phafd_hip_fix_conjugate": ([V, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int], ctypes.c_int),
        "phafd_hip_set_noise": ([V, _dp], ctypes.c_int),
        "phafd_hip_fix_floryhuggins": ([V] + [ctypes.c_double] * 3, ctypes.c_int),
        "phafd_hip_fix_gradphi": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_fix_drag": ([V, ctypes.c_int, _ip, _ip, ctypes.c_double], ctypes.c_int),
        "phafd_hip_fix_chain": ([V, ctypes.c_int, _ip, _ip] + [ctypes.c_double] * 4, ctypes.c_int),
        "phafd_hip_fix_semiflexible": ([V, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int] + [ctypes.c_double] * 5,
                                       ctypes.c_int),
        "phafd_hip_neighbor": ([V, ctypes.c_double], ctypes.c_int),
        "phafd_hip_setup": ([V, ctypes.c_double], ctypes.c_int),
        "phafd_hip_run": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_phase": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_counts": ([V, _lp], ctypes.c_int),
        "phafd_hip_gridlist": ([V, ctypes.c_int64, _ip, _ip, _ip, _lp, _lp], ctypes.c_int),
        "phafd_hip_halflist": ([V, ctypes.c_int64, _ip, _ip, _ip, _lp, _lp], ctypes.c_int),
        "phafd_hip_gathered": ([V, _dp, _ip, _ip, _ip], ctypes.c_int),
        "phafd_hip_bins": ([V, _ip, _ip, _ip, _ip], ctypes.c_int),
        "phafd_hip_energy": ([V, _dp], ctypes.c_int),
        "phafd_hip_timers": ([V, _dp, ctypes.c_int], ctypes.c_int),
        "phafd_hip_probe": ([V, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_set_option": ([V, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "phafd_hip_fft_r2c": ([ctypes.c_int] * 3 + [_dp, _dp], ctypes.c_int),
        "phafd_hip_fft_c2r": ([ctypes.c_int] * 3 + [_dp, _dp], ctypes.c_int),
        "phafd_hip_gradphi": ([ctypes.c_int] * 3 + [_dp] * 5, ctypes.c_int),
        "phafd_hip_gradphi_noalloc": ([ctypes.c_int] * 3 + [_dp] + [V] * 5, ctypes.c_int),
        "phafd_hip_floryhuggins": ([ctypes.c_int64, _dp, _dp] + [ctypes.c_double] * 3 + [_lp, _lp], ctypes.c_int),
        "phafd_hip_spectral_update": ([ctypes.c_int] * 3 + [_dp, _dp, _dp] + [ctypes.c_double] * 5 + [_dp],
                                      ctypes.c_int),
        "phafd_hip_batch_create": ([ctypes.POINTER(Desc), ctypes.c_int, ctypes.POINTER(V)], ctypes.c_int),
        "phafd_hip_batch_member": ([V, ctypes.c_int, ctypes.POINTER(V)], ctypes.c_int),
        "phafd_hip_batch_setup": ([V, ctypes.c_double], ctypes.c_int),
        "phafd_hip_batch_run": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_batch_set_graphs": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_batch_info": ([V, _dp], ctypes.c_int),
        "phafd_hip_batch_set_fused": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_batch_fused_info": ([V, _dp], ctypes.c_int),
        "phafd_hip_batch_destroy": ([V], ctypes.c_int),
        "phafd_hip_batch_last_error": ([V], ctypes.c_char_p),
        "phafd_unet_create": ([ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(V)], ctypes.c_int),
        "phafd_unet_destroy": ([V], ctypes.c_int),
        "phafd_unet_last_error": ([V], ctypes.c_char_p),
        "phafd_unet_nparams": ([V, _lp], ctypes.c_int),
        "phafd_unet_get_params": ([V, _fp], ctypes.c_int),
        "phafd_unet_set_params": ([V, _fp], ctypes.c_int),
        "phafd_unet_set_norm": ([V, _dp], ctypes.c_int),
        "phafd_unet_get_norm": ([V, _dp], ctypes.c_int),
        "phafd_unet_forward": ([V, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp], ctypes.c_int),
        "phafd_unet_backward": ([V, _fp], ctypes.c_int),
        "phafd_unet_get_grads": ([V, _fp, ctypes.c_int], ctypes.c_int),
        "phafd_unet_data_alloc": ([V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "phafd_unet_data_put": ([V, ctypes.c_int, _fp, _fp], ctypes.c_int),
        "phafd_unet_train": ([V, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp], ctypes.c_int),
        "phafd_unet_eval": ([V, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_density": ([V, _fp], ctypes.c_int),
        "phafd_hip_block_residual": ([V, _dp, _dp, _fp, _fp, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_surrogate_base": ([V, _dp, _fp, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_run_surrogate": ([V, V, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp], ctypes.c_int),
        "phafd_make_unique_seed": ([ctypes.c_int], ctypes.c_int),
        "phafd_ic_noisy_constant": ([ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp],
                                    ctypes.c_int),
        "phafd_ic_flat_interface": ([ctypes.c_int] * 3 + [ctypes.c_double] * 5 + [ctypes.c_int, _dp],
                                    ctypes.c_int),
        "phafd_ic_sphere": ([ctypes.c_int] * 3 + [_dp, _dp] + [ctypes.c_double] * 4 + [_dp], ctypes.c_int),
        "phafd_noise_create": ([ctypes.c_int], V),
        "phafd_noise_destroy": ([V], None),
        "phafd_noise_next": ([V] + [ctypes.c_int] * 3 + [_dp] + [ctypes.c_double] * 3 + [_dp], ctypes.c_int),
        "phafd_ic_random_chains": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                    ctypes.c_uint64, _dp, _ip], ctypes.c_int),
        "phafd_host_last_error": ([], ctypes.c_char_p),
        "phafd_vtk_write_grid": (_GEO + [ctypes.POINTER(_dp)], ctypes.c_int),
        "phafd_vtk_write_pgrid": (_GEO + [_cp], ctypes.c_int),
        "phafd_vtk_write_atoms": ([_cp, ctypes.c_int, ctypes.c_int, _pp, _vpp, ctypes.c_int, ctypes.c_int],
                                  ctypes.c_int),
        "phafd_vtk_write_patoms": ([_cp, ctypes.c_int, _pp, _cp], ctypes.c_int),
        "phafd_vtk_collection": ([_cp, ctypes.c_int, ctypes.c_double, _cp], ctypes.c_int),
        "phafd_lammps_write_step": ([_cp, ctypes.c_int64, ctypes.c_int, _dp, _dp, ctypes.c_int, _pp, _vpp,
                                     ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "phafd_vtk_read_grid": ([_cp] + [ctypes.c_int] * 4 + [_pp, ctypes.POINTER(_dp), ctypes.c_int],
                                ctypes.c_int),
        "phafd_vtk_read_atoms": ([_cp, ctypes.c_int, ctypes.c_int, _pp, ctypes.c_int, _vpp], ctypes.c_int),
        "phafd_hip_dump_grid": ([V, _cp, ctypes.c_int, _pp, ctypes.c_int], ctypes.c_int),
        "phafd_hip_dump_atoms": ([V, _cp, ctypes.c_int, _pp] + [ctypes.c_int] * 3, ctypes.c_int),
        "phafd_hip_dump_wait": ([V], ctypes.c_int),
        "phafd_hip_compute_pair": ([V, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_fourier_capture": ([V, ctypes.c_int], ctypes.c_int),
        "phafd_hip_compute_grid_clusters": ([V, ctypes.c_double, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_run_until_touching": ([V, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _lp, _ip, _lp, _ip, _ip], ctypes.c_int),
        "phafd_hip_compute_complex": ([V, ctypes.c_int, ctypes.c_int, _dp], ctypes.c_int),
        "phafd_hip_grid_ave": ([V] + [ctypes.c_int] * 6 + [_dp], ctypes.c_int),
        "phafd_hip_read_dump_grid": ([V, _cp, ctypes.c_int, _pp, ctypes.c_int], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names of every entry point the headers declare (checked by the CPU tests)."""
    names = []
    for h in ("phafd_hip.h", "phafd_host.h"):
        path = os.path.join(os.path.dirname(HERE), "include", h)
        import re
        txt = open(path).read()
        names += re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(phafd_[a-z0-9_]+)\s*\(", txt, re.M)
    return sorted(set(names))


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f(a):
    return None if a is None else a.ctypes.data_as(_fp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _c3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _check(rc, ctx=None):
    if rc < 0:
        msg = lib().phafd_hip_last_error(ctx)
        raise PhafdError(rc, (msg or b"").decode())
    return rc


# --------------------------------------------------------------- host helpers
def make_unique_seed(seed):
    return lib().phafd_make_unique_seed(seed)


def ic_noisy_constant(shape, average, variance, seed):
    """grid_populate constant concentration <average> <variance> <seed>."""
    out = np.empty(int(np.prod(shape)))
    _check(lib().phafd_ic_noisy_constant(out.size, average, variance, seed, _d(out)))
    return out.reshape(shape)


def ic_random_chains(nchains, nbeads, bond, minsep, boxlo, period, seed):
    n = nchains * nbeads
    xyz = np.zeros((n, 3))
    mol = np.zeros(n, dtype=np.int32)
    lo = np.ascontiguousarray(boxlo, dtype=np.float64)
    pe = np.ascontiguousarray(period, dtype=np.float64)
    _check(lib().phafd_ic_random_chains(nchains, nbeads, bond, minsep, _d(lo), _d(pe), seed, _d(xyz), _i(mol)))
    return xyz, mol


# ------------------------------------------------------------ dump files
CPX_REAL, CPX_IMAG, CPX_MODULUS, CPX_NORM = 0, 1, 2, 3
CLUSTER_LT, CLUSTER_GT = 0, 1
VTK_FMT_LF, VTK_FMT_E = 0, 1
PVD_HEADER, PVD_DATASET, PVD_FOOTER = 0, 1, 2
GRID_ATTRIBUTES = ("phi", "chempot", "gradphi_x", "gradphi_y", "gradphi_z")
ATOM_VEC3 = ("x", "F", "ux")
ATOM_INT = ("ix", "ID", "molID", "type", "label")


def _names(names):
    if isinstance(names, str):
        names = [names]
    return len(names), (ctypes.c_char_p * len(names))(*[n.encode() for n in names])


def _host(rc):
    if rc < 0:
        raise PhafdError(rc, (lib().phafd_host_last_error() or b"").decode())
    return rc


def _fmt(fmt):
    if fmt in ("lf", VTK_FMT_LF):
        return VTK_FMT_LF
    if fmt in ("e", VTK_FMT_E):
        return VTK_FMT_E
    raise ValueError("Dump fmt must be either lf or e.")


def vtk_write_grid(path, arrays, names, origin, spacing):
    """Dump::write_grid_timestep of host arrays ([Nz][Ny][Nx] each)."""
    arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    nz, ny, nx = arrays[0].shape
    n, nm = _names(names)
    ptrs = (_dp * len(arrays))(*[_d(a) for a in arrays])
    _host(lib().phafd_vtk_write_grid(str(path).encode(), nx, ny, nz, _c3(origin), _c3(spacing), n, nm, ptrs))


def vtk_write_pgrid(path, shape, names, origin, spacing, source):
    nz, ny, nx = shape
    n, nm = _names(names)
    _host(lib().phafd_vtk_write_pgrid(str(path).encode(), nx, ny, nz, _c3(origin), _c3(spacing), n, nm,
                                      source.encode()))


def vtk_write_atoms(path, arrays, names, precision=-1, fmt="lf"):
    """Dump::write_atom_timestep (atom/owned); arrays per name: (n, 3) float64
    for x, F, ux and a c_<id> compute pair, (n,) int32 for ix, ID, molID,
    type, label."""
    conv = [np.ascontiguousarray(a, dtype=np.float64 if (w in ATOM_VEC3 or w.startswith("c_")) else np.int32)
            for a, w in zip(arrays, names)]
    n, nm = _names(names)
    ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in conv])
    _host(lib().phafd_vtk_write_atoms(str(path).encode(), len(conv[0]), n, nm, ptrs, precision, _fmt(fmt)))


def lammps_write_step(path, timestep, arrays, names, boxlo, boxhi, precision=-1, fmt="lf"):
    """Dump::write_lammps_dump: one timestep appended to a LAMMPS-style file."""
    conv = [np.ascontiguousarray(a, dtype=np.float64 if (w in ATOM_VEC3 or w.startswith("c_")) else np.int32)
            for a, w in zip(arrays, names)]
    n, nm = _names(names)
    ptrs = (ctypes.c_void_p * n)(*[a.ctypes.data for a in conv])
    _host(lib().phafd_lammps_write_step(str(path).encode(), int(timestep), len(conv[0]), _c3(boxlo), _c3(boxhi),
                                        n, nm, ptrs, precision, _fmt(fmt)))


def vtk_write_patoms(path, names, source):
    n, nm = _names(names)
    _host(lib().phafd_vtk_write_patoms(str(path).encode(), n, nm, source.encode()))


def vtk_collection(path, part, time=0.0, file=""):
    _host(lib().phafd_vtk_collection(str(path).encode(), part, time, file.encode()))


def vtk_read_grid(path, shape, names, no_padding=False):
    """ReadDump grid: the requested attributes of a .vti file, in request order."""
    n, nm = _names(names)
    out = [np.empty(shape) for _ in range(n)]
    ptrs = (_dp * n)(*[_d(a) for a in out])
    nz, ny, nx = shape
    _host(lib().phafd_vtk_read_grid(str(path).encode(), nx, ny, nz, n, nm, ptrs, int(no_padding)))
    return out


def read_dump_atoms(path, natoms, names, molecule=True):
    """ReadDump atoms (atom_populate read_dump): {name: array} of natoms atoms."""
    n, nm = _names(names)
    out = {w: (np.empty((natoms, 3)) if w in ATOM_VEC3 else np.empty(natoms, np.int32)) for w in names}
    ptrs = (ctypes.c_void_p * n)(*[out[w].ctypes.data for w in names])
    _host(lib().phafd_vtk_read_atoms(str(path).encode(), natoms, n, nm, int(molecule), ptrs))
    return out


class NoiseGen:
    """Reference-order mt19937 thermal noise of grid/conjugate/volfrac."""

    def __init__(self, seed):
        self.h = lib().phafd_noise_create(seed)

    def next(self, nx, ny, nz, period, mobility, temp, dt):
        out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
        pe = np.ascontiguousarray(period, dtype=np.float64)
        _check(lib().phafd_noise_next(self.h, nx, ny, nz, _d(pe), mobility, temp, dt,
                                      out.ctypes.data_as(_dp)))
        return out

    def __del__(self):
        try:
            lib().phafd_noise_destroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------ stateless kernels
def fft_r2c(phi):
    nz, ny, nx = phi.shape
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    _check(lib().phafd_hip_fft_r2c(nx, ny, nz, _d(phi), out.ctypes.data_as(_dp)))
    return out


def fft_c2r(spec, nx):
    nz, ny, hx = spec.shape
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.zeros((nz, ny, nx))
    _check(lib().phafd_hip_fft_c2r(nx, ny, nz, spec.ctypes.data_as(_dp), _d(out)))
    return out


def gradphi(phi, period):
    nz, ny, nx = phi.shape
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    g = [np.zeros_like(phi) for _ in range(3)]
    pe = np.ascontiguousarray(period, dtype=np.float64)
    _check(lib().phafd_hip_gradphi(nx, ny, nz, _d(pe), _d(phi), _d(g[0]), _d(g[1]), _d(g[2])))
    return g


def floryhuggins(phi, mu, temp, volFH, chi):
    phi = np.array(phi, dtype=np.float64, order="C")
    mu = np.array(mu, dtype=np.float64, order="C")
    a = ctypes.c_int64(0)
    b = ctypes.c_int64(0)
    _check(lib().phafd_hip_floryhuggins(phi.size, _d(phi), _d(mu), temp, volFH, chi,
                                        ctypes.byref(a), ctypes.byref(b)))
    return phi, mu, a.value, b.value


def noise_field(shape, period, mobility, temp, dt, key0, step):
    """The thermal noise the update of timestep `step` adds in NOISE_DEVICE
    mode (key0 = make_unique_seed(seed) of the fix): (Nz, Ny, Nx/2+1) complex."""
    nz, ny, nx = shape
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    pe = np.ascontiguousarray(period, dtype=np.float64)
    _check(lib().phafd_hip_noise_field(nx, ny, nz, _d(pe), mobility, temp, dt, int(key0), int(step),
                                       out.ctypes.data_as(_dp)))
    return out


def spectral_update(phihat, muhat, nx, period, mobility, temp, volFH, gamma, dt, noise=None):
    nz, ny, hx = phihat.shape
    ph = np.array(phihat, dtype=np.complex128, order="C")
    mu = np.ascontiguousarray(muhat, dtype=np.complex128)
    pe = np.ascontiguousarray(period, dtype=np.float64)
    nzz = None if noise is None else np.ascontiguousarray(noise, dtype=np.complex128)
    _check(lib().phafd_hip_spectral_update(nx, ny, nz, _d(pe), ph.ctypes.data_as(_dp), mu.ctypes.data_as(_dp),
                                           mobility, temp, volFH, gamma, dt,
                                           None if nzz is None else nzz.ctypes.data_as(_dp)))
    return ph


# ------------------------------------------------------------------- system
class System:
    """One whole-box subdomain on one GPU; method names follow the script
    commands of the reference (src/input.cpp:48-299)."""

    def __init__(self, nx, ny, nz, boxlo, period, zghost=ZGHOST_REFERENCE, precision=FP64, device=0,
                 _member=None):
        self.shape = (nz, ny, nx)
        self.boxlo = [float(v) for v in boxlo]
        self.period = [float(v) for v in period]
        self.natoms = 0
        if _member is not None:   # a batch member: the batch owns the context
            self.h, self.owned = _member, False
            return
        d = Desc(nx, ny, nz, _c3(boxlo), _c3(period), zghost, precision, device)
        h = ctypes.c_void_p()
        _check(lib().phafd_hip_create(ctypes.byref(d), ctypes.byref(h)))
        self.h, self.owned = h, True

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                lib().phafd_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return _check(rc, self.h)

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        self._c(lib().phafd_hip_set_field(self.h, which, _d(a)))

    def get_field(self, which):
        a = np.zeros(self.shape)
        self._c(lib().phafd_hip_get_field(self.h, which, _d(a)))
        return a

    def set_field_ld(self, which, buf, ld):
        """set_field from a host buffer whose x rows are ld doubles apart."""
        a = np.ascontiguousarray(buf, dtype=np.float64)
        nz, ny, nx = self.shape
        if a.size < ((nz * ny - 1) * ld + nx):
            raise ValueError("buffer too small for the padded rows")
        self._c(lib().phafd_hip_set_field_ld(self.h, which, _d(a), ld))

    def get_field_ld(self, which, buf, ld):
        """get_field into a host buffer (modified in place) whose x rows are ld doubles apart."""
        nz, ny, nx = self.shape
        if not (buf.flags["C_CONTIGUOUS"] and buf.dtype == np.float64) or buf.size < ((nz * ny - 1) * ld + nx):
            raise ValueError("need a contiguous float64 buffer holding the padded rows")
        self._c(lib().phafd_hip_get_field_ld(self.h, which, _d(buf), ld))

    def noise_queued(self):
        n = ctypes.c_int64()
        self._c(lib().phafd_hip_noise_queued(self.h, ctypes.byref(n)))
        return n.value

    def get_spectrum(self, which):
        nz, ny, nx = self.shape
        a = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
        self._c(lib().phafd_hip_get_spectrum(self.h, which, a.ctypes.data_as(_dp)))
        return a

    def set_atoms(self, xyz, types, tags=None, mols=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = xyz.shape[0]
        t = np.ascontiguousarray(types, dtype=np.int32)
        tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.int32)
        ml = None if mols is None else np.ascontiguousarray(mols, dtype=np.int32)
        self.natoms = n
        self._c(lib().phafd_hip_set_atoms(self.h, n, _d(xyz), _i(t), _i(tg), _i(ml)))
        self.atom_ints = {"type": t, "ID": np.arange(n, dtype=np.int32) if tg is None else tg,
                          "molID": np.zeros(n, np.int32) if ml is None else ml}

    def set_images(self, image):
        a = np.ascontiguousarray(image, dtype=np.int32)
        self._c(lib().phafd_hip_set_images(self.h, _i(a)))

    def get_atoms(self):
        x = np.zeros((self.natoms, 3))
        u = np.zeros((self.natoms, 3))
        im = np.zeros(self.natoms, dtype=np.int32)
        self._c(lib().phafd_hip_get_atoms(self.h, _d(x), _d(u), _i(im)))
        return x, u, im

    def get_forces(self):
        f = np.zeros((self.natoms, 3))
        self._c(lib().phafd_hip_get_forces(self.h, _d(f)))
        return f

    def compute_pair(self, ipair):
        """compute <id> pair <style>: owned forces of pair style ipair alone."""
        f = np.zeros((self.natoms, 3))
        self._c(lib().phafd_hip_compute_pair(self.h, ipair, _d(f)))
        return f

    def run_until_touching(self, nsteps, touch_point, continue_running=False, delay_step=0, cap=1000):
        """run_until_touching: {"stopped", "steps", "occurred", "ended"}."""
        occ, end = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        no, ne, done = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = self._c(lib().phafd_hip_run_until_touching(
            self.h, nsteps, touch_point, int(continue_running), delay_step, cap,
            occ.ctypes.data_as(_lp), ctypes.byref(no), end.ctypes.data_as(_lp), ctypes.byref(ne),
            ctypes.byref(done)))
        return {"stopped": rc == 1, "steps": done.value, "occurred": occ[:min(no.value, cap)].tolist(),
                "ended": end[:min(ne.value, cap)].tolist()}

    def compute_grid_clusters(self, threshold, condition="lt"):
        """compute grid/clusters phi <threshold> lt|gt: (Nz, Ny, Nx) labels."""
        cond = {"lt": CLUSTER_LT, "gt": CLUSTER_GT}.get(condition, condition)
        out = np.zeros(self.shape)
        self._c(lib().phafd_hip_compute_grid_clusters(self.h, threshold, cond, _d(out)))
        return out

    def fourier_capture(self, on=True):
        """Keep ft_phi / ft_chempot of the last step of each later run."""
        self._c(lib().phafd_hip_fourier_capture(self.h, int(on)))

    def _spec_shape(self):
        nz, ny, nx = self.shape
        return (ny, nz, nx // 2 + 1)

    def compute_complex(self, which, quant):
        """compute complex ft_phi|ft_chempot <quant>: (Ny, Nz, Nx/2+1), the
        reference's ft array order."""
        out = np.zeros(self._spec_shape())
        self._c(lib().phafd_hip_compute_complex(self.h, which, quant, _d(out)))
        return out

    def grid_ave(self, slot, which, quant, first, last, repeat, want=False):
        out = np.zeros(self._spec_shape()) if want else None
        self._c(lib().phafd_hip_grid_ave(self.h, slot, which, quant, int(first), int(last), repeat, _d(out)))
        return out

    def dump_grid(self, path, names, async_=False):
        """dump ... grid: the current fields to a .vti file (async_: the write
        overlaps later steps; dump_wait() or the next dump reports errors)."""
        n, nm = _names(names)
        self._c(lib().phafd_hip_dump_grid(self.h, str(path).encode(), n, nm, int(async_)))

    def dump_atoms(self, path, names, precision=-1, fmt="lf", async_=False):
        """dump ... atom/owned: the owned atoms to a .vtp file."""
        n, nm = _names(names)
        self._c(lib().phafd_hip_dump_atoms(self.h, str(path).encode(), n, nm, precision, _fmt(fmt), int(async_)))

    def dump_wait(self):
        self._c(lib().phafd_hip_dump_wait(self.h))

    def read_dump_grid(self, path, names, no_padding=False):
        """grid_populate read_dump: fields from a .vti file."""
        n, nm = _names(names)
        self._c(lib().phafd_hip_read_dump_grid(self.h, str(path).encode(), n, nm, int(no_padding)))

    def pair_gridatom(self, style, cut):
        return self._c(lib().phafd_hip_pair_gridatom(self.h, style, cut))

    def pair_coeff(self, ipair, itype, eps=float("nan"), phit=float("nan"), nucwidth=float("nan")):
        self._c(lib().phafd_hip_pair_coeff(self.h, ipair, itype, eps, phit, nucwidth))

    def pair_ljcut(self, cut, eps, sigma):
        return self._c(lib().phafd_hip_pair_ljcut(self.h, cut, eps, sigma))

    def pair_harmoniccut(self, cut, spring_k):
        return self._c(lib().phafd_hip_pair_harmoniccut(self.h, cut, spring_k))

    def fix_conjugate(self, seed, mobility=1.0, temp=1.0, volFH=1.0, gamma=1.0, noise=NOISE_OFF):
        self._c(lib().phafd_hip_fix_conjugate(self.h, seed, mobility, temp, volFH, gamma, noise))

    def set_noise(self, xi):
        if xi is None:
            self._c(lib().phafd_hip_set_noise(self.h, None))
        else:
            a = np.ascontiguousarray(xi, dtype=np.complex128)
            self._c(lib().phafd_hip_set_noise(self.h, a.ctypes.data_as(_dp)))

    def fix_floryhuggins(self, temp=1.0, volFH=1.0, chi=1.0):
        self._c(lib().phafd_hip_fix_floryhuggins(self.h, temp, volFH, chi))

    def fix_gradphi(self, once):
        self._c(lib().phafd_hip_fix_gradphi(self.h, 1 if once else 0))

    def fix_drag(self, starts, ends, drag):
        s = np.ascontiguousarray(starts, dtype=np.int32)
        e = np.ascontiguousarray(ends, dtype=np.int32)
        self._c(lib().phafd_hip_fix_drag(self.h, len(s), _i(s), _i(e), drag))

    def fix_chain(self, starts, ends, r0, kbond, kbend, drag):
        s = np.ascontiguousarray(starts, dtype=np.int32)
        e = np.ascontiguousarray(ends, dtype=np.int32)
        self._c(lib().phafd_hip_fix_chain(self.h, len(s), _i(s), _i(e), r0, kbond, kbend, drag))

    def fix_semiflexible(self, tether, starts, ends, seed, bondlength, zeta_para, zeta_perp, bending, temp):
        """fix atom/semiflexible/{notether,singletether,doubletether} (tether 0, 1, 2) over group chunks."""
        s = np.ascontiguousarray(starts, dtype=np.int32)
        e = np.ascontiguousarray(ends, dtype=np.int32)
        self._c(lib().phafd_hip_fix_semiflexible(self.h, tether, len(s), _i(s), _i(e), seed, bondlength, zeta_para,
                                                 zeta_perp, bending, temp))

    def neighbor(self, skin):
        self._c(lib().phafd_hip_neighbor(self.h, skin))

    def setup(self, dt):
        self._c(lib().phafd_hip_setup(self.h, dt))

    def run(self, nsteps):
        self._c(lib().phafd_hip_run(self.h, nsteps))

    def phase(self, ph):
        self._c(lib().phafd_hip_phase(self.h, ph))

    def counts(self):
        o = np.zeros(8, dtype=np.int64)
        self._c(lib().phafd_hip_counts(self.h, o.ctypes.data_as(_lp)))
        keys = ["nowned", "ngathered", "nlocal", "nghost", "ncalls", "nless", "ngreat", "overflow"]
        return dict(zip(keys, [int(v) for v in o]))

    def _list(self, fn):
        inum = ctypes.c_int64(0)
        nent = ctypes.c_int64(0)
        self._c(fn(self.h, 0, None, None, None, ctypes.byref(inum), ctypes.byref(nent)))
        cap = max(inum.value, nent.value, 1)
        il = np.zeros(max(inum.value, 1), dtype=np.int32)
        nn = np.zeros(max(inum.value, 1), dtype=np.int32)
        nb = np.zeros(max(nent.value, 1), dtype=np.int32)
        self._c(fn(self.h, cap, _i(il), _i(nn), _i(nb), ctypes.byref(inum), ctypes.byref(nent)))
        return il[:inum.value], nn[:inum.value], nb[:nent.value]

    def gridlist(self):
        return self._list(lib().phafd_hip_gridlist)

    def halflist(self):
        return self._list(lib().phafd_hip_halflist)

    def gathered(self):
        n = self.counts()["ngathered"]
        x = np.zeros((max(n, 1), 3))
        t = np.zeros(max(n, 1), dtype=np.int32)
        lab = np.zeros(max(n, 1), dtype=np.int32)
        own = np.zeros(max(n, 1), dtype=np.int32)
        self._c(lib().phafd_hip_gathered(self.h, _d(x), _i(t), _i(lab), _i(own)))
        return x[:n], t[:n], lab[:n], own[:n]

    def bins(self):
        info = np.zeros(12, dtype=np.int32)
        self._c(lib().phafd_hip_bins(self.h, _i(info), None, None, None))
        n = self.counts()["ngathered"]
        bh = np.zeros(info[9], dtype=np.int32)
        b = np.zeros(max(n, 1), dtype=np.int32)
        a2b = np.zeros(max(n, 1), dtype=np.int32)
        self._c(lib().phafd_hip_bins(self.h, _i(info), _i(bh), _i(b), _i(a2b)))
        return info, bh, b[:n], a2b[:n]

    def energy(self):
        o = np.zeros(4)
        self._c(lib().phafd_hip_energy(self.h, _d(o)))
        return o

    def density(self):
        """Bead density (sum of the grid-atom kernel values), float32 (Nz, Ny, Nx)."""
        r = np.zeros(self.shape, dtype=np.float32)
        self._c(lib().phafd_hip_density(self.h, _f(r)))
        return r

    def block_residual(self, phi0, phi1, rho0, rho1, K):
        """(residual, mass drift) of phi1 as phi0 advanced by K steps, bead
        densities rho0 and rho1 at the block's start and end."""
        a = np.ascontiguousarray(phi0, dtype=np.float64)
        b = np.ascontiguousarray(phi1, dtype=np.float64)
        r0 = np.ascontiguousarray(rho0, dtype=np.float32)
        r1 = np.ascontiguousarray(rho1, dtype=np.float32)
        o = np.zeros(2)
        self._c(lib().phafd_hip_block_residual(self.h, _d(a), _d(b), _f(r0), _f(r1), int(K), _d(o)))
        return float(o[0]), float(o[1])

    def surrogate_base(self, phi0, rho0, K):
        """phi0 advanced K steps by exponential Euler (block-start density)."""
        a = np.ascontiguousarray(phi0, dtype=np.float64)
        r0 = np.ascontiguousarray(rho0, dtype=np.float32)
        out = np.zeros(self.shape)
        self._c(lib().phafd_hip_surrogate_base(self.h, _d(a), _f(r0), int(K), _d(out)))
        return out

    def run_surrogate(self, net, nsteps, K, tol):
        """Blocks of K steps advanced by the network with the plain step as the
        check and fallback; returns the block statistics."""
        o = np.zeros(8)
        self._c(lib().phafd_hip_run_surrogate(self.h, net.h, int(nsteps), int(K), float(tol), _d(o)))
        keys = ["blocks", "accepted", "fallbacks", "max_residual", "last_residual", "max_mass_drift",
                "ms_network_path", "ms_total"]
        return dict(zip(keys, o.tolist()))

    def set_option(self, key, value):
        self._c(lib().phafd_hip_set_option(self.h, key, float(value)))

    KERNELS = ["r2c_x", "y_fwd", "y_inv", "z_fwd", "z_inv", "c2r_x", "grad_multiply", "update", "fh",
               "ga_deposit", "ga_force", "grid_count", "grid_fill", "half_count", "half_fill", "lj",
               "z_grad", "z_update", "unet_conv", "c2r_xd"]

    def probe(self, kernel=None):
        """Start (kernel name) or stop (None) HIP-event timing of one kernel;
        returns (ms, launches, algorithmic bytes) accumulated so far."""
        kid = -1 if kernel is None else self.KERNELS.index(kernel)
        o = np.zeros(3)
        self._c(lib().phafd_hip_probe(self.h, kid, _d(o)))
        return o[0], int(o[1]), o[2]

    def timers(self, reset=False):
        o = np.zeros(8)
        self._c(lib().phafd_hip_timers(self.h, _d(o), 1 if reset else 0))
        keys = ["gradphi", "neighbor", "pair", "integrate", "fh_r2c", "update", "c2r_phi", "total"]
        return dict(zip(keys, o.tolist()))


# ------------------------------------------------------------------ batches
class Batch:
    """Independent replicas of one grid advanced together (phafd_hip_batch_*);
    member(i) is a System configured like any other."""

    def __init__(self, nx, ny, nz, boxlo, period, nmembers, zghost=ZGHOST_REFERENCE, device=0):
        d = Desc(nx, ny, nz, _c3(boxlo), _c3(period), zghost, FP64, device)
        h = ctypes.c_void_p()
        rc = lib().phafd_hip_batch_create(ctypes.byref(d), nmembers, ctypes.byref(h))
        if rc < 0:
            raise PhafdError(rc, (lib().phafd_hip_batch_last_error(None) or b"").decode())
        self.h = h
        self.members = []
        for i in range(nmembers):
            m = ctypes.c_void_p()
            self._c(lib().phafd_hip_batch_member(self.h, i, ctypes.byref(m)))
            self.members.append(System(nx, ny, nz, boxlo, period, _member=m))

    def _c(self, rc):
        if rc < 0:
            raise PhafdError(rc, (lib().phafd_hip_batch_last_error(self.h) or b"").decode())
        return rc

    def member(self, i):
        return self.members[i]

    def setup(self, dt):
        self._c(lib().phafd_hip_batch_setup(self.h, dt))

    def run(self, nsteps):
        self._c(lib().phafd_hip_batch_run(self.h, nsteps))

    def set_graphs(self, on):
        self._c(lib().phafd_hip_batch_set_graphs(self.h, 1 if on else 0))

    def set_fused(self, mode):
        """-1: autotuned (default), 0: per-member pair launches, 1: one launch per pass for every member."""
        self._c(lib().phafd_hip_batch_set_fused(self.h, int(mode)))

    def fused_info(self):
        o = np.zeros(14)
        self._c(lib().phafd_hip_batch_fused_info(self.h, _d(o)))
        return {"fused": bool(o[0]), "tune_fused_ms": float(o[1]), "tune_member_ms": float(o[2]),
                "atoms_per_item": int(o[3]), "grid_cap": int(o[4]),
                "tune_shape_ms": {f"{a}x{c}": float(v) for (a, c), v in
                                  zip([(256, 1024), (256, 2048), (256, 4096), (128, 2048), (128, 4096), (64, 4096)],
                                      o[5:11])},
                "step_choice_decided": bool(o[11]), "step_ms_per_member": float(o[12]),
                "step_ms_one_launch_per_pass": float(o[13])}

    def info(self):
        o = np.zeros(12)
        self._c(lib().phafd_hip_batch_info(self.h, _d(o)))
        return {"members": int(o[0]), "members_per_launch": int(o[1]), "member_streams": int(o[2]),
                "graphs": bool(o[3]), "tune_grad_ms": list(o[4:8]), "tune_pair_ms": list(o[8:12])}

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            lib().phafd_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- surrogate
class UNet:
    """The surrogate network (phafd_unet_*): parameters in PyTorch Conv3d
    order per layer (LAYERS), training and inference on the GPU."""

    # (name, cin, cout, kernel size, resolution level)
    LAYERS = [("e1", 4, 16, 3, 1), ("e1b", 16, 16, 3, 1), ("e2", 16, 32, 3, 2), ("b", 32, 64, 3, 3),
              ("q2", 64, 32, 1, 3), ("d2", 32, 32, 3, 2), ("q1", 32, 16, 1, 2), ("d1", 16, 16, 3, 1),
              ("head", 16, 1, 1, 1)]

    def __init__(self, seed=1, device=0):
        h = ctypes.c_void_p()
        rc = lib().phafd_unet_create(device, seed, ctypes.byref(h))
        if rc < 0:
            raise PhafdError(rc, (lib().phafd_unet_last_error(None) or b"").decode())
        self.h = h
        n = ctypes.c_int64()
        self._c(lib().phafd_unet_nparams(self.h, ctypes.byref(n)))
        self.nparams = int(n.value)

    def _c(self, rc):
        if rc < 0:
            raise PhafdError(rc, (lib().phafd_unet_last_error(self.h) or b"").decode())
        return rc

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        self._c(lib().phafd_unet_get_params(self.h, _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        self._c(lib().phafd_unet_set_params(self.h, _f(p)))

    def set_norm(self, norm5):
        a = np.ascontiguousarray(norm5, dtype=np.float64)
        self._c(lib().phafd_unet_set_norm(self.h, _d(a)))

    def get_norm(self):
        a = np.zeros(5)
        self._c(lib().phafd_unet_get_norm(self.h, _d(a)))
        return a

    def forward(self, in4):
        """in4: float32 (4, Nz, Ny, Nx) -> float32 (Nz, Ny, Nx)."""
        x = np.ascontiguousarray(in4, dtype=np.float32)
        _, nz, ny, nx = x.shape
        out = np.zeros((nz, ny, nx), dtype=np.float32)
        self._c(lib().phafd_unet_forward(self.h, nx, ny, nz, _f(x), _f(out)))
        return out

    def backward(self, dout):
        d = np.ascontiguousarray(dout, dtype=np.float32)
        self._c(lib().phafd_unet_backward(self.h, _f(d)))

    def get_grads(self, zero=True):
        g = np.zeros(self.nparams, dtype=np.float32)
        self._c(lib().phafd_unet_get_grads(self.h, _f(g), 1 if zero else 0))
        return g

    def data_alloc(self, nx, ny, nz, nsamples):
        self._c(lib().phafd_unet_data_alloc(self.h, nx, ny, nz, nsamples))

    def data_put(self, i, in4, target):
        x = np.ascontiguousarray(in4, dtype=np.float32)
        y = np.ascontiguousarray(target, dtype=np.float32)
        self._c(lib().phafd_unet_data_put(self.h, i, _f(x), _f(y)))

    def train(self, nsteps, batch, ntrain, lr):
        losses = np.zeros(nsteps)
        self._c(lib().phafd_unet_train(self.h, nsteps, batch, ntrain, lr, _d(losses)))
        return losses

    def eval(self, i0, count):
        losses = np.zeros(count)
        self._c(lib().phafd_unet_eval(self.h, i0, count, _d(losses)))
        return losses

    def save(self, path):
        np.savez(path, params=self.get_params(), norm=self.get_norm())

    def load(self, path):
        z = np.load(path)
        self.set_params(z["params"])
        self.set_norm(z["norm"])

    def close(self):
        if getattr(self, "h", None):
            lib().phafd_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def surrogate_inputs(phi, rho0, rho1, norm):
    """The network's four input channels (float32) from phi and the bead
    densities at the block's start and end, as the library forms them on
    the device."""
    phi = np.asarray(phi, dtype=np.float64)
    mean = phi.mean()
    c0 = (phi - mean) / norm[0]
    c1 = np.asarray(rho0, dtype=np.float64) / norm[1]
    c2 = np.asarray(rho1, dtype=np.float64) / norm[1]
    c3 = np.full(phi.shape, (mean - norm[2]) / norm[3])
    return np.stack([c0, c1, c2, c3]).astype(np.float32)
