"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PHI, MU, GRADX, GRADY, GRADZ = 0, 1, 2, 3, 4
ZGHOST_REFERENCE, ZGHOST_PERIODIC = 0, 1
NOISE_OFF, NOISE_MT19937, NOISE_FIELD = 0, 1, 2
PAIR_GAUSSIAN, PAIR_LJISH, PAIR_LJCUT, PAIR_LJISH_LINEAR, PAIR_HARMONIC = 0, 1, 2, 3, 4
PH_START, PH_NEIGH, PH_PAIR, PH_INITIAL, PH_POSTFORCE, PH_PREFINAL, PH_FINAL, PH_POSTFINAL = range(8)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_last_error.argtypes = [ctypes.c_void_p]
        for name in ("orc_set_field", "orc_get_field"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_set_atoms.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip, _ip, _ip]
        L.orc_get_atoms.argtypes = [ctypes.c_void_p, _dp, _dp, _ip]
        L.orc_get_forces.argtypes = [ctypes.c_void_p, _dp]
        L.orc_add_pair.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 3
        L.orc_pair_coeff.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 3
        L.orc_add_fix_conjugate.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int]
        L.orc_set_noise.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_semi_implicit.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_add_fix_floryhuggins.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
        L.orc_add_fix_gradphi.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_add_fix_drag.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, ctypes.c_double]
        L.orc_add_fix_chain.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip] + [ctypes.c_double] * 4
        L.orc_add_fix_semiflexible.argtypes = ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int]
                                               + [ctypes.c_double] * 5)
        L.orc_semiflexible_seeds.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_philox4x64_10.argtypes = [_u64p, _u64p, _u64p]
        L.orc_beadrod_step.argtypes = ([ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_uint64, ctypes.c_int64]
                                       + [ctypes.c_double] * 6 + [ctypes.c_int, _ip])
        L.orc_neighbor.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_setup.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_phase.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_compute_pair.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_compute_complex.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_compute_grid_clusters.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_run_until_touching.argtypes = ([ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, _lp, _ip, _lp, _ip, _ip])
        L.orc_counts.argtypes = [ctypes.c_void_p, _lp]
        for name in ("orc_gridlist", "orc_halflist"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int64, _ip, _ip, _ip, _lp, _lp]
        L.orc_gathered.argtypes = [ctypes.c_void_p, _dp, _ip, _ip, _ip]
        L.orc_bins.argtypes = [ctypes.c_void_p, _ip, _ip, _ip, _ip]
        L.orc_energy.argtypes = [ctypes.c_void_p, _dp]
        L.orc_fft_r2c.argtypes = [ctypes.c_int] * 3 + [_dp, _dp]
        L.orc_fft_c2r.argtypes = [ctypes.c_int] * 3 + [_dp, _dp]
        L.orc_gradphi.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, _dp, _dp, _dp]
        L.orc_floryhuggins.argtypes = [ctypes.c_int64, _dp, _dp] + [ctypes.c_double] * 3 + [_lp, _lp]
        L.orc_spectral_update.argtypes = ([ctypes.c_int] * 3 + [_dp, _dp, _dp] + [ctypes.c_double] * 5
                                          + [ctypes.c_int, ctypes.c_int, _dp])
        L.orc_noise_field.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_int] + [ctypes.c_double] * 3 + [_dp]
        L.orc_mt19937_uniform.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_make_unique_seed.argtypes = [ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_mt19937_raw.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
        _cp, _pp = ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)
        _vpp = ctypes.POINTER(ctypes.c_void_p)
        L.orc_vtk_last_error.restype = ctypes.c_char_p
        _geo = [_cp] + [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_int, _pp]
        L.orc_vtk_write_grid.argtypes = _geo + [ctypes.POINTER(_dp)]
        L.orc_vtk_write_pgrid.argtypes = _geo + [_cp]
        L.orc_vtk_write_atoms.argtypes = [_cp, ctypes.c_int, ctypes.c_int, _pp, _vpp, ctypes.c_int, ctypes.c_int]
        L.orc_vtk_write_patoms.argtypes = [_cp, ctypes.c_int, _pp, _cp]
        L.orc_vtk_collection.argtypes = [_cp, ctypes.c_int, ctypes.c_double, _cp]
        L.orc_lammps_write_step.argtypes = [_cp, ctypes.c_int64, ctypes.c_int, _dp, _dp, ctypes.c_int, _pp, _vpp,
                                            ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def _c3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


# ------------------------------------------------------------ stateless
def fft_r2c(phi):
    nz, ny, nx = phi.shape
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    rc = lib().orc_fft_r2c(nx, ny, nz, dptr(phi), out.ctypes.data_as(_dp))
    assert rc == 0
    return out


def fft_c2r(spec, nx):
    nz, ny, hx = spec.shape
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.zeros((nz, ny, nx), dtype=np.float64)
    rc = lib().orc_fft_c2r(nx, ny, nz, spec.ctypes.data_as(_dp), dptr(out))
    assert rc == 0
    return out


def gradphi(phi, period):
    nz, ny, nx = phi.shape
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    g = [np.zeros_like(phi) for _ in range(3)]
    rc = lib().orc_gradphi(nx, ny, nz, _c3(period), dptr(phi), dptr(g[0]), dptr(g[1]), dptr(g[2]))
    assert rc == 0
    return g


def floryhuggins(phi, mu, temp, volFH, chi):
    phi = np.array(phi, dtype=np.float64, order="C")
    mu = np.array(mu, dtype=np.float64, order="C")
    a = ctypes.c_int64(0)
    b = ctypes.c_int64(0)
    lib().orc_floryhuggins(phi.size, dptr(phi), dptr(mu), temp, volFH, chi, ctypes.byref(a), ctypes.byref(b))
    return phi, mu, a.value, b.value


def spectral_update(phihat, muhat, nx, period, mobility, temp, volFH, gamma, dt,
                    noise_mode=NOISE_OFF, seed=0, noise=None):
    nz, ny, hx = phihat.shape
    ph = np.array(phihat, dtype=np.complex128, order="C")
    mu = np.ascontiguousarray(muhat, dtype=np.complex128)
    nz_ = None if noise is None else np.ascontiguousarray(noise, dtype=np.complex128)
    rc = lib().orc_spectral_update(nx, ny, nz, _c3(period), ph.ctypes.data_as(_dp), mu.ctypes.data_as(_dp),
                                   mobility, temp, volFH, gamma, dt, noise_mode, seed,
                                   None if nz_ is None else nz_.ctypes.data_as(_dp))
    assert rc == 0
    return ph


def noise_field(nx, ny, nz, period, seed, mobility, temp, dt):
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    rc = lib().orc_noise_field(nx, ny, nz, _c3(period), seed, mobility, temp, dt, out.ctypes.data_as(_dp))
    assert rc == 0
    return out


def mt19937_uniform(seed, n):
    out = np.zeros(n)
    lib().orc_mt19937_uniform(seed, n, dptr(out))
    return out


def mt19937_raw(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt19937_raw(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def make_unique_seed(seed):
    return lib().orc_make_unique_seed(seed)


def philox4x64_10(counter, key):
    """The four output words of Philox4x64-10 for one counter and key."""
    c = np.ascontiguousarray(counter, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    p = ctypes.POINTER(ctypes.c_uint64)
    lib().orc_philox4x64_10(c.ctypes.data_as(p), k.ctypes.data_as(p), out.ctypes.data_as(p))
    return out


def beadrod_step(R, Fext, tether=0, seed=1, step=1, b=1.0, zpar=1.0, zperp=1.0, bending=0.0, temp=0.0, dt=1e-3,
                 itermax=50):
    """One bead-rod timestep of one polymer; returns (new R, corrector iterations)."""
    R = np.array(R, dtype=np.float64, order="C")
    F = np.ascontiguousarray(Fext, dtype=np.float64)
    it = np.zeros(1, dtype=np.int32)
    rc = lib().orc_beadrod_step(tether, len(R), dptr(R), dptr(F), seed, step, b, zpar, zperp, bending, temp, dt,
                                itermax, iptr(it))
    if rc != 0:
        raise RuntimeError(lib().orc_last_error(None).decode())
    return R, int(it[0])


def set_threads(n):
    """OpenMP threads of the oracle's FFT lines and pointwise loops."""
    if lib().orc_set_threads(int(n)) != 0:
        raise ValueError("thread count must be >= 1")


def threads():
    return int(lib().orc_threads())


# ------------------------------------------------------------ system
class System:
    """Whole-box single-rank restatement of PHAFD's Integrate::setup/run."""

    def __init__(self, nx, ny, nz, boxlo, period, zghost=ZGHOST_REFERENCE):
        self.shape = (nz, ny, nx)
        self.h = lib().orc_create(nx, ny, nz, _c3(boxlo), _c3(period), zghost)
        if not self.h:
            raise ValueError("oracle: invalid grid")
        self.pair_styles = []

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise RuntimeError(lib().orc_last_error(self.h).decode())
        return rc

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        self._chk(lib().orc_set_field(self.h, which, dptr(a)))

    def get_field(self, which):
        a = np.zeros(self.shape)
        self._chk(lib().orc_get_field(self.h, which, dptr(a)))
        return a

    def set_atoms(self, xyz, types, tags=None, mols=None):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = xyz.shape[0]
        t = np.ascontiguousarray(types, dtype=np.int32)
        tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.int32)
        ml = None if mols is None else np.ascontiguousarray(mols, dtype=np.int32)
        self.natoms = n
        self._chk(lib().orc_set_atoms(self.h, n, dptr(xyz), iptr(t), iptr(tg), iptr(ml)))

    def get_atoms(self):
        x = np.zeros((self.natoms, 3))
        u = np.zeros((self.natoms, 3))
        im = np.zeros(self.natoms, dtype=np.int32)
        lib().orc_get_atoms(self.h, dptr(x), dptr(u), iptr(im))
        return x, u, im

    def get_forces(self):
        f = np.zeros((self.natoms, 3))
        lib().orc_get_forces(self.h, dptr(f))
        return f

    def compute_complex(self, which, quant):
        """(Ny, Nz, Nx/2+1) array; between PH_PREFINAL and PH_FINAL."""
        nz, ny, nx = self.shape
        out = np.zeros((ny, nz, nx // 2 + 1))
        self._chk(lib().orc_compute_complex(self.h, which, quant, dptr(out)))
        return out

    def run_until_touching(self, nsteps, touch_point, continue_running=False, delay_step=0, cap=1000):
        occ, end = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        no, ne, done = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        rc = self._chk(lib().orc_run_until_touching(
            self.h, nsteps, touch_point, int(continue_running), delay_step, cap,
            occ.ctypes.data_as(_lp), ctypes.byref(no), end.ctypes.data_as(_lp), ctypes.byref(ne),
            ctypes.byref(done)))
        return {"stopped": rc == 1, "steps": done.value, "occurred": occ[:no.value].tolist(),
                "ended": end[:ne.value].tolist()}

    def compute_grid_clusters(self, threshold, condition=0):
        out = np.zeros(self.shape)
        self._chk(lib().orc_compute_grid_clusters(self.h, threshold, condition, dptr(out)))
        return out

    def compute_pair(self, ipair):
        f = np.zeros((self.natoms, 3))
        self._chk(lib().orc_compute_pair(self.h, ipair, dptr(f)))
        return f

    def add_pair(self, style, a, b=0.0, c=0.0):
        return self._chk(lib().orc_add_pair(self.h, style, a, b, c))

    def pair_coeff(self, ipair, itype, eps=float("nan"), phit=float("nan"), nucwidth=float("nan")):
        self._chk(lib().orc_pair_coeff(self.h, ipair, itype, eps, phit, nucwidth))

    def fix_conjugate(self, seed, mobility=1.0, temp=1.0, volFH=1.0, gamma=1.0, noise=NOISE_OFF):
        self._chk(lib().orc_add_fix_conjugate(self.h, seed, mobility, temp, volFH, gamma, noise))

    def set_noise(self, xi):
        if xi is None:
            lib().orc_set_noise(self.h, None)
        else:
            a = np.ascontiguousarray(xi, dtype=np.complex128)
            lib().orc_set_noise(self.h, a.ctypes.data_as(_dp))

    def set_semi_implicit(self, on=True):
        lib().orc_set_semi_implicit(self.h, int(on))

    def fix_floryhuggins(self, temp=1.0, volFH=1.0, chi=1.0):
        self._chk(lib().orc_add_fix_floryhuggins(self.h, temp, volFH, chi))

    def fix_gradphi(self, once):
        self._chk(lib().orc_add_fix_gradphi(self.h, 1 if once else 0))

    def fix_drag(self, starts, ends, drag):
        s = np.ascontiguousarray(starts, dtype=np.int32)
        e = np.ascontiguousarray(ends, dtype=np.int32)
        self._chk(lib().orc_add_fix_drag(self.h, len(s), iptr(s), iptr(e), drag))

    def fix_chain(self, starts, ends, r0, kbond, kbend, drag):
        s = np.ascontiguousarray(starts, dtype=np.int32)
        e = np.ascontiguousarray(ends, dtype=np.int32)
        self._chk(lib().orc_add_fix_chain(self.h, len(s), iptr(s), iptr(e), r0, kbond, kbend, drag))

    def fix_semiflexible(self, tether, starts, ends, seed, bondlength, zeta_para, zeta_perp, bending, temp):
        """fix atom/semiflexible/{notether,singletether,doubletether} (tether 0, 1, 2)."""
        s = np.ascontiguousarray(starts, dtype=np.int32)
        e = np.ascontiguousarray(ends, dtype=np.int32)
        self._chk(lib().orc_add_fix_semiflexible(self.h, tether, len(s), iptr(s), iptr(e), seed, bondlength,
                                                 zeta_para, zeta_perp, bending, temp))

    def semiflexible_seeds(self, ifix, npoly):
        out = np.zeros(npoly, dtype=np.int32)
        self._chk(lib().orc_semiflexible_seeds(self.h, ifix, iptr(out)))
        return out

    def neighbor(self, skin):
        self._chk(lib().orc_neighbor(self.h, skin))

    def setup(self, dt):
        self._chk(lib().orc_setup(self.h, dt))

    def run(self, nsteps):
        self._chk(lib().orc_run(self.h, nsteps))

    def phase(self, ph):
        self._chk(lib().orc_phase(self.h, ph))

    def counts(self):
        o = np.zeros(8, dtype=np.int64)
        lib().orc_counts(self.h, o.ctypes.data_as(_lp))
        keys = ["nowned", "ngathered", "nlocal", "nghost", "ncalls", "nless", "ngreat", "overflow"]
        return dict(zip(keys, [int(v) for v in o]))

    def _list(self, fn):
        inum = ctypes.c_int64(0)
        nent = ctypes.c_int64(0)
        fn(self.h, 0, None, None, None, ctypes.byref(inum), ctypes.byref(nent))
        cap = max(inum.value, nent.value, 1)
        il = np.zeros(inum.value, dtype=np.int32)
        nn = np.zeros(inum.value, dtype=np.int32)
        nb = np.zeros(max(nent.value, 1), dtype=np.int32)
        self._chk(fn(self.h, cap, iptr(il), iptr(nn), iptr(nb), ctypes.byref(inum), ctypes.byref(nent)))
        return il, nn, nb[:nent.value]

    def gridlist(self):
        return self._list(lib().orc_gridlist)

    def halflist(self):
        return self._list(lib().orc_halflist)

    def gathered(self):
        n = self.counts()["ngathered"]
        x = np.zeros((n, 3))
        t = np.zeros(n, dtype=np.int32)
        lab = np.zeros(n, dtype=np.int32)
        own = np.zeros(n, dtype=np.int32)
        lib().orc_gathered(self.h, dptr(x), iptr(t), iptr(lab), iptr(own))
        return x, t, lab, own

    def bins(self):
        info = np.zeros(12, dtype=np.int32)
        lib().orc_bins(self.h, iptr(info), None, None, None)
        n = self.counts()["ngathered"]
        bh = np.zeros(info[9], dtype=np.int32)
        b = np.zeros(max(n, 1), dtype=np.int32)
        a2b = np.zeros(max(n, 1), dtype=np.int32)
        lib().orc_bins(self.h, iptr(info), iptr(bh), iptr(b), iptr(a2b))
        return info, bh, b[:n], a2b[:n]

    def energy(self):
        o = np.zeros(4)
        self._chk(lib().orc_energy(self.h, dptr(o)))
        return o


# ------------------------------------------------------------ dump writers
def _names(names):
    return (ctypes.c_char_p * len(names))(*[n.encode() for n in names])


def _vtk(rc):
    if rc < 0:
        raise RuntimeError(lib().orc_vtk_last_error().decode())


def vtk_write_grid(path, arrays, names, origin, spacing):
    """Dump::write_grid_timestep restated with std::ostream (one rank)."""
    arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    nz, ny, nx = arrays[0].shape
    ptrs = (_dp * len(arrays))(*[a.ctypes.data_as(_dp) for a in arrays])
    _vtk(lib().orc_vtk_write_grid(path.encode(), nx, ny, nz, (ctypes.c_double * 3)(*origin),
                                  (ctypes.c_double * 3)(*spacing), len(names), _names(names), ptrs))


def vtk_write_pgrid(path, shape, names, origin, spacing, source):
    nz, ny, nx = shape
    _vtk(lib().orc_vtk_write_pgrid(path.encode(), nx, ny, nz, (ctypes.c_double * 3)(*origin),
                                   (ctypes.c_double * 3)(*spacing), len(names), _names(names), source.encode()))


def vtk_write_atoms(path, arrays, names, precision=-1, fmt=0):
    """arrays: per name, (n, 3) float64 for x/F/ux, (n,) int32 otherwise."""
    conv = [np.ascontiguousarray(a, dtype=np.float64 if a.ndim == 2 else np.int32) for a in arrays]
    n = len(conv[0])
    ptrs = (ctypes.c_void_p * len(conv))(*[a.ctypes.data for a in conv])
    _vtk(lib().orc_vtk_write_atoms(path.encode(), n, len(names), _names(names), ptrs, precision, fmt))


def vtk_write_patoms(path, names, source):
    _vtk(lib().orc_vtk_write_patoms(path.encode(), len(names), _names(names), source.encode()))


def vtk_collection(path, part, time=0.0, file=""):
    _vtk(lib().orc_vtk_collection(path.encode(), part, time, file.encode()))


def lammps_write_step(path, timestep, arrays, names, boxlo, boxhi, precision=-1, fmt=0):
    conv = [np.ascontiguousarray(a, dtype=np.float64 if a.ndim == 2 else np.int32) for a in arrays]
    ptrs = (ctypes.c_void_p * len(conv))(*[a.ctypes.data for a in conv])
    _vtk(lib().orc_lammps_write_step(path.encode(), timestep, len(conv[0]), (ctypes.c_double * 3)(*boxlo),
                                     (ctypes.c_double * 3)(*boxhi), len(names), _names(names), ptrs, precision, fmt))
