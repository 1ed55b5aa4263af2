"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Tolerances (relative L2 unless noted):
  FFT / grad(phi) ....... 1e-12   (different FFT algorithms, FP64)
  spectral update ....... bit-exact given the same spectra
  Flory-Huggins ......... 1e-15   (device log vs glibc log, <= 1 ulp)
  gathered atoms, bins,
  grid and half lists ... bit-exact (integer / index results)
  mu, forces, one step .. 1e-10   (the stated FP64 tolerance)
  200-step trajectory ... 1e-10 for phi, positions
  bead pair styles ...... lj/cut and harmonic/cut at the same 1e-10; the
                          pruned list against the full half list bitwise
"""
import os

import numpy as np
import pytest

import phafd_amd as P
from oracle import oracle_ctypes as O
from tests.helpers import workload, make_gpu, make_oracle, rel_l2

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def _tags():
    return sorted({k.split("_", 2)[2] for k in GOLD.files if min(map(int, k.split("_", 2)[2].split("x"))) >= 4})


@pytest.mark.parametrize("tag", _tags())
def test_fft_golden(tag):
    a = GOLD[f"r2c_in_{tag}"]
    ref = GOLD[f"r2c_out_{tag}"]
    assert rel_l2(P.fft_r2c(a).view(float), ref.view(float)) < 1e-13
    c = GOLD[f"c2r_in_{tag}"]
    nx = int(tag.split("x")[2])
    assert rel_l2(P.fft_c2r(c, nx), GOLD[f"c2r_out_{tag}"]) < 1e-13


@pytest.mark.parametrize("shape", [(64, 64, 64), (32, 16, 128), (128, 8, 4), (4, 256, 512), (16, 1024, 8),
                                   (512, 8, 16), (8, 512, 16), (4, 8, 1024)])
def test_fft_vs_oracle(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    ref = O.fft_r2c(a)
    got = P.fft_r2c(a)
    assert rel_l2(got.view(float), ref.view(float)) < 1e-13
    c = ref * (1 + 0.1j)
    assert rel_l2(P.fft_c2r(c, shape[2]), O.fft_c2r(c, shape[2])) < 1e-13


def test_gradphi_vs_oracle():
    shape = (32, 16, 64)
    L = [64.0, 12.0, 40.0]
    phi = np.random.default_rng(1).standard_normal(shape)
    for g, r in zip(P.gradphi(phi, L), O.gradphi(phi, L)):
        assert rel_l2(g, r) < 1e-12


def test_floryhuggins_vs_oracle():
    rng = np.random.default_rng(2)
    phi = rng.uniform(-0.2, 1.2, 100000)
    mu = rng.standard_normal(100000)
    g = P.floryhuggins(phi, mu, 1.1, 0.9, 2.7)
    r = O.floryhuggins(phi, mu, 1.1, 0.9, 2.7)
    assert np.array_equal(g[0], r[0])
    assert g[2:] == r[2:]
    assert rel_l2(g[1], r[1]) < 1e-15


@pytest.mark.parametrize("noise", [False, True])
def test_spectral_update_bit_exact(noise):
    nx, ny, nz = 16, 8, 32
    L = [16.0, 12.0, 40.0]
    rng = np.random.default_rng(3)
    ph = O.fft_r2c(rng.standard_normal((nz, ny, nx)))
    mh = O.fft_r2c(rng.standard_normal((nz, ny, nx)))
    xi = O.noise_field(nx, ny, nz, L, 77, 1.0, 1.0, 1e-3) if noise else None
    mode = O.NOISE_FIELD if noise else O.NOISE_OFF
    r = O.spectral_update(ph, mh, nx, L, 1.2, 0.8, 1.1, 0.7, 1e-3, mode, 0, xi)
    g = P.spectral_update(ph, mh, nx, L, 1.2, 0.8, 1.1, 0.7, 1e-3, xi)
    assert np.array_equal(g, r)


def _lists_equal(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("zghost", [0, 1])
def test_setup_ghosts_bins_lists_bit_exact(zghost):
    w = workload(n=32, nchains=12, nbeads=25, zghost=zghost, seed=3)
    g = make_gpu(w)
    o = make_oracle(w)
    cg, co = g.counts(), o.counts()
    for k in ("nowned", "ngathered", "nlocal", "nghost", "ncalls"):
        assert cg[k] == co[k], k
    for a, b in zip(g.gathered(), o.gathered()):
        assert np.array_equal(a, b)
    for a, b in zip(g.bins(), o.bins()):
        assert np.array_equal(a, b)
    assert _lists_equal(g.gridlist(), o.gridlist())
    assert _lists_equal(g.halflist(), o.halflist())


def test_list_overflow_is_reported_like_the_reference(capfd):
    # more than Neighbor::oneatom (2000) atoms within reach of a grid point:
    # the reference prints a warning and goes on
    # (src/npair_grid_bin.cpp:118-121); the list stays complete.  The step
    # walks its own grid-atom entries; the reference's grid list (and its
    # warning) is built when it is asked for.
    w = workload(n=16, nchains=2, nbeads=5, lj=False, chain=False, seed=5)
    rng = np.random.default_rng(5)
    w["xyz"] = 7.0 + 2.0 * rng.random((2100, 3))
    w["mol"] = np.zeros(2100, np.int32)
    w["types"] = np.zeros(2100, np.int32)
    g = make_gpu(w)
    o = make_oracle(w)
    try:
        capfd.readouterr()
        gl = g.gridlist()
        err = capfd.readouterr().err
        assert "Neighbor list overflow, terrible things are going to happen." in err
        ol = o.gridlist()
        assert gl[1].max() > 2000   # numneigh of the densest grid point
        assert _lists_equal(gl, ol)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("ljish", [False, True, "linear"])
def test_one_step_phases(ljish):
    w = workload(n=32, nchains=10, nbeads=30, ljish=ljish, seed=4, nucwidth=2.0)
    g = make_gpu(w)
    o = make_oracle(w)
    for ph in range(8):
        g.phase(ph)
        o.phase(ph)
        if ph == P.PH_START:
            for f in (P.GRADX, P.GRADY, P.GRADZ):
                assert rel_l2(g.get_field(f), o.get_field(f)) < 1e-12
        if ph == P.PH_PAIR:
            assert rel_l2(g.get_field(P.MU), o.get_field(P.MU)) < 1e-10
            assert rel_l2(g.get_forces(), o.get_forces()) < 1e-10
        if ph == P.PH_POSTFORCE:
            assert rel_l2(g.get_field(P.MU), o.get_field(P.MU)) < 1e-10
        if ph in (P.PH_INITIAL, P.PH_FINAL):
            assert rel_l2(g.get_atoms()[0], o.get_atoms()[0]) < 1e-12
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-12


def test_pair_pass_bit_exact_with_identical_inputs():
    # With grad(phi) taken from the oracle, mu and forces must match the
    # reference loop order exactly (exp() aside: compared at 1e-15).
    w = workload(n=32, nchains=10, nbeads=30, seed=5, nucwidth=2.5)
    g = make_gpu(w)
    o = make_oracle(w)
    for s in (g, o):
        s.phase(P.PH_START)
        s.phase(P.PH_NEIGH)
    for f in (P.GRADX, P.GRADY, P.GRADZ):
        g.set_field(f, o.get_field(f))
    g.phase(P.PH_PAIR)
    o.phase(P.PH_PAIR)
    assert rel_l2(g.get_field(P.MU), o.get_field(P.MU)) < 1e-15
    assert rel_l2(g.get_forces(), o.get_forces()) < 1e-14


@pytest.mark.parametrize("zghost", [0, 1])
def test_trajectory_200_steps(zghost):
    w = workload(n=64, nchains=20, nbeads=50, seed=42, zghost=zghost, chi=3.0)
    g = make_gpu(w)
    o = make_oracle(w)
    g.run(200)
    o.run(200)
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
    assert rel_l2(g.get_atoms()[0], o.get_atoms()[0]) < 1e-10
    assert rel_l2(g.get_forces(), o.get_forces()) < 1e-8
    assert g.counts()["ncalls"] == o.counts()["ncalls"]
    assert _lists_equal(g.gridlist(), o.gridlist())
    ge, oe = g.energy(), o.energy()
    assert np.allclose(ge, oe, rtol=1e-9, atol=1e-12)


def test_grid_only_run_and_pair_before_atoms():
    # no atoms: the field alone (FH + conjugate) against the oracle; a pair
    # style before any atom is the reference's error (src/input.cpp:103)
    n = 32
    phi = P.ic_noisy_constant((n, n, n), 0.5, 0.02, 77)
    runs = []
    for is_oracle in (False, True):
        s = (O.System if is_oracle else P.System)(n, n, n, [0, 0, 0], [float(n)] * 3)
        s.set_field(P.PHI, phi)
        with pytest.raises((RuntimeError, P.PhafdError), match="Cannot create pair style before atoms are made."):
            if is_oracle:
                s.add_pair(0, 3.0)
            else:
                s.pair_gridatom(0, 3.0)
        s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, P.NOISE_OFF)
        s.fix_floryhuggins(1.0, 1.0, 3.0)
        s.setup(1e-3)
        s.run(100)
        runs.append((s.get_field(P.PHI), s.get_field(P.MU)))
        s.close()
    assert rel_l2(runs[0][0], runs[1][0]) < 1e-10
    assert rel_l2(runs[0][1], runs[1][1]) < 1e-10


@pytest.mark.parametrize("zghost", [0, 1])
def test_atoms_on_and_beyond_the_box_faces(zghost):
    # Domain::pbc / image counters (src/domain.cpp:102-208): beads exactly on
    # the upper faces, on the lower ones, and one period outside (given
    # unwrapped; pbc shifts by one period, as the reference's does); wrapped
    # positions, images and the gathered atoms bitwise at set-up, then a
    # short trajectory against the oracle
    w = workload(n=32, nchains=2, nbeads=5, lj=False, chain=False, drag=True, seed=8, zghost=zghost)
    L = w["box"]
    rng = np.random.default_rng(8)
    base = 4.0 + 24.0 * rng.random((40, 3))
    base[0] = [L, 5.0, 6.0]
    base[1] = [7.0, L, 8.0]
    base[2] = [9.0, 10.0, L]
    base[3] = [0.0, 0.0, 0.0]
    base[4] = [L, L, L]
    shift = rng.integers(-1, 2, (40, 3)) * L
    shift[:5] = 0
    w["xyz"] = base + shift
    w["mol"] = np.zeros(40, np.int32)
    w["types"] = np.zeros(40, np.int32)
    w["starts"], w["ends"] = np.arange(40, dtype=np.int32), np.arange(1, 41, dtype=np.int32)
    g = make_gpu(w)
    o = make_oracle(w)
    try:
        for a, b in zip(g.get_atoms(), o.get_atoms()):
            assert np.array_equal(a, b)
        for a, b in zip(g.gathered(), o.gathered()):
            assert np.array_equal(a, b)
        assert _lists_equal(g.gridlist(), o.gridlist())
        g.run(20)
        o.run(20)
        assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
        assert rel_l2(g.get_atoms()[0], o.get_atoms()[0]) < 1e-12   # atom/drag moves x
        assert np.array_equal(g.get_atoms()[2], o.get_atoms()[2])
    finally:
        g.close()
        o.close()


def test_reference_noise_field_steps():
    w = workload(n=32, nchains=5, nbeads=20, seed=6, noise=P.NOISE_FIELD)
    g = make_gpu(w)
    w2 = dict(w)
    w2["noise"] = O.NOISE_MT19937
    o = make_oracle(w2)
    gen = P.NoiseGen(5678)
    for _ in range(3):
        g.set_noise(gen.next(32, 32, 32, [32.0] * 3, 1.0, 1.0, 1e-3))
        g.run(1)
        o.run(1)
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-12


def test_nan_is_reported():
    w = workload(n=16, nchains=2, nbeads=5, seed=8)
    w["phi"] = w["phi"].copy()
    w["phi"][0, 0, 0] = np.nan
    g = make_gpu(w)
    with pytest.raises(P.PhafdError) as e:
        g.run(1)
    assert e.value.code == -3


def test_list_free_deposit_matches_stored_list_bitwise():
    # the step's list-free mu deposit against the stored NPairGridBin list walk,
    # over several neighbour rebuilds
    w = workload(n=32, nchains=12, nbeads=30, seed=12, nucwidth=2.0, zghost=1)
    a = make_gpu(w)
    b = make_gpu(w)
    b.set_option(P.OPT_EAGER_GRIDLIST, 1)
    for _ in range(4):
        a.run(10)
        b.run(10)
        assert np.array_equal(a.get_field(P.MU), b.get_field(P.MU))
        assert np.array_equal(a.get_forces(), b.get_forces())
        assert np.array_equal(a.get_field(P.PHI), b.get_field(P.PHI))
    assert a.counts()["ncalls"] > 1
    assert _lists_equal(a.gridlist(), b.gridlist())


def test_linear_style_trajectory_and_stored_list():
    # gridatom/LJish/linear (the style of the reference's example script,
    # src/pair_gridatom_lj_ish_linear.cpp): 100 steps against the oracle with
    # rebuilds, and the default deposit against the stored-list walk bitwise
    # epsilon 0.2 keeps phi inside (0, 1) over the run (at 1.0 it crosses 1
    # near step 90, where log(phi / (1 - phi)) amplifies rounding)
    w = workload(n=32, nchains=12, nbeads=30, seed=12, ljish="linear", gauss_cut=2.5, nucwidth=2.5, zghost=1,
                 eps=0.2)
    g = make_gpu(w)
    b = make_gpu(w)
    b.set_option(P.OPT_EAGER_GRIDLIST, 1)
    o = make_oracle(w)
    for s in (g, b, o):
        s.run(100)
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
    assert rel_l2(g.get_field(P.MU), o.get_field(P.MU)) < 1e-10
    assert rel_l2(g.get_atoms()[0], o.get_atoms()[0]) < 1e-10
    assert rel_l2(g.get_forces(), o.get_forces()) < 1e-8
    assert g.counts()["ncalls"] == o.counts()["ncalls"] > 1
    assert np.allclose(g.energy(), o.energy(), rtol=1e-9, atol=1e-12)
    assert np.array_equal(g.get_field(P.MU), b.get_field(P.MU))
    assert np.array_equal(g.get_forces(), b.get_forces())
    assert g.get_field(P.MU).min() < -0.2   # the linear deposit, -e per in-cut bead


@pytest.mark.parametrize("lj", [True, False])
def test_harmonic_cut_one_step_and_trajectory(lj):
    # pair harmonic/cut (src/pair_harmonic_cut.cpp) beside lj/cut or alone:
    # the pair phase against the oracle, 100 steps with rebuilds, and the
    # pruned list (cut = the larger of the two bead cuts) against the full
    # NPairHalfBinNewton list bitwise
    w = workload(n=32, nchains=12, nbeads=30, seed=14, zghost=1, lj=lj, harmonic=(1.3, 40.0))
    g = make_gpu(w)
    o = make_oracle(w)
    for ph in range(P.PH_PAIR + 1):
        g.phase(ph)
        o.phase(ph)
    assert rel_l2(g.get_forces(), o.get_forces()) < 1e-10
    g = make_gpu(w)
    o = make_oracle(w)
    full = make_gpu(w)
    full.set_option(P.OPT_EAGER_HALFLIST, 1)
    for s in (g, o, full):
        s.run(100)
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
    assert rel_l2(g.get_atoms()[0], o.get_atoms()[0]) < 1e-10
    assert rel_l2(g.get_forces(), o.get_forces()) < 1e-8
    assert g.counts()["ncalls"] == o.counts()["ncalls"] > 1
    assert np.allclose(g.energy(), o.energy(), rtol=1e-9, atol=1e-12)
    assert np.array_equal(g.get_atoms()[0], full.get_atoms()[0])
    assert np.array_equal(g.get_forces(), full.get_forces())


@pytest.mark.parametrize("ljish", [False, "linear"])
def test_compute_pair_end_of_step(ljish):
    # ComputePair after a run (the end of its last step): the forces of each
    # pair style alone, from the step's gathered positions and the current
    # fields, against the oracle; the owned forces stay as they were and the
    # grid-atom style's deposit lands on mu once more
    w = workload(n=32, nchains=10, nbeads=25, seed=33, zghost=1, ljish=ljish, harmonic=(1.3, 40.0), eps=0.2)
    g = make_gpu(w)
    o = make_oracle(w)
    g.run(25)
    o.run(25)
    F = g.get_forces()
    for ip in range(3):   # grid-atom, lj/cut, harmonic/cut
        fg, fo = g.compute_pair(ip), o.compute_pair(ip)
        assert rel_l2(fg, fo) < 1e-10, ip
        assert np.abs(fg).max() > 0
    assert np.array_equal(g.get_forces(), F)
    assert rel_l2(g.get_field(P.MU), o.get_field(P.MU)) < 1e-10
    g.run(5)
    o.run(5)
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
    with pytest.raises(P.PhafdError):
        g.compute_pair(3)


def _fields_only(cls, shape, period, noise, seed=0):
    nz, ny, nx = shape
    s = cls(nx, ny, nz, [0.0, 0.0, 0.0], period)
    s.set_field(0, P.ic_noisy_constant(shape, 0.5, 0.05, 99 + seed))
    z = np.zeros(0, np.int32)
    s.set_atoms(np.zeros((0, 3)), z, tags=z, mols=z)
    s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, noise)
    s.fix_floryhuggins(1.0, 1.0, 3.0)
    s.setup(1e-3)
    return s


# grids whose z pass takes the fused kernels (Nz in {16, ..., 512}; 32, 128
# and 512 split unevenly, 8 x 4, 16 x 8 and 16 x 32); the anisotropic
# periods check that the wave-number scales go to the right axes
@pytest.mark.parametrize("shape,period", [((16, 64, 32), [32.0, 96.0, 12.0]),
                                          ((256, 16, 8), [8.0, 16.0, 256.0]),
                                          ((64, 64, 64), [64.0, 64.0, 64.0]),
                                          ((128, 32, 16), [16.0, 40.0, 128.0]),
                                          ((32, 128, 64), [64.0, 128.0, 32.0]),
                                          ((512, 16, 32), [32.0, 16.0, 512.0]),
                                          ((16, 512, 32), [32.0, 512.0, 16.0])])
def test_fused_spectral_path(shape, period):
    nz, ny, nx = shape
    fused = _fields_only(P.System, shape, period, P.NOISE_FIELD)
    plain = _fields_only(P.System, shape, period, P.NOISE_FIELD)
    plain.set_option(P.OPT_FUSE_SPECTRAL, 0)
    o = _fields_only(O.System, shape, period, O.NOISE_MT19937)
    gen = P.NoiseGen(5678)
    for _ in range(4):
        xi = gen.next(nx, ny, nz, period, 1.0, 1.0, 1e-3)
        fused.set_noise(xi)
        plain.set_noise(xi)
        fused.run(1)
        plain.run(1)
        o.run(1)
        # mu = (T/v)(ln phi - ln(1 - phi)) + chi (1 - 2 phi) is small at
        # phi ~ 0.5 while its slope is ~2: a 1e-13 change of phi moves it by
        # a few 1e-13 relative, so mu's bar is 1e-12 (measured 3.4e-13 on
        # the 32 x 128 x 64 grid); an indexing mistake would show at O(1)
        for f in (P.PHI, P.MU, P.GRADX, P.GRADY, P.GRADZ):
            assert rel_l2(fused.get_field(f), plain.get_field(f)) < (1e-12 if f == P.MU else 1e-13), f
        assert rel_l2(fused.get_spectrum(P.MUHAT), plain.get_spectrum(P.MUHAT)) < 1e-12
    for f in (P.PHI, P.MU, P.GRADX, P.GRADY, P.GRADZ):
        assert rel_l2(fused.get_field(f), o.get_field(f)) < (1e-10 if f == P.MU else 1e-12), f


def test_fused_spectral_probe_counts():
    # the fused kernels are the ones that run on a 64^3 grid
    w = workload(n=64, nchains=4, nbeads=10, seed=9)
    g = make_gpu(w)
    g.set_option(P.OPT_NEXT_GRADIENT, 0)
    g.probe("z_update")
    g.run(3)
    ms, launches, nbytes = g.probe(None)
    # per bulk mode: mu and phihat read and written (64 B); the first two
    # steps leave muhat unwritten (48 B), only the run's last step stores it
    assert launches == 3 and nbytes == (2 * 48.0 + 64.0) * 31 * 64 * 64
    g.probe("z_grad")
    g.run(2)
    assert g.probe(None)[1] == 2
    # with the next step's gradient made by the update, only a run's first
    # step transforms phi forward; its update also writes the z passes of
    # phi and of the z gradient (+32 B per mode) on every step but the last
    g.set_option(P.OPT_NEXT_GRADIENT, 1)
    g.probe("z_grad")
    g.run(3)
    assert g.probe(None)[1] == 1
    g.probe("z_update")
    g.run(3)
    ms, launches, nbytes = g.probe(None)
    assert launches == 3 and nbytes == (2 * 80.0 + 64.0) * 31 * 64 * 64


def test_fused_kernels_run_at_128():
    # 128 = 16 x 8 (an uneven split): the fused z update, the separable
    # gradient and the next step's gradient all run on a 128^3 grid
    w = workload(n=128, nchains=4, nbeads=10, seed=9)
    g = make_gpu(w)
    try:
        g.set_option(P.OPT_NEXT_GRADIENT, 0)
        g.probe("z_update")
        g.run(2)
        assert g.probe(None)[1] == 2
        g.probe("z_grad")
        g.run(2)
        assert g.probe(None)[1] == 2
        g.set_option(P.OPT_NEXT_GRADIENT, 1)
        g.probe("c2r_xd")
        g.run(3)
        assert g.probe(None)[1] == 2   # every step of a run but its last
    finally:
        g.close()


def test_pruned_lj_list_matches_full_list_bitwise():
    # the step's pruned lj/cut list against the full NPairHalfBinNewton list,
    # over several neighbour rebuilds; the lazily materialised full list
    # equals the eager one
    w = workload(n=32, nchains=12, nbeads=30, seed=13, zghost=1)
    a = make_gpu(w)
    b = make_gpu(w)
    b.set_option(P.OPT_EAGER_HALFLIST, 1)
    for _ in range(4):
        a.run(10)
        b.run(10)
        assert np.array_equal(a.get_forces(), b.get_forces())
        assert np.array_equal(a.get_atoms()[0], b.get_atoms()[0])
    assert a.counts()["ncalls"] > 1
    assert _lists_equal(a.halflist(), b.halflist())
    a.run(5)
    b.run(5)
    assert np.array_equal(a.get_atoms()[0], b.get_atoms()[0])


@pytest.mark.parametrize("ljish", [False, True, "linear"])
def test_dense_tile_mu_terms(ljish):
    # 400 atoms packed into a 5^3 corner (across the periodic boundary) with a
    # long grid-atom cut: the tiles of the mu deposit there stage more atoms
    # than fit in LDS at once and take them in several rounds
    w = workload(n=32, nchains=40, nbeads=10, seed=21, lj=False, chain=False, gauss_cut=3.0,
                 nucwidth=3.0, ljish=ljish, zghost=1)
    rng = np.random.default_rng(5)
    w["xyz"] = np.mod(rng.uniform(-2.5, 2.5, size=w["xyz"].shape), 32.0)
    a = make_gpu(w)
    b = make_gpu(w)
    b.set_option(P.OPT_EAGER_GRIDLIST, 1)
    o = make_oracle(w)
    for s in (a, b, o):
        s.run(2)
    assert np.array_equal(a.get_field(P.MU), b.get_field(P.MU))
    assert np.array_equal(a.get_forces(), b.get_forces())
    assert rel_l2(a.get_field(P.MU), o.get_field(P.MU)) < 1e-10
    assert rel_l2(a.get_forces(), o.get_forces()) < 1e-10
    assert rel_l2(a.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10


# ---------------------------------------------------------------- FP32 form
# FP32 FFT, fields and grid-atom terms (Kahan-compensated force folds and mu
# sums); atoms, lists, lj/cut and chains stay FP64, and so do the k = 0 mode
# and the energies.  Its oracle is the FP64 path (pinned above to the CPU
# oracle); the bounds below are the measured FP32 error with headroom.
def test_fp32_form_against_fp64():
    w = workload(n=64, nchains=20, nbeads=50, seed=42, zghost=1, chi=3.0)
    a = make_gpu(w, precision=P.FP32)
    b = make_gpu(w)
    phi0 = w["phi"].mean()
    a.run(1)
    b.run(1)
    m1 = a.get_field(P.PHI).mean()
    a.run(99)
    b.run(99)
    e_phi = rel_l2(a.get_field(P.PHI), b.get_field(P.PHI))
    e_mu = rel_l2(a.get_field(P.MU), b.get_field(P.MU))
    e_x = rel_l2(a.get_atoms()[0], b.get_atoms()[0])
    e_f = rel_l2(a.get_forces(), b.get_forces())
    ea, eb = np.array(a.energy()), np.array(b.energy())
    print(f"fp32 vs fp64 after 100 steps: phi {e_phi:.2e} mu {e_mu:.2e} x {e_x:.2e} F {e_f:.2e} "
          f"energies {np.abs(ea - eb) / np.maximum(np.abs(eb), 1e-30)}")
    assert e_phi < 1e-5
    assert e_mu < 1e-4
    assert e_x < 1e-6
    assert e_f < 1e-3
    assert np.allclose(ea, eb, rtol=1e-4, atol=1e-6)
    # mass: the k = 0 mode comes from the FP64 sum of phi, so the mean stays
    # at the FP32 value nearest the initial mean (half an ulp of 0.5 away at
    # most); what moves it is only the rounding of the c2r output around
    # that mode (measured 3.7e-9 over 99 steps), below one FP32 ulp
    m100 = a.get_field(P.PHI).mean()
    assert abs(m1 - phi0) < 3.0e-8
    assert abs(m100 - m1) < 2e-8
    assert a.counts()["ncalls"] == b.counts()["ncalls"]


@pytest.mark.parametrize("shape,period", [((16, 64, 32), [32.0, 96.0, 12.0]), ((32, 32, 32), [32.0, 32.0, 32.0]),
                                          ((512, 16, 16), [16.0, 16.0, 512.0])])
def test_fp32_fields_vs_oracle(shape, period):
    # fused and unfused FP32 spectral paths against the CPU oracle
    nz, ny, nx = shape
    o = _fields_only(O.System, shape, period, O.NOISE_OFF)
    a = _fields_only(lambda *args: P.System(*args, precision=P.FP32), shape, period, P.NOISE_OFF)
    o.run(20)
    a.run(20)
    errs = {f: rel_l2(a.get_field(f), o.get_field(f)) for f in (P.PHI, P.MU, P.GRADX, P.GRADY, P.GRADZ)}
    print(f"fp32 fields vs oracle {shape}: {errs}")
    assert errs[P.PHI] < 1e-5
    assert all(e < 2e-4 for e in errs.values())
