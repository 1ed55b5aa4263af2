"""The oracle's bead-rod polymers (fix atom/semiflexible/*,
src/fixatom_semiflexible.cpp:27-191) pinned by what the mathematics fixes.

The reference hands the dynamics to beadrodpmer, which is absent (no
version pinned, no tests or fixtures), so the oracle restates the published
bead-rod Brownian dynamics of Montesi, Morse & Pasquali (J. Chem. Phys. 122,
084903, 2005).  Pinned here: the per-polymer seeds (the reference's own
draws), Philox4x64-10 (Random123's known answer and numpy), the bond check
and its message, exact rigid translation with anisotropic friction, the
second-order rotation of a dimer, rod lengths after the Newton corrector,
T = 0 bending relaxation, tethers, the itermax failure, and the equilibrium
bond-angle statistics of free trimers (the Langevin function).  Parity with
beadrodpmer itself is unpinned.  CPU only.
"""
import numpy as np
import pytest

from oracle import oracle_ctypes as O


def _line(n, d=(1.0, 0.0, 0.0), b=1.0, o=(5.0, 6.0, 7.0)):
    d = np.asarray(d, float)
    d /= np.linalg.norm(d)
    return np.array([np.asarray(o) + i * b * d for i in range(n)])


def test_philox_known_answer_and_numpy():
    # Random123 kat_vectors: philox4x64-10, zero counter and key
    want = [0x16554d9eca36314c, 0xdb20fe9d672d0fdc, 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    assert [int(v) for v in O.philox4x64_10([0, 0, 0, 0], [0, 0])] == want
    # numpy's Philox advances its counter before each block
    key = np.array([7, 0x4245414452], dtype=np.uint64)
    g = np.random.Philox(key=key, counter=np.array([2, 5, 0, 0], dtype=np.uint64))
    assert [int(v) for v in g.random_raw(4)] == [int(v) for v in O.philox4x64_10([3, 5, 0, 0], key)]


def test_polymer_seeds_are_the_reference_draws():
    # uniform_int_distribution<int>(0, INT_MAX) on mt19937(make_unique_seed(seed))
    # (src/fixatom_semiflexible.cpp:44-80); libstdc++ (GCC >= 11) maps a full
    # 32-bit draw g to g >> 1 for that range
    xyz = _line(12)
    s = O.System(8, 8, 8, [0, 0, 0], [32.0] * 3)
    s.set_atoms(xyz, np.zeros(12, np.int32), tags=np.arange(12, dtype=np.int32), mols=np.zeros(12, np.int32))
    s.fix_semiflexible(0, [0, 4, 8], [4, 8, 12], 1234, 1.0, 1.0, 1.0, 0.0, 0.0)
    raw = O.mt19937_raw(O.make_unique_seed(1234), 3)
    assert np.array_equal(s.semiflexible_seeds(0, 3), (raw >> 1).astype(np.int32))
    s.close()


def test_bond_check_message():
    xyz = _line(5)
    xyz[3:] += [0.01, 0.0, 0.0]   # one bond 1e-2 too long: beyond 1e-3
    s = O.System(8, 8, 8, [0, 0, 0], [32.0] * 3)
    s.set_atoms(xyz, np.zeros(5, np.int32), tags=np.arange(5, dtype=np.int32), mols=np.zeros(5, np.int32))
    s.fix_semiflexible(0, [0], [5], 1, 1.0, 1.0, 1.0, 0.0, 0.0)
    with pytest.raises(RuntimeError, match="Adjacent atom displacements do not match the FixAtomSemiFlexible "
                                           "bondlength specified."):
        s.setup(1e-3)
    s.close()


def test_rigid_translation_is_exact():
    # a straight rod under one uniform force translates rigidly:
    # v = (f.d) d / zeta_para + (f - (f.d) d) / zeta_perp, no tension
    rng = np.random.default_rng(0)
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    R = _line(6, d, b=1.3)
    f = rng.standard_normal(3)
    zpar, zperp, dt = 0.4, 1.7, 1e-2
    R1, it = O.beadrod_step(R, np.tile(f, (6, 1)), b=1.3, zpar=zpar, zperp=zperp, dt=dt)
    v = (f @ d) * d / zpar + (f - (f @ d) * d) / zperp
    assert np.abs(R1 - R - dt * v).max() < 1e-14
    assert it == 0


def test_dimer_rotation_is_second_order():
    # opposite perpendicular forces on a dimer: angular velocity 2 f / (zeta_perp b);
    # one step of a second-order scheme errs by O(dt^3)
    R = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    F = np.array([[0, -1.0, 0], [0, 1.0, 0]])
    err = []
    for dt in (1e-2, 5e-3):
        R1, it = O.beadrod_step(R, F, zpar=0.5, zperp=2.0, dt=dt)
        assert abs(np.linalg.norm(R1[1] - R1[0]) - 1.0) < 1e-12
        ang = np.arctan2(R1[1, 1] - R1[0, 1], R1[1, 0] - R1[0, 0])
        err.append(abs(ang - 2 * dt / 2.0))
    assert 7.0 < err[0] / err[1] < 9.0


def test_bending_relaxes_at_zero_temperature_and_rods_keep_length():
    R = np.array([[0.0, 0, 0], [1.0, 0, 0], [1.0 + np.cos(1.0), np.sin(1.0), 0], [0.0, 0.0, 0.0]])
    d = np.array([0.5 - R[2][0], 1.5 - R[2][1], 0.3])
    R[3] = R[2] + d / np.linalg.norm(d)   # three unit bonds
    E = []
    for k in range(300):
        R, it = O.beadrod_step(R, np.zeros((4, 3)), bending=2.0, zpar=0.5, zperp=1.0, dt=1e-2, step=k + 1)
        u = np.diff(R, axis=0)
        assert np.abs(np.linalg.norm(u, axis=1) - 1.0).max() < 1e-12
        un = u / np.linalg.norm(u, axis=1)[:, None]
        E.append(2.0 * np.sum(1.0 - np.sum(un[:-1] * un[1:], axis=1)))
    assert all(np.diff(E) <= 1e-14)
    assert E[-1] < 1e-6 * E[0]


@pytest.mark.parametrize("tether", [1, 2])
def test_tethered_beads_stay(tether):
    rng = np.random.default_rng(3)
    R = np.cumsum(np.vstack([[0, 0, 0], rng.standard_normal((5, 3))]), axis=0)
    d = np.diff(R, axis=0)
    R = np.vstack([[0, 0, 0], np.cumsum(d / np.linalg.norm(d, axis=1)[:, None], axis=0)])
    R0 = R.copy()
    F = rng.standard_normal((6, 3))
    for k in range(20):
        R, it = O.beadrod_step(R, F, tether=tether, temp=1.0, bending=1.0, dt=1e-3, step=k + 1)
    assert np.array_equal(R[0], R0[0])
    if tether == 2:
        assert np.array_equal(R[-1], R0[-1])
    assert np.abs(np.linalg.norm(np.diff(R, axis=0), axis=1) - 1.0).max() < 1e-12
    assert np.abs(R[1:-1] - R0[1:-1]).max() > 1e-3


def test_itermax_failure_is_reported():
    # a step far too long for the corrector: more than itermax iterations
    R = _line(4)
    F = np.array([[0, 50.0, 0], [0, -50.0, 0], [0, 50.0, 0], [0, -50.0, 0]])
    _, it = O.beadrod_step(R, F, dt=0.5, itermax=3)
    assert it == 4


def _trimer_stats(kb, zpar, zperp, nch=400, nsteps=6000, dt=2.5e-3):
    xyz = []
    side = int(np.ceil(nch ** (1 / 3)))
    for c in range(nch):
        o = np.array([c % side, (c // side) % side, c // side ** 2]) * 5.0 + 1.0
        xyz += [o + [i, 0, 0] for i in range(3)]
    xyz = np.array(xyz)
    s = O.System(8, 8, 8, [0, 0, 0], [400.0] * 3)
    n = len(xyz)
    s.set_atoms(xyz, np.zeros(n, np.int32), tags=np.arange(n, dtype=np.int32), mols=np.zeros(n, np.int32))
    st = np.arange(nch) * 3
    s.fix_semiflexible(0, st, st + 3, 11, 1.0, zpar, zperp, kb, 1.0)
    s.setup(dt)
    s.run(nsteps // 10)
    c1, c2 = [], []
    every = int(0.2 / dt)
    for _ in range((nsteps - nsteps // 10) // every):
        s.run(every)
        b = np.diff(s.get_atoms()[1].reshape(nch, 3, 3), axis=1)
        c = np.sum(b[:, 0] * b[:, 1], axis=1)
        c1.append(c.mean())
        c2.append((c * c).mean())
    s.close()
    return float(np.mean(c1)), float(np.mean(c2))


def test_free_trimer_bond_angles_follow_boltzmann():
    # with the metric pseudo-potential, bond orientations are independent
    # with weight exp(kb c / kT): <c> = coth(kb) - 1/kb (0 for kb = 0) and
    # <c^2> = 1 - 2 coth(kb)/kb + 2/kb^2 (1/3 for kb = 0); the rigid chain
    # without it would give <c^2> = 0.3217 at kb = 0 (weight sqrt(1 - c^2/4))
    m, m2 = _trimer_stats(0.0, 0.5, 1.0)
    assert abs(m) < 0.015 and abs(m2 - 1 / 3) < 0.008, (m, m2)
    m, _ = _trimer_stats(2.0, 0.5, 1.0)
    assert abs(m - (1 / np.tanh(2.0) - 0.5)) < 0.02, m
