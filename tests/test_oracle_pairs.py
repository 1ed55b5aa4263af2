"""Pins of the oracle's pair styles and chain fix (CPU only)."""
import numpy as np
import pytest

from oracle import oracle_ctypes as O


def one_atom(pos, phi, style=O.PAIR_GAUSSIAN, cut=3.0, nucwidth=1.5, n=8, L=8.0):
    s = O.System(n, n, n, [0, 0, 0], [L] * 3)
    s.set_field(O.PHI, phi)
    s.set_atoms(np.array([pos]), np.zeros(1, np.int32))
    p = s.add_pair(style, cut)
    s.pair_coeff(p, 0, 2.0, 0.1, nucwidth)
    s.fix_gradphi(False)
    s.neighbor(0.5)
    s.setup(1e-3)
    s.phase(O.PH_START)
    s.phase(O.PH_NEIGH)
    s.phase(O.PH_PAIR)
    return s


def brute(pos, phi, grads, style, w, eps=2.0, phit=0.1, n=8, L=8.0):
    g = np.arange(n) * L / n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    mu = np.zeros_like(phi)
    F = np.zeros(3)
    # the central copy plus its periodic images in x and y (no z images on one rank)
    for sx in (-L, 0, L):
        for sy in (-L, 0, L):
            p = np.array(pos) + [sx, sy, 0]
            r2 = (X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2
            s = r2 / w ** 2
            if style == O.PAIR_GAUSSIAN:
                m = r2 <= w ** 2
                e = (np.exp(-s / 2) - np.exp(-0.5)) * np.exp(0.5) * eps
            else:
                m = r2 < w ** 2
                with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
                    e = np.exp(-1.0 / (1 - s) + 1) * eps
            e = np.where(m, e, 0.0)
            # the linear style (pair_gridatom_lj_ish_linear.cpp:46-179) deposits
            # -e itself and weights the force by it; the others by 2 (phi - phit) e
            wgt = -e if style == O.PAIR_LJISH_LINEAR else 2 * (phi - phit) * e
            mu += wgt
            for c in range(3):
                F[c] -= (grads[c] * wgt * (L / n) ** 3).sum()
    return mu, F


def test_gridatom_deposit_and_force_bruteforce():
    n, L = 8, 8.0
    g = np.arange(n) * L / n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    phi = 0.4 + 0.1 * np.sin(2 * np.pi * X / L) + 0.05 * np.cos(2 * np.pi * Y / L)
    for style in (O.PAIR_GAUSSIAN, O.PAIR_LJISH, O.PAIR_LJISH_LINEAR):
        s = one_atom([0.7, 7.6, 3.3], phi, style)
        grads = [s.get_field(O.GRADX), s.get_field(O.GRADY), s.get_field(O.GRADZ)]
        mu_ref, F_ref = brute([0.7, 7.6, 3.3], phi, grads, style, 1.5)
        assert np.allclose(s.get_field(O.MU), mu_ref, rtol=1e-12, atol=1e-14)
        assert np.allclose(s.get_forces()[0], F_ref, rtol=1e-10, atol=1e-14)


def test_gridatom_zero_force_for_constant_phi():
    s = one_atom([3.3, 4.1, 2.2], np.full((8, 8, 8), 0.37))
    assert np.abs(s.get_forces()).max() == 0.0
    assert s.get_field(O.MU).max() > 0
    s = one_atom([3.3, 4.1, 2.2], np.full((8, 8, 8), 0.37), O.PAIR_LJISH_LINEAR)
    assert np.abs(s.get_forces()).max() == 0.0
    assert s.get_field(O.MU).min() < 0 and s.get_field(O.MU).max() == 0.0


def test_linear_style_deposit_does_not_depend_on_phi():
    # mu of gridatom/LJish/linear is -e at every in-cut point whatever phi is
    # (no (phi - phi_t) factor), and its peak is -eps at the atom's grid point
    rng = np.random.default_rng(5)
    a = one_atom([3.0, 4.0, 2.0], 0.4 + 0.1 * rng.random((8, 8, 8)), O.PAIR_LJISH_LINEAR)
    b = one_atom([3.0, 4.0, 2.0], np.full((8, 8, 8), 0.9), O.PAIR_LJISH_LINEAR)
    assert np.array_equal(a.get_field(O.MU), b.get_field(O.MU))
    assert a.get_field(O.MU)[2, 4, 3] == -2.0   # exp(1 - 1/(1 - 0)) eps = eps = 2


def two_lj(p0, p1, L=16.0, pairs=((O.PAIR_LJCUT, 1.5, 1.0, 1.0),)):
    s = O.System(16, 16, 16, [0, 0, 0], [L] * 3)
    s.set_field(O.PHI, np.full((16, 16, 16), 0.5))
    s.set_atoms(np.array([p0, p1]), np.zeros(2, np.int32))
    for pr in pairs:
        s.add_pair(*pr)
    s.neighbor(0.5)
    s.setup(1e-3)
    s.phase(O.PH_START)
    s.phase(O.PH_NEIGH)
    s.phase(O.PH_PAIR)
    return s.get_forces()


def lj_force(d):
    r = np.linalg.norm(d)
    return 24 * (2 * r ** -12 - r ** -6) / r * (d / r)


def test_lj_two_body_and_newton():
    F = two_lj([5.0, 5.0, 5.0], [6.1, 5.2, 4.9])
    d = np.array([5.0, 5.0, 5.0]) - np.array([6.1, 5.2, 4.9])
    assert np.allclose(F[0], lj_force(d), rtol=1e-13)
    assert np.allclose(F[0], -F[1], rtol=0, atol=0)


def test_lj_across_x_boundary_and_not_z():
    F = two_lj([0.3, 5.0, 5.0], [15.5, 5.0, 5.0])    # x images interact
    assert np.allclose(F[0], lj_force(np.array([0.8, 0.0, 0.0])), rtol=1e-12)
    assert np.allclose(F[0], -F[1], atol=1e-15)
    F = two_lj([5.0, 5.0, 0.3], [5.0, 5.0, 15.5])    # no z images on one rank
    assert np.abs(F).max() == 0.0


HARM = (O.PAIR_HARMONIC, 1.3, 40.0)


def harmonic_force(d, cut=1.3, k=40.0):
    # U = k/2 (cut - r)^2 for r < cut, so F_i = -dU/dx_i = k (cut - r) d / r
    r = np.linalg.norm(d)
    return k * (cut - r) * d / r if r < cut else np.zeros(3)


def test_harmonic_two_body_and_newton():
    for p1 in ([6.1, 5.2, 4.9], [5.3, 5.0, 5.4], [5.0, 6.29, 5.0]):
        F = two_lj([5.0, 5.0, 5.0], p1, pairs=(HARM,))
        d = np.array([5.0, 5.0, 5.0]) - np.array(p1)
        assert np.allclose(F[0], harmonic_force(d), rtol=1e-13, atol=0)
        assert np.array_equal(F[0], -F[1])
    # the force is repulsive: atom 0 is pushed away from atom 1
    F = two_lj([5.0, 5.0, 5.0], [5.5, 5.0, 5.0], pairs=(HARM,))
    assert F[0][0] < 0 and F[0][1] == 0 and F[0][2] == 0


def test_harmonic_force_is_minus_energy_gradient():
    p0, p1 = np.array([5.0, 5.0, 5.0]), np.array([5.7, 5.4, 4.8])
    F = two_lj(p0, p1, pairs=(HARM,))

    def energy(x0):
        r = np.linalg.norm(x0 - p1)
        return 0.5 * 40.0 * (1.3 - r) ** 2 if r < 1.3 else 0.0
    h = 1e-6
    g = np.array([(energy(p0 + h * e) - energy(p0 - h * e)) / (2 * h) for e in np.eye(3)])
    assert np.allclose(F[0], -g, rtol=1e-7)


def test_harmonic_cut_is_strict_and_coincident_atoms_separate():
    F = two_lj([5.0, 5.0, 5.0], [6.5, 5.0, 5.0], pairs=((O.PAIR_HARMONIC, 1.5, 40.0),))
    assert np.abs(F).max() == 0.0                      # r == cut: no force
    F = two_lj([5.0, 5.0, 5.0], [6.6, 5.0, 5.0], pairs=(HARM,))
    assert np.abs(F).max() == 0.0
    # coincident atoms: pushed apart along (1, 1, 1) as if 0.001 apart on each axis
    F = two_lj([5.0, 5.0, 5.0], [5.0, 5.0, 5.0], pairs=(HARM,))
    assert np.allclose(F[0], harmonic_force(np.full(3, 0.001)), rtol=1e-13)
    assert np.array_equal(F[0], -F[1])


def test_harmonic_across_x_boundary_and_sum_with_lj():
    F = two_lj([0.3, 5.0, 5.0], [15.5, 5.0, 5.0], pairs=(HARM,))
    assert np.allclose(F[0], harmonic_force(np.array([0.8, 0.0, 0.0])), rtol=1e-12)
    F = two_lj([5.0, 5.0, 0.3], [5.0, 5.0, 15.5], pairs=(HARM,))   # no z images on one rank
    assert np.abs(F).max() == 0.0
    both = two_lj([5.0, 5.0, 5.0], [6.1, 5.2, 4.9], pairs=((O.PAIR_LJCUT, 1.5, 1.0, 1.0), HARM))
    d = np.array([5.0, 5.0, 5.0]) - np.array([6.1, 5.2, 4.9])
    assert np.allclose(both[0], lj_force(d) + harmonic_force(d), rtol=1e-13)


def test_chain_force_is_minus_energy_gradient():
    rng = np.random.default_rng(2)
    nb = 7
    xyz = np.cumsum(rng.normal(size=(nb, 3)) * 0.6, 0) + 8.0
    k, r0, kb = 100.0, 1.0, 10.0

    def energy(u):
        s = O.System(8, 8, 8, [0, 0, 0], [16.0] * 3)
        s.set_field(O.PHI, np.full((8, 8, 8), 0.5))
        s.set_atoms(u, np.zeros(nb, np.int32))
        s.fix_chain([0], [nb], r0, k, kb, 1.0)
        s.setup(1.0)
        return s.energy()[3]

    s = O.System(8, 8, 8, [0, 0, 0], [16.0] * 3)
    s.set_field(O.PHI, np.full((8, 8, 8), 0.5))
    s.set_atoms(xyz, np.zeros(nb, np.int32))
    s.fix_chain([0], [nb], r0, k, kb, 1.0)
    s.setup(1.0)     # F = 0, so one half step moves u by drag * Fchain * dt / 2
    s.phase(O.PH_INITIAL)
    _, u1, _ = s.get_atoms()
    fchain = (u1 - xyz) / 0.5
    h = 1e-6
    for b in range(nb):
        for c in range(3):
            up = xyz.copy(); up[b, c] += h
            dn = xyz.copy(); dn[b, c] -= h
            g = (energy(up) - energy(dn)) / (2 * h)
            assert abs(fchain[b, c] + g) < 1e-5 * max(1.0, abs(g))


def test_compute_pair_splits_the_pair_forces():
    # ComputePair (src/compute_pair.cpp:43-85) right after the pair phase sees
    # the phase's inputs: the per-style forces add up to the phase's force
    # (superposition), the owned forces are left untouched, and the grid-atom
    # style's second compute deposits onto mu once more
    from tests.helpers import workload, make_oracle
    w = workload(n=32, nchains=8, nbeads=20, seed=31, harmonic=(1.3, 40.0))
    o = make_oracle(w)
    o.run(5)
    for ph in (O.PH_START, O.PH_NEIGH, O.PH_PAIR):
        o.phase(ph)
    F, mu = o.get_forces(), o.get_field(O.MU)
    parts = [o.compute_pair(i) for i in range(3)]   # gridatom, lj/cut, harmonic/cut
    assert np.array_equal(o.get_forces(), F)
    assert np.allclose(sum(parts), F, rtol=0, atol=1e-13 * np.abs(F).max())
    assert all(np.abs(p).max() > 0 for p in parts)
    assert np.allclose(o.get_field(O.MU), 2 * mu, rtol=1e-13, atol=1e-15)
    with pytest.raises(RuntimeError):
        o.compute_pair(3)


def test_compute_pair_two_body_closed_form():
    s = O.System(16, 16, 16, [0, 0, 0], [16.0] * 3)
    s.set_field(O.PHI, np.full((16, 16, 16), 0.5))
    s.set_atoms(np.array([[5.0, 5.0, 5.0], [6.1, 5.2, 4.9]]), np.zeros(2, np.int32))
    s.add_pair(O.PAIR_LJCUT, 1.5, 1.0, 1.0)
    s.add_pair(*HARM)
    s.neighbor(0.5)
    s.setup(1e-3)
    for ph in (O.PH_START, O.PH_NEIGH, O.PH_PAIR):
        s.phase(ph)
    d = np.array([5.0, 5.0, 5.0]) - np.array([6.1, 5.2, 4.9])
    assert np.allclose(s.compute_pair(0)[0], lj_force(d), rtol=1e-13)
    assert np.allclose(s.compute_pair(1)[0], harmonic_force(d), rtol=1e-13)
