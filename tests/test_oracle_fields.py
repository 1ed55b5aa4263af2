"""Pins of the oracle's Flory-Huggins and spectral update (CPU only)."""
import numpy as np

from oracle import oracle_ctypes as O


def test_floryhuggins_closed_form_and_reflection():
    phi = np.array([0.3, 0.5, 0.9, -0.1, 1.2, 0.7])
    mu = np.array([0.1, -0.2, 0.0, 0.5, 0.0, 1.0])
    T, v, chi = 1.3, 0.7, 2.5
    p2, m2, nl, ng = O.floryhuggins(phi, mu, T, v, chi)
    refl = np.array([0.3, 0.5, 0.9, 0.1, 0.8, 0.7])
    assert np.array_equal(p2, refl)
    expect = mu + T / v * (np.log(refl / (1 - refl)) + chi * (1 - 2 * refl))
    assert np.allclose(m2, expect, rtol=1e-15, atol=1e-15)
    assert (nl, ng) == (1, 1)


def _mesh(n, L):
    x = [np.arange(n[d]) * L[d] / n[d] for d in range(3)]
    return np.meshgrid(x[2], x[1], x[0], indexing="ij")   # Z, Y, X


def test_update_single_mode_recurrence():
    # linearised Model B: with mu = c*phi for one Fourier mode, one explicit
    # Euler step multiplies its amplitude by 1 - M q^2 dt (c + T/v gamma q^2)
    n = (16, 8, 32)          # nx, ny, nz
    L = (16.0, 12.0, 40.0)
    Z, Y, X = _mesh(n, L)
    a, b, c_ = 2, 1, 3
    k = 2 * np.pi * (a * X / L[0] + b * Y / L[1] + c_ * Z / L[2])
    phi = 0.4 + 0.01 * np.cos(k)
    cmu = 1.7
    mu = cmu * (phi - 0.4)
    M, T, v, g, dt = 1.1, 0.9, 1.2, 0.8, 1e-3
    ph = O.fft_r2c(phi)
    mh = O.fft_r2c(mu)
    ph2 = O.spectral_update(ph, mh, n[0], L, M, T, v, g, dt)
    new = O.fft_c2r(ph2, n[0])
    qr = lambda m, Ld: 2 * 3.141592 * m / Ld
    q2 = qr(a, L[0]) ** 2 + qr(b, L[1]) ** 2 + qr(c_, L[2]) ** 2
    fac = 1 - M * q2 * dt * (cmu + T / v * g * q2)
    assert np.abs(new - (0.4 + 0.01 * fac * np.cos(k))).max() < 1e-14


def test_update_conserves_mass_exactly_and_enforces_hermitian_planes():
    rng = np.random.default_rng(11)
    nx, ny, nz = 8, 8, 16
    hx = nx // 2 + 1
    ph = rng.standard_normal((nz, ny, hx)) + 1j * rng.standard_normal((nz, ny, hx))
    mh = rng.standard_normal((nz, ny, hx)) + 1j * rng.standard_normal((nz, ny, hx))
    out = O.spectral_update(ph, mh, nx, [8.0, 8.0, 16.0], 1.0, 1.0, 1.0, 1.0, 1e-3)
    N = nx * ny * nz
    assert out[0, 0, 0] == ph[0, 0, 0] * (1.0 / N)   # origin_update only rescales
    for kx in (0, nx // 2):
        P = out[:, :, kx]
        mirror = np.conj(P[(-np.arange(nz)) % nz][:, (-np.arange(ny)) % ny])
        # every non-self-conjugate mode equals the conj of its mirror exactly
        self_mask = np.zeros((nz, ny), bool)
        for j in (0, nz // 2):
            for i in (0, ny // 2):
                self_mask[j, i] = True
        assert np.array_equal(P[~self_mask], mirror[~self_mask])


def test_noise_draw_order_and_count():
    nx, ny, nz = 8, 4, 4
    L = [8.0, 4.0, 4.0]
    seed = 99
    xi = O.noise_field(nx, ny, nz, L, seed, 1.0, 1.0, 1e-3)
    hx = nx // 2 + 1
    # bulk modes draw first, two uniforms each, in (ky, kz, kx) loop order
    u = O.mt19937_uniform(seed, 2 * ny * nz * (hx - 2) + 64)
    pre = np.sqrt(12 * 1.0 * 1.0 / np.prod(L)) * np.sqrt(1e-3)
    q = lambda i, n, Ld: 2 * 3.141592 * (i if i <= n // 2 else i - n) / Ld
    w = 0
    for ky in range(ny):
        for kz in range(nz):
            for kx in range(1, hx - 1):
                qq = np.sqrt(q(kx, nx, L[0]) ** 2 + q(kz, nz, L[2]) ** 2 + q(ky, ny, L[1]) ** 2)
                assert np.isclose(xi[kz, ky, kx].real, pre * qq * u[w], rtol=1e-14, atol=0)
                assert np.isclose(xi[kz, ky, kx].imag, pre * qq * u[w + 1], rtol=1e-14, atol=0)
                w += 2
    # mirrors and the origin consume no draw
    assert xi[0, 0, 0] == 0
    nonzero_planes = sum(int(np.count_nonzero(xi[:, :, k])) for k in (0, hx - 1))
    # per plane: (ny*nz - 4)/2 complex modes + 4 self-conjugate (3 real draws
    # for kx = 0, since the origin has none; 4 for kx = Nx/2)
    assert nonzero_planes == 2 * ((ny * nz - 4) // 2) + 7


def test_mt19937_standard_kat():
    # [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937 (seed 5489) produces 4123659995
    r = O.mt19937_raw(5489, 10000)
    assert int(r[-1]) == 4123659995
    assert list(r[:2]) == [3499211612, 581869302]
    # uniform_real_distribution(-0.5, 0.5) = (lo + hi*2^32)/2^64 - 0.5 (libstdc++)
    u = O.mt19937_uniform(5489, 1)[0]
    assert u == (3499211612 + 581869302 * 2.0 ** 32) / 2.0 ** 64 - 0.5


def _modes(shape, amps):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    (a1, n1), (a2, n2), (a3, n3) = amps
    return (0.5 + a1 * np.cos(2 * np.pi * n1 * x / nx) + a2 * np.sin(2 * np.pi * n2 * y / ny)
            + a3 * np.cos(2 * np.pi * n3 * z / nz))


def test_semi_implicit_and_explicit_closed_form_decay():
    # mu = 0 (no Flory-Huggins, no atoms): K steps scale each Fourier mode by
    # (1 - c s)^K (explicit, src/conjugate_volfrac.cpp:106-109) or
    # (1 + c s)^-K (semi-implicit, src/conjugate.cpp:356), c s = M q^2 dt T/v
    # gamma q^2 with q = 2*3.141592 n / L; a dropped factor, the wrong axis
    # scale or the wrong sign fails the 1e-12 bar
    shape, L = (16, 8, 32), [32.0, 8.0, 16.0]
    nz, ny, nx = shape
    amps = [(0.01, 3), (0.02, 1), (0.015, 2)]
    Mb, T, v, gam = 1.3, 0.9, 1.1, 0.8
    for semi, dt, K in ((False, 1e-3, 50), (True, 0.1, 40)):
        s = O.System(nx, ny, nz, [0.0] * 3, L)
        s.set_field(0, _modes(shape, amps))
        z = np.zeros(0, np.int32)
        s.set_atoms(np.zeros((0, 3)), z, tags=z, mols=z)
        s.fix_conjugate(5678, Mb, T, v, gam, O.NOISE_OFF)
        s.setup(dt)
        s.set_semi_implicit(semi)
        s.run(K)
        fac = []
        for (a, n), Lax in zip(amps, L):
            q = 2 * 3.141592 / Lax * n
            cs = Mb * q * q * dt * (T / v * gam * q * q)
            fac.append((a * (1.0 / (1.0 + cs) if semi else 1.0 - cs) ** K, n))
        got = s.get_field(0)
        assert np.max(np.abs(got - _modes(shape, fac))) < 1e-12
        assert np.max(np.abs(got - _modes(shape, amps))) > 1e-5
