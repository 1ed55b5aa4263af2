"""Pins of the oracle's FFT and grad(phi) (CPU only).

r2c / c2r are checked against numpy rfftn / irfftn golden vectors (a library
routine; non-Hermitian c2r inputs included) and against a brute-force DFT;
grad(phi) against closed forms of single modes on an anisotropic grid, which
fails if the y/z swap of the reference's transposed layout (src/grid.cpp:269-280,
src/domain.cpp:206-208) or the 3.141592 wavenumber is lost.
"""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle_ctypes as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
SHAPES = sorted({k.split("_", 2)[2] for k in GOLD.files})


@pytest.mark.parametrize("tag", SHAPES)
def test_r2c_golden(tag):
    a = GOLD[f"r2c_in_{tag}"]
    ref = GOLD[f"r2c_out_{tag}"]
    out = O.fft_r2c(a)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("tag", SHAPES)
def test_c2r_golden_non_hermitian(tag):
    c = GOLD[f"c2r_in_{tag}"]
    ref = GOLD[f"c2r_out_{tag}"]
    nx = int(tag.split("x")[2])
    out = O.fft_c2r(c, nx)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def brute_dft_r2c(a):
    nz, ny, nx = a.shape
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=complex)
    for kz, ky, kx in itertools.product(range(nz), range(ny), range(nx // 2 + 1)):
        s = 0j
        for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
            s += a[z, y, x] * np.exp(-2j * np.pi * (kz * z / nz + ky * y / ny + kx * x / nx))
        out[kz, ky, kx] = s
    return out


@pytest.mark.parametrize("shape", [(4, 4, 4), (2, 4, 8)])
def test_r2c_brute_force(shape):
    a = np.random.default_rng(3).standard_normal(shape)
    assert np.allclose(O.fft_r2c(a), brute_dft_r2c(a), atol=1e-12)


def test_roundtrip():
    a = np.random.default_rng(4).standard_normal((16, 8, 32))
    b = O.fft_c2r(O.fft_r2c(a), 32) / a.size
    assert np.abs(a - b).max() < 1e-14


def test_gradphi_single_modes_anisotropic():
    nx, ny, nz = 16, 8, 32
    L = np.array([16.0, 12.0, 40.0])
    x = np.arange(nx) * L[0] / nx
    y = np.arange(ny) * L[1] / ny
    z = np.arange(nz) * L[2] / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    a, b, c = 2, 3, 5
    phi = (np.sin(2 * np.pi * a * X / L[0]) + 0.5 * np.cos(2 * np.pi * b * Y / L[1])
           + 0.25 * np.sin(2 * np.pi * c * Z / L[2]))
    gx, gy, gz = O.gradphi(phi, L)
    q = lambda m, Ld: 2 * 3.141592 * m / Ld   # truncated pi of Domain::dq*
    assert np.abs(gx - q(a, L[0]) * np.cos(2 * np.pi * a * X / L[0])).max() < 1e-12
    assert np.abs(gy + 0.5 * q(b, L[1]) * np.sin(2 * np.pi * b * Y / L[1])).max() < 1e-12
    assert np.abs(gz - 0.25 * q(c, L[2]) * np.cos(2 * np.pi * c * Z / L[2])).max() < 1e-12
    # with the exact pi the result would differ at ~2e-7 relative
    assert np.abs(gx - 2 * np.pi * a / L[0] * np.cos(2 * np.pi * a * X / L[0])).max() > 1e-8


def test_gradphi_nyquist_modes_vanish():
    # the Nyquist cosine along each axis has zero derivative at grid points;
    # keeping Im() at kx = 0 / Nx/2 in the c2r would break this
    n = 8
    idx = np.arange(n)
    Z, Y, X = np.meshgrid(idx, idx, idx, indexing="ij")
    phi = np.cos(np.pi * X) + np.cos(np.pi * Y) + np.cos(np.pi * Z) + 0.3
    g = O.gradphi(phi, [8.0, 8.0, 8.0])
    for gi in g:
        assert np.abs(gi).max() < 1e-13
