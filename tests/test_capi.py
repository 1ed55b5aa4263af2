"""The C-ABI library loads and exports every declared entry point, and the
product's host-side input generators agree with the oracle (CPU only)."""
import numpy as np

import phafd_amd as P
from oracle import oracle_ctypes as O


def test_library_exports_every_declared_symbol():
    L = P.lib()
    names = P.exported_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n


def test_missing_gpu_fails_loudly_not_silently():
    import phafd_amd.api as A
    # without a GPU the create call must return an error code, never fall back
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    try:
        A.System(8, 8, 8, [0, 0, 0], [8.0] * 3)
    except A.PhafdError as e:
        assert e.code in (-1, -2)
    else:
        raise AssertionError("System() succeeded without a GPU")


def test_noisy_constant_matches_reference_rng():
    seed = 1234
    got = P.ic_noisy_constant((4, 4, 4), 0.5, 0.02, seed).ravel()
    u = O.mt19937_uniform(O.make_unique_seed(seed), 64)
    assert np.array_equal(got, 0.5 + 0.02 * u)


def test_host_noise_generator_matches_oracle_two_steps():
    nx, ny, nz = 8, 8, 4
    L = [8.0, 8.0, 4.0]
    g = P.NoiseGen(5678)
    a1 = g.next(nx, ny, nz, L, 1.0, 1.0, 1e-3)
    a2 = g.next(nx, ny, nz, L, 1.0, 1.0, 1e-3)
    # the oracle in mt19937 mode draws the same stream; compare via its
    # per-mode record of one step (fresh generator -> step 1)
    ref1 = O.noise_field(nx, ny, nz, L, O.make_unique_seed(5678), 1.0, 1.0, 1e-3)
    assert np.array_equal(a1, ref1)
    assert not np.array_equal(a1, a2)


def test_random_chains_deterministic_bonds():
    x1, m1 = P.ic_random_chains(3, 10, 1.0, 0.9, [0, 0, 0], [20.0] * 3, 42)
    x2, _ = P.ic_random_chains(3, 10, 1.0, 0.9, [0, 0, 0], [20.0] * 3, 42)
    assert np.array_equal(x1, x2)
    b = np.linalg.norm(np.diff(x1.reshape(3, 10, 3), axis=1), axis=-1)
    assert np.allclose(b, 1.0)
    assert list(m1[::10]) == [0, 1, 2]
