"""Pins of the oracle's compute complex and of the grid/ave schedule (CPU).

compute complex is checked against numpy: the spectrum at the in_fourier
hook is numpy's rfftn of phi at the start of the step (of mu after the
Flory-Huggins post_force for ft_chempot), laid out in the reference's ft
array order [ky][kz][kx], with numpy's own modulus and norm."""
import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phafd_amd.computes import save_steps


def _numpy_quant(spec, quant, prefac):
    t = np.transpose(spec, (1, 0, 2))   # (kz, ky, kx) -> (ky, kz, kx)
    return {0: t.real * prefac, 1: t.imag * prefac, 2: np.abs(t) * prefac, 3: np.abs(t) ** 2 * prefac ** 2}[quant]


@pytest.mark.parametrize("shape,box", [((16, 8, 32), (12.0, 8.0, 20.0)), ((8, 8, 8), (8.0, 8.0, 8.0))])
def test_compute_complex_against_numpy_rfftn(shape, box):
    nz, ny, nx = shape
    s = O.System(nx, ny, nz, [0, 0, 0], list(box))
    phi0 = 0.5 + 0.05 * np.random.default_rng(41).standard_normal(shape)
    s.set_field(O.PHI, phi0)
    z = np.zeros(0, np.int32)
    s.set_atoms(np.zeros((0, 3)), z, tags=z, mols=z)
    s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, O.NOISE_OFF)
    s.fix_floryhuggins(1.0, 1.0, 3.0)
    s.setup(1e-3)
    s.run(2)
    phi = s.get_field(O.PHI)
    for ph in (O.PH_START, O.PH_NEIGH, O.PH_PAIR, O.PH_INITIAL, O.PH_POSTFORCE, O.PH_PREFINAL):
        s.phase(ph)
    mu = s.get_field(O.MU)
    prefac = 1.0 / np.prod([b / n for b, n in zip(box, (nx, ny, nz))])
    for which, field in ((0, phi), (1, mu)):
        ref_spec = np.fft.rfftn(field)
        for quant in range(4):
            got = s.compute_complex(which, quant)
            ref = _numpy_quant(ref_spec, quant, prefac)
            assert got.shape == (ny, nz, nx // 2 + 1)
            assert np.allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max()), (which, quant)


def test_grid_ave_schedule():
    # every 2, repeat 3, freq 10: samples 6 8 10, 16 18 20, ...
    assert save_steps(2, 3, 10, 2) == [6, 8, 10, 16, 18, 20]
    assert save_steps(5, 1, 5, 3) == [5, 10, 15]
    assert save_steps(1, 4, 4, 1) == [1, 2, 3, 4]


class _FakeSystem:
    """Records the run driver's calls (no GPU): run lengths, the capture flag
    of each run, and the grid/ave samples."""

    def __init__(self):
        self.runs, self.capture, self.samples = [], False, []
        self.shape = (4, 4, 4)

    def fourier_capture(self, on=True):
        self.capture = bool(on)

    def run(self, n):
        self.runs.append((n, self.capture))

    def grid_ave(self, slot, which, quant, first, last, repeat, want=False):
        self.samples.append((first, last))
        return np.zeros(1) if want else None

    def dump_wait(self):
        pass


def test_run_driver_stops_and_captures_at_samples():
    from phafd_amd.computes import ComputeComplex, FixGridAve
    from phafd_amd.dump import run_with_dumps
    s = _FakeSystem()
    ave = FixGridAve(s, "ave", 2, 3, 10, ComputeComplex(s, "sq", "ft_phi", "norm"))
    end = run_with_dumps(s, 25, [], 1e-3, first_step=100, aves=[ave])
    assert end == 125
    # a run restarts the schedule: samples at 106 108 110, 116 118 120
    steps, t = [], 100
    for n, cap in s.runs:
        t += n
        steps.append((t, cap))
    assert steps == [(106, True), (108, True), (110, True), (116, True), (118, True), (120, True), (125, False)]
    assert s.samples == [(True, False), (False, False), (False, True)] * 2
    assert s.capture is False
