"""compute complex and fix grid/ave on the GPU against the oracle, through
the C ABI.

  compute complex ... ft_phi / ft_chempot at the in_fourier hook of a run's
                      last step, all four quantities, in the reference's ft
                      order: 1e-10 relative L2 (FFT rounding), on a grid whose
                      z pass is the fused one (64) and one that is not (32)
  capture ........... keeping the spectra changes no step (bit-identical)
  grid/ave .......... every 2, repeat 3, freq 10 over 20 steps: the two window
                      averages against the oracle's samples averaged in the
                      reference's order of operations, 1e-10
"""
import numpy as np
import pytest

import phafd_amd as P
from phafd_amd.computes import ComputeComplex, FixGridAve, run_with_aves
from oracle import oracle_ctypes as O
from tests.helpers import workload, make_gpu, make_oracle, rel_l2

pytestmark = pytest.mark.gpu
PRE_IN_FOURIER = (O.PH_START, O.PH_NEIGH, O.PH_PAIR, O.PH_INITIAL, O.PH_POSTFORCE, O.PH_PREFINAL)


def _oracle_step_with(o, fn):
    """One oracle step with fn() called at its in_fourier hook."""
    for ph in PRE_IN_FOURIER:
        o.phase(ph)
    r = fn()
    o.phase(O.PH_FINAL)
    o.phase(O.PH_POSTFINAL)
    return r


@pytest.mark.parametrize("n", [32, 64])
def test_compute_complex_in_fourier(n):
    w = workload(n=n, nchains=8, nbeads=20, seed=42, zghost=1)
    g, o = make_gpu(w), make_oracle(w)
    g.fourier_capture(True)
    g.run(12)
    o.run(11)
    ref = _oracle_step_with(o, lambda: {(wh, q): o.compute_complex(wh, q) for wh in (0, 1) for q in range(4)})
    for (wh, q), r in ref.items():
        got = g.compute_complex(P.PHIHAT if wh == 0 else P.MUHAT, q)
        assert got.shape == r.shape
        assert rel_l2(got, r) < 1e-10, (wh, q)
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10


def test_capture_changes_no_step():
    w = workload(n=64, nchains=8, nbeads=20, seed=43, zghost=1)
    a, b = make_gpu(w), make_gpu(w)
    a.fourier_capture(True)
    for _ in range(3):
        a.run(4)
        b.run(4)
    assert np.array_equal(a.get_field(P.PHI), b.get_field(P.PHI))
    assert np.array_equal(a.get_atoms()[0], b.get_atoms()[0])


def test_grid_ave_windows_against_oracle():
    w = workload(n=32, nchains=8, nbeads=20, seed=44, zghost=1)
    g, o = make_gpu(w), make_oracle(w)
    cmod = ComputeComplex(g, "sq", "ft_phi", "norm")
    ave = FixGridAve(g, "sqave", 2, 3, 10, cmod, slot=1)
    windows = []
    orig = ave.end_of_step

    def keep(step):
        orig(step)
        if (step - ave.first_step) % ave.freq == 0:
            windows.append(ave.array.copy())
    ave.end_of_step = keep
    run_with_aves(g, 20, [ave])
    assert len(windows) == 2
    # the oracle: samples at steps 6 8 10 and 16 18 20, summed from zero and
    # divided by repeat after the last addition (src/fixgrid_ave.cpp:170-215)
    acc, got = None, []
    for step in range(1, 21):
        if step in (6, 8, 10, 16, 18, 20):
            x = _oracle_step_with(o, lambda: o.compute_complex(0, 3))
            acc = x.copy() if step in (6, 16) else acc + x
            if step in (10, 20):
                acc = acc / 3
                got.append(acc)
        else:
            o.run(1)
    for a, b in zip(windows, got):
        assert rel_l2(a, b) < 1e-10
    assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10


def test_compute_complex_fp32_form():
    # the FP32 form keeps float spectra: the capture widens them.  Against the
    # oracle at 2e-4: the non-mean modes carry FP32 rounding relative to the
    # fluctuation, not to phi (imag, which has no mean mode, measures 6e-5;
    # the FP32 form's grad phi is at 3e-5 for the same reason)
    w = workload(n=64, nchains=8, nbeads=20, seed=45, zghost=1)
    g, o = make_gpu(w, P.FP32), make_oracle(w)
    g.fourier_capture(True)
    g.run(6)
    o.run(5)
    ref = _oracle_step_with(o, lambda: {q: o.compute_complex(0, q) for q in range(4)})
    for q, r in ref.items():
        assert rel_l2(g.compute_complex(P.PHIHAT, q), r) < 2e-4, q
