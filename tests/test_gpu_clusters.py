"""compute grid/clusters and run_until_touching on the GPU, through the C ABI.

  clusters ........ against the oracle's sweeps bit for bit (integer labels)
                    on random fields of several shapes and on a phase-separated
                    field; at 256^3 through the properties that fix the
                    answer: one label per connected component, the label is
                    the id of a member and no member's id is smaller
  touching ........ the events and stop step of run_until_touching against the
                    oracle on the same trajectory (touch point away from every
                    distance by more than the trajectory's 1e-10 agreement)
"""
import numpy as np
import pytest

import phafd_amd as P
from oracle import oracle_ctypes as O
from tests.helpers import workload, make_gpu, make_oracle

pytestmark = pytest.mark.gpu


def _gpu_field(phi):
    nz, ny, nx = phi.shape
    s = P.System(nx, ny, nz, [0, 0, 0], [float(nx), float(ny), float(nz)])
    s.set_field(P.PHI, phi)
    return s


def _oracle_field(phi):
    nz, ny, nx = phi.shape
    s = O.System(nx, ny, nz, [0, 0, 0], [float(nx), float(ny), float(nz)])
    s.set_field(O.PHI, phi)
    return s


@pytest.mark.parametrize("shape,thr,cond", [((8, 8, 8), 0.45, "lt"), ((16, 4, 8), 0.5, "gt"),
                                            ((4, 16, 32), 0.4, "lt"), ((32, 8, 4), 0.55, "gt"),
                                            ((32, 32, 32), 0.5, "lt"), ((4, 64, 8), 0.5, "lt")])
def test_clusters_match_oracle(shape, thr, cond):
    phi = 0.5 + 0.1 * np.random.default_rng(sum(shape) + 1).standard_normal(shape)
    got = _gpu_field(phi).compute_grid_clusters(thr, cond)
    ref = _oracle_field(phi).compute_grid_clusters(thr, 0 if cond == "lt" else 1)
    assert np.array_equal(got, ref)


def test_clusters_of_a_phase_separated_field():
    w = workload(n=32, nchains=4, nbeads=20, seed=51, chi=3.0, var=0.05)
    o = make_oracle(w)
    o.run(300)
    phi = o.get_field(O.PHI)
    thr = float(phi.mean())
    for cond in ("lt", "gt"):
        got = _gpu_field(phi).compute_grid_clusters(thr, cond)
        ref = _oracle_field(phi).compute_grid_clusters(thr, 0 if cond == "lt" else 1)
        assert np.array_equal(got, ref)


def test_clusters_256_properties():
    n = 256
    rng = np.random.default_rng(52)
    # smooth random field: large clusters that wrap the box
    k = np.fft.fftfreq(n)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    phi = np.fft.ifftn(np.fft.fftn(rng.standard_normal((n, n, n))) * np.exp(-kk * 400)).real
    thr = 0.0
    lab = _gpu_field(phi).compute_grid_clusters(thr, "lt")
    inside = phi < thr
    assert np.array_equal(lab >= 0, inside)
    i, j, kx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ids = (i + (j * n + kx) * n).astype(np.float64)
    for ax in range(3):   # neighbours inside the set share the label
        nb_in = np.roll(inside, -1, axis=ax)
        both = inside & nb_in
        assert np.array_equal(lab[both], np.roll(lab, -1, axis=ax)[both])
    assert np.all(lab[inside] <= ids[inside])                      # no member has a smaller id
    labels = np.unique(lab[inside]).astype(np.int64)
    li, lk, lj = labels % n, (labels // n) % n, labels // (n * n)   # id = i + (j*n + k)*n
    assert np.array_equal(lab[li, lj, lk], labels.astype(np.float64))   # each label is a member's id
    assert 1 < len(labels) < inside.sum() / 100


def test_run_until_touching_matches_oracle():
    w = workload(n=32, nchains=1, nbeads=12, seed=50, lj=True)
    ref = make_oracle(w)
    d = []
    for _ in range(60):
        ref.run(1)
        x = ref.get_atoms()[0]
        d.append(np.linalg.norm(x[-1] - x[0]))
    d = np.sort(np.array(d))
    gaps = np.diff(d)
    m = int(np.argmax(gaps[len(d) // 4: 3 * len(d) // 4])) + len(d) // 4
    tp = float(0.5 * (d[m] + d[m + 1]))          # midway in a wide gap of the distances
    assert gaps[m] > 1e-6
    for cont, delay in ((True, 2), (True, 0), (False, 0)):
        g = make_gpu(w).run_until_touching(60, tp, continue_running=cont, delay_step=delay)
        o = make_oracle(w).run_until_touching(60, tp, continue_running=cont, delay_step=delay)
        assert g == o, (cont, delay)
    assert g["stopped"] and g["steps"] < 60


def test_clusters_fp32_form():
    # FP32 fields: the labels of the float field (read back widened, exactly)
    # equal the oracle's on the same values
    phi = 0.5 + 0.1 * np.random.default_rng(53).standard_normal((16, 32, 32))
    nz, ny, nx = phi.shape
    s = P.System(nx, ny, nz, [0, 0, 0], [float(nx), float(ny), float(nz)], precision=P.FP32)
    s.set_field(P.PHI, phi)
    got = s.compute_grid_clusters(0.5, "gt")
    ref = _oracle_field(s.get_field(P.PHI)).compute_grid_clusters(0.5, 1)
    assert np.array_equal(got, ref)
