"""Pin of the oracle's compute grid/clusters (CPU): an independent
breadth-first search over the periodic 6-neighbour graph gives each cluster
the smallest of the reference's ids i + (j*Nx + k)*Ny of its points."""
from collections import deque

import numpy as np
import pytest

from oracle import oracle_ctypes as O


def bfs_labels(phi, thr, gt):
    nz, ny, nx = phi.shape
    inside = phi > thr if gt else phi < thr
    lab = np.full(phi.shape, -1.0)
    seen = np.zeros(phi.shape, bool)
    for start in zip(*np.nonzero(inside)):
        if seen[start]:
            continue
        comp, q = [], deque([start])
        seen[start] = True
        while q:
            i, j, k = q.popleft()
            comp.append((i, j, k))
            for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                n = ((i + di) % nz, (j + dj) % ny, (k + dk) % nx)
                if inside[n] and not seen[n]:
                    seen[n] = True
                    q.append(n)
        m = min(i + (j * nx + k) * ny for i, j, k in comp)
        for c in comp:
            lab[c] = m
    return lab


def _sys(phi):
    nz, ny, nx = phi.shape
    s = O.System(nx, ny, nz, [0, 0, 0], [float(nx), float(ny), float(nz)])
    s.set_field(O.PHI, phi)
    return s


@pytest.mark.parametrize("shape,thr,gt", [((8, 8, 8), 0.45, False), ((16, 4, 8), 0.5, True),
                                          ((4, 16, 32), 0.4, False), ((32, 8, 4), 0.55, True)])
def test_clusters_against_bfs(shape, thr, gt):
    rng = np.random.default_rng(sum(shape))
    phi = 0.5 + 0.1 * rng.standard_normal(shape)
    got = _sys(phi).compute_grid_clusters(thr, 1 if gt else 0)
    assert np.array_equal(got, bfs_labels(phi, thr, gt))


def test_clusters_special_cases():
    phi = np.full((4, 4, 4), 0.3)
    assert np.array_equal(_sys(phi).compute_grid_clusters(0.5, 0), np.zeros((4, 4, 4)))   # one cluster, id 0
    assert np.array_equal(_sys(phi).compute_grid_clusters(0.5, 1), np.full((4, 4, 4), -1.0))
    # a slab across the z boundary is one cluster through the periodic wrap
    phi = np.full((8, 4, 4), 0.9)
    phi[0] = phi[7] = 0.1
    lab = _sys(phi).compute_grid_clusters(0.5, 0)
    assert len(np.unique(lab[lab >= 0])) == 1 and lab[7, 0, 0] == 0
