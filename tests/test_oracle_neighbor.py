"""Pins of the oracle's ghosts, bins and lists against first principles (CPU).

* ghost images are enumerated by hand near the faces (the single-rank copy
  swap makes no z images, finding F3 of the survey; periodic mode does);
* every grid-list row equals the brute-force set of gathered atoms within
  cutneigh, and its order is the reference's linked-list order: stencil bins
  in stencil order, each bin LOCAL atoms ascending then GHOST atoms ascending
  (src/nbin_standard.cpp:181-229, src/npair_grid_bin.cpp:74-116);
* the half list covers every physical pair within cutneigh exactly once in
  periodic mode, and none across z in reference mode.
"""
import numpy as np
import pytest

from oracle import oracle_ctypes as O
from tests.helpers import chains


def small_system(zghost, xyz, box=16.0, n=16, gauss=3.0, lj=True):
    s = O.System(n, n, n, [0, 0, 0], [box] * 3, zghost=zghost)
    s.set_field(O.PHI, np.full((n, n, n), 0.5))
    s.set_atoms(xyz, np.zeros(len(xyz), np.int32))
    s.add_pair(O.PAIR_GAUSSIAN, gauss)
    if lj:
        s.add_pair(O.PAIR_LJCUT, 1.122462048, 1.0, 1.0)
    s.neighbor(0.5)
    s.setup(1e-3)
    return s


def test_ghost_images_by_hand():
    xyz = np.array([[0.2, 0.3, 5.0], [8.0, 15.5, 0.1], [8.0, 8.0, 8.0]])
    s = small_system(O.ZGHOST_REFERENCE, xyz)
    gx, tag, lab, own = s.gathered()
    cut = 3.5
    # copies first, in owner order, all LOCAL
    assert list(own[:3]) == [0, 1, 2] and list(lab[:3]) == [1, 1, 1]
    imgs = {(int(o), tuple(np.round(p, 9))) for o, p in zip(own[3:], gx[3:])}
    expect = {(0, (0.2, 16.3, 5.0)), (1, (8.0, -0.5, 0.1)),   # y images
              (0, (16.2, 0.3, 5.0)), (0, (16.2, 16.3, 5.0))}  # x image, corner image
    assert imgs == expect
    assert all(l == 2 for l in lab[3:])
    assert cut == s_cutneigh(s)
    # periodic mode adds the z image of atom 1 (z = 0.1 <= cutghost)
    sp = small_system(O.ZGHOST_PERIODIC, xyz)
    gxp, _, labp, ownp = sp.gathered()
    imgsp = {(int(o), tuple(np.round(p, 9))) for o, p in zip(ownp[3:], gxp[3:])}
    assert (1, (8.0, 15.5, 16.1)) in imgsp and (1, (8.0, -0.5, 16.1)) in imgsp


def s_cutneigh(s):
    info, *_ = s.bins()
    return 3.5


def test_noninteger_origin_quirk():
    # int() truncation of the slab bounds (src/comm_brick.cpp:162-169) leaves
    # atoms with z in [-0.5, 0) without a LOCAL copy: only a +Lz GHOST image
    s = O.System(16, 16, 16, [-0.5, -0.5, -0.5], [16.0] * 3)
    s.set_field(O.PHI, np.full((16, 16, 16), 0.5))
    s.set_atoms(np.array([[3.0, 3.0, -0.25], [3.0, 3.0, 4.0]]), np.zeros(2, np.int32))
    s.add_pair(O.PAIR_GAUSSIAN, 3.0)
    s.neighbor(0.5)
    s.setup(1e-3)
    gx, _, lab, own = s.gathered()
    assert list(own[:2]) == [0, 1] and list(lab[:2]) == [2, 1]
    assert np.isclose(gx[0, 2], 15.75)


def _bins_geometry(s):
    info, bh, b, a2b = s.bins()
    return info, bh, b, a2b


@pytest.mark.parametrize("zghost", [O.ZGHOST_REFERENCE, O.ZGHOST_PERIODIC])
def test_gridlist_bruteforce_and_order(zghost):
    xyz, _ = chains(4, 12, 16.0, seed=5)
    s = small_system(zghost, xyz)
    il, nn, nb = s.gridlist()
    gx, _, lab, _ = s.gathered()
    nowned = len(xyz)
    info, bh, bins, a2b = s.bins()
    nbinx, nbiny, nbinz, mbx, mby, mbz, mlx, mly, mlz = info[:9]
    cut2 = 3.5 ** 2
    n = 16
    g = np.arange(n)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(float)   # dx = 1, origin 0
    d2 = ((pts[:, None, :] - gx[None, :, :]) ** 2).sum(-1)
    off = np.concatenate([[0], np.cumsum(nn)])
    rows = {int(p): nb[off[r]:off[r + 1]] for r, p in enumerate(il)}
    expect_il = np.nonzero((d2 <= cut2).any(1))[0]
    assert np.array_equal(il, expect_il)
    binsize = 16.0 / nbinx
    def bin3(p):
        return np.floor(np.asarray(p) / binsize).astype(int)
    for p in expect_il[:: max(1, len(expect_il) // 200)]:
        want = set((np.nonzero(d2[p] <= cut2)[0] + nowned).tolist())
        got = rows[int(p)]
        assert set(got.tolist()) == want and len(got) == len(want)
        # order: (stencil rank of the atom's bin, ghost flag, index) ascending
        gb = bin3(pts[p])
        keys = []
        for a in got:
            ab = bin3(gx[a - nowned])
            d = ab - gb
            keys.append((d[2], d[1], d[0], int(lab[a - nowned] == 2), int(a)))
        assert keys == sorted(keys)


@pytest.mark.parametrize("zghost", [O.ZGHOST_REFERENCE, O.ZGHOST_PERIODIC])
def test_halflist_exactly_once(zghost):
    xyz, _ = chains(5, 15, 16.0, seed=9)
    s = small_system(zghost, xyz)
    il, nn, nb = s.halflist()
    gx, _, lab, own = s.gathered()
    nowned = len(xyz)
    L = 16.0
    cut2 = 3.5 ** 2
    seen = {}
    off = np.concatenate([[0], np.cumsum(nn)])
    for r, i in enumerate(il):
        for j in nb[off[r]:off[r + 1]]:
            a, b = own[i - nowned], own[j - nowned]
            d = gx[j - nowned] - gx[i - nowned]
            key = (min(a, b), max(a, b)) + tuple(np.round(d if a <= b else -d, 6))
            seen[key] = seen.get(key, 0) + 1
    assert all(v == 1 for v in seen.values())
    # brute force over minimum images
    xs = gx[:0]
    owned = np.array([gx[np.nonzero((own == k) & (lab == 1))[0][0]] for k in range(nowned)])
    want = set()
    for a in range(nowned):
        for b in range(a + 1, nowned):
            d = owned[b] - owned[a]
            crosses_z = abs(d[2]) > L / 2
            d = d - L * np.round(d / L)
            if (d ** 2).sum() <= cut2:
                if zghost == O.ZGHOST_REFERENCE and crosses_z:
                    continue
                want.add((a, b) + tuple(np.round(d, 6)))
    assert set(seen) == want


def test_stencils_by_definition():
    xyz, _ = chains(2, 5, 16.0, seed=1)
    s = small_system(O.ZGHOST_REFERENCE, xyz)
    info, *_ = s.bins()
    nbinx = info[0]
    b = 16.0 / nbinx
    cut = 3.5
    sx = int(np.ceil(cut / b - 1e-12))
    full = half = 0
    for k in range(-sx, sx + 1):
        for j in range(-sx, sx + 1):
            for i in range(-sx, sx + 1):
                d = [max(abs(t) - 1, 0) * b for t in (i, j, k)]
                if sum(t * t for t in d) < cut * cut:
                    full += 1
                    if k > 0 or j > 0 or (j == 0 and i > 0):
                        if k >= 0:
                            half += 1
    assert info[10] == full and info[11] == half
