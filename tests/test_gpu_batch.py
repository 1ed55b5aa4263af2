"""Batched replicas (phafd_hip_batch_*) through the C ABI.

A batch advances independent systems of one grid together: the grid passes
run once over all members (a slab of fields, captured in hipGraphs), the atom
side per member.  The arithmetic of a member is the arithmetic of the same
system run alone (each step's gradient from phi: PHAFD_OPT_NEXT_GRADIENT
off), so the bar is bit-identity with the solo run (phi,
positions, images, forces, list counters) on every path the grid can take:
  64^3 -- separable gradient and the fused z pass of the update;
  32^3 -- whole r2c / c2r transforms and the pointwise kernels.
One member is also checked against the CPU oracle (relative L2 1e-10, the
FP64 tolerance of the solo trajectory tests).
"""
import numpy as np
import pytest

import phafd_amd as P
from tests.helpers import workload, apply, make_gpu, make_oracle, rel_l2

pytestmark = pytest.mark.gpu


def _members(n, k=3):
    # different chi, mean phi, atom counts and grid-atom styles per member;
    # one dynamics (mobility, temperature, volFH, gamma) for all
    return [workload(n=n, nchains=3 + 4 * i, nbeads=20, seed=11 + i, chi=2.5 + 0.5 * i,
                     phi0=0.4 + 0.05 * i, ljish=(i % 2 == 1)) for i in range(k)]


def _batch(ws, graphs=True):
    n = ws[0]["n"]
    b = P.Batch(n, n, n, [0.0] * 3, [float(n)] * 3, len(ws))
    for i, w in enumerate(ws):
        apply(b.member(i), w, False, setup=False)
    b.setup(ws[0]["dt"])
    b.set_graphs(graphs)
    return b


def _state(s):
    x, u, im = s.get_atoms()
    c = s.counts()
    return dict(phi=s.get_field(P.PHI), x=x, u=u, im=im, f=s.get_forces(),
                calls=c["ncalls"], nless=c["nless"], ngreat=c["ngreat"])


def _assert_same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"


@pytest.mark.parametrize("n", [64, 32])
def test_members_match_solo_runs_bitwise(n):
    ws = _members(n)
    b = _batch(ws)
    info = b.info()
    assert info["members"] == 3 and 1 <= info["members_per_launch"] <= 3
    b.run(12)
    for i, w in enumerate(ws):
        s = make_gpu(w)
        s.set_option(P.OPT_NEXT_GRADIENT, 0)   # the batch takes every step's gradient from phi
        s.run(12)
        _assert_same(_state(b.member(i)), _state(s), f"member {i}")
    b.close()


@pytest.mark.parametrize("n", [64, 32])
def test_one_launch_per_pass_matches_per_member_launches_bitwise(n):
    # the members' pair phase in one launch per pass (grid-atom force walk,
    # mu sums, lj/cut, fold levels, chain steps over every member) against
    # per-member launches and against the solo runs, through rebuilds
    ws = _members(n, 4)
    f = _batch(ws)
    f.set_fused(1)
    m = _batch(ws)
    m.set_fused(0)
    assert f.fused_info()["fused"] and not m.fused_info()["fused"]
    for _ in range(3):
        f.run(5)
        m.run(5)
        for i in range(len(ws)):
            _assert_same(_state(f.member(i)), _state(m.member(i)), f"member {i}")
    assert f.member(3).counts()["ncalls"] > 1
    for i, w in enumerate(ws):
        s = make_gpu(w)
        s.set_option(P.OPT_NEXT_GRADIENT, 0)
        s.run(15)
        _assert_same(_state(f.member(i)), _state(s), f"member {i} vs solo")
    f.close()
    m.close()


def test_graphs_match_eager_launches_bitwise():
    ws = _members(64, 4)
    g = _batch(ws, graphs=True)
    e = _batch(ws, graphs=False)
    for _ in range(3):   # graph replays with the members' eager work between
        g.run(4)
        e.run(4)
        for i in range(len(ws)):
            _assert_same(_state(g.member(i)), _state(e.member(i)), f"member {i}")
    g.close()
    e.close()


def test_member_against_oracle():
    ws = _members(32, 2)
    b = _batch(ws)
    b.run(20)
    o = make_oracle(ws[1])
    o.run(20)
    m = b.member(1)
    assert rel_l2(m.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
    assert rel_l2(m.get_atoms()[0], o.get_atoms()[0]) < 1e-10
    assert rel_l2(m.get_forces(), o.get_forces()) < 1e-8
    assert m.counts()["ncalls"] == o.counts()["ncalls"]
    b.close()


def test_batch_rejects_members_it_cannot_share_launches_for():
    ws = _members(32, 2)
    b = P.Batch(32, 32, 32, [0.0] * 3, [32.0] * 3, 2)
    apply(b.member(0), ws[0], False, setup=False)
    ws[1]["noise"] = P.NOISE_FIELD
    apply(b.member(1), ws[1], False, setup=False)
    with pytest.raises(P.PhafdError, match="noise off"):
        b.setup(1e-3)
    with pytest.raises(P.PhafdError, match="setup"):
        b.run(1)
    b.close()
