"""Bead-rod polymers (fix atom/semiflexible/{notether,singletether,
doubletether}, src/fixatom_semiflexible.cpp:27-191) on the GPU against the
oracle, and their equilibrium statistics.

The oracle's bead-rod dynamics is pinned by closed forms and statistics in
test_oracle_beadrod.py (beadrodpmer itself is absent: parity with it is
unpinned).  Here: whole trajectories with grid-atom and lj/cut forces and
thermal noise against the oracle at 1e-10 in both z modes, tethers, the
energy, the bond check, the itermax failure, and free-chain bond-angle
statistics (the Langevin function) with anisotropic friction.
"""
import numpy as np
import pytest

import phafd_amd as P
from helpers import rel_l2, workload

pytestmark = pytest.mark.gpu

BR = dict(seed=99, bondlength=1.0, zeta_para=0.5, zeta_perp=1.0, bending=2.0, temp=1.0)


def _pair(w, tether=0, **kw):
    """GPU and oracle systems of workload w with the bead-rod fix on its chains."""
    import helpers as H
    args = dict(BR, **kw)
    out = []
    for is_oracle in (False, True):
        if is_oracle:
            from oracle import oracle_ctypes as O
            s = O.System(w["n"], w["n"], w["n"], [w["origin"]] * 3, [w["box"]] * 3, zghost=w["zghost"])
        else:
            s = P.System(w["n"], w["n"], w["n"], [w["origin"]] * 3, [w["box"]] * 3, zghost=w["zghost"])
        H.apply(s, w, is_oracle, setup=False)
        s.fix_semiflexible(tether, w["starts"], w["ends"], args["seed"], args["bondlength"], args["zeta_para"],
                           args["zeta_perp"], args["bending"], args["temp"])
        s.setup(w["dt"])
        out.append(s)
    return out


def _chains_workload(zghost=1, seed=21, nchains=12, nbeads=30, lj=True):
    w = workload(n=32, nchains=nchains, nbeads=nbeads, seed=seed, zghost=zghost, chain=False, lj=lj, nucwidth=2.0)
    return w


@pytest.mark.parametrize("zghost", [0, 1])
def test_beadrod_trajectory_vs_oracle(zghost):
    w = _chains_workload(zghost=zghost)
    g, o = _pair(w)
    for _ in range(4):
        g.run(25)
        o.run(25)
        assert rel_l2(g.get_atoms()[1], o.get_atoms()[1]) < 1e-10
        assert rel_l2(g.get_field(P.PHI), o.get_field(P.PHI)) < 1e-10
        assert rel_l2(g.get_forces(), o.get_forces()) < 1e-10
    assert g.counts()["ncalls"] == o.counts()["ncalls"] > 1
    u = g.get_atoms()[1].reshape(w["starts"].size, -1, 3)
    assert np.abs(np.linalg.norm(np.diff(u, axis=1), axis=2) - 1.0).max() < 1e-11
    eg, eo = g.energy(), o.energy()
    assert abs(eg[3] - eo[3]) <= 1e-10 * abs(eo[3])


@pytest.mark.parametrize("tether", [1, 2])
def test_beadrod_tethers_vs_oracle(tether):
    w = _chains_workload(seed=22, nchains=6, nbeads=20)
    g, o = _pair(w, tether=tether)
    u0 = g.get_atoms()[1].reshape(6, 20, 3)
    g.run(40)
    o.run(40)
    u = g.get_atoms()[1].reshape(6, 20, 3)
    assert rel_l2(u, o.get_atoms()[1].reshape(6, 20, 3)) < 1e-10
    assert np.array_equal(u[:, 0], u0[:, 0])
    if tether == 2:
        assert np.array_equal(u[:, -1], u0[:, -1])
    assert np.abs(u[:, 1:-1] - u0[:, 1:-1]).max() > 1e-3


def test_beadrod_bond_check_and_itermax_errors():
    w = _chains_workload(seed=23, nchains=4, nbeads=10, lj=False)
    w["xyz"] = w["xyz"].copy()
    w["xyz"][5:10] += w["xyz"][5] - w["xyz"][4]   # bond 4-5 of chain 0 doubled: beyond 1e-3
    import helpers as H
    s = P.System(32, 32, 32, [0, 0, 0], [32.0] * 3, zghost=1)
    H.apply(s, w, False, setup=False)
    s.fix_semiflexible(0, w["starts"], w["ends"], 1, 1.0, 0.5, 1.0, 2.0, 1.0)
    with pytest.raises(P.PhafdError, match="Adjacent atom displacements do not match the FixAtomSemiFlexible "
                                          "bondlength specified."):
        s.setup(1e-3)
    s.close()
    # a timestep far too long for the corrector: the step returns the
    # reference's error after at most 50 iterations per polymer
    w = _chains_workload(seed=23, nchains=4, nbeads=10, lj=False)
    s = P.System(32, 32, 32, [0, 0, 0], [32.0] * 3, zghost=1)
    H.apply(s, w, False, setup=False)
    s.fix_semiflexible(0, w["starts"], w["ends"], 1, 1.0, 0.5, 1.0, 2.0, 50.0)
    s.setup(2.0)
    with pytest.raises(P.PhafdError, match="Failed to maintain polymer bond-length constraints."):
        s.run(3)
    s.close()


def test_beadrod_rejected_where_unsupported():
    w = _chains_workload(seed=24, nchains=4, nbeads=10)
    g, _ = _pair(w)
    net = P.UNet(seed=1)
    with pytest.raises(P.PhafdError, match="atom/semiflexible"):
        g.run_surrogate(net, 10, 10, 0.1)
    net.close()


def _free_chains(nb, nch, kb, zpar, zperp, dt, nsteps, every):
    side = int(np.ceil(nch ** (1 / 3)))
    xyz = []
    for c in range(nch):
        o = np.array([c % side, (c // side) % side, c // side ** 2]) * (nb + 3.0) + 1.0
        xyz += [o + [i, 0, 0] for i in range(nb)]
    xyz = np.array(xyz)
    box = side * (nb + 3.0) + 2.0
    s = P.System(8, 8, 8, [0, 0, 0], [box] * 3)
    s.set_field(P.PHI, np.full((8, 8, 8), 0.5))
    n = len(xyz)
    s.set_atoms(xyz, np.zeros(n, np.int32), tags=np.arange(n, dtype=np.int32), mols=np.zeros(n, np.int32))
    st = np.arange(nch, dtype=np.int32) * nb
    s.fix_semiflexible(0, st, st + nb, 5, 1.0, zpar, zperp, kb, 1.0)
    s.setup(dt)
    s.run(nsteps // 10)
    c1, c2 = [], []
    for _ in range((nsteps - nsteps // 10) // every):
        s.run(every)
        b = np.diff(s.get_atoms()[1].reshape(nch, nb, 3), axis=1)
        c = np.sum(b[:, :-1] * b[:, 1:], axis=2)
        c1.append(c.mean())
        c2.append((c * c).mean())
    s.close()
    return float(np.mean(c1)), float(np.mean(c2))


@pytest.mark.parametrize("zpar,zperp", [(0.5, 1.0), (2.0, 1.0)])
def test_beadrod_free_chain_bond_angles(zpar, zperp):
    # independent bond orientations weighted by exp(kb c / kT) (the metric
    # pseudo-potential cancels the rigid-rod metric): <c> = coth(kb) - 1/kb,
    # <c^2> = 1 - 2 coth(kb)/kb + 2/kb^2 (0 and 1/3 at kb = 0), for every
    # friction anisotropy; dt = 1e-3 leaves an O(dt) bias of ~1e-3
    m, m2 = _free_chains(3, 8000, 0.0, zpar, zperp, 1e-3, 11000, 200)
    print(f"kb 0 zpar {zpar}: <c> {m:.4f} <c^2> {m2:.4f}", flush=True)
    assert abs(m) < 0.006 and abs(m2 - 1 / 3) < 0.006
    m, m2 = _free_chains(4, 6000, 2.0, zpar, zperp, 1e-3, 11000, 200)
    L = 1 / np.tanh(2.0) - 0.5
    L2 = 1 - 2 / np.tanh(2.0) / 2.0 + 2 / 4.0
    print(f"kb 2 zpar {zpar}: <c> {m:.4f} ({L:.4f}) <c^2> {m2:.4f} ({L2:.4f})", flush=True)
    assert abs(m - L) < 0.008 and abs(m2 - L2) < 0.008
