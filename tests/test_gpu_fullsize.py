"""Parity at the benchmark workloads' full sizes (HIP path through the C ABI
against the FP64 oracle, and the FP32 / surrogate / batched forms against
the plain FP64 GPU path).

- 256^3 grid, 5000 x 100-bead chains (500k beads), gridatom/gaussian width 1
  cut 3, lj/cut 2^(1/6), chi 2.5: gathered atoms, bins, grid list and half
  list bit for bit at set-up and after the first rebuild (step 11 of this
  workload; the gathered positions there at 1e-12, since 11 steps of two
  FFT implementations agree to rounding), in both z ghost modes; phi, mu,
  positions and forces after 3 steps and at the rebuild
  (src/npair_grid_bin.cpp:74-125, src/nbin_standard.cpp:181-229).
- The same workload under the learned surrogate for 1000 steps against the
  plain FP64 run, for the K = 100 network (trained at 64^3) and the K = 10
  one (trained at 128^3): phi within 1e-3, bead positions bounded.
- The 512^3 / 2M-bead LJish workload under the LJish network (K = 10) for
  1000 steps against plain FP64: phi within 1e-3, bead positions bounded.
- 512^3 grid, 20000 x 100-bead chains (2M beads), gridatom/LJish cut 2.5:
  lists and one FP64 step against the oracle; the FP32 form against FP64
  over 500 steps (src/pair_gridatom_lj_ish.cpp:44-181).
- 64 batched 128^3 replicas: every member bit for bit against its solo
  run; member 0 against the oracle over 20 steps.

The workloads are bench.py's (same generators, seeds and parameters).
"""
import numpy as np
import pytest

import bench
import phafd_amd as P
from helpers import rel_l2
from oracle import oracle_ctypes as O

pytestmark = pytest.mark.gpu


def _inputs(name, zghost=None):
    wl = dict(bench.WORKLOADS[name])
    if zghost is not None:
        wl["zghost"] = zghost
    phi, xyz, mol = bench.setup_inputs(wl)
    return wl, phi, xyz, mol


def _gpu(wl, phi, xyz, mol, precision=None):
    n = wl["n"]
    s = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"],
                 precision=P.FP64 if precision is None else precision)
    bench.configure(s, wl, phi, xyz, mol)
    return s


def _oracle(wl, phi, xyz, mol):
    n = wl["n"]
    s = O.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
    bench.configure(s, wl, phi, xyz, mol, is_oracle=True)
    return s


def _same_lists(g, o, where, exact_positions=True):
    """Counts, gathered atoms (positions, types, labels, owners), bins, grid
    and half lists bit for bit.  After steps the two trajectories agree to
    rounding only, so the gathered positions are then compared at 1e-12 and
    every integer array still bit for bit."""
    cg, co = g.counts(), o.counts()
    for k in ("nowned", "ngathered", "nlocal", "nghost", "ncalls"):
        assert cg[k] == co[k], (where, k, cg[k], co[k])
    for i, (a, b) in enumerate(zip(g.gathered(), o.gathered())):
        if i == 0 and not exact_positions:
            assert rel_l2(a, b) < 1e-12, (where, "gathered positions", rel_l2(a, b))
            continue
        assert np.array_equal(a, b), (where, "gathered", i)
    for i, (a, b) in enumerate(zip(g.bins(), o.bins())):
        assert np.array_equal(a, b), (where, "bins", i)
    for i, (a, b) in enumerate(zip(g.gridlist(), o.gridlist())):
        assert np.array_equal(a, b), (where, "gridlist", i)
    for i, (a, b) in enumerate(zip(g.halflist(), o.halflist())):
        assert np.array_equal(a, b), (where, "halflist", i)


def _same_state(g, o, where, tol=1e-10):
    for f in (P.PHI, P.MU):
        e = rel_l2(g.get_field(f), o.get_field(f))
        assert e < tol, (where, f, e)
    xg, xo = g.get_atoms()[0], o.get_atoms()[0]
    assert rel_l2(xg, xo) < 1e-12, (where, "positions", rel_l2(xg, xo))
    assert np.max(np.abs(xg - xo)) < 1e-9, (where, "positions max", float(np.max(np.abs(xg - xo))))
    e = rel_l2(g.get_forces(), o.get_forces())
    assert e < tol, (where, "forces", e)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("zghost", [0, 1])
def test_c2_lists_and_steps_vs_oracle(zghost):
    print(f"c2 256^3, 500k beads, zghost {zghost}", flush=True)
    wl, phi, xyz, mol = _inputs("c2", zghost)
    g = _gpu(wl, phi, xyz, mol)
    o = _oracle(wl, phi, xyz, mol)
    try:
        _same_lists(g, o, "setup")
        g.run(3)
        o.run(3)
        _same_state(g, o, "step 3")
        step = 3
        while g.counts()["ncalls"] == 1:
            assert step < 40, "no rebuild within 40 steps"
            g.run(1)
            o.run(1)
            step += 1
        print(f"first rebuild at step {step}", flush=True)
        _same_lists(g, o, f"rebuild at step {step}", exact_positions=False)
        _same_state(g, o, f"step {step}")
    finally:
        g.close()
        o.close()


def _surrogate_vs_plain(name, weights, steps=1000):
    """steps of the workload by the plain FP64 path and by the surrogate
    (network + check + fallback): phi error, bead error, block statistics."""
    wl, phi, xyz, mol = _inputs(name)
    z = np.load(weights)
    K, tol = int(z["K"][0]), float(z["tol"][0])
    a = _gpu(wl, phi, xyz, mol)
    u0 = a.get_atoms()[1]
    a.run(steps)
    ref, uref = a.get_field(P.PHI), a.get_atoms()[1]
    a.close()
    net = P.UNet(seed=1)
    net.load(weights)
    s = _gpu(wl, phi, xyz, mol)
    try:
        st = s.run_surrogate(net, steps, K, tol)
        got, u = s.get_field(P.PHI), s.get_atoms()[1]
    finally:
        s.close()
        net.close()
    e = rel_l2(got, ref)
    beads = bench.bead_error(u, uref, u0)
    print(f"{name} surrogate ({weights.rsplit('/', 1)[-1]}, K {K}): phi rel L2 {e:.3e}, beads {beads}, "
          f"blocks {st['blocks']}, accepted {st['accepted']}", flush=True)
    return e, beads, st


@pytest.mark.timeout(600)
@pytest.mark.parametrize("weights", [bench.SURROGATE, bench.SURROGATE_K10], ids=["K100_64", "K10_128"])
def test_c2_surrogate_1000_steps_vs_plain(weights):
    e, beads, st = _surrogate_vs_plain("c2", weights)
    assert st["accepted"] > 0
    assert e <= 1e-3
    # bead positions (unwrapped, bead diameter 1): the beads feel the
    # grid-atom force of each block's start for the block's K steps
    # (measured: K = 100 rms 9.2e-4, max 0.057; K = 10 rms 7.7e-5, max
    # 0.0057; against an rms displacement of 1.18)
    assert beads["rms"] <= 0.01 and beads["max"] <= 0.25, beads


@pytest.mark.timeout(900)
def test_c3_ljish_surrogate_1000_steps_vs_plain():
    # the LJish-coupled network (K = 10, trained at 128^3) on the 512^3 /
    # 2M-bead workload; measured phi 9.0e-4 after 1000 steps (96 of 100
    # blocks accepted), beads rms 2.0e-3, max 0.057 (rms displacement 1.13)
    e, beads, st = _surrogate_vs_plain("c3", bench.SURROGATE_LJISH_K10)
    assert st["accepted"] >= 0.9 * st["blocks"]
    assert e <= 1e-3
    assert beads["rms"] <= 0.01 and beads["max"] <= 0.25, beads


@pytest.mark.timeout(900)
def test_c3_lists_and_one_step_vs_oracle():
    print("c3 512^3, 2M beads, LJish cut 2.5", flush=True)
    wl, phi, xyz, mol = _inputs("c3")
    g = _gpu(wl, phi, xyz, mol)
    o = _oracle(wl, phi, xyz, mol)
    try:
        _same_lists(g, o, "setup")
        g.run(1)
        o.run(1)
        _same_state(g, o, "step 1")
        eg, eo = g.energy(), o.energy()
        for i in range(4):
            assert abs(eg[i] - eo[i]) <= 1e-10 * max(abs(eo[i]), 1.0), (i, eg[i], eo[i])
    finally:
        g.close()
        o.close()


@pytest.mark.timeout(900)
def test_c3_fp32_form_500_steps_vs_fp64():
    wl, phi, xyz, mol = _inputs("c3")
    snaps = []
    for prec in (P.FP64, P.FP32):
        s = _gpu(wl, phi, xyz, mol, precision=prec)
        try:
            s.run(500)
            snaps.append(bench.snapshot(s, P))
        finally:
            s.close()
    ref, got = snaps
    err = {k: bench.rel_err(got[k], ref[k]) for k in ref}
    print(f"c3 FP32 vs FP64 after 500 steps: {err}", flush=True)
    # relative L2 (energies: relative per term); measured phi 8.0e-6, mu
    # 3.7e-5, grad 4.6e-5, positions 3.9e-9, forces 3.4e-6, energies <= 4.2e-5
    assert err["phi"] <= 1e-4
    assert err["mu"] <= 5e-4
    assert err["gradphi"] <= 5e-4
    assert err["positions"] <= 1e-7
    assert err["forces"] <= 1e-4
    for i, v in enumerate(err["energies"]):
        assert v <= 5e-4, (i, v)


@pytest.mark.timeout(900)
def test_c4_members_bitwise_vs_solo_and_member0_vs_oracle():
    bw = bench.BATCH
    n, box = bw["n"], bw["box"]
    members = bench.batch_members(bw)
    assert len(members) == 64
    b = P.Batch(n, n, n, [0, 0, 0], [box] * 3, len(members), zghost=bw["zghost"])
    try:
        for i, (wl, phi, xyz, mol) in enumerate(members):
            bench.configure(b.member(i), wl, phi, xyz, mol, setup=False)
        b.setup(1e-3)
        b.run(20)
        got = []
        for i in range(len(members)):
            m = b.member(i)
            got.append((m.get_field(P.PHI), m.get_field(P.MU), m.get_atoms()[0], m.get_forces()))
    finally:
        b.close()
    for i, (wl, phi, xyz, mol) in enumerate(members):
        s = _gpu(wl, phi, xyz, mol)
        try:
            s.set_option(P.OPT_NEXT_GRADIENT, 0)   # the batch takes every step's gradient from phi
            s.run(20)
            ref = (s.get_field(P.PHI), s.get_field(P.MU), s.get_atoms()[0], s.get_forces())
        finally:
            s.close()
        for k, (a, r) in enumerate(zip(got[i], ref)):
            assert np.array_equal(a, r), (i, k)
    wl, phi, xyz, mol = members[0]
    o = _oracle(wl, phi, xyz, mol)
    try:
        o.run(20)
        ph, mu, x, f = got[0]
        assert rel_l2(ph, o.get_field(P.PHI)) < 1e-10
        assert rel_l2(mu, o.get_field(P.MU)) < 1e-10
        assert rel_l2(x, o.get_atoms()[0]) < 1e-12
        assert rel_l2(f, o.get_forces()) < 1e-10
    finally:
        o.close()
