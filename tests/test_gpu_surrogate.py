"""The learned surrogate on the GPU, through the C ABI.

  forward ............ against the PyTorch CPU FP32 restatement with the same
                       bf16 storage points (tests/unet_ref.py): rel. L2 5e-3
                       (FP32 sums in another order flip a few bf16 roundings)
  parameter grads .... against torch autograd of that restatement, per layer
                       rel. L2 3e-2 (gradients are stored in bf16)
  finite differences . directional derivative along each layer's bias
                       gradient (biases are FP32 masters) and the head
                       weights: 3e-2 relative
  training ........... loss falls on a learnable map; two runs from one seed
                       give the same losses bit for bit
  block check, base . residual, mass drift and the linearised K-step base
                       against numpy (rfftn) on the same inputs: 1e-9
  bead density ....... against a periodic brute-force sum: 1e-5 (FP32 output)
  reproducibility .... two surrogate runs from one state: identical phi and
                       positions (the density is summed in fixed point)
  atom windows ....... speculative windows of atom steps against one step per
                       read-back, bit for bit (with lj/cut, and with
                       harmonic/cut as the only bead pair)
  fallback ........... a rejected block is K plain steps from the restored
                       atoms (1e-13: re-binned lists may reorder sums)
  gain ............... the committed trained network against the same block
                       scheme without it (the exponential-Euler base alone)
                       on a held-out 64^3 production-case trajectory
"""
import os

import numpy as np
import pytest
import torch

import phafd_amd as P
from phafd_amd import surrogate as S
from tests.helpers import workload, make_gpu, rel_l2
from tests import unet_ref as R

pytestmark = pytest.mark.gpu
SHAPE = (8, 32, 64)   # (nz, ny, nx)


def _random_params(net, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for name, cin, cout, ks in R.LAYERS:
        bound = scale * np.sqrt(6.0 / (cin * ks ** 3))
        parts.append(rng.uniform(-bound, bound, cout * cin * ks ** 3))
        parts.append(rng.uniform(-0.1, 0.1, cout))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == net.nparams
    return p


def _inputs(seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((4,) + SHAPE).astype(np.float32)


def _layer_slices():
    out, k = [], 0
    for name, cin, cout, ks in R.LAYERS:
        nw = cout * cin * ks ** 3
        out.append((name, slice(k, k + nw), slice(k + nw, k + nw + cout)))
        k += nw + cout
    return out


def test_forward_matches_torch_restatement():
    net = P.UNet(seed=3)
    p = _random_params(net, 1)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    x = _inputs(2)
    got = net.forward(x)
    ref = R.forward(R.split_params(p), x).detach().numpy()
    assert rel_l2(got, ref) < 5e-3
    net.close()


def test_forward_on_a_tiled_grid_is_the_tiled_forward():
    """On a 256^3 grid holding a 64^3 input tiled 4 x 4 x 4, the output is
    the 64^3 output tiled, bit for bit: every point sees the same inputs
    through the same arithmetic, while the large grid takes the launch
    shapes of large grids (several z slices per block at every level)."""
    net = P.UNet(seed=3)
    net.set_params(_random_params(net, 13))
    x = np.random.default_rng(14).standard_normal((4, 64, 64, 64)).astype(np.float32)
    small = net.forward(x)
    big = net.forward(np.tile(x, (1, 4, 4, 4)))
    assert np.array_equal(big, np.tile(small, (4, 4, 4)))
    net.close()


def test_parameter_gradients_match_torch_autograd():
    net = P.UNet(seed=3)
    p = _random_params(net, 4)
    net.set_params(p)
    x = _inputs(5)
    dout = (1.0 + 0.5 * np.random.default_rng(6).standard_normal(SHAPE)).astype(np.float32)
    net.forward(x)
    net.backward(dout)
    g = net.get_grads(zero=True)
    tp = R.split_params(p, requires_grad=True)
    out = R.forward(tp, x)
    (out * torch.as_tensor(dout)).sum().backward()
    gr = R.flat_grads(tp)
    for name, ws, bs in _layer_slices():
        assert rel_l2(g[ws], gr[ws]) < 3e-2, f"{name} weight grad"
        assert rel_l2(g[bs], gr[bs]) < 3e-2, f"{name} bias grad"
    net.close()


def test_gradients_against_finite_differences():
    net = P.UNet(seed=3)
    p = _random_params(net, 7)
    x = _inputs(8)
    # a cotangent with a non-zero mean: the derivative is a coherent sum over
    # points, while the bf16 roundings of the two evaluations add incoherently
    dout = (1.0 + 0.5 * np.random.default_rng(9).standard_normal(SHAPE)).astype(np.float32)

    def loss(q):
        net.set_params(q)
        return float(np.dot(net.forward(x).ravel().astype(np.float64), dout.ravel()))

    net.set_params(p)
    net.forward(x)
    net.backward(dout)
    g = net.get_grads(zero=True)
    sl = _layer_slices()
    checks = [(name, bs) for name, _, bs in sl] + [("head weight", sl[-1][1])]
    for name, s in checks:
        gs = g[s].astype(np.float64)
        n = np.linalg.norm(gs)
        assert n > 0, name
        d = np.zeros_like(p)
        d[s] = (gs / n).astype(np.float32)
        h = 0.1 if name != "head weight" else 1e-2
        fd = (loss(p + h * d) - loss(p - h * d)) / (2 * h)
        assert abs(fd - n) < 3e-2 * n, f"{name}: finite difference {fd} vs gradient norm {n}"
    net.close()


def _learnable_data(n, seed):
    """Inputs and a fixed local linear map of them (a smoothed channel 0 plus
    a multiple of channel 1): a target the network can represent."""
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for _ in range(n):
        x = rng.standard_normal((4,) + SHAPE).astype(np.float32) * np.array([1, 0.5, 0.5, 0], np.float32)[:, None, None, None]
        c0 = x[0].astype(np.float64)
        lap = sum(np.roll(c0, s, a) for a in range(3) for s in (1, -1)) - 6 * c0
        ys.append((0.3 * lap + 0.5 * x[1]).astype(np.float32))
        xs.append(x)
    return xs, ys


def test_training_reduces_loss_reproducibly():
    xs, ys = _learnable_data(8, 10)
    runs = []
    for _ in range(2):
        net = P.UNet(seed=11)
        p = _random_params(net, 12, scale=0.5)
        net.set_params(p)
        net.data_alloc(SHAPE[2], SHAPE[1], SHAPE[0], len(xs))
        for i, (x, y) in enumerate(zip(xs, ys)):
            net.data_put(i, x, y)
        before = net.eval(0, len(xs)).mean()
        losses = net.train(150, 2, len(xs), 2e-3)
        after = net.eval(0, len(xs)).mean()
        runs.append((before, losses, after))
        net.close()
    before, losses, after = runs[0]
    assert after < 0.3 * before, (before, after)
    assert np.array_equal(runs[0][1], runs[1][1])


def _np_modes(shape, L, f2, K, M=1.0, T=1.0, v=1.0, gamma=1.0, dt=1e-3):
    """Per-mode c and G of the plain update linearised about the mean."""
    pi_ref = 3.141592
    nz, ny, nx = shape
    kz = np.fft.fftfreq(nz) * nz
    ky = np.fft.fftfreq(ny) * ny
    kx = np.arange(nx // 2 + 1)
    q2 = ((2 * pi_ref / L * kz)[:, None, None] ** 2 + (2 * pi_ref / L * ky)[None, :, None] ** 2
          + (2 * pi_ref / L * kx)[None, None, :] ** 2)
    c = M * q2 * dt
    G = 1 - c * (f2 + T / v * gamma * q2)
    return c, G


def _np_block(phi0, phi1, rho0, rho1, K, w, phit=0.1, chi=3.0, tv=1.0):
    """The block check and the exponential-Euler base in numpy (rfftn)."""
    r0, r1 = rho0.astype(np.float64), rho1.astype(np.float64)
    m, rbar = phi0.mean(), r0.mean()
    f2 = tv * (1 / m + 1 / (1 - m) - 2 * chi) + 2 * rbar
    c, G = _np_modes(phi0.shape, w["box"], f2, K)
    mu = lambda p, r: tv * (np.log(p / (1 - p)) + chi * (1 - 2 * p)) + 2 * (p - phit) * r
    f = np.fft.rfftn
    p0, p1, m0, m1 = f(phi0), f(phi1), f(mu(phi0, r0)), f(mu(phi1, r1))
    A = sum(G ** (K - 1 - k) * (1 - k / K) for k in range(K))
    B = sum(G ** (K - 1 - k) * (k / K) for k in range(K))
    Rk = p1 - G ** K * p0 + c * (A * (m0 - f2 * p0) + B * (m1 - f2 * p1))
    wt = np.full(p0.shape[-1], 2.0)
    wt[0] = wt[-1] = 1.0
    res = np.sqrt(np.sum(wt * np.abs(Rk) ** 2) / np.sum(wt * np.abs(p1 - p0) ** 2))
    drift = (phi1.sum() - phi0.sum()) / phi0.sum()
    S = sum(G ** k for k in range(K))
    base = np.fft.irfftn(G ** K * p0 - c * S * (m0 - f2 * p0), s=phi0.shape, axes=(0, 1, 2))
    return res, drift, base


def test_block_check_and_base_against_numpy():
    w = workload(n=32, nchains=6, nbeads=20, seed=21, chi=3.0)
    g = make_gpu(w)
    phi0 = g.get_field(P.PHI)
    rho0 = g.density()
    g.run(10)
    phi1 = g.get_field(P.PHI)
    rho1 = g.density()
    base = g.surrogate_base(phi0, rho0, 10)
    res, drift = g.block_residual(phi0, phi1, rho0, rho1, 10)
    ref_res, ref_drift, ref_base = _np_block(phi0, phi1, rho0, rho1, 10, w)
    assert abs(res - ref_res) < 1e-9 * ref_res
    assert abs(drift - ref_drift) < 1e-12
    assert rel_l2(base - phi0, ref_base - phi0) < 1e-10
    # with both densities the plain method's own block has a small residual
    # (the non-linear part of mu interpolated across the block); a doubled
    # change does not
    assert res < 0.05
    bad, _ = g.block_residual(phi0, phi0 + 2 * (phi1 - phi0), rho0, rho1, 10)
    assert bad > 10 * res
    # the base is much closer to the plain result than no change
    assert np.linalg.norm(phi1 - base) < 0.3 * np.linalg.norm(phi1 - phi0)


def _periodic_density(w, xyz, n):
    """Brute force: sum of the Gaussian kernel value over every atom image
    within the cut, on the n^3 grid of spacing box/n."""
    box, cut = w["box"], w["nucwidth"]   # the kernel's range is its width
    d = box / n
    rho = np.zeros((n, n, n))
    cs = cut * cut
    for x in xyz:
        lo = np.floor((x - cut) / d).astype(int)
        hi = np.ceil((x + cut) / d).astype(int)
        ii = [np.arange(lo[a], hi[a] + 1) for a in range(3)]
        dx = ii[0] * d - x[0]
        dy = ii[1] * d - x[1]
        dz = ii[2] * d - x[2]
        r2 = dz[:, None, None] ** 2 + dy[None, :, None] ** 2 + dx[None, None, :] ** 2
        e = np.where(r2 <= cs, (np.exp(-r2 / cs / 2) - np.exp(-0.5)) * np.exp(0.5) * 1.0, 0.0)
        np.add.at(rho, (np.mod(ii[2], n)[:, None, None], np.mod(ii[1], n)[None, :, None],
                        np.mod(ii[0], n)[None, None, :]), e)
    return rho


def test_bead_density_against_brute_force():
    w = workload(n=32, nchains=5, nbeads=20, seed=22, zghost=1, lj=False)
    g = make_gpu(w)
    rho = g.density()
    ref = _periodic_density(w, g.get_atoms()[0], 32)
    assert rel_l2(rho, ref) < 1e-5


def test_rejected_blocks_are_plain_steps():
    """A rejected block puts the atoms back and runs K plain steps: the same
    arithmetic as the plain run, except that lists rebuilt during the atoms'
    K steps are rebuilt at the restored positions, which can reorder the
    grid-atom sums (rounding-level differences)."""
    w = workload(n=64, nchains=8, nbeads=20, seed=23, chi=3.0)
    a = make_gpu(w)
    b = make_gpu(w)
    net = P.UNet(seed=1)
    st = a.run_surrogate(net, 20, 10, 0.0)   # tolerance 0: every block falls back
    assert st["blocks"] == 2 and st["fallbacks"] == 2 and st["accepted"] == 0
    b.run(20)
    assert rel_l2(a.get_field(P.PHI), b.get_field(P.PHI)) < 1e-13
    assert rel_l2(a.get_atoms()[0], b.get_atoms()[0]) < 1e-13
    assert rel_l2(a.get_forces(), b.get_forces()) < 1e-11
    net.close()


WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "phafd_amd", "data",
                       "surrogate_gaussian64.npz")


def test_surrogate_runs_are_bit_reproducible():
    # every kernel of a block sums in a fixed order (the bead density in
    # 64-bit fixed point), so two runs from one state agree bit for bit
    z = np.load(WEIGHTS)
    K, tol = int(z["K"][0]), float(z["tol"][0])
    out = []
    for _ in range(2):
        s = S.make_system(P, 64, S.FAMILIES["production"], seed=501)
        net = P.UNet(seed=1)
        net.load(WEIGHTS)
        st = s.run_surrogate(net, 2 * K, K, tol)
        out.append((s.get_field(P.PHI), s.get_atoms()[0], st["accepted"]))
        net.close()
        s.close()
    assert out[0][2] == out[1][2] == 2
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_atom_step_windows_match_single_steps():
    # speculative windows of atom steps (one read-back per window, steps
    # after a rebuild decision skipped and redone) against one read-back per
    # step: the same trajectory bit for bit, over blocks with rebuilds
    z = np.load(WEIGHTS)
    K, tol = int(z["K"][0]), float(z["tol"][0])
    out = []
    for windows in (1, 0):
        s = S.make_system(P, 64, S.FAMILIES["production"], seed=502)
        s.set_option(P.OPT_ATOM_WINDOWS, windows)
        net = P.UNet(seed=1)
        net.load(WEIGHTS)
        s.run_surrogate(net, 2 * K, K, tol)
        out.append((s.get_field(P.PHI), s.get_atoms()[0], s.counts()["ncalls"]))
        net.close()
        s.close()
    assert out[0][2] == out[1][2] and out[0][2] > 2   # the same rebuilds, several of them
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_surrogate_atom_steps_with_harmonic_cut():
    # the surrogate's atom steps with pair harmonic/cut as the only bead pair
    # (it then writes the pair force from zero): windows against single steps
    # bit for bit, and the block trajectory stays within the acceptance test
    z = np.load(WEIGHTS)
    K, tol = int(z["K"][0]), float(z["tol"][0])
    _, phi0, var, chi, nch, nb, style, cut, width = S.FAMILIES["production"]
    n = 64
    out = []
    for windows in (1, 0):
        s = P.System(n, n, n, [0, 0, 0], [float(n)] * 3, zghost=P.ZGHOST_PERIODIC)
        s.set_field(P.PHI, P.ic_noisy_constant((n, n, n), phi0, var, 1234 + 503))
        xyz, mol = P.ic_random_chains(nch, nb, 1.0, 1.0, [0, 0, 0], [float(n)] * 3, 42 + 503)
        s.set_atoms(xyz, np.zeros(len(xyz), np.int32), tags=np.arange(len(xyz), dtype=np.int32), mols=mol)
        ip = s.pair_gridatom(style, cut)
        s.pair_coeff(ip, 0, 1.0, 0.1, width)
        s.pair_harmoniccut(1.2, 30.0)
        s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, P.NOISE_OFF)
        s.fix_floryhuggins(1.0, 1.0, chi)
        st = (np.arange(nch) * nb).astype(np.int32)
        s.fix_chain(st, st + nb, 1.0, 100.0, 10.0, 1.0)
        s.neighbor(0.5)
        s.setup(1e-3)
        s.set_option(P.OPT_ATOM_WINDOWS, windows)
        net = P.UNet(seed=1)
        net.load(WEIGHTS)
        s.run_surrogate(net, 2 * K, K, tol)
        out.append((s.get_field(P.PHI), s.get_atoms()[0], s.counts()["ncalls"]))
        net.close()
        s.close()
    assert out[0][2] == out[1][2] and out[0][2] > 1
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.isfinite(out[0][1]).all()


def test_trained_network_gains_over_the_base_alone():
    z = np.load(WEIGHTS)
    K, tol = int(z["K"][0]), float(z["tol"][0])
    steps = 5 * K
    fam = S.FAMILIES["production"]
    plain = S.make_system(P, 64, fam, seed=500)   # a seed no training trajectory used
    plain.run(steps)
    ref = plain.get_field(P.PHI)
    err, stats = {}, {}
    for name in ("network", "base only", "network, checked"):
        s = S.make_system(P, 64, fam, seed=500)
        net = P.UNet(seed=1)
        if name == "base only":
            net.set_norm(z["norm"])   # a fresh network's head is zero: the base alone
        else:
            net.load(WEIGHTS)
        t = tol if name == "network, checked" else 1e9   # 1e9: every block accepted
        stats[name] = s.run_surrogate(net, steps, K, t)
        err[name] = rel_l2(s.get_field(P.PHI), ref)
        net.close()
    print(f"surrogate errors vs plain after {steps} steps:", err, stats)
    assert stats["network"]["accepted"] == 5 and stats["base only"]["accepted"] == 5
    assert err["network"] < 0.5 * err["base only"]
    assert err["network"] < 1e-3
    # with the calibrated tolerance the error stays within the same bound
    assert err["network, checked"] < 1e-3
