"""Training data and training of the surrogate network on the GPU.

Trajectories of the plain solver at 64^3 (dx = 1, periodic ghosts along z)
in three families -- the 64^3 reference case (phi0 0.5, chi 3, 20 x 50-bead
chains, gridatom/gaussian cut 3), the 256^3 production case's bead density
and composition at 64^3 (phi0 0.4, chi 2.5, 80 x 100 beads, gaussian cut 3)
and the 512^3 case's coupling at 64^3 (LJish cut 2.5, width 2.5) -- are cut
into blocks of K steps.  A sample is the network input (phi at the block
start, the bead density at its start and end) and the target (phi after K
steps - the exponential-Euler base) / norm[4].  The
network is trained on the GPU (phafd_unet_train); the tolerance of the
block check is set above the residuals of the trained network's own
predictions on the held-out trajectories.

    python -m phafd_amd.surrogate --out weights.npz [--K 100 --trajectories 100 --blocks 10]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

FAMILIES = {
    # name: (name, phi0, var, chi, nchains, nbeads, style, cut, width)
    "reference64": ("reference64", 0.5, 0.02, 3.0, 20, 50, 0, 3.0, 1.0),
    "production": ("production", 0.4, 0.04, 2.5, 80, 100, 0, 3.0, 1.0),
    "ljish": ("ljish", 0.4, 0.04, 2.5, 80, 100, 1, 2.5, 2.5),
}
# trajectory mixes: the Gaussian-coupled network (the 64^3 reference case
# and the production case's composition) and the LJish-coupled one
MIXES = {
    "gaussian": ["production", "reference64", "production", "production"],
    "ljish": ["ljish"],
}


def make_system(P, n, fam, seed):
    """A plain-solver system of one family (the same script commands as the
    benchmark workloads) on an n^3 grid; the chain count scales with the
    volume, so the bead density is the family's at 64^3."""
    _, phi0, var, chi, nch, nb, style, cut, width = fam
    nch = nch * (n // 64) ** 3 if n >= 64 else max(1, nch * n ** 3 // 64 ** 3)
    s = P.System(n, n, n, [0, 0, 0], [float(n)] * 3, zghost=P.ZGHOST_PERIODIC)
    s.set_field(P.PHI, P.ic_noisy_constant((n, n, n), phi0, var, 1234 + seed))
    xyz, mol = P.ic_random_chains(nch, nb, 1.0, 1.0, [0, 0, 0], [float(n)] * 3, 42 + seed)
    s.set_atoms(xyz, np.zeros(len(xyz), np.int32), tags=np.arange(len(xyz), dtype=np.int32), mols=mol)
    ip = s.pair_gridatom(style, cut)
    s.pair_coeff(ip, 0, 1.0, 0.1, width)
    s.pair_ljcut(1.122462048, 1.0, 1.0)
    s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, P.NOISE_OFF)
    s.fix_floryhuggins(1.0, 1.0, chi)
    st = (np.arange(nch) * nb).astype(np.int32)
    s.fix_chain(st, st + nb, 1.0, 100.0, 10.0, 1.0)
    s.neighbor(0.5)
    s.setup(1e-3)
    return s


def blocks(P, s, nblocks, K):
    """(phi0, rho0, rho1, base, phi1, plain residual) of consecutive K-step
    blocks (rho0, rho1: bead densities at the block's start and end)."""
    rho1 = None
    for _ in range(nblocks):
        phi0 = s.get_field(P.PHI)
        rho0 = s.density() if rho1 is None else rho1
        s.run(K)
        phi1 = s.get_field(P.PHI)
        rho1 = s.density()
        base = s.surrogate_base(phi0, rho0, K)
        res, _ = s.block_residual(phi0, phi1, rho0, rho1, K)
        yield phi0, rho0, rho1, base, phi1, res


def build_dataset(P, net, mix="gaussian", n=64, ntraj=20, nblocks=100, K=10, log=print):
    """Fill the network's device training set; returns per-sample metadata
    (|phi1 - phi0|, |phi1 - base|, plain residual, trajectory) and the norm
    used (its scales taken from the first trajectory's first block)."""
    names = MIXES[mix]
    fams = [FAMILIES[names[i % len(names)]] for i in range(ntraj)]
    norm = None
    meta = []
    net.data_alloc(n, n, n, ntraj * nblocks)
    k = 0
    t0 = time.time()
    for t, fam in enumerate(fams):
        s = make_system(P, n, fam, seed=100 + t)
        for phi0, rho0, rho1, base, phi1, res in blocks(P, s, nblocks, K):
            d = phi1 - base
            if norm is None:   # scales from the first block: O(1) inputs and targets
                norm = np.array([max(float(np.std(phi0)), 1e-6), 1.0, 0.4, 0.1, max(float(np.std(d)), 1e-12)])
                net.set_norm(norm)
            x = P.surrogate_inputs(phi0, rho0, rho1, norm)
            net.data_put(k, x, (d / norm[4]).astype(np.float32))
            meta.append((float(np.linalg.norm(phi1 - phi0)), float(np.linalg.norm(d)), res, t))
            k += 1
        s.close()
        log(f"trajectory {t + 1}/{ntraj} ({fam[0]}) done, {k} samples, {time.time() - t0:.1f} s")
    return np.array(meta), norm


def calibrate(P, net, norm, mix, first_seed, ntraj, nblocks, K, n=64):
    """Residuals of the network's own predictions on trajectories it was not
    trained on (regenerated from their seeds) and the phi error per block;
    the block check's tolerance is set above the largest residual, so that
    it rejects what the network does not usually produce."""
    names = MIXES[mix]
    res, err = [], []
    for t in range(first_seed, first_seed + ntraj):
        s = make_system(P, n, FAMILIES[names[t % len(names)]], seed=100 + t)
        for phi0, rho0, rho1, base, phi1, _ in blocks(P, s, nblocks, K):
            out = net.forward(P.surrogate_inputs(phi0, rho0, rho1, norm)).astype(np.float64)
            pred = base + norm[4] * (out - out.mean())
            r, _ = s.block_residual(phi0, pred, rho0, rho1, K)
            res.append(r)
            err.append(float(np.linalg.norm(pred - phi1) / np.linalg.norm(phi1 - phi0)))
        s.close()
    return np.array(res), np.array(err)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    # defaults: the committed network (phafd_amd/data/surrogate_gaussian64.npz)
    ap.add_argument("--trajectories", type=int, default=100)
    ap.add_argument("--blocks", type=int, default=10)
    ap.add_argument("--K", type=int, default=100)
    ap.add_argument("--steps", type=int, default=9000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--holdout", type=int, default=4, help="trajectories kept for validation")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mix", default="gaussian", choices=sorted(MIXES))
    ap.add_argument("--n", type=int, default=64, help="training grid (n^3, dx = 1)")
    ap.add_argument("--recalibrate", metavar="NPZ", default=None,
                    help="only set the block-check tolerance of a trained network again, on its held-out "
                         "trajectories (--trajectories, --holdout, --blocks as trained, --blocks possibly longer)")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import phafd_amd as P

    def log(msg):
        print(msg, flush=True)

    if args.recalibrate:
        z = np.load(args.recalibrate)
        K, n, mix, norm = int(z["K"][0]), int(z["n"][0]), str(z["mix"][0]), z["norm"]
        net = P.UNet(seed=args.seed)
        net.load(args.recalibrate)
        vres, verr = calibrate(P, net, norm, mix, args.trajectories - args.holdout, args.holdout, args.blocks, K, n=n)
        tol = 1.25 * float(vres.max())
        np.savez(args.out, params=z["params"], norm=norm, tol=np.array([tol]), K=z["K"], n=z["n"], mix=z["mix"])
        log(json.dumps({"recalibrated": args.recalibrate, "old_tol": float(z["tol"][0]), "tol": tol,
                        "held_out_blocks": int(vres.size), "network_residual_validation_max": float(vres.max()),
                        "network_residual_validation_mean": float(vres.mean()),
                        "phi_err_per_block_max": float(verr.max())}))
        net.close()
        return

    net = P.UNet(seed=args.seed)
    meta, norm = build_dataset(P, net, mix=args.mix, n=args.n, ntraj=args.trajectories, nblocks=args.blocks, K=args.K,
                               log=log)
    ntrain = (args.trajectories - args.holdout) * args.blocks
    nval = len(meta) - ntrain
    t0 = time.time()
    # learning rate per chunk: a warmup over the first half chunk (Adam's first
    # updates have full size; at full rate they can leave every ReLU of a layer
    # inactive, after which the network predicts a constant), then decay
    per = max(1, args.steps // 6)
    w = max(1, per // 6)
    schedule = [(w, 0.1), (w, 0.3), (w, 0.6), (per - 3 * w, 1.0)] + [(per, c) for c in (0.7, 0.5, 0.3, 0.15, 0.05)]
    hist = []
    # the validation error of predicting no correction at all: a network
    # that stays at this level has lost every ReLU of a layer and only
    # predicts a constant (which the mean removal cancels)
    # (the targets d / norm[4] have zero mean: both phi1 and the base keep the mass)
    null_mse = float(np.mean((meta[ntrain:, 1] / norm[4]) ** 2 / args.n ** 3)) if nval else None
    for k, (nst, c) in enumerate(schedule):
        if nst <= 0:
            continue
        losses = net.train(nst, args.batch, ntrain, args.lr * c)
        val = net.eval(ntrain, nval) if nval else np.zeros(0)
        hist.append((float(losses.mean()), float(val.mean()) if nval else None))
        log(f"lr {args.lr * c:.2e}: train mse {hist[-1][0]:.4e}, validation mse {hist[-1][1]}, "
            f"{time.time() - t0:.1f} s")
        if nval and k >= 4 and hist[-1][1] >= 0.9 * null_mse:
            net.close()
            raise SystemExit(f"training collapsed: validation mse {hist[-1][1]:.4e} stays at the no-correction "
                             f"level {null_mse:.4e} (a layer's ReLUs all inactive); rerun with a lower --lr or "
                             f"another --seed")
    # relative error of the predicted change on the validation blocks
    val = net.eval(ntrain, nval) if nval else np.zeros(0)
    N = args.n ** 3   # the training grid
    err_pred = np.sqrt(val * N) * norm[4]
    rel = err_pred / meta[ntrain:, 0] if nval else np.zeros(0)
    base_only = meta[ntrain:, 1] / meta[ntrain:, 0] if nval else np.zeros(0)
    vres, verr = calibrate(P, net, norm, args.mix, args.trajectories - args.holdout, args.holdout, args.blocks,
                           args.K, n=args.n)
    tol = 1.25 * float(vres.max())
    summary = {"mix": args.mix, "n": args.n, "norm": norm.tolist(), "tol": tol, "K": args.K, "train_samples": ntrain,
               "validation_samples": nval,
               "history": hist, "validation_rel_err_change_mean": float(rel.mean()) if nval else None,
               "validation_rel_err_change_max": float(rel.max()) if nval else None,
               "base_only_rel_err_change_mean": float(base_only.mean()) if nval else None,
               "plain_residual_max": float(meta[:, 2].max()),
               "network_residual_validation_max": float(vres.max()),
               "network_residual_validation_mean": float(vres.mean()), "train_seconds": time.time() - t0}
    np.savez(args.out, params=net.get_params(), norm=norm, tol=np.array([tol]), K=np.array([args.K]),
             n=np.array([args.n]), mix=np.array([args.mix]))
    log(json.dumps(summary))
    net.close()


if __name__ == "__main__":
    main()
