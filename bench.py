#!/usr/bin/env python3
"""Benchmark of the PHAFD per-timestep hot path on one MI355X.

One "step" is one full timestep of the reference's Integrate::run hook
sequence (src/integrate.cpp:105-203) over the whole box: grad(phi) by FFT,
neighbour decision (rebuild: pbc, ghost images, bins, grid and half lists),
grid-atom + lj/cut pair pass and ghost force fold-back, chain half step,
Flory-Huggins + r2c(mu), spectral update, chain half step, c2r(phi), NaN check.

Default workload (the plain-FP64-on-GPU configuration): 256^3 periodic grid,
box 256, phi0 = 0.4 + U(-0.02, 0.02), chi = 2.5, noise off, 5000 chains x 100
beads (random walks, bond 1, no two beads closer than 1), lj/cut sigma 1 cut 2^(1/6), gridatom/gaussian
cut 3 width 1 epsilon 1, neighbour skin 0.5, semiflexible harmonic chains
(k = 100, kb = 10), periodic ghost images along z, dt = 1e-3.  Inputs are generated on the host, uploaded,
and resident in HBM before the timed region starts.

Prints ONE JSON line (rank 0).  `value` = grid points x steps / s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-pt·steps/s + beads·steps/s, surrogate/FP32 vs FP64 at ≤1e-3 φ err"

WORKLOADS = {
    # zghost=1: periodic ghost images along z.  One reference rank makes none
    # (its copy swap marks every in-box atom LOCAL), which drops lj/cut across
    # the z boundary; with stiff bonds at dt = 1e-3 a bond straddling z = 0
    # then compresses below the WCA stability limit and the run blows up at
    # its first re-wrap.  Periodic images are what two or more ranks compute.
    "c2": dict(n=256, box=256.0, nchains=5000, nbeads=100, phi0=0.4, var=0.04, chi=2.5,
               cut=3.0, nucwidth=1.0, ljish=False, zghost=1),
    "c1": dict(n=64, box=64.0, nchains=20, nbeads=50, phi0=0.5, var=0.02, chi=3.0,
               cut=3.0, nucwidth=1.0, ljish=False, zghost=1),
    # the mixed-precision workload: its figure of merit is the FP32 leg
    "c3": dict(n=512, box=512.0, nchains=20000, nbeads=100, phi0=0.4, var=0.04, chi=2.5,
               cut=2.5, nucwidth=2.5, ljish=True, zghost=1),
}
DESCR = {
    "c2": "plain FP64 on GPU: 256^3 grid, 5000x100-bead semiflexible chains (500k beads), "
          "gridatom/gaussian width 1 cut 3, lj/cut 2^(1/6), chi 2.5, noise off",
    "c1": "plain FP64: 64^3 grid, 20x50-bead chains, gridatom/gaussian width 1 cut 3, chi 3, noise off",
    "c3": "512^3 grid, 20000x100-bead chains (2M beads), gridatom/LJish cut 2.5, lj/cut 2^(1/6), chi 2.5, "
          "noise off; FP64 and the FP32 form",
    "c4": "batched replicas: 64 independent 128^3 systems, chi U[2,4], bead counts log-uniform 10k-200k in "
          "chains of 50-200, phi0 mean U[0.3,0.5]; gridatom/gaussian width 1 cut 3, lj/cut 2^(1/6), noise off; "
          "one batched launch per grid pass in hipGraphs, members per launch and atom streams autotuned",
}
# batched replicas (phafd_hip_batch_*): ranges the members are drawn from
BATCH = dict(n=128, box=128.0, members=64, chi=(2.0, 4.0), beads=(1.0e4, 2.0e5), chain=(50, 200),
             phi0=(0.3, 0.5), var=0.04, cut=3.0, nucwidth=1.0, ljish=False, zghost=1)


def setup_inputs(wl):
    import phafd_amd as P
    n = wl["n"]
    phi = P.ic_noisy_constant((n, n, n), wl["phi0"], wl["var"], 1234)
    xyz, mol = P.ic_random_chains(wl["nchains"], wl["nbeads"], 1.0, 1.0, [0, 0, 0], [wl["box"]] * 3, 42)
    return phi, xyz, mol


def configure(s, wl, phi, xyz, mol, is_oracle=False, setup=True, noise=0, noise_temp=1.0):
    nch, nb = wl["nchains"], wl["nbeads"]
    s.set_field(0, phi)
    s.set_atoms(xyz, np.zeros(len(xyz), np.int32), tags=np.arange(len(xyz), dtype=np.int32), mols=mol)
    style = 1 if wl["ljish"] else 0
    ip = s.add_pair(style, wl["cut"]) if is_oracle else s.pair_gridatom(style, wl["cut"])
    s.pair_coeff(ip, 0, 1.0, 0.1, wl["nucwidth"])
    if is_oracle:
        s.add_pair(2, 1.122462048, 1.0, 1.0)
    else:
        s.pair_ljcut(1.122462048, 1.0, 1.0)
    s.fix_conjugate(5678, 1.0, noise_temp, 1.0, 1.0, noise)
    s.fix_floryhuggins(1.0, 1.0, wl["chi"])
    st = (np.arange(nch) * nb).astype(np.int32)
    s.fix_chain(st, st + nb, 1.0, 100.0, 10.0, 1.0)
    s.neighbor(0.5)
    if setup:
        s.setup(1e-3)


def batch_members(bw, seed=2024):
    """Per-member workloads of the batched-replica run (seeded draws)."""
    import phafd_amd as P
    rng = np.random.default_rng(seed)
    n, box = bw["n"], bw["box"]
    out = []
    for i in range(bw["members"]):
        chi = float(rng.uniform(*bw["chi"]))
        target = float(np.exp(rng.uniform(np.log(bw["beads"][0]), np.log(bw["beads"][1]))))
        nb = int(rng.integers(bw["chain"][0], bw["chain"][1] + 1))
        nch = max(1, int(round(target / nb)))
        phi0 = float(rng.uniform(*bw["phi0"]))
        wl = dict(bw, nchains=nch, nbeads=nb, chi=chi, phi0=phi0)
        phi = P.ic_noisy_constant((n, n, n), phi0, bw["var"], 1234 + i)
        xyz, mol = P.ic_random_chains(nch, nb, 1.0, 1.0, [0, 0, 0], [box] * 3, 42 + i)
        out.append((wl, phi, xyz, mol))
    return out


def host_cpu():
    """Model name and logical CPU counts of the host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_threads():
    """Threads given to the CPU oracle: OMP_NUM_THREADS if set, else the
    CPUs this process may run on."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    return int(v) if v.isdigit() and int(v) > 0 else len(os.sched_getaffinity(0))


def cpu_baseline(wl, phi, xyz, mol, steps):
    """The CPU oracle on a bounded sample of the same workload: `steps` full
    timesteps after setup.  Its FFT lines and pointwise loops run on
    cpu_threads() OpenMP threads; the pair pass, lists and update traversal
    run on one.  Returns (seconds, threads)."""
    from oracle import oracle_ctypes as O
    th = cpu_threads()
    O.set_threads(th)
    n = wl["n"]
    s = O.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
    configure(s, wl, phi, xyz, mol, is_oracle=True)
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    s.close()
    return dt, th


def cpu_baseline_record(value, t, th, steps, what):
    return dict({"value": value, "unit": "grid-pt·steps/s", "cores": th, "kind": "port",
                 "sample": f"{steps} full timesteps of {what} after setup, CPU FP64 oracle, {th} OpenMP threads for "
                           f"the FFT lines and pointwise loops, one for the rest ({t:.1f} s)"}, **host_cpu())


def bead_error(u, uref, u0):
    """Unwrapped bead positions u against the plain FP64 run's uref (both
    from u0): RMS and max distance, and the RMS displacement of the plain
    run over the same steps as the scale (lengths in the box's units, bead
    diameter 1)."""
    d = np.linalg.norm(u - uref, axis=1)
    return {"rms": float(np.sqrt(np.mean(d * d))), "max": float(d.max()),
            "rms_displacement_fp64": float(np.sqrt(np.mean(np.sum((uref - u0) ** 2, axis=1))))}


def snapshot(s, P):
    """State compared between the FP64 run and the FP32 form."""
    return {"phi": s.get_field(P.PHI), "mu": s.get_field(P.MU), "gradphi": np.stack(
        [s.get_field(P.GRADX), s.get_field(P.GRADY), s.get_field(P.GRADZ)]),
        "positions": s.get_atoms()[0], "forces": s.get_forces(), "energies": np.array(s.energy())}


def rel_err(a, b):
    """Relative L2 error (elementwise relative for the energy vector)."""
    if a.ndim == 1 and a.size == 4:
        return [float(abs(x - y) / max(abs(y), 1e-300)) for x, y in zip(a, b)]
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def copy_calibration(nbytes, reps):
    """A plain device-to-device copy of `nbytes`, `reps` times: the reference
    kernel whose FETCH_SIZE / WRITE_SIZE counters scale the ones of the FFT
    kernels into bytes."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "reps": reps}), flush=True)


def device_copy_gbs(nbytes=1 << 30, reps=5):
    """Rate of a device-to-device hipMemcpyAsync of 1 GiB (read + write bytes
    per second, timed with HIP events), reported beside the roofline's
    nominal peak as a measured streaming rate of this device.  Calls the HIP
    runtime the library loaded (found in the process's mappings), not
    torch's own copy of it."""
    import ctypes
    import phafd_amd as P
    P.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    if path is None:
        return None
    hip = ctypes.CDLL(path)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    n = ctypes.c_size_t(nbytes)
    try:
        if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
            return None
        if hip.hipEventCreate(ctypes.byref(e0)) or hip.hipEventCreate(ctypes.byref(e1)):
            return None
        hip.hipMemcpyAsync(b, a, n, 3, None)   # hipMemcpyDeviceToDevice, null stream
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, n, 3, None)
        hip.hipEventRecord(e1, None)
        if hip.hipEventSynchronize(e1):
            return None
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9 if ms.value > 0 else None
    finally:
        for e in (e0, e1):
            if e.value:
                hip.hipEventDestroy(e)
        for p in (a, b):
            if p.value:
                hip.hipFree(p)


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/<kernel>_<workload>_pmc.md), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


DATA = os.path.join(ROOT, "phafd_amd", "data")
SURROGATE = os.path.join(DATA, "surrogate_gaussian64.npz")          # K = 100, trained at 64^3
SURROGATE_K10 = os.path.join(DATA, "surrogate_gaussian128_k10.npz")  # K = 10, trained at 128^3
SURROGATE_LJISH = os.path.join(DATA, "surrogate_ljish64.npz")       # K = 100, LJish coupling, 64^3
SURROGATE_LJISH_K10 = os.path.join(DATA, "surrogate_ljish128_k10.npz")  # K = 10, LJish coupling, 128^3


def surrogate_networks(wl):
    """Committed networks for the workload's grid-atom coupling (the first
    is the headline one)."""
    names = [SURROGATE_LJISH_K10, SURROGATE_LJISH] if wl["ljish"] else [SURROGATE, SURROGATE_K10]
    return [p for p in names if os.path.exists(p)]


def surrogate_leg(args, wl, phi, xyz, mol, weights):
    """The surrogate-accelerated form on the same workload and initial state:
    A steps (accuracy_steps, a multiple of every network's K; default 1000)
    run by the plain FP64 path once and, per network, by blocks of K
    (network + check, plain fallback), all timed; phi and the bead positions
    compared at the end.  One result per network."""
    import phafd_amd as P
    Ks = [int(np.load(w)["K"][0]) for w in weights]
    L = int(np.lcm.reduce(Ks))
    A = args.accuracy_steps if args.accuracy_steps else args.steps
    A = max(L, (A + L - 1) // L * L)
    n = wl["n"]

    def system():
        s = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
        configure(s, wl, phi, xyz, mol)
        return s

    a = system()
    u0 = a.get_atoms()[1]
    t0 = time.perf_counter()
    a.run(A)
    plain = time.perf_counter() - t0
    ref = a.get_field(P.PHI)
    uref = a.get_atoms()[1]
    a.close()
    return [surrogate_run(P, system, w, A, plain, ref, uref, u0, n) for w in weights]


def surrogate_run(P, system, weights, A, plain, ref, uref, u0, n):
    """A steps of one network's surrogate form from a fresh system, timed,
    against the plain run's phi (ref) and bead positions (uref)."""
    z = np.load(weights)
    K, tol = int(z["K"][0]), float(z["tol"][0])
    trained = int(z["n"][0]) if "n" in z.files else 64
    net = P.UNet(seed=1)
    net.load(weights)
    s = system()
    s.run_surrogate(net, 0, K, tol)   # allocations outside the timed region
    s.probe("unet_conv")
    t0 = time.perf_counter()
    st = s.run_surrogate(net, A, K, tol)
    el = time.perf_counter() - t0
    pms, plaunch, pflops = s.probe(None)
    got = s.get_field(P.PHI)
    beads = bead_error(s.get_atoms()[1], uref, u0)
    s.close()
    net.close()
    conv_tflops = pflops / (pms * 1e-3) / 1e12 if plaunch else None
    return {"network": os.path.basename(weights), "trained_grid": [trained] * 3,
            "value": n ** 3 * A / el, "unit": "grid-pt·steps/s", "ms_per_step": 1e3 * el / A, "steps": A,
            "plain_fp64_ms_per_step_same_run": 1e3 * plain / A, "speedup_vs_fp64": plain / el,
            f"phi_rel_l2_vs_fp64_after_{A}_steps": rel_err(got, ref), "K": K, "tol": tol,
            "blocks": int(st["blocks"]), "accepted": int(st["accepted"]), "fallbacks": int(st["fallbacks"]),
            "max_residual": st["max_residual"], f"bead_position_err_vs_fp64_after_{A}_steps": beads,
            "ms_network_path_per_block": st["ms_network_path"] / max(st["blocks"], 1),
            "roofline": {"bound": "mfma", "kernel": "k_conv<16,16,27> (3^3 conv 16->16, full resolution, BF16 MFMA)",
                         "achieved": conv_tflops, "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": conv_tflops / 2500.0 if conv_tflops else None,
                         "avg_launch_ms": pms / max(plaunch, 1), "launches": plaunch,
                         "algorithmic_flops_per_launch": pflops / max(plaunch, 1)},
            "form": "per block of K steps: atoms K steps with the block-start grid-atom force held, bead densities "
                    "at both ends, exponential-Euler spectral base + BF16 U-Net correction, K-step residual check, "
                    "plain fallback; FP64 fields"}


def dump_leg(args, wl, phi, xyz, mol, plain_ms):
    """The workload with the reference example's dumps (grid: phi, chempot,
    gradphi_x/y/z; atom/owned: x ux ix F ID molID type) every K steps,
    written synchronously and asynchronously (the snapshot in HBM, the copy
    to the host and the file write overlapping the next steps)."""
    import shutil
    import tempfile
    import phafd_amd as P
    from phafd_amd.dump import Dump, run_with_dumps
    n = wl["n"]
    res = {"every": args.dump_every, "steps": args.steps, "ms_per_step_no_dumps": plain_ms}
    for mode in ("sync", "async"):
        d = tempfile.mkdtemp(prefix="phafd_dump_", dir=args.dump_dir)
        try:
            s = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
            configure(s, wl, phi, xyz, mol)
            s.run(args.warmup)
            grid = Dump(s, "realgrid", "grid", os.path.join(d, "field"), args.dump_every,
                        ["phi", "chempot", "gradphi_x", "gradphi_y", "gradphi_z"], async_=(mode == "async"))
            atoms = Dump(s, "atoms", "atom/owned", os.path.join(d, "atom"), args.dump_every,
                         ["x", "ux", "ix", "F", "ID", "molID", "type"], async_=(mode == "async"))
            for dm in (grid, atoms):
                dm.write_collection_header()
            # one untimed dump of each kind: the pinned host buffers and the
            # device stages are allocated once per run, at its first dump
            for dm in (grid, atoms):
                dm.write_collection_middle(args.warmup, 1e-3)
            s.dump_wait()
            t0 = time.perf_counter()
            run_with_dumps(s, args.steps, [grid, atoms], 1e-3, first_step=args.warmup)
            s.dump_wait()
            el = time.perf_counter() - t0
            for dm in (grid, atoms):
                dm.write_collection_footer()
            files = [os.path.join(d, f) for f in os.listdir(d)]
            res[f"ms_per_step_{mode}"] = 1e3 * el / args.steps
            res["bytes_written"] = sum(os.path.getsize(f) for f in files)
            if mode == "sync":   # one dump of each kind alone, buffers allocated
                t1 = time.perf_counter()
                s.dump_grid(os.path.join(d, "one.vti"), grid.attributes)
                t2 = time.perf_counter()
                s.dump_atoms(os.path.join(d, "one.vtp"), atoms.attributes)
                t3 = time.perf_counter()
                res["grid_dump_ms"] = 1e3 * (t2 - t1)
                res["atom_dump_ms"] = 1e3 * (t3 - t2)
            s.close()
        finally:
            shutil.rmtree(d, ignore_errors=True)
    return res


def diagnostics_leg(args, wl, phi, xyz, mol):
    """The computes beside the step on the same workload: grid clusters of
    phi (union-find), compute complex of ft_phi at in_fourier, and
    run_until_touching's per-step cost."""
    import phafd_amd as P
    n = wl["n"]
    s = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
    configure(s, wl, phi, xyz, mol)
    s.run(args.warmup)
    out = {}
    lab = s.compute_grid_clusters(wl["phi0"], "lt")   # allocates the work arrays
    t = time.perf_counter()
    for _ in range(5):
        lab = s.compute_grid_clusters(wl["phi0"], "lt")
    out["grid_clusters_ms_incl_d2h"] = 1e3 * (time.perf_counter() - t) / 5
    out["clusters"] = int(len(np.unique(lab[lab >= 0])))
    s.fourier_capture(True)
    s.run(1)
    s.compute_complex(P.PHIHAT, P.CPX_NORM)
    t = time.perf_counter()
    for _ in range(5):
        s.compute_complex(P.PHIHAT, P.CPX_NORM)
    out["compute_complex_ms_incl_d2h"] = 1e3 * (time.perf_counter() - t) / 5
    s.fourier_capture(False)
    t = time.perf_counter()
    r = s.run_until_touching(args.steps, 0.0, continue_running=True)
    out["run_until_touching_ms_per_step"] = 1e3 * (time.perf_counter() - t) / r["steps"]
    s.close()
    return out


def bench_batch(args):
    """Batched replicas: the batch timed over `steps`, then the same members
    run one after another as solo systems (the state of every member is
    compared bit for bit with its solo run), then a few eager steps of the
    batch under the kernel probe for the roofline."""
    import phafd_amd as P
    bw = BATCH
    n, box = bw["n"], bw["box"]
    members = batch_members(bw)
    b = P.Batch(n, n, n, [0, 0, 0], [box] * 3, len(members), zghost=bw["zghost"])
    for i, (wl, phi, xyz, mol) in enumerate(members):
        configure(b.member(i), wl, phi, xyz, mol, setup=False)
    b.setup(1e-3)
    info = b.info()
    b.run(args.warmup)
    t0 = time.perf_counter()
    b.run(args.steps)
    elapsed = time.perf_counter() - t0
    got = []
    for i in range(len(members)):
        m = b.member(i)
        got.append((m.get_field(P.PHI), m.get_atoms()[0], m.get_forces()))
    # the other pair-phase mode (one launch per pass for every member vs
    # per-member launches on the member streams), timed over as many steps
    # both modes at the same point of the run: blocks of 10 steps ordered
    # A B B A (a drift of the step time over the blocks cancels), the first
    # step of each block untimed
    fi = b.fused_info()
    mode_ms = {0: 0.0, 1: 0.0}
    for mode in (1, 0, 0, 1):
        b.set_fused(mode)
        b.run(1)
        t1 = time.perf_counter()
        b.run(10)
        mode_ms[mode] += 1e3 * (time.perf_counter() - t1) / 20
    b.set_fused(1 if fi["fused"] else 0)
    b.set_graphs(False)
    m0 = b.member(0)
    m0.probe(args.probe)
    b.run(5)
    pms, plaunch, pbytes = m0.probe(None)
    b.close()

    solo = 0.0
    same = 0
    for i, (wl, phi, xyz, mol) in enumerate(members):
        # timed: the solo run as it runs by default; compared bit for bit:
        # the same run taking every step's gradient from phi, as the batch does
        s = P.System(n, n, n, [0, 0, 0], [box] * 3, zghost=bw["zghost"])
        configure(s, wl, phi, xyz, mol)
        s.run(args.warmup)
        t1 = time.perf_counter()
        s.run(args.steps)
        solo += time.perf_counter() - t1
        s.close()
        s = P.System(n, n, n, [0, 0, 0], [box] * 3, zghost=bw["zghost"])
        configure(s, wl, phi, xyz, mol)
        s.set_option(P.OPT_NEXT_GRADIENT, 0)
        s.run(args.warmup + args.steps)
        ref = (s.get_field(P.PHI), s.get_atoms()[0], s.get_forces())
        same += all(np.array_equal(a, r) for a, r in zip(got[i], ref))
        s.close()

    N = n ** 3 * len(members)
    beads = [wl["nchains"] * wl["nbeads"] for wl, _, _, _ in members]
    avg_ms = pms / max(plaunch, 1)
    bytes_per_launch = pbytes / max(plaunch, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if plaunch else None
    peak = 8.0e3
    out = {
        "metric": METRIC,
        "value": N * args.steps / elapsed,
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": DESCR["c4"], "members": len(members), "grid": [n, n, n],
                   "beads_total": int(sum(beads)), "beads_min": int(min(beads)), "beads_max": int(max(beads)),
                   "members_per_launch": info["members_per_launch"], "member_streams": info["member_streams"],
                   "autotune_grad_ms": info["tune_grad_ms"], "autotune_pair_ms": info["tune_pair_ms"]},
        "beads_steps_per_s": sum(beads) * args.steps / elapsed,
        "sequential_solo": {"value": N * args.steps / solo, "ms_per_step": 1e3 * solo / args.steps,
                            "batch_speedup": solo / elapsed},
        "members_bit_identical_to_solo": f"{same}/{len(members)}",
        "pair_phase": {"mode": "one launch per pass for every member" if fi["fused"] else "per-member launches",
                       "picked_by": "autotuner: whole steps 16-47 timed, the modes in blocks of 8 ordered A B B A",
                       "step_ms_per_member_launches_trial": fi["step_ms_per_member"],
                       "step_ms_one_launch_per_pass_trial": fi["step_ms_one_launch_per_pass"],
                       "setup_ms_one_launch_per_pass": fi["tune_fused_ms"],
                       "setup_ms_per_member_launches": fi["tune_member_ms"],
                       "tuned_shape": {"atoms_per_item": fi["atoms_per_item"], "grid_cap": fi["grid_cap"],
                                       "setup_ms_by_shape": fi["tune_shape_ms"]},
                       "ms_per_step_abba": {"one_launch_per_pass": mode_ms[1], "per_member_launches": mode_ms[0]}},
        "roofline": {"bound": "hbm", "kernel": args.probe, "achieved": achieved, "peak": peak,
                     "unit": "GB/s", "frac": (achieved / peak) if achieved else None, "traffic": None,
                     "avg_launch_ms": avg_ms, "launches": plaunch,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "note": "eager batched launches (graphs off) after the timed run"},
    }
    if not args.no_cpu_baseline:
        wl, phi, xyz, mol = members[0]
        t, th = cpu_baseline(wl, phi, xyz, mol, args.cpu_steps)
        out["cpu_baseline"] = cpu_baseline_record(n ** 3 * args.cpu_steps / t, t, th, args.cpu_steps,
                                                  f"member 0 ({beads[0]} beads)")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS) + ["c4"])
    ap.add_argument("--probe", default=None,
                    help="kernel timed with HIP events for the roofline (default: the step's longest kernel, "
                         "ga_force; c2r_x for the batched workload)")
    ap.add_argument("--fft-probe", default="z_update",
                    help="FFT-side kernel also timed, over extra steps after the timed region")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (counter calibration) and exit")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-form leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--no-noise", action="store_true", help="skip the thermal-noise leg")
    ap.add_argument("--surrogate-net", action="append", default=None, metavar="NPZ",
                    help="network(s) of the surrogate leg (default: the committed ones for the workload's coupling)")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the timed plain-vs-surrogate comparison (0: --steps); rounded up to a multiple "
                         "of the network's K")
    ap.add_argument("--timers", action="store_true", help="print per-hook GPU times to stderr")
    ap.add_argument("--dump-every", type=int, default=0, metavar="K",
                    help="also time the run with grid and atom dumps every K steps (sync and async)")
    ap.add_argument("--dump-dir", default=None, help="directory for the dump leg's files (default: system temp)")
    ap.add_argument("--diagnostics", action="store_true",
                    help="also time grid clusters, compute complex and run_until_touching on the workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not built", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 5)
        return

    if args.workload == "c4":
        args.probe = args.probe or "c2r_x"
        bench_batch(args)
        return
    args.probe = args.probe or "ga_force"

    import phafd_amd as P
    wl = WORKLOADS[args.workload]
    n = wl["n"]
    phi, xyz, mol = setup_inputs(wl)
    s = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
    configure(s, wl, phi, xyz, mol)
    s.run(args.warmup)
    s.timers(reset=True)
    s.probe(args.probe)
    t0 = time.perf_counter()
    s.run(args.steps)          # every step ends with a stream synchronise (NaN check)
    elapsed = time.perf_counter() - t0
    pms, plaunch, pbytes = s.probe(None)
    timers = s.timers()
    counts = s.counts()
    ref = snapshot(s, P)
    fft_roof = None
    if args.fft_probe and args.fft_probe != "none":
        # the longest FFT-side kernel, over 20 further steps (outside the
        # timed region; the fields compared below are those of the timed run)
        s.probe(args.fft_probe)
        s.run(20)
        fms, flaunch, fbytes = s.probe(None)
        if flaunch:
            favg = fms / flaunch
            fach = fbytes / flaunch / (favg * 1e-3) / 1e9
            fft_roof = {"bound": "hbm", "kernel": args.fft_probe, "achieved": fach, "peak": 8.0e3, "unit": "GB/s",
                        "frac": fach / 8.0e3, "avg_launch_ms": favg, "launches": flaunch,
                        "algorithmic_bytes_per_launch": fbytes / flaunch,
                        "traffic": read_traffic(args.fft_probe, args.workload)}
    s.close()

    N = n ** 3
    fp32 = None
    if not args.no_fp32:
        # the same workload, steps and initial state in the FP32 form; its phi
        # error is taken against the FP64 run above
        s32 = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"], precision=P.FP32)
        configure(s32, wl, phi, xyz, mol)
        s32.run(args.warmup)
        t1 = time.perf_counter()
        s32.run(args.steps)
        el32 = time.perf_counter() - t1
        got = snapshot(s32, P)
        s32.close()
        err = {k: rel_err(got[k], ref[k]) for k in ref}
        fp32 = {"value": N * args.steps / el32, "unit": "grid-pt·steps/s", "ms_per_step": 1e3 * el32 / args.steps,
                "speedup_vs_fp64": elapsed / el32, "phi_rel_l2_vs_fp64": err["phi"],
                "rel_err_vs_fp64": err, "steps_compared": args.warmup + args.steps,
                "form": "FP32 FFT + fields, Kahan-compensated grid-atom sums, FP64 k=0 mode, atoms and energies"}

    # thermal noise on (grid/conjugate/volfrac temp 0.01, generated on the
    # device per update; at temp 1 the noise of a unit-spacing grid, about
    # 0.14 per point per step at dt 1e-3, drives phi out of [-1, 2] and the
    # FH logarithm to NaN): the same workload, timed over as many steps
    noise_leg = None
    if not args.no_noise:
        sn = P.System(n, n, n, [0, 0, 0], [wl["box"]] * 3, zghost=wl["zghost"])
        configure(sn, wl, phi, xyz, mol, noise=P.NOISE_DEVICE, noise_temp=0.01)
        sn.run(args.warmup)
        t2 = time.perf_counter()
        sn.run(args.steps)
        eln = time.perf_counter() - t2
        sn.close()
        noise_leg = {"ms_per_step": 1e3 * eln / args.steps, "value": N * args.steps / eln, "unit": "grid-pt·steps/s",
                     "form": "device Philox noise (PHAFD_NOISE_DEVICE), grid/conjugate/volfrac temp 0.01, mobility 1"}

    sur = None
    nets = args.surrogate_net if args.surrogate_net else surrogate_networks(wl)
    if not args.no_surrogate and nets:
        sur = surrogate_leg(args, wl, phi, xyz, mol, nets)

    nbeads = wl["nchains"] * wl["nbeads"]
    ms_per_step = 1e3 * elapsed / args.steps
    peak = 8.0e3   # GB/s, MI355X HBM3E
    avg_ms = pms / max(plaunch, 1)
    bytes_per_launch = pbytes / max(plaunch, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if plaunch else None
    traffic = read_traffic(args.probe, args.workload)
    copy_gbs = device_copy_gbs()
    out = {
        "metric": METRIC,
        "value": N * args.steps / elapsed,
        "unit": "grid-pt·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": DESCR[args.workload], "grid": [n, n, n], "beads": nbeads,
                   "dt": 1e-3, "neighbor_builds_in_run": counts["ncalls"]},
        "beads_steps_per_s": nbeads * args.steps / elapsed,
        "roofline": {"bound": "hbm", "kernel": args.probe, "achieved": achieved, "peak": peak,
                     "unit": "GB/s", "frac": (achieved / peak) if achieved else None,
                     "traffic": traffic, "avg_launch_ms": avg_ms, "launches": plaunch,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "hipmemcpy_d2d_GBs": copy_gbs,
                     "frac_of_hipmemcpy_d2d": (achieved / copy_gbs) if (achieved and copy_gbs) else None},
    }
    if fft_roof is not None:
        out["roofline_fft"] = fft_roof
    if noise_leg is not None:
        out["noise_device"] = noise_leg
    if fp32 is not None:
        out["fp32"] = fp32
    if sur is not None:
        out["surrogate"] = sur[0]
        if len(sur) > 1:
            out["surrogate_other_networks"] = sur[1:]
    if args.dump_every > 0:
        out["dumps"] = dump_leg(args, wl, phi, xyz, mol, ms_per_step)
    if args.diagnostics:
        out["diagnostics"] = diagnostics_leg(args, wl, phi, xyz, mol)
    if args.timers:
        print(json.dumps({"per_hook_ms_total": timers}), file=sys.stderr)
    if not args.no_cpu_baseline:
        t, th = cpu_baseline(wl, phi, xyz, mol, args.cpu_steps)
        out["cpu_baseline"] = cpu_baseline_record(N * args.cpu_steps / t, t, th, args.cpu_steps, "the same workload")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
