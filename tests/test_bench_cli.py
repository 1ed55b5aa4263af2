"""The benchmark's command line and the committed surrogate networks, on the
CPU: one GPU only (a request for more exits with a message), and every
network the bench deploys exists with the fields the loader reads."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_committed_networks_per_coupling():
    for ljish, names in ((False, ["surrogate_gaussian64.npz", "surrogate_gaussian128_k10.npz"]),
                         (True, ["surrogate_ljish128_k10.npz", "surrogate_ljish64.npz"])):
        nets = bench.surrogate_networks(dict(bench.WORKLOADS["c2"], ljish=ljish))
        assert [os.path.basename(p) for p in nets] == names
        for p in nets:
            z = np.load(p)   # no pickles: plain arrays only
            assert {"params", "norm", "tol", "K"} <= set(z.files)
            assert z["norm"].shape == (5,) and np.all(np.isfinite(z["params"]))
            K, tol = int(z["K"][0]), float(z["tol"][0])
            assert K in (10, 100) and 0 < tol < 1
            if "mix" in z.files:
                assert str(z["mix"][0]) == ("ljish" if ljish else "gaussian")
