"""Pins of the oracle's run_until_touching (CPU): the event bookkeeping of
src/integrate.cpp:355-395 against closed forms (a pair of beads always or
never within touch_point), and its distances and step numbers against the
positions read back after single steps."""
import numpy as np

from oracle import oracle_ctypes as O
from tests.helpers import workload, make_oracle


def _pair(sep):
    s = O.System(16, 16, 16, [0, 0, 0], [16.0] * 3)
    s.set_field(O.PHI, np.full((16, 16, 16), 0.5))
    s.set_atoms(np.array([[5.0, 5.0, 5.0], [5.0 + sep, 5.0, 5.0]]), np.zeros(2, np.int32),
                tags=np.arange(2, dtype=np.int32), mols=np.zeros(2, np.int32))
    s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, O.NOISE_OFF)
    s.fix_chain(np.array([0], np.int32), np.array([2], np.int32), 1.0, 10.0, 0.0, 1.0)
    s.neighbor(0.5)
    s.setup(1e-3)
    return s


def test_touching_closed_forms():
    r = _pair(1.0).run_until_touching(20, 1.5)                       # touches at once: stops after step 1
    assert r == {"stopped": True, "steps": 1, "occurred": [1], "ended": []}
    r = _pair(1.0).run_until_touching(20, 0.5)                       # never touches
    assert r == {"stopped": False, "steps": 20, "occurred": [], "ended": []}
    r = _pair(1.0).run_until_touching(6, 1.5, continue_running=True, delay_step=0)
    assert r["occurred"] == [1, 2, 3, 4, 5, 6] and r["ended"] == []  # delay 0: every touching step counts
    r = _pair(1.0).run_until_touching(6, 1.5, continue_running=True, delay_step=3)
    assert r["occurred"] == [1] and r["ended"] == []                 # never apart long enough


def test_touching_events_follow_the_distances():
    w = workload(n=32, nchains=1, nbeads=12, seed=50, lj=True)
    ref = make_oracle(w)
    d = []
    for _ in range(60):
        ref.run(1)
        x = ref.get_atoms()[0]
        d.append(np.linalg.norm(x[-1] - x[0]))
    d = np.array(d)
    tp = float(np.median(d))
    delay = 2
    occ, end, since, brk = [], [], delay, False
    for t, di in enumerate(d, start=1):
        if di < tp:
            if since >= delay:
                occ.append(t)
            since, brk = 0, True
        else:
            if brk and since == delay:
                end.append(t - delay)
            since += 1
    r = make_oracle(w).run_until_touching(60, tp, continue_running=True, delay_step=delay)
    assert r["occurred"] == occ and r["ended"] == end and r["steps"] == 60
    assert len(occ) >= 1
    first = next(t for t, di in enumerate(d, start=1) if di < tp)
    r = make_oracle(w).run_until_touching(60, tp)
    assert r["stopped"] and r["steps"] == first and r["occurred"] == [first]
