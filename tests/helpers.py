"""Shared workload definitions for the parity tests.

A workload is a plain dict; `apply` configures either the CPU oracle
(oracle.oracle_ctypes.System) or the HIP path (phafd_amd.System) with the
same inputs, the way a reference input script would (src/input.cpp:48-299).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle_ctypes as O  # noqa: E402


def chains(nchains, nbeads, box, seed, bond=1.0, minsep=0.9, origin=0.0):
    """Random-walk chains from the product's host generator (input data only)."""
    import phafd_amd as P
    lo = [origin] * 3
    xyz, mol = P.ic_random_chains(nchains, nbeads, bond, minsep, lo, [box] * 3, seed)
    return xyz, mol


def workload(n=32, box=None, nchains=6, nbeads=20, phi0=0.5, var=0.02, chi=3.0, seed=7,
             gauss_cut=3.0, nucwidth=1.0, ljish=False, lj=True, chain=True, drag=False,
             skin=0.5, dt=1e-3, zghost=0, origin=0.0, types=1, noise=0, eps=1.0, harmonic=None):
    box = float(n if box is None else box)
    import phafd_amd as P
    phi = P.ic_noisy_constant((n, n, n), phi0, var, 1234 + seed)
    xyz, mol = chains(nchains, nbeads, box, seed, origin=origin)
    typ = (np.arange(len(xyz)) % types).astype(np.int32)
    starts = np.arange(nchains, dtype=np.int32) * nbeads
    ends = starts + nbeads
    return dict(n=n, box=box, origin=origin, phi=phi, xyz=xyz, mol=mol, types=typ, chi=chi,
                gauss_cut=gauss_cut, nucwidth=nucwidth, ljish=ljish, lj=lj, chain=chain, drag=drag,
                skin=skin, dt=dt, zghost=zghost, starts=starts, ends=ends, ntypes=types, noise=noise, eps=eps,
                harmonic=harmonic)


def apply(sysobj, w, is_oracle, setup=True):
    s = sysobj
    s.set_field(0, w["phi"])
    s.set_atoms(w["xyz"], w["types"], tags=np.arange(len(w["xyz"]), dtype=np.int32), mols=w["mol"])
    # ljish: False (gaussian), True (LJish) or "linear" (LJish/linear)
    style = 3 if w["ljish"] == "linear" else (1 if w["ljish"] else 0)
    if is_oracle:
        ip = s.add_pair(style, w["gauss_cut"])
    else:
        ip = s.pair_gridatom(style, w["gauss_cut"])
    for t in range(w["ntypes"]):
        s.pair_coeff(ip, t, w["eps"] * (1.0 + 0.5 * t), 0.1, w["nucwidth"])
    if w["lj"]:
        if is_oracle:
            s.add_pair(O.PAIR_LJCUT, 1.122462048, 1.0, 1.0)
        else:
            s.pair_ljcut(1.122462048, 1.0, 1.0)
    if w["harmonic"]:   # (cut, spring constant) of pair harmonic/cut
        if is_oracle:
            s.add_pair(O.PAIR_HARMONIC, *w["harmonic"])
        else:
            s.pair_harmoniccut(*w["harmonic"])
    s.fix_conjugate(5678, 1.0, 1.0, 1.0, 1.0, w["noise"])
    s.fix_floryhuggins(1.0, 1.0, w["chi"])
    # chain_range / drag_range: the chains (index range) each fix acts on
    if w["chain"]:
        a, b = w.get("chain_range", (0, len(w["starts"])))
        s.fix_chain(w["starts"][a:b], w["ends"][a:b], 1.0, 100.0, 10.0, 1.0)
    if w["drag"]:
        a, b = w.get("drag_range", (0, len(w["starts"])))
        s.fix_drag(w["starts"][a:b], w["ends"][a:b], 1.0)
    s.neighbor(w["skin"])
    if setup:
        s.setup(w["dt"])
    return s


def make_oracle(w):
    s = O.System(w["n"], w["n"], w["n"], [w["origin"]] * 3, [w["box"]] * 3, zghost=w["zghost"])
    return apply(s, w, True)


def make_gpu(w, precision=None):
    import phafd_amd as P
    s = P.System(w["n"], w["n"], w["n"], [w["origin"]] * 3, [w["box"]] * 3, zghost=w["zghost"],
                 precision=P.FP64 if precision is None else precision)
    return apply(s, w, False)


def rel_l2(a, b):
    # complex inputs (spectra) are compared as complex: both parts count
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b)
    a = np.asarray(a, dtype=np.complex128 if cplx else np.float64)
    b = np.asarray(b, dtype=np.complex128 if cplx else np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))
