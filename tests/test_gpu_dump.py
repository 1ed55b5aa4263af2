"""Dumps and restart of the GPU state, through the C ABI.

  grid dump ........ the file of phafd_hip_dump_grid equals the host writer's
                     file of the same fields read back with get_field (whose
                     bytes equal the oracle's, tests/test_dump_format.py),
                     FP64 and FP32 forms
  atom dump ........ likewise for the owned atoms, %.17e
  asynchronous ..... a dump overlapping later steps writes the snapshot of its
                     call, and the steps are bit-identical to a run without it
  restart .......... fields and atoms read back from the dumps continue the
                     run: 1e-10 of the uninterrupted run (the rebuilt lists
                     reorder sums)
  dump command ..... the collection, file names and per-step files of the
                     dump command over a run
"""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import phafd_amd as P
from phafd_amd.dump import Dump, run_with_dumps
from oracle import oracle_ctypes as O
from tests.helpers import workload, make_gpu, apply, rel_l2

pytestmark = pytest.mark.gpu
GRID_NAMES = ["phi", "chempot", "gradphi_x", "gradphi_y", "gradphi_z"]
FIELDS = [P.PHI, P.MU, P.GRADX, P.GRADY, P.GRADZ]
ATOM_NAMES = ["x", "ux", "ix", "F", "ID", "molID", "type"]


def _read(p):
    with open(p, "rb") as f:
        return f.read()


def _host_grid(s, path, w):
    spacing = [w["box"] / w["n"]] * 3
    P.vtk_write_grid(path, [s.get_field(f) for f in FIELDS], GRID_NAMES, [w["origin"]] * 3, spacing)


def _host_atoms(s, path, w, precision=17, fmt="e"):
    x, ux, img = s.get_atoms()
    n = len(x)
    vals = {"x": x, "ux": ux, "ix": img, "F": s.get_forces(), "ID": np.arange(n, dtype=np.int32),
            "molID": w["mol"], "type": w["types"]}
    P.vtk_write_atoms(path, [vals[k] for k in ATOM_NAMES], ATOM_NAMES, precision, fmt)


@pytest.mark.parametrize("precision", [P.FP64, P.FP32])
def test_grid_and_atom_dumps_match_host_writer(tmp_path, precision):
    w = workload(n=32, nchains=8, nbeads=20, seed=21, zghost=1, origin=-16.0)
    s = make_gpu(w, precision)
    s.run(15)
    a, b = str(tmp_path / "g.vti"), str(tmp_path / "h.vti")
    s.dump_grid(a, GRID_NAMES)
    _host_grid(s, b, w)
    assert _read(a) == _read(b)
    a, b = str(tmp_path / "g.vtp"), str(tmp_path / "h.vtp")
    s.dump_atoms(a, ATOM_NAMES, 17, "e")
    _host_atoms(s, b, w)
    assert _read(a) == _read(b)
    s.dump_atoms(a, ["ID", "x", "label", "type"])   # default format, another order
    got = P.read_dump_atoms(a, len(w["xyz"]), ["x", "ID", "type", "label"], molecule=False)
    assert np.array_equal(got["label"], np.zeros(len(w["xyz"]), np.int32))
    assert np.array_equal(got["type"], w["types"])
    with pytest.raises(P.PhafdError, match="Attribute does not exist"):
        s.dump_grid(a, ["phi", "density"])


def test_async_dumps_overlap_steps_without_changing_them(tmp_path):
    w = workload(n=32, nchains=8, nbeads=20, seed=22, zghost=1)
    s, t = make_gpu(w), make_gpu(w)
    s.run(5)
    t.run(5)
    sync_g, async_g = str(tmp_path / "s.vti"), str(tmp_path / "a.vti")
    sync_a, async_a = str(tmp_path / "s.vtp"), str(tmp_path / "a.vtp")
    s.dump_grid(sync_g, GRID_NAMES)
    s.dump_atoms(sync_a, ATOM_NAMES, 17, "e")
    s.dump_grid(async_g, GRID_NAMES, async_=True)
    s.run(10)                        # runs while the copy and the write proceed
    s.dump_atoms(async_a, ATOM_NAMES, 17, "e", async_=True)   # beside the grid dump if still in flight
    s.run(10)
    s.dump_wait()
    t.run(20)
    assert _read(sync_g) == _read(async_g)
    assert np.array_equal(s.get_field(P.PHI), t.get_field(P.PHI))
    assert np.array_equal(s.get_atoms()[0], t.get_atoms()[0])
    ref = str(tmp_path / "r.vtp")
    u = make_gpu(w)
    u.run(15)
    _host_atoms(u, ref, w)
    assert _read(async_a) == _read(ref)
    assert _read(sync_a) != _read(ref)


def test_restart_from_dumps_continues_the_run(tmp_path):
    w = workload(n=32, nchains=10, nbeads=25, seed=23, zghost=1)
    a = make_gpu(w)
    a.run(30)
    g, at = str(tmp_path / "r.vti"), str(tmp_path / "r.vtp")
    a.dump_grid(g, ["phi", "chempot"])
    a.dump_atoms(at, ATOM_NAMES, 17, "e")
    phi30 = a.get_field(P.PHI)
    d = P.read_dump_atoms(at, len(w["xyz"]), ["x", "type", "ID", "molID", "ix"])
    w2 = dict(w, xyz=d["x"], types=d["type"], mol=d["molID"], phi=np.zeros_like(w["phi"]))
    b = P.System(w["n"], w["n"], w["n"], [w["origin"]] * 3, [w["box"]] * 3, zghost=w["zghost"])
    apply(b, w2, False, setup=False)
    b.read_dump_grid(g, ["phi"])
    assert np.array_equal(b.get_field(P.PHI), phi30)
    b.set_images(d["ix"])
    b.setup(w["dt"])
    # setup wraps atoms that left the box since the last rebuild (and bumps
    # their images), so positions agree modulo the period and unwrapped ones
    # agree to rounding
    def wrapped_diff(p, q):
        d = p - q
        return np.abs(d - w["box"] * np.round(d / w["box"])).max()
    x, ux, img = b.get_atoms()
    ax, aux, aimg = a.get_atoms()
    assert wrapped_diff(x, ax) < 1e-12
    assert np.allclose(ux, aux, rtol=0, atol=1e-12)
    a.run(40)
    b.run(40)
    assert rel_l2(b.get_field(P.PHI), a.get_field(P.PHI)) < 1e-10
    assert wrapped_diff(b.get_atoms()[0], a.get_atoms()[0]) < 1e-9
    assert rel_l2(b.get_atoms()[1], a.get_atoms()[1]) < 1e-10


@pytest.mark.parametrize("async_", [False, True])
def test_dump_command_over_a_run(tmp_path, async_):
    w = workload(n=32, nchains=6, nbeads=20, seed=24)
    s = make_gpu(w)
    base = str(tmp_path / "out" / "run1field")
    os.makedirs(os.path.dirname(base))
    grid = Dump(s, "realgrid", "grid", base, 10, GRID_NAMES, async_=async_)
    atoms = Dump(s, "atoms", "atom/owned", base + "atom", 15, ATOM_NAMES, decimalplaces=8, fmt="e", async_=async_)
    for d in (grid, atoms):
        d.write_collection_header()
        d.write_collection_middle(0, w["dt"])
    end = run_with_dumps(s, 30, [grid, atoms], w["dt"])
    for d in (grid, atoms):
        d.write_collection_footer()
    assert end == 30
    root = ET.parse(base + ".pvd").getroot()
    files = [e.get("file") for e in root.iter("DataSet")]
    assert files == [f"run1field_{t}.pvti" for t in (0, 10, 20, 30)]
    assert [e.get("timestep") for e in root.iter("DataSet")] == ["0", "0.01", "0.02", "0.03"]
    for t in (0, 10, 20, 30):
        assert os.path.exists(f"{base}_p0_{t}.vti")
        src = ET.parse(f"{base}_{t}.pvti").getroot().find("PImageData/Piece").get("Source")
        assert src == f"run1field_p0_{t}.vti"
    for t in (0, 15, 30):
        assert os.path.exists(f"{base}atom_p0_{t}.vtp") and os.path.exists(f"{base}atom_{t}.pvtp")
    ref = str(tmp_path / "ref.vti")
    _host_grid(s, ref, w)
    assert _read(ref) == _read(f"{base}_p0_30.vti")


def test_example_dumps_with_compute_pair_against_oracle(tmp_path):
    # the reference example's dump lines: an atom/owned dump holding
    # c_droppolyforces (compute pair of the gridatom/LJish/linear style) and a
    # grid dump of phi and chempot at the same step; the compute runs before
    # the dumps, so the dumped chempot carries its second deposit
    from tests.helpers import make_oracle
    w = workload(n=32, nchains=8, nbeads=20, seed=25, ljish="linear", eps=0.2)
    s, o = make_gpu(w), make_oracle(w)
    base = str(tmp_path / "run")
    attrs = ["x", "ux", "ix", "F", "c_droppolyforces", "ID", "molID", "type"]
    grid = Dump(s, "realgrid", "grid", base + "field", 10, ["phi", "chempot"], async_=True)
    atoms = Dump(s, "atoms", "atom/owned", base + "atom", 10, attrs, decimalplaces=17, fmt="e",
                 computes={"droppolyforces": 0})
    for d in (grid, atoms):
        d.write_collection_header()
    run_with_dumps(s, 20, [grid, atoms], w["dt"])
    s.dump_wait()
    o.run(20)
    f_o = o.compute_pair(0)
    piece = ET.parse(f"{base}atom_p0_20.vtp").getroot().find("PolyData/Piece")
    arrays = {e.get("Name"): e for e in piece.iter("DataArray")}
    f_g = np.array(arrays["c_droppolyforces"].text.split(), dtype=np.float64).reshape(-1, 3)
    assert arrays["c_droppolyforces"].get("NumberOfComponents") == "3"
    assert rel_l2(f_g, f_o) < 1e-10
    x_g = np.array(arrays["x"].text.split(), dtype=np.float64).reshape(-1, 3)
    assert rel_l2(x_g, o.get_atoms()[0]) < 1e-10
    phi, mu = P.vtk_read_grid(f"{base}field_p0_20.vti", s.shape, ["phi", "chempot"])
    assert rel_l2(phi, o.get_field(P.PHI)) < 1e-10
    assert rel_l2(mu, o.get_field(P.MU)) < 1e-10


def test_dumps_and_grid_ave_in_one_run(tmp_path):
    # a grid dump every 5 steps beside fix grid/ave every 2 repeat 3 freq 10:
    # the average equals the one of a run with the average alone, the dumped
    # phi that of a run with the dumps alone
    from phafd_amd.computes import ComputeComplex, FixGridAve, run_with_aves
    w = workload(n=32, nchains=6, nbeads=20, seed=26)
    a, b, c = make_gpu(w), make_gpu(w), make_gpu(w)
    ave_a = FixGridAve(a, "ave", 2, 3, 10, ComputeComplex(a, "sq", "ft_phi", "modulus"))
    grid_a = Dump(a, "g", "grid", str(tmp_path / "a"), 5, ["phi"], async_=True)
    grid_a.write_collection_header()
    run_with_dumps(a, 20, [grid_a], w["dt"], aves=[ave_a])
    a.dump_wait()
    ave_b = FixGridAve(b, "ave", 2, 3, 10, ComputeComplex(b, "sq", "ft_phi", "modulus"))
    run_with_aves(b, 20, [ave_b])
    assert np.array_equal(ave_a.array, ave_b.array)
    grid_c = Dump(c, "g", "grid", str(tmp_path / "c"), 5, ["phi"])
    grid_c.write_collection_header()
    run_with_dumps(c, 20, [grid_c], w["dt"])
    c.dump_wait()   # dumps are asynchronous by default: the last one may still be writing
    for t in (5, 10, 15, 20):
        assert _read(str(tmp_path / f"a_p0_{t}.vti")) == _read(str(tmp_path / f"c_p0_{t}.vti"))


def test_lammpsatom_dump_over_a_run(tmp_path):
    # dump <name> lammpsatom: one text file, a block per dumped step, values
    # against the oracle's state at the same steps
    from tests.helpers import make_oracle
    w = workload(n=32, nchains=6, nbeads=20, seed=27, ljish="linear", eps=0.2)
    s, o = make_gpu(w), make_oracle(w)
    attrs = ["x", "F", "c_pf", "ID", "type"]
    d = Dump(s, "lmp", "lammpsatom", str(tmp_path / "traj"), 10, attrs, decimalplaces=17, fmt="e",
             computes={"pf": 0})
    d.write_collection_header()
    run_with_dumps(s, 20, [d], w["dt"])
    d.write_collection_footer()
    text = open(str(tmp_path / "traj_p0.lmmpstrj")).read().split("ITEM: TIMESTEP\n")[1:]
    assert [blk.split("\n")[0] for blk in text] == ["10", "20"]
    n = len(w["xyz"])
    o.run(20)
    rows = np.array([r.split() for r in text[1].split("\n")[8:8 + n]], dtype=np.float64)
    assert rel_l2(rows[:, 0:3], o.get_atoms()[0]) < 1e-10
    assert rel_l2(rows[:, 3:6], o.get_forces()) < 1e-8
    assert rel_l2(rows[:, 6:9], o.compute_pair(0)) < 1e-10
    assert np.array_equal(rows[:, 9], np.arange(n)) and np.array_equal(rows[:, 10], w["types"])
    assert not (tmp_path / "traj.pvd").exists()


def test_ftgrid_dump_of_computes(tmp_path):
    # dump <name> ftgrid: a compute complex (its step's in_fourier spectrum)
    # and a fix grid/ave over it, in the ft order with the reference's dq
    # spacings; the compute against the oracle at the dumped step
    import re
    import struct
    from tests.helpers import make_oracle
    from phafd_amd.computes import ComputeComplex, FixGridAve
    w = workload(n=32, nchains=6, nbeads=20, seed=28, box=40.0)
    s, o = make_gpu(w), make_oracle(w)
    sq = ComputeComplex(s, "sq", "ft_phi", "norm")
    ave = FixGridAve(s, "sqave", 2, 3, 10, sq)
    d = Dump(s, "ft", "ftgrid", str(tmp_path / "sq"), 10, ["c_sq", "f_sqave"], computes={"sq": sq, "sqave": ave})
    d.write_collection_header()
    run_with_dumps(s, 20, [d], w["dt"], aves=[ave])
    d.write_collection_footer()
    raw = open(str(tmp_path / "sq_p0_20.vti"), "rb").read()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n_')
    dq = [2 * 3.141592 / 40.0] * 3
    assert f'Spacing="{dq[0]:.6g} {dq[1]:.6g} {dq[2]:.6g} "'.encode() in head
    (nb,) = struct.unpack_from("<I", tail, 0)
    got = np.frombuffer(tail, "<f8", nb // 8, 4).reshape(32, 32, 17)
    o.run(19)
    for ph in (O.PH_START, O.PH_NEIGH, O.PH_PAIR, O.PH_INITIAL, O.PH_POSTFORCE, O.PH_PREFINAL):
        o.phase(ph)
    assert rel_l2(got, o.compute_complex(0, 3)) < 1e-10
    (nb2,) = struct.unpack_from("<I", tail, 4 + nb)
    got_ave = np.frombuffer(tail, "<f8", nb2 // 8, 8 + nb).reshape(32, 32, 17)
    assert np.array_equal(got_ave, ave.array)
    assert re.search(rb'DataArray Name="f_sqave"', head)


def test_grid_dump_with_cluster_labels(tmp_path):
    # a grid dump holding phi and a grid/clusters compute (the labels of the
    # dumped step's phi), read back with the format's own layout
    import struct
    from phafd_amd.computes import ComputeGridClusters
    w = workload(n=32, nchains=6, nbeads=20, seed=29, chi=3.0, var=0.05)
    s = make_gpu(w)
    cl = ComputeGridClusters(s, "clusters", "phi", 0.5, "lt")
    d = Dump(s, "g", "grid", str(tmp_path / "cl"), 10, ["phi", "c_clusters"], computes={"clusters": cl})
    d.write_collection_header()
    run_with_dumps(s, 20, [d], w["dt"])
    d.write_collection_footer()
    raw = open(str(tmp_path / "cl_p0_20.vti"), "rb").read()
    tail = raw.partition(b'<AppendedData encoding="raw">\n_')[2]
    (nb,) = struct.unpack_from("<I", tail, 0)
    phi = np.frombuffer(tail, "<f8", nb // 8, 4).reshape(32, 32, 32)
    lab = np.frombuffer(tail, "<f8", nb // 8, 8 + nb).reshape(32, 32, 32)
    assert np.array_equal(phi, s.get_field(P.PHI))
    assert np.array_equal(lab, s.compute_grid_clusters(0.5, "lt"))
    assert np.array_equal(lab >= 0, phi < 0.5)
