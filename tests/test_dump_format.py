"""Dump file formats on the CPU (no GPU work): the oracle's std::ostream
restatement of the reference's writers (src/dump.cpp) pinned by an
independent reading of the files (Python's XML parser, the raw appended
block decoded with numpy, numbers compared with Python's own %g / %f / %e
formatting), then the product's writer (phafd_amd/host/vtk.cpp, std::to_chars)
against the oracle byte for byte, and the product's reader (ReadDump,
src/read_dump.cpp) on the oracle's files."""
import re
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import phafd_amd as P
from oracle import oracle_ctypes as O

GRID_NAMES = ["phi", "chempot", "gradphi_x", "gradphi_y", "gradphi_z"]


def _grid_arrays(shape, seed, n=5):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) * 10.0 ** rng.integers(-3, 4) for _ in range(n)]


def _atoms(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-128, 128, (n, 3))
    x[0] = [0.0, -0.0, 1e-300]
    x[1] = [123456789.125, -1e-7, 5e20]
    F = rng.standard_normal((n, 3)) * 1e3
    ux = x + rng.integers(-2, 3, (n, 3)) * 256.0
    ix = rng.integers(0, 1 << 30, n).astype(np.int32)
    ID = np.arange(n, dtype=np.int32)
    mol = (np.arange(n) // 7).astype(np.int32)
    typ = (np.arange(n) % 3).astype(np.int32) - 1
    return {"x": x, "F": F, "ux": ux, "ix": ix, "ID": ID, "molID": mol, "type": typ,
            "label": np.zeros(n, np.int32)}


def _read(p):
    with open(p, "rb") as f:
        return f.read()


def _parse_vti(path):
    """Header fields and arrays of a .vti with raw appended data."""
    raw = _read(path)
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n_')
    header = head.decode()
    names = re.findall(r'<DataArray Name="([^"]+)" type="Float64" format="appended" offset="(\d+)"/>', header)
    arrays, pos = [], 0
    for name, off in names:
        assert int(off) == pos
        (nbytes,) = struct.unpack_from("<I", tail, pos)
        arrays.append(np.frombuffer(tail, dtype="<f8", count=nbytes // 8, offset=pos + 4))
        pos += 4 + nbytes
    assert tail[pos:] == b"\n</AppendedData>\n</VTKFile>\n"
    return header, [n for n, _ in names], arrays


@pytest.mark.parametrize("shape,origin,spacing", [((8, 6, 4), [-128.0, -128.0, -128.0], [4.0, 4.0, 4.0]),
                                                  ((3, 5, 7), [0.0, 0.5, -1.25], [10 / 3, 0.1, 2.0 ** -20])])
def test_oracle_grid_file_reads_back(tmp_path, shape, origin, spacing):
    arrs = _grid_arrays(shape, 1)
    p = str(tmp_path / "o.vti")
    O.vtk_write_grid(p, arrs, GRID_NAMES, origin, spacing)
    header, names, got = _parse_vti(p)
    assert names == GRID_NAMES
    for a, g in zip(arrs, got):
        assert np.array_equal(a.ravel(), g)
    nz, ny, nx = shape
    g6 = lambda v: f"{v:.6g}"   # noqa: E731 -- the stream default
    assert f'WholeExtent="0 {nx - 1} 0 {ny - 1} 0  {nz - 1}"' in header
    assert f'Origin="{g6(origin[0])} {g6(origin[1])} {g6(origin[2])} "' in header
    assert f'Spacing="{g6(spacing[0])} {g6(spacing[1])} {g6(spacing[2])} "' in header
    assert f'<Piece Extent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}">' in header
    q = str(tmp_path / "o.pvti")
    O.vtk_write_pgrid(q, shape, GRID_NAMES, origin, spacing, "o_p0_0.vti")
    root = ET.parse(q).getroot()
    pim = root.find("PImageData")
    assert pim.get("WholeExtent") == f"0 {nx - 1} 0 {ny - 1}  0 {nz - 1}"
    assert [e.get("Name") for e in pim.find("PPointData")] == GRID_NAMES
    assert pim.find("Piece").get("Extent") == f"0 {nx} 0 {ny} 0 {nz - 1}"
    assert pim.find("Piece").get("Source") == "o_p0_0.vti"


def _py_fmt(precision, fmt):
    if precision > 0 and fmt == 0:
        return lambda v: f"{v:.{precision}f}"
    if precision > 0 and fmt == 1:
        return lambda v: f"{v:.{precision}e}"
    if fmt == 1:
        return lambda v: f"{v:e}"
    return lambda v: f"{v:.6g}"


@pytest.mark.parametrize("precision,fmt", [(-1, 0), (-1, 1), (3, 0), (12, 0), (4, 1), (17, 1)])
def test_oracle_atom_file_reads_back(tmp_path, precision, fmt):
    a = _atoms(40, 2)
    names = ["x", "ux", "ix", "F", "ID", "molID", "type"]
    p = str(tmp_path / "o.vtp")
    O.vtk_write_atoms(p, [a[w] for w in names], names, precision, fmt)
    root = ET.parse(p).getroot()
    piece = root.find("PolyData/Piece")
    assert piece.get("NumberOfPoints") == "40"
    arrays = [piece.find("Points/DataArray")] + list(piece.find("PointData"))
    assert [e.get("Name") for e in arrays] == names
    f = _py_fmt(precision, fmt)
    for e in arrays:
        w = e.get("Name")
        toks = e.text.split()
        if w in P.ATOM_VEC3:
            assert e.get("type") == "Float64" and e.get("NumberOfComponents") == "3"
            assert toks == [f(v) for v in a[w].ravel()]
        else:
            assert e.get("type") == "Int64" and e.get("NumberOfComponents") == "1"
            assert toks == [str(v) for v in a[w]]
    q = str(tmp_path / "o.pvtp")
    O.vtk_write_patoms(q, names, "o_p0_0.vtp")
    r = ET.parse(q).getroot().find("PPolyData")
    assert r.find("PPoints/PDataArray").get("Name") == "x"
    assert [e.get("Name") for e in r.find("PPointData")] == names[1:]
    assert r.find("Piece").get("Source") == "o_p0_0.vtp"


def test_oracle_collection_file(tmp_path):
    p = str(tmp_path / "c.pvd")
    O.vtk_collection(p, 0)
    for t in (0, 100, 200):
        O.vtk_collection(p, 1, t * 1e-3, f"run_{t}.pvti")
    O.vtk_collection(p, 2)
    root = ET.parse(p).getroot()
    ds = root.find("Collection").findall("DataSet")
    assert [d.get("timestep") for d in ds] == ["0", "0.1", "0.2"]
    assert [d.get("file") for d in ds] == ["run_0.pvti", "run_100.pvti", "run_200.pvti"]
    assert _read(p).endswith(b"</Collection>\n</VTKFile>")


@pytest.mark.parametrize("shape,origin,spacing,nattr", [((8, 6, 4), [-128.0, -128.0, -128.0], [4.0, 4.0, 4.0], 5),
                                                        ((3, 5, 7), [0.0, 0.5, -1.25], [10 / 3, 0.1, 1e-9], 1),
                                                        ((16, 16, 16), [1e6, -3.5e-5, 0.0], [0.25, 0.5, 7.0], 2)])
def test_product_grid_writer_matches_oracle_bytes(tmp_path, shape, origin, spacing, nattr):
    arrs = _grid_arrays(shape, 3, nattr)
    names = GRID_NAMES[:nattr]
    o, g = str(tmp_path / "o.vti"), str(tmp_path / "g.vti")
    O.vtk_write_grid(o, arrs, names, origin, spacing)
    P.vtk_write_grid(g, arrs, names, origin, spacing)
    assert _read(o) == _read(g)
    O.vtk_write_pgrid(o + ".p", shape, names, origin, spacing, "src.vti")
    P.vtk_write_pgrid(g + ".p", shape, names, origin, spacing, "src.vti")
    assert _read(o + ".p") == _read(g + ".p")


@pytest.mark.parametrize("precision,fmt", [(-1, 0), (-1, 1), (1, 0), (3, 0), (12, 0), (4, 1), (17, 1), (25, 0)])
def test_product_atom_writer_matches_oracle_bytes(tmp_path, precision, fmt):
    a = _atoms(300, 4)
    a["c_droppolyforces"] = a["F"][::-1] * 0.37
    for names in (["x", "ux", "ix", "F", "c_droppolyforces", "ID", "molID", "type"], ["ID", "x", "label"]):
        o, g = str(tmp_path / "o.vtp"), str(tmp_path / "g.vtp")
        O.vtk_write_atoms(o, [a[w] for w in names], names, precision, fmt)
        P.vtk_write_atoms(g, [a[w] for w in names], names, precision, "e" if fmt else "lf")
        assert _read(o) == _read(g)
        O.vtk_write_patoms(o + ".p", names, "s.vtp")
        P.vtk_write_patoms(g + ".p", names, "s.vtp")
        assert _read(o + ".p") == _read(g + ".p")


def test_product_atom_writer_in_chunks_matches_oracle_bytes(tmp_path):
    # large enough that the product formats the arrays in chunks on threads
    a = _atoms(50000, 8)
    names = ["x", "ux", "ix", "F", "ID", "molID", "type"]
    for precision, fmt in ((-1, 0), (17, 1)):
        o, g = str(tmp_path / "o.vtp"), str(tmp_path / "g.vtp")
        O.vtk_write_atoms(o, [a[w] for w in names], names, precision, fmt)
        P.vtk_write_atoms(g, [a[w] for w in names], names, precision, "e" if fmt else "lf")
        assert _read(o) == _read(g)


def test_product_collection_matches_oracle_bytes(tmp_path):
    o, g = str(tmp_path / "o.pvd"), str(tmp_path / "g.pvd")
    O.vtk_collection(o, 0)
    P.vtk_collection(g, P.PVD_HEADER)
    for t, dt in ((0, 1e-3), (100, 1e-3), (12345, 2.5e-4), (7, 1 / 3)):
        O.vtk_collection(o, 1, t * dt, f"a/b_{t}.pvti")
        P.vtk_collection(g, P.PVD_DATASET, t * dt, f"a/b_{t}.pvti")
    O.vtk_collection(o, 2)
    P.vtk_collection(g, P.PVD_FOOTER)
    assert _read(o) == _read(g)


def test_writer_errors():
    with pytest.raises(P.PhafdError, match="positions"):
        P.vtk_write_atoms("/tmp/none.vtp", [np.zeros(2, np.int32)], ["ID"])
    with pytest.raises(P.PhafdError, match="Invalid per atom quantity"):
        P.vtk_write_atoms("/tmp/none.vtp", [np.zeros((2, 3)), np.zeros(2, np.int32)], ["x", "velocity"])
    with pytest.raises(P.PhafdError, match="not dumped here"):
        P.vtk_write_atoms("/tmp/none.vtp", [np.zeros((2, 3)), np.zeros(2, np.int32)], ["x", "f_ave"])
    with pytest.raises(P.PhafdError, match="Cannot open file"):
        P.vtk_write_grid("/nonexistent_dir/x.vti", [np.zeros((2, 2, 2))], ["phi"], [0] * 3, [1] * 3)


def test_reader_takes_requested_attributes_from_oracle_file(tmp_path):
    shape = (6, 5, 4)
    arrs = _grid_arrays(shape, 5)
    p = str(tmp_path / "o.vti")
    O.vtk_write_grid(p, arrs, GRID_NAMES, [0.0] * 3, [1.0] * 3)
    got = P.vtk_read_grid(p, shape, ["gradphi_y", "phi"])   # other arrays skipped
    assert np.array_equal(got[0], arrs[3]) and np.array_equal(got[1], arrs[0])
    got = P.vtk_read_grid(p, shape, GRID_NAMES, no_padding=True)
    assert all(np.array_equal(g, a) for g, a in zip(got, arrs))


def test_reader_errors(tmp_path):
    shape = (6, 5, 4)
    arrs = _grid_arrays(shape, 6, 2)
    p = str(tmp_path / "o.vti")
    O.vtk_write_grid(p, arrs, ["phi", "chempot"], [0.0] * 3, [1.0] * 3)
    with pytest.raises(P.PhafdError, match="needs attribute phi"):
        P.vtk_read_grid(p, shape, ["chempot"])
    with pytest.raises(P.PhafdError, match="Not all attributes requested"):
        P.vtk_read_grid(p, shape, ["phi", "gradphi_x"])
    with pytest.raises(P.PhafdError, match="incorrect size phi"):
        P.vtk_read_grid(p, (6, 5, 5), ["phi"])
    with pytest.raises(P.PhafdError, match="file not found"):
        P.vtk_read_grid(str(tmp_path / "missing.vti"), shape, ["phi"])


def test_atom_reader_round_trip_and_required_attributes(tmp_path):
    a = _atoms(50, 7)
    names = ["x", "ux", "ix", "F", "ID", "molID", "type"]
    p = str(tmp_path / "o.vtp")
    O.vtk_write_atoms(p, [a[w] for w in names], names, 17, 1)   # %.17e: every double exactly
    got = P.read_dump_atoms(p, 50, names)
    for w in names:
        assert np.array_equal(got[w], a[w]), w
    q = str(tmp_path / "d.vtp")
    O.vtk_write_atoms(q, [a[w] for w in names], names, -1, 0)   # the default: 6 significant digits
    got = P.read_dump_atoms(q, 50, ["x", "ID", "type", "molID", "ix"])
    assert np.array_equal(got["x"], np.array([float(f"{v:.6g}") for v in a["x"].ravel()]).reshape(-1, 3))
    with pytest.raises(P.PhafdError, match="needs attribute molID"):
        P.read_dump_atoms(p, 50, ["x", "ID", "type"])
    assert P.read_dump_atoms(p, 50, ["x", "ID", "type"], molecule=False)["ID"][-1] == 49
    with pytest.raises(P.PhafdError, match="Only found attributes"):
        P.read_dump_atoms(p, 50, ["x", "ID", "type", "label"], molecule=False)


@pytest.mark.parametrize("precision,fmt", [(-1, 0), (-1, 1), (4, 0), (17, 1)])
def test_lammps_dump_matches_oracle_and_reads_back(tmp_path, precision, fmt):
    a = _atoms(200, 9)
    a["c_pf"] = a["F"] * -0.5
    names = ["x", "F", "c_pf", "ID", "molID", "type", "ix"]
    lo, hi = [-128.0, -128.0, -128.0], [128.0, 128.0, 128.0]
    o, g = str(tmp_path / "o.lmmpstrj"), str(tmp_path / "g.lmmpstrj")
    for t in (0, 100):   # two steps appended to one file
        O.lammps_write_step(o, t, [a[w] for w in names], names, lo, hi, precision, fmt)
        P.lammps_write_step(g, t, [a[w] for w in names], names, lo, hi, precision, "e" if fmt else "lf")
    assert _read(o) == _read(g)
    lines = _read(o).decode().split("\n")
    f = _py_fmt(precision, fmt)
    assert lines[:4] == ["ITEM: TIMESTEP", "0", "ITEM: NUMBER OF ATOMS", "200"]
    assert lines[5] == f"{f(-128.0)} {f(128.0)}"
    assert lines[8] == "ITEM: ATOMS x y z Fx Fy Fz c_pf_0 c_pf_1 c_pf_2 ID molID type ix "
    row = lines[9].split()
    assert row[:3] == [f(v) for v in a["x"][0]] and row[6:9] == [f(v) for v in a["c_pf"][0]]
    assert row[9:] == [str(a[w][0]) for w in ("ID", "molID", "type", "ix")]
    assert lines[9 + 200] == "ITEM: TIMESTEP" and lines[10 + 200] == "100"


def test_grid_ave_dump_frequency_must_match():
    # FixGridAve::setup refuses a dump of the fix whose every differs from
    # the fix's freq (src/fixgrid_ave.cpp:101-107); no GPU work happens first
    from phafd_amd.computes import ComputeComplex, FixGridAve
    from phafd_amd.dump import Dump, run_with_dumps
    sq = ComputeComplex(None, "sq", "ft_phi", "norm")
    ave = FixGridAve(None, "sqave", 2, 5, 10, sq)
    bad = Dump(None, "ft", "ftgrid", "/nonexistent/sq", 20, ["c_sq", "f_sqave"], computes={"sq": sq, "sqave": ave})
    with pytest.raises(ValueError, match="doesn't match dump frequency"):
        run_with_dumps(None, 100, [bad], 1e-3, aves=[ave])
