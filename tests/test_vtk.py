"""VTK files of the export pipeline: the layout the reference's reader parses
(src/VtkReader/VtkReader.jl) and the names its writer gives (src/auxiliary/io.jl:143-179).

The checks parse the files independently of peridynamics.vtk.read_vtk, following the
reader's own steps: the raw appended block between `<AppendedData encoding="raw">` + `_`
and `\\n  </AppendedData>`; per DataArray a header of four UInt64 whose fourth is the
compressed size (get_decompressed_data, :241-254), then zlib data; the attributes
type/Name/NumberOfComponents/offset/format="appended" it requires (DataArray, :17-40);
Points, PointData under Piece and FieldData under UnstructuredGrid (find_data_xml,
:99-111); Piece Source paths relative to the .pvtu (get_vtu_files, :266-278)."""
import os
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import pytest

import peridynamics as pd
from peridynamics import vtk

DT = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}


def parse_like_reference(fn):
    raw = open(fn, "rb").read()
    m = raw.index(b'<AppendedData encoding="raw">')
    begin = raw.index(b"_", m + len(b'<AppendedData encoding="raw">')) + 1
    end = raw.index(b"\n  </AppendedData>", begin)
    data = raw[begin:end]
    root = ET.fromstring(raw[:begin] + raw[end:])
    assert root.tag == "VTKFile" and root.get("header_type") == "UInt64"
    grid = root.find("UnstructuredGrid")
    piece = grid.find("Piece")

    def get(da):
        for k in ("type", "Name", "NumberOfComponents", "offset", "format"):
            assert da.get(k) is not None, (k, da.attrib)
        assert da.get("format") == "appended"
        off = int(da.get("offset"))
        hdr = np.frombuffer(data[off:off + 32], np.uint64)
        assert hdr[0] == 1
        arr = np.frombuffer(zlib.decompress(data[off + 32:off + 32 + int(hdr[3])]), DT[da.get("type")])
        assert arr.nbytes == hdr[1]
        nc = int(da.get("NumberOfComponents"))
        return arr.reshape(-1, nc) if nc > 1 else arr

    out = {"position": get(piece.find("Points").find("DataArray"))}
    for da in piece.find("PointData"):
        out[da.get("Name")] = get(da)
    for da in grid.find("FieldData"):
        out[da.get("Name")] = get(da)
    cells = {da.get("Name"): da for da in piece.find("Cells")}
    assert set(cells) == {"connectivity", "offsets", "types"}
    return out


def sample(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 3)), {"displacement": rng.standard_normal((n, 3)),
                                         "damage": rng.uniform(0, 1, n)}


def test_vtu_layout_and_round_trip(tmp_path):
    pos, data = sample(1000)
    fn = str(tmp_path / "a.vtu")
    vtk.write_vtu(fn, pos, data, 1.25e-5)
    r = parse_like_reference(fn)
    assert np.array_equal(r["position"], pos)
    assert np.array_equal(r["displacement"], data["displacement"])
    assert np.array_equal(r["damage"], data["damage"])
    assert np.array_equal(r["time"], [1.25e-5])
    r2 = pd.read_vtk(fn)
    for k in r:
        assert np.array_equal(r2[k], r[k])


@pytest.mark.parametrize("parts", [1, 3, 16])
def test_pvtu_parts_and_names(tmp_path, parts):
    pos, data = sample(5003, seed=parts)
    base = vtk.vtk_filebase(str(tmp_path))
    assert base == os.path.join(str(tmp_path), "vtk", "timestep")
    assert vtk.vtk_filebase("r", "my body") == os.path.join("r", "vtk", "my_body_timestep")
    os.makedirs(os.path.dirname(base))
    vtk.write_pvtu(base, 40, pos, data, 0.5, parts=parts)
    fn = base + "_40.pvtu"
    root = ET.parse(fn).getroot()
    srcs = [p.get("Source") for p in root.find("PUnstructuredGrid").findall("Piece")]
    assert srcs == [f"timestep_40/timestep_40_{k + 1}.vtu" for k in range(parts)]
    pieces = [parse_like_reference(os.path.join(os.path.dirname(fn), s)) for s in srcs]
    assert np.array_equal(np.concatenate([p["position"] for p in pieces]), pos)
    assert np.array_equal(np.concatenate([p["damage"] for p in pieces]), data["damage"])
    r = pd.read_vtk(fn)
    assert np.array_equal(r["position"], pos)
    assert np.array_equal(r["displacement"], data["displacement"])
    assert np.array_equal(r["time"], [0.5])


def test_process_each_export_order(tmp_path):
    base = vtk.vtk_filebase(str(tmp_path))
    os.makedirs(os.path.dirname(base))
    for n in (0, 10, 100, 20):
        pos, data = sample(50, seed=n)
        vtk.write_pvtu(base, n, pos, data, n * 1e-3)
    seen = pd.process_each_export(lambda r0, r, i: (i, float(r["time"][0])),
                                  os.path.join(str(tmp_path), "vtk"), default_value=(0, 0.0))
    assert seen == [(1, 0.0), (2, 0.01), (3, 0.02), (4, 0.1)]


def test_reader_errors(tmp_path):
    with pytest.raises(ValueError):
        pd.read_vtk(str(tmp_path / "missing.pvtu"))
    p = tmp_path / "x.txt"
    p.write_text("x")
    with pytest.raises(ValueError):
        pd.read_vtk(str(p))


def test_compression_choice(tmp_path):
    """Random FP64 data is stored (zlib level 0), a mostly-zero field is deflated; both are
    zlib streams the reader decodes."""
    n = 100_000
    rng = np.random.default_rng(1)
    dmg = np.where(rng.uniform(size=n) < 0.95, 0.0, rng.uniform(size=n))
    fn = str(tmp_path / "c.vtu")
    vtk.write_vtu(fn, rng.standard_normal((n, 3)), {"damage": dmg}, 0.0)
    raw = open(fn, "rb").read()
    begin = raw.index(b"_", raw.index(b'<AppendedData encoding="raw">') + 10) + 1
    sizes = []
    off = 0
    for _ in range(5):  # Points, connectivity, offsets, types, damage
        hdr = np.frombuffer(raw[begin + off:begin + off + 32], np.uint64)
        sizes.append((int(hdr[1]), int(hdr[3])))
        off += 32 + int(hdr[3])
    assert sizes[0][1] >= sizes[0][0]          # positions: stored
    assert sizes[4][1] < 0.3 * sizes[4][0]     # damage: deflated
    r = parse_like_reference(fn)
    assert np.array_equal(r["damage"], dmg)
