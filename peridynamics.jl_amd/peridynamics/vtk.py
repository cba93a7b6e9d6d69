"""VTK export of simulation results, and reading them back.

The files follow what the reference writes with WriteVTK (_export_results,
src/auxiliary/io.jl:157-167; get_vtk_filebase, io.jl:143-148; get_filename, io.jl:169-179) and
what its reader accepts (src/VtkReader/VtkReader.jl): per export a parallel unstructured
grid `<root>/vtk/timestep_<n>.pvtu` whose pieces `timestep_<n>/timestep_<n>_<part>.vtu`
hold the current positions as points, one VTK_VERTEX cell per point, the exported fields as
point data and the time as field data, every array zlib-compressed in raw appended form
with UInt64 block headers (the deflate level is chosen per array, see _block: the writing
has to keep up with the device, and FP64 field data hardly compresses).  A body is split into
`parts` contiguous pieces written in parallel, as the reference's threads backend writes
one piece per chunk.

VTKExporter is the host half of the export pipeline: at an export step it asks libpdx for
an asynchronous snapshot (pdx_export_begin) and returns; the previous snapshot, by then in
pinned host memory, is compressed and written by worker threads while the device runs the
next steps."""
import functools
import os
import re
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

VTK_VERTEX = 1
_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64", np.dtype(np.uint8): "UInt8"}
_DTYPES = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}


def vtk_filebase(root, body_name=""):
    """get_vtk_filebase (io.jl:143-148)."""
    name = str(body_name).replace(" ", "_")
    return os.path.join(root, "vtk", f"{name}_timestep" if name else "timestep")


def _block(a, level):
    """One appended DataArray: UInt64 header [blocks = 1, size, size, compressed size] and
    the zlib stream.  level None picks per array: deflate level 1 where a 64 KiB probe
    shrinks below 90 % (damage, counts), stored blocks (level 0, still a zlib stream) where
    it does not (positions, displacements: FP64 mantissas do not compress)."""
    raw = memoryview(np.ascontiguousarray(a)).cast("B")
    if level is None:
        probe = raw[:65536]
        level = 1 if len(zlib.compress(probe, 1)) < 0.9 * max(probe.nbytes, 1) else 0
    comp = zlib.compress(raw, level)  # releases the GIL: pieces compress in parallel
    return np.array([1, raw.nbytes, raw.nbytes, len(comp)], np.uint64).tobytes() + comp


@functools.lru_cache(maxsize=64)
def _cell_blocks(n, level):
    """connectivity, offsets and types of n VTK_VERTEX cells: the same for every export of
    a piece, so compressed once."""
    conn = np.arange(n, dtype=np.int64)
    return (_block(conn, level), _block(conn + 1, level),
            _block(np.full(n, VTK_VERTEX, np.uint8), level))


def write_vtu(filename, position, point_data, time, level=None):
    """One piece: position (n, 3), point_data {name: (n,) or (n, c)}, time (field data)."""
    position = np.asarray(position, np.float64)
    n = position.shape[0]
    arrays = []  # (xml section, attributes, data or a compressed block)
    arrays.append(("Points", dict(type="Float64", Name="Points", NumberOfComponents="3"), position))
    cb = _cell_blocks(n, level)
    arrays.append(("Cells", dict(type="Int64", Name="connectivity"), cb[0]))
    arrays.append(("Cells", dict(type="Int64", Name="offsets"), cb[1]))
    arrays.append(("Cells", dict(type="UInt8", Name="types"), cb[2]))
    for name, a in point_data.items():
        a = np.asarray(a)
        nc = 1 if a.ndim == 1 else a.shape[1]
        arrays.append(("PointData", dict(type=_TYPES[a.dtype], Name=name, NumberOfComponents=str(nc)), a))
    tarr = np.array([float(time)], np.float64)
    arrays.append(("FieldData", dict(type="Float64", Name="time", NumberOfTuples="1",
                                     NumberOfComponents="1"), tarr))
    blobs, offset, xml = [], 0, {}
    for sec, attrs, a in arrays:
        b = a if isinstance(a, bytes) else _block(a, level)
        attrs = dict(attrs, format="appended", offset=str(offset))
        xml.setdefault(sec, []).append(
            "<DataArray " + " ".join(f'{k}="{v}"' for k, v in attrs.items()) + "/>")
        blobs.append(b)
        offset += len(b)

    def section(sec, indent):
        pad = " " * indent
        return (f"{pad}<{sec}>\n" + "".join(f"{pad}  {e}\n" for e in xml.get(sec, [])) + f"{pad}</{sec}>\n")

    head = ('<?xml version="1.0" encoding="utf-8"?>\n'
            '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64" compressor="vtkZLibDataCompressor">\n'
            "  <UnstructuredGrid>\n"
            f'    <Piece NumberOfPoints="{n}" NumberOfCells="{n}">\n'
            + section("Points", 6) + section("Cells", 6) + section("PointData", 6)
            + "    </Piece>\n" + section("FieldData", 4)
            + "  </UnstructuredGrid>\n"
            '  <AppendedData encoding="raw">\n   _')
    with open(filename, "wb") as f:
        f.write(head.encode())
        for b in blobs:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def write_pvtu(filebase, n_step, position, point_data, time, parts=1, level=None, pool=None):
    """The files of one export (_export_results): `<filebase>_<n>.pvtu` and its `parts`
    pieces.  With a thread pool the pieces are written by its workers; returns their
    futures (or None when written here)."""
    name = f"{filebase}_{n_step}"
    base = os.path.basename(name)
    os.makedirs(name, exist_ok=True)
    n = np.asarray(position).shape[0]
    parts = max(1, min(int(parts), max(n, 1)))
    bounds = np.linspace(0, n, parts + 1).astype(np.int64)
    srcs = []
    jobs = []
    for k in range(parts):
        lo, hi = bounds[k], bounds[k + 1]
        fn = os.path.join(name, f"{base}_{k + 1}.vtu")
        srcs.append(f"{base}/{base}_{k + 1}.vtu")
        pd_k = {f: a[lo:hi] for f, a in point_data.items()}
        jobs.append((fn, position[lo:hi], pd_k))
    pdas = "".join(
        f'      <PDataArray type="{_TYPES[np.asarray(a).dtype]}" Name="{f}" '
        f'NumberOfComponents="{1 if np.asarray(a).ndim == 1 else np.asarray(a).shape[1]}"/>\n'
        for f, a in point_data.items())
    pieces = "".join(f'    <Piece Source="{s}"/>\n' for s in srcs)
    with open(name + ".pvtu", "w") as f:
        f.write('<?xml version="1.0" encoding="utf-8"?>\n'
                '<VTKFile type="PUnstructuredGrid" version="1.0" byte_order="LittleEndian" '
                'header_type="UInt64">\n'
                '  <PUnstructuredGrid GhostLevel="0">\n'
                "    <PPoints>\n"
                '      <PDataArray type="Float64" Name="Points" NumberOfComponents="3"/>\n'
                "    </PPoints>\n"
                "    <PPointData>\n" + pdas + "    </PPointData>\n"
                + pieces + "  </PUnstructuredGrid>\n</VTKFile>\n")
    if pool is None:
        for fn, pos, pdk in jobs:
            write_vtu(fn, pos, pdk, time, level)
        return None
    return [pool.submit(write_vtu, fn, pos, pdk, time, level) for fn, pos, pdk in jobs]


# ---------------------------------------------------------------------------------- reading
_DA = re.compile(r"<DataArray\s+([^>]*?)/?>")
_ATTR = re.compile(r'(\w+)="([^"]*)"')


def _read_vtu(filename):
    raw = open(filename, "rb").read()
    marker = raw.find(b'<AppendedData encoding="raw">')
    if marker < 0:
        raise ValueError(f"{filename}: no raw appended data")
    begin = raw.find(b"_", marker) + 1
    end = raw.find(b"\n  </AppendedData>", begin)
    data = raw[begin:end]
    xml = (raw[:begin] + raw[end:]).decode()
    out = {}

    def arr(attrs):
        off = int(attrs["offset"])
        hdr = np.frombuffer(data[off:off + 32], np.uint64)
        comp = data[off + 32:off + 32 + int(hdr[3])]
        a = np.frombuffer(zlib.decompress(comp), _DTYPES[attrs["type"]])
        nc = int(attrs.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    def section(tag):
        i = xml.find(f"<{tag}>")
        if i < 0:
            return []
        j = xml.find(f"</{tag}>", i)
        return [dict(_ATTR.findall(m)) for m in _DA.findall(xml[i:j])]

    out["position"] = arr(section("Points")[0])
    for a in section("PointData"):
        out[a["Name"]] = arr(a)
    fields = {a["Name"]: arr(a) for a in section("FieldData")}
    return out, fields


def read_vtk(filename):
    """read_vtk (src/VtkReader/VtkReader.jl:296-323): position and point data of a .vtu or
    .pvtu (pieces concatenated in order) as (n,) or (n, c) arrays, field data (time) with
    duplicates across pieces removed."""
    if not os.path.isfile(filename):
        raise ValueError(f"cannot find file {filename}!")
    if filename.endswith(".vtu"):
        points, fields = _read_vtu(filename)
        return {**points, **fields}
    if not filename.endswith(".pvtu"):
        raise ValueError(f"cannot read file with extension {os.path.splitext(filename)[1]}")
    xml = open(filename).read()
    srcs = re.findall(r'<Piece\s+Source="([^"]+)"', xml)
    parts = [_read_vtu(os.path.join(os.path.dirname(filename), s)) for s in srcs]
    out = {k: np.concatenate([p[0][k] for p in parts], axis=0) for k in parts[0][0]}
    for k in parts[0][1]:
        out[k] = np.unique(np.concatenate([p[1][k] for p in parts]))
    return out


def export_files(vtk_path):
    """The .pvtu files of a results folder in time-step order (the reference file first)."""
    fs = [f for f in os.listdir(vtk_path) if f.endswith(".pvtu")]

    def step(f):
        m = re.search(r"_(\d+)\.pvtu$", f)
        return int(m.group(1)) if m else -1

    return [os.path.join(vtk_path, f) for f in sorted(fs, key=step)]


def process_each_export(f, vtk_path, default_value=None, serial=True):
    """process_each_export(f, vtk_path, default_value) (src/auxiliary/process_each_export.jl):
    f(r0, r, id) for every export file, r0 the reference results, id counted from 1.
    Returns the list of f's results when default_value is given, else None."""
    files = export_files(vtk_path)
    if not files:
        raise ValueError(f"no export files in {vtk_path}")
    r0 = read_vtk(files[0])
    results = [] if default_value is not None else None
    for i, fn in enumerate(files):
        res = f(r0, read_vtk(fn), i + 1)
        if results is not None:
            results.append(res if res is not None else default_value)
    return results


# ---------------------------------------------------------------------------------- pipeline
class VTKExporter:
    """Host half of the export pipeline for one data handler.  capture(dh, n, t) is an
    on_export callback: it starts the device snapshot of step n and hands the previous
    snapshot to the writer threads; close() drains and waits for the last files."""

    def __init__(self, dh, root, fields, body_name="", parts=None, level=None, workers=None):
        self.dh = dh
        self.filebase = vtk_filebase(root, body_name)
        os.makedirs(os.path.dirname(self.filebase), exist_ok=True)
        self.fields = [f for f in fields if f != "position"]
        cpus = os.cpu_count() or 1
        self.parts = int(parts) if parts else max(1, min(16, cpus, dh.n // 100_000 + 1))
        self.level = level
        self.pool = ThreadPoolExecutor(max_workers=workers or min(16, cpus))
        self.times = {}
        self.slot_jobs = {0: [], 1: []}
        self.begun = 0
        self.pending = 0
        self.files = []

    def _drain_one(self):
        tag, slot, arrays = self.dh.ctx.export_wait()
        self.pending -= 1
        pos = arrays.pop("position")
        futs = write_pvtu(self.filebase, tag, pos, arrays, self.times.pop(tag), self.parts,
                          self.level, self.pool)
        self.slot_jobs[slot] = futs
        self.files.append(f"{self.filebase}_{tag}.pvtu")

    def capture(self, dh, n, t):
        slot = self.begun % 2
        for fut in self.slot_jobs[slot]:  # the copy will overwrite this slot's host memory
            fut.result()
        self.slot_jobs[slot] = []
        self.times[n] = t
        dh.ctx.export_begin(["position"] + self.fields, n)
        self.begun += 1
        self.pending += 1
        while self.pending > 1:  # the previous snapshot is on the host by now
            self._drain_one()

    def close(self):
        while self.pending:
            self._drain_one()
        for lst in self.slot_jobs.values():
            for fut in lst:
                fut.result()
        self.pool.shutdown()
        return self.files
