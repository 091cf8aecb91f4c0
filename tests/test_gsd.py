"""GSD files of the C++ facade (fly::io::BinaryFile, gsd.hpp:118-327, gsd.cpp:63-253; host code only).

tests/cpp/gsd_check.cpp runs the reference's own round trip (tests/IO/gsd.cpp:32-111) and the cases it leaves
open (header offsets, index and name-list growth, reopen and append, clear, per-species values, a triclinic
box, error paths).  Here a second reader, written with `struct` from the published GSD 2.0 layout
(256-byte header, 32-byte index entries ordered by frame then name id, NUL-separated name list), reads the
file that program leaves and checks every chunk's bytes, and a second writer built the same way makes a
file the facade must read, so a layout mistake made the same way in the writer and the reader of the
facade still fails.  Against the gsd library itself (absent here, an empty
submodule in the reference) the format is **parity unpinned**."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SIZES = {1: 1, 2: 2, 3: 4, 4: 8, 5: 1, 6: 2, 7: 4, 8: 8, 9: 4, 10: 8}
DTYPES = {1: "<u1", 2: "<u2", 3: "<u4", 4: "<u8", 5: "<i1", 6: "<i2", 7: "<i4", 8: "<i8", 9: "<f4", 10: "<f8"}


def read_gsd(path):
    """Frames of a GSD 2.0 file as a list of {name: array}."""
    b = open(path, "rb").read()
    (magic, idx_loc, idx_n, nl_loc, nl_bytes, schema_v, gsd_v, app, schema, _) = struct.unpack_from(
        "<QQQQQII64s64s80s", b, 0)
    assert magic == 0x65DF65DF65DF65DF
    assert gsd_v >> 16 == 2
    names = b[nl_loc:nl_loc + nl_bytes].split(b"\0")
    names = [n.decode() for n in names[:names.index(b"")]]
    frames = []
    for k in range(idx_n):
        frame, N, loc, M, nid, typ, flags = struct.unpack_from("<QQqIHBB", b, idx_loc + 32 * k)
        if loc == 0:
            break
        while len(frames) <= frame:
            frames.append({})
        a = np.frombuffer(b, DTYPES[typ], N * M, loc).reshape(N, M)
        frames[frame][names[nid]] = a
    return dict(app=app.rstrip(b"\0").decode(), schema=schema.rstrip(b"\0").decode(), schema_version=schema_v,
                names=names, frames=frames)


def write_gsd(path, frames, index_entries=8):
    """A GSD 2.0 file written with `struct` from the published layout: frames is a list of
    {name: (type code, 2-D array)}; inside each frame the index entries go in descending name id (the
    lookup must not depend on their order); the index holds `index_entries` slots."""
    names = []
    for fr in frames:
        for k in fr:
            if k not in names:
                names.append(k)
    nl = b"".join(n.encode() + b"\0" for n in names)
    nl_bytes = max(64, len(nl) + 1)
    idx_loc = 256
    nl_loc = idx_loc + 32 * index_entries
    data_loc = nl_loc + nl_bytes
    blobs, entries, at = [], [], data_loc
    for f, fr in enumerate(frames):
        for name in sorted(fr, key=names.index, reverse=True):
            typ, a = fr[name]
            a = np.ascontiguousarray(a, dtype=DTYPES[typ])
            entries.append(struct.pack("<QQqIHBB", f, a.shape[0], at, a.shape[1], names.index(name), typ, 0))
            blobs.append(a.tobytes())
            at += len(blobs[-1])
    assert len(entries) <= index_entries
    head = struct.pack("<QQQQQII64s64s80s", 0x65DF65DF65DF65DF, idx_loc, index_entries, nl_loc, nl_bytes,
                       (1 << 16) | 4, 2 << 16, b"pytest", b"hoomd", b"")
    idx = b"".join(entries).ljust(32 * index_entries, b"\0")
    with open(path, "wb") as fh:
        fh.write(head + idx + nl.ljust(nl_bytes, b"\0") + b"".join(blobs))


def build_checker(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    exe = os.path.join(tmp_path, "gsd_check")
    src = os.path.join(ROOT, "tests", "cpp", "gsd_check.cpp")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", exe, src], check=True)
    return exe


def test_reads_a_file_written_from_the_published_layout(tmp_path):
    # The converse check: a file written here with struct (small index, entries of a frame in descending
    # name id, a frame without positions) read through io::BinaryFile.
    exe = build_checker(tmp_path)
    pos = np.array([[0.5, -1.25, 3.0], [1e-3, 2.0, -7.5], [4.25, 0.0, 1.0]])
    path = os.path.join(tmp_path, "foreign.gsd")
    write_gsd(path, [{"particles/N": (3, [[3]]), "particles/position": (10, pos), "log/value": (7, [[7]])},
                     {"particles/N": (3, [[3]]), "log/value": (7, [[9]])}])
    r = subprocess.run([exe, "--read", path], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "frames 2"
    for f, value in ((0, 7), (1, 9)):
        head = lines[1 + f].split(" pos ")
        assert head[0] == f"frame {f} N 3 value {value}"
        assert np.array_equal(np.array([float(x) for x in head[1].split()]).reshape(3, 3), pos)
    # A file that is not GSD, and one whose index points past its end, are refused.
    bad = os.path.join(tmp_path, "bad.gsd")
    with open(bad, "wb") as fh:
        fh.write(b"\0" * 300)
    assert "Not a GSD file" in subprocess.run([exe, "--read", bad], capture_output=True, text=True).stdout
    with open(path, "r+b") as fh:
        fh.seek(8)
        fh.write(struct.pack("<Q", 1 << 40))
    assert "File corrupt" in subprocess.run([exe, "--read", path], capture_output=True, text=True).stdout


def test_gsd_program(tmp_path):
    exe = build_checker(tmp_path)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "gsd ok" in r.stdout

    g = read_gsd(os.path.join(tmp_path, "layout.gsd"))
    assert (g["app"], g["schema"], g["schema_version"]) == ("openFLY", "hoomd", (1 << 16) | 4)
    assert g["names"] == ["particles/N", "particles/position", "particles/typeid", "log/energy"]
    f0, f1 = g["frames"]
    assert f0["particles/N"].dtype == np.uint32 and f0["particles/N"].tolist() == [[3]]
    want = np.array([[0.25 * i, -1.5 * i, 3.0 + i] for i in range(3)])
    assert f0["particles/position"].dtype == np.float64 and np.array_equal(f0["particles/position"], want)
    assert f0["particles/typeid"].dtype == np.uint32 and f0["particles/typeid"].ravel().tolist() == [2, 1, 0]
    assert set(f1) == {"log/energy"} and f1["log/energy"].tolist() == [[-123.456]]

    # The reference round trip's file: box in HOOMD's six floats, names as 32-byte rows, fallback frame.
    g = read_gsd(os.path.join(tmp_path, "BinaryFile_test.gsd"))
    f0, f1 = g["frames"]
    assert f0["configuration/box"].dtype == np.float32 and f0["configuration/box"].ravel().tolist() == [1, 1, 1, 0, 0, 0]
    assert f0["log/periodicity"].ravel().tolist() == [1, 1, 1]
    assert f0["log/typemap/N"].dtype == np.int64 and f0["log/typemap/N"].tolist() == [[2]]
    t = f0["particles/types"]
    assert t.dtype == np.int8 and t.shape == (2, 32)
    assert [bytes(r.astype(np.uint8)).rstrip(b"\0").decode() for r in t] == ["Fe", "H"]
    assert f0["log/typemap/particles/index"].ravel().tolist() == [6, 1]
    assert np.array_equal(f1["particles/position"], f0["particles/position"] + 1)
    assert "particles/typeid" not in f1

    # 400 frames: index and name list were moved to the end of the file; entries stay readable in order.
    g = read_gsd(os.path.join(tmp_path, "growth.gsd"))
    assert len(g["frames"]) == 400
    for f, fr in enumerate(g["frames"]):
        assert fr["frame"].tolist() == [[f]]
        assert fr[f"log/value/a_fairly_long_chunk_name_{f % 60}"].tolist() == [[0.5 * f]]
