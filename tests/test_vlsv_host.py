"""The VLSV format pinned against the reference's own reader without a GPU.

tests/vlsv_host_writer.cpp writes a file with the library's host-side writer
(rhybrid_amd/csrc/rh_vlsv.hpp: the grid part every rh_write_vlsv file starts
with, and cell variables).  tests/golden/vlsv_reference_reader.* hold what
the reference's VlsvReader (tools/rhybridpy.py) read from that file when the
fixture was made.  Here the writer is rebuilt under AddressSanitizer and
UndefinedBehaviorSanitizer, must write the same bytes, and the tests' reader
must read the same mesh, parameters and values from them."""
import hashlib
import json
import os
import subprocess

import numpy as np

from vlsv_read import read_vlsv

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def write_file(tmp_path):
    exe = str(tmp_path / "vlsv_host_writer")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-o", exe, os.path.join(HERE, "vlsv_host_writer.cpp")])
    path = str(tmp_path / "host.vlsv")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    subprocess.check_call([exe, path], env=env)
    return path


def test_host_writer_matches_the_reference_reader_fixture(tmp_path):
    fx = json.load(open(os.path.join(GOLDEN, "vlsv_reference_reader.json")))
    arr = np.load(os.path.join(GOLDEN, "vlsv_reference_reader.npz"))
    path = write_file(tmp_path)
    raw = open(path, "rb").read()
    assert hashlib.sha256(raw).hexdigest() == fx["sha256"]
    f = read_vlsv(path)
    # mesh: MESH_BBOX gives the cells per axis and the block, the node coordinates the extent
    bbox = f[("MESH_BBOX", "")][0]
    assert bbox[:3].tolist() == fx["mesh_size"] and bbox[3:].tolist() == fx["block_size"]
    x, y, z = (f[(t, "")][0] for t in ("MESH_NODE_CRDS_X", "MESH_NODE_CRDS_Y", "MESH_NODE_CRDS_Z"))
    assert [x[0], y[0], z[0], x[-1], y[-1], z[-1]] == fx["extent"]
    for k, v in fx["parameters"].items():
        assert float(f[("PARAMETER", k)][0][0]) == v, k
    assert [k[1] for k in f if k[0] == "VARIABLE"] == fx["variables"]
    assert int(f[("VARIABLE", "cellB")][1]["vectorsize"]) == fx["vectorsize_cellB"]
    for name in ("cellRhoQi", "cellB", "cellNPles", "CellID"):
        assert np.array_equal(f[("VARIABLE", name)][0], arr[name]), name
    # selections by CellID (the cell index) and the magnitude operator
    ids = f[("VARIABLE", "CellID")][0]
    pos = {int(c): q for q, c in enumerate(ids)}
    B = f[("VARIABLE", "cellB")][0]
    assert np.array_equal(B[[pos[5], pos[17], pos[0]]], arr["cellB_at_5_17_0"])
    assert np.allclose(np.linalg.norm(B[[pos[3]]], axis=1), arr["cellB_mag_at_3"], rtol=0, atol=0)
    # the reader's geometric CellIDs are one-based: CellID c is cell c - 1
    nx, ny, nz = fx["mesh_size"]
    ext = np.array(fx["extent"])
    d = (ext[3:] - ext[:3]) / np.array(fx["mesh_size"])
    for c, want in zip((1, 2, 120), arr["coords_of_1_2_120"]):
        n = c - 1
        ijk = np.array([n % nx, (n // nx) % ny, n // (nx * ny)])
        assert np.allclose(ext[:3] + (ijk + 0.5) * d, want, rtol=1e-15)
    # the one-byte flag array, which the reference's reader cannot read
    assert fx["flag_read_error"] == "TypeError"
    flag = f[("VARIABLE", "innerFlagField")][0]
    assert flag.dtype == np.uint8 and np.array_equal(flag, (np.arange(ids.size) % 3 == 0).astype(np.uint8))
