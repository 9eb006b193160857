"""Writes tests/golden/vlsv_reference_reader.{json,npz}: what the reference's
own VLSV reader (tools/rhybridpy.py VlsvReader, read in the build container
from the reference tree) reads from a file written by the library's host-side
writer (tests/vlsv_host_writer.cpp, rhybrid_amd/csrc/rh_vlsv.hpp).

The fixture records the file's SHA-256 and the reader's results: the mesh
(size, block size, extent), the parameters, the variable list, whole
variables, selections by CellID, the magnitude operator, the vector size,
cell-centre coordinates of given CellIDs (the reader's one-based geometric
convention), and the error the reader raises for the one-byte flag array
(its type table has no one-byte entry).  tests/test_vlsv_host.py rebuilds the
writer, checks that it writes the same bytes, and reads them back with the
tests' own reader against these values.

    python tests/golden/make_vlsv_fixture.py      (needs the reference tree)
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFERENCE_TOOLS = "/root/reference/tools"


def build_writer(out_dir: str, sanitize: bool = False) -> str:
    exe = os.path.join(out_dir, "vlsv_host_writer")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-o", exe, os.path.join(ROOT, "tests", "vlsv_host_writer.cpp")]
    if sanitize:
        cmd[3:3] = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    subprocess.check_call(cmd)
    return exe


def main():
    sys.path.insert(0, REFERENCE_TOOLS)
    import rhybridpy as rhb
    with tempfile.TemporaryDirectory() as d:
        exe = build_writer(d)
        path = os.path.join(d, "host.vlsv")
        subprocess.check_call([exe, path])
        raw = open(path, "rb").read()
        r = rhb.VlsvReader(path)
        scal = {"sha256": hashlib.sha256(raw).hexdigest(), "bytes": len(raw),
                "mesh_size": r.get_spatial_mesh_size().tolist(),
                "block_size": r.get_spatial_block_size().tolist(),
                "extent": r.get_spatial_mesh_extent().tolist(),
                "variables": r.get_all_variables(),
                "vectorsize_cellB": r.read_variable_vectorsize("cellB"),
                "parameters": {k: float(r.read_parameter(k)) for k in
                               ("time", "t", "timestep", "dt", "xmin", "ymin", "zmin", "xmax", "ymax", "zmax",
                                "xcells_ini", "ycells_ini", "zcells_ini")}}
        try:
            r.read_variable("innerFlagField")
            scal["flag_read_error"] = None
        except Exception as e:  # the reader's type table has no one-byte types
            scal["flag_read_error"] = type(e).__name__
        arrays = {"cellRhoQi": np.asarray(r.read_variable("cellRhoQi")),
                  "cellB": np.asarray(r.read_variable("cellB")),
                  "cellNPles": np.asarray(r.read_variable("cellNPles")),
                  "CellID": np.asarray(r.read_variable("CellID")),
                  "cellB_at_5_17_0": np.asarray(r.read_variable("cellB", cellids=[5, 17, 0])),
                  "cellB_mag_at_3": np.asarray(r.read_variable("cellB", cellids=[3], operator="magnitude")),
                  "coords_of_1_2_120": np.asarray(r.get_cell_coordinates(np.array([1, 2, 120])))}
    with open(os.path.join(HERE, "vlsv_reference_reader.json"), "w") as f:
        json.dump(scal, f, indent=1)
    np.savez(os.path.join(HERE, "vlsv_reference_reader.npz"), **arrays)
    print(json.dumps(scal))


if __name__ == "__main__":
    main()
