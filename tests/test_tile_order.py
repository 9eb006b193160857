"""The XCD-aware launch orders (rhybrid_amd/csrc/rh_geom.hpp) visit every tile
exactly once for every tile-grid shape up to 13 x 11 x 9 and the production
shapes: tests/tile_order_check.cpp compiled on the host (no GPU)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_launch_orders_cover_every_tile_once(tmp_path):
    exe = str(tmp_path / "tile_order_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-o", exe, os.path.join(HERE, "tile_order_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
