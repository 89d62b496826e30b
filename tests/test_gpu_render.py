"""The raytracer on the GPU (sph_render) against the oracle's restatement of
DEM::render (src/raytracer.cuh:14-553), byte for byte, and the `sphere -r`
command line against the same image."""

import os
import subprocess
import sys

import numpy as np
import pytest

from oracle_lib import Oracle
from sphere_amd import configs
from sphere_amd.dem import DEM, SphereInvalid
from sphere_amd.spherebin import Spherebin

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("oracle_built")]

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sphere_amd")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _coloured_packing(n=2000, seed=9, sid="rt-packing"):
    """Random spheres with every colour field set; the last 20 repeat the
    first 20 exactly (same hit distance: the lower index must win) with
    other colour values, and every 7th sphere is fixed."""
    rng = np.random.default_rng(seed)
    sb = Spherebin(n, nw=1, sid=sid)
    sb.radius = rng.uniform(0.01, 0.04, n)
    sb.x = np.column_stack([rng.uniform(0.05, 0.95, n), rng.uniform(0.05, 0.95, n), rng.uniform(0.05, 0.7, n)])
    sb.x[-20:] = sb.x[:20]
    sb.radius[-20:] = sb.radius[:20]
    sb.L = np.array([1.0, 1.0, 1.0])
    sb.num = np.array([12, 12, 12], dtype=np.uint32)
    sb.set_walls(1)
    sb.w_x[0] = 0.8
    sb.vel = rng.normal(0.0, 1.0, (n, 3))
    sb.fixvel[::7] = 1.0
    sb.angvel = rng.normal(0.0, 5.0, (n, 3))
    sb.angpos = rng.normal(0.0, 2.0, (n, 3))
    sb.xysum = rng.uniform(0.0, 1e-2, (n, 2))
    sb.p = rng.uniform(0.0, 1e4, n)
    sb.k_n = sb.k_t = 1e5
    sb.rho = 2600.0
    sb.time_dt = 1e-7
    sb.time_total = 1.0
    return sb


CASES = [(0, 1e3, 0.0), (1, 1e4, 0.0), (1, 1e4, 3e3), (2, 3.0, 0.0), (2, 3.0, 1.0), (3, 10.0, 4.0),
         (4, 1e-2, 0.0), (5, 4.0, 2.0)]


def test_every_method_matches_the_oracle(tmp_path):
    sb = _coloured_packing()
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0, checkVals=0)
    for method, mx, cut in CASES:
        want = orc.render(method, mx, cut, 128, 128)
        for cull in (True, False):
            got = dem.render(method, mx, cut, 128, 128, cull=cull)
            assert np.array_equal(got, want), (method, mx, cut, cull, int((got != want).any(-1).sum()))
        assert (want != 255).any()


@pytest.mark.parametrize("shape", [(96, 64), (64, 96), (1, 1), (50, 17)])
def test_image_shapes(tmp_path, shape):
    """Wide images (zero-height plane), tall images (rows repeat modulo the
    width), a single pixel, widths off the 16-pixel tile."""
    w, h = shape
    sb = _coloured_packing(n=600, sid=f"rt-shape{w}x{h}")
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0, checkVals=0)
    for method, mx, cut in ((0, 1e3, 0.0), (2, 3.0, 0.5)):
        want = orc.render(method, mx, cut, w, h)
        assert np.array_equal(dem.render(method, mx, cut, w, h), want), (shape, method)


def test_after_steps_in_sorted_order(tmp_path):
    _say("render after steps")
    sb = configs.lattice_settle(nside=(8, 8, 8), spacing=2.15e-2, jitter=5e-4, periodic=1,
                                gamma_n=0.0, gamma_t=10.0, sid="rt-settle")
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    dem.step(300)
    assert orc.step(300) == 0
    for method, mx, cut in ((0, 1e3, 0.0), (2, 0.5, 0.0), (3, 50.0, 1.0)):
        want = orc.render(method, mx, cut, 160, 160)
        assert np.array_equal(dem.render(method, mx, cut, 160, 160), want), method


def test_refusals(tmp_path):
    sb = _coloured_packing(n=50, sid="rt-refuse")
    dem = DEM(sb, verbosity=0, checkVals=0)
    with pytest.raises(SphereInvalid):
        dem.render(6)
    with pytest.raises(SphereInvalid):
        dem.render(0, width=0)
    with pytest.raises(SphereInvalid):
        dem.render(0, width=8192, height=4096)


def test_full_size_tiles_match_every_sphere_per_pixel():
    """1,048,576 spheres (the random compression packing): the tiled
    render against the untiled one (every sphere tested for every pixel,
    the reference's loop), byte for byte."""
    _say("render full size")
    sb = configs.random_compression()
    dem = DEM(sb, verbosity=0)
    dem.step(10)
    for method, mx, cut in ((0, 1e3, 0.0), (2, 0.05, 0.0)):
        tiled = dem.render(method, mx, cut, 320, 320, cull=True)
        brute = dem.render(method, mx, cut, 320, 320, cull=False)
        assert np.array_equal(tiled, brute), method
        assert (tiled != 255).any(-1).mean() > 0.2


def test_sphere_render_command_line(tmp_path):
    sb = _coloured_packing(n=800, sid="rt-cli")
    path = tmp_path / "rt-cli.bin"
    sb.writebin(str(path))
    orc = Oracle(str(path))
    for args, (method, mx, cut) in ((["-r"], (0, 0.0, 0.0)), (["-m", "vel", "3", "-l", "1"], (2, 3.0, 1.0))):
        out = subprocess.run([os.path.join(PKG, "sphere"), "-q", *args, str(path)], cwd=tmp_path,
                             capture_output=True, timeout=120)
        assert out.returncode == 0, out.stderr
        data = (tmp_path / "img_out" / "rt-cli.ppm").read_bytes()
        header = b"P6 800 800 255\n"
        assert data[:len(header)] == header
        img = np.frombuffer(data[len(header):], dtype=np.uint8).reshape(800, 800, 3)
        assert np.array_equal(img, orc.render(method, mx, cut, 800, 800)), args
