"""The reference's analysis members on the GPU (sph_porosity,
sph_find_overlaps) against the oracle's restatement of DEM::porosity
(src/sphere.cpp:363-483) and DEM::findOverlaps (src/sphere.cpp:553-604),
and the `porosity` / `forcechains` tools against the same numbers."""

import math
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

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "sphere_amd")

# porosity sums are taken in another order than the oracle's running
# subtraction: rounding only
POROSITY_RTOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _random_overlapping(n=3000, seed=3, sid="ovl-random"):
    rng = np.random.default_rng(seed)
    sb = Spherebin(n, nw=1, sid=sid)
    sb.radius = rng.uniform(0.01, 0.03, n)
    sb.x = np.column_stack([rng.uniform(0.03, 0.97, n), rng.uniform(0.03, 0.97, n), rng.uniform(0.03, 0.8, n)])
    sb.L = np.array([1.0, 1.0, 1.0])
    sb.num = np.array([16, 16, 16], dtype=np.uint32)
    sb.periodic = 0
    sb.set_walls(1)
    sb.w_x[0] = 0.85
    sb.k_n = sb.k_t = 1e5
    sb.rho = 2600.0
    sb.time_dt = 1e-7
    sb.time_total = 1.0
    return sb


def _compare(dem, orc, slices=(1, 7, 10)):
    for s in slices:
        for exact in (False, True):
            want = orc.porosity(s, 1 if exact else 0)
            if isinstance(want, int):
                with pytest.raises(SphereInvalid):
                    dem.porosity(s, exact=exact)
                continue
            z, p = dem.porosity(s, exact=exact)
            np.testing.assert_array_equal(z, want[0])
            np.testing.assert_allclose(p, want[1], rtol=POROSITY_RTOL, atol=1e-15)
    ij, dn = dem.findOverlaps()
    wij, wdn = orc.find_overlaps()
    np.testing.assert_array_equal(ij, wij)
    np.testing.assert_array_equal(dn, wdn)
    return len(ij)


def test_random_overlapping_packing(tmp_path):
    """Unsorted state straight from the file, thousands of overlaps."""
    sb = _random_overlapping()
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    n = _compare(dem, orc, slices=(1, 3, 10, 25))
    assert n > 1000


@pytest.mark.parametrize("periodic", [0, 1])
def test_after_steps_in_sorted_order(tmp_path, periodic):
    """After steps the device keeps the spheres in cell order and the wall
    has moved: the analysis must follow both."""
    _say(f"analysis after steps periodic={periodic}")
    sb = configs.lattice_settle(nside=(8, 8, 8), spacing=2.15e-2, jitter=5e-4, periodic=periodic,
                                gamma_n=0.0, gamma_t=10.0, sid=f"ana-settle{periodic}")
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    for k in (0, 300, 900):
        if k:
            dem.step(k)
            assert orc.step(k) == 0
        np.testing.assert_array_equal(dem.get("x"), orc.get("x"))
        n = _compare(dem, orc)
    assert n > 0


def test_edge_cases(tmp_path):
    # one sphere: no pairs; too many slices for the reference's rule
    sb = configs.two_spheres(distance=-1e-3, radius=0.1, sid="ana-pair")
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    assert _compare(dem, orc, slices=(1, 2, 50)) == 1
    with pytest.raises(SphereInvalid):
        dem.porosity(0)
    # no overlaps at all
    sb2 = configs.two_spheres(distance=1e-3, radius=0.1, sid="ana-apart")
    dem2 = DEM(sb2, verbosity=0)
    ij, dn = dem2.findOverlaps()
    assert ij.shape == (0, 2) and dn.shape == (0,)


def _slab_volumes(z, r, lo, hi):
    t1 = np.maximum(lo - z, -r)
    t2 = np.minimum(hi - z, r)
    v = math.pi * (r * r * (t2 - t1) - (t2 * t2 * t2 - t1 * t1 * t1) / 3.0)
    return np.where(t2 > t1, v, 0.0)


def _reference_rule(z, r, lo, hi):
    """The reference's branches (src/sphere.cpp:408-454), vectorised."""
    zl, zh = z - r, z + r
    Vs = 4.0 / 3.0 * math.pi * r * r * r
    cap = lambda h: math.pi * h * h / 3.0 * (3.0 * r - h)  # noqa: E731
    inside = (lo <= zl) & (zh <= hi)
    low_a = (lo >= z) & (lo <= zh)
    low_b = ~low_a & (lo < z) & (lo > zl)
    up_a = (hi <= z) & (hi >= zl)
    up_b = ~up_a & (hi > z) & (hi < zh)
    out = np.where(low_a, cap(zh - lo), 0.0) + np.where(low_b, Vs + cap(lo - zl), 0.0)
    out = out + np.where(up_a, cap(hi - zl), 0.0) + np.where(up_b, Vs + cap(zh - hi), 0.0)
    return np.where(inside, Vs, out)


def test_full_size_dense_lattice():
    """1,048,576 spheres (FCC, twelve overlaps per interior sphere) after a
    few steps: the pair set and every delta_n against scipy's k-d tree on
    the downloaded positions, the porosity of both rules against a
    vectorised evaluation."""
    from scipy.spatial import cKDTree

    _say("analysis full size dense lattice")
    sb = configs.dense_lattice()
    dem = DEM(sb, verbosity=0)
    dem.step(20)
    x = dem.get("x")
    ij, dn = dem.findOverlaps()
    pos, r = x[:, :3], x[:, 3]
    cand = cKDTree(pos).query_pairs(2.0 * r.max(), output_type="ndarray")
    cand.sort(axis=1)
    d = pos[cand[:, 1]] - pos[cand[:, 0]]
    delta = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) - (r[cand[:, 0]] + r[cand[:, 1]])
    keep = delta < 0.0
    want = cand[keep]
    order = np.lexsort((want[:, 1], want[:, 0]))
    assert len(ij) > 5_000_000
    np.testing.assert_array_equal(ij, want[order])
    np.testing.assert_array_equal(dn, delta[keep][order])
    top = dem.get("walls_nx")[0, 3]
    for slices in (10, 64):
        h = (top - 0.0) / slices
        V = h * sb.L[0] * sb.L[1]
        for exact, rule in ((True, _slab_volumes), (False, _reference_rule)):
            z, p = dem.porosity(slices, exact=exact)
            removed = np.array([rule(pos[:, 2], r, k * h, k * h + h).sum() for k in range(slices)])
            np.testing.assert_allclose(p, (V - removed) / V, rtol=1e-10)
    # the exact rule's slices hold every sphere below the top once
    z, p = dem.porosity(10, exact=True)
    assert ((1.0 - p) * (top / 10) * sb.L[0] * sb.L[1]).sum() == pytest.approx(
        (4.0 / 3.0 * math.pi * r ** 3).sum(), rel=1e-3)


def test_porosity_tool_prints_the_same_numbers(tmp_path):
    sb = _random_overlapping(n=2000, sid="tool-por")
    path = tmp_path / "tool-por.bin"
    sb.writebin(str(path))
    orc = Oracle(str(path))
    for args, rule in ((["-s", "8"], 0), (["-s", "8", "--exact"], 1)):
        out = subprocess.run([os.path.join(PKG, "porosity"), *args, str(path)], capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stderr
        rows = [line.split("\t") for line in out.stdout.strip().splitlines()]
        z, p = orc.porosity(8, rule)
        assert [float(a) for a, _ in rows] == [float(f"{v:g}") for v in z]
        np.testing.assert_allclose([float(b) for _, b in rows], p, rtol=1e-5)
    bad = subprocess.run([os.path.join(PKG, "porosity"), "-s", "0", str(path)], capture_output=True, text=True)
    assert bad.returncode == 1


def test_forcechains_tool(tmp_path):
    sb = _random_overlapping(n=1500, sid="tool-fc")
    path = tmp_path / "tool-fc.bin"
    sb.writebin(str(path))
    orc = Oracle(str(path))
    wij, wdn = orc.find_overlaps()
    f_n = -sb.k_n * wdn
    x = orc.get("x")
    lc = float(np.median(f_n))
    out = subprocess.run([os.path.join(PKG, "forcechains"), "-f", "txt", "-lc", str(lc), str(path)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0].startswith("x_1, [m]\ty_1, [m]\tz_1, [m]\tx_2, [m]")
    rows = np.array([[float(v) for v in line.split("\t")] for line in lines[1:]])
    sel = f_n >= lc
    assert rows.shape == (int(sel.sum()), 7)
    want = np.column_stack([x[wij[sel, 0], :3], x[wij[sel, 1], :3], f_n[sel]])
    np.testing.assert_allclose(rows, want, rtol=1e-5)
    # the gnuplot script: one arrow per contact above the cutoff (no fixed spheres here)
    gp = subprocess.run([os.path.join(PKG, "forcechains"), "-lc", str(lc), "-2d", str(path)],
                        capture_output=True, text=True, timeout=120)
    assert gp.returncode == 0, gp.stderr
    arrows = [line for line in gp.stdout.splitlines() if line.startswith("set arrow ")]
    assert len(arrows) == int((f_n > lc).sum())
    assert gp.stdout.rstrip().splitlines()[-1].startswith("plot [")


def test_reference_porosity_cases(tmp_path):
    """The packings of tests/porosity_tests.py (simple cubic 10^3, a grid
    under a wall) on the GPU: both rules against the oracle, and the exact
    rule's slice mean against the bulk porosity."""
    from test_oracle_analysis import _bulk, _packing

    n = 10
    idx = np.arange(n ** 3)
    cubic = _packing(np.column_stack([idx % n, (idx // n) % n, idx // (n * n)]) * 2.0 + 1.0, np.ones(n ** 3),
                     L=(20.0, 20.0, 20.0), sid="gpu-por-cubic")
    rng = np.random.default_rng(4)
    g = np.arange(5)
    gx, gy, gz = np.meshgrid(g, g, np.arange(4), indexing="ij")
    x = np.column_stack([gx.ravel() * 0.2 + 0.1, gy.ravel() * 0.2 + 0.1, gz.ravel() * 0.18 + 0.15])
    x += rng.uniform(-0.04, 0.04, x.shape)
    grid = _packing(x, rng.uniform(0.03, 0.05, len(x)), top=0.8, sid="gpu-por-grid")
    for sb, top in ((cubic, 20.0), (grid, 0.8)):
        orc = Oracle.from_spherebin(sb, tmp_path)
        dem = DEM(sb, verbosity=0, checkVals=0)
        _compare(dem, orc, slices=(1, 2, 4))
        for slices in (1, 2, 4):
            _, p = dem.porosity(slices, exact=True)
            assert np.mean(p) == pytest.approx(_bulk(sb, top), abs=1e-12)


def test_spherebin_porosity_matches_the_handle(tmp_path):
    """Spherebin.porosity returns (porosity, depth) as python/sphere.py's
    porosity() does, from the same GPU computation."""
    sb = _random_overlapping(n=500, sid="sb-porosity")
    p, z = sb.porosity(slices=6)
    orc = Oracle.from_spherebin(sb, tmp_path)
    wz, wp = orc.porosity(6, 0)
    np.testing.assert_array_equal(z, wz)
    np.testing.assert_allclose(p, wp, rtol=POROSITY_RTOL)
    pe, _ = sb.porosity(slices=6, exact=True)
    assert np.mean(pe) == pytest.approx(sb.bulkPorosity(), abs=1e-12)
