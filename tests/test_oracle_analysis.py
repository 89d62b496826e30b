"""Pins of the oracle's host analysis routines: DEM::porosity
(src/sphere.cpp:363-483) and DEM::findOverlaps (src/sphere.cpp:553-604),
against textbook cap volumes, numerical integration and scipy's k-d tree."""

import math

import numpy as np
import pytest
from scipy import integrate
from scipy.spatial import cKDTree

from oracle_lib import Oracle
from sphere_amd.spherebin import Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _cap(h, r):
    """Textbook volume of a spherical cap of height h on a sphere of radius r."""
    return math.pi * h * h * (3.0 * r - h) / 3.0


def _ball(r):
    return 4.0 * math.pi * r ** 3 / 3.0


def _packing(x, r, L=(1.0, 1.0, 1.0), top=None, sid="analysis"):
    n = len(r)
    sb = Spherebin(n, nw=0 if top is None else 1, sid=sid)
    if top is None:
        sb.set_walls(0)
    else:
        sb.w_x[0] = top
    sb.L = np.asarray(L, dtype=float)
    rmax = float(np.max(r))
    sb.num = np.maximum(1, np.floor(sb.L / (2.0 * rmax))).astype(np.uint32)
    sb.periodic = 0
    sb.radius[:] = r
    sb.x[:] = x
    sb.k_n = sb.k_t = 1e5
    sb.rho = 2600.0
    sb.time_dt = 1e-6
    sb.time_total = 1.0
    return sb


def test_porosity_sphere_inside_one_slice(tmp_path):
    r = 0.05
    sb = _packing([[0.5, 0.5, 0.375]], [r], sid="por-inside")
    z, p = Oracle.from_spherebin(sb, tmp_path).porosity(4, rule=0)
    np.testing.assert_allclose(z, [0.125, 0.375, 0.625, 0.875], rtol=0, atol=1e-15)
    V = 0.25 * 1.0 * 1.0
    np.testing.assert_allclose(p, [1.0, 1.0 - _ball(r) / V, 1.0, 1.0], rtol=1e-14)


@pytest.mark.parametrize("above", [True, False])
def test_porosity_reference_rule_at_a_plane(tmp_path, above):
    """A sphere crossing the plane z = 0.25 between slices 0 and 1.  The
    slice that holds the centre removes V_sphere + cap (the reference's
    sign), the other slice the cap on its side; the exact rule removes
    V_sphere - cap and cap."""
    r, d = 0.05, 0.02
    zc = 0.25 + d if above else 0.25 - d
    sb = _packing([[0.5, 0.5, zc]], [r], sid=f"por-plane{int(above)}")
    o = Oracle.from_spherebin(sb, tmp_path)
    V = 0.25
    cap_small = _cap(r - d, r)
    own, other = (1, 0) if above else (0, 1)
    _, p0 = o.porosity(4, rule=0)
    removed0 = (1.0 - p0) * V
    np.testing.assert_allclose(removed0[own], _ball(r) + cap_small, rtol=1e-13)
    np.testing.assert_allclose(removed0[other], cap_small, rtol=1e-13)
    _, p1 = o.porosity(4, rule=1)
    removed1 = (1.0 - p1) * V
    np.testing.assert_allclose(removed1[own], _ball(r) - cap_small, rtol=1e-12)
    np.testing.assert_allclose(removed1[other], cap_small, rtol=1e-12)
    assert removed0[2] == removed0[3] == removed1[2] == removed1[3] == 0.0


def test_porosity_top_is_wall_zero(tmp_path):
    """With walls the slices end at wall 0's position, not at L_z."""
    sb = _packing([[0.5, 0.5, 0.1]], [0.02], top=0.6, sid="por-wall")
    z, p = Oracle.from_spherebin(sb, tmp_path).porosity(3, rule=0)
    np.testing.assert_allclose(z, [0.1, 0.3, 0.5], rtol=1e-15)
    np.testing.assert_allclose(p[0], 1.0 - _ball(0.02) / 0.2, rtol=1e-15)


def test_porosity_refusal_uses_last_radius(tmp_path):
    """The reference's r_max returns the last particle's radius: a thin slice
    is refused when the last sphere is large, accepted when only an earlier
    one is."""
    big, small = 0.1, 0.01
    sb = _packing([[0.3, 0.5, 0.5], [0.7, 0.5, 0.5]], [small, big], sid="por-refuse")
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.porosity(5, rule=0) == 1  # h = 0.2 <= 2 * 0.1
    assert isinstance(o.porosity(5, rule=1), tuple)  # the exact rule has no such limit
    sb2 = _packing([[0.3, 0.5, 0.5], [0.7, 0.5, 0.5]], [big, small], sid="por-accept")
    assert isinstance(Oracle.from_spherebin(sb2, tmp_path).porosity(5, rule=0), tuple)
    assert o.porosity(0, rule=0) == 2


def test_porosity_exact_rule_against_quadrature(tmp_path):
    rng = np.random.default_rng(11)
    n = 40
    r = rng.uniform(0.02, 0.06, n)
    x = np.column_stack([rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n), rng.uniform(0.08, 0.92, n)])
    sb = _packing(x, r, sid="por-quad")
    slices = 7
    z, p = Oracle.from_spherebin(sb, tmp_path).porosity(slices, rule=1)
    h = 1.0 / slices
    removed = np.zeros(slices)
    for zc, ri in zip(x[:, 2], r):
        for k in range(slices):
            a, b = max(k * h, zc - ri), min(k * h + h, zc + ri)
            if b > a:
                removed[k] += integrate.quad(lambda t: math.pi * (ri * ri - (t - zc) ** 2), a, b, epsabs=0, epsrel=1e-13)[0]
    np.testing.assert_allclose((1.0 - p) * h, removed, rtol=1e-10, atol=1e-16)
    # every sphere lies inside [0, L_z]: the slices remove the total volume
    assert (1.0 - p).sum() * h == pytest.approx(sum(_ball(v) for v in r), rel=1e-12)


def test_find_overlaps_against_kdtree(tmp_path):
    rng = np.random.default_rng(5)
    n = 600
    r = rng.uniform(0.01, 0.03, n)
    x = rng.uniform(0.0, 1.0, (n, 3))
    x[:5] = x[5:10] + 1e-3  # a few near-coincident pairs
    sb = _packing(x, r, sid="ovl-kdtree")
    ij, dn = Oracle.from_spherebin(sb, tmp_path).find_overlaps()
    cand = cKDTree(x).query_pairs(2.0 * r.max(), output_type="ndarray")
    d = x[cand[:, 1]] - x[cand[:, 0]]
    delta = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) - (r[cand[:, 0]] + r[cand[:, 1]])
    keep = delta < 0.0
    want = cand[keep]
    order = np.lexsort((want[:, 1], want[:, 0]))
    assert len(ij) > 20
    np.testing.assert_array_equal(ij, want[order])
    np.testing.assert_array_equal(dn, delta[keep][order])
    assert np.all(ij[:, 0] < ij[:, 1])


def test_find_overlaps_ignores_periodic_images(tmp_path):
    """Two spheres touching only across the periodic x boundary."""
    sb = _packing([[0.01, 0.5, 0.5], [0.99, 0.5, 0.5]], [0.02, 0.02], sid="ovl-periodic")
    sb.periodic = 1
    ij, dn = Oracle.from_spherebin(sb, tmp_path).find_overlaps()
    assert len(ij) == 0


def _bulk(sb, top):
    """Spherebin.bulkPorosity (python/sphere.py:1579-1591)."""
    V_total = sb.L[0] * sb.L[1] * top
    return (V_total - np.sum(4.0 / 3.0 * math.pi * sb.radius ** 3)) / V_total


def test_simple_cubic_mean_equals_bulk(tmp_path):
    """tests/porosity_tests.py:38-59: 10^3 spheres of radius 1 in simple cubic
    packing; the mean over 1, 2 and 4 slices equals the bulk porosity
    1 - pi/6 for both rules (the planes at 5, 10, 15 pass through centres or
    touching points, where the reference's branches are exact)."""
    n = 10
    idx = np.arange(n ** 3)
    ix, iy, iz = idx % n, (idx // n) % n, idx // (n * n)
    x = np.column_stack([ix, iy, iz]) * 2.0 + 1.0
    sb = _packing(x, np.ones(n ** 3), L=(20.0, 20.0, 20.0), sid="por-cubic")
    o = Oracle.from_spherebin(sb, tmp_path)
    bulk = _bulk(sb, 20.0)
    assert bulk == pytest.approx(1.0 - math.pi / 6.0, abs=1e-12)
    for rule in (0, 1):
        for slices in (1, 2, 4):
            _, p = o.porosity(slices, rule)
            assert np.mean(p) == pytest.approx(bulk, abs=1e-12), (rule, slices)


def test_random_grid_mean_against_bulk(tmp_path):
    """tests/porosity_tests.py:7-36 on a seeded grid of 100 spheres under a
    wall: the exact rule's mean equals the bulk porosity; the reference's rule
    falls short by 2 sum(cap) / V_total, a cap of the part below (above) the
    plane for every sphere that crosses a slice plane with its centre above
    (below) it."""
    rng = np.random.default_rng(4)
    g = np.arange(5)
    gx, gy, gz = np.meshgrid(g, g, np.arange(4), indexing="ij")
    x = np.column_stack([gx.ravel() * 0.2 + 0.1, gy.ravel() * 0.2 + 0.1, gz.ravel() * 0.18 + 0.15])
    x += rng.uniform(-0.04, 0.04, x.shape)
    r = rng.uniform(0.03, 0.05, len(x))
    top = 0.8
    sb = _packing(x, r, top=top, sid="por-grid")
    o = Oracle.from_spherebin(sb, tmp_path)
    bulk = _bulk(sb, top)
    for slices in (1, 2, 4):
        h = top / slices
        _, p1 = o.porosity(slices, 1)
        assert np.mean(p1) == pytest.approx(bulk, abs=1e-12)
        extra = 0.0
        for zc, ri in zip(x[:, 2], r):
            for k in range(1, slices):
                plane = k * h
                if zc - ri < plane < zc:
                    extra += 2.0 * _cap(plane - (zc - ri), ri)
                elif zc < plane < zc + ri:
                    extra += 2.0 * _cap(zc + ri - plane, ri)
        _, p0 = o.porosity(slices, 0)
        assert np.mean(p0) == pytest.approx(bulk - extra / (sb.L[0] * sb.L[1] * top), abs=1e-12)
        assert (extra > 0.0) == (slices > 1)


def test_spherebin_bulk_porosity():
    """Spherebin.bulkPorosity against its definition (python/sphere.py:1579-1591)."""
    sb = _packing([[0.5, 0.5, 0.5], [0.2, 0.2, 0.2]], [0.1, 0.05], L=(1.0, 2.0, 3.0), sid="bulk")
    V = 1.0 * 2.0 * 3.0
    assert sb.bulkPorosity() == pytest.approx((V - _ball(0.1) - _ball(0.05)) / V, rel=1e-14)
    sb2 = _packing([[0.5, 0.5, 0.5]], [0.1], L=(1.0, 2.0, 3.0), top=0.75, sid="bulk-wall")
    assert sb2.bulkPorosity() == pytest.approx((1.5 - _ball(0.1)) / 1.5, rel=1e-14)
