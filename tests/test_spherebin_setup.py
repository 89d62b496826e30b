"""The Spherebin experiment set-up methods (python/sphere.py:783-822,
1058-1115, 1199-1355, 1629-1655, 2892-2907), seeded: radii distributions
against their moments, the random grid placement against a pair search, the
wall fits, the shear layers, the run wrapper (dry) and readlast."""

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from sphere_amd import configs
from sphere_amd.spherebin import Spherebin, status


def test_lognormal_radii_have_the_requested_mean_and_variance():
    sb = Spherebin(200000, nw=0, sid="psd")
    m, v = 440e-6, 8.8e-9
    sb.generateRadii(radius_mean=m, radius_variance=v, seed=11)
    assert abs(sb.radius.mean() - m) < 3e-3 * m
    assert abs(sb.radius.var() - v) < 2e-2 * v
    assert np.all(sb.radius > 0)
    again = Spherebin(200000, nw=0, sid="psd")
    again.generateRadii(radius_mean=m, radius_variance=v, seed=11)
    assert np.array_equal(again.radius, sb.radius)


def test_uniform_radii_use_the_variance_as_half_width():
    sb = Spherebin(50000, nw=0, sid="uni")
    sb.generateRadii(psd="uni", radius_mean=1e-2, radius_variance=2e-3, seed=3)
    assert sb.radius.min() >= 8e-3 and sb.radius.max() <= 1.2e-2
    assert abs(sb.radius.mean() - 1e-2) < 2e-5
    with pytest.raises(ValueError):
        sb.generateRadii(psd="bimodal")


def test_random_grid_positions_do_not_overlap_and_fit_the_world():
    n = 1500
    sb = Spherebin(n, nw=0, sid="grid")
    sb.generateRadii(radius_mean=1e-2, radius_variance=4e-6, seed=5)
    sb.initRandomGridPos(gridnum=(24, 24, 40), seed=6, contactmodel=1, periodic=2)
    r, x = sb.radius, sb.x
    cell = 2.1 * r.max() * 2
    assert sb.contactmodel == 1 and sb.periodic == 2
    assert np.all(x - r[:, None] >= 0.0)
    assert np.all(x + r[:, None] <= sb.L)
    np.testing.assert_array_equal(sb.num, np.ceil(np.max(x + r[:, None], axis=0) / cell))
    np.testing.assert_allclose(sb.L, sb.num * cell)
    pairs = cKDTree(x).query_pairs(2.0 * r.max())
    for i, j in pairs:
        assert np.linalg.norm(x[i] - x[j]) > r[i] + r[j]
    # each particle stays inside its own coarse cell (index from the linear
    # index; the y index wraps with the x cell count, as in the reference)
    c = np.floor(np.array([24, 24, 40]) / 2)
    i = np.arange(n)
    g = np.stack([i % c[0], (i // c[0]) % c[0], i // (c[0] * c[1])], axis=1)
    assert np.all(np.floor(x / cell) == g)
    small = Spherebin(10000, nw=0, sid="too-many")
    small.radius[:] = 1e-2
    with pytest.raises(ValueError, match="not sufficiently large"):
        small.initRandomGridPos(gridnum=(12, 12, 32))


def _packed():
    sb = configs.lattice_settle(nside=(6, 6, 8), sid="setup")
    sb.vel[:] = 0.1
    sb.angvel[:] = 2.0
    sb.es[:] = 1.0
    return sb


def test_consolidate_fits_the_upper_wall_and_holds_the_stress():
    sb = _packed()
    zmin = float(np.min(sb.x[:, 2] - sb.radius))
    zmax = float(np.max(sb.x[:, 2] + sb.radius))
    cell = float(sb.L[0] / sb.num[0])
    sb.consolidate(deviatoric_stress=12e3)
    assert sb.nw == 1 and list(sb.wmode) == [1] and sb.w_devs[0] == 12e3
    np.testing.assert_array_equal(sb.w_n[0], [0.0, 0.0, -1.0])
    assert sb.w_x[0] == zmax
    assert sb.L[2] == (zmax - zmin) * 1.1 + zmin
    assert sb.num[2] == math.ceil(((zmax - zmin) * 1.1 + zmin) / cell)
    assert sb.w_m[0] == sb.rho * sb.np * math.pi * (cell / 2.0) ** 3
    assert not sb.vel.any() and not sb.angvel.any() and not sb.es.any()


def test_uniaxial_strain_rate_moves_the_wall_at_fixed_velocity():
    sb = _packed()
    sb.uniaxialStrainRate(wvel=-0.002)
    assert list(sb.wmode) == [2] and sb.w_vel[0] == -0.002
    assert sb.w_x[0] == float(np.max(sb.x[:, 2] + sb.radius))


def test_shear_fixes_the_bottom_and_drives_the_top_layer():
    sb = _packed()
    sb.consolidate()
    num2, L2 = int(sb.num[2]), float(sb.L[2])
    z = sb.x[:, 2].copy()
    z_min = float(np.min(z - sb.radius))
    z_max = float(np.max(z + sb.radius))
    sb.shear(shear_strain_rate=0.5)
    cell = float(sb.L[0] / sb.num[0])
    assert sb.num[2] == num2 + 1 and sb.L[2] == (num2 + 1) * cell and sb.L[2] > L2
    d_bot = 2.0 * sb.radius[z < (z_max - z_min) * 0.3].max()
    d_top = 2.0 * sb.radius[z > (z_max - z_min) * 0.7].max()
    low, high = z < z_min + d_bot, z > z_max - d_top
    assert low.any() and high.any()
    assert np.all(sb.fixvel[low | high] == 1.0) and np.all(sb.fixvel[~(low | high)] == 0.0)
    assert np.all(sb.vel[high, 0] == (z_max - z_min) * 0.5) and not sb.vel[~high].any()
    assert sb.gamma_wn == 0 and sb.gamma_wt == 0 and sb.mu_ws == 0 and sb.mu_wd == 0


def test_run_dry_and_readlast(tmp_path):
    sb = _packed()
    assert sb.run(dry=True, verbose=False, root=str(tmp_path)) == 0
    assert (tmp_path / "input" / "setup.bin").exists()
    (tmp_path / "output").mkdir()
    later = sb.copy()
    later.x[:, 2] += 0.01
    later.writebin(str(tmp_path / "output" / "setup.output00003.bin"))
    (tmp_path / "output" / "setup.status.dat").write_text("1.0000e-02 5.0000e+01 3\n")
    assert status("setup", str(tmp_path)) == 3
    sb.readlast(verbose=False, root=str(tmp_path))
    assert np.array_equal(sb.x, later.x) and sb.sid == "setup"
