"""Ensembles (sph_ens_*): many independent systems stepped by one launch of
each kernel.  Every member must evolve exactly as it does in a handle of its
own (which the other parity tests pin to the CPU oracle), so members are
compared bit for bit with single-handle runs, and one member directly with
the oracle."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from sphere_amd import configs

MEMBER_FIELDS = ["x", "vel", "angvel", "force", "torque", "angpos", "xysum", "vel0", "angvel0", "es_dot", "es",
                 "ev_dot", "ev", "p", "contacts", "delta_t", "time"]
WALL_FIELDS = ["walls_nx", "walls_mvfd", "walls_vel0", "wall_pp"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _members():
    """Uneven members: different sizes, periodicities, walls, contact
    parameters and time steps."""
    return [
        configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=1, gamma_n=0.0,
                               gamma_t=10.0, sid="m0"),
        configs.lattice_settle(nside=(5, 4, 6), spacing=2.15e-2, jitter=5e-4, periodic=0, k_n=3e5, k_t=3e5,
                               gamma_n=20.0, mu=0.3, seed=11, sid="m1"),
        configs.two_spheres(distance=-0.02, v=(-0.3, 0.1, 0.0), w=(0.0, 0.0, 2.0), k=1e5, sid="m2"),
        configs.lattice_settle(nside=(7, 3, 4), spacing=2.25e-2, jitter=5e-4, periodic=2, nw=1, mu=0.7, seed=12,
                               sid="m3"),
        configs.ensemble_member(n=400, k_n=2e5, mu=0.4, packing=0.34, seed=13, sid="m4"),
        configs.lattice_settle(nside=(6, 6, 6), spacing=2.15e-2, jitter=5e-4, periodic=1, nw=1, gamma_n=5.0,
                               k_n=2e6, k_t=2e6, seed=14, sid="m5"),
    ]


def _fields(sb):
    return MEMBER_FIELDS + (WALL_FIELDS if sb.nw else [])


def _compare(ens, k, dem, fields):
    errs = []
    for f in fields:
        a, b = ens.get(k, f), dem.get(f)
        if a.shape != b.shape:
            errs.append(f"member {k} {f}: shape {a.shape} vs {b.shape}")
            continue
        bad = np.argwhere(a.view(np.uint64) != b.view(np.uint64)) if a.dtype == np.float64 else np.argwhere(a != b)
        if bad.size:
            i = tuple(bad[0])
            errs.append(f"member {k} {f}: {len(bad)} differ, first at {i}: ensemble {a[i]!r} alone {b[i]!r}")
    return errs


def regridded(sb, factor):
    """The grid sph_ens_open gives a member for cell_factor >= 1."""
    edge = factor * 2.0 * float(np.max(sb.radius))
    num = []
    for a in range(3):
        c = math.floor(sb.L[a] / edge)
        per = (a == 0 and sb.periodic in (1, 2)) or (a == 1 and sb.periodic == 1)
        num.append(max(c, 3 if per else 1))
    sb.num = np.array(num, dtype=np.uint32)
    return sb


def test_calibration_ensemble_is_seeded():
    a = configs.calibration_ensemble(members=3, n_lo=64, n_hi=128, steps=10)
    b = configs.calibration_ensemble(members=3, n_lo=64, n_hi=128, steps=10)
    for x, y in zip(a, b):
        assert x.np == y.np and 64 <= x.np <= 128
        assert np.array_equal(x.x, y.x) and x.k_n == y.k_n and x.mu_s == y.mu_s
        assert 1e5 <= x.k_n <= 1e7 and 0.2 <= x.mu_s <= 0.8
        assert x.periodic == 1 and x.nw == 1 and x.wmode[0] == 1
        assert np.all(x.L / x.num >= 2.0 * np.max(x.radius))


def test_regrid_rule_keeps_cells_wide_enough():
    for f in (1.0, 1.25, 1.5, 2.0):
        g = regridded(configs.ensemble_member(n=200, seed=5), f)
        edge = f * 2.0 * float(np.max(g.radius))
        cell = g.L / g.num
        floor_cells = np.floor(g.L / edge)
        # cells at least the requested edge wherever the minimum count did not force more
        assert np.all((cell >= edge) | (floor_cells < np.array([3, 3, 1])))
        assert np.all(g.num >= np.array([3, 3, 1]))


@pytest.mark.gpu
@pytest.mark.parametrize("search", [2, 4, 5, 0])
def test_members_match_single_handles(search):
    from sphere_amd.dem import DEM, Ensemble

    _say(f"ensemble of 6 vs single handles, search {search}")
    ms = _members()
    ens = Ensemble(ms)
    ens.set_search(search)
    if search == 2:
        ens.set_blocks(128, 64)
    singles = [DEM(sb, verbosity=0) for sb in ms]
    errs = []
    for n in (1, 30, 200):
        ens.step(n)
        for d in singles:
            d.step(n)
        for k, (sb, d) in enumerate(zip(ms, singles)):
            errs += [f"after +{n}: {e}" for e in _compare(ens, k, d, _fields(sb))]
            if ens.member_info(k).t_current != d.info().t_current:
                errs.append(f"after +{n}: member {k} time {ens.member_info(k).t_current} vs {d.info().t_current}")
        if errs:
            break
    for d in singles:
        d.close()
    ens.close()
    assert not errs, errs[:8]


@pytest.mark.gpu
def test_mixed_members_match_single_handles():
    from sphere_amd.dem import DEM, Ensemble

    _say("mixed-precision ensemble vs single handles")
    ms = _members()
    ens = Ensemble(ms)
    ens.set_precision("mixed")
    singles = []
    for sb in ms:
        d = DEM(sb, verbosity=0)
        d.set_precision("mixed")
        singles.append(d)
    ens.step(150)
    errs = []
    for k, (sb, d) in enumerate(zip(ms, singles)):
        d.step(150)
        errs += _compare(ens, k, d, _fields(sb))
        d.close()
    ens.close()
    assert not errs, errs[:8]


@pytest.mark.gpu
def test_autotuned_grid_and_blocks():
    """cell_factor 0: the autotuner picks the grid; members then match single
    handles opened on the same regridded input."""
    from sphere_amd.dem import DEM, Ensemble

    _say("ensemble autotuner")
    ms = _members()
    ens = Ensemble(ms, cell_factor=0.0)
    info = ens.info()
    assert info.cell_factor in (1.0, 1.25, 1.5, 2.0)
    assert all(t > 0.0 for t in info.tuned_ms)
    ens.step(60)  # the first call also tunes the search kernel and block sizes
    info = ens.info()
    assert info.search_mode in (2, 4, 5) and info.force_block in (64, 128, 256)
    assert info.search_block in (64, 128, 256)
    errs = []
    for k, sb in enumerate(ms):
        g = regridded(sb, info.cell_factor)
        assert tuple(ens.member_info(k).num) == tuple(int(v) for v in g.num)
        d = DEM(g, verbosity=0)
        d.step(60)
        errs += _compare(ens, k, d, _fields(sb))
        d.close()
    ens.close()
    assert not errs, errs[:8]


@pytest.mark.gpu
def test_member_matches_oracle(tmp_path):
    """One member of an ensemble against the CPU oracle directly."""
    from oracle_lib import Oracle
    from parity import STATE_FIELDS, first_mismatch
    from sphere_amd.dem import Ensemble

    _say("ensemble member vs oracle")
    ms = _members()
    k = 5
    orc = Oracle.from_spherebin(ms[k], tmp_path)
    ens = Ensemble(ms)
    ens.set_search(5)
    errs = []
    for n in (1, 120):
        ens.step(n)
        assert orc.step(n) == 0
        for f in STATE_FIELDS + ["contacts", "delta_t", "walls_nx", "walls_mvfd", "wall_pp"]:
            m = first_mismatch(f, ens.get(k, f), orc.get(f))
            if m:
                errs.append(m)
    ens.close()
    orc.close()
    assert not errs, errs


@pytest.mark.gpu
def test_ensemble_refuses_bonds_and_grid_walking_kernels():
    from sphere_amd.dem import Ensemble, SphereInvalid, Unsupported

    with pytest.raises(Unsupported):
        Ensemble([configs.two_spheres(sid="a"), configs.two_spheres(bonded=True, sid="b")])
    m1 = configs.two_spheres(sid="c")
    m1.contactmodel = 1
    with pytest.raises(Unsupported):
        Ensemble([configs.two_spheres(sid="a"), m1])
    ens = Ensemble([configs.two_spheres(sid="a")])
    for mode in (1, 3):
        with pytest.raises(SphereInvalid):
            ens.set_search(mode)
    with pytest.raises(Unsupported):
        ens.get(0, "cell_keys")
    ens.close()


@pytest.mark.gpu
def test_calibration_ensemble_members_match_single_handles():
    """A 48-member slice of the calibration ensemble (uneven counts up to
    32,768, k_n and mu spread, mixed precision, autotuned grid), 120 steps:
    a sample of members bit-identical to their own handles on the same
    regridded inputs."""
    from sphere_amd.dem import DEM, Ensemble

    _say("calibration ensemble sample")
    ms = configs.calibration_ensemble(members=48, steps=1000)
    ens = Ensemble(ms, cell_factor=0.0)
    ens.set_precision("mixed")
    ens.step(120)
    info = ens.info()
    assert info.members == 48
    errs = []
    for k in (0, 7, 23, 47):
        d = DEM(regridded(ms[k], info.cell_factor), verbosity=0)
        d.set_precision("mixed")
        d.step(120)
        errs += _compare(ens, k, d, _fields(ms[k]))
        d.close()
    ens.close()
    assert not errs, errs[:8]


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_calibration_ensemble_full_size(tmp_path):
    """The calibration ensemble at its stated size: 512 members (uneven
    counts in [1,024, 32,768], about 8.5M spheres), autotuned grid, 1,000
    mixed-precision steps in one handle -- every member bit-identical to its
    own handle on the regridded input after the same steps.  Then the same
    ensemble in FP64 for 1,000 steps, its two largest members against the
    CPU oracle."""
    from oracle_lib import Oracle
    from parity import STATE_FIELDS, first_mismatch
    from sphere_amd.dem import DEM, Ensemble

    steps = 1000
    ms = configs.calibration_ensemble(members=512, steps=steps)
    n_total = sum(m.np for m in ms)
    _say(f"calibration ensemble full size: {len(ms)} members, {n_total} spheres")
    ens = Ensemble(ms, cell_factor=0.0)
    factor = ens.info().cell_factor
    ens.set_precision("mixed")
    ens.step(steps)
    errs = []
    for k, m in enumerate(ms):
        d = DEM(regridded(m.copy(), factor), verbosity=0)
        d.set_precision("mixed")
        # every kernel variant gives the same bits: the ensemble's choice,
        # fixed, spares each single handle its timing trials
        d.set_search(5)
        d.set_force_kernel(1)
        d.step(steps)
        errs += _compare(ens, k, d, _fields(m))
        d.close()
        if k % 16 == 15:
            _say(f"  {k + 1} members compared")
    ens.close()
    assert not errs, errs[:8]
    _say("FP64 ensemble against the oracle, two largest members")
    ens = Ensemble(ms, cell_factor=factor)
    ens.step(steps)
    for k in np.argsort([m.np for m in ms])[-2:]:
        orc = Oracle.from_spherebin(regridded(ms[k].copy(), factor), tmp_path)
        assert orc.step(steps) == 0
        for f in STATE_FIELDS + ["contacts", "delta_t", "walls_nx", "walls_mvfd", "wall_pp"]:
            mm = first_mismatch(f, ens.get(int(k), f), orc.get(f))
            if mm:
                errs.append(f"member {k}: {mm}")
        orc.close()
    ens.close()
    assert not errs, errs[:8]
