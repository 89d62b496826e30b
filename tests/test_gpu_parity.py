"""GPU vs CPU-oracle parity of the FP64 step, bit for bit, through the C-ABI."""

import math
import sys

import numpy as np
import pytest

from parity import run_pair
from sphere_amd import configs

pytestmark = pytest.mark.gpu


# steps before the hand-over to the oracle, and steps checked after it, of
# the bonded solid's breakage window
BONDED_WINDOW = (4800, 200)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_fp64_sqrt_div_correctly_rounded(tmp_path):
    """The parity of the whole path rests on IEEE-correct f64 sqrt and divide on
    gfx950: a two-sphere contact's overlap goes through both; compare one step
    over many random separations and radii against the oracle."""
    rng = np.random.default_rng(7)
    for trial in range(8):
        d = float(rng.uniform(-0.2, -1e-6))
        sb = configs.two_spheres(distance=d, radius=float(rng.uniform(0.1, 0.5)),
                                 v=tuple(rng.uniform(-1e-3, 1e-3, 3)), w=tuple(rng.uniform(-1, 1, 3)),
                                 gamma=float(rng.uniform(0, 10)), k=1e5, sid=f"pair{trial}")
        errs = run_pair(sb, tmp_path, [1, 3])
        assert not errs, errs


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_small_settle_parity(tmp_path, periodic):
    _say(f"small settle periodic={periodic}")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=periodic,
                                gamma_n=0.0, gamma_t=10.0, sid=f"settle-p{periodic}")
    errs = run_pair(sb, tmp_path, [1, 10, 200, 800])
    assert not errs, errs


@pytest.mark.parametrize("search", [0, 5])
def test_c0_settle_parity(tmp_path, search):
    """4,096-sphere x/y-periodic settling case, 2,000 steps, checked at
    intervals; with the automatic kernel choice and with neighbour lists
    forced (the falling spheres trigger list rebuilds)."""
    _say(f"c0 settle 4096 search={search}")
    sb = configs.lattice_settle()
    errs = run_pair(sb, tmp_path, [1, 99, 400, 1500], setup=lambda d: d.set_search(search))
    assert not errs, errs


@pytest.mark.parametrize("periodic,L", [(0, 20.0), (1, 4.22)], ids=["shell-pass", "small-periodic-grid"])
def test_lists_cover_pairs_beyond_the_build_scan(tmp_path, periodic, L):
    """Neighbour lists when a cell is barely wider than the largest pair: two
    spheres two cells apart along z at the list build (outside its 27-cell
    scan) but within r_a + r_b + 0.9 skin close in on each other.  The build
    must list them (its pass over the cells two away), or, on a periodic axis
    of 4 cells where that pass would visit a cell twice, rebuild before they
    can overlap."""
    _say(f"lists: pair two cells apart at the build, periodic={periodic}")
    sb = configs.two_spheres(k=1e7, periodic=periodic, sid=f"list-gap-{periodic}")
    sb.L = np.array([L, L, 20.0])
    sb.num = np.maximum(1, np.floor(sb.L / (2.0 * 0.5 * 1.0001))).astype(np.uint32)
    assert periodic == 0 or sb.num[0] == 4
    cell = sb.L[2] / sb.num[2]
    # a at the top of its cell, b 1.08 (r_a + r_b = 1, skin = cell / 10) away,
    # in the cell after next; they approach at 10 m/s each
    za = (np.floor(10.0 / cell) + 1.0) * cell - 0.02
    sb.x[0] = (0.5 * L, 0.5 * L, za)
    sb.x[1] = (0.5 * L, 0.5 * L, za + 1.08)
    assert int(sb.x[1, 2] // cell) - int(sb.x[0, 2] // cell) == 2
    sb.vel[0] = (0.0, 0.0, 10.0)
    sb.vel[1] = (0.0, 0.0, -10.0)
    steps = [1] + [5] * 20
    errs = run_pair(sb, tmp_path, steps, setup=lambda d: d.set_search(5))
    assert not errs, errs


def test_lists_with_contacts_crossing_the_periodic_faces(tmp_path):
    """A crystal (every nearest-neighbour pair in contact) drifting along x
    and y through its periodic faces under the neighbour-list search: the
    list steps skip held contacts, except after a step in which a sphere
    crossed a face, when the image codes of its contacts change.  Contacts,
    image codes and springs against the oracle, which re-tests every pair."""
    _say("lists: drifting crystal through the periodic faces")
    sb = configs.dense_lattice(cells=(6, 6, 4), sid="drift")
    sb.vel[:, 0], sb.vel[:, 1] = 5.0, 3.0  # 0.09 and 0.05 mm per step; faces 7 mm from the nearest spheres
    errs = run_pair(sb, tmp_path, [1, 49, 100, 150, 100], setup=lambda d: d.set_search(5))
    assert not errs, errs


@pytest.mark.parametrize("search", [2, 4, 5])
def test_position_update_midway(tmp_path, search):
    """sph_set of positions between steps (cell ids and neighbour lists must
    follow): a sphere is moved into contact with another one."""
    from oracle_lib import Oracle
    from parity import all_fields, compare_all
    from sphere_amd.dem import DEM

    _say(f"position update search={search}")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=1,
                                gamma_n=0.0, gamma_t=10.0, sid=f"setx-{search}")
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    dem.set_search(search)
    dem.step(40)
    assert orc.step(40) == 0
    x = dem.get("x")
    assert np.array_equal(x, orc.get("x"))
    # sphere 7 slides along x onto its neighbour 8, overlapping it slightly
    x[7, :3] = x[8, :3] - np.array([0.999 * (x[7, 3] + x[8, 3]), 0.0, 0.0])
    dem.set("x", x)
    orc.set("x", x)
    dem.step(60)
    assert orc.step(60) == 0
    errs = compare_all(dem, orc, all_fields(orc))
    dem.close()
    orc.close()
    assert not errs, errs


def test_c0_blowup_same_step(tmp_path):
    """The reference's model-2 normal dashpot adds energy (vel_n = -dot(v, n),
    src/contactmodels.cuh:341,366), so the 4,096-sphere settling case leaves
    the grid once the lowest layer collides.  The GPU must fail at the same
    step as the oracle, with the same state up to that step."""
    from oracle_lib import Oracle
    from sphere_amd.dem import DEM, OutOfGrid

    _say("c0 blow-up step")
    sb = configs.lattice_settle()
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    dem.step(2000)
    assert orc.step(2000) == 0
    gpu_fail = orc_fail = None
    for k in range(2001, 3001):
        if orc_fail is None and orc.step(1):
            orc_fail = k
        if gpu_fail is None:
            try:
                dem.step(1)
            except OutOfGrid:
                gpu_fail = k
        if gpu_fail and orc_fail:
            break
    dem.close()
    assert orc_fail is not None
    assert gpu_fail == orc_fail


def test_top_wall_stress_parity(tmp_path):
    _say("stress wall")
    sb = configs.lattice_settle(nside=(6, 6, 6), spacing=2.25e-2, jitter=5e-4, periodic=1, nw=1,
                                gamma_n=0.0, gamma_t=10.0, sid="wall")
    sb.wmode = np.array([1], dtype=np.int32)
    sb.w_devs[0] = 5e3
    sb.w_x[0] = float(np.max(sb.x[:, 2] + sb.radius))
    sb.w_m[0] = 10.0
    errs = run_pair(sb, tmp_path, [1, 50, 500])
    assert not errs, errs


def test_bond_pair_parity(tmp_path):
    _say("bond pair")
    sb = configs.two_spheres(distance=0.0, radius=0.5, v=(1e-4, 2e-5, -3e-5), w=(1e-4, -2e-4, 5e-5),
                             bonded=True, k=1.16e9, sid="bondpair")
    errs = run_pair(sb, tmp_path, [1, 100, 1000])
    assert not errs, errs


def test_bonded_lattice_parity(tmp_path):
    """Bond strengths low enough that bonds break from about step 100 on
    (32 of 620 broken by step 300 in the oracle): breakage indices and the
    surviving bonds' state bit for bit."""
    _say("bonded fcc")
    sb = configs.fcc_bonded(cells=(3, 3, 4), sigma_b=3e4, tau_b=3e4)
    errs = run_pair(sb, tmp_path, [1, 50, 150, 150])
    assert not errs, errs


def test_bonded_solid_full_size(tmp_path):
    """The 262,144-sphere bonded FCC solid (about 1.5M bonds) at full size,
    with strengths that break bonds within the first steps, against the
    oracle for 25 steps."""
    from oracle_lib import Oracle

    _say("bonded solid 262,144 spheres")
    sb = configs.fcc_bonded(cells=(32, 32, 64), sigma_b=2e3, tau_b=2e3)
    assert sb.np == 262144 and sb.nb0 > 1_400_000
    errs = run_pair(sb, tmp_path, [1, 9, 15])
    assert not errs, errs


def test_bonded_solid_breakage_window_full_size(tmp_path):
    """The benchmark's bonded solid at full size (262,144 spheres, about 1.5M
    bonds, strengths 1e5 Pa, top and bottom layers held at +-0.005 m/s):
    the GPU runs the first BONDED_WINDOW[0] steps, hands its whole state to
    the oracle, and both run on through the window in which bonds break --
    breakage indices and every array bit for bit."""
    from oracle_lib import Oracle
    from parity import all_fields, compare_all, hand_over
    from sphere_amd.dem import DEM

    start, length = BONDED_WINDOW
    _say(f"bonded solid breakage window {start}..{start + length}")
    sb = configs.fcc_bonded(cells=(32, 32, 64), sigma_b=1e5, tau_b=1e5)
    dem = DEM(sb, verbosity=0)
    dem.step(start)
    orc = Oracle.from_spherebin(sb, tmp_path)
    hand_over(dem, orc)
    broken0 = int((dem.get("bonds")[:, 0] >= sb.np).sum())
    done = 0
    while done < length:
        k = min(50, length - done)
        dem.step(k)
        assert orc.step(k) == 0
        done += k
        errs = compare_all(dem, orc, all_fields(orc))
        assert not errs, (start + done, errs)
    broken = int((orc.get("bonds")[:, 0] >= sb.np).sum())
    _say(f"broken bonds {broken0} -> {broken}")
    dem.close()
    orc.close()
    assert broken > broken0


def test_set_bonds_checks_pairs_and_rebuilds_the_bond_lists(tmp_path):
    """sph_set(SPH_BONDS) mid-run: a pair naming a particle >= np is refused
    and leaves the bonds as they were; a valid new pair list (one bond
    broken by hand, one re-pointed at another neighbour) is what the
    following steps use, bond forces included -- GPU and oracle, given the
    same pairs, stay bit-identical."""
    from oracle_lib import Oracle
    from parity import all_fields, compare_all
    from sphere_amd.dem import DEM, SphereInvalid

    _say("sph_set bonds")
    sb = configs.fcc_bonded(cells=(3, 3, 4), sigma_b=1e9, tau_b=1e9)
    orc = Oracle.from_spherebin(sb, tmp_path)
    dem = DEM(sb, verbosity=0)
    dem.step(20)
    assert orc.step(20) == 0
    pairs = dem.get("bonds")
    assert np.array_equal(pairs, orc.get("bonds"))
    bad = pairs.copy()
    bad[3, 1] = sb.np
    with pytest.raises(SphereInvalid):
        dem.set("bonds", bad)
    assert np.array_equal(dem.get("bonds"), pairs)
    new = pairs.copy()
    new[0, 0] = sb.nb0  # broken
    i = int(new[1, 0])
    others = [int(j) for j in new[new[:, 0] == i][:, 1] if int(j) != int(new[1, 1])]
    new[1, 1] = others[0]  # bond 1 now joins i to another of its neighbours
    dem.set("bonds", new)
    orc.set("bonds", new)
    dem.step(30)
    assert orc.step(30) == 0
    errs = compare_all(dem, orc, all_fields(orc))
    dem.close()
    orc.close()
    assert not errs, errs


def test_first_call_tunes_on_a_copy(tmp_path):
    """With the kernel choice automatic, the first call of any length times
    the kernels on a copy of the state and restores it: a 2-step call runs 2
    steps of the chosen kernels, and the results stay the oracle's."""
    from sphere_amd.dem import DEM

    _say("first call tunes on a copy")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=1, gamma_n=0.0,
                                gamma_t=10.0, sid="tune-copy")
    errs = run_pair(sb, tmp_path, [2, 1, 40])
    assert not errs, errs
    dem = DEM(sb, verbosity=0)
    dem.step(2)
    info = dem.info()
    assert info.steps_done == 2
    assert abs(dem.get("time")[1] - 2 * sb.time_dt) < 1e-18
    dem.set_profiling(True)
    dem.step(1)
    assert dem.timing()["search_mode"] in (2, 4, 5)
    dem.close()


SEARCH_VARIANTS = [
    ("rows", lambda d: d.set_search(2)),
    ("rows-vw", lambda d: (d.set_search(2), d.set_force_kernel(2))),
    ("lists-vw", lambda d: (d.set_search(5), d.set_force_kernel(2))),
    ("wave", lambda d: d.set_search(3)),
    ("half", lambda d: d.set_search(4)),
    ("lists", lambda d: d.set_search(5)),
    ("tiles-auto", lambda d: d.set_search(1)),
    ("tiles-1x1", lambda d: (d.set_search(1), d.set_tile(1, 1))),
    ("tiles-wide", lambda d: (d.set_search(1), d.set_tile(6, 3))),
]


COARSE_GRIDS = [((2, 2, 2), 0), ((1, 3, 2), 0), ((2, 1, 1), 0), ((3, 3, 2), 1), ((4, 3, 1), 1),
                ((3, 1, 2), 2), ((3, 2, 2), 2)]


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("num,periodic", COARSE_GRIDS, ids=[f"{g[0]}-p{g[1]}" for g in COARSE_GRIDS])
def test_coarse_grid_search_parity(tmp_path, mode, num, periodic):
    """Cells far larger than the spheres: 10-40 members per cell, more than
    128 candidates per neighbourhood, one- and two-cell axes, and three-cell
    periodic axes where a neighbourhood wraps onto both ends of the row."""
    _say(f"coarse grid num={num} periodic={periodic} search={mode}")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=periodic,
                                gamma_n=0.0, gamma_t=10.0, sid=f"coarse-{mode}")
    sb.num = np.array(num, dtype=np.uint32)
    errs = run_pair(sb, tmp_path, [1, 50, 200], setup=lambda d: d.set_search(mode))
    assert not errs, errs


@pytest.mark.parametrize("name,setup", SEARCH_VARIANTS, ids=[v[0] for v in SEARCH_VARIANTS])
@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_search_variants_parity(tmp_path, name, setup, periodic):
    """Both contact-search kernels (LDS tiles of any shape, one thread per
    sorted particle) give the oracle's slots bit for bit, including tiles
    wider than the grid where periodic images repeat staged cells."""
    _say(f"search {name} periodic={periodic}")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=periodic,
                                gamma_n=0.0, gamma_t=10.0, sid=f"search-{name}-{periodic}")
    errs = run_pair(sb, tmp_path, [1, 50, 300], setup=setup)
    assert not errs, errs


def test_readbin_replaces_the_simulation(tmp_path):
    """sph_readbin on an open handle (DEM::readbin): the new file's state,
    step count reset, the handle's settings kept; a missing file leaves the
    handle as it was."""
    from sphere_amd.dem import DEM, SphereIOError

    _say("readbin on an open handle")
    a = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=1, gamma_n=0.0,
                               gamma_t=10.0, sid="rb-a")
    b = configs.lattice_settle(nside=(5, 5, 4), spacing=2.15e-2, jitter=5e-4, periodic=0, nw=1, seed=3, sid="rb-b")
    pa, pb = str(tmp_path / "rb-a.bin"), str(tmp_path / "rb-b.bin")
    a.writebin(pa)
    b.writebin(pb)
    d = DEM(pa, verbosity=0)
    d.set_precision("mixed")
    d.set_search(5)
    d.step(20)
    d.readbin(pb)
    info = d.info()
    assert info.np == b.np and info.steps_done == 0 and info.precision == 1
    ref = DEM(pb, verbosity=0)
    ref.set_precision("mixed")
    ref.set_search(5)
    d.step(50)
    ref.step(50)
    for f in ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t", "walls_nx", "walls_mvfd", "time"):
        assert np.array_equal(d.get(f), ref.get(f)), f
    x = d.get("x")
    with pytest.raises(SphereIOError):
        d.readbin(str(tmp_path / "missing.bin"))
    assert np.array_equal(d.get("x"), x)
    d.step(5)
    ref.step(5)
    assert np.array_equal(d.get("x"), ref.get("x"))
    d.close()
    ref.close()


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_contact_model_1_parity(tmp_path, periodic):
    """Contact model 1 (contactLinearViscous, found and processed in the
    27-cell scan, no slots): every array bit for bit, the slots untouched."""
    _say(f"contact model 1 periodic={periodic}")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=periodic, gamma_n=5.0,
                                gamma_t=10.0, sid=f"m1-{periodic}")
    sb.contactmodel = 1
    errs = run_pair(sb, tmp_path, [1, 50, 300])
    assert not errs, errs


def test_capillary_cohesion_parity(tmp_path):
    """Contact model 1 with the capillary bond between spheres closer than
    db (capillaryCohesion_exp and its exponential) against the oracle."""
    _say("capillary cohesion")
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.3e-2, jitter=5e-4, periodic=1, gamma_n=5.0, gamma_t=10.0,
                                sid="m1-capillary")
    sb.contactmodel = 1
    sb.kappa, sb.db, sb.V_b = 0.05, 2e-3, 1e-9
    errs = run_pair(sb, tmp_path, [1, 50, 300])
    assert not errs, errs


def _hertz(sb, gamma_r=2e-3, mu_r=0.05):
    """Contact model 3 with its rolling resistance on; Hertz stiffness k_n =
    k_t = 1e7 N/m^(3/2) (as stiff as the linear 1e6 N/m at 1e-2 m of overlap)."""
    sb.contactmodel = 3
    sb.k_n = sb.k_t = 1e7
    sb.gamma_r, sb.mu_r = gamma_r, mu_r
    sb.time_dt *= 0.5  # the walls keep the linear law, now with k_n = 1e7 N/m
    return sb


@pytest.mark.parametrize("periodic", [0, 1, 2])
def test_contact_model_3_parity(tmp_path, periodic):
    """Contact model 3 (contactHertz for a positive overlap, with rolling
    resistance) on a settling lattice: every array bit for bit against the
    oracle."""
    _say(f"contact model 3 periodic={periodic}")
    sb = _hertz(configs.lattice_settle(nside=(6, 6, 5), spacing=2.25e-2, jitter=5e-4, periodic=periodic,
                                       gamma_n=5.0, gamma_t=10.0, sid=f"m3-{periodic}"))
    errs = run_pair(sb, tmp_path, [1, 50, 300, 1200])
    assert not errs, errs


@pytest.mark.parametrize("search", [2, 4, 5])
def test_contact_model_3_sheared_layer_parity(tmp_path, search):
    """A sheared FCC block (top layer driven, spheres rolling against each
    other) under contact model 3 with each search kernel."""
    _say(f"contact model 3 shear search={search}")
    sb = _hertz(configs.fcc_shear(cells=(4, 4, 8), gamma_n=20.0, sid=f"m3-shear-{search}"), gamma_r=5e-3)
    errs = run_pair(sb, tmp_path, [1, 20, 40], setup=lambda d: d.set_search(search))
    assert not errs, errs


def test_contact_model_3_is_fp64_only():
    """The mixed mode has no FP32 form of the Hertz law: refused."""
    from sphere_amd.dem import DEM, Unsupported

    dem = DEM(_hertz(configs.two_spheres(sid="hertz")), verbosity=0)
    with pytest.raises(Unsupported, match="FP64"):
        dem.set_precision("mixed")
    dem.close()


@pytest.mark.parametrize("workload", ["c1-compression", "dense-crystal"])
def test_full_size_steps_match_oracle(tmp_path, workload):
    """The benchmark's 1,048,576-sphere inputs at full size: every array of
    the step bit-identical to the oracle after 1 and 3 steps (the random
    compression packing with its stress wall, and the FCC crystal with twelve
    contacts per sphere)."""
    _say(f"full size parity {workload}")
    sb = configs.random_compression() if workload == "c1-compression" else configs.dense_lattice()
    errs = run_pair(sb, tmp_path, [1, 2])
    assert not errs, errs


def _builds(dem):
    return int(dem.timing()["list_builds"])


@pytest.mark.parametrize("workload", ["c1-compression", "dense-crystal"])
def test_list_search_across_a_rebuild_full_size(tmp_path, workload):
    """Neighbour lists (search mode 5) forced on the 1,048,576-sphere inputs,
    bit-identical to the oracle's full search across a list rebuild.  The
    random compression packing falls under gravity until a sphere has moved
    far enough to request a rebuild: a first handle finds the step of the
    second build, a second one runs to three steps before it, hands its state
    to the oracle, and both go on through the rebuild.  The crystal barely
    moves, so there the rebuild is requested midway (sph_set_search(5))."""
    from oracle_lib import Oracle
    from parity import all_fields, compare_all, hand_over
    from sphere_amd.dem import DEM

    _say(f"list search across a rebuild, {workload}")
    if workload == "dense-crystal":
        sb = configs.dense_lattice()
        dem = DEM(sb, verbosity=0)
        dem.set_search(5)
        dem.set_force_kernel(1)
        orc = Oracle.from_spherebin(sb, tmp_path)
        for k, rebuild in ((2, False), (1, True), (2, False)):
            if rebuild:
                dem.set_search(5)
            dem.step(k)
            assert orc.step(k) == 0
            errs = compare_all(dem, orc, all_fields(orc))
            assert not errs, errs
        assert _builds(dem) == 2
        dem.close()
        orc.close()
        return

    def lists_only(d):  # lists and one force kernel: no timing trials
        d.set_search(5)
        d.set_force_kernel(1)

    sb = configs.random_compression()
    probe = DEM(sb, verbosity=0)
    lists_only(probe)
    done = 0
    while _builds(probe) < 2:
        probe.step(200)
        done += 200
        assert done <= 20000, "no list rebuild within 20,000 steps"
    probe.close()
    probe = DEM(sb, verbosity=0)
    lists_only(probe)
    probe.step(done - 200)
    at = done - 200
    while _builds(probe) < 2:
        probe.step(1)
        at += 1
    probe.close()
    _say(f"second list build during step {at}")
    assert at > 3
    dem = DEM(sb, verbosity=0)
    lists_only(dem)
    dem.step(at - 3)
    assert _builds(dem) == 1
    orc = Oracle.from_spherebin(sb, tmp_path)
    hand_over(dem, orc)
    for k in (2, 2, 2):
        dem.step(k)
        assert orc.step(k) == 0
        errs = compare_all(dem, orc, all_fields(orc))
        assert not errs, errs
    assert _builds(dem) == 2
    dem.close()
    orc.close()


def test_mixed_precision_stress_error_at_full_size():
    """Mixed precision (FP32 pair forces) on the 1,048,576-sphere dense
    crystal, 300 steps: the normal stress on the top wall (its per-particle
    contact forces summed over the wall area) within 1e-6 relative of the
    FP64 run, positions within 1e-4 of the smallest radius."""
    from sphere_amd.dem import DEM

    _say("mixed precision error at full size")
    sb = configs.dense_lattice()
    area = float(sb.L[0] * sb.L[1])
    out = {}
    for prec in ("fp64", "mixed"):
        dem = DEM(sb, verbosity=0)
        dem.set_precision(prec)
        dem.step(300)
        out[prec] = (float(np.sum(dem.get("wall_pp")[0])) / area, dem.get("x")[:, :3].copy())
        dem.close()
    s64, x64 = out["fp64"]
    smx, xmx = out["mixed"]
    assert s64 > 1e4  # the crystal presses on its wall
    assert abs(smx - s64) / abs(s64) <= 1e-6, (s64, smx)
    assert np.abs(xmx - x64).max() <= 1e-4 * float(sb.radius.min())
