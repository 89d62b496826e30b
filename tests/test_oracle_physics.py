"""Pins of the oracle's force laws and integrator against closed forms and an
independent ODE solution (src/contactmodels.cuh:10-109,283-493,
src/integration.cuh:9-295)."""

import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle_lib import Oracle
from sphere_amd.spherebin import PI_REF, Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _box(n, L=1.0, num=10, periodic=0, sid="phys"):
    sb = Spherebin(n, nw=0, sid=sid)
    sb.set_walls(0)
    sb.L = np.full(3, float(L))
    sb.num = np.full(3, num, dtype=np.uint32)
    sb.periodic = periodic
    sb.k_n = sb.k_t = 1e5
    sb.gamma_n = sb.gamma_t = 0.0
    sb.gamma_wn = sb.gamma_wt = 0.0
    sb.mu_s = sb.mu_d = 0.5
    sb.mu_ws = sb.mu_wd = 0.5
    sb.rho = 2600.0
    sb.time_dt = 1e-5
    sb.time_total = 10.0
    return sb


def _mass(r, rho=2600.0):
    return 4.0 / 3.0 * PI_REF * r ** 3 * rho


def test_free_fall_leapfrog_closed_form(tmp_path):
    """x_n = x0 + v0 n dt + g dt^2 n(n+1)/2 and v_n = v0 + g dt (n + 1/2)."""
    sb = _box(1, sid="fall")
    sb.x[0] = (0.5, 0.5, 0.9)
    sb.radius[0] = 0.01
    sb.vel[0] = (0.3, -0.2, 0.5)
    sb.g = np.array([0.0, 0.0, -9.81])
    o = Oracle.from_spherebin(sb, tmp_path)
    n = 700
    assert o.step(n) == 0
    dt = sb.time_dt
    x = o.get("x")[0, :3]
    v = o.get("vel")[0, :3]
    want_x = sb.x[0] + sb.vel[0] * n * dt + sb.g * dt * dt * n * (n + 1) / 2
    want_v = sb.vel[0] + sb.g * dt * (n + 0.5)
    np.testing.assert_allclose(x, want_x, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(v, want_v, rtol=1e-12, atol=1e-15)
    xy = o.get("xysum")[0]
    np.testing.assert_allclose(xy, sb.vel[0, :2] * n * dt, rtol=1e-12)


def test_zero_g_no_contact_is_bit_invariant(tmp_path):
    """tests/io_tests.py:8-34: a free, contact-free, gravity-free system does
    not change by a single bit."""
    rng = np.random.default_rng(2)
    sb = _box(64, sid="still")
    g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    sb.x = 0.2 + g * 0.2 + rng.uniform(-0.01, 0.01, size=(64, 3))
    sb.radius = rng.uniform(0.02, 0.04, size=64)
    o = Oracle.from_spherebin(sb, tmp_path)
    before = {k: o.get(k).copy() for k in ("x", "vel", "angvel", "angpos", "xysum")}
    assert o.step(50) == 0
    for k, a in before.items():
        assert np.array_equal(o.get(k).view(np.uint64), a.view(np.uint64)), k
    assert not o.get("force").any() and not o.get("torque").any()


def _head_on(tmp_path, gamma, k=1e5, r=0.01, v0=0.1, dt=None):
    sb = _box(2, sid="collide")
    sb.k_n = sb.k_t = k
    sb.gamma_n = sb.gamma_t = gamma
    sb.radius[:] = r
    sb.x[0] = (0.5 - r - 1e-5, 0.5, 0.5)
    sb.x[1] = (0.5 + r + 1e-5, 0.5, 0.5)
    sb.vel[0, 0] = v0 / 2
    sb.vel[1, 0] = -v0 / 2
    m = _mass(r)
    m_eff = m / 2
    w0 = math.sqrt(k / m_eff)
    sb.time_dt = dt or (2 * math.pi / w0) / 2000
    o = Oracle.from_spherebin(sb, tmp_path)
    steps_in = 0
    for i in range(20000):
        o.step(1)
        c = o.get("contacts")
        if (c[0] != 2).any():
            steps_in += 1
        elif steps_in:
            break
    v = o.get("vel")
    return (v[1, 0] - v[0, 0]) / v0, steps_in * sb.time_dt, w0, m_eff, sb


def test_elastic_collision_duration_and_restitution(tmp_path):
    """Linear spring: contact time pi/omega0 with omega0 = sqrt(k/m_eff),
    restitution 1."""
    e, t_c, w0, _, sb = _head_on(tmp_path, gamma=0.0)
    assert abs(e - 1.0) < 2e-3
    assert abs(t_c - math.pi / w0) < 3 * sb.time_dt


def test_normal_dashpot_matches_reference_sign(tmp_path):
    """The reference's model-2 dashpot uses vel_n = -dot(v_ab, n)
    (src/contactmodels.cuh:341,366): approaching pairs are pulled in and
    separating pairs pushed out, so e > 1.  Compared with an independent ODE
    solution of m_eff s'' = max(0, -k s + gamma s')."""
    gamma = 20.0
    e, t_c, w0, m_eff, sb = _head_on(tmp_path, gamma=gamma)
    k = sb.k_n

    def rhs(t, y):
        s, sd = y
        f = max(0.0, -k * s + gamma * sd)
        return [sd, f / m_eff]

    def leave(t, y):
        return y[0]
    leave.terminal = True
    leave.direction = 1
    sol = solve_ivp(rhs, [0, 1.0], [-1e-12, -0.1], events=leave, rtol=1e-10, atol=1e-14, max_step=1e-5)
    e_ref = sol.y[1, -1] / 0.1
    assert e_ref > 1.05
    assert abs(e - e_ref) < 5e-3 * e_ref


def test_rest_on_floor_overlap(tmp_path):
    """Equilibrium overlap with the floor is m g / k_n."""
    sb = _box(1, sid="rest")
    r = 0.02
    sb.radius[0] = r
    sb.x[0] = (0.5, 0.5, r)
    sb.g = np.array([0.0, 0.0, -9.81])
    sb.gamma_wn = 5.0
    m = _mass(r)
    sb.time_dt = 0.05 * math.sqrt(m / sb.k_n)
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(20000) == 0
    z = o.get("x")[0, 2]
    want = r - m * 9.81 / sb.k_n
    assert abs(z - want) < 1e-6 * (m * 9.81 / sb.k_n)


def _two_overlapping(tmp_path, vt, k=1e5, r=0.01, overlap=1e-4, mu=0.4):
    sb = _box(2, sid="slip")
    sb.k_n = sb.k_t = k
    sb.gamma_n, sb.gamma_t = 0.0, 3.0
    sb.mu_s = sb.mu_d = mu
    sb.radius[:] = r
    sb.x[0] = (0.5, 0.5, 0.5)
    sb.x[1] = (0.5 + 2 * r - overlap, 0.5, 0.5)
    sb.vel[1] = (0.0, vt, 0.0)
    o = Oracle.from_spherebin(sb, tmp_path)
    for st in ("bonds", "cellid", "sort", "reorder", "topology", "interact"):
        o.stage(st)
    return o, sb


def test_coulomb_slip_limit(tmp_path):
    """Sliding contact: |f_t| = mu_s |f_n| and the spring is reset to the
    Coulomb-consistent length |k_t delta_t + gamma_t v_t| = mu_d |f_n|."""
    o, sb = _two_overlapping(tmp_path, vt=5.0)
    F = o.get("force")[0]
    n = np.array([-1.0, 0.0, 0.0])
    fn = F @ n
    ft = np.linalg.norm(F - fn * n)
    assert fn > 0
    assert abs(ft - sb.mu_s * fn) < 1e-9 * fn
    dtt = o.get("delta_t")[0]
    k = int(np.nonzero(o.get("contacts")[0] != 2)[0][0])
    vt = np.array([0.0, -5.0, 0.0])  # velocity of 0 relative to 1 (no spin)
    assert abs(np.linalg.norm(sb.k_t * dtt[k] + sb.gamma_t * vt) - sb.mu_d * fn) < 1e-9 * fn
    assert o.get("es_dot")[0] > 0


def test_coulomb_stick_below_limit(tmp_path):
    o, sb = _two_overlapping(tmp_path, vt=1e-3)
    F = o.get("force")[0]
    n = np.array([-1.0, 0.0, 0.0])
    fn = F @ n
    ft = np.linalg.norm(F - fn * n)
    assert 0 < ft < sb.mu_s * fn
    assert o.get("es_dot")[0] == 0.0


def test_pair_forces_conserve_momentum_and_angular_momentum(tmp_path):
    """Newton's third law for the pair force and zero net moment about the
    origin, in a dense spinning cluster away from all walls."""
    rng = np.random.default_rng(5)
    n = 300
    sb = _box(n, L=2.0, num=20, sid="cluster")
    sb.gamma_n, sb.gamma_t = 0.5, 2.0
    pts = []
    while len(pts) < n:
        p = rng.uniform(0.8, 1.2, 3)
        pts.append(p)
    sb.x = np.array(pts)
    sb.radius = rng.uniform(0.02, 0.03, n)
    sb.vel = rng.normal(0, 0.05, (n, 3))
    sb.angvel = rng.normal(0, 5.0, (n, 3))
    sb.time_dt = 1e-6
    o = Oracle.from_spherebin(sb, tmp_path)
    o.step(30)
    for st in ("bonds", "cellid", "sort", "reorder", "topology", "interact"):
        o.stage(st)
    F, T, x = o.get("force"), o.get("torque"), o.get("x")[:, :3]
    scale = np.abs(F).sum()
    assert scale > 0
    assert np.abs(F.sum(0)).max() < 1e-10 * scale
    L = (T + np.cross(x, F)).sum(0)
    assert np.abs(L).max() < 1e-10 * (np.abs(T).sum() + np.abs(np.cross(x, F)).sum())


def _wall_box(tmp_path, wmode, vel=0.0, devs=0.0, mass=1.0, gamma_wn=0.0):
    sb = _box(1, sid="wall")
    sb.set_walls(1)
    r = 0.05
    sb.radius[0] = r
    sb.x[0] = (0.5, 0.5, r)
    sb.wmode = np.array([wmode], dtype=np.int32)
    sb.w_x[0] = 2 * r
    sb.w_vel[0] = vel
    sb.w_m[0] = mass
    sb.w_devs[0] = devs
    sb.gamma_wn = gamma_wn
    m = _mass(r)
    sb.time_dt = 0.02 * math.sqrt(min(m, mass) / sb.k_n)
    return sb


def test_velocity_wall_moves_uniformly(tmp_path):
    sb = _wall_box(tmp_path, wmode=2, vel=-1e-3)
    sb.x[0, 2] = 0.02  # below the wall, not touching the floor
    sb.radius[0] = 0.01
    o = Oracle.from_spherebin(sb, tmp_path)
    n = 500
    o.step(n)
    pos = o.get("walls_nx")[0, 3]
    assert abs(pos - (0.1 + n * sb.time_dt * -1e-3)) < 1e-12
    assert o.get("walls_mvfd")[0, 1] == -1e-3


def test_stress_wall_static_equilibrium(tmp_path):
    """A stress-controlled wall on a sphere squeezed against the floor (g = 0):
    placed where both overlaps equal sigma_0 L0 L1 / k_n it stays there, with
    the summed wall force equal to the load.  (A moving wall is not damped
    toward this state: contactLinear_wall adds the wall velocity to the
    particle's, src/contactmodels.cuh:33, so the relative velocity of a
    moving wall has the wrong sign.)"""
    sigma = 2.0
    sb = _wall_box(tmp_path, wmode=1, devs=sigma, mass=1.0)
    r = sb.radius[0]
    F = sigma * sb.L[0] * sb.L[1]
    d = F / sb.k_n
    sb.x[0, 2] = r - d
    sb.w_x[0] = 2 * r - 2 * d
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(2000) == 0
    fz = o.get("walls_mvfd")[0, 2]
    assert abs(fz - F) < 1e-9 * F
    assert abs(o.get("walls_nx")[0, 3] - (2 * r - 2 * d)) < 1e-9 * d
    assert abs(o.get("wall_pp")[0, 0] - F) < 1e-9 * F


# ---- contactLinear_wall tangential friction (src/contactmodels.cuh:47-80)

def _floor_sphere(sid, r=0.01, g=0.0, periodic=1):
    sb = _box(1, sid=sid, periodic=periodic)
    sb.radius[0] = r
    sb.g = np.array([0.0, 0.0, g])
    return sb


def test_sliding_sphere_on_floor_rolls_at_five_sevenths(tmp_path):
    """A sphere launched sliding without spin on the floor: Coulomb friction
    mu_wd N decelerates it at mu_wd g while its torque about the centre
    spins it up, until it rolls.  Angular momentum about the contact point
    is conserved, so it rolls at 5/7 of the launch speed with angular
    velocity v / r, reached at t = 2 v0 / (7 mu_wd g) (textbook sliding-to-
    rolling).  Pins the friction's sign, the dynamic coefficient above 1 mm/s
    and the torque's sign and lever arm (-r n x f_t)."""
    r, g, v0 = 0.01, -9.81, 0.5
    sb = _floor_sphere("roll", r=r, g=g)
    m = _mass(r)
    sb.x[0] = (0.5, 0.5, r - m * abs(g) / sb.k_n)  # at the equilibrium overlap
    sb.vel[0] = (v0, 0.0, 0.0)
    sb.gamma_wn = 30.0           # under critical (2 sqrt(k m) = 66)
    sb.gamma_wt = 1e3            # viscous force far above the Coulomb cap while sliding
    sb.mu_wd, sb.mu_ws = 0.3, 0.6
    sb.time_dt = 0.05 * math.sqrt(m / sb.k_n)
    o = Oracle.from_spherebin(sb, tmp_path)
    t_roll = 2 * v0 / (7 * sb.mu_wd * abs(g))
    half = int(0.5 * t_roll / sb.time_dt)
    assert o.step(half) == 0
    t = half * sb.time_dt
    v = o.get("vel")[0, 0]
    # leapfrog velocity at step n carries the half step: v0 - a dt (n + 1/2)
    want = v0 - sb.mu_wd * abs(g) * sb.time_dt * (half + 0.5)
    assert abs(v - want) < 2e-3 * (v0 - v)
    w = o.get("angvel")[0]
    assert w[1] > 0 and abs(w[0]) < 1e-12 and abs(w[2]) < 1e-12
    # spin-up: torque mu N r over I = 2/5 m r^2 -> omega = 5 mu g t / (2 r)
    assert abs(w[1] - 2.5 * sb.mu_wd * abs(g) * sb.time_dt * (half + 0.5) / r) < 2e-3 * w[1]
    assert o.step(int(1.5 * t_roll / sb.time_dt)) == 0
    v = o.get("vel")[0, :3]
    w = o.get("angvel")[0]
    assert abs(v[0] - 5.0 / 7.0 * v0) < 1e-3 * v0
    assert abs(w[1] * r - v[0]) < 1e-3 * v0
    assert abs(v[1]) < 1e-12 and abs(w[0]) < 1e-9 and abs(w[2]) < 1e-9


@pytest.mark.parametrize("vt,gamma_wt,regime", [
    (0.0011, 1e5, "dynamic"),   # just above 1 mm/s: capped at mu_wd N
    (0.0009, 1e5, "static"),    # just below: capped at mu_ws N
    (0.01, 1.0, "viscous"),     # gamma_wt |v_t| below the cap
])
def test_floor_friction_regimes_one_step(tmp_path, vt, gamma_wt, regime):
    """One step of a sphere pressed into the floor (overlap 1e-4 m, no
    gravity, no normal damping) sliding along (cos 30, sin 30, 0): the
    tangential force opposes the slip with magnitude gamma_wt |v_t| below the
    Coulomb cap and the cap above it, mu_wd N for |v_t| > 1 mm/s and mu_ws N
    below; the torque is (contact point - centre) x f_t."""
    r, ov = 0.01, 1e-4
    sb = _floor_sphere("fric-" + regime, r=r)
    sb.x[0] = (0.5, 0.5, r - ov)
    d = np.array([math.cos(math.pi / 6), math.sin(math.pi / 6), 0.0])
    sb.vel[0] = vt * d
    sb.gamma_wn, sb.gamma_wt = 0.0, gamma_wt
    sb.mu_wd, sb.mu_ws = 0.3, 0.6
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    F, T = o.get("force")[0], o.get("torque")[0]
    N = sb.k_n * ov
    assert abs(F[2] - N) < 1e-12 * N
    ft = F.copy()
    ft[2] = 0.0
    mag = {"dynamic": sb.mu_wd * N, "static": sb.mu_ws * N, "viscous": gamma_wt * vt}[regime]
    np.testing.assert_allclose(ft, -mag * d, rtol=1e-12, atol=1e-18)
    arm = np.array([0.0, 0.0, -r])
    np.testing.assert_allclose(T, np.cross(arm, ft), rtol=1e-12, atol=1e-18)


def test_moving_wall_adds_its_scalar_velocity_to_every_component(tmp_path):
    """A sphere at rest under a top wall moving at w (wall velocity wmode 2):
    the reference adds the scalar w to all three components of the contact
    velocity (operator+(Float3, Float), src/vector_arithmetic.h:61,
    src/contactmodels.cuh:33), so the tangential slip is (w, w, 0) -- the
    sphere feels friction along -(1, 1, 0) sign(w) although nothing slides
    sideways -- and the normal damping sees v_n = -w."""
    r, ov, w = 0.01, 1e-4, 0.02
    sb = _box(1, sid="wvel", periodic=1)
    sb.set_walls(1)
    sb.wmode = np.array([2], dtype=np.int32)
    sb.w_n[0] = (0.0, 0.0, -1.0)
    sb.radius[0] = r
    sb.x[0] = (0.5, 0.5, 0.5)
    sb.w_x[0] = 0.5 + r - ov
    sb.w_vel[0] = w
    sb.w_m[0] = 1.0
    sb.gamma_wn, sb.gamma_wt = 2.0, 3.0
    sb.mu_wd = sb.mu_ws = 0.5
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    F, T = o.get("force")[0], o.get("torque")[0]
    N = sb.k_n * ov + sb.gamma_wn * w      # f_n = (-k_n delta - gamma_wn v_n) n with v_n = -w
    slip = math.sqrt(2.0) * w
    mag = min(sb.gamma_wt * slip, sb.mu_wd * N)
    ft = -mag * np.array([1.0, 1.0, 0.0]) / math.sqrt(2.0)
    np.testing.assert_allclose(F, ft + np.array([0.0, 0.0, -N]), rtol=1e-12, atol=1e-18)
    assert F[0] == F[1] and F[0] < 0
    np.testing.assert_allclose(T, np.cross(np.array([0.0, 0.0, r]), ft), rtol=1e-12, atol=1e-18)
    assert abs(o.get("wall_pp")[0, 0] - N) < 1e-12 * N


# ---- integrate's fixvel branch (src/integration.cuh:65-77)

def test_fixvel_free_flight_keeps_velocities_to_the_bit(tmp_path):
    """A fixvel sphere in free flight under gravity with a sideways
    component: no horizontal acceleration, gravity cancelled on z, no angular
    acceleration -- velocity and angular velocity stay bit-identical, the
    position advances by v dt per step; a free sphere beside it falls."""
    sb = _box(2, L=4.0, num=10, sid="fixvel")
    sb.g = np.array([1.5, -2.0, -9.81])
    sb.radius[:] = 0.01
    sb.x[0] = (1.0, 1.0, 2.0)
    sb.x[1] = (3.0, 3.0, 2.0)
    sb.vel[:] = (0.3, -0.2, 0.1)
    sb.angvel[:] = (1.0, 2.0, 3.0)
    sb.fixvel[0] = 1.0
    o = Oracle.from_spherebin(sb, tmp_path)
    n = 400
    assert o.step(n) == 0
    v, w, x = o.get("vel"), o.get("angvel"), o.get("x")
    assert np.array_equal(v[0, :3], sb.vel[0]) and v[0, 3] == 1.0
    assert np.array_equal(w[0], sb.angvel[0])
    np.testing.assert_allclose(x[0, :3], sb.x[0] + n * sb.time_dt * sb.vel[0], rtol=1e-12)
    np.testing.assert_allclose(o.get("angpos")[0], n * sb.time_dt * sb.angvel[0], rtol=1e-12)
    np.testing.assert_allclose(v[1, :3], sb.vel[1] + sb.g * sb.time_dt * (n + 0.5), rtol=1e-12)


def test_fixvel_sphere_moves_in_z_by_contact_force_only(tmp_path):
    """A fixvel sphere pressed into the floor under gravity: its z
    acceleration is the contact force over its mass with gravity removed
    (free to dilate), its horizontal velocity is held although the floor
    pushes it sideways by friction, and it does not spin up."""
    r, ov = 0.01, 1e-4
    sb = _floor_sphere("fixvel-z", r=r, g=-9.81)
    sb.x[0] = (0.5, 0.5, r - ov)
    sb.vel[0] = (0.05, 0.0, 0.0)
    sb.fixvel[0] = 1.0
    sb.gamma_wn, sb.gamma_wt = 0.0, 1e3
    sb.mu_wd = sb.mu_ws = 0.5
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    F = o.get("force")[0]
    N = sb.k_n * ov
    assert abs(F[2] - N) < 1e-12 * N and F[0] < 0
    m = _mass(r)
    v = o.get("vel")[0]
    # one leapfrog step from rest in z with acceleration N / m: 1.5 a dt
    assert abs(v[2] - 1.5 * N / m * sb.time_dt) < 1e-12 * abs(v[2])
    assert v[0] == 0.05 and v[1] == 0.0
    assert not o.get("angvel")[0].any()


# ---- periodic wrap before the cell id (src/sorting.cuh:7-19, src/integration.cuh:80-95)

def test_sphere_past_the_periodic_boundary_is_hashed_wrapped(tmp_path):
    """A sphere whose integrated position lies just past x = L (integrate
    wraps the position before moving it, so it can end up to v dt outside)
    is hashed at its wrapped position, cell (0, y, z) of its own row, and
    meets the sphere just inside x = 0.  The reference hashes the unwrapped
    position: floor(x / cell) = num_x puts it in cell (0, y + 1, z) of the
    next row, whose neighbourhood misses that partner for this step; this
    build defines the hash on the wrapped position instead."""
    r = 0.01
    sb = _box(2, L=1.0, num=10, periodic=1, sid="wrap")
    sb.radius[:] = r
    sb.x[0] = (1.0 + 1e-3, 0.55, 0.55)
    sb.x[1] = (0.015, 0.55, 0.55)
    o = Oracle.from_spherebin(sb, tmp_path)
    o.stage("bonds")
    o.stage("cellid")
    keys = o.get("cell_keys")
    own_row = 5 * 100 + 5 * 10 + 0
    assert keys[0] == own_row and keys[1] == own_row
    assert keys[0] != own_row + 10  # the reference's next-row cell
    for st in ("sort", "reorder", "topology"):
        o.stage(st)
    c = o.get("contacts")
    assert 1 in c[0] and 0 in c[1]
    assert abs(o.get("x")[0, 0] - 1e-3) < 1e-15


def test_fixvel_hold_z_extension_keeps_the_whole_velocity(tmp_path):
    """This build's extension: fixvel >= 2 holds vel_z as well (a velocity-
    controlled layer); fixvel = 1 beside it is the reference's rule and is
    pushed in z by the same floor contact."""
    from sphere_amd.spherebin import FIXVEL_HOLD_Z

    r, ov = 0.01, 1e-4
    sb = _box(2, sid="hold-z", periodic=1)
    sb.g = np.array([0.0, 0.0, -9.81])
    sb.radius[:] = r
    sb.x[0] = (0.3, 0.5, r - ov)
    sb.x[1] = (0.7, 0.5, r - ov)
    sb.vel[:] = (0.02, -0.01, 0.003)
    sb.fixvel[:] = (FIXVEL_HOLD_Z, 1.0)
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(50) == 0
    v = o.get("vel")
    assert np.array_equal(v[0, :3], sb.vel[0])
    np.testing.assert_allclose(o.get("x")[0, :3], sb.x[0] + 50 * sb.time_dt * sb.vel[0], rtol=1e-12)
    assert np.array_equal(v[1, :2], sb.vel[1, :2]) and v[1, 2] > sb.vel[1, 2]
    assert o.get("force")[0, 2] > 0  # the floor pushes, the held layer ignores it
