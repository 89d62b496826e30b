"""Pins of the oracle's contact model 3 (contactHertz, src/contactmodels.cuh:
499-666) for a positive overlap xi: Hertzian normal force k_n xi^(3/2)
(static overlap under a load, elastic collision time and peak overlap),
Mindlin-type tangential force with its Coulomb cap and spring reset, and the
rolling resistance in both of its regimes."""

import math

import numpy as np
import pytest
from scipy.integrate import quad, solve_ivp

from oracle_lib import Oracle
from sphere_amd.spherebin import PI_REF, Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _pair(r=0.01, gap=0.0, k_n=1e6, k_t=1e6, gamma_n=0.0, gamma_t=0.0, mu=0.5, gamma_r=0.0, mu_r=0.0,
          dt=1e-6, sid="hertz"):
    """Two spheres on the x axis of an open 1 m box, no gravity; `gap` < 0
    is an overlap."""
    sb = Spherebin(2, nw=0, sid=sid)
    sb.set_walls(0)
    sb.L = np.full(3, 1.0)
    sb.num = np.full(3, 10, dtype=np.uint32)
    sb.periodic = 0
    sb.contactmodel = 3
    sb.k_n, sb.k_t = k_n, k_t
    sb.gamma_n, sb.gamma_t = gamma_n, gamma_t
    sb.gamma_r, sb.mu_r = gamma_r, mu_r
    sb.gamma_wn = sb.gamma_wt = 0.0
    sb.mu_s = sb.mu_d = mu
    sb.mu_ws = sb.mu_wd = 0.0
    sb.rho = 2600.0
    sb.radius[:] = r
    sb.x[0] = (0.5 - r - gap / 2.0, 0.5, 0.5)
    sb.x[1] = (0.5 + r + gap / 2.0, 0.5, 0.5)
    sb.g = np.zeros(3)
    sb.time_dt = dt
    sb.time_total = 10.0
    return sb


def _mass(r, rho=2600.0):
    return 4.0 / 3.0 * PI_REF * r ** 3 * rho


def test_static_overlap_under_a_load(tmp_path):
    """A sphere resting on a held one under gravity settles where
    k_n xi^(3/2) = m g, with the contact force equal to its weight."""
    r, k_n = 0.01, 1e6
    sb = _pair(r=r, k_n=k_n, gamma_n=100.0, dt=2e-6, sid="hertz-rest")
    m = _mass(r)
    g = 9.81
    xi = (m * g / k_n) ** (2.0 / 3.0)
    # vertical stack: sphere 1 held below, sphere 0 above, 10 % off equilibrium
    sb.x[1] = (0.5, 0.5, 0.4)
    sb.x[0] = (0.5, 0.5, 0.4 + 2 * r - 0.9 * xi)
    sb.fixvel[1] = 2.0  # velocity held in all three components
    sb.g = np.array([0.0, 0.0, -g])
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(40000) == 0
    x = o.get("x")
    got = 2 * r - (x[0, 2] - x[1, 2])
    assert abs(got - xi) / xi < 1e-6, (got, xi)
    f = o.get("force")
    assert abs(f[0, 2] - m * g) / (m * g) < 1e-6
    assert np.allclose(f[0, :2], 0.0, atol=1e-15)


def test_elastic_collision_time_and_peak_overlap(tmp_path):
    """Undamped head-on collision of two equal spheres closing at v: peak
    overlap xi_max = (5 m_eff v^2 / (4 k_n))^(2/5) and contact time
    2 (xi_max / v) * integral_0^1 (1 - u^(5/2))^(-1/2) du; they separate at
    their approach speed."""
    r, k_n, v0 = 0.01, 1e6, 0.05
    dt = 2e-7
    sb = _pair(r=r, k_n=k_n, gap=1e-7, dt=dt, sid="hertz-bump")
    sb.vel[0, 0], sb.vel[1, 0] = v0, -v0
    m = _mass(r)
    meff = m / 2.0
    v = 2 * v0
    xi_max = (5.0 * meff * v * v / (4.0 * k_n)) ** 0.4
    t_c = 2.0 * xi_max / v * quad(lambda u: 1.0 / math.sqrt(1.0 - u ** 2.5), 0.0, 1.0)[0]
    o = Oracle.from_spherebin(sb, tmp_path)
    peak, steps, in_contact, t_in, t_out = 0.0, 0, False, None, None
    while steps < int(3 * t_c / dt) + 2000:
        assert o.step(1) == 0
        steps += 1
        x = o.get("x")
        xi = 2 * r - (x[1, 0] - x[0, 0])
        if xi > 0 and not in_contact:
            in_contact, t_in = True, steps
        if in_contact:
            peak = max(peak, xi)
            if xi <= 0:
                t_out = steps
                break
    assert t_out is not None
    assert abs(peak - xi_max) / xi_max < 2e-3, (peak, xi_max)
    assert abs((t_out - t_in) * dt - t_c) / t_c < 5e-3, ((t_out - t_in) * dt, t_c)
    vel = o.get("vel")
    assert abs(vel[1, 0] - vel[0, 0] - v) / v < 5e-3  # restitution 1 without damping


def test_damped_collision_against_the_ode(tmp_path):
    """Head-on collision with the normal dashpot: the separation speed
    matches the ODE m_eff xi'' = -k_n xi^(3/2) - gamma_n xi^(1/4) xi'
    (solved independently) -- the dashpot's exponent and its sign."""
    r, k_n, gamma_n, v0 = 0.01, 1e6, 2.0, 0.05
    dt = 2e-7
    sb = _pair(r=r, k_n=k_n, gamma_n=gamma_n, gap=1e-7, dt=dt, sid="hertz-damped")
    sb.vel[0, 0], sb.vel[1, 0] = v0, -v0
    m = _mass(r)
    meff = m / 2.0
    v = 2 * v0

    def rhs(t, y):
        xi, u = y
        return [u, (-k_n * max(xi, 0.0) ** 1.5 - gamma_n * max(xi, 0.0) ** 0.25 * u) / meff]

    leave = lambda t, y: y[0]  # noqa: E731
    leave.terminal, leave.direction = True, -1
    sol = solve_ivp(rhs, (0.0, 1.0), [1e-15, v], events=leave, rtol=1e-11, atol=1e-16, max_step=1e-6)
    v_out = -sol.y_events[0][0][1]
    assert v_out < 0.98 * v  # the dashpot takes energy out
    o = Oracle.from_spherebin(sb, tmp_path)
    for _ in range(int(2 * sol.t_events[0][0] / dt) + 1000):
        assert o.step(1) == 0
        x = o.get("x")
        if 2 * r - (x[1, 0] - x[0, 0]) < -1e-6:
            break
    vel = o.get("vel")
    assert abs((vel[1, 0] - vel[0, 0]) - v_out) / v_out < 5e-3, (vel[1, 0] - vel[0, 0], v_out)


def _one_step(tmp_path, sb):
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    return o.get("force"), o.get("torque"), o.get("es_dot"), o.get("delta_t")


@pytest.mark.parametrize("gamma_r,regime", [(1e-4, "viscous"), (10.0, "capped")])
def test_rolling_resistance(tmp_path, gamma_r, regime):
    """Sphere 0 spins about the contact normal (no sliding at the contact):
    its torque is the rolling resistance alone, -min(gamma_r r |w|, mu_r r
    |f_n|) along w, and the partner feels the opposite one."""
    r, k_n, xi, w = 0.01, 1e6, 2e-5, 3.0
    sb = _pair(r=r, k_n=k_n, gap=-xi, gamma_r=gamma_r, mu_r=0.1, sid=f"hertz-roll-{regime}")
    sb.angvel[0] = (w, 0.0, 0.0)
    f, t, _, _ = _one_step(tmp_path, sb)
    f_n = k_n * xi ** 1.5
    want = min(gamma_r * r * w, 0.1 * r * f_n)
    if regime == "viscous":
        assert gamma_r * r * w < 0.1 * r * f_n
    else:
        assert gamma_r * r * w > 0.1 * r * f_n
    np.testing.assert_allclose(t[0], (-want, 0.0, 0.0), rtol=1e-6, atol=1e-18)
    np.testing.assert_allclose(t[1], (want, 0.0, 0.0), rtol=1e-6, atol=1e-18)
    np.testing.assert_allclose(f[0], (-f_n, 0.0, 0.0), rtol=1e-6)  # x_ab = x_0 - x_1 points along -x


def test_tangential_stick_and_slip(tmp_path):
    """Sliding along y at v_t: below the cap the force is the viscous
    -gamma_t xi^(1/4) v_t and the spring grows by v_t dt; above it the force
    is mu_d |f_n| against the motion, the spring is reset to
    (f_t + gamma_t xi^(1/4) v_t) / (-k_t xi^(1/2)) and es_dot = |spring| k_t / dt."""
    r, k_n, xi, dt = 0.01, 1e6, 2e-5, 1e-6
    f_n = k_n * xi ** 1.5
    for gamma_t, vt, mu in ((1.0, 1e-3, 10.0), (50.0, 0.5, 0.3)):
        sb = _pair(r=r, k_n=k_n, k_t=1e6, gap=-xi, gamma_t=gamma_t, mu=mu, dt=dt, sid="hertz-slide")
        sb.vel[0, 1] = vt
        f, t, es_dot, dlt = _one_step(tmp_path, sb)
        visc = gamma_t * xi ** 0.25 * vt
        if visc < mu * f_n:  # stick
            np.testing.assert_allclose(f[0, 1], -visc, rtol=1e-6)
            np.testing.assert_allclose(dlt[0, :, 1][dlt[0, :, 1] != 0], [vt * dt], rtol=1e-6)
            assert es_dot[0] == 0.0
        else:  # slip
            np.testing.assert_allclose(f[0, 1], -mu * f_n, rtol=1e-6)
            spring = (-mu * f_n + visc) / (-1e6 * xi ** 0.5)
            np.testing.assert_allclose(dlt[0, :, 1][dlt[0, :, 1] != 0], [spring], rtol=1e-6)
            np.testing.assert_allclose(es_dot[0], abs(spring) * 1e6 / dt, rtol=1e-6)
        # the sliding sphere's torque is (contact point - centre) x f_t, the
        # contact point (r - xi/2) along +x from sphere 0's centre
        lever = r - xi / 2.0
        np.testing.assert_allclose(t[0, 2], lever * f[0, 1], rtol=1e-6)
