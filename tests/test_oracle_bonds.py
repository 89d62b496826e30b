"""Pins of the oracle's parallel-bond model (src/cohesion.cuh:8-233) with
the reference's two-particle scenarios (tests/bond_tests.py:33-192, including
the contraction, shear, twist and bend cases its `continue` skips) and closed
forms."""

import math

import numpy as np
import pytest

from oracle_lib import Oracle
from sphere_amd import configs
from sphere_amd.spherebin import Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _state(o, sb):
    s = sb.copy()
    s.x = o.get("x")[:, :3]
    s.vel = o.get("vel")[:, :3]
    s.angvel = o.get("angvel")
    bd, bo = o.get("bonds_delta"), o.get("bonds_omega")
    s.bonds = o.get("bonds")
    s.bonds_delta_n, s.bonds_delta_t = bd[:, 3], bd[:, :3]
    s.bonds_omega_n, s.bonds_omega_t = bo[:, 3], bo[:, :3]
    return s


def _energy(s):
    return s.energy("kin") + s.energy("rot") + s.energy("bondpot")


def _pair(distance, v=(0, 0, 0), w=(0, 0, 0)):
    sb = configs.two_spheres(distance=distance, radius=0.5, bonded=True, k=1.16e9, gamma=0.0)
    sb.vel[1] = v
    sb.angvel[1] = w
    return sb


@pytest.mark.parametrize("distance", [0.2, 0.0])
def test_stability_stays_exactly_at_rest(tmp_path, distance):
    sb = _pair(distance)
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(500) == 0
    assert not o.get("vel")[:, :3].any() and not o.get("angvel").any()
    assert _energy(_state(o, sb)) == 0.0


CASES = {
    "expansion": dict(v=(1e-4, 0, 0)),
    "contraction": dict(v=(-1e-4, 0, 0)),
    "shear": dict(v=(0, 0, 1e-4)),
    "twist": dict(w=(1e-4, 0, 0)),
    "bend": dict(w0=(0, -1e-4, 0), w=(0, 1e-4, 0)),
}


def _loss(tmp_path, case, distance, shrink, horizon=0.02):
    kw = dict(CASES[case])
    w0 = kw.pop("w0", None)
    sb = _pair(distance, **kw)
    if w0 is not None:
        sb.angvel[0] = w0
    sb.time_dt /= shrink
    e0 = _energy(sb)
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(int(horizon / sb.time_dt)) == 0
    return (e0 - _energy(_state(o, sb))) / e0


@pytest.mark.parametrize("distance", [0.2, 0.0])
@pytest.mark.parametrize("case", list(CASES))
def test_energy_converges_first_order(tmp_path, case, distance):
    """Kinetic + rotational + bond potential energy (python/sphere.py:1515-1562).
    The bond displacement is advanced with the end-of-step velocity estimate
    vel = vel0 + acc dt/2 (src/integration.cuh:120-127, read by
    src/cohesion.cuh:98-105), which makes the scheme first order and
    dissipative: the energy lost over a fixed time halves when dt halves and
    vanishes as dt -> 0.  (The reference's own check uses a 1e-5 J absolute
    tolerance, tests/pytestutils.py:20, larger than the energies involved.)"""
    l4 = _loss(tmp_path, case, distance, 4)
    l8 = _loss(tmp_path, case, distance, 8)
    l16 = _loss(tmp_path, case, distance, 16)
    assert 0 < l16 < l8 < l4
    assert 1.7 < l4 / l8 < 2.3 and 1.8 < l8 / l16 < 2.2
    assert l16 < 0.05


def _period(signal, dt):
    s = np.asarray(signal)
    zc = np.nonzero((s[:-1] < 0) & (s[1:] >= 0))[0]
    assert len(zc) >= 3
    # linear interpolation of the upward zero crossings
    t = zc + (-s[zc]) / (s[zc + 1] - s[zc])
    return float(np.mean(np.diff(t))) * dt


def test_axial_frequency(tmp_path):
    """Normal bond stiffness k_n A with A = pi (lambda r)^2: the pair
    oscillates at omega = sqrt(k_n A / m_eff), m_eff = m/2."""
    sb = _pair(0.2, v=(1e-4, 0, 0))
    sb.time_dt /= 10
    o = Oracle.from_spherebin(sb, tmp_path)
    r = 0.5
    A = math.pi * (sb.lambda_bar * r) ** 2
    m = 4 / 3 * math.pi * r ** 3 * sb.rho
    w = math.sqrt(sb.k_n * A / (m / 2))
    dt = sb.time_dt
    steps = int(4 * 2 * math.pi / w / dt)
    sig = []
    for _ in range(steps):
        o.step(1)
        sig.append(o.get("bonds_delta")[0, 3])
    T = _period(sig, dt)
    assert abs(T - 2 * math.pi / w) < 2e-3 * T


def test_twist_frequency(tmp_path):
    """Twisting stiffness k_t J with J = pi (lambda r)^4 / 2 about the bond
    axis: omega = sqrt(k_t J / (I/2)), I = 2/5 m r^2."""
    sb = _pair(0.2, w=(1e-4, 0, 0))
    sb.time_dt /= 10
    o = Oracle.from_spherebin(sb, tmp_path)
    r = 0.5
    J = math.pi * (sb.lambda_bar * r) ** 4 / 2
    m = 4 / 3 * math.pi * r ** 3 * sb.rho
    I = 0.4 * m * r * r
    w = math.sqrt(sb.k_t * J / (I / 2))
    dt = sb.time_dt
    steps = int(4 * 2 * math.pi / w / dt)
    sig = []
    for _ in range(steps):
        o.step(1)
        sig.append(o.get("bonds_omega")[0, 3])
    T = _period(sig, dt)
    assert abs(T - 2 * math.pi / w) < 2e-3 * T


def test_breakage_at_tensile_strength(tmp_path):
    """Pulled apart, the bond breaks on the first step whose normal stress
    k_n |delta_n| reaches sigma_b; afterwards the pair flies apart freely and
    the bond's first index holds max(nb0, np)."""
    sb = _pair(0.2, v=(1e-2, 0, 0))
    sb.sigma_b = 5.8e3  # k_n x 5e-6 m: reached before the 8.6e-6 m peak extension
    sb.tau_b = math.inf
    o = Oracle.from_spherebin(sb, tmp_path)
    hist = []
    broke = None
    for k in range(5000):
        o.step(1)
        b = o.get("bonds")[0]
        if b[0] >= sb.np:
            broke = k
            break
        hist.append(o.get("bonds_delta")[0, 3])
    assert broke is not None and broke > 2
    assert o.get("bonds")[0, 0] == max(sb.nb0, sb.np)
    last, prev = abs(hist[-1]), abs(hist[-2])
    assert sb.k_n * last < sb.sigma_b <= sb.k_n * (last + 2 * (last - prev))
    v0 = o.get("vel")[:, 0].copy()
    o.step(50)
    assert np.array_equal(o.get("vel")[:, 0], v0)
