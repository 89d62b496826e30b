"""Pins of the oracle's contact model 1 (contactLinearViscous,
src/contactmodels.cuh:114-253, found and processed in the 27-cell scan of
src/contactsearch.cuh:76-173,548-576) and of the capillary bond
(capillaryCohesion_exp, src/cohesion.cuh:314-350) against closed forms."""

import math
import struct

import numpy as np
import pytest

from oracle_lib import Oracle, lib
from sphere_amd.spherebin import PI_REF, Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _ulp_diff(a: float, b: float) -> int:
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def test_exp_within_two_ulp():
    """The oracle's exponential (Cody-Waite reduction, degree-13 Taylor
    series, 2^k scaling) against the C library's exp."""
    f = lib().orc_exp
    assert f(0.0) == 1.0
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-700.0, 700.0, 2000), rng.uniform(-1.0, 1.0, 2000),
                         [0.5 * math.log(2.0), -0.5 * math.log(2.0), 1.0, -1.0, 1e-300, -1e-12, 30.0]])
    worst = max(_ulp_diff(f(float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 2, worst


def _pair(distance, v_a=(0.0, 0.0, 0.0), w_a=(0.0, 0.0, 0.0), r=0.01, k=1e5, gamma_n=0.0, gamma_t=0.0,
          mu=0.5, model=1, kappa=0.0, db=0.0, V_b=0.0, sid="m1pair"):
    sb = Spherebin(2, nw=0, sid=sid)
    sb.set_walls(0)
    sb.L = np.full(3, 1.0)
    sb.num = np.full(3, 10, dtype=np.uint32)
    sb.periodic = 0
    sb.radius[:] = r
    sb.x[0] = (0.5, 0.5, 0.5)
    sb.x[1] = (0.5 + 2.0 * r + distance, 0.5, 0.5)
    sb.vel[0] = v_a
    sb.angvel[0] = w_a
    sb.k_n = sb.k_t = k
    sb.gamma_n, sb.gamma_t = gamma_n, gamma_t
    sb.gamma_wn = sb.gamma_wt = 0.0
    sb.mu_s = sb.mu_d = mu
    sb.mu_ws = sb.mu_wd = 0.0
    sb.rho = 2600.0
    sb.contactmodel = model
    sb.kappa, sb.db, sb.V_b = kappa, db, V_b
    sb.time_dt = 1e-6
    sb.time_total = 1.0
    return sb


def test_model1_normal_force_and_no_torque(tmp_path):
    """f_n = (-k_n delta - gamma_n v_n) n with v_n = (v_a - v_b) . n (model
    1's sign), n from b to a; equal and opposite on b; no torque; the
    pressure |f_n| / (4 pi r^2); the contact slots stay empty."""
    d, vn, k, gn, r = -2e-4, -0.05, 1e5, 3.0, 0.01
    sb = _pair(d, v_a=(0.05, 0.0, 0.0), gamma_n=gn, k=k, r=r)  # a moves towards +x, i.e. towards b
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    F = o.get("force")
    n = np.array([-1.0, 0.0, 0.0])  # x_ab = x_a - x_b points along -x
    v_n = float(np.dot(np.array([0.05, 0.0, 0.0]), n))
    assert v_n == pytest.approx(vn)
    want = (-k * d - gn * v_n) * n
    np.testing.assert_allclose(F[0], want, rtol=1e-12)
    np.testing.assert_allclose(F[1], -want, rtol=1e-12)
    assert np.all(o.get("torque") == 0.0)
    np.testing.assert_allclose(o.get("p")[0], np.linalg.norm(want) / (4.0 * PI_REF * r * r), rtol=1e-12)
    assert np.all(o.get("contacts") == 2)
    assert np.all(o.get("delta_t") == 0.0)


@pytest.mark.parametrize("slip", [False, True])
def test_model1_viscous_shear_and_coulomb_cap(tmp_path, slip):
    """Tangential force -gamma_t v_t below the cap, -mu |f_n - f_c| v_t/|v_t|
    above it (no spring: the force follows the sliding velocity)."""
    d, k, gt, mu = -1e-4, 1e5, (2.0 if not slip else 400.0), 0.3
    vt = 0.02
    sb = _pair(d, v_a=(0.0, vt, 0.0), gamma_t=gt, k=k, mu=mu, sid=f"m1shear{int(slip)}")
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    F = o.get("force")[0]
    fn = -k * d  # along n = -x
    cap = mu * fn
    ft = min(gt * vt, cap)
    assert (gt * vt > cap) == slip
    np.testing.assert_allclose(F, [-fn, -ft, 0.0], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(o.get("force")[1], [fn, ft, 0.0], rtol=1e-12, atol=1e-18)
    assert np.all(o.get("torque") == 0.0)


def test_capillary_bond_closed_form(tmp_path):
    """For a gap 0 <= d < db: F_a = -kappa R_geo exp(-d / lambda) n with
    lambda = 0.9 h sqrt(V_b / R_har), h = -sqrt(max(r_a/r_b, r_b/r_a)) (the
    reference's float literal 0.9f), R_har = 2 r_a r_b / (r_a + r_b)."""
    gap, r, kappa, db, V_b = 1e-4, 0.01, 0.05, 1e-3, 1e-9
    sb = _pair(gap, r=r, kappa=kappa, db=db, V_b=V_b, sid="capillary")
    o = Oracle.from_spherebin(sb, tmp_path)
    assert o.step(1) == 0
    x = sb.x[0] - sb.x[1]
    length = float(np.linalg.norm(x))
    delta = length - 2.0 * r
    n = x / length
    h = -math.sqrt(1.0)
    R_har = 2.0 * r * r / (2.0 * r)
    lam = float(np.float32(0.9)) * h * math.sqrt(V_b / R_har)
    want = -kappa * math.sqrt(r * r) * math.exp(-delta / lam) * n
    np.testing.assert_allclose(o.get("force")[0], want, rtol=1e-13)
    np.testing.assert_allclose(o.get("force")[1], -want, rtol=1e-13)
    # beyond db no force
    sb2 = _pair(2.0 * db, r=r, kappa=kappa, db=db, V_b=V_b, sid="capillary-far")
    o2 = Oracle.from_spherebin(sb2, tmp_path)
    assert o2.step(1) == 0
    assert np.all(o2.get("force") == 0.0)
