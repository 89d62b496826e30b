"""The hand-written adjoint (sph_adjoint_forward / sph_adjoint_backward and
the torch.autograd.Function over it) against central finite differences of
the same objective, relative step 1e-6."""

from __future__ import annotations

import numpy as np
import pytest

from sphere_amd import configs

THETA0 = (1e6, 1e6, 20.0, 0.5)  # k_n, k_t, gamma_n, mu
# Central differences of a dense granular run are only as good as the run is
# smooth in theta: over many steps the roundoff of J(theta +- h) is amplified
# by the chaotic dynamics and some contact or slip branch flips, so the check
# uses a short horizon (the adjoint itself has no horizon limit).
FD_STEPS = 30


def _case():
    return configs.fcc_shear(cells=(4, 4, 8), gamma_n=THETA0[2], sid="adj-fcc-shear")


def _relerr(g, fd):
    return np.abs(g - fd) / np.abs(fd)


def _say(msg):
    print(f"[adjoint] {msg}", flush=True)


def test_shear_weights_select_driven_layer():
    from sphere_amd.adjoint import shear_stress_weights

    sb = _case()
    w = shear_stress_weights(sb)
    driven = (sb.fixvel > 0) & (sb.vel[:, 0] > 0)
    assert driven.sum() == 32
    assert np.all(w[driven, 0] == -1.0 / (sb.L[0] * sb.L[1]))
    assert np.all(w[~driven] == 0.0) and np.all(w[:, 1:] == 0.0)


def _fd(sb, w, nsteps, precision="fp64", rel=1e-6, theta0=THETA0):
    from sphere_amd.adjoint import objective_value

    th0 = np.array(theta0)
    fd = np.zeros(4)
    for k in range(4):
        h = rel * abs(th0[k])
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (objective_value(sb, tp, nsteps, w, precision) - objective_value(sb, tm, nsteps, w, precision)) / (2 * h)
    return fd


@pytest.mark.gpu
def test_adjoint_matches_central_differences():
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    _say(f"fp64 adjoint vs central differences, 512 spheres, {FD_STEPS} steps")
    sb = _case()
    w = shear_stress_weights(sb)
    dem = DEM(sb, verbosity=0)
    dem.set_params(k_n=THETA0[0], k_t=THETA0[1], gamma_n=THETA0[2], mu_s=THETA0[3], mu_d=THETA0[3])
    J = dem.adjoint_forward(FD_STEPS, w, seg_len=8)  # 4 segments, the last partial
    g = dem.adjoint_backward()
    dem.close()
    fd = _fd(sb, w, FD_STEPS)
    err = _relerr(g, fd)
    _say(f"J={J:.6e} adjoint={g} fd={fd} relative error={err}")
    assert np.all(err <= 1e-4), f"relative error {err}"


@pytest.mark.gpu
def test_autograd_function_gives_the_adjoint():
    import torch

    from sphere_amd.adjoint import mean_shear_stress, shear_stress_weights
    from sphere_amd.dem import DEM

    _say("autograd wrapper")
    sb = _case()
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    J = mean_shear_stress(theta, sb, 120, seg_len=50)
    J.backward()
    dem = DEM(sb, verbosity=0)
    dem.set_params(k_n=THETA0[0], k_t=THETA0[1], gamma_n=THETA0[2], mu_s=THETA0[3], mu_d=THETA0[3])
    J2 = dem.adjoint_forward(120, shear_stress_weights(sb), seg_len=50)
    g2 = dem.adjoint_backward()
    dem.close()
    # the theta block sums are FP64 atomics, so the last bits depend on
    # their order
    assert abs(float(J) - J2) <= 1e-12 * abs(J2)
    np.testing.assert_allclose(theta.grad.numpy(), g2, rtol=1e-10, atol=0.0)


@pytest.mark.gpu
def test_adjoint_leaves_the_forward_end_state():
    """Forward + backward leave the handle exactly where a plain run of the
    same steps ends (checkpoints restore every array bit for bit)."""
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    _say("end state after backward")
    sb = _case()
    a = DEM(sb, verbosity=0)
    a.adjoint_forward(90, shear_stress_weights(sb), seg_len=40)
    a.adjoint_backward()
    b = DEM(sb, verbosity=0)
    b.step(90)
    for f in ("x", "vel", "angvel", "force", "torque", "contacts", "delta_t", "walls_nx", "walls_mvfd"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.step(10)
    b.step(10)
    assert np.array_equal(a.get("x"), b.get("x"))
    a.close()
    b.close()


@pytest.mark.gpu
def test_mixed_adjoint_close_to_fp64_differences():
    """Mixed mode (FP32 pair forces): its adjoint against the FP64 central
    differences; FP32 rounding bounds the agreement, not the adjoint."""
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    _say("mixed adjoint vs fp64 central differences")
    sb = _case()
    w = shear_stress_weights(sb)
    dem = DEM(sb, verbosity=0)
    dem.set_precision("mixed")
    dem.set_params(k_n=THETA0[0], k_t=THETA0[1], gamma_n=THETA0[2], mu_s=THETA0[3], mu_d=THETA0[3])
    dem.adjoint_forward(FD_STEPS, w, seg_len=8)
    g = dem.adjoint_backward()
    dem.close()
    fd = _fd(sb, w, FD_STEPS)
    err = _relerr(g, fd)
    _say(f"mixed adjoint={g} fp64 fd={fd} rel err={err}")
    assert np.all(err <= 1e-2), f"relative error {err}"


@pytest.mark.gpu
def test_bonded_input_is_unsupported():
    from sphere_amd.dem import DEM, Unsupported

    sb = configs.two_spheres(bonded=True)
    dem = DEM(sb, verbosity=0)
    with pytest.raises(Unsupported):
        dem.adjoint_forward(5, np.zeros((2, 3)))
    dem.close()


@pytest.mark.gpu
def test_distinct_static_and_dynamic_friction_is_refused():
    """theta holds one friction coefficient, mu = mu_s = mu_d: a handle with
    mu_s != mu_d is refused rather than given d/dmu_s + d/dmu_d."""
    from sphere_amd.dem import DEM, SphereInvalid

    sb = configs.two_spheres(sid="mu-split")
    sb.mu_s, sb.mu_d = 0.5, 0.4
    dem = DEM(sb, verbosity=0)
    with pytest.raises(SphereInvalid, match="mu_s"):
        dem.adjoint_forward(5, np.zeros((2, 3)))
    dem.close()


@pytest.mark.gpu
def test_adjoint_at_full_size_matches_central_differences():
    """The benchmark's adjoint workload at its full size (the 65,536-sphere
    dense shear cell, top layer driven at 0.1 m/s under 10 kPa), over a
    10-step horizon in 3 segments, against central differences."""
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    _say("fp64 adjoint vs central differences, 65,536 spheres, 10 steps")
    sb = configs.fcc_shear(gamma_n=THETA0[2], sid="adj-fcc-shear-full")
    assert sb.np == 65536
    w = shear_stress_weights(sb)
    dem = DEM(sb, verbosity=0)
    dem.set_params(k_n=THETA0[0], k_t=THETA0[1], gamma_n=THETA0[2], mu_s=THETA0[3], mu_d=THETA0[3])
    J = dem.adjoint_forward(10, w, seg_len=4)
    g = dem.adjoint_backward()
    dem.close()
    fd = _fd(sb, w, 10)
    err = _relerr(g, fd)
    _say(f"J={J:.6e} adjoint={g} fd={fd} relative error={err}")
    assert np.all(err <= 1e-4), f"relative error {err}"


# The differentiable shear test as specified: the 65,536-sphere polydisperse
# cell (config-1 lognormal radii) compacted under its 10 kPa wall, then
# sheared at 0.1 m/s; checkpoints every 100 steps.
C3_THETA = (1e6, 1e6, 2.0, 0.5)
C3_SEG = 100
C3_FD_STEPS = 2 * C3_SEG  # two full checkpoint segments


@pytest.fixture(scope="module")
def c3_cell():
    _say("consolidating the 65,536-sphere polydisperse shear cell")
    return configs.consolidated_shear_cell(gamma_n=C3_THETA[2], mu=C3_THETA[3], log=_say)


@pytest.mark.gpu
def test_consolidated_shear_cell_is_dense(c3_cell):
    """The recipe's packing: every sphere inside the box, the driven and the
    held layers in place, about four or more contacts per sphere."""
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    sb = c3_cell
    assert sb.np == 65536 and sb.contactmodel == 2
    assert float(np.ptp(sb.radius)) > 0.5 * float(np.median(sb.radius))  # polydisperse
    w = shear_stress_weights(sb)
    assert int((w[:, 0] != 0.0).sum()) > 100
    dem = DEM(sb, verbosity=0)
    ij, _ = dem.findOverlaps()
    dem.close()
    cn = 2.0 * len(ij) / sb.np
    _say(f"C3 packing: C/N = {cn:.2f}, driven {int((w[:, 0] != 0).sum())}")
    assert cn >= 4.0


@pytest.mark.gpu
def test_adjoint_on_the_consolidated_shear_cell(c3_cell):
    """FP64 adjoint of the mean shear stress on the compacted polydisperse
    cell with checkpoints every 100 steps, over two full segments, against
    central differences with relative step 1e-6 (bar 1e-4 per component)."""
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    sb = c3_cell
    w = shear_stress_weights(sb)
    _say(f"fp64 adjoint vs central differences, consolidated cell, {C3_FD_STEPS} steps, segments of {C3_SEG}")
    dem = DEM(sb, verbosity=0)
    J = dem.adjoint_forward(C3_FD_STEPS, w, seg_len=C3_SEG)
    g = dem.adjoint_backward()
    dem.close()
    fd = _fd(sb, w, C3_FD_STEPS, theta0=C3_THETA)
    err = _relerr(g, fd)
    _say(f"J={J:.6e} adjoint={g} fd={fd} relative error={err}")
    assert np.all(err <= 1e-4), f"relative error {err}"
