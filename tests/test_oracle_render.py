"""Pins of the oracle's raytracer (DEM::render, src/raytracer.cuh:14-553)
against an FP64 restatement from the textbook: the ray-sphere quadratic in
canonical form, nearest |t| of the smaller root, Lambertian shading of that
point, the colour bar and the reference's cutoff and fixed-sphere rules."""

import math

import numpy as np
import pytest

from oracle_lib import Oracle
from sphere_amd.spherebin import Spherebin

pytestmark = pytest.mark.usefixtures("oracle_built")


def _scene(centres, radii, L=(1.0, 1.0, 1.0), sid="render", **fields):
    n = len(radii)
    sb = Spherebin(n, nw=0, sid=sid)
    sb.set_walls(0)
    sb.L = np.asarray(L, dtype=float)
    sb.num = np.maximum(1, np.floor(sb.L / (2.0 * max(radii)))).astype(np.uint32)
    sb.x[:] = centres
    sb.radius[:] = radii
    sb.k_n = sb.k_t = 1e5
    sb.rho = 2600.0
    sb.time_dt = 1e-6
    sb.time_total = 1.0
    for k, v in fields.items():
        getattr(sb, k)[...] = v
    return sb


def _camera(L, maxpos, W, H):
    eye = np.array([2.3 * L[0], -5.0 * L[1], 1.3 * L[2]])
    look = np.asarray(maxpos) / 2.0
    view = eye - look
    w = -view / np.linalg.norm(view)
    u = np.cross([0.0, 0.0, 1.0], w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    D = 0.8 * np.linalg.norm(view)
    imgw = 1.35 * L[0]
    hw = float(H // W)
    i = np.arange(W)[None, :] + 0.5
    j = np.arange(H)[:, None] + 0.5
    pu = -0.5 * imgw + imgw * i / W
    pv = (-0.5 * imgw * hw + imgw * hw * j / H) * np.ones_like(pu)
    pu = pu * np.ones_like(pv)
    d = -D * w[None, None, :] + pu[..., None] * u + pv[..., None] * v
    light = -eye * np.array([1.0, 0.2, 0.6])
    return eye, d, light / np.linalg.norm(light)


def _textbook(sb, W, H, colour=None, keep=None):
    """FP64 render: returns (rgb, hit mask, marginal mask)."""
    L = sb.L
    maxpos = np.maximum(0.0, sb.x.max(axis=0))
    e, d, light = _camera(L, maxpos, W, H)
    best = np.full((H, W), np.inf)
    who = np.full((H, W), -1)
    normal = np.zeros((H, W, 3))
    marginal = np.zeros((H, W), dtype=bool)
    a = (d * d).sum(-1)
    for k in range(sb.np):
        if keep is not None and not keep[k]:
            continue
        c, R = sb.x[k], sb.radius[k]
        oc = e - c
        b = 2.0 * (d * oc).sum(-1)
        cc = oc @ oc - R * R
        disc = b * b - 4.0 * a * cc
        # rays whose distance to the centre is within 0.1 % of R: the FP32
        # test may go either way
        marginal |= np.abs(disc) < 2e-3 * 4.0 * a * R * R
        hit = disc > 0
        t = (-b - np.sqrt(np.where(hit, disc, 0.0))) / (2.0 * a)
        better = hit & (np.abs(t) < best)
        p = e + t[..., None] * d
        n = (p - c) / R
        best = np.where(better, np.abs(t), best)
        who = np.where(better, k, who)
        normal = np.where(better[..., None], n, normal)
    shade = 200.0 * np.maximum(0.0, (normal * light).sum(-1)) + 50.0
    if colour is None:
        factors = np.broadcast_to(np.array([0.48, 0.41, 0.27]), (H, W, 3))
    else:
        factors = colour[np.maximum(who, 0)]
    rgb = np.where((who >= 0)[..., None], np.floor(shade[..., None] * factors), 255.0)
    return rgb, who >= 0, marginal


def _agree(img, want, hit, marginal, tol=1):
    ok = ~marginal
    got_hit = ~(img == 255).all(-1)
    assert np.array_equal(got_hit[ok], hit[ok])
    diff = np.abs(img.astype(float) - want)[ok & hit]
    assert diff.max() <= tol, diff.max()
    assert hit.sum() > 20


def test_single_sphere_shading(tmp_path):
    sb = _scene([[0.5, 0.5, 0.5]], [0.15], sid="rt-one")
    img = Oracle.from_spherebin(sb, tmp_path).render(0, width=64, height=64)
    _agree(img, *_textbook(sb, 64, 64))


def _bar(r):
    r = min(max(r, 0.01), 0.99)
    return np.array([min(1.0, 0.209 * r ** 3 - 2.49 * r * r + 3.0 * r + 0.0109),
                     min(1.0, -2.44 * r * r + 2.15 * r + 0.369),
                     min(1.0, -2.21 * r * r + 1.61 * r + 0.573)])


def test_occlusion_picks_nearest_root(tmp_path):
    """Three overlapping silhouettes coloured by x displacement (method 4):
    each pixel shows the sphere whose smaller root has the least |t|."""
    cs = [[0.45, 0.35, 0.5], [0.5, 0.55, 0.52], [0.56, 0.75, 0.49]]
    xs = [0.0, 500.0, 1000.0]
    sb = _scene(cs, [0.12, 0.12, 0.12], sid="rt-occl", xysum=np.array([[x, 0.0] for x in xs]))
    img = Oracle.from_spherebin(sb, tmp_path).render(4, max_val=1000.0, width=96, height=96)
    colour = np.array([_bar(x / 1000.0) for x in xs])
    rgb, hit, marg = _textbook(sb, 96, 96, colour=colour)
    _agree(img, rgb, hit, marg)
    seen = {tuple(p) for p in img.reshape(-1, 3)}
    assert len(seen) > 3


def test_cutoff_and_fixed_spheres(tmp_path):
    """Method 2 (|vel|): with a lower cutoff, slow and fixed spheres are not
    drawn; without it a fixed sphere is drawn grey (0.5 x shading)."""
    cs = [[0.3, 0.5, 0.5], [0.5, 0.5, 0.5], [0.7, 0.5, 0.5]]
    vel = np.array([[0.1, 0.0, 0.0], [3.0, 4.0, 0.0], [0.0, 8.0, 6.0]])
    fixvel = np.array([0.0, 0.0, 1.0])
    sb = _scene(cs, [0.08, 0.08, 0.08], sid="rt-cut", vel=vel, fixvel=fixvel)
    o = Oracle.from_spherebin(sb, tmp_path)
    speeds = np.linalg.norm(vel, axis=1)
    cut = o.render(2, max_val=10.0, lower_cutoff=1.0, width=80, height=80)
    colour = np.array([_bar(s / 10.0) for s in speeds])
    rgb, hit, marg = _textbook(sb, 80, 80, colour=colour, keep=[False, True, False])
    _agree(cut, rgb, hit, marg)
    full = o.render(2, max_val=10.0, width=80, height=80)
    colour[2] = 0.5
    rgb, hit, marg = _textbook(sb, 80, 80, colour=colour)
    _agree(full, rgb, hit, marg)


def test_wide_image_has_a_flat_plane(tmp_path):
    """height / width is an integer division in cameraInit: an image wider
    than tall gets a zero-height image plane, every row the same."""
    sb = _scene([[0.5, 0.5, 0.3]], [0.3], sid="rt-wide")
    img = Oracle.from_spherebin(sb, tmp_path).render(0, width=64, height=32)
    assert (img == img[0]).all()
    assert (img[0] != 255).any()
