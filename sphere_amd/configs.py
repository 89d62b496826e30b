"""Seeded synthetic inputs for the benchmark and parity workloads.

Each builder returns a :class:`~sphere_amd.spherebin.Spherebin` that writes a
reference-layout ``.bin``.  Seeds default to ``20261202 + index``.  Setup
conventions follow the reference's Python front end where it has one:
grid cells at least one maximum diameter (python/sphere.py:909-937), an upper
wall placed on the highest particle with the default wall mass
rho*np*pi*(cell/2)^3 (python/sphere.py:1199-1242), stress-controlled
consolidation (python/sphere.py:1243-1258) and fixed-velocity shear layers
(python/sphere.py:1297-1356).
"""

from __future__ import annotations

import math

import numpy as np

from .spherebin import FIXVEL_HOLD_Z, Spherebin

BASE_SEED = 20261202


def _dt(sb: Spherebin, factor: float = 0.17) -> float:
    r_min = float(np.min(sb.radius))
    m_min = 4.0 / 3.0 * math.pi * r_min ** 3 * sb.rho
    return factor * math.sqrt(m_min / max(sb.k_n, sb.k_t))


def _contact_params(sb: Spherebin, k_n=1e6, k_t=1e6, gamma_n=100.0, gamma_t=100.0, mu=0.5, rho=2600.0):
    sb.k_n, sb.k_t = float(k_n), float(k_t)
    sb.gamma_n, sb.gamma_t = float(gamma_n), float(gamma_t)
    sb.mu_s = sb.mu_d = float(mu)
    sb.gamma_wn, sb.gamma_wt = float(gamma_n), float(gamma_t)
    sb.mu_ws = sb.mu_wd = float(mu)
    sb.rho = float(rho)
    sb.contactmodel = 2


def _grid(sb: Spherebin, L, cell_min: float):
    sb.origo = np.zeros(3)
    sb.L = np.asarray(L, dtype=np.float64)
    sb.num = np.maximum(1, np.floor(sb.L / cell_min)).astype(np.uint32)


def lognormal_radii(rng: np.random.Generator, n: int, median=1e-2, sigma=0.2, lo=0.5, hi=2.0):
    r = median * np.exp(sigma * rng.standard_normal(n))
    return np.clip(r, lo * median, hi * median)


def lattice_settle(nside=(16, 16, 16), spacing=2.4e-2, jitter=1e-3, r_lo=0.9e-2, r_hi=1.1e-2,
                   periodic=1, gravity=-9.81, steps=10000, seed=BASE_SEED, nw=0,
                   sid="c0-settle", k_n=1e6, k_t=1e6, gamma_n=100.0, gamma_t=100.0,
                   mu=0.5) -> Spherebin:
    """Jittered cubic lattice settling under gravity onto the floor.

    With the defaults this is the 4,096-sphere, x/y-periodic settling case:
    radii U[0.9, 1.1]e-2 m, rho 2600, k_n = k_t = 1e6, gamma = 100, mu = 0.5,
    dt = 0.17 sqrt(m_min/k_n), 10,000 steps.
    """
    rng = np.random.default_rng(seed)
    nx, ny, nz = nside
    n = nx * ny * nz
    sb = Spherebin(n, nw=0, sid=sid)
    sb.set_walls(nw)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sites = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    # original order: x fastest, then y, then z
    sites = sites[np.lexsort((sites[:, 0], sites[:, 1], sites[:, 2]))]
    sb.x = (sites + 0.5) * spacing + rng.uniform(-jitter, jitter, size=(n, 3))
    sb.radius = rng.uniform(r_lo, r_hi, size=n)
    _contact_params(sb, k_n=k_n, k_t=k_t, gamma_n=gamma_n, gamma_t=gamma_t, mu=mu)
    L = np.array([nx * spacing, ny * spacing, nz * spacing])
    _grid(sb, L, 2.0 * r_hi)
    sb.periodic = int(periodic)
    sb.g = np.array([0.0, 0.0, gravity])
    sb.time_dt = _dt(sb)
    sb.time_current = 0.0
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    if nw:
        sb.w_x[0] = L[2]
        sb.w_m[0] = sb.rho * n * math.pi * (L[0] / sb.num[0] / 2.0) ** 3
    return sb


def random_compression(n=1048576, box=(2.07, 2.07, 4.13), periodic=1, stress=10e3, gravity=-9.81,
                       steps=2000, seed=BASE_SEED + 1, sid="c1-compress", wall_mass_scale=1.0,
                       height_margin=1.1) -> Spherebin:
    """Random non-overlapping lognormal packing under a stress-controlled top wall.

    Radii: lognormal, median 1e-2 m, sigma 0.2, clipped to [0.5, 2] x median.
    The 0.8 x 0.8 x 1.6 m box of the nominal workload cannot hold 1,048,576
    such spheres (5.26 m^3 of solid), so the default box is 2.07 x 2.07 x
    4.13 m (packing fraction about 0.3); the grid height is 1.1 x the packing
    height as adjustUpperWall leaves it.
    """
    from .dem import pack_rsa

    rng = np.random.default_rng(seed)
    sb = Spherebin(n, nw=1, sid=sid)
    sb.radius = lognormal_radii(rng, n)
    _contact_params(sb)
    box = np.asarray(box, dtype=np.float64)
    sb.x = pack_rsa(sb.radius, box, periodic, seed)
    r_max = float(np.max(sb.radius))
    L = np.array([box[0], box[1], box[2] * height_margin])
    _grid(sb, L, 2.0 * r_max)
    sb.periodic = int(periodic)
    sb.g = np.array([0.0, 0.0, gravity])
    sb.set_walls(1)
    sb.wmode = np.array([1], dtype=np.int32)
    sb.w_x[0] = float(np.max(sb.x[:, 2] + sb.radius))
    cell = L[0] / sb.num[0]
    sb.w_m[0] = wall_mass_scale * sb.rho * n * math.pi * (cell / 2.0) ** 3
    sb.w_devs[0] = float(stress)
    sb.time_dt = _dt(sb)
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    return sb


def ensemble_member(n=4096, k_n=1e6, mu=0.5, stress=10e3, gravity=-9.81, steps=1000, packing=0.3,
                    gamma=None, seed=BASE_SEED + 4, sid="c4-member") -> Spherebin:
    """One packing of the calibration ensemble: ``n`` lognormal spheres (as
    :func:`random_compression`) placed at random without overlap in an
    x/y-periodic box of aspect 1:1:2 sized for the packing fraction, settling
    under gravity below a top wall held at ``stress`` (10 kPa); contact
    stiffness ``k_n`` (= k_t) and friction ``mu`` per member, dt = 0.17
    sqrt(m_min / k_n).  The dashpots default to 100 sqrt(k_n / 1e6) N s/m:
    the damping ratio of the 1e6 N/m, 100 N s/m contacts of the other
    workloads at every stiffness.  (A fixed 100 N s/m on the softest members,
    k_n = 1e5, gives gamma dt / m_min = 1.5; the compressed top layer then
    blows apart within about 1,100 steps, in the oracle as on the GPU.)"""
    if gamma is None:
        gamma = 100.0 * math.sqrt(k_n / 1e6)
    from .dem import pack_rsa

    rng = np.random.default_rng(seed)
    sb = Spherebin(n, nw=1, sid=sid)
    sb.radius = lognormal_radii(rng, n)
    _contact_params(sb, k_n=k_n, k_t=k_n, gamma_n=gamma, gamma_t=gamma, mu=mu)
    solid = float(np.sum(4.0 / 3.0 * math.pi * sb.radius ** 3))
    side = (solid / packing / 2.0) ** (1.0 / 3.0)
    r_max = float(np.max(sb.radius))
    side = max(side, 3.0 * 2.0 * r_max)  # at least three cells across the periodic axes
    box = np.array([side, side, 2.0 * side])
    sb.x = pack_rsa(sb.radius, box, 1, seed)
    L = np.array([box[0], box[1], box[2] * 1.1])
    _grid(sb, L, 2.0 * r_max)
    sb.periodic = 1
    sb.g = np.array([0.0, 0.0, gravity])
    sb.set_walls(1)
    sb.wmode = np.array([1], dtype=np.int32)
    sb.w_x[0] = float(np.max(sb.x[:, 2] + sb.radius))
    cell = L[0] / sb.num[0]
    sb.w_m[0] = sb.rho * n * math.pi * (cell / 2.0) ** 3
    sb.w_devs[0] = float(stress)
    sb.time_dt = _dt(sb)
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    return sb


def calibration_ensemble(members=512, n_lo=1024, n_hi=32768, k_lo=1e5, k_hi=1e7, mu_lo=0.2, mu_hi=0.8,
                         steps=1000, seed=BASE_SEED + 4):
    """The calibration ensemble: ``members`` packings with seeded particle
    counts uniform in [n_lo, n_hi], k_n uniform in [k_lo, k_hi] and mu
    uniform in [mu_lo, mu_hi] (512 members hold about 8.6M spheres)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_lo, n_hi + 1, size=members)
    ks = rng.uniform(k_lo, k_hi, size=members)
    mus = rng.uniform(mu_lo, mu_hi, size=members)
    return [ensemble_member(n=int(ns[i]), k_n=float(ks[i]), mu=float(mus[i]), steps=steps,
                            seed=seed + 1000 + i, sid=f"c4-{i:03d}") for i in range(members)]


def fcc_bonded(cells=(4, 4, 8), radius=1e-2, jitter_frac=0.01, bond_reach=1.05, k_bond=1e7,
               lambda_bar=0.5, sigma_b=math.inf, tau_b=math.inf, pull_velocity=0.01,
               gravity=0.0, steps=5000, seed=BASE_SEED + 2, sid="c2-bonded") -> Spherebin:
    """Jittered FCC solid with parallel bonds between near neighbours.

    Bonds join every pair with centre distance <= 1.05 (r_i + r_j).  The bond
    stiffness enters through the contact k_n = k_t (the reference scales it by
    the bond area, src/cohesion.cuh:124-177).  The top and bottom layers are
    pulled apart at +-pull_velocity / 2 by holding their whole velocity:
    fixvel = FIXVEL_HOLD_Z (2), this build's extension of the reference's
    fixed horizontal velocity (fixvel = 1 keeps vel_x and vel_y only and
    lets the bonds decelerate the layers in z, src/integration.cuh:65-77).
    No periodic axes, no walls beyond the reference defaults.
    """
    rng = np.random.default_rng(seed)
    a = 2.0 * math.sqrt(2.0) * radius
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cx, cy, cz = cells
    ii, jj, kk = np.meshgrid(np.arange(cx), np.arange(cy), np.arange(cz), indexing="ij")
    corners = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    pts = (corners[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    pts = pts[np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))]
    n = pts.shape[0]
    margin = 2.0 * radius
    pts = pts + margin + rng.uniform(-jitter_frac * radius, jitter_frac * radius, size=pts.shape)
    sb = Spherebin(n, nw=0, sid=sid)
    sb.set_walls(0)
    sb.x = pts
    sb.radius = np.full(n, radius)
    _contact_params(sb, k_n=k_bond, k_t=k_bond, gamma_n=0.0, gamma_t=0.0)
    ext = np.array([cx * a, cy * a, cz * a]) + 2 * margin
    L = ext + np.array([0.0, 0.0, 0.2 * ext[2]])
    _grid(sb, L, 2.0 * radius * 1.0001)
    sb.periodic = 0
    sb.g = np.array([0.0, 0.0, gravity])
    # bonds: neighbour pairs within reach (cell grid search)
    from scipy.spatial import cKDTree

    tree = cKDTree(sb.x)
    pairs = np.array(sorted(tree.query_pairs(r=bond_reach * 2.0 * radius)), dtype=np.uint32).reshape(-1, 2)
    sb.set_bonds(pairs)
    sb.lambda_bar = float(lambda_bar)
    sb.sigma_b = float(sigma_b)
    sb.tau_b = float(tau_b)
    z = sb.x[:, 2]
    zlo, zhi = z.min(), z.max()
    bottom = z < zlo + 0.5 * a
    top = z > zhi - 0.5 * a
    sb.fixvel[bottom | top] = FIXVEL_HOLD_Z
    sb.vel[top, 2] = pull_velocity / 2.0
    sb.vel[bottom, 2] = -pull_velocity / 2.0
    sb.time_dt = _dt(sb, 0.17) * 0.5
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    return sb


def shear_cell(n=65536, box=None, stress=10e3, shear_velocity=0.1, gravity=-9.81, steps=2000,
               seed=BASE_SEED + 3, sid="c3-shear", packing_fraction=0.3) -> Spherebin:
    """Polydisperse packing (config-1 radii), x/y periodic, top wall at 10 kPa,
    the top layer moved at a fixed horizontal velocity and the bottom layer
    held (python/sphere.py:1297-1356)."""
    from .dem import pack_rsa

    rng = np.random.default_rng(seed)
    r = lognormal_radii(rng, n)
    if box is None:
        vol = float(np.sum(4.0 / 3.0 * math.pi * r ** 3)) / packing_fraction
        side = (vol / 2.0) ** (1.0 / 3.0)
        box = (side, side, 2.0 * side)
    sb = random_compression(n=n, box=box, periodic=1, stress=stress, gravity=gravity, steps=steps,
                            seed=seed, sid=sid)
    sb.radius = r
    sb.x = pack_rsa(r, np.asarray(box, float), 1, seed)
    sb.w_x[0] = float(np.max(sb.x[:, 2] + sb.radius))
    z = sb.x[:, 2]
    zmin, zmax = float(np.min(z - r)), float(np.max(z + r))
    d_bot = 2.0 * float(np.max(r[z < zmin + 0.3 * (zmax - zmin)]))
    d_top = 2.0 * float(np.max(r[z > zmin + 0.7 * (zmax - zmin)]))
    bot = z < zmin + d_bot
    top = z > zmax - d_top
    sb.fixvel[bot | top] = 1.0
    sb.vel[bot, :2] = 0.0
    sb.vel[top, 0] = shear_velocity
    sb.vel[top, 1] = 0.0
    sb.gamma_wn = sb.gamma_wt = 0.0
    sb.mu_ws = sb.mu_wd = 0.0
    sb.time_dt = _dt(sb)
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    return sb


def consolidated_shear_cell(n=65536, stress=10e3, shear_velocity=0.1, gamma_n=2.0, gamma_t=10.0, mu=0.5,
                            steps=2000, seed=BASE_SEED + 3, sid="c3-shear", fall_steps=40000, settle_steps=160000,
                            settle_gamma=10.0, wall_steps=20000, wall_mass_scale=1e-3, chunk=5000,
                            log=None) -> Spherebin:
    """The differentiable shear test's input: shear_cell's seeded lognormal
    packing (config-1 radii, x/y periodic), compacted on the GPU under its
    10 kPa top wall, then set up for shearing.

    1. Settling under gravity with contact model 1 (contactLinearViscous,
       src/contactmodels.cuh:114-253, whose normal dashpot damps: model 2's
       feeds energy in on approach) and no friction (a frictionless packing
       jams with about six contacts per sphere), the top wall parked at the
       top of the grid: `fall_steps` steps of free fall onto the floor, then
       every `chunk` steps up to `settle_steps` every sphere keeps only its
       downward velocity (horizontal and upward components of the velocity
       and half-step velocity zeroed, spins stopped), so that the small
       spheres the collapse throws up fall back and the bed comes to rest.
       The
       dashpot is `settle_gamma` = 10 N s/m: the explicit update of a dashpot
       is stable only while dt * (sum of the dashpots on a sphere) / m stays
       below 2, and with the smallest spheres (m = 1.4e-3 kg) touching a
       dozen others at dt = 6.3e-6 s, 100 N s/m is past that.
    2. The top wall placed just above the packing and held at `stress`
       (wmode 1, the reference's consolidate, python/sphere.py:1243-1258)
       for `wall_steps` steps, its mass the reference's default times
       `wall_mass_scale` (0.001, as the reference's own shear example scales
       it, python/shear-test.py:77), so that it reaches the packing early.
    3. Contact model 2 with friction `mu`, a normal dashpot `gamma_n`
       (2 N s/m: model 2's normal dashpot feeds energy in on approach, so it
       stays small) and a tangential dashpot `gamma_t` within the same
       stability limit (10 N s/m: the tangential force also spins the sphere
       up, 3.5 times its effect on the sliding speed alone),
       frictionless undamped walls, and the shear set-up of shear_cell: the
       top layer driven at `shear_velocity` in x and the bottom layer held
       (fixvel, python/sphere.py:1297-1356).

    Runs on the GPU (sphere_amd.dem).  `log(msg)` receives progress lines.
    """
    from .dem import DEM

    say = log or (lambda msg: None)
    sb = shear_cell(n=n, stress=stress, shear_velocity=shear_velocity, steps=steps, seed=seed, sid=sid)
    r = np.asarray(sb.radius)
    dt = float(sb.time_dt)

    def fresh_clock(s, nsteps):
        s.time_current = 0.0
        s.time_total = (nsteps + 1) * dt
        s.time_file_dt = s.time_total
        s.time_step_count = 0

    def fastest(d):
        return float(np.max(np.linalg.norm(d.get("vel")[:, :3], axis=1)))

    def calm(d):  # keep only downward motion: spheres thrown up fall back, the rest rests
        v = d.get("vel")
        v[:, :2] = 0.0
        v[:, 2] = np.minimum(v[:, 2], 0.0)
        d.set("vel", v)
        v0 = d.get("vel0")
        v0[:, :2] = 0.0
        v0[:, 2] = np.minimum(v0[:, 2], 0.0)
        d.set("vel0", v0)
        z3 = np.zeros((sb.np, 3))
        d.set("angvel", z3)
        d.set("angvel0", z3)

    sb.fixvel[:] = 0.0
    sb.vel[:] = 0.0
    sb.contactmodel = 1
    sb.gamma_n = sb.gamma_t = float(settle_gamma)
    sb.gamma_wn, sb.gamma_wt = float(settle_gamma), 0.0
    sb.mu_s = sb.mu_d = 0.0
    sb.mu_ws = sb.mu_wd = 0.0
    sb.wmode = np.array([0], dtype=np.int32)
    sb.w_x[0] = float(sb.L[2])
    fresh_clock(sb, settle_steps)
    dem = DEM(sb, verbosity=0)
    done = 0
    while done < settle_steps:
        k = min(chunk, settle_steps - done)
        dem.step(k)
        done += k
        if done % (4 * chunk) == 0 or done == settle_steps:
            x = dem.get("x")
            say(f"settling: {done} steps, fastest sphere {fastest(dem):.3g} m/s, "
                f"highest top {float(np.max(x[:, 2] + x[:, 3])):.4f} m")
        if done >= fall_steps:
            calm(dem)
    sb = dem.to_spherebin()
    dem.close()
    tops = sb.x[:, 2] + sb.radius
    top = float(np.max(tops))
    say(f"settled: highest top {top:.4f} m, 99.9th percentile {float(np.percentile(tops, 99.9)):.4f} m")
    sb.wmode = np.array([1], dtype=np.int32)
    sb.w_x[0] = top + 0.1 * float(np.median(r))
    sb.w_vel[0] = 0.0
    sb.w_devs[0] = float(stress)
    sb.w_m[0] = float(sb.w_m[0]) * wall_mass_scale
    fresh_clock(sb, wall_steps)
    dem = DEM(sb, verbosity=0)
    done = 0
    while done < wall_steps:
        k = min(chunk, wall_steps - done)
        dem.step(k)
        done += k
    ij, _ = dem.findOverlaps()
    cn = 2.0 * len(ij) / sb.np
    wall = dem.get("walls_nx")[0, 3]
    say(f"under the wall: {wall_steps} steps, wall at {wall:.4f} m (packing top {top:.4f}), C/N = {cn:.2f}, "
        f"fastest sphere {fastest(dem):.3g} m/s")
    sb = dem.to_spherebin()
    dem.close()
    sb.contactmodel = 2
    sb.gamma_n, sb.gamma_t = float(gamma_n), float(gamma_t)
    sb.mu_s = sb.mu_d = float(mu)
    sb.gamma_wn = sb.gamma_wt = 0.0
    sb.mu_ws = sb.mu_wd = 0.0
    z = sb.x[:, 2]
    rr = sb.radius
    zmin, zmax = float(np.min(z - rr)), float(np.max(z + rr))
    d_bot = 2.0 * float(np.max(rr[z < zmin + 0.3 * (zmax - zmin)]))
    d_top = 2.0 * float(np.max(rr[z > zmin + 0.7 * (zmax - zmin)]))
    bot = z < zmin + d_bot
    topl = z > zmax - d_top
    sb.vel[:] = 0.0
    sb.angvel[:] = 0.0
    sb.fixvel[:] = 0.0
    sb.fixvel[bot | topl] = 1.0
    sb.vel[topl, 0] = shear_velocity
    sb.time_current = 0.0
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    sb.time_step_count = 0
    sb.sid = sid
    return sb


def dense_lattice(cells=(64, 64, 64), radius=1e-2, overlap=2e-5, jitter=1e-6, gravity=0.0, steps=2000,
                  seed=BASE_SEED + 5, sid="dense-fcc") -> Spherebin:
    """Confined FCC crystal, 4 x 64^3 = 1,048,576 spheres, every nearest
    neighbour pair overlapping by `overlap`: twelve contacts per sphere in the
    interior for the whole run, the contact load the loose random packing
    does not reach within its time window.  x/y periodic at the lattice
    period; the floor and a fixed top wall overlap the outer layers by
    overlap/2.  The normal dashpot is off (gamma_n = 0): model 2's normal
    dashpot adds energy on approach (vel_n = -v_ab . n,
    src/contactmodels.cuh:283-493) and blows a compressed crystal apart."""
    rng = np.random.default_rng(seed)
    r_eff = radius - overlap / 2.0
    a = 2.0 * math.sqrt(2.0) * r_eff
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cx, cy, cz = cells
    kk, jj, ii = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    corners = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    pts = ((corners[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    pts = pts[np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))]
    pts[:, 0] += a / 4.0
    pts[:, 1] += a / 4.0
    pts[:, 2] += r_eff
    n = pts.shape[0]
    sb = Spherebin(n, nw=1, sid=sid)
    sb.x = pts + rng.uniform(-jitter, jitter, size=pts.shape)
    sb.radius = np.full(n, radius)
    _contact_params(sb, gamma_n=0.0)
    top = float(pts[:, 2].max()) + r_eff
    L = np.array([cx * a, cy * a, top * 1.05])
    _grid(sb, L, 2.0 * radius * 1.0001)
    sb.periodic = 1
    sb.g = np.array([0.0, 0.0, gravity])
    sb.set_walls(1)
    sb.wmode = np.array([0], dtype=np.int32)
    sb.w_x[0] = top
    sb.w_m[0] = sb.rho * n * math.pi * (L[0] / sb.num[0] / 2.0) ** 3
    sb.time_dt = _dt(sb)
    sb.time_total = steps * sb.time_dt
    sb.time_file_dt = sb.time_total
    return sb


def fcc_shear(cells=(16, 16, 64), shear_velocity=0.1, stress=10e3, gamma_n=0.0, sid="fcc-shear",
              **kw) -> Spherebin:
    """Dense shear cell: the confined FCC crystal of dense_lattice with its
    top layer driven at a fixed horizontal velocity and its bottom layer held
    (fixvel, python/sphere.py:1297-1356), under a stress-controlled top wall.
    4 x 16 x 16 x 64 = 65,536 spheres by default.  Stands in for a compacted
    random packing, which the seeded random placement cannot produce without
    a long settling run."""
    sb = dense_lattice(cells=cells, sid=sid, **kw)
    sb.gamma_n = float(gamma_n)
    sb.gamma_wn = 0.0
    z = sb.x[:, 2]
    a = 2.0 * math.sqrt(2.0) * float(sb.radius[0])
    top = z > z.max() - 0.25 * a
    bot = z < z.min() + 0.25 * a
    sb.fixvel[top | bot] = 1.0
    sb.vel[top, 0] = shear_velocity
    sb.wmode = np.array([1], dtype=np.int32)
    sb.w_devs[0] = float(stress)
    return sb


def two_spheres(distance=0.0, radius=0.5, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0), bonded=False,
                gamma=0.0, k=1.16e9, mu=0.4, L=20.0, periodic=1, gravity=0.0, sid="pair") -> Spherebin:
    """Two spheres on the x axis, the layout of tests/bond_tests.py:42-50."""
    sb = Spherebin(2, nw=0, sid=sid)
    sb.set_walls(0)
    sb.x[0] = (10.0, 10.0, 10.0)
    sb.x[1] = (10.0 + 2.0 * radius + distance, 10.0, 10.0)
    sb.radius[:] = radius
    sb.vel[1] = v
    sb.angvel[1] = w
    _contact_params(sb, k_n=k, k_t=k, gamma_n=gamma, gamma_t=gamma, mu=mu)
    sb.gamma_wn = sb.gamma_wt = 0.0
    _grid(sb, (L, L, L), 2.0 * radius * 1.0001)
    sb.periodic = int(periodic)
    sb.g = np.array([0.0, 0.0, gravity])
    if bonded:
        sb.bond(0, 1)
        sb.lambda_bar = 1.0
    sb.time_dt = _dt(sb, 0.075)
    sb.time_total = 0.2
    sb.time_file_dt = 0.01
    return sb
