"""In-memory sphere state and the little-endian ``.bin`` codec.

``Spherebin`` mirrors the attribute names of the reference's Python front end
(python/sphere.py:14-140) so that setup and analysis scripts read the same, and
``readbin``/``writebin`` follow the on-disk layout that the C++ side reads and
writes (src/file_io.cpp:29-301 read, :304-490 write; python/sphere.py:254-529).

Layout (no padding, no magic):

====================  =====================================================
header                nd u32 (=3), np u32
time                  dt, current, total, file_dt f64; step_count u32
grid                  origo[3] f64, L[3] f64, num[3] u32, periodic i32
particles             (x, y, z, r) x np; xysum[2] x np; (vx, vy, vz, fixvel)
                      x np; force[3], angpos[3], angvel[3], torque[3] x np
                      each; es_dot, es, ev_dot, ev, p x np each (f64)
params                g[3], k_n, k_t, k_r, gamma_n, gamma_t, gamma_r, mu_s,
                      mu_d, mu_r, gamma_wn, gamma_wt, mu_ws, mu_wd, rho f64;
                      contactmodel u32; kappa, db, V_b f64
walls                 nw u32; wmode[nw] i32; (nx, ny, nz, pos) x nw f64;
                      (mass, vel, force, devs) x nw f64; devs_A, devs_f f64
bonds                 lambda_bar f64, nb0 u32, sigma_b, tau_b f64;
                      (i, j) x nb0 u32; delta_n[nb0]; delta_t[nb0][3];
                      omega_n[nb0]; omega_t[nb0][3] f64
fluid                 nu f64 (fluid cell data when nu > 0 is not supported)
====================  =====================================================
"""

from __future__ import annotations

import io
import math
import os

import numpy as np

MAXWALLS = 6
PI_REF = 3.14159265358979  # src/constants.h:14
# fixvel > 0 keeps a particle's x and y velocity (src/integration.cuh:65-77);
# this build's extension: fixvel >= FIXVEL_HOLD_Z keeps its z velocity too
FIXVEL_HOLD_Z = 2.0


class BinFormatError(ValueError):
    """Raised for a truncated or malformed ``.bin`` file."""


class Spherebin:
    """Particle, wall, bond and parameter arrays of one simulation.

    Defaults follow python/sphere.py:22-140.
    """

    def __init__(self, np_: int = 1, nd: int = 3, nw: int = 1, sid: str = "unnamed"):
        n = int(np_)
        self.sid = sid
        self.nd = int(nd)
        self.np = n
        self.time_dt = 0.0
        self.time_current = 0.0
        self.time_total = 0.0
        self.time_file_dt = 0.0
        self.time_step_count = 0
        self.origo = np.zeros(3)
        self.L = np.zeros(3)
        self.num = np.zeros(3, dtype=np.uint32)
        self.periodic = 0
        self.x = np.zeros((n, 3))
        self.radius = np.ones(n)
        self.xysum = np.zeros((n, 2))
        self.vel = np.zeros((n, 3))
        self.fixvel = np.zeros(n)
        self.force = np.zeros((n, 3))
        self.angpos = np.zeros((n, 3))
        self.angvel = np.zeros((n, 3))
        self.torque = np.zeros((n, 3))
        self.es_dot = np.zeros(n)
        self.es = np.zeros(n)
        self.ev_dot = np.zeros(n)
        self.ev = np.zeros(n)
        self.p = np.zeros(n)
        self.g = np.zeros(3)
        self.k_n = 1.16e9
        self.k_t = 1.16e9
        self.k_r = 0.0
        self.gamma_n = 0.0
        self.gamma_t = 0.0
        self.gamma_r = 0.0
        self.mu_s = 1.0
        self.mu_d = 1.0
        self.mu_r = 0.0
        self.gamma_wn = 1.0e3
        self.gamma_wt = 1.0e3
        self.mu_ws = 1.0
        self.mu_wd = 1.0
        self.rho = 2600.0
        self.contactmodel = 2
        self.kappa = 0.0
        self.db = 0.0
        self.V_b = 0.0
        self.nw = int(nw)
        self.wmode = np.zeros(self.nw, dtype=np.int32)
        self.w_n = np.zeros((self.nw, 3))
        default_normals = [(0, 0, -1), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0)]
        for i in range(min(self.nw, 5)):
            self.w_n[i] = default_normals[i]
        self.w_x = np.ones(self.nw)
        self.w_m = np.zeros(self.nw)
        self.w_vel = np.zeros(self.nw)
        self.w_force = np.zeros(self.nw)
        self.w_devs = np.zeros(self.nw)
        self.w_devs_A = 0.0
        self.w_devs_f = 0.0
        self.lambda_bar = 1.0
        self.sigma_b = math.inf
        self.tau_b = math.inf
        self.bonds = np.zeros((0, 2), dtype=np.uint32)
        self.bonds_delta_n = np.zeros(0)
        self.bonds_delta_t = np.zeros((0, 3))
        self.bonds_omega_n = np.zeros(0)
        self.bonds_omega_t = np.zeros((0, 3))
        self.nu = 0.0

    # ------------------------------------------------------------------ sizes
    @property
    def nb0(self) -> int:
        return int(self.bonds.shape[0])

    # ------------------------------------------------------------------ codec
    def to_bytes(self) -> bytes:
        """Serialise in the layout of src/file_io.cpp:319-452."""
        n = self.np
        nw = self.nw
        nb = self.nb0
        if nw > MAXWALLS:
            raise BinFormatError(f"nw={nw} exceeds MAXWALLS={MAXWALLS}")
        out = io.BytesIO()
        w = out.write
        w(np.array([self.nd, n], dtype="<u4").tobytes())
        w(np.array([self.time_dt, self.time_current, self.time_total,
                    self.time_file_dt], dtype="<f8").tobytes())
        w(np.array([self.time_step_count], dtype="<u4").tobytes())
        w(np.asarray(self.origo, dtype="<f8").tobytes())
        w(np.asarray(self.L, dtype="<f8").tobytes())
        w(np.asarray(self.num, dtype="<u4").tobytes())
        w(np.array([self.periodic], dtype="<i4").tobytes())
        xr = np.empty((n, 4))
        xr[:, :3] = self.x
        xr[:, 3] = self.radius
        w(xr.astype("<f8").tobytes())
        w(np.asarray(self.xysum, dtype="<f8").tobytes())
        vf = np.empty((n, 4))
        vf[:, :3] = self.vel
        vf[:, 3] = self.fixvel
        w(vf.astype("<f8").tobytes())
        for a in (self.force, self.angpos, self.angvel, self.torque):
            w(np.asarray(a, dtype="<f8").reshape(n, 3).tobytes())
        for a in (self.es_dot, self.es, self.ev_dot, self.ev, self.p):
            w(np.asarray(a, dtype="<f8").reshape(n).tobytes())
        w(np.asarray(self.g, dtype="<f8").tobytes())
        w(np.array([self.k_n, self.k_t, self.k_r, self.gamma_n, self.gamma_t,
                    self.gamma_r, self.mu_s, self.mu_d, self.mu_r, self.gamma_wn,
                    self.gamma_wt, self.mu_ws, self.mu_wd, self.rho],
                   dtype="<f8").tobytes())
        w(np.array([self.contactmodel], dtype="<u4").tobytes())
        w(np.array([self.kappa, self.db, self.V_b], dtype="<f8").tobytes())
        w(np.array([nw], dtype="<u4").tobytes())
        w(np.asarray(self.wmode, dtype="<i4").reshape(nw).tobytes())
        nx = np.empty((nw, 4))
        nx[:, :3] = np.asarray(self.w_n).reshape(nw, 3)
        nx[:, 3] = self.w_x
        w(nx.astype("<f8").tobytes())
        mv = np.stack([np.asarray(self.w_m, float), np.asarray(self.w_vel, float),
                       np.asarray(self.w_force, float), np.asarray(self.w_devs, float)],
                      axis=1) if nw else np.zeros((0, 4))
        w(mv.astype("<f8").tobytes())
        w(np.array([self.w_devs_A, self.w_devs_f], dtype="<f8").tobytes())
        w(np.array([self.lambda_bar], dtype="<f8").tobytes())
        w(np.array([nb], dtype="<u4").tobytes())
        w(np.array([self.sigma_b, self.tau_b], dtype="<f8").tobytes())
        w(np.asarray(self.bonds, dtype="<u4").reshape(nb, 2).tobytes())
        w(np.asarray(self.bonds_delta_n, dtype="<f8").reshape(nb).tobytes())
        w(np.asarray(self.bonds_delta_t, dtype="<f8").reshape(nb, 3).tobytes())
        w(np.asarray(self.bonds_omega_n, dtype="<f8").reshape(nb).tobytes())
        w(np.asarray(self.bonds_omega_t, dtype="<f8").reshape(nb, 3).tobytes())
        if self.nu > 0.0:
            raise BinFormatError("fluid cell data (nu > 0) is not supported")
        w(np.array([self.nu], dtype="<f8").tobytes())
        return out.getvalue()

    def writebin(self, path: str) -> str:
        data = self.to_bytes()
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "wb") as fh:
            fh.write(data)
        return path

    @classmethod
    def from_bytes(cls, data: bytes, sid: str = "unnamed") -> "Spherebin":
        """Parse the layout of src/file_io.cpp:45-237."""
        mv = memoryview(data)
        pos = 0

        def take(dtype, count):
            nonlocal pos
            size = np.dtype(dtype).itemsize * count
            if pos + size > len(mv):
                raise BinFormatError(f"truncated .bin: need {size} bytes at offset {pos}")
            arr = np.frombuffer(mv[pos:pos + size], dtype=dtype, count=count).copy()
            pos += size
            return arr

        nd, n = (int(v) for v in take("<u4", 2))
        if nd != 3:
            raise BinFormatError(f"nd={nd}; only 3-D data is supported")
        sb = cls.__new__(cls)
        sb.sid = sid
        sb.nd = nd
        sb.np = n
        t = take("<f8", 4)
        sb.time_dt, sb.time_current, sb.time_total, sb.time_file_dt = (float(v) for v in t)
        sb.time_step_count = int(take("<u4", 1)[0])
        sb.origo = take("<f8", 3)
        sb.L = take("<f8", 3)
        sb.num = take("<u4", 3)
        sb.periodic = int(take("<i4", 1)[0])
        xr = take("<f8", 4 * n).reshape(n, 4)
        sb.x = xr[:, :3].copy()
        sb.radius = xr[:, 3].copy()
        sb.xysum = take("<f8", 2 * n).reshape(n, 2)
        vf = take("<f8", 4 * n).reshape(n, 4)
        sb.vel = vf[:, :3].copy()
        sb.fixvel = vf[:, 3].copy()
        sb.force = take("<f8", 3 * n).reshape(n, 3)
        sb.angpos = take("<f8", 3 * n).reshape(n, 3)
        sb.angvel = take("<f8", 3 * n).reshape(n, 3)
        sb.torque = take("<f8", 3 * n).reshape(n, 3)
        sb.es_dot = take("<f8", n)
        sb.es = take("<f8", n)
        sb.ev_dot = take("<f8", n)
        sb.ev = take("<f8", n)
        sb.p = take("<f8", n)
        sb.g = take("<f8", 3)
        (sb.k_n, sb.k_t, sb.k_r, sb.gamma_n, sb.gamma_t, sb.gamma_r, sb.mu_s,
         sb.mu_d, sb.mu_r, sb.gamma_wn, sb.gamma_wt, sb.mu_ws, sb.mu_wd,
         sb.rho) = (float(v) for v in take("<f8", 14))
        sb.contactmodel = int(take("<u4", 1)[0])
        sb.kappa, sb.db, sb.V_b = (float(v) for v in take("<f8", 3))
        nw = int(take("<u4", 1)[0])
        if nw > MAXWALLS:
            raise BinFormatError(f"nw={nw} exceeds MAXWALLS={MAXWALLS}")
        sb.nw = nw
        sb.wmode = take("<i4", nw)
        nx = take("<f8", 4 * nw).reshape(nw, 4)
        sb.w_n = nx[:, :3].copy()
        sb.w_x = nx[:, 3].copy()
        mvfd = take("<f8", 4 * nw).reshape(nw, 4)
        sb.w_m = mvfd[:, 0].copy()
        sb.w_vel = mvfd[:, 1].copy()
        sb.w_force = mvfd[:, 2].copy()
        sb.w_devs = mvfd[:, 3].copy()
        sb.w_devs_A, sb.w_devs_f = (float(v) for v in take("<f8", 2))
        sb.lambda_bar = float(take("<f8", 1)[0])
        nb = int(take("<u4", 1)[0])
        sb.sigma_b, sb.tau_b = (float(v) for v in take("<f8", 2))
        sb.bonds = take("<u4", 2 * nb).reshape(nb, 2)
        sb.bonds_delta_n = take("<f8", nb)
        sb.bonds_delta_t = take("<f8", 3 * nb).reshape(nb, 3)
        sb.bonds_omega_n = take("<f8", nb)
        sb.bonds_omega_t = take("<f8", 3 * nb).reshape(nb, 3)
        sb.nu = float(take("<f8", 1)[0])
        if sb.nu > 0.0:
            raise BinFormatError("fluid cell data (nu > 0) is not supported")
        if pos != len(mv):
            raise BinFormatError(f"{len(mv) - pos} trailing bytes after the fluid block")
        return sb

    @classmethod
    def readbin(cls, path: str) -> "Spherebin":
        with open(path, "rb") as fh:
            data = fh.read()
        sid = os.path.basename(path).rsplit(".", 1)[0]
        return cls.from_bytes(data, sid=sid)

    def copy(self) -> "Spherebin":
        return Spherebin.from_bytes(self.to_bytes(), sid=self.sid)

    # --------------------------------------------------------- setup helpers
    def defaultParams(self, mu_s=0.4, mu_d=0.4, mu_r=0.0, rho=2600.0,
                      k_n=1.16e9, k_t=1.16e9, k_r=0.0, gamma_n=0.0,
                      gamma_t=0.0, gamma_r=0.0, gamma_wn=1.0e4,
                      gamma_wt=1.0e4, capillaryCohesion=0):
        """Contact parameters; defaults of python/sphere.py:1391-1474."""
        self.rho = float(rho)
        self.k_n, self.k_t, self.k_r = float(k_n), float(k_t), float(k_r)
        self.gamma_n, self.gamma_t, self.gamma_r = float(gamma_n), float(gamma_t), float(gamma_r)
        self.mu_s, self.mu_d, self.mu_r = float(mu_s), float(mu_d), float(mu_r)
        self.gamma_wn, self.gamma_wt = float(gamma_wn), float(gamma_wt)
        self.mu_ws, self.mu_wd = float(mu_s), float(mu_d)
        self.kappa = 0.0 if not capillaryCohesion else self.kappa

    def zeroKinematics(self):
        """python/sphere.py:1185-1198."""
        self.vel[:] = 0.0
        self.angvel[:] = 0.0
        self.angpos[:] = 0.0
        self.es[:] = 0.0
        self.ev[:] = 0.0
        self.xysum[:] = 0.0

    def initTemporal(self, total, current=0.0, file_dt=0.05, step_count=0,
                     dt_factor=0.075):
        """Time step dt = f*sqrt(m_min/max(k_n,k_t)) (python/sphere.py:1357-1377).

        The reference defaults to f=0.075; f=0.17 (O'Sullivan et al. 2003) is
        the commented-out alternative at python/sphere.py:1369.
        """
        r_min = float(np.min(self.radius))
        m_min = 4.0 / 3.0 * math.pi * r_min ** 3 * self.rho
        self.time_dt = dt_factor * math.sqrt(m_min / max(self.k_n, self.k_t))
        self.time_current = float(current)
        self.time_total = float(total)
        self.time_file_dt = float(file_dt)
        self.time_step_count = int(step_count)

    def bond(self, i: int, j: int):
        """Register an intact bond with zero displacement (python/sphere.py:1476-1509)."""
        self.bonds = np.vstack([self.bonds, np.array([[i, j]], dtype=np.uint32)]).astype(np.uint32)
        self.bonds_delta_n = np.append(self.bonds_delta_n, 0.0)
        self.bonds_delta_t = np.vstack([self.bonds_delta_t, np.zeros((1, 3))])
        self.bonds_omega_n = np.append(self.bonds_omega_n, 0.0)
        self.bonds_omega_t = np.vstack([self.bonds_omega_t, np.zeros((1, 3))])

    def set_bonds(self, pairs):
        pairs = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
        nb = pairs.shape[0]
        self.bonds = pairs.copy()
        self.bonds_delta_n = np.zeros(nb)
        self.bonds_delta_t = np.zeros((nb, 3))
        self.bonds_omega_n = np.zeros(nb)
        self.bonds_omega_t = np.zeros((nb, 3))

    def set_walls(self, nw: int):
        self.nw = int(nw)
        self.wmode = np.zeros(nw, dtype=np.int32)
        self.w_n = np.zeros((nw, 3))
        default_normals = [(0, 0, -1), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, 1)]
        for i in range(nw):
            self.w_n[i] = default_normals[i]
        self.w_x = np.ones(nw)
        self.w_m = np.zeros(nw)
        self.w_vel = np.zeros(nw)
        self.w_force = np.zeros(nw)
        self.w_devs = np.zeros(nw)

    # --------------------------------------------------------------- analysis
    def mass(self) -> np.ndarray:
        """Particle mass with the constant and evaluation order of src/integration.cuh:46."""
        r = self.radius
        return 4.0 / 3.0 * PI_REF * r * r * r * self.rho

    def energy(self, method: str) -> float:
        """Energy sums of python/sphere.py:1515-1562."""
        m = 4.0 / 3.0 * math.pi * self.radius ** 3 * self.rho
        if method == "pot":
            return float(np.sum(m * math.sqrt(float(np.dot(self.g, self.g))) * self.x[:, 2]))
        if method == "kin":
            return float(np.sum(0.5 * m * np.sum(self.vel ** 2, axis=1)))
        if method == "rot":
            return float(np.sum(0.5 * 0.4 * m * self.radius ** 2 * np.sum(self.angvel ** 2, axis=1)))
        if method == "shear":
            return float(np.sum(self.es))
        if method == "shearrate":
            return float(np.sum(self.es_dot))
        if method == "visc_n":
            return float(np.sum(self.ev))
        if method == "visc_n_rate":
            return float(np.sum(self.ev_dot))
        if method == "bondpot":
            if self.nb0 == 0:
                return 0.0
            intact = self.bonds[:, 0] < self.np
            b = self.bonds[intact]
            R = self.lambda_bar * np.minimum(self.radius[b[:, 0]], self.radius[b[:, 1]])
            A = math.pi * R ** 2
            I = 0.25 * math.pi * R ** 4
            J = 2.0 * I
            dn = self.bonds_delta_n[intact]
            dt_ = self.bonds_delta_t[intact]
            on = self.bonds_omega_n[intact]
            ot = self.bonds_omega_t[intact]
            return float(np.sum(0.5 * A * self.k_n * dn ** 2)
                         + np.sum(0.5 * A * self.k_t * np.sum(dt_ ** 2, axis=1))
                         + np.sum(0.5 * J * self.k_t * on ** 2)
                         + np.sum(0.5 * I * self.k_n * np.sum(ot ** 2, axis=1)))
        raise ValueError(f"unknown energy method {method!r}")

    def bulkPorosity(self) -> float:
        """(V_total - V_solid) / V_total with V_total = L0 L1 L2, or L0 L1 w_x[0]
        under a wall (python/sphere.py:1579-1591)."""
        if self.nw == 0:
            V_total = self.L[0] * self.L[1] * self.L[2]
        else:
            V_total = self.L[0] * self.L[1] * self.w_x[0]
            if V_total <= 0.0:
                raise ValueError("Could not determine total volume")
        V_solid = float(np.sum(4.0 / 3.0 * math.pi * self.radius ** 3))
        return float((V_total - V_solid) / V_total)

    def porosity(self, slices: int = 10, exact: bool = False):
        """(porosity, depth) per horizontal slice, as python/sphere.py:1593-1627
        returns them from the `porosity` tool; computed by DEM::porosity on
        the GPU (exact=True: exact sphere volumes per slice)."""
        from .dem import DEM

        with DEM(self, verbosity=0, checkVals=0) as dem:
            z, p = dem.porosity(slices, exact=exact)
        return p, z

    # ------------------------------------------- experiment set-up (seeded)
    # The reference draws from the global, unseeded numpy.random; these take
    # a seed or a numpy Generator so that a set-up is reproducible.

    @staticmethod
    def _rng(seed=None, rng=None) -> np.random.Generator:
        return rng if rng is not None else np.random.default_rng(seed)

    def generateRadii(self, psd: str = "logn", radius_mean: float = 440e-6, radius_variance: float = 8.8e-9,
                      histogram: bool = False, seed=None, rng=None):
        """python/sphere.py:789-822: radii from a log-normal distribution of
        the given mean and variance (mu = ln(m^2 / sqrt(v + m^2)), sigma =
        sqrt(ln(v / m^2 + 1))), or ``psd='uni'``: uniform in [mean -
        variance, mean + variance] (the reference uses the variance as the
        half-width).  The histogram plot is not drawn."""
        g = self._rng(seed, rng)
        if psd == "logn":
            mu = math.log(radius_mean ** 2 / math.sqrt(radius_variance + radius_mean ** 2))
            sigma = math.sqrt(math.log(radius_variance / radius_mean ** 2 + 1.0))
            self.radius = g.lognormal(mu, sigma, self.np)
        elif psd == "uni":
            self.radius = g.uniform(radius_mean - radius_variance, radius_mean + radius_variance, self.np)
        else:
            raise ValueError(f"unknown particle size distribution {psd!r}")

    def initRandomGridPos(self, g=(0.0, 0.0, -9.80665), gridnum=(12, 12, 32), periodic: int = 1,
                          contactmodel: int = 2, seed=None, rng=None):
        """python/sphere.py:1058-1115: every particle in its own cell of a
        coarse grid of gridnum/2 cells of 4.2 r_max, offset at random inside
        the cell without touching its faces; the world is then the smallest
        whole number of such cells holding every sphere.  Radii must be set
        first.  As in the reference, a particle's y cell index wraps with the
        x cell count."""
        gen = self._rng(seed, rng)
        self.g = np.asarray(g, dtype=np.float64).copy()
        self.periodic = int(periodic)
        coarse = np.floor(np.asarray(gridnum, dtype=np.float64) / 2.0)
        r_max = float(np.max(self.radius))
        cellsize = 2.1 * r_max * 2
        if (coarse[0] - 1) * (coarse[1] - 1) * (coarse[2] - 1) < self.np:
            raise ValueError("Error! The grid is not sufficiently large.")
        i = np.arange(self.np, dtype=np.float64)
        gridpos = np.stack([i % coarse[0], np.floor(i / coarse[0]) % coarse[0],
                            np.floor(i / (coarse[0] * coarse[1]))], axis=1)
        r = (self.radius * 1.05)[:, None]
        u = gen.random_sample((self.np, 3)) if hasattr(gen, "random_sample") else gen.random((self.np, 3))
        self.x = gridpos * cellsize + ((cellsize - r) - r) * u + r
        self.contactmodel = int(contactmodel)
        top = np.max(self.x + self.radius[:, None], axis=0)
        self.num = np.ceil(top / cellsize).astype(np.uint32)
        self.L = self.num.astype(np.float64) * cellsize

    def adjustWall(self, idx: int, adjust: float = 1.1):
        """python/sphere.py:1214-1240: grid and wall `idx` fitted to the
        particles along the wall's axis (0: z, 1-2: x, 3-4: y); the wall
        sits at the particles' extreme and weighs rho np pi (cell/2)^3."""
        dim = {0: 2, 1: 0, 2: 0, 3: 1, 4: 1}.get(idx)
        if dim is None:
            raise ValueError("adjustWall: idx value not understood")
        xmin = float(np.min(self.x[:, dim] - self.radius))
        xmax = float(np.max(self.x[:, dim] + self.radius))
        cellsize = float(self.L[0] / self.num[0])
        self.num[dim] = np.uint32(math.ceil(((xmax - xmin) * adjust + xmin) / cellsize))
        self.L[dim] = (xmax - xmin) * adjust + xmin
        self.w_x[idx] = xmax if idx in (0, 1, 3) else xmin
        self.w_m[idx] = self.rho * self.np * math.pi * (cellsize / 2.0) ** 3

    def adjustUpperWall(self, z_adjust: float = 1.1):
        """python/sphere.py:1199-1212: one fixed upper wall (normal -z) at the
        top of the packing, the grid 10 % higher."""
        self.set_walls(1)
        self.wmode = np.zeros(1, dtype=np.int32)
        self.w_n = np.array([[0.0, 0.0, -1.0]])
        self.adjustWall(0, adjust=z_adjust)

    def consolidate(self, deviatoric_stress: float = 10e3, periodic: int = 1):
        """python/sphere.py:1243-1257: zero the kinematics, fit the upper
        wall to the packing and hold it at `deviatoric_stress` (wmode 1).
        (`periodic` is accepted and unused, as in the reference.)"""
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = np.array([1], dtype=np.int32)
        self.w_devs = np.ones(1) * float(deviatoric_stress)

    def uniaxialStrainRate(self, wvel: float = -0.001, periodic: int = 1):
        """python/sphere.py:1259-1271: zero the kinematics, fit the upper
        wall and move it at the fixed velocity `wvel` (wmode 2)."""
        self.zeroKinematics()
        self.adjustUpperWall()
        self.wmode = np.array([2], dtype=np.int32)
        self.w_vel = np.array([float(wvel)])

    def shear(self, shear_strain_rate: float = 1.0, periodic: int = 1):
        """python/sphere.py:1297-1355: one more cell of grid height for
        dilation, zero kinematics, the lowest particles (centres below z_min
        + the largest diameter among those below 0.3 (z_max - z_min)) held
        at zero horizontal velocity and the highest ones (above z_max - the
        largest diameter among those above 0.7 (z_max - z_min)) moved at
        (z_max - z_min) * shear_strain_rate in x, with fixvel = 1; frictionless
        undamped walls.  The 0.3 / 0.7 selections compare absolute heights,
        as the reference does."""
        z = self.x[:, 2]
        z_min = float(np.min(z - self.radius))
        z_max = float(np.max(z + self.radius))
        cellsize = float(self.L[0] / self.num[0])
        self.num[2] += 1
        self.L[2] = self.num[2] * cellsize
        self.zeroKinematics()
        d_bot = float(np.max(self.radius[z < (z_max - z_min) * 0.3])) * 2.0
        low = z < (z_min + d_bot)
        self.fixvel[low] = 1.0
        self.angvel[low] = 0.0
        self.vel[low, 0] = 0.0
        self.vel[low, 1] = 0.0
        d_top = float(np.max(self.radius[z > (z_max - z_min) * 0.7])) * 2.0
        high = z > (z_max - d_top)
        self.fixvel[high] = 1.0
        self.angvel[high] = 0.0
        self.vel[high, 0] = (z_max - z_min) * shear_strain_rate
        self.vel[high, 1] = 0.0
        if self.nw > 0:
            self.gamma_wn = 0.0
            self.gamma_wt = 0.0
            self.mu_ws = 0.0
            self.mu_wd = 0.0

    def run(self, verbose: bool = True, hideinputfile: bool = False, dry: bool = False, root: str = ".") -> int:
        """python/sphere.py:1629-1655: run the `sphere` executable on
        ``<root>/input/<sid>.bin`` from ``root`` (outputs in
        ``<root>/output/``); the input is written first.  Returns the exit
        status, with the reference's warning when it is not 0.  Valgrind,
        cuda-memcheck and the Darcy solver are not part of this path."""
        import subprocess

        exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere")
        os.makedirs(os.path.join(root, "input"), exist_ok=True)
        self.writebin(os.path.join(root, "input", f"{self.sid}.bin"))
        cmd = [exe]
        if not verbose:
            cmd.append("-q")
        if dry:
            cmd.append("--dry")
        cmd.append(os.path.join("input", f"{self.sid}.bin"))
        out = subprocess.DEVNULL if hideinputfile else None
        status = subprocess.call(cmd, cwd=root, stdout=out)
        if status != 0:
            print("Warning: the sphere run returned with status " + str(status))
        return status

    def readlast(self, verbose: bool = True, root: str = "."):
        """python/sphere.py:783-787: read the last output file of this sid,
        the file number taken from ``<root>/output/<sid>.status.dat``."""
        last = status(self.sid, root)
        fn = os.path.join(root, "output", "{0}.output{1:0=5}.bin".format(self.sid, last))
        other = Spherebin.readbin(fn)
        other.sid = self.sid
        self.__dict__.update(other.__dict__)
        if verbose:
            print("Read " + fn)

    def wall_stress(self, idx: int = 0) -> float:
        """Normal stress on a wall, w_force / (L0*L1) (python/sphere.py:2594,2724)."""
        return float(self.w_force[idx] / (self.L[0] * self.L[1]))

    def shear_stress(self) -> float:
        """tau = sum over fixvel>0, vel_x>0 of -force_x, per area.

        python/sphere.py:2718-2721 sums -force_x without dividing by area; the
        division by L0*L1 makes it a stress.
        """
        sel = (self.fixvel > 0.0) & (self.vel[:, 0] > 0.0)
        return float(-np.sum(self.force[sel, 0]) / (self.L[0] * self.L[1]))


def status(project: str, root: str = ".") -> int:
    """python/sphere.py:2892-2907: the last output file number written to
    ``<root>/output/<project>.status.dat`` (its third field)."""
    with open(os.path.join(root, "output", f"{project}.status.dat")) as fh:
        return int(fh.read().split()[2])
