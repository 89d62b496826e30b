"""ctypes access to the CPU oracle (oracle/dem_oracle.h) for the tests.

Test infrastructure: the product never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "build", "liboracle_dem.so")
# a planted-mistake build of the oracle, for the pin tests' own check
MUTANT_LIB = os.environ.get("SPH_ORACLE_MUTANT_LIB")

FIELDS = {
    "x": 0, "vel": 1, "angvel": 2, "force": 3, "torque": 4, "angpos": 5, "xysum": 6,
    "vel0": 7, "angvel0": 8, "es_dot": 9, "es": 10, "ev_dot": 11, "ev": 12, "p": 13,
    "contacts": 14, "distmod": 15, "delta_t": 16, "cell_keys": 17, "sort_idx": 18,
    "cell_start": 19, "cell_end": 20, "bonds": 21, "bonds_delta": 22, "bonds_omega": 23,
    "walls_nx": 24, "walls_mvfd": 25, "walls_vel0": 26, "wall_pp": 27, "time": 28,
    "x_sorted": 29,
}
U32 = {"contacts", "cell_keys", "sort_idx", "cell_start", "cell_end", "bonds"}
STAGES = ["bonds", "cellid", "sort", "reorder", "topology", "interact", "bond_apply",
          "integrate", "summation", "walls", "time"]
F_OUT_OF_GRID, F_CONTACT_OVERFLOW, F_NAN = 1, 2, 4

_lib = None


def ensure_built():
    srcs = [os.path.join(ORACLE_DIR, f) for f in ("dem_oracle.cpp", "dem_oracle.h")]
    if (not os.path.exists(LIB)) or any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in srcs):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if MUTANT_LIB:
            L = C.CDLL(MUTANT_LIB)
        else:
            ensure_built()
            L = C.CDLL(LIB)
        L.orc_open.restype = C.c_void_p
        L.orc_open.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_size_t]
        L.orc_close.argtypes = [C.c_void_p]
        L.orc_writebin.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int]
        L.orc_field_bytes.restype = C.c_size_t
        L.orc_field_bytes.argtypes = [C.c_void_p, C.c_int]
        L.orc_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        L.orc_set.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        L.orc_dims.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        L.orc_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_exp.restype = C.c_double
        L.orc_exp.argtypes = [C.c_double]
        L.orc_porosity.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_find_overlaps.restype = C.c_uint64
        L.orc_find_overlaps.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_render.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_void_p]
        _lib = L
    return _lib


class Oracle:
    def __init__(self, path: str, nc: int = 32):
        self.L = lib()
        err = C.create_string_buffer(512)
        self.h = self.L.orc_open(os.fspath(path).encode(), nc, err, 512)
        if not self.h:
            raise RuntimeError(f"oracle open failed: {err.value.decode()}")
        d = (C.c_uint64 * 5)()
        self.L.orc_dims(self.h, d)
        self.np, self.nb0, self.nw, self.nc, self.K = (int(v) for v in d)

    @classmethod
    def from_spherebin(cls, sb, tmpdir, nc: int = 32):
        p = os.path.join(str(tmpdir), f"{sb.sid}.oracle-in.bin")
        sb.writebin(p)
        return cls(p, nc)

    def close(self):
        if self.h:
            self.L.orc_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1) -> int:
        return int(self.L.orc_step(self.h, n))

    def stage(self, name: str) -> int:
        return int(self.L.orc_stage(self.h, STAGES.index(name)))

    def shape(self, name):
        n, nc, nb, nw, K = self.np, self.nc, self.nb0, self.nw, self.K
        return {
            "x": (n, 4), "vel": (n, 4), "x_sorted": (n, 4), "angvel": (n, 3), "force": (n, 3),
            "torque": (n, 3), "angpos": (n, 3), "xysum": (n, 2), "vel0": (n, 3), "angvel0": (n, 3),
            "es_dot": (n,), "es": (n,), "ev_dot": (n,), "ev": (n,), "p": (n,),
            "contacts": (n, nc), "distmod": (n, nc, 4), "delta_t": (n, nc, 3), "cell_keys": (n,),
            "sort_idx": (n,), "cell_start": (K,), "cell_end": (K,), "bonds": (nb, 2),
            "bonds_delta": (nb, 4), "bonds_omega": (nb, 4), "walls_nx": (nw, 4),
            "walls_mvfd": (nw, 4), "walls_vel0": (nw,), "wall_pp": (nw, n), "time": (4,),
        }[name]

    def get(self, name: str) -> np.ndarray:
        dt = np.uint32 if name in U32 else np.float64
        out = np.empty(self.shape(name), dtype=dt)
        rc = self.L.orc_get(self.h, FIELDS[name], out.ctypes.data_as(C.c_void_p), out.nbytes)
        if rc != 0:
            raise RuntimeError(f"orc_get({name}) -> {rc}")
        return out

    def set(self, name: str, value):
        dt = np.uint32 if name in U32 else np.float64
        arr = np.ascontiguousarray(np.asarray(value, dtype=dt).reshape(self.shape(name)))
        rc = self.L.orc_set(self.h, FIELDS[name], arr.ctypes.data_as(C.c_void_p), arr.nbytes)
        if rc != 0:
            raise RuntimeError(f"orc_set({name}) -> {rc}")

    def params(self):
        p = (C.c_double * 6)()
        self.L.orc_get_params(self.h, p)
        return list(p)

    def set_params(self, vals):
        p = (C.c_double * 6)(*vals)
        self.L.orc_set_params(self.h, p)

    def porosity(self, slices: int, rule: int = 0):
        """(z_pos, porosity), or the oracle's status code (1: slices too
        thin for the reference's rule, 2: slices < 1) as an int."""
        z = np.empty(max(slices, 0))
        p = np.empty(max(slices, 0))
        rc = self.L.orc_porosity(self.h, slices, rule, z.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p))
        return rc if rc else (z, p)

    def find_overlaps(self):
        n = int(self.L.orc_find_overlaps(self.h, None, None, 0))
        ij = np.empty((n, 2), dtype=np.uint32)
        dn = np.empty(n)
        got = int(self.L.orc_find_overlaps(self.h, ij.ctypes.data_as(C.c_void_p), dn.ctypes.data_as(C.c_void_p), n))
        assert got == n
        return ij, dn

    def render(self, method: int = 0, max_val: float = 1e3, lower_cutoff: float = 0.0, width: int = 800,
               height: int = 800) -> np.ndarray:
        img = np.empty((height, width, 3), dtype=np.uint8)
        rc = self.L.orc_render(self.h, method, max_val, lower_cutoff, width, height, img.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError(f"orc_render -> {rc}")
        return img

    def writebin(self, path):
        rc = self.L.orc_writebin(self.h, os.fspath(path).encode())
        if rc != 0:
            raise RuntimeError("orc_writebin failed")
