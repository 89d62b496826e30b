"""Python mirror of the reference's ``DEM`` class over the gfx950 C-ABI.

``DEM(inputbin, verbosity, checkVals, ...)`` opens a ``.bin`` file onto the
GPU exactly where the reference constructs its ``DEM`` (src/sphere.cpp:16-77)
and exposes the same members (src/sphere.h:248-321): ``readbin``,
``writebin``, ``checkValues``, ``reportValues`` and ``startTime``.  Errors
surface as :class:`SphereError` subclasses carrying the library's message;
the reference exits the process instead (src/utility.cu:9-28).

The shared library must have been built (``python -m sphere_amd.build``);
there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Dict

import numpy as np

from .spherebin import Spherebin

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphere_dem.so")

NC = 32

# status codes (include/sphere_dem.h)
SPH_OK = 0
SPH_E_IO = -1
SPH_E_INVALID = -2
SPH_E_OUT_OF_GRID = -3
SPH_E_CONTACT_OVERFLOW = -4
SPH_E_NAN = -5
SPH_E_HIP = -6
SPH_E_STATE = -7
SPH_E_UNSUPPORTED = -8

PREC_FP64 = 0
PREC_MIXED = 1

FIELDS: Dict[str, int] = {
    "x": 0, "vel": 1, "angvel": 2, "force": 3, "torque": 4, "angpos": 5,
    "xysum": 6, "vel0": 7, "angvel0": 8, "es_dot": 9, "es": 10, "ev_dot": 11,
    "ev": 12, "p": 13, "contacts": 14, "distmod": 15, "delta_t": 16,
    "cell_keys": 17, "sort_idx": 18, "cell_start": 19, "cell_end": 20,
    "bonds": 21, "bonds_delta": 22, "bonds_omega": 23, "walls_nx": 24,
    "walls_mvfd": 25, "walls_vel0": 26, "wall_pp": 27, "time": 28,
    "x_sorted": 29,
}

_U32_FIELDS = {"contacts", "cell_keys", "sort_idx", "cell_start", "cell_end", "bonds"}


class SphereError(RuntimeError):
    code = 0


class SphereIOError(SphereError):
    code = SPH_E_IO


class SphereInvalid(SphereError):
    code = SPH_E_INVALID


class OutOfGrid(SphereError):
    code = SPH_E_OUT_OF_GRID


class ContactOverflow(SphereError):
    code = SPH_E_CONTACT_OVERFLOW


class NonFinite(SphereError):
    code = SPH_E_NAN


class HipError(SphereError):
    code = SPH_E_HIP


class Unsupported(SphereError):
    code = SPH_E_UNSUPPORTED


_ERRORS = {cls.code: cls for cls in (SphereIOError, SphereInvalid, OutOfGrid, ContactOverflow,
                                     NonFinite, HipError, Unsupported)}


class SphInfo(C.Structure):
    _fields_ = [("np", C.c_uint64), ("nb0", C.c_uint64), ("nw", C.c_uint64),
                ("nc", C.c_uint64), ("ncells", C.c_uint64), ("num", C.c_uint32 * 3),
                ("periodic", C.c_int32), ("dt", C.c_double), ("t_current", C.c_double),
                ("t_total", C.c_double), ("file_dt", C.c_double), ("steps_done", C.c_uint64),
                ("precision", C.c_int32)]


class SphParams(C.Structure):
    _fields_ = [("k_n", C.c_double), ("k_t", C.c_double), ("gamma_n", C.c_double),
                ("gamma_t", C.c_double), ("mu_s", C.c_double), ("mu_d", C.c_double)]


class SphTiming(C.Structure):
    _fields_ = [("bonds_ms", C.c_double), ("cellid_ms", C.c_double), ("sort_ms", C.c_double),
                ("reorder_ms", C.c_double), ("contact_ms", C.c_double), ("walls_ms", C.c_double),
                ("total_ms", C.c_double), ("steps", C.c_uint64), ("search_ms", C.c_double),
                ("force_ms", C.c_double), ("tile_x", C.c_uint32), ("tile_y", C.c_uint32),
                ("search_mode", C.c_int32), ("search_auto", C.c_int32), ("force_variant", C.c_int32),
                ("list_builds", C.c_uint32)]


class SphEnsInfo(C.Structure):
    _fields_ = [("members", C.c_uint32), ("np", C.c_uint64), ("ncells", C.c_uint64), ("steps_done", C.c_uint64),
                ("cell_factor", C.c_double), ("tuned_ms", C.c_double * 4), ("search_mode", C.c_int32),
                ("force_block", C.c_uint32), ("search_block", C.c_uint32), ("precision", C.c_int32)]


EXPORTS = {
    "sph_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "sph_open_mem": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "sph_close": (None, [C.c_void_p]),
    "sph_last_error": (C.c_char_p, [C.c_void_p]),
    "sph_step": (C.c_int, [C.c_void_p, C.c_uint64]),
    "sph_run": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
    "sph_writebin": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sph_readbin": (C.c_int, [C.c_void_p, C.c_char_p]),
    "sph_check_values": (C.c_int, [C.c_void_p]),
    "sph_field_bytes": (C.c_size_t, [C.c_void_p, C.c_int]),
    "sph_get": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "sph_set": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "sph_info": (C.c_int, [C.c_void_p, C.POINTER(SphInfo)]),
    "sph_get_params": (C.c_int, [C.c_void_p, C.POINTER(SphParams)]),
    "sph_set_params": (C.c_int, [C.c_void_p, C.POINTER(SphParams)]),
    "sph_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_get_timing": (C.c_int, [C.c_void_p, C.POINTER(SphTiming)]),
    "sph_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_set_tile": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "sph_set_search": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_set_force_kernel": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_bandwidth_probe": (C.c_int, [C.c_uint64, C.c_int, C.POINTER(C.c_double)]),
    "sph_flops_probe": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_double)]),
    "sph_dry_run": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t,
                              C.POINTER(C.c_size_t), C.c_char_p, C.c_size_t]),
    "sph_pack_rsa": (C.c_int, [C.c_uint64, C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_uint64,
                               C.c_void_p]),
    "sph_adjoint_forward": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p, C.POINTER(C.c_double)]),
    "sph_adjoint_backward": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "sph_porosity": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "sph_porosity_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "sph_find_overlaps": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "sph_render": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_uint32, C.c_uint32, C.c_int,
                             C.c_void_p]),
    "sph_ens_open_mem": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_uint32, C.c_double, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    "sph_ens_open": (C.c_int, [C.POINTER(C.c_char_p), C.c_uint32, C.c_double, C.c_int, C.POINTER(C.c_void_p)]),
    "sph_ens_close": (None, [C.c_void_p]),
    "sph_ens_last_error": (C.c_char_p, [C.c_void_p]),
    "sph_ens_step": (C.c_int, [C.c_void_p, C.c_uint64]),
    "sph_ens_field_bytes": (C.c_size_t, [C.c_void_p, C.c_uint32, C.c_int]),
    "sph_ens_get": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t]),
    "sph_ens_member_info": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(SphInfo)]),
    "sph_ens_info": (C.c_int, [C.c_void_p, C.POINTER(SphEnsInfo)]),
    "sph_ens_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_ens_set_search": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_ens_set_blocks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "sph_ens_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_ens_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "sph_ens_get_timing": (C.c_int, [C.c_void_p, C.POINTER(SphTiming)]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing; build it with `python -m sphere_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _raise(rc: int, msg: str):
    cls = _ERRORS.get(rc, SphereError)
    err = cls(f"[{rc}] {msg}")
    err.code = rc
    raise err


class DEM:
    """One simulation on the GPU (src/sphere.h:248-321)."""

    def __init__(self, inputbin, verbosity: int = 1, checkVals: int = 1, dry: int = 0,
                 initCuda: int = 1, transferConstMem: int = 1, darcyflow: int = 0):
        self._lib = load_library()
        self._h = C.c_void_p()
        if darcyflow:
            raise Unsupported("Darcy flow is outside this path")
        if isinstance(inputbin, Spherebin):
            self.sid = inputbin.sid
            data = inputbin.to_bytes()
            buf = C.create_string_buffer(data, len(data))
            rc = self._lib.sph_open_mem(buf, len(data), int(verbosity), int(checkVals), C.byref(self._h))
        else:
            path = os.fspath(inputbin)
            self.sid = os.path.basename(path).rsplit(".", 1)[0]
            rc = self._lib.sph_open(path.encode(), int(verbosity), int(checkVals), C.byref(self._h))
        if rc != SPH_OK:
            _raise(rc, self._lib.sph_last_error(None).decode())
        if dry:
            self.reportValues()

    # -------------------------------------------------------------- plumbing
    def _check(self, rc: int):
        if rc != SPH_OK:
            _raise(rc, self._lib.sph_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.sph_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --------------------------------------------------- reference members
    def readbin(self, target: str):
        """Replaces the simulation with the contents of ``target`` (settings
        kept; on a read or check error the state is unchanged)."""
        self._check(self._lib.sph_readbin(self._h, os.fspath(target).encode()))
        self.sid = os.path.basename(os.fspath(target)).rsplit(".", 1)[0]

    def writebin(self, target: str):
        self._check(self._lib.sph_writebin(self._h, os.fspath(target).encode()))

    def checkValues(self):
        self._check(self._lib.sph_check_values(self._h))

    def reportValues(self):
        i = self.info()
        print(f"  np={i.np} nb0={i.nb0} nw={i.nw} grid={tuple(i.num)} periodic={i.periodic}\n"
              f"  dt={i.dt:g} s, current={i.t_current:g} s, total={i.t_total:g} s, file_dt={i.file_dt:g} s")

    def startTime(self, outdir: str = "output"):
        os.makedirs(outdir, exist_ok=True)
        self._check(self._lib.sph_run(self._h, os.fspath(outdir).encode(), self.sid.encode()))

    # ------------------------------------------------------------- extras
    def step(self, n: int = 1):
        self._check(self._lib.sph_step(self._h, int(n)))

    def info(self) -> SphInfo:
        out = SphInfo()
        self._check(self._lib.sph_info(self._h, C.byref(out)))
        return out

    @property
    def np(self) -> int:
        return int(self.info().np)

    def field_shape(self, name: str):
        i = self.info()
        n, nb, nw = int(i.np), int(i.nb0), int(i.nw)
        shapes = {
            "x": (n, 4), "vel": (n, 4), "x_sorted": (n, 4), "angvel": (n, 3), "force": (n, 3),
            "torque": (n, 3), "angpos": (n, 3), "vel0": (n, 3), "angvel0": (n, 3), "xysum": (n, 2),
            "es_dot": (n,), "es": (n,), "ev_dot": (n,), "ev": (n,), "p": (n,),
            "contacts": (n, NC), "distmod": (n, NC, 4), "delta_t": (n, NC, 3),
            "cell_keys": (n,), "sort_idx": (n,), "cell_start": (int(i.ncells),),
            "cell_end": (int(i.ncells),), "bonds": (nb, 2), "bonds_delta": (nb, 4),
            "bonds_omega": (nb, 4), "walls_nx": (nw, 4), "walls_mvfd": (nw, 4),
            "walls_vel0": (nw,), "wall_pp": (nw, n), "time": (4,),
        }
        return shapes[name]

    def get(self, name: str) -> np.ndarray:
        fid = FIELDS[name]
        dtype = np.uint32 if name in _U32_FIELDS else np.float64
        out = np.empty(self.field_shape(name), dtype=dtype)
        nbytes = self._lib.sph_field_bytes(self._h, fid)
        if nbytes != out.nbytes:
            raise SphereError(f"size mismatch for {name}: {nbytes} vs {out.nbytes}")
        self._check(self._lib.sph_get(self._h, fid, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def set(self, name: str, value):
        fid = FIELDS[name]
        dtype = np.uint32 if name in _U32_FIELDS else np.float64
        arr = np.ascontiguousarray(np.asarray(value, dtype=dtype).reshape(self.field_shape(name)))
        self._check(self._lib.sph_set(self._h, fid, arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def get_params(self) -> Dict[str, float]:
        p = SphParams()
        self._check(self._lib.sph_get_params(self._h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in SphParams._fields_}

    def set_params(self, **kw):
        cur = self.get_params()
        cur.update({k: float(v) for k, v in kw.items()})
        p = SphParams(**cur)
        self._check(self._lib.sph_set_params(self._h, C.byref(p)))

    def adjoint_forward(self, nsteps: int, weights, seg_len: int = 100) -> float:
        """Runs `nsteps` steps keeping checkpoints every `seg_len` steps and
        returns J = mean over the steps of sum_i weights[i] . F_i."""
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(self.np, 3))
        J = C.c_double()
        self._check(self._lib.sph_adjoint_forward(self._h, int(nsteps), int(seg_len),
                                                  w.ctypes.data_as(C.c_void_p), C.byref(J)))
        return J.value

    def adjoint_backward(self) -> np.ndarray:
        """dJ/d(k_n, k_t, gamma_n, mu) of the last adjoint_forward run."""
        g = (C.c_double * 4)()
        self._check(self._lib.sph_adjoint_backward(self._h, g))
        return np.array(list(g), dtype=np.float64)

    def porosity(self, slices: int = 10, exact: bool = False):
        """DEM::porosity (src/sphere.cpp:363-483): (z_pos, porosity) of
        `slices` horizontal slices; exact=True removes the exact volume of
        each sphere inside a slice instead of the reference's rule."""
        z = np.empty(max(int(slices), 0))
        p = np.empty_like(z)
        self._check(self._lib.sph_porosity(self._h, int(slices), 1 if exact else 0,
                                           z.ctypes.data_as(C.c_void_p), p.ctypes.data_as(C.c_void_p)))
        return z, p

    def findOverlaps(self):
        """DEM::findOverlaps (src/sphere.cpp:553-604): (ij, delta_n), the
        pairs i < j of overlapping spheres (no periodic images), i major."""
        n = C.c_uint64()
        self._check(self._lib.sph_find_overlaps(self._h, None, None, 0, C.byref(n)))
        ij = np.empty((n.value, 2), dtype=np.uint32)
        dn = np.empty(n.value)
        if n.value:
            self._check(self._lib.sph_find_overlaps(self._h, ij.ctypes.data_as(C.c_void_p),
                                                    dn.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
        return ij, dn

    def render(self, method: int = 0, maxval: float = 1.0e3, lower_cutoff: float = 0.0, width: int = 800,
               height: int = 800, cull: bool = True) -> np.ndarray:
        """DEM::render (src/raytracer.cuh:374-553): the image as a
        (height, width, 3) uint8 array in the reference's pixel order.
        method 0 normal, 1 pressure, 2 velocity, 3 angular velocity,
        4 x displacement, 5 angular position."""
        img = np.empty((int(height), int(width), 3), dtype=np.uint8)
        self._check(self._lib.sph_render(self._h, int(method), float(maxval), float(lower_cutoff), int(width),
                                         int(height), 1 if cull else 0, img.ctypes.data_as(C.c_void_p)))
        return img

    def set_precision(self, prec):
        if isinstance(prec, str):
            prec = {"fp64": PREC_FP64, "mixed": PREC_MIXED}[prec]
        self._check(self._lib.sph_set_precision(self._h, int(prec)))

    def set_profiling(self, on: bool = True):
        self._check(self._lib.sph_set_profiling(self._h, 1 if on else 0))

    def set_graph(self, on: bool = True):
        self._check(self._lib.sph_set_graph(self._h, 1 if on else 0))

    def set_tile(self, tile_x: int = 0, tile_y: int = 0):
        self._check(self._lib.sph_set_tile(self._h, int(tile_x), int(tile_y)))

    def set_search(self, mode: int = 0):
        self._check(self._lib.sph_set_search(self._h, int(mode)))

    def set_force_kernel(self, variant: int = 0):
        self._check(self._lib.sph_set_force_kernel(self._h, int(variant)))

    def timing(self) -> Dict[str, float]:
        t = SphTiming()
        self._check(self._lib.sph_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in SphTiming._fields_}

    def to_spherebin(self) -> Spherebin:
        """Snapshot of the current state (through writebin's layout)."""
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, f"{self.sid}.bin")
            self.writebin(p)
            sb = Spherebin.readbin(p)
        sb.sid = self.sid
        return sb


class Ensemble:
    """Many independent systems stepped together (sph_ens_*; no counterpart
    in the reference, which runs one DEM per process).  ``members`` are
    :class:`Spherebin` objects or ``.bin`` paths; ``cell_factor`` < 0 keeps
    their grids, >= 1 regrids every member to cells of cell_factor x 2 r_max,
    0 lets the autotuner choose.  Member k evolves as ``DEM(members[k])``
    would on the same grid."""

    def __init__(self, members, cell_factor: float = -1.0, verbosity: int = 0):
        self._lib = load_library()
        self._h = C.c_void_p()
        n = len(members)
        if all(isinstance(m, Spherebin) for m in members):
            data = [m.to_bytes() for m in members]
            self._bufs = [C.create_string_buffer(b, len(b)) for b in data]
            ptrs = (C.c_void_p * n)(*[C.cast(b, C.c_void_p) for b in self._bufs])
            sizes = (C.c_size_t * n)(*[len(b) for b in data])
            rc = self._lib.sph_ens_open_mem(ptrs, sizes, n, float(cell_factor), int(verbosity), C.byref(self._h))
            self._bufs = None
        else:
            paths = (C.c_char_p * n)(*[os.fspath(m).encode() for m in members])
            rc = self._lib.sph_ens_open(paths, n, float(cell_factor), int(verbosity), C.byref(self._h))
        if rc != SPH_OK:
            _raise(rc, self._lib.sph_ens_last_error(None).decode())
        self.members = n

    def _check(self, rc: int):
        if rc != SPH_OK:
            _raise(rc, self._lib.sph_ens_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.sph_ens_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1):
        self._check(self._lib.sph_ens_step(self._h, int(n)))

    def info(self) -> SphEnsInfo:
        out = SphEnsInfo()
        self._check(self._lib.sph_ens_info(self._h, C.byref(out)))
        return out

    def member_info(self, k: int) -> SphInfo:
        out = SphInfo()
        self._check(self._lib.sph_ens_member_info(self._h, int(k), C.byref(out)))
        return out

    def get(self, k: int, name: str) -> np.ndarray:
        i = self.member_info(k)
        n, nw = int(i.np), int(i.nw)
        shapes = {
            "x": (n, 4), "vel": (n, 4), "angvel": (n, 3), "force": (n, 3), "torque": (n, 3),
            "angpos": (n, 3), "vel0": (n, 3), "angvel0": (n, 3), "xysum": (n, 2), "es_dot": (n,),
            "es": (n,), "ev_dot": (n,), "ev": (n,), "p": (n,), "contacts": (n, NC), "delta_t": (n, NC, 3),
            "walls_nx": (nw, 4), "walls_mvfd": (nw, 4), "walls_vel0": (nw,), "wall_pp": (nw, n), "time": (4,),
        }
        if name not in shapes:
            raise Unsupported(f"field {name} is not available per ensemble member")
        out = np.empty(shapes[name], dtype=np.uint32 if name in _U32_FIELDS else np.float64)
        self._check(self._lib.sph_ens_get(self._h, int(k), FIELDS[name], out.ctypes.data_as(C.c_void_p),
                                          out.nbytes))
        return out

    def set_precision(self, prec):
        if isinstance(prec, str):
            prec = {"fp64": PREC_FP64, "mixed": PREC_MIXED}[prec]
        self._check(self._lib.sph_ens_set_precision(self._h, int(prec)))

    def set_search(self, mode: int = 0):
        self._check(self._lib.sph_ens_set_search(self._h, int(mode)))

    def set_blocks(self, force_block: int, search_block: int):
        self._check(self._lib.sph_ens_set_blocks(self._h, int(force_block), int(search_block)))

    def set_graph(self, on: bool = True):
        self._check(self._lib.sph_ens_set_graph(self._h, 1 if on else 0))

    def set_profiling(self, on: bool = True):
        self._check(self._lib.sph_ens_set_profiling(self._h, 1 if on else 0))

    def timing(self) -> Dict[str, float]:
        t = SphTiming()
        self._check(self._lib.sph_ens_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in SphTiming._fields_}


def pack_rsa(radii, box, periodic: int, seed: int, max_tries: int = 100000) -> np.ndarray:
    """Seeded random sequential packing (host code in the same library)."""
    lib = load_library()
    r = np.ascontiguousarray(radii, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64)
    out = np.zeros((r.size, 3))
    rc = lib.sph_pack_rsa(r.size, r.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p),
                          int(periodic), int(seed) & 0xFFFFFFFFFFFFFFFF, int(max_tries),
                          out.ctypes.data_as(C.c_void_p))
    if rc != 0:
        raise SphereError("random packing failed: box too small for the radii")
    return out


def dry_run(data: bytes, check_vals: int = 1, roundtrip: bool = False):
    """Parses (and checks) a .bin image on the host; returns the re-serialised
    bytes when ``roundtrip`` is set (no GPU involved)."""
    lib = load_library()
    src = C.create_string_buffer(bytes(data), len(data))
    err = C.create_string_buffer(1024)
    n = C.c_size_t(0)
    rc = lib.sph_dry_run(src, len(data), int(check_vals), None, 0, C.byref(n) if roundtrip else None, err, 1024)
    if rc != SPH_OK:
        _raise(rc, err.value.decode())
    if not roundtrip:
        return None
    out = C.create_string_buffer(n.value)
    rc = lib.sph_dry_run(src, len(data), int(check_vals), out, n.value, C.byref(n), err, 1024)
    if rc != SPH_OK:
        _raise(rc, err.value.decode())
    return out.raw[: n.value]


def bandwidth_probe(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Read+write GB/s of a plain device-to-device copy (k_copy_probe)."""
    lib = load_library()
    out = C.c_double(0.0)
    rc = lib.sph_bandwidth_probe(int(nbytes), int(reps), C.byref(out))
    if rc != SPH_OK:
        _raise(rc, "bandwidth probe failed")
    return out.value


def flops_probe(fp64: bool = True, blocks: int = 4096, iters: int = 4096, reps: int = 10) -> float:
    """TFLOP/s of k_flops_probe: 16 x iters fused-multiply-add FLOPs per
    thread, 256 threads per block."""
    lib = load_library()
    out = C.c_double(0.0)
    rc = lib.sph_flops_probe(1 if fp64 else 0, int(blocks), int(iters), int(reps), C.byref(out))
    if rc != SPH_OK:
        _raise(rc, "flops probe failed")
    return out.value
