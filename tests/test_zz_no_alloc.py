"""sph_porosity_noalloc (device work space in a static per-GPU array, no
allocation in the call) against sph_porosity (work space allocated per
call), through the C-ABI: the same bits on two calls in a row, the second
on a smaller input, and requests beyond the array's size refused."""

import ctypes as C

import numpy as np
import pytest

from sphere_amd import configs
from sphere_amd.dem import DEM

pytestmark = pytest.mark.gpu

SPH_E_INVALID = -2
NOALLOC_MAX = 1048576  # SPH_POROSITY_NOALLOC_MAX, doubles


def _call(dem, fn, slices, rule):
    z = np.full(slices, np.nan)
    p = np.full(slices, np.nan)
    rc = getattr(dem._lib, fn)(dem._h, int(slices), int(rule), z.ctypes.data_as(C.c_void_p),
                               p.ctypes.data_as(C.c_void_p))
    return rc, z, p


def _same(dem, slices, rule):
    """Both calls on the same handle: the same status and message, and on
    success the same bits."""
    rc0, z0, p0 = _call(dem, "sph_porosity", slices, rule)
    msg0 = dem._lib.sph_last_error(dem._h) if rc0 else b""
    rc1, z1, p1 = _call(dem, "sph_porosity_noalloc", slices, rule)
    msg1 = dem._lib.sph_last_error(dem._h) if rc1 else b""
    assert (rc0, msg0) == (rc1, msg1)
    if rc0 == 0:
        assert np.array_equal(z0.view(np.uint64), z1.view(np.uint64))
        assert np.array_equal(p0.view(np.uint64), p1.view(np.uint64))
    return rc1, p1


def test_noalloc_porosity_matches_the_allocating_call():
    print("[no-alloc] porosity without allocation vs sph_porosity", flush=True)
    big = DEM(configs.dense_lattice(cells=(24, 24, 24), sid="noalloc-big"), verbosity=0)
    big.step(5)
    # first call: a large input (55,296 spheres, 400 slices, exact rule)
    rc, p_big = _same(big, 400, 1)
    assert rc == 0 and np.all((p_big > 0.0) & (p_big < 1.0))
    small = DEM(configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, sid="noalloc-small"),
                verbosity=0)
    small.step(3)
    # second call: a smaller input and fewer slices, on the array the first
    # call left filled; the reference rule's slice limit refuses 7 slices of
    # this 0.1075 m box alike in both calls
    assert _same(small, 5, 0)[0] == 0
    assert _same(small, 3, 1)[0] == 0
    assert _same(small, 7, 0)[0] == -2
    # and the large one again after the small one
    rc, p_again = _same(big, 400, 1)
    assert rc == 0 and np.array_equal(p_again, p_big)
    big.close()
    small.close()


def test_noalloc_porosity_refuses_requests_beyond_its_array():
    print("[no-alloc] size limit", flush=True)
    dem = DEM(configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, sid="noalloc-limit"),
              verbosity=0)
    nblk = -(-dem.np // 2048)
    too_many = (NOALLOC_MAX - 1793) // (nblk + 2) + 1
    rc, _, _ = _call(dem, "sph_porosity_noalloc", too_many, 1)
    assert rc == SPH_E_INVALID
    assert b"above its maximum" in dem._lib.sph_last_error(dem._h)
    # the largest request that fits still runs, and matches the allocating call
    assert _same(dem, too_many - 1, 1)[0] == 0
    dem.close()
