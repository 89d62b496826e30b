"""Committed golden fixtures (tests/golden): the oracle on the CPU and the HIP
path on the GPU must both reproduce them bit for bit.

The fixtures are inputs plus expected arrays after a fixed number of steps,
written by tests/golden/make_golden.py; the 4,096-sphere settling case is
stored as SHA-256 digests of every field."""

from __future__ import annotations

import glob
import hashlib
import json
import os

import numpy as np
import pytest

from parity import first_mismatch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[: -len(".in.bin")] for p in glob.glob(os.path.join(GOLDEN, "*.in.bin")))


def _expected(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.out.npz"))  # plain arrays, allow_pickle stays False
    steps = int(z["steps"][0])
    return steps, {k: z[k] for k in z.files if k != "steps"}


def _compare(sim, want):
    errs = []
    for f, w in want.items():
        m = first_mismatch(f, sim.get(f), w)
        if m:
            errs.append(m)
    return errs


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_fixture_set_is_complete():
    assert len(CASES) >= 6, CASES
    for name in CASES:
        assert os.path.exists(os.path.join(GOLDEN, f"{name}.out.npz")), name
    assert os.path.exists(os.path.join(GOLDEN, "c0_settle.json"))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(oracle_built, name):
    from oracle_lib import Oracle

    steps, want = _expected(name)
    orc = Oracle(os.path.join(GOLDEN, f"{name}.in.bin"))
    assert orc.step(steps) == 0
    errs = _compare(orc, want)
    orc.close()
    assert not errs, errs


def test_c0_input_is_reproducible():
    """The seeded 4,096-sphere settling input still has the digest the fixture
    was made from (the generator did not drift)."""
    import tempfile

    from sphere_amd import configs

    rec = json.load(open(os.path.join(GOLDEN, "c0_settle.json")))
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "c0.bin")
        configs.lattice_settle(sid="c0_settle").writebin(p)
        assert _digest(np.fromfile(p, dtype=np.uint8)) == rec["input_sha256"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_reproduces_fixture(name):
    from sphere_amd.dem import DEM

    print(f"[golden] {name}", flush=True)
    steps, want = _expected(name)
    dem = DEM(os.path.join(GOLDEN, f"{name}.in.bin"), verbosity=0)
    dem.step(steps)
    errs = _compare(dem, want)
    dem.close()
    assert not errs, errs


@pytest.mark.gpu
def test_gpu_reproduces_c0_digests(tmp_path):
    from sphere_amd import configs
    from sphere_amd.dem import DEM

    print("[golden] c0_settle digests", flush=True)
    rec = json.load(open(os.path.join(GOLDEN, "c0_settle.json")))
    p = os.path.join(str(tmp_path), "c0.bin")
    configs.lattice_settle(sid="c0_settle").writebin(p)
    dem = DEM(p, verbosity=0)
    dem.step(int(rec["steps"]))
    bad = [f for f, h in rec["fields"].items() if _digest(dem.get(f)) != h]
    dem.close()
    assert not bad, f"fields differing from the fixture: {bad}"
