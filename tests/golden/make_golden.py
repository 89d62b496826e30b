"""Writes the golden fixtures of this directory from the CPU oracle.

Each scenario is an input `.bin` (`<name>.in.bin`) plus the expected arrays
after a fixed number of steps (`<name>.out.npz`, plain arrays, no pickles).
The 4,096-sphere settling case keeps its input reproducible from the seeded
generator and stores SHA-256 digests of the input and of every output field
(`c0_settle.json`) instead of the arrays.

    python tests/golden/make_golden.py

The fixtures are data: the tests compare the oracle (CPU) and the HIP path
(GPU) against them bit for bit.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

# fields stored per scenario (the parity set that does not depend on layout)
FIELDS = ["x", "vel", "angvel", "force", "torque", "angpos", "xysum", "vel0", "angvel0",
          "es_dot", "es", "ev_dot", "ev", "p", "contacts", "delta_t", "cell_keys", "sort_idx",
          "cell_start", "cell_end", "time"]
BOND_FIELDS = ["bonds", "bonds_delta", "bonds_omega"]
WALL_FIELDS = ["walls_nx", "walls_mvfd"]


def scenarios():
    from sphere_amd import configs

    out = []
    sb = configs.two_spheres(distance=-2e-2, radius=0.5, v=(-0.01, 0.002, 0.0), w=(0.0, 0.0, 3.0),
                             gamma=2.0, k=1e5, sid="pair_oblique")
    out.append(("pair_oblique", sb, 12))
    for periodic in (0, 1, 2):
        sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=periodic,
                                    gamma_n=0.0, gamma_t=10.0, sid=f"lattice_p{periodic}")
        out.append((f"lattice_p{periodic}", sb, 300))
    sb = configs.lattice_settle(nside=(4, 4, 6), spacing=2.15e-2, jitter=5e-4, periodic=1, nw=1,
                                gamma_n=0.0, gamma_t=10.0, sid="lattice_wall")
    sb.wmode[0] = 1
    sb.w_devs[0] = 50.0
    sb.w_x[0] = float(np.max(sb.x[:, 2] + sb.radius))
    out.append(("lattice_wall", sb, 300))
    sb = configs.fcc_bonded(cells=(2, 2, 3), sigma_b=5.8e3, tau_b=1e9, pull_velocity=0.5, sid="fcc_break")
    out.append(("fcc_break", sb, 300))
    return out


def fields_for(orc):
    f = list(FIELDS)
    if orc.nb0:
        f += BOND_FIELDS
    if orc.nw:
        f += WALL_FIELDS
    return f


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def c0_input():
    from sphere_amd import configs

    return configs.lattice_settle(sid="c0_settle")


def main():
    from oracle_lib import Oracle

    tmp = tempfile.mkdtemp(prefix="golden-")
    for name, sb, steps in scenarios():
        path = os.path.join(HERE, f"{name}.in.bin")
        sb.writebin(path)
        orc = Oracle(path)
        flags = orc.step(steps)
        if flags:
            raise SystemExit(f"{name}: oracle flags {flags}")
        arrays = {f: orc.get(f) for f in fields_for(orc)}
        arrays["steps"] = np.array([steps], dtype=np.uint64)
        orc.close()
        np.savez_compressed(os.path.join(HERE, f"{name}.out.npz"), **arrays)
        nc = int((arrays["contacts"] != sb.np).sum())
        print(f"{name}: np={sb.np} steps={steps} active slots={nc}")

    sb = c0_input()
    path = os.path.join(tmp, "c0.bin")
    sb.writebin(path)
    steps = 200
    orc = Oracle(path)
    flags = orc.step(steps)
    if flags:
        raise SystemExit(f"c0: oracle flags {flags}")
    rec = {"steps": steps, "input_sha256": digest(np.fromfile(path, dtype=np.uint8)),
           "fields": {f: digest(orc.get(f)) for f in FIELDS}}
    orc.close()
    with open(os.path.join(HERE, "c0_settle.json"), "w") as fh:
        json.dump(rec, fh, indent=1, sort_keys=True)
    print("c0_settle: digests written")


if __name__ == "__main__":
    main()
