"""Helpers that run one input on the GPU (through the C-ABI) and on the CPU
oracle and compare every array the step produces."""

from __future__ import annotations

import numpy as np

# fields compared after a run; all must be bit-identical in FP64 mode
STATE_FIELDS = ["x", "vel", "angvel", "force", "torque", "angpos", "xysum", "vel0", "angvel0",
                "es_dot", "es", "ev_dot", "ev", "p"]
SORT_FIELDS = ["cell_keys", "sort_idx", "cell_start", "cell_end", "x_sorted"]
CONTACT_FIELDS = ["contacts", "delta_t"]
BOND_FIELDS = ["bonds", "bonds_delta", "bonds_omega"]
WALL_FIELDS = ["walls_nx", "walls_mvfd", "wall_pp"]


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.view(np.uint64)
    return a


def first_mismatch(name, got, want):
    g, w = bits(got), bits(want)
    if g.shape != w.shape:
        return f"{name}: shape {g.shape} vs {w.shape}"
    bad = np.argwhere(g != w)
    if bad.size == 0:
        return None
    i = tuple(bad[0])
    return (f"{name}: {len(bad)} entries differ; first at {i}: gpu={got[i]!r} oracle={want[i]!r}")


def compare_all(dem, orc, fields, extra_masks=None):
    errors = []
    for f in fields:
        g = dem.get(f)
        w = orc.get(f)
        if f == "distmod":
            occ = orc.get("contacts") != orc.np
            g = g[occ]
            w = w[occ]
        m = first_mismatch(f, g, w)
        if m:
            errors.append(m)
    return errors


def all_fields(orc):
    fields = STATE_FIELDS + SORT_FIELDS + CONTACT_FIELDS + ["distmod"]
    if orc.nb0:
        fields += BOND_FIELDS
    if orc.nw:
        fields += WALL_FIELDS
    return fields


def run_pair(sb, tmpdir, steps_list, precision="fp64", setup=None):
    """Steps GPU and oracle side by side; returns a list of mismatch messages.
    `setup(dem)` may choose kernel variants before the first step."""
    from oracle_lib import Oracle
    from sphere_amd.dem import DEM

    orc = Oracle.from_spherebin(sb, tmpdir)
    dem = DEM(sb, verbosity=0)
    if setup is not None:
        setup(dem)
    if precision != "fp64":
        dem.set_precision(precision)
    errors = []
    done = 0
    for k in steps_list:
        dem.step(k)
        flags = orc.step(k)
        done += k
        assert flags == 0, f"oracle flags {flags}"
        errs = compare_all(dem, orc, all_fields(orc))
        if errs:
            errors += [f"after {done} steps: {e}" for e in errs]
            break
    dem.close()
    orc.close()
    return errors


# every field that carries state from one step to the next (per-step
# outputs and sort fields are recomputed by the next step)
HANDOVER_FIELDS = ["x", "vel", "angvel", "force", "torque", "angpos", "xysum", "vel0", "angvel0", "es_dot", "es",
                   "ev_dot", "ev", "p", "contacts", "distmod", "delta_t"]


def hand_over(dem, orc):
    """Copies the GPU handle's whole state into the oracle (which must have
    been opened on the same input), so that the oracle continues the GPU's
    trajectory from this step on."""
    fields = list(HANDOVER_FIELDS)
    if orc.nb0:
        fields += BOND_FIELDS
    if orc.nw:
        fields += ["walls_nx", "walls_mvfd", "walls_vel0"]
    for f in fields:
        orc.set(f, dem.get(f))
    orc.set("time", dem.get("time"))
