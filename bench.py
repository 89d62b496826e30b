"""Benchmark of the gfx950 DEM time step.

Default workload: 1,048,576 spheres, lognormal radii (median 1e-2 m, sigma
0.2, clipped to [0.5, 2] x median), seeded random non-overlapping packing in
a 2.07 x 2.07 x 4.13 m x/y-periodic box (the nominal 0.8 x 0.8 x 1.6 m box
cannot hold that much solid), gravity, top wall at 10 kPa, the contact
parameters k_n = k_t = 1e6 N/m, gamma_n = gamma_t = 100 N s/m, mu = 0.5.

A step is one full DEM time step (bonds, cell ids, sort, reorder,
contact search + forces + integration, wall reduction + wall update) with
all state resident in HBM.  Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def contact_bytes(n, c, w):
    """Algorithmic HBM bytes of one contact-kernel launch (topology + interact +
    integrate fused): 32N + 32C (topology), 196N + 76C + 8WN (interact),
    388N (integrate)."""
    return 616 * n + 108 * c + 8 * w * n


def step_bytes(n, c, b, w, k):
    """Algorithmic bytes of one FP64 step: 848N + 16WN + 108C + 136B + 8K."""
    return 848 * n + 16 * w * n + 108 * c + 136 * b + 8 * k


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(bin_path, n, sample_steps):
    """Serial CPU oracle (one core) on the first `sample_steps` steps of the
    same input; returns particle-steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    t0 = time.perf_counter()
    orc = Oracle(bin_path)
    log(f"oracle loaded in {time.perf_counter() - t0:.1f}s")
    orc.step(1)  # first touch of the slot arrays, untimed
    t0 = time.perf_counter()
    flags = orc.step(sample_steps)
    dt = time.perf_counter() - t0
    orc.close()
    if flags:
        log(f"oracle flags {flags}")
    return n * sample_steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU scaling is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from sphere_amd import configs
    from sphere_amd.dem import DEM

    t0 = time.perf_counter()
    sb = configs.random_compression(n=args.n)
    log(f"generated {args.n} spheres in {time.perf_counter() - t0:.1f}s, grid {list(sb.num)}")
    tmp = tempfile.mkdtemp(prefix="sphere-bench-")
    bin_path = os.path.join(tmp, "c1.bin")
    sb.writebin(bin_path)
    n = sb.np
    K = int(sb.num[0]) * int(sb.num[1]) * int(sb.num[2])

    def timed_run(precision):
        dem = DEM(bin_path, verbosity=0)
        dem.set_precision(precision)
        dem.step(args.warmup)
        import ctypes as C
        from sphere_amd.dem import load_library
        lib = load_library()
        # barrier + sync on both sides of the timed region: sph_step returns
        # only after the device finished (it reads back the error flags)
        t0 = time.perf_counter()
        dem.step(args.steps)
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        contacts = dem.get("contacts")
        c = int((contacts != n).sum())
        # per-kernel hipEvent timing on the handle's stream over separate steps
        dem.set_profiling(True)
        dem.step(min(args.steps, 50))
        tim = dem.timing()
        dem.set_profiling(False)
        wall = dem.get("walls_mvfd")
        dem.close()
        return ms, c, tim, wall

    ms, c, tim, wall = timed_run(args.precision)
    log(f"{args.precision}: {ms:.3f} ms/step, C/N={c / n:.3f}, per-kernel {tim}")
    value = n / (ms * 1e-3)
    steps_prof = tim["steps"] or 1
    contact_ms = tim["contact_ms"] / steps_prof
    achieved = contact_bytes(n, c, 1) / (contact_ms * 1e-3) / 1e9
    result = {
        "metric": "particle-steps/s at 1M spheres (FP64 vs mixed); adjoint/forward step time ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic (seeded lognormal random packing)",
        "config": {"workload": "1,048,576-sphere lognormal packing under a 10 kPa top wall (2.07x2.07x4.13 m, x/y periodic)",
                   "n": n, "cells": K, "contacts_per_particle": c / n, "precision": args.precision},
        "per_kernel_ms": {k: (v / steps_prof if k.endswith("_ms") else v) for k, v in tim.items()},
        "roofline": {"bound": "hbm", "kernel": "k_contact", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": contact_bytes(n, c, 1)},
        "step_roofline": {"algorithmic_bytes": step_bytes(n, c, 0, 1, K),
                          "achieved_GBs": step_bytes(n, c, 0, 1, K) / (ms * 1e-3) / 1e9},
    }
    if not args.no_mixed and args.precision == "fp64":
        ms_m, c_m, tim_m, _ = timed_run("mixed")
        log(f"mixed: {ms_m:.3f} ms/step")
        result["mixed"] = {"value": n / (ms_m * 1e-3), "ms_per_step": ms_m, "speedup_vs_fp64": ms / ms_m,
                           "contact_ms": tim_m["contact_ms"] / (tim_m["steps"] or 1)}
    result["adjoint_forward_ratio"] = None
    if not args.no_cpu:
        rate, secs = cpu_baseline(bin_path, n, args.cpu_steps)
        result["cpu_baseline"] = {"value": rate, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                                  "sample": f"same 1,048,576-sphere input, {args.cpu_steps} steps after one untimed step ({secs:.1f}s)",
                                  "cpu": cpu_model()}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
