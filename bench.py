"""Benchmark of the gfx950 DEM time step.

Default workload: 1,048,576 spheres, lognormal radii (median 1e-2 m, sigma
0.2, clipped to [0.5, 2] x median), seeded random non-overlapping packing in
a 2.07 x 2.07 x 4.13 m x/y-periodic box (the nominal 0.8 x 0.8 x 1.6 m box
cannot hold that much solid), gravity, top wall at 10 kPa, contact
parameters k_n = k_t = 1e6 N/m, gamma_n = gamma_t = 100 N s/m, mu = 0.5.

A step is one full DEM time step (bonds, sort, reorder, contact search,
contact forces + integration + next cell ids, wall update) with all state
resident in HBM.  Prints ONE JSON line on stdout.

Also measured: the mixed-precision mode on the same input, and both modes on
a dense 1,048,576-sphere FCC crystal (twelve contacts per sphere), with the
mixed mode's error against the FP64 path per output; the adjoint on a
65,536-sphere shear cell; the 262,144-sphere bonded solid (about 1.5M
bonds) over 5,000 steps; and the 512-member calibration ensemble stepped in
one handle against the same members one at a time; and the reference's
analysis members (porosity, pair overlaps) on the dense crystal.  Per-kernel time comes
from hipEvents recorded on the simulation's own stream around each stage,
over a separate run of the same steps; the roofline entry is for the stage
with the largest share of the step.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 1M spheres (FP64 vs mixed); adjoint/forward step time ratio"
STAGES = ("bonds", "cellid", "sort", "reorder", "search", "force", "walls")
STAGE_KERNELS = {
    "bonds": "k_bonds", "cellid": "(ids, counts and ranks from k_force)", "sort": "k_scan_reduce/k_scan_down/k_scatter/k_cellsort",
    "reorder": "k_reorder", "search": "k_topo_rows", "force": "k_force", "walls": "k_wallsum/k_walls",
}
PROFILE_KERNEL = {"reorder": "k_reorder<false>", "search": "k_topo_rows<false, false>",
                  "force": "k_force<double, false, false, false>"}
FORCE_PROFILE = {1: "k_force<double, false, false, false>", 2: "k_force<double, false, true, false>"}
SEARCH_KERNELS = {1: "k_topo", 2: "k_topo_rows", 3: "k_topo_wave", 4: "k_topo_half/k_topo_apply",
                  5: "k_list_search (k_topo_rows<true> on rebuild steps)"}
SEARCH_PROFILE = {1: "k_topo", 2: "k_topo_rows<false, false>", 3: "k_topo_wave", 4: "k_topo_half<false>",
                  5: "k_list_search<false>"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes(n, c, b, w, k):
    """Algorithmic HBM bytes per launch of each stage: FP64 compact records,
    8 B per double, 4 B per index, every required array read or written once,
    neighbour and bond-endpoint gathers counted as cache resident."""
    return {
        "bonds": 136 * b,                          # pair 8, delta 32, omega 32 read; delta, omega 64 write
        "cellid": 0,                               # keys, counts and ranks are written by the force kernel (below)
        "sort": 16 * n + 8 * k,                    # key + index read and write, cell bounds
        "reorder": 184 * n,                        # key, index, 11 doubles gathered and written
        "search": 32 * n + 32 * c,                 # own x,y,z,r; old + new slot, image per active slot
        # interact (196N + 76C + 8WN) + integrate (388N) + next cell id and rank (8N)
        "force": 592 * n + 76 * c + 8 * w * n,
        "walls": 8 * w * n,
    }


def roofline_entry(kernel, algorithmic_bytes, ms, traffic=None):
    """HBM roofline of one kernel: algorithmic bytes per launch over its
    measured time against the 8 TB/s peak."""
    achieved = algorithmic_bytes / (ms * 1e-3) / 1e9 if ms > 0 else None
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
            "algorithmic_bytes_per_launch": algorithmic_bytes, "ms": ms}


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def measured_traffic(kernel, summary="c1_1M_pmc_traffic.md"):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", summary)
    if not kernel or not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\|\s*%s\s*\|.*\|\s*([0-9.eE+]+)\s*\|\s*$" % re.escape(kernel), line)
        if m:
            return float(m.group(1))
    return None


def cpu_baseline(bin_path, n, sample_steps):
    """Serial CPU oracle (one core) on `sample_steps` steps of the same input
    after one untimed step; returns particle-steps/s and seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle

    t0 = time.perf_counter()
    orc = Oracle(bin_path)
    log(f"oracle loaded in {time.perf_counter() - t0:.1f}s")
    orc.step(1)
    t0 = time.perf_counter()
    flags = orc.step(sample_steps)
    dt = time.perf_counter() - t0
    orc.close()
    if flags:
        log(f"oracle flags {flags}")
    return n * sample_steps / dt, dt


def timed_run(bin_path, precision, steps, warmup, prof_steps):
    """Warm-up, K timed steps (HIP graph replays), then a profiled segment
    with per-stage hipEvents.  Returns timings and the final state."""
    from sphere_amd.dem import DEM

    dem = DEM(bin_path, verbosity=0)
    n = dem.np
    dem.set_precision(precision)
    dem.step(warmup)
    # sph_step returns after the device finished the steps (it reads back the
    # error flags), so the timed region is bracketed by device syncs
    t0 = time.perf_counter()
    dem.step(steps)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    c = int((dem.get("contacts") != n).sum())
    dem.set_profiling(True)
    dem.step(prof_steps)
    tim = dem.timing()
    dem.set_profiling(False)
    info = dem.info()
    state = {"x": dem.get("x")[:, :3].copy(), "mvfd": dem.get("walls_mvfd") if info.nw else None,
             "wall_pp0": dem.get("wall_pp")[0].copy() if info.nw else None}
    dem.close()
    stage_ms = {k: tim[f"{k}_ms"] / max(1, tim["steps"]) for k in STAGES}
    return {"ms": ms, "contacts": c, "stage_ms": stage_ms, "tile": [tim["tile_x"], tim["tile_y"]],
            "search_mode": int(tim["search_mode"]), "force_variant": int(tim["force_variant"]), "state": state,
            "steps_total": warmup + steps + prof_steps}


def mixed_error(ref, mix, area, r_min):
    """Error of the mixed run against the FP64 run after the same steps."""
    dx = np.abs(mix["state"]["x"] - ref["state"]["x"])
    out = {"steps": ref["steps_total"], "pos_max_abs_m": float(dx.max()),
           "pos_max_rel_to_r_min": float(dx.max() / r_min)}
    if ref["state"]["wall_pp0"] is not None:
        # normal stress on the top wall: its per-particle contact forces summed
        # (the same sum the wall update uses), over the wall area
        s_ref = float(np.sum(ref["state"]["wall_pp0"])) / area
        s_mix = float(np.sum(mix["state"]["wall_pp0"])) / area
        out["wall_stress_fp64_Pa"] = s_ref
        out["wall_stress_mixed_Pa"] = s_mix
        out["wall_stress_rel"] = abs(s_mix - s_ref) / abs(s_ref) if s_ref != 0.0 else None
    return out


def adjoint_run(tmp, nsteps, seg=100):
    """Mean-shear-stress gradient on the differentiable shear test: 65,536
    polydisperse spheres (config-1 radii) compacted under the 10 kPa wall,
    then sheared at 0.1 m/s (configs.consolidated_shear_cell): plain forward
    time against the adjoint's forward (with checkpoints every `seg` steps)
    and backward (recompute with tape + reverse sweep)."""
    from sphere_amd import configs
    from sphere_amd.adjoint import shear_stress_weights
    from sphere_amd.dem import DEM

    t0 = time.perf_counter()
    sb = configs.consolidated_shear_cell()
    t_recipe = time.perf_counter() - t0
    path = os.path.join(tmp, "shear.bin")
    sb.writebin(path)
    w = shear_stress_weights(sb)
    dem = DEM(path, verbosity=0)
    dem.step(20)
    dem.close()
    dem = DEM(path, verbosity=0)
    t0 = time.perf_counter()
    dem.step(nsteps)
    t_plain = time.perf_counter() - t0
    dem.close()
    dem = DEM(path, verbosity=0)
    t0 = time.perf_counter()
    J = dem.adjoint_forward(nsteps, w, seg_len=seg)
    t1 = time.perf_counter()
    g = dem.adjoint_backward()
    t2 = time.perf_counter()
    dem.close()
    log(f"adjoint: packing {t_recipe:.1f} s; plain {t_plain * 1e3 / nsteps:.4f} ms/step, forward+checkpoints "
        f"{(t1 - t0) * 1e3 / nsteps:.4f}, backward {(t2 - t1) * 1e3 / nsteps:.4f}; J={J:.6e} grad={g}")
    return {"workload": f"{sb.np:,}-sphere polydisperse shear cell (lognormal radii), compacted under 10 kPa, "
                        f"top layer at 0.1 m/s, {nsteps} steps, checkpoints every {seg}",
            "driven_spheres": int((w[:, 0] != 0.0).sum()), "objective": "mean shear stress (Pa)", "J": J,
            "grad": dict(zip(("k_n", "k_t", "gamma_n", "mu"), [float(v) for v in g])),
            "plain_ms_per_step": t_plain * 1e3 / nsteps, "forward_ms_per_step": (t1 - t0) * 1e3 / nsteps,
            "backward_ms_per_step": (t2 - t1) * 1e3 / nsteps,
            "backward_over_forward": (t2 - t1) / t_plain, "total_over_forward": (t2 - t0) / t_plain}


def bonded_run(tmp, steps, warmup, prof_steps):
    """The 262,144-sphere bonded FCC solid (about 1.5M bonds, strengths 1e5 Pa
    so that bonds break during the run) pulled apart for `steps` FP64 steps."""
    from sphere_amd import configs
    from sphere_amd.dem import DEM

    sb = configs.fcc_bonded(cells=(32, 32, 64), sigma_b=1e5, tau_b=1e5, steps=steps)
    path = os.path.join(tmp, "bonded.bin")
    sb.writebin(path)
    n, nb = sb.np, sb.nb0
    dem = DEM(path, verbosity=0)
    dem.step(warmup)
    t0 = time.perf_counter()
    dem.step(steps)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    broken = int((dem.get("bonds")[:, 0] >= n).sum())
    dem.set_profiling(True)
    dem.step(prof_steps)
    tim = dem.timing()
    dem.close()
    stage_ms = {k: tim[f"{k}_ms"] / max(1, tim["steps"]) for k in STAGES}
    bonds_bytes = 136 * nb
    log(f"bonded: {ms:.4f} ms/step, {nb} bonds, {broken} broken after {warmup + steps} steps, stages {stage_ms}")
    return {"workload": f"{n:,}-sphere jittered FCC solid, {nb:,} bonds (d <= 1.05 (r_i + r_j)), k = 1e7, "
                        f"sigma_b = tau_b = 1e5 Pa, top and bottom layers pulled apart at 0.01 m/s, FP64",
            "n": n, "bonds": nb, "steps": steps, "warmup": warmup, "value": n / (ms * 1e-3),
            "unit": "particle-steps/s", "ms_per_step": ms, "broken_bonds": broken, "stage_ms": stage_ms,
            "k_bonds": {"algorithmic_bytes": bonds_bytes, "GBs": bonds_bytes / (stage_ms["bonds"] * 1e-3) / 1e9
                        if stage_ms["bonds"] > 0 else None}}


def analysis_run(dpath, reps=5, with_cpu=True):
    """DEM::porosity and DEM::findOverlaps (src/sphere.cpp:363-604) on the
    dense crystal after 20 steps: wall-clock per call (device work, the
    pair download and the call's buffer allocations), and the serial oracle
    on the same porosity and on the pair search of a 16,384-sphere crystal."""
    from sphere_amd import configs
    from sphere_amd.dem import DEM

    dem = DEM(dpath, verbosity=0)
    n = dem.np
    dem.step(20)
    out = {"workload": "1,048,576-sphere FCC crystal after 20 steps (sorted on the device)"}
    for slices, exact in ((10, False), (10, True), (1000, True)):
        dem.porosity(slices, exact=exact)
        t0 = time.perf_counter()
        for _ in range(reps):
            dem.porosity(slices, exact=exact)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[f"porosity_{slices}_{'exact' if exact else 'reference'}_ms"] = ms
    dem.findOverlaps()
    t0 = time.perf_counter()
    for _ in range(reps):
        ij, dn = dem.findOverlaps()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    out["overlaps"] = int(len(ij))
    out["find_overlaps_ms"] = ms
    out["find_overlaps_pairs_per_s"] = len(ij) / (ms * 1e-3)
    # DEM::render (src/raytracer.cuh:374-553) at the reference's 800 x 800
    img = dem.render(0, 1e3, 0.0, 800, 800)
    t0 = time.perf_counter()
    for _ in range(reps):
        img = dem.render(0, 1e3, 0.0, 800, 800)
    out["render_800_ms"] = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    brute = dem.render(0, 1e3, 0.0, 800, 800, cull=False)
    out["render_800_every_sphere_per_pixel_ms"] = (time.perf_counter() - t0) * 1e3
    out["render_tiles_equal_every_sphere"] = bool(np.array_equal(img, brute))
    log(f"analysis: render 800x800 {out['render_800_ms']:.2f} ms tiled, "
        f"{out['render_800_every_sphere_per_pixel_ms']:.1f} ms testing every sphere per pixel")
    dem.close()
    log(f"analysis: porosity(10) {out['porosity_10_reference_ms']:.2f} ms, find_overlaps {ms:.2f} ms "
        f"({len(ij):,} pairs)")
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle

        orc = Oracle(dpath)
        t0 = time.perf_counter()
        orc.porosity(10, 0)
        out["cpu_porosity_10_reference_ms"] = (time.perf_counter() - t0) * 1e3
        orc.close()
        small = configs.dense_lattice(cells=(16, 16, 16), sid="dense-small")
        spath = os.path.join(os.path.dirname(dpath), "dense-small.bin")
        small.writebin(spath)
        orc = Oracle(spath)
        t0 = time.perf_counter()
        npairs = len(orc.find_overlaps()[0])
        secs = time.perf_counter() - t0
        orc.close()
        out["cpu_find_overlaps"] = {"n": small.np, "pairs": npairs, "seconds": secs,
                                    "note": "the reference's O(N^2) double loop, serial, one core"}
        orc = Oracle(spath)
        t0 = time.perf_counter()
        orc.render(0, 1e3, 0.0, 200, 200)
        out["cpu_render"] = {"n": small.np, "pixels": 200 * 200, "seconds": time.perf_counter() - t0,
                             "note": "the reference's every-sphere-per-pixel loop, serial, one core"}
        orc.close()
        log(f"analysis cpu: porosity(10) {out['cpu_porosity_10_reference_ms']:.1f} ms at {n:,}; "
            f"find_overlaps {secs:.2f} s at {small.np:,}")
    return out


def regrid(sb, factor):
    """A member on the grid sph_ens_open gives it for cell_factor >= 1
    (cells of factor x 2 r_max, at least 3 on a periodic axis)."""
    import math

    sb = sb.copy()
    if factor >= 1.0:
        edge = factor * 2.0 * float(np.max(sb.radius))
        num = []
        for a in range(3):
            per = (a == 0 and sb.periodic in (1, 2)) or (a == 1 and sb.periodic == 1)
            num.append(max(math.floor(sb.L[a] / edge), 3 if per else 1))
        sb.num = np.array(num, dtype=np.uint32)
    return sb


def ensemble_run(members, steps, warmup, sample_every=1, sample_steps=None):
    """The calibration ensemble in one handle (one launch of each kernel per
    step for every member, graph-replayed), mixed precision, grid and block
    sizes from the autotuner; against the same members run one at a time,
    each in its own handle for the same steps (every member by default;
    `sample_every` > 1 extrapolates a sample by member count)."""
    sample_steps = sample_steps or steps
    from sphere_amd.dem import DEM, Ensemble

    n_total = sum(m.np for m in members)
    t0 = time.perf_counter()
    ens = Ensemble(members, cell_factor=0.0)
    t_open = time.perf_counter() - t0
    ens.set_precision("mixed")
    ens.step(warmup)  # the first call also tunes the search kernel and the block sizes (on a copy)
    t0 = time.perf_counter()
    ens.step(steps)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    info = ens.info()
    ens.set_profiling(True)
    ens.step(20)
    tim = ens.timing()
    ens.close()
    stage_ms = {k: tim[f"{k}_ms"] / max(1, tim["steps"]) for k in STAGES}
    log(f"ensemble: {len(members)} members, {n_total} spheres, {ms:.4f} ms/step, cell factor {info.cell_factor}, "
        f"search {info.search_mode}, blocks force {info.force_block} search {info.search_block}, open {t_open:.1f}s")
    sample = members[::sample_every]
    per = []
    for sb in sample:
        d = DEM(regrid(sb, info.cell_factor), verbosity=0)
        d.set_precision("mixed")
        d.step(warmup)
        t0 = time.perf_counter()
        d.step(sample_steps)
        per.append((time.perf_counter() - t0) * 1e3 / sample_steps)
        d.close()
    ms_one = float(np.sum(per)) * sample_every
    log(f"one at a time: {np.mean(per):.4f} ms/step per member ({len(sample)} members timed), {ms_one:.2f} ms per "
        f"ensemble step")
    return {"workload": f"{len(members)} packings, {n_total:,} spheres (counts uniform in [1024, 32768]), k_n uniform "
                        "in [1e5, 1e7], mu uniform in [0.2, 0.8], C1 radii, settling under a 10 kPa top wall, mixed "
                        "precision",
            "members": len(members), "n": n_total, "steps": steps, "warmup": warmup,
            "value": n_total / (ms * 1e-3), "unit": "particle-steps/s", "ms_per_step": ms, "stage_ms": stage_ms,
            "autotuned": {"cell_factor": info.cell_factor, "trial_ms_per_step": list(info.tuned_ms),
                          "search_kernel": SEARCH_KERNELS.get(info.search_mode),
                          "force_block": info.force_block, "search_block": info.search_block},
            "one_at_a_time": {"value": n_total / (ms_one * 1e-3), "ms_per_ensemble_step": ms_one,
                              "ms_per_member_step": float(np.mean(per)), "members_timed": len(sample),
                              "steps": sample_steps, "grid": "each member on the ensemble's autotuned grid"},
            "speedup_vs_one_at_a_time": ms_one / ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # (the first sph_step call of any length times the search and force
    # kernel variants on a copy of the state and restores it, so every step
    # of every call runs the chosen kernels whatever the warm-up)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--cpu-steps", type=int, default=13)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-dense", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-bonded", action="store_true")
    ap.add_argument("--no-analysis", action="store_true")
    ap.add_argument("--bonded-steps", type=int, default=5000)
    ap.add_argument("--ensemble-members", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--probe-only", action="store_true",
                    help="run only the copy-kernel bandwidth probe (counter calibration)")
    ap.add_argument("--profile-leg", choices=["bonded", "ensemble"],
                    help="run only this leg (for kernel profiles): the bonded solid for --bonded-steps, or the "
                         "calibration ensemble for --ensemble-steps without the one-at-a-time comparison")
    ap.add_argument("--profile-dense", choices=["fp64", "mixed"],
                    help="run only the dense crystal in this precision (kernel profiles); prints its timings")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU decomposition is a design note",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from sphere_amd import configs
    from sphere_amd.dem import bandwidth_probe, flops_probe

    if args.probe_only:
        gbs = bandwidth_probe(1 << 30, 20)
        print(json.dumps({"probe": "k_copy_probe", "bytes_per_launch": 2 * (1 << 30), "GBs": gbs}))
        for fp64 in (True, False):
            tf = flops_probe(fp64, 4096, 4096, 10)
            print(json.dumps({"probe": "k_flops_probe<%s>" % ("double" if fp64 else "float"),
                              "flops_per_launch": 16 * 4096 * 4096 * 256, "TFLOPs": tf}))
        return

    prof_steps = min(args.steps, 50)
    tmp = tempfile.mkdtemp(prefix="sphere-bench-")
    if args.profile_leg == "bonded":
        print(json.dumps(bonded_run(tmp, args.bonded_steps, args.warmup, prof_steps)), flush=True)
        return
    if args.profile_leg == "ensemble":
        members = configs.calibration_ensemble(members=args.ensemble_members, steps=args.ensemble_steps)
        print(json.dumps(ensemble_run(members, args.ensemble_steps, max(args.warmup, 1), sample_every=len(members),
                                      sample_steps=1)), flush=True)
        return
    if args.profile_dense:
        sbd = configs.dense_lattice()
        dpath = os.path.join(tmp, "dense.bin")
        sbd.writebin(dpath)
        run = timed_run(dpath, args.profile_dense, args.steps, args.warmup, prof_steps)
        print(json.dumps({"workload": "dense crystal", "n": sbd.np, "precision": args.profile_dense,
                          "ms_per_step": run["ms"], "contacts_per_particle": run["contacts"] / sbd.np,
                          "stage_ms": run["stage_ms"], "search_mode": run["search_mode"],
                          "force_variant": run["force_variant"]}), flush=True)
        return
    t0 = time.perf_counter()
    sb = configs.random_compression(n=args.n)
    log(f"generated {args.n} spheres in {time.perf_counter() - t0:.1f}s, grid {[int(v) for v in sb.num]}")
    bin_path = os.path.join(tmp, "c1.bin")
    sb.writebin(bin_path)
    n = sb.np
    K = int(sb.num[0]) * int(sb.num[1]) * int(sb.num[2])
    area = float(sb.L[0] * sb.L[1])
    r_min = float(sb.radius.min())

    main_run = timed_run(bin_path, args.precision, args.steps, args.warmup, prof_steps)
    ms, c, stage_ms = main_run["ms"], main_run["contacts"], main_run["stage_ms"]
    log(f"{args.precision}: {ms:.4f} ms/step, C/N={c / n:.3f}, stages {stage_ms}")
    kb = kernel_bytes(n, c, sb.nb0, sb.nw, K)
    stage_kernels = dict(STAGE_KERNELS, search=SEARCH_KERNELS.get(main_run["search_mode"], "?"))
    profile_kernel = dict(PROFILE_KERNEL, search=SEARCH_PROFILE.get(main_run["search_mode"]),
                          force=FORCE_PROFILE[main_run["force_variant"]])
    per_kernel = {k: {"ms": stage_ms[k], "kernels": stage_kernels[k], "algorithmic_bytes": kb[k],
                      "GBs": (kb[k] / (stage_ms[k] * 1e-3) / 1e9) if stage_ms[k] > 0 and kb[k] else None}
                  for k in STAGES}
    dom = max(stage_ms, key=lambda k: stage_ms[k])
    achieved = kb[dom] / (stage_ms[dom] * 1e-3) / 1e9
    step_b = sum(kb.values())
    result = {
        "metric": METRIC,
        "value": n / (ms * 1e-3),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 state, f32 pair forces",
        "data": "synthetic (seeded lognormal random packing)",
        "config": {"workload": "1,048,576 lognormal spheres, random packing, 10 kPa top wall, 2.07x2.07x4.13 m x/y-periodic box",
                   "n": n, "cells": K, "contacts_per_particle": c / n, "precision": args.precision,
                   "search_kernel": stage_kernels["search"], "search_mode": main_run["search_mode"]},
        "per_kernel": per_kernel,
        "roofline": {"bound": "hbm", "kernel": stage_kernels[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(profile_kernel.get(dom)),
                     "algorithmic_bytes_per_launch": kb[dom]},
        "step_roofline": {"algorithmic_bytes": step_b, "achieved_GBs": step_b / (ms * 1e-3) / 1e9,
                          "frac": step_b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_mixed and args.precision == "fp64":
        mix = timed_run(bin_path, "mixed", args.steps, args.warmup, prof_steps)
        log(f"mixed: {mix['ms']:.4f} ms/step")
        result["mixed"] = {"value": n / (mix["ms"] * 1e-3), "ms_per_step": mix["ms"],
                           "speedup_vs_fp64": ms / mix["ms"], "contacts_per_particle": mix["contacts"] / n,
                           "error_vs_fp64": mixed_error(main_run, mix, area, r_min)}
    if not args.no_dense and args.precision == "fp64":
        sbd = configs.dense_lattice()
        dpath = os.path.join(tmp, "dense.bin")
        sbd.writebin(dpath)
        dn = sbd.np
        d64 = timed_run(dpath, "fp64", args.steps, args.warmup, prof_steps)
        dense = {"workload": "1,048,576-sphere FCC crystal (4 x 64^3), nearest neighbours overlapping by 2e-5 m, "
                             "x/y periodic, floor and fixed top wall, gamma_n = 0",
                 "contacts_per_particle": d64["contacts"] / dn, "fp64": {"value": dn / (d64["ms"] * 1e-3),
                 "ms_per_step": d64["ms"], "stage_ms": d64["stage_ms"],
                 "search_kernel": SEARCH_KERNELS.get(d64["search_mode"])}}
        kd = kernel_bytes(dn, d64["contacts"], 0, sbd.nw, int(np.prod(sbd.num.astype(np.int64))))
        dsk = SEARCH_PROFILE.get(d64["search_mode"])
        dense["fp64"]["roofline"] = {
            "force": roofline_entry(FORCE_PROFILE[d64["force_variant"]], kd["force"], d64["stage_ms"]["force"],
                                    measured_traffic(FORCE_PROFILE[d64["force_variant"]], "dense_1M_fp64_pmc_traffic.md")),
            "search": roofline_entry(dsk, kd["search"], d64["stage_ms"]["search"],
                                     measured_traffic(dsk, "dense_1M_fp64_pmc_traffic.md")),
        }
        log(f"dense fp64: {d64['ms']:.4f} ms/step, C/N={d64['contacts'] / dn:.2f}")
        if not args.no_mixed:
            dmx = timed_run(dpath, "mixed", args.steps, args.warmup, prof_steps)
            log(f"dense mixed: {dmx['ms']:.4f} ms/step")
            dense["mixed"] = {"value": dn / (dmx["ms"] * 1e-3), "ms_per_step": dmx["ms"],
                              "speedup_vs_fp64": d64["ms"] / dmx["ms"], "stage_ms": dmx["stage_ms"],
                              "error_vs_fp64": mixed_error(d64, dmx, float(sbd.L[0] * sbd.L[1]),
                                                           float(sbd.radius.min()))}
        result["dense"] = dense
        if not args.no_analysis:
            result["analysis"] = analysis_run(dpath, with_cpu=not args.no_cpu)
    result["adjoint_forward_ratio"] = None
    if not args.no_adjoint:
        result["adjoint"] = adjoint_run(tmp, args.adjoint_steps)
        result["adjoint_forward_ratio"] = result["adjoint"]["backward_over_forward"]
    if not args.no_bonded and args.precision == "fp64":
        result["bonded"] = bonded_run(tmp, args.bonded_steps, args.warmup, prof_steps)
    if not args.no_ensemble:
        t0 = time.perf_counter()
        members = configs.calibration_ensemble(members=args.ensemble_members, steps=args.ensemble_steps)
        log(f"generated the ensemble in {time.perf_counter() - t0:.1f}s")
        result["ensemble"] = ensemble_run(members, args.ensemble_steps, max(args.warmup, 1))
        del members
    if not args.no_cpu:
        rate, secs = cpu_baseline(bin_path, n, args.cpu_steps)
        result["cpu_baseline"] = {"value": rate, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                                  "sample": f"same 1,048,576-sphere input, {args.cpu_steps} steps after one untimed step ({secs:.1f} s)",
                                  "cpu": cpu_model(), "host_nproc": os.cpu_count()}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
