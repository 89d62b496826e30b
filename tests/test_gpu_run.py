"""DEM::startTime (src/device.cu:576-1299) through sph_run and the `sphere`
command line: which steps write output files (the FP32 file clock of
src/device.cu:594,1109-1123), every output file against the oracle stepped
to the same step, and status.dat (src/device.cu:655-662,1178-1185)."""

import os
import subprocess

import numpy as np
import pytest

from oracle_lib import Oracle
from sphere_amd import configs
from sphere_amd.dem import DEM
from sphere_amd.spherebin import Spherebin

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("oracle_built")]

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sphere_amd")


def _cadence(t0, dt, total, file_dt):
    """The reference's loop: step, time += dt, filetimeclock (float) += dt,
    output (file and status.dat) when filetimeclock >= file_dt.  Returns the
    step counts after which files 1, 2, ... are written, the number of steps
    run and the time of the last status.dat."""
    t, clock, steps, outs, t_status = t0, np.float32(0.0), 0, [], t0
    while t <= total:
        steps += 1
        t += dt
        clock = np.float32(float(clock) + dt)
        if float(clock) >= file_dt:
            outs.append(steps)
            t_status = t
            clock = np.float32(0.0)
    return outs, steps, t_status


def _same_state(path, orc, k, tmp_path):
    got = Spherebin.readbin(str(path))
    ref_path = tmp_path / f"oracle-{k}.bin"
    orc.writebin(str(ref_path))
    want = Spherebin.readbin(str(ref_path))
    for name in ("x", "radius", "vel", "fixvel", "angvel", "angpos", "xysum", "force", "torque", "es_dot", "es",
                 "ev_dot", "ev", "p", "w_x", "w_vel", "w_force"):
        a, b = np.asarray(getattr(got, name)), np.asarray(getattr(want, name))
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (path.name, name)
    assert got.time_current == want.time_current
    assert got.time_step_count == k


@pytest.mark.parametrize("file_steps", [7.3, 10.0])
def test_output_cadence_and_files(tmp_path, file_steps):
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=1, nw=1,
                                gamma_n=0.0, gamma_t=10.0, sid=f"run-cadence{int(file_steps)}")
    sb.time_total = 61.5 * sb.time_dt
    sb.time_file_dt = file_steps * sb.time_dt
    outs, nsteps, t_status = _cadence(sb.time_current, sb.time_dt, sb.time_total, sb.time_file_dt)
    assert len(outs) >= 5
    dem = DEM(sb, verbosity=0)
    outdir = tmp_path / "output"
    dem.startTime(str(outdir))
    files = sorted(p.name for p in outdir.glob(f"{sb.sid}.output*.bin"))
    assert files == [f"{sb.sid}.output{k:05d}.bin" for k in range(len(outs) + 1)]
    orc = Oracle.from_spherebin(sb, tmp_path)
    done = 0
    for k, at in enumerate([0] + outs):
        assert orc.step(at - done) == 0
        done = at
        _same_state(outdir / files[k], orc, k, tmp_path)
    status = (outdir / f"{sb.sid}.status.dat").read_text()
    # status.dat is rewritten with each output file only
    assert status == "%2.4e %2.4e %d\n" % (t_status, 100.0 * t_status / sb.time_total, len(outs))
    # the handle has run every step of the loop
    assert orc.step(nsteps - done) == 0
    assert np.array_equal(dem.get("x"), orc.get("x"))


def test_sphere_command_line_runs_the_same_files(tmp_path):
    sb = configs.lattice_settle(nside=(6, 6, 5), spacing=2.15e-2, jitter=5e-4, periodic=1,
                                gamma_n=0.0, gamma_t=10.0, sid="run-cli")
    sb.time_total = 30.5 * sb.time_dt
    sb.time_file_dt = 10.0 * sb.time_dt
    (tmp_path / "input").mkdir()
    sb.writebin(str(tmp_path / "input" / "run-cli.bin"))
    out = subprocess.run([os.path.join(PKG, "sphere"), "-q", "input/run-cli.bin"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    dem = DEM(sb, verbosity=0)
    dem.startTime(str(tmp_path / "api"))
    for name in sorted(os.listdir(tmp_path / "api")):
        assert (tmp_path / "output" / name).read_bytes() == (tmp_path / "api" / name).read_bytes(), name
    status = subprocess.run([os.path.join(PKG, "sphere_status"), "run-cli"], cwd=tmp_path, capture_output=True,
                            text=True, timeout=60)
    outs, _, _ = _cadence(sb.time_current, sb.time_dt, sb.time_total, sb.time_file_dt)
    assert status.returncode == 0 and f"run-cli.output{len(outs):05d}.bin" in status.stdout
