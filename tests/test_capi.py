"""The C-ABI library loads and exports every entry point its header declares
(no compute calls: they need a GPU)."""

import ctypes
import os
import re
import subprocess

from sphere_amd import dem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = dem.load_library()
    declared = _declared()
    assert len(declared) >= 20, sorted(declared)
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_exports_have_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", dem.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(sph_[a-z0-9_]+)$", out, flags=re.M))
    assert _declared() <= exported


def test_wrapper_binds_every_export():
    assert _declared() <= set(dem.EXPORTS)


def test_library_is_gfx950_only():
    data = open(dem.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa-+(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}, targets


def test_analysis_tools_parse_arguments_without_a_gpu():
    """`porosity` and `forcechains` (src/porosity.cpp, src/forcechains.cpp):
    help, version and the errors raised before any input is opened."""
    pkg = os.path.join(ROOT, "sphere_amd")
    for tool, flag in (("porosity", "--slices"), ("forcechains", "--lower-cutoff")):
        exe = os.path.join(pkg, tool)
        out = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0 and flag in out.stdout
        assert subprocess.run([exe, "-V"], capture_output=True, text=True, timeout=60).returncode == 0
        missing = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert missing.returncode == 1 and "missing input binary file" in missing.stderr
    bad = subprocess.run([os.path.join(pkg, "porosity"), "-s", "0"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "number of slices must be a positive" in bad.stderr


def test_sphere_status_reads_a_status_file(tmp_path):
    """`sphere_status <sid>` (src/sphere_status.cpp) on the "%2.4e %2.4e %d"
    file startTime writes."""
    (tmp_path / "output").mkdir()
    (tmp_path / "output" / "demo.status.dat").write_text("%2.4e %2.4e %d\n" % (0.0125, 12.5, 3))
    exe = os.path.join(ROOT, "sphere_amd", "sphere_status")
    out = subprocess.run([exe, "demo"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "Current simulation time:  0.0125 s" in out.stdout
    assert "Percentage completed:     12.5%" in out.stdout
    assert "demo.output00003.bin" in out.stdout
    assert subprocess.run([exe], cwd=tmp_path, capture_output=True, timeout=60).returncode == 1
    assert subprocess.run([exe, "absent"], cwd=tmp_path, capture_output=True, timeout=60).returncode == 1


def test_sphere_render_flags_parse_without_a_gpu():
    """`sphere -m <method>` (src/main.cpp:97-131): an unknown method is
    refused before any input is opened; -h lists the raytracer's flags."""
    exe = os.path.join(ROOT, "sphere_amd", "sphere")
    out = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "--render" in out.stdout and "angpos" in out.stdout
    bad = subprocess.run([exe, "-m", "colour", "1", "x.bin"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "Visualization method not understood" in bad.stderr
    short = subprocess.run([exe, "-m", "vel"], capture_output=True, text=True, timeout=60)
    assert short.returncode == 1 and "needs a maximum value" in short.stderr
