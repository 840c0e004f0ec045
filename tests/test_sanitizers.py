"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

* The file formats (form_factors_amd/csrc/ffio.cpp) and the conduction graph
  builder (form_factors_amd/csrc/ffk_graph.cpp), which read external files,
  are linked with tests/sanitize/io_harness.cpp and driven over every model
  and edge file of tests/golden/models (scene import with and without inline
  materials, export, CSV export, conduction graph; every .task case).
* The oracle (oracle/ff_oracle.cpp) is built as a sanitized shared library
  and tests/test_oracle.py runs against it in a child process.

Any sanitizer report aborts the run (-fno-sanitize-recover); host code only,
built with g++ here."""
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "tests", "golden", "models")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]

needs_gxx = pytest.mark.skipif(not shutil.which("g++"), reason="g++ not available")

TASKS = [("parallel_planes.task", 2, 1), ("grid_10x10_heat.task", 100, 1), ("tgrid_10x10_heat.task", 200, 1),
         ("confetti_2k.task", 2000, 3), ("patch_sphere_2k_sun.task", 2000, 1), ("edge/frames.task", 3, 3),
         ("edge/frames.task", 2, 4), ("edge/bad_step.task", 1, 1), ("edge/bad_tmprt.task", 1, 1)]


def _runtime_libs():
    libs = []
    for name in ("libasan.so", "libubsan.so"):
        p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
        if not os.path.isabs(p) or not os.path.exists(p):
            pytest.skip(f"{name} not available")
        libs.append(p)
    return libs


@needs_gxx
def test_file_formats_and_graph_under_sanitizers(tmp_path):
    _runtime_libs()
    exe = str(tmp_path / "io_harness")
    src = [os.path.join(ROOT, "tests", "sanitize", "io_harness.cpp"),
           os.path.join(ROOT, "form_factors_amd", "csrc", "ffio.cpp"),
           os.path.join(ROOT, "form_factors_amd", "csrc", "ffk_graph.cpp")]
    r = subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include"), *src, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    objs = sorted(glob.glob(os.path.join(MODELS, "*.obj")) + glob.glob(os.path.join(MODELS, "edge", "*.obj")))
    assert len(objs) >= 20
    args = []
    for o in objs:
        args += ["obj", o]
    for name, n, calls in TASKS:
        args += ["task", os.path.join(MODELS, name), str(n), str(calls)]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, *args], capture_output=True, text=True, env=env, cwd=str(tmp_path), timeout=600)
    report = r.stderr[-4000:]
    assert r.returncode == 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert sum(line.startswith("obj ") for line in r.stdout.splitlines()) == 2 * len(objs)


@needs_gxx
def test_oracle_under_sanitizers(tmp_path):
    libs = _runtime_libs()
    so = str(tmp_path / "liboracle_san.so")
    r = subprocess.run(["g++", "-std=c++17", *SAN, "-ffp-contract=off", "-fPIC", "-shared", "-pthread",
                        os.path.join(ROOT, "oracle", "ff_oracle.cpp"), "-o", so], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, FF_ORACLE_SO=so, LD_PRELOAD=":".join(libs),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py")], capture_output=True, text=True, env=env,
                       cwd=ROOT, timeout=1200)
    out = r.stdout[-3000:] + r.stderr[-3000:]
    assert r.returncode == 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, out
    assert " passed" in r.stdout
