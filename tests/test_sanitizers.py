"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer:
libdfx's host setup (dolfinx_amd/csrc/setup.cpp: meshes, RCM, numbering,
entities, dofmaps, boundary dofs, facet domains, sparsity, including the
threaded paths above 200,000 cells) and the CPU oracle (oracle/*.c), each
built with gcc/g++ -fsanitize=address,undefined into a small driver
(tests/sanitize/) and run on small meshes. A sanitizer report, a leak or a
failed call fails the test. CPU only."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
       "-fno-sanitize-recover=undefined"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
           UBSAN_OPTIONS="print_stacktrace=1")


def _run(build, exe):
    b = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    report = r.stdout[-2000:] + r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "LeakSanitizer" not in r.stderr, report
    return r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_setup_under_sanitizers(tmp_path):
    exe = str(tmp_path / "setup_asan")
    out = _run(["g++", "-std=c++17", *SAN, "-pthread", "-I", os.path.join(ROOT, "include"),
                "-I", os.path.join(ROOT, "dolfinx_amd", "csrc"),
                os.path.join(ROOT, "dolfinx_amd", "csrc", "setup.cpp"),
                os.path.join(ROOT, "tests", "sanitize", "setup_driver.cpp"), "-o", exe], exe)
    assert "setup ok" in out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_asan")
    src = [os.path.join(ROOT, "oracle", f) for f in ("mesh.c", "fem.c", "la.c")]
    out = _run(["gcc", "-std=c11", *SAN, "-D_DEFAULT_SOURCE", "-I", os.path.join(ROOT, "oracle"),
                *src, os.path.join(ROOT, "tests", "sanitize", "oracle_driver.c"), "-lm", "-o", exe],
               exe)
    assert "oracle ok" in out
