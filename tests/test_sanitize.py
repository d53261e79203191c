"""Host code under AddressSanitizer and UBSan (CPU only).

The loaders, format converters and the CPU oracle are compiled with
``-fsanitize=address,undefined -fno-sanitize-recover=all`` into small drivers
(``tests/sanitize/``) that feed them random and mutated inputs; a sanitizer
report, a leak or a count that differs from brute force fails the test.  No
GPU code is built here: the device kernels cannot run under a sanitizer.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
# The sanitizer runtime must come first in the library list; a preloaded
# library of the environment may precede it, which is harmless here.
ENV = dict(os.environ, ASAN_OPTIONS="verify_asan_link_order=0:detect_leaks=1:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")


def _build(tmp_path, compiler, sources, flags, name):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not available")
    exe = str(tmp_path / name)
    cmd = [compiler] + SAN + flags + sources + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def test_loaders_and_converters_under_asan(tmp_path):
    src = os.path.join(ROOT, "kcgpu_amd", "csrc")
    exe = _build(tmp_path, "g++",
                 [os.path.join(ROOT, "tests", "sanitize", "loaders_fuzz.cpp"),
                  os.path.join(src, "kc_io.cpp"), os.path.join(src, "kc_convert.cpp")],
                 ["-std=c++17", "-fopenmp"], "loaders_fuzz")
    scratch = tmp_path / "files"
    scratch.mkdir()
    r = subprocess.run([exe, str(scratch), "2500"], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "loaded=" in r.stdout


def test_oracle_vs_brute_force_under_asan(tmp_path):
    exe = _build(tmp_path, "gcc",
                 [os.path.join(ROOT, "tests", "sanitize", "oracle_vs_brute.c"),
                  os.path.join(ROOT, "oracle", "kc_oracle.c")],
                 ["-fopenmp"], "oracle_vs_brute")
    r = subprocess.run([exe, "120"], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "bad=0" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
