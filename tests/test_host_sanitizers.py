"""The host C++ (YAML parser, FTOD / cc4s readers, step runner, Davidson
driver and the EOM packed spaces) built with AddressSanitizer and
UndefinedBehaviorSanitizer, and the CPU tests that exercise it run against
that build: the reader and runner tests with the sanitized step runner, the
Davidson tests with the sanitized harness.  Any sanitizer report fails the
run (non-zero exit)."""

import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "sisi4s_amd")


@pytest.fixture(scope="module")
def sanitized_runner(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    if not os.path.exists(os.path.join(LIB_DIR, "libsisi_engine.so")):
        pytest.skip("engine library not built")
    host = os.path.join(LIB_DIR, "host")
    srcs = sorted(os.path.join(host, f) for f in os.listdir(host) if f.endswith(".cpp"))
    exe = str(tmp_path_factory.mktemp("asan") / "sisi4s-mi355x-asan")
    cmd = [gxx, "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-std=c++17", "-Wall",
           "-I" + os.path.join(ROOT, "include")] + srcs + [
        "-L" + LIB_DIR, "-lsisi_engine", "-Wl,-rpath," + LIB_DIR, "-L/opt/rocm/lib",
        "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def _pytest(files, env_extra):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", **env_extra)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
                       + [os.path.join(ROOT, "tests", f) for f in files],
                       capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stdout + r.stderr
    assert "runtime error:" not in r.stdout + r.stderr


def test_readers_and_runner_under_sanitizers(sanitized_runner):
    _pytest(["test_readers.py", "test_host_runner.py"], {"SISI_RUNNER": sanitized_runner})


def test_davidson_driver_under_sanitizers():
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    _pytest(["test_davidson.py"], {"SISI_SANITIZE": "1"})
