"""Build the HIP engine library in-tree for gfx950.

    python -m sisi4s_amd.build        (or sisi4s_amd.build.build())

Compiles every source under sisi4s_amd/csrc with hipcc into objects under
sisi4s_amd/_build and links sisi4s_amd/libsisi_engine.so.
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsisi_engine.so")
RUNNER = os.path.join(HERE, "sisi4s-mi355x")
HOST = os.path.join(HERE, "host")
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(HERE, "..", "include")]
# Device code: no packed-FP32 instructions (v_pk_{add,mul,fma}_f32).
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OUT, src + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-pthread", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []
    errors = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for src, obj, rc, log in ex.map(_compile, _sources()):
            objs.append(obj)
            if rc != 0:
                errors.append("%s:\n%s" % (src, log))
            elif verbose and log.strip():
                print(log)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs + \
        ["-L/opt/rocm/lib", "-lhipblaslt", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_runner()
    return LIB


def build_runner():
    """Host step runner (C++ mirror of the plugin surface) linked to the engine."""
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    tmp = RUNNER + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include")] + srcs + \
        ["-L" + HERE, "-lsisi_engine", "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib",
         "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("runner build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, RUNNER)
    return RUNNER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
