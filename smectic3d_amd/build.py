"""Build libsmectic.so in-tree for gfx950 with hipcc.

    python -m smectic3d_amd.build            # build (always recompiles)
    python -m smectic3d_amd.build --check    # build only if sources are newer

Every translation unit is compiled with -Xclang -target-feature -Xclang
-packed-fp32-ops (no packed FP32 instructions are emitted).
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsmectic.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

UNITS = ["smx_inst_f32.hip", "smx_inst_f64.hip", "smx_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", INCLUDE]


def _sources():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(INCLUDE, "smectic.h")]


def up_to_date():
    if not os.path.exists(LIB):
        return False
    t = os.path.getmtime(LIB)
    return all(os.path.getmtime(s) <= t for s in _sources())


def _compile(unit):
    src = os.path.join(CSRC, unit)
    obj = os.path.join(OBJ, unit.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32-ops feature and says so; that is harmless.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {unit}:\n{err}")
    return obj


DRIVER_SRC = os.path.join(HERE, "driver", "smx_run.cpp")
DRIVER = os.path.join(HERE, "driver", "smx_run")


def build_driver():
    """The C++ driver (reference main() loops over the C ABI), linked against libsmectic.so."""
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, DRIVER_SRC,
           "-L", HERE, "-lsmectic", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n" + r.stderr)
    return DRIVER


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(UNITS), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print(f"built {LIB} and {DRIVER}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true", help="rebuild only when a source is newer")
    a = ap.parse_args()
    if a.check and up_to_date():
        print("up to date")
        sys.exit(0)
    build()
