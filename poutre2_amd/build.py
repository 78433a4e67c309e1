"""Build the gfx950 library libpoutre_amd.so in-tree with hipcc.

Every .hip / .cpp file under csrc/ is compiled for gfx950 only, in parallel, and
linked into poutre2_amd/libpoutre_amd.so, which exports the C ABI of
include/poutre_amd.h and the C++ API of include/poutre_amd.hpp.

Usage: python -m poutre2_amd.build   (or poutre2_amd.build.build())
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpoutre_amd.so")
# C++ API test program (tests/cpp/api_kats.cpp), linked against the library
API_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "api_kats.cpp")
API_TEST_BIN = os.path.join(HERE, "api_kats")
# host-only C++ API checks (tests/cpp/host_kats.cpp): fixture format and SE model, no GPU needed
HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "host_kats.cpp")
HOST_TEST_BIN = os.path.join(HERE, "host_kats")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# packed-fp32 instructions are disabled: hipcc would otherwise form them from plain float code
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lrocprofiler-sdk-roctx"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for src, out in ((API_TEST_SRC, API_TEST_BIN), (HOST_TEST_SRC, HOST_TEST_BIN)):
        if not os.path.exists(src):
            continue
        cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", out,
               "-L", HERE, "-lpoutre_amd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ test program build failed ({os.path.basename(src)}):\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
