"""Build the native libraries in-tree (no build system needed).

  genie_amd/lib/libgenie_hip.so    HIP engine for gfx950 (C ABI: include/genie_hip.h)
  genie_amd/lib/libgenie.so        C++ host library (GENIE API, include/genie/) + C entry
                                   points (include/genie_table.h); links libgenie_hip.so
  genie_amd/lib/libgenie_synth.so  deterministic synthetic workloads
  genie_amd/lib/test_api           C++ API test driver (tests/cpp/test_api.cc)

Every call recompiles everything (a few seconds per file).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "genie_amd")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("GENIE_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed FP32 instructions stay off on this target
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result", f"-I{INC}"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{INC}"]

HIP_SOURCES = ["hip/engine.hip", "hip/server.hip"]
HOST_SOURCES = ["host/codecs.cc", "host/inv_list.cc", "host/inv_table.cc", "host/query.cc",
                "host/knn.cc", "host/original.cc", "host/interface.cc", "host/file_reader.cc",
                "host/capi.cc", "host/serialization.cc", "host/json.cc", "host/wire.cc"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _obj(src):
    return os.path.join(OBJ, src.replace("/", "_") + ".o")


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in HOST_SOURCES:
        jobs.append([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    jobs.append([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "synth/synth.cc"), "-o", _obj("synth/synth.cc")])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    hip_so = os.path.join(LIB, "libgenie_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", hip_so,
          *[_obj(s) for s in HIP_SOURCES]])
    _run([CXX, "-shared", "-fPIC", "-pthread", "-o", os.path.join(LIB, "libgenie.so"),
          *[_obj(s) for s in HOST_SOURCES], f"-L{LIB}", "-lgenie_hip", "-Wl,-rpath,$ORIGIN"])
    _run([CXX, "-shared", "-fPIC", "-pthread", "-o", os.path.join(LIB, "libgenie_synth.so"),
          _obj("synth/synth.cc")])
    test_src = os.path.join(ROOT, "tests", "cpp", "test_api.cc")
    if os.path.exists(test_src):
        # the C++ API test driver; it links the CPU oracle as its checker
        odir = os.path.join(ROOT, "oracle")
        _run(["make", "-C", odir, "-s", "liboracle.so"])
        _run([CXX, *CXX_FLAGS, f"-I{odir}", "-o", os.path.join(LIB, "test_api"), test_src,
              f"-L{LIB}", "-lgenie", "-lgenie_hip", f"-L{odir}", "-loracle",
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/../../oracle"])
    return LIB


def build_oracle():
    """The CPU checker (oracle/liboracle.so) and, where the reference tree is
    present, the reference's own inv_list (oracle/_ref/)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-s", "liboracle.so"])
    if os.path.isdir("/root/reference/src/genie/table"):
        _run(["make", "-C", odir, "-s", "ref"])


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
    print("built", LIB)
