"""Build the native libraries in-tree.

  dynamap_amd/_lib/libdynamap_hip.so  HIP kernels + C ABI (gfx950)
  dynamap_amd/_lib/libdm_synth.so     synthetic RGB-D generator (host, OpenMP)

Every HIP translation unit is compiled with
  -ffp-contract=off                           no FMA contraction (FP32 results
                                              reproducible on the CPU)
  -fhip-fp32-correctly-rounded-divide-sqrt    IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                              no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "_lib")
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["volume.hip", "track.hip", "scan.hip", "preprocess.hip", "mesh.hip", "nonrigid.hip", "api.hip", "server.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-but-set-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(jobs=8):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        cmd = ["hipcc", *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", path, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
        if len(procs) >= jobs:
            p, s = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s}")
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s}")
    out = os.path.join(LIB, "libdynamap_hip.so")
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libdm_synth.so")
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-Wall",
          os.path.join(HERE, "synth", "synth.c"), "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_cpp_tests():
    """The C++ program that drives the reference-shaped classes
    (tests/cpp/test_builder.cpp), linked against the HIP library."""
    src = os.path.join(ROOT, "tests", "cpp", "test_builder.cpp")
    if not os.path.exists(src):
        return None
    out_dir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "test_builder")
    _run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", LIB, "-ldynamap_hip",
          "-Wl,-rpath," + LIB, "-Wl,-rpath,$ORIGIN/../../../dynamap_amd/_lib", "-o", out])
    return out


def build_all():
    build_synth()
    build_hip()
    build_cpp_tests()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "synth" in what:
            build_synth()
        if "hip" in what:
            build_hip()
