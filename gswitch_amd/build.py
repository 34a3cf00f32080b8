"""Build libgswitch_amd.so (HIP, gfx950) and the CPU oracle.

    python gswitch_amd/build.py            # library + oracle
    python gswitch_amd/build.py --lib      # library only

The library is compiled in-tree with hipcc for --offload-arch=gfx950 so the
built .so travels with the source tree. Packed-FP32 instructions are disabled
for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgswitch_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GSW_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    os.path.join(CSRC, "capi.hip"),
    os.path.join(CSRC, "host", "graphgen.cpp"),
    os.path.join(CSRC, "host", "mtx.cpp"),
]

DEVICE_FLAGS = [
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib(jobs=None):
    objs = []
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
        objs.append(obj)
        if src.endswith(".hip"):
            cmd = [HIPCC, "-x", "hip", "--offload-arch=" + ARCH, "-O3",
                   "-std=c++17", "-fPIC", "-Wno-unused-value", "-Wno-unused-result", "-c", src, "-o", obj,
                   "-I", CSRC, "-I", os.path.join(REPO, "include")] + DEVICE_FLAGS
        else:
            cmd = [HIPCC, "-x", "c++", "-O3", "-std=c++17", "-fPIC", "-c", src,
                   "-o", obj, "-I", CSRC]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    rc = [p.wait() for p in procs]
    if any(rc):
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all"])
    if os.path.isdir("/root/reference/src/model"):
        _run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"])


def main(argv):
    if "--oracle" not in argv:
        build_lib()
    if "--lib" not in argv:
        build_oracle()


if __name__ == "__main__":
    main(sys.argv[1:])
