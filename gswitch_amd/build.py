"""Build libgswitch_amd.so (HIP, gfx950) and the CPU oracle.

    python gswitch_amd/build.py            # library, applications, oracle
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


APPS_DIR = os.path.join(REPO, "tests", "apps")
APPS = ["bfs_app", "sssp_app", "cc_app"]


def build_apps():
    """Applications written against the template API (gswitch.h) the way a
    GSWITCH application is written; they link the library for file input
    and find it through their rpath."""
    out_dir = os.path.join(APPS_DIR, "build")
    os.makedirs(out_dir, exist_ok=True)
    procs = []
    for app in APPS:
        cmd = [HIPCC, "-x", "hip", "--offload-arch=" + ARCH, "-O3", "-std=c++17",
               "-Wno-unused-value", "-Wno-unused-result",
               os.path.join(APPS_DIR, app + ".hip"), "-o", os.path.join(out_dir, app),
               "-I", os.path.join(CSRC, "gswitch"), "-I", os.path.join(REPO, "include"),
               "-L", HERE, "-lgswitch_amd", "-Wl,-rpath,$ORIGIN/../../../gswitch_amd"
               ] + DEVICE_FLAGS
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    if any(p.wait() for p in procs):
        raise RuntimeError("application build failed")


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all"])
    if os.path.isdir("/root/reference/src/model"):
        _run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"])


def main(argv):
    if "--oracle" not in argv:
        build_lib()
        if "--lib" not in argv:
            build_apps()
    if "--lib" not in argv:
        build_oracle()


if __name__ == "__main__":
    main(sys.argv[1:])
