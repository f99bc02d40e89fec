"""In-tree build of the native libraries (gfx950 HIP + host C++).

``python -m sem2dpack_amd.build`` compiles
  * ``sem2dpack_amd/lib/libsem2d_host.so`` from ``csrc/host_*.cpp`` (g++),
  * ``sem2dpack_amd/lib/libsem2d_hip.so``  from ``csrc/*.hip`` (hipcc, gfx950).
Every HIP source is compiled with ``-Xclang -target-feature -Xclang
-packed-fp32-ops`` so that hipcc emits no packed-FP32 instructions.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=fast", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build failed: " + os.path.basename(cmd[-1]))
    return r


def build_host():
    os.makedirs(LIB, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "host_*.cpp")))
    out = os.path.join(LIB, "libsem2d_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, "-o", out] + srcs)
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not srcs:
        return None
    out = os.path.join(LIB, "libsem2d_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_all():
    return [build_host(), build_hip()]


if __name__ == "__main__":
    for p in build_all():
        if p:
            print("built", p)
