"""Build recipe of the phx_elph HIP library (gfx950), in-tree.

    python -m phoebe_amd.build

compiles phoebe_amd/csrc/*.hip with hipcc and the host-only
phoebe_amd/csrc/*.cpp (the el-ph file reader / writer) with g++ into
phoebe_amd/libphx_elph.so.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libphx_elph.so")
ARCH = os.environ.get("PHX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def build(verbose: bool = False) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objs = []
    jobs = []
    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    for s in srcs:
        o = os.path.join(odir, os.path.basename(s) + ".o")
        cmd = ["hipcc", "-c", *[f for f in FLAGS if f != "-shared"], s, "-o", o]
        jobs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
        objs.append(o)
    for s in sorted(glob.glob(os.path.join(HERE, "csrc", "*.cpp"))):
        o = os.path.join(odir, os.path.basename(s) + ".o")
        cmd = ["g++", "-c", *HOST_FLAGS, s, "-o", o]
        jobs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
        objs.append(o)
        srcs = srcs + [s]
    err = None
    for j, s in zip(jobs, srcs):
        out, _ = j.communicate()
        if verbose and out:
            print(out)
        if j.returncode != 0:
            err = f"compile failed on {s}:\n{out}"
    if err:
        raise RuntimeError(err)
    tmp = LIB + ".tmp"
    subprocess.run(["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
                    "-lpthread"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
