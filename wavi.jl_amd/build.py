"""Build libwavi_hip.so for gfx950 with hipcc (in-tree, into wavi.jl_amd/lib/).

-ffp-contract=off keeps every a*b+c of the FP64 path as two roundings, as the
reference evaluates it; -packed-fp32-ops is turned off (hipcc's host pass
warns that it does not know the feature, which is harmless).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "wavi_api.hip")
OUT = os.path.join(HERE, "lib", "libwavi_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-pthread", "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libwavi_hip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
