"""Build the gfx950 HIP library in-tree (siminf_amd/libsiminf_hip.so).

hipcc cross-compiles for gfx950 without a GPU.  -ffp-contract=off keeps
every double operation unfused so device results round exactly like the
host build of the same expressions; packed FP32 instructions are switched
off.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "siminf_hip.hip")
OUT = os.path.join(HERE, "libsiminf_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    """Compile the library; raises CalledProcessError on failure."""
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=None if verbose else subprocess.DEVNULL,
                   stderr=subprocess.PIPE if not verbose else None)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
