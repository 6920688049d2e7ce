"""Build the gfx950 HIP library in-tree (siminf_amd/libsiminf_hip.so).

hipcc cross-compiles for gfx950 without a GPU.  -ffp-contract=off keeps
every double operation unfused so device results round exactly like the
host build of the same expressions; packed FP32 instructions are switched
off.

The library is several translation units: the host side with the
model-independent kernels (siminf_hip.hip) and one unit per built-in
model (builtin_*.hip), each instantiating the solver kernels for its
model.  They compile in parallel and link into one shared object.
"""
import concurrent.futures
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SRC = os.path.join(CSRC, "siminf_hip.hip")
OUT = os.path.join(HERE, "libsiminf_hip.so")
OBJ_DIR = os.path.join(HERE, "_obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# The models' translation units (node and coupled kernels) are scheduled
# for instruction-level parallelism: their wavefronts run one or two per
# SIMD, each a long dependent chain (the SSA step) beside independent work
# (the next attempt's Philox block and logarithm), so interleaving the two
# matters more than the register pressure the default scheduler minimises
# (measured on the coupled kernel: SISe3 1600 herds x 1460 days 138.0 to
# 132.2 ms per run).
ILP_FLAGS = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]


def unit_flags(src):
    """Compile flags of one translation unit."""
    base = os.path.basename(src)
    if base.startswith("builtin_"):
        return FLAGS + ILP_FLAGS
    return FLAGS


def sources():
    """The library's translation units, host side first."""
    return [SRC] + sorted(glob.glob(os.path.join(CSRC, "builtin_*.hip")))


def _headers():
    return glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(
        os.path.join(HERE, "..", "include", "*.h"))


def _compile(src, verbose, incremental=False):
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + ".o")
    if incremental and os.path.exists(obj):
        newest = max(os.path.getmtime(f) for f in [src] + _headers())
        if os.path.getmtime(obj) > newest:
            return obj
    cmd = [HIPCC] + unit_flags(src) + ["-c", "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-4000:]))
    return obj


def build(verbose=False, jobs=None, incremental=False):
    """Compile and link the library; raises RuntimeError or
    CalledProcessError on failure.  incremental skips units whose object
    is newer than their source and every header."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1), 16)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, incremental), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
