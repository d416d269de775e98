"""Build the gfx950 scoring library (saber_amd/libsaber_hip.so) in-tree.

Every HIP source under saber_amd/csrc is compiled by hipcc for gfx950 (one
process per translation unit, in parallel) and linked into one shared
library exporting the C ABI of include/saber_hip.h.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsaber_hip.so")
WLIB = os.path.join(HERE, "libsaber_workload.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed fp32 VALU ops are kept off: the compiler would otherwise form
# v_pk_*_f32 from ordinary float code. Calls are never marked as tail calls:
# a non-inlined device function reached by a tail call has to save the
# callee-saved registers it uses on the stack, one spill per register and
# call (the per-shape bodies of the band and full-matrix kernels).
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fno-optimize-sibling-calls",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return obj


def _compile_host(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stderr[-4000:]}")
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    # host-only parts of the library (the host pipeline around the kernels)
    hsrcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.startswith("host_") and f.endswith(".cpp"))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs)) + list(ex.map(_compile_host, hsrcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    # Host-only synthetic workload generator (not on the scoring path).
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", WLIB + ".tmp",
                        os.path.join(CSRC, "workload.cpp")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"workload build failed:\n{r.stderr[-4000:]}")
    os.replace(WLIB + ".tmp", WLIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
