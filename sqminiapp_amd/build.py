"""Build libsqamd.so (host setup + gfx950 kernels + C ABI + the reference-named
host layer) and the sq_main driver in-tree with hipcc.

    python -m sqminiapp_amd.build          # compile every time
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsqamd.so")
EXE = os.path.join(HERE, "sq_main")
ARCH = os.environ.get("SQ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["sq_setup.cpp", "sq_api.cpp", "sq_lsdft.cpp", "sq_kernels.hip", "sq_zmarch.hip", "sq_scf.hip", "sq_setup.hip",
           "sq_forces.hip", "sq_mlp.hip", "sq_zbatch.hip", "sq_scfstep.hip"]
DRIVER = "sq_main.cpp"

# -ffp-contract=off: products and sums round as written (the FP64 form follows
# the reference's operation order; the FP32 form asks for fma explicitly).
# Packed FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-result"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; libsqamd.so cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        jobs.append(([cc, *COMMON, "-x", "hip", *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        list(ex.map(lambda j: _run(j[0], verbose), jobs))
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", *DEVICE, "-o", tmp, *(o for _, o in jobs), "-lpthread"], verbose)
    os.replace(tmp, LIB)
    # the driver is host code only; it finds libsqamd.so next to itself
    tmp = EXE + ".tmp"
    _run([cc, *COMMON, os.path.join(CSRC, DRIVER), "-o", tmp, "-L" + HERE, "-lsqamd", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
