"""Build librfhip.so (HIP kernels + C ABI + host builder) for gfx950, in-tree.

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops). Objects go to
rapidfem_amd/_build/, the library to rapidfem_amd/librfhip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "librfhip.so")

HIP_SOURCES = ["rfhip_api.hip", "rfhip_plain.hip", "rfhip_fused.hip", "rfhip_vec.hip", "rfhip_debye.hip",
               "rfhip_sparse.hip", "rfhip_xfer.hip", "rfhip_mlp.hip", "rfhip_noalloc.hip"]
CXX_SOURCES = ["host_builder.cpp", "krylov_host.cpp"]

ARCH = os.environ.get("RFHIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build librfhip.so")


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    cc = hipcc()
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    common = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
    hip_flags = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                 "-munsafe-fp-atomics"]
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([cc, "-x", "hip", *hip_flags, *common, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(OUT_DIR, src.replace(".cpp", ".o"))
        objs.append(obj)
        jobs.append(["g++", *common, "-ffp-contract=off", *inc, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for log in ex.map(_compile, jobs):
            if verbose and log.strip():
                print(log, file=sys.stderr)
    tmp = LIB + ".tmp"
    _compile([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    print(build(verbose="-v" in sys.argv))
