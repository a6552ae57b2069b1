"""Build liblbm.so (gfx950 HIP kernels + C-ABI) in-tree with hipcc.

Every call recompiles.  Flags:
  --offload-arch=gfx950               CDNA4 only
  -ffp-contract=off                   no implicit FMA contraction (fused steps must equal single steps,
                                      and the FP32 CPU mirror must reproduce the kernels bit for bit)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                      no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
LIB_PATH = os.path.join(HERE, "liblbm.so")

SOURCES = ["lbm_step.hip", "lbm_fused.hip", "lbm_fused_f32.hip", "lbm_fused_f16.hip", "lbm_fused_bf16.hip",
           "lbm_fused_sums.hip", "lbm_ensemble_sums.hip", "lbm_ensemble_fast.hip",
           "lbm_ensemble.hip", "lbm_mlp.hip", "lbm_closure.hip",
           "lbm_pair.hip", "lbm_pair_rb.hip", "lbm_quad.hip", "lbm_selftest.hip", "lbm_probe.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose: bool = False, extra_flags=None, jobs=None) -> str:
    """Compile every source to an object in parallel, then link liblbm.so."""
    import concurrent.futures as cf
    objdir = os.path.join(PKG_ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"] + list(extra_flags or [])

    def compile_one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc()] + flags + ["-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        return obj

    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking liblbm.so")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
