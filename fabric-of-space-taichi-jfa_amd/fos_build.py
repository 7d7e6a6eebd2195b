"""Builds libfos.so (HIP, gfx950) in place from csrc/*.hip.

Flags that carry the numerics contract:
  -ffp-contract=off   no FMA contraction: each fp32 expression rounds as written
  -packed-fp32-ops    off: scalar fp32 only (no v_pk_* fp32 instructions)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfos.so")
SOURCES = ["fos_ctx.hip", "fos_jfa.hip", "fos_grid.hip", "fos_dynamics.hip", "fos_topology.hip",
           "fos_predict.hip", "fos_batch.hip"]
ARCH = os.environ.get("FOS_OFFLOAD_ARCH", "gfx950")


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def build(verbose=False):
    """Compiles every HIP source into libfos.so (always rebuilds)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-o", tmp,
    ] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed building {LIB} (exit {proc.returncode})")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
