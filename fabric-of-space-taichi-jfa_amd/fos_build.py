"""Builds libfos.so (HIP, gfx950) in place from csrc/*.hip.

Each source is compiled to an object in its own hipcc process (in
parallel), then the objects are linked into the shared library.

Flags that carry the numerics contract:
  -ffp-contract=off   no FMA contraction: each fp32 expression rounds as written
  -packed-fp32-ops    off: scalar fp32 only (no v_pk_* fp32 instructions)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfos.so")
OBJ = os.path.join(HERE, "_build")
SOURCES = ["fos_ctx.hip", "fos_jfa.hip", "fos_grid.hip", "fos_dynamics.hip", "fos_topology.hip",
           "fos_predict.hip", "fos_batch.hip", "fos_static.hip",
           "fos_frame.hip"]
ARCH = os.environ.get("FOS_OFFLOAD_ARCH", "gfx950")
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = os.path.join(rocm, "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False, jobs=None):
    """Compiles every HIP source into libfos.so (always rebuilds)."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[hipcc()] + FLAGS + ["-c", "-o", o, s] for s, o in zip(srcs, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        procs = list(ex.map(lambda c: _run(c, verbose), cmds))
    for cmd, proc in zip(cmds, procs):
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError(f"hipcc failed compiling {cmd[-1]} (exit {proc.returncode})")
    tmp = LIB + ".tmp"
    proc = _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed linking {LIB} (exit {proc.returncode})")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
