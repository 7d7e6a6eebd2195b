"""CPU checks of the boundary: libfos.so loads and exports every symbol
include/fos.h declares, and the ctypes table matches the header."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fos.h")
LIB = os.path.join(ROOT, "fabric-of-space-taichi-jfa_amd", "libfos.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fos_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("fos_jfa_init", "fos_jfa_step", "fos_jfa_passes", "fos_majority",
                 "fos_face_accumulate", "fos_finalize_faces", "fos_grid_build",
                 "fos_pbd_pass", "fos_pressure_iter", "fos_fsc_targets", "fos_nudge_radii"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfos.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.fos_abi_version.restype = ctypes.c_int
    assert lib.fos_abi_version() == 2


def test_ctypes_table_covers_header():
    import fos_lib
    names = set(declared()) - {"fos_abi_version", "fos_last_error"}
    assert names == set(fos_lib.SIGNATURES), names ^ set(fos_lib.SIGNATURES)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "fabric-of-space-taichi-jfa_amd")
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            src = open(os.path.join(pkg, fn)).read()
            assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, re.M), fn


def test_calls_fail_loudly_without_a_gpu():
    import torch
    import fos_lib
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(fos_lib.FosError):
        fos_lib.Context(10, 8, 27)


def test_bench_rejects_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with status 2 and says so
    (before touching a GPU)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
