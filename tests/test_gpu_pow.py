"""The pressure kernels' x^(1/3) fast path (a Newton cube root with an
unambiguous-rounding test, falling back to the double pow) against the double
pow it replaces, on the GPU: it must never return another float.  dynamics.py:
1540 (`Vi_new ** (1.0 / P_exp)`) is the reference's use of the power."""
import struct

import pytest
import torch

pytestmark = pytest.mark.gpu


def bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def check(gpu, first, count, stride):
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    gpu.call("fos_pow_third_check", first, count, stride, gpu.ptr(cnt), gpu.stream())
    torch.cuda.synchronize()
    bad, slow = (int(v) for v in cnt.cpu())
    return bad, slow


def test_pow_third_fast_path_over_every_positive_float(gpu):
    """All 2^31 positive float patterns (zero, denormals, infinity and NaNs
    included), in eight launches: no argument gets another float than the
    double pow's; the fast path refuses the 2^24 non-normal patterns and
    only the rare ambiguous roundings besides."""
    bad = slow = 0
    for c in range(8):
        b, s = check(gpu, c << 28, 1 << 28, 1)
        print(f"[pow third] patterns {c << 28:#010x}.. bad={b} slow={s}", flush=True)
        bad, slow = bad + b, slow + s
    assert bad == 0
    non_normal = 2 * (1 << 23)
    assert non_normal <= slow < non_normal + (1 << 31) // 100000


def test_pow_third_fast_path_over_the_pressure_range(gpu):
    """Every float in [1e-9, 1e-3) (the volumes r^3 of radii 1e-3 .. 0.1)."""
    first, last = bits(1e-9), bits(1e-3)
    bad, slow = check(gpu, first, last - first, 1)
    print(f"[pow third] range {last - first} bad={bad} slow={slow}", flush=True)
    assert bad == 0 and slow < (last - first) // 100000


def test_pow_third_rejects_patterns_past_the_positive_floats(gpu):
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    with pytest.raises(gpu.FosError):
        gpu.call("fos_pow_third_check", 0x7fffffff, 2, 1, gpu.ptr(cnt), gpu.stream())
