"""sadcs_hip_check_device_noalloc (the device check without allocation) against
sadcs_hip_check_device, through the C ABI: the same verdict and index on correct outputs and on
planted errors, over two calls in a row with the second input smaller, and the size limit."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

LIMIT = 1 << 23


def verdicts(gpu, d_text, n, d_sa, d_bwt, primary):
    both = []
    for name in ("sadcs_hip_check_device", "sadcs_hip_check_device_noalloc"):
        result = (ctypes.c_int64 * 2)(-7, -7)
        status = getattr(gpu.lib, name)(ctypes.c_void_p(d_text), n, ctypes.c_void_p(d_sa),
                                        ctypes.c_void_p(d_bwt) if d_bwt else None, primary, result)
        assert status == 0, f"{name}: {gpu.gpu_last_error()}"
        both.append((result[0], result[1]))
    return both


def cases(text, sa, bwt, primary):
    """(label, sa, bwt, primary): the outputs, then one planted error of each kind."""
    n = sa.size
    yield "correct", sa, bwt, primary
    swapped = sa.copy()
    swapped[[n // 3, n // 3 + 1]] = swapped[[n // 3 + 1, n // 3]]
    yield "neighbours swapped", swapped, bwt, primary
    repeated = sa.copy()
    repeated[n // 2] = repeated[n // 5]
    yield "entry repeated", repeated, bwt, primary
    outside = sa.copy()
    outside[n - 2] = n + 5
    yield "entry out of range", outside, bwt, primary
    changed = bwt.copy()
    changed[n // 7] ^= 1
    yield "BWT byte changed", sa, changed, primary
    yield "primary moved", sa, bwt, primary + 1 if primary < n else primary - 1


def test_noalloc_check_equals_allocating_check(gpu):
    for n, seed in [(1_000_003, 3), (70_001, 4)]:  # the second call smaller than the first
        text = oracle.Rng(seed).random_text(n, 4)
        sa, bwt, primary = gpu.gpu_build(text, True, True)
        d_text = gpu.DeviceBuffer.from_host(np.frombuffer(text, dtype=np.uint8), extra=64)
        for label, s, b, p in cases(text, sa, np.frombuffer(bwt, dtype=np.uint8), primary):
            print(f"n={n}: {label}", flush=True)
            d_sa = gpu.DeviceBuffer.from_host(np.ascontiguousarray(s, dtype=np.int32))
            d_bwt = gpu.DeviceBuffer.from_host(np.ascontiguousarray(b, dtype=np.uint8))
            allocating, noalloc = verdicts(gpu, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, p)
            assert noalloc == allocating, (n, label)
            if label == "correct":
                assert noalloc == (0, -1)
            else:
                assert noalloc[0] != 0, (n, label)
            without_bwt = verdicts(gpu, d_text.ptr, n, d_sa.ptr, None, p)
            assert without_bwt[0] == without_bwt[1], (n, label, "no BWT")
            d_sa.free()
            d_bwt.free()
        d_text.free()


def test_noalloc_check_refuses_beyond_its_limit(gpu):
    n = LIMIT + 1
    d_text = gpu.DeviceBuffer(n + 64)
    d_sa = gpu.DeviceBuffer(4 * n)
    result = (ctypes.c_int64 * 2)(-7, -7)
    status = gpu.lib.sadcs_hip_check_device_noalloc(ctypes.c_void_p(d_text.ptr), n, ctypes.c_void_p(d_sa.ptr), None, 0, result)
    assert status == gpu.GpuStatus.NotEnoughMemory
    assert "at most 8388608" in gpu.gpu_last_error()
    d_text.free()
    d_sa.free()
