"""The C ABI without a GPU: the library loads, exports every symbol include/sadcs_hip.h declares,
plans memory, and reports Unavailable (not a crash) where there is no device."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import sadcs_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sadcs_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sadcs_hip_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(sadcs_amd.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sadcs_amd.LIBRARY_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_sample_size_matches_cover():
    lib = oracle.oracle_lib()
    for n in [0, 1, 2, 31, 32, 33, 1000, 1 << 24, 2**31 - 1025]:
        assert sadcs_amd.gpu_sample_size(n, 32) == (lib.oracle_pack(32, n) if n > 0 else 0)
        assert sadcs_amd.gpu_sample_size(n, 128) == (lib.oracle_pack(128, n) if n > 0 else 0)


def test_memory_plan_within_ten_bytes_per_input_byte():
    for n in [1 << 20, 1 << 24, 1 << 28, 1 << 30]:
        per_byte = sadcs_amd.gpu_required_bytes(n) / n
        assert 8.0 < per_byte <= 10.0, (n, per_byte)
    assert sadcs_amd.gpu_required_bytes(0) == 0


def test_batch_plan_positive():
    lengths = np.array([1000, 5000, 1 << 20], dtype=np.int64)
    plan = sadcs_amd.lib.sadcs_hip_batch_required_bytes(lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 3)
    assert plan > 10 * int(lengths.sum())


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK), reason="a GPU is present")
def test_no_device_is_unavailable():
    status, _ = sadcs_amd.gpu_query_device()
    assert status == sadcs_amd.GpuStatus.Unavailable
    assert sadcs_amd.gpu_build_status(b"banana") == sadcs_amd.GpuStatus.Unavailable
    with pytest.raises(sadcs_amd.GpuError) as info:
        sadcs_amd.GpuContext(1000)
    assert info.value.status == sadcs_amd.GpuStatus.Unavailable
    # n <= 1 needs no device work but still reports the device status, as gpusa.cu:528-542.
    assert sadcs_amd.gpu_build_status(b"", n=0) == sadcs_amd.GpuStatus.Unavailable


def test_device_call_memory_plan():
    """A device-resident fused build needs 4n of its own (the class lists, live together at the
    peak) beyond the caller's text (n + 64), SA (4n) and BWT (n) buffers: the caller's outputs serve
    as scratch before the merge (the tile-row tables in the BWT buffer's tail), so the total is 10
    bytes per input byte plus a few KiB of fixed tables."""
    for n in (1 << 20, 1 << 26, 1 << 30, (1 << 31) - 1025):
        own = sadcs_amd.lib.sadcs_hip_device_required_bytes(n, 1, 1)
        assert 0 < own <= 4 * n + (64 << 10), (n, own / n)


def test_gather_calibration_refuses_bad_arguments():
    """sadcs_hip_gather_lines checks its arguments before any device work."""
    lib = sadcs_amd.lib
    assert lib.sadcs_hip_gather_lines(None, 1024, 10, 1) == sadcs_amd.GpuStatus.Failed
    assert lib.sadcs_hip_gather_lines(ctypes.c_void_p(16), 0, 10, 1) == sadcs_amd.GpuStatus.Failed
    assert lib.sadcs_hip_gather_lines(ctypes.c_void_p(16), 1 << 32, 10, 1) == sadcs_amd.GpuStatus.Failed
    assert "gather_lines" in sadcs_amd.gpu_last_error()
