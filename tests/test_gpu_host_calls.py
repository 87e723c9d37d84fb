"""Host-memory calls large enough that the top-level merge runs in row slices and the outputs are
downloaded while later slices merge (more than 2^24 output rows).

Whatever the slicing, a host call must return exactly what a device-resident build leaves in HBM
(itself checked row by row on the device, and against the reference CPU build for the 40 MiB
text): for the SA and the BWT together and alone, into
pageable and into pinned host memory, for cover 32 and cover 128, and for a batch of texts whose
slices cut across text boundaries (the reference's host API: src/cuda/gpusa.h:101-108).
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

N = 40 << 20  # three slices of the top-level merge


@pytest.fixture(scope="module")
def reference_build(gpu):
    """The text and the device-resident build's outputs (checked on the device)."""
    from sadcs_amd import workloads
    text = workloads.markov(N, 23, padding=64)
    d_text = gpu.DeviceBuffer.from_host(text)
    d_sa = gpu.DeviceBuffer(4 * N)
    d_bwt = gpu.DeviceBuffer(N)
    with gpu.GpuContext(N) as ctx:
        primary = gpu.gpu_build_device(ctx, d_text.ptr, N, d_sa.ptr, d_bwt.ptr)
    assert gpu.gpu_check_device(d_text.ptr, N, d_sa.ptr, d_bwt.ptr, primary) == (0, -1)
    sa, bwt = d_sa.to_host(np.int32), d_bwt.to_host(np.uint8).tobytes()
    for buf in (d_text, d_sa, d_bwt):
        buf.free()
    # And equal to the reference CPU build of the same text, so the host calls below are compared
    # with the reference as well.
    want_sa = oracle.ref_suffix_array(text[:N].tobytes())
    assert np.array_equal(sa, want_sa)
    assert (bwt, primary) == oracle.bwt_from_sa(text[:N].tobytes(), want_sa)
    return text[:N].copy(), sa, bwt, primary


def test_pageable_outputs(gpu, reference_build):
    text, sa, bwt, primary = reference_build
    stats = gpu.GpuStats()
    got_sa, got_bwt, got_primary = gpu.gpu_build(text, stats=stats)
    print(f"host call: upload {stats.upload_seconds * 1e3:.1f} ms, device {stats.device_seconds * 1e3:.1f} ms, "
          f"download {stats.download_seconds * 1e3:.1f} ms", flush=True)
    assert np.array_equal(got_sa, sa)
    assert got_bwt == bwt and got_primary == primary
    assert np.array_equal(gpu.gpu_suffix_array(text), sa)
    assert gpu.gpu_bwt(text) == (bwt, primary)


def test_pinned_outputs(gpu, reference_build):
    text, sa, bwt, primary = reference_build
    sa_ptr, sa_view = gpu.gpu_allocate_pinned(4 * N)
    bwt_ptr, bwt_view = gpu.gpu_allocate_pinned(N)
    assert sa_ptr and bwt_ptr
    try:
        for want_sa, want_bwt in [(True, True), (True, False), (False, True)]:
            print(f"pinned outputs: sa={want_sa} bwt={want_bwt}", flush=True)
            sa_view[:] = 0xEE
            bwt_view[:] = 0xEE
            got_primary = ctypes.c_int32(-1)
            stats = gpu.GpuStats()
            status = gpu.lib.sadcs_hip_build(None, text.ctypes.data, N, sa_ptr if want_sa else None,
                                             bwt_ptr if want_bwt else None, ctypes.byref(got_primary), ctypes.byref(stats))
            assert status == 0, gpu.gpu_last_error()
            if want_sa:
                assert np.array_equal(sa_view.view(np.int32), sa)
            if want_bwt:
                assert bwt_view.tobytes() == bwt and got_primary.value == primary
    finally:
        gpu.gpu_free_pinned(sa_ptr)
        gpu.gpu_free_pinned(bwt_ptr)


def test_cover_128_context(gpu, reference_build):
    text, sa, bwt, primary = reference_build
    with gpu.GpuContext(N, cover=128) as ctx:
        got_sa, got_bwt, got_primary = gpu.gpu_build(text, context=ctx)
        assert np.array_equal(got_sa, sa)
        assert got_bwt == bwt and got_primary == primary
        # The same context again, for a text too short to slice.
        short = text[:1 << 20]
        assert np.array_equal(gpu.gpu_suffix_array(short, context=ctx), gpu.gpu_suffix_array(short))


def test_batch_slices_cut_across_texts(gpu):
    """A 48 MiB batch of 23 texts: slice boundaries fall inside texts, each text's BWT shifted by
    its own first row. Every text's outputs equal its single-text build and the reference CPU
    build's."""
    from sadcs_amd import workloads
    rng = np.random.default_rng(11)
    lengths = [int(x) for x in rng.integers(1 << 20, 3 << 20, 23)]
    texts = [workloads.markov(n, 200 + i) if i % 3 else workloads.dna_like(n, 300 + i) for i, n in enumerate(lengths)]
    print(f"batch of {len(texts)} texts, {sum(lengths)} bytes", flush=True)
    results = gpu.gpu_batch_build(texts)
    reference = oracle.ref_outputs([t.tobytes() for t in texts])
    for i, (t, (sa, bwt, primary), (ref_sa, ref_bwt, ref_primary)) in enumerate(zip(texts, results, reference)):
        want_sa, want_bwt, want_primary = gpu.gpu_build(t)
        assert np.array_equal(sa, want_sa), i
        assert bwt == want_bwt and primary == want_primary, i
        assert np.array_equal(sa, ref_sa), i
        assert (bwt, primary) == (ref_bwt, ref_primary), i


def test_two_slices_odd_length(gpu):
    """Just over 2^24 rows (two slices, the second short) and an odd length: SA, BWT and primary
    from a host call equal the device-resident build's, for both outputs and for each alone."""
    from sadcs_amd import workloads
    n = (1 << 24) + 12345
    text = workloads.dna_like(n, 31, padding=64)
    d_text = gpu.DeviceBuffer.from_host(text)
    d_sa = gpu.DeviceBuffer(4 * n)
    d_bwt = gpu.DeviceBuffer(n)
    with gpu.GpuContext(n) as ctx:
        primary = gpu.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr)
    assert gpu.gpu_check_device(d_text.ptr, n, d_sa.ptr, d_bwt.ptr, primary) == (0, -1)
    sa, bwt = d_sa.to_host(np.int32), d_bwt.to_host(np.uint8).tobytes()
    host = text[:n].copy()
    got_sa, got_bwt, got_primary = gpu.gpu_build(host)
    assert np.array_equal(got_sa, sa)
    assert got_bwt == bwt and got_primary == primary
    assert np.array_equal(gpu.gpu_suffix_array(host), sa)
    assert gpu.gpu_bwt(host) == (bwt, primary)
