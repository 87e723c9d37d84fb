"""The device checker (sadcs_hip_check_device) against the oracle's checkers.

The planted mutations are the reference's checker selftest (src/cli/selftest.cpp:309-385): a
correct suffix array passes, and swapped neighbours, a far swap, a duplicate, an out-of-range and a
negative entry each fail. Beyond pass/fail, the first failing index must be the one the oracle's
fast check reports (src/common/checker.cpp:53-109), and for the BWT the first index where it
differs from the BWT read off the suffix array; random mutations widen that to many positions.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def device_check(gpu, text, sa, bwt=None, primary=-1):
    raw = np.frombuffer(text, dtype=np.uint8) if isinstance(text, bytes) else text
    d_text = gpu.DeviceBuffer.from_host(raw)
    d_sa = gpu.DeviceBuffer.from_host(np.ascontiguousarray(sa, dtype=np.int32))
    d_bwt = gpu.DeviceBuffer.from_host(np.frombuffer(bwt, dtype=np.uint8)) if bwt is not None else None
    return gpu.gpu_check_device(d_text.ptr, raw.size, d_sa.ptr, d_bwt.ptr if d_bwt else None, primary)


def test_reference_mutations(gpu):
    text = oracle.Rng(7).random_text(3000, 3)
    good = oracle.naive_sa(text)
    assert device_check(gpu, text, good) == (0, -1)

    def swap(i, j):
        def apply(sa):
            sa[[i, j]] = sa[[j, i]]
        return apply

    def assign(i, value):
        def apply(sa):
            sa[i] = value(sa)
        return apply

    mutations = [
        ("swapped neighbours", swap(1000, 1001), 3),
        ("swapped far apart", swap(10, 2900), 3),
        ("duplicate entry", assign(500, lambda sa: sa[1500]), 2),
        ("out of range", assign(42, lambda sa: sa.size), 1),
        ("negative entry", assign(0, lambda sa: -1), 1),
    ]
    for name, apply, category in mutations:
        print(f"mutation {name}", flush=True)
        bad = good.copy()
        apply(bad)
        want = oracle.check_sa(text, bad)
        assert want >= 0
        assert device_check(gpu, text, bad) == (category, want), name


def test_bwt_mutations(gpu):
    text = oracle.Rng(7).random_text(3000, 3)
    sa = oracle.naive_sa(text)
    good, primary = oracle.bwt_from_sa(text, sa)
    assert device_check(gpu, text, sa, good, primary) == (0, -1)
    arr = bytearray(good)
    i = next(k for k in range(1000, len(arr) - 1) if arr[k] != arr[k + 1])
    arr[i], arr[i + 1] = arr[i + 1], arr[i]
    swapped = bytes(arr)
    arr = bytearray(good)
    arr[1234] ^= 1
    changed = bytes(arr)
    cases = [
        ("swapped neighbours", swapped, primary, (5, i)),
        ("changed character", changed, primary, (5, 1234)),
        ("primary index off by one", good, primary + 1 if primary < 3000 else primary - 1, (4, -1)),
        ("primary index 0", good, 0, (4, -1)),
    ]
    for name, b, p, want in cases:
        print(f"BWT mutation {name}", flush=True)
        assert oracle.check_bwt(text, b, p) != -1, name
        assert device_check(gpu, text, sa, b, p) == want, name


def test_random_mutations(gpu):
    """Random swaps and overwrites at random places: the device verdict and first failing index
    equal the oracle's, for the suffix array and the BWT."""
    rng = np.random.default_rng(5)
    for case in range(24):
        n = int(rng.integers(2, 5000))
        sigma = int(rng.choice([1, 2, 4, 26, 256]))
        text = oracle.Rng(100 + case).random_text(n, sigma)
        sa = oracle.naive_sa(text)
        bwt, primary = oracle.bwt_from_sa(text, sa)
        bad = sa.copy()
        kind = case % 4
        i, j = (int(x) for x in rng.integers(0, n, 2))
        if kind == 0:
            bad[[i, j]] = bad[[j, i]]
        elif kind == 1:
            bad[i] = bad[j]
        elif kind == 2:
            bad[i] = int(rng.choice([n, n + 7, -1, -(1 << 31)]))
        else:
            k = min(n, i + 64)
            bad[i:k] = bad[i:k][::-1].copy()
        want = oracle.check_sa(text, bad)
        got = device_check(gpu, text, bad)
        print(f"case {case}: n={n} sigma={sigma} kind={kind} oracle={want} device={got}", flush=True)
        assert got[1] == want and (got[0] == 0) == (want < 0)
        # BWT: flip one byte (or none) against the correct suffix array.
        arr = bytearray(bwt)
        flip = int(rng.integers(0, n)) if kind != 3 else -1
        if flip >= 0:
            arr[flip] ^= 0x40
        got = device_check(gpu, text, sa, bytes(arr), primary)
        want_bwt = oracle.check_bwt(text, bytes(arr), primary)
        assert (got[0] == 0) == (want_bwt == -1)
        assert got == ((5, flip) if flip >= 0 else (0, -1))


def test_edge_lengths(gpu):
    for text in [b"a", b"ab", b"ba", b"aaaa", bytes(range(256))]:
        sa = oracle.naive_sa(text)
        bwt, primary = oracle.bwt_from_sa(text, sa)
        assert device_check(gpu, text, sa, bwt, primary) == (0, -1), text
    duplicate = np.array([0, 0], dtype=np.int32)
    assert oracle.check_sa(b"ab", duplicate) == 0
    assert device_check(gpu, b"ab", duplicate) == (2, 0)
    assert device_check(gpu, b"ab", np.array([1, 0], dtype=np.int32)) == (3, 0)


def test_check_of_a_device_build(gpu):
    """A 16 MiB device build passes the device check; one swap planted in the device suffix array
    fails at the index the oracle names."""
    from sadcs_amd import workloads
    text = workloads.markov(1 << 24, 19)
    n = text.size
    d_text = gpu.DeviceBuffer.from_host(text, extra=64)
    d_sa = gpu.DeviceBuffer(4 * n)
    d_bwt = gpu.DeviceBuffer(n)
    with gpu.GpuContext(n) as ctx:
        primary = gpu.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, gpu.GpuStats())
    assert gpu.gpu_check_device(d_text.ptr, n, d_sa.ptr, d_bwt.ptr, primary) == (0, -1)
    sa = d_sa.to_host(np.int32)
    assert oracle.check_sa(text.tobytes(), sa) == -1
    bad = sa.copy()
    bad[[123_456, 9_000_001]] = bad[[9_000_001, 123_456]]
    want = oracle.check_sa(text.tobytes(), bad)
    d_bad = gpu.DeviceBuffer.from_host(bad)
    assert gpu.gpu_check_device(d_text.ptr, n, d_bad.ptr) == (3, want)


def test_largest_text_passes_the_device_check(gpu):
    """The largest text the library takes (2^31 - 1025 bytes: positions near 2^31 in 32-bit
    entries, 8 GiB suffix array): SA + BWT on the device, every row checked there."""
    from sadcs_amd import workloads
    n = (1 << 31) - 1025
    text = workloads.markov(n, 99, padding=64)
    d_text = gpu.DeviceBuffer.from_host(text)
    del text
    d_sa = gpu.DeviceBuffer(4 * n)
    d_bwt = gpu.DeviceBuffer(n)
    with gpu.GpuContext(n) as ctx:
        stats = gpu.GpuStats()
        primary = gpu.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, stats)
    print(f"n={n}: device {stats.device_seconds * 1e3:.1f} ms, peak {stats.device_bytes / n:.2f} B/byte", flush=True)
    # The 10n budget: text, SA and BWT (6n, the caller's) plus the library's own 4n (the class lists
    # at the peak), and a few KiB of fixed tables (text padding, segment table, error flags).
    assert stats.device_bytes <= 10 * n + (64 << 10)
    assert gpu.gpu_check_device(d_text.ptr, n, d_sa.ptr, d_bwt.ptr, primary) == (0, -1)


def test_gather_calibration_runs(gpu):
    """The known-size line gather used to scale HBM counters runs in both modes on a small array."""
    lines = gpu.DeviceBuffer(1 << 20)
    lines.zero()
    for whole in (True, False):
        gpu.gpu_gather_lines(lines.ptr, (1 << 20) // 128, 100_000, whole)
    lines.free()
