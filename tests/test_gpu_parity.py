"""GPU parity: the HIP construction against the oracle, byte for byte.

The cases are the reference's GPU selftest (src/cli/selftest.cpp:432-606), with its generators
and seeds: every suffix array must equal the reference CPU build's (oracle/_ref) or, for small
inputs, the naive suffix array of the oracle; every BWT must equal the BWT read off that suffix
array (selftest.cpp:141-159), built in place of a text copy as `build --bwt --gpu` does; sample
ranks must equal the ranks the suffix array implies (gpusa.h:62-73).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

LARGE = 1 << 22


def expected_sa(text):
    if len(text) <= 5000 or not oracle.ref_available():
        return oracle.naive_sa(text) if len(text) <= 20000 else oracle.suffix_array(text)
    return oracle.ref_suffix_array(text)


def check_case(gpu, name, text, context=None, cover=None):
    print(f"case {name} (n={len(text)})", flush=True)
    sa = expected_sa(text)
    got = gpu.gpu_suffix_array(text, context=context)
    assert got.shape == sa.shape, name
    mismatch = np.flatnonzero(got != sa)
    assert mismatch.size == 0, f"{name}: SA differs first at {mismatch[0]}"
    want_bwt, want_primary = oracle.bwt_from_sa(text, sa)
    b, primary = gpu.gpu_bwt(text, context=context)
    assert primary == want_primary, f"{name}: primary {primary}, expected {want_primary}"
    assert b == want_bwt, f"{name}: BWT differs"


def check_ranks(gpu, name, text):
    print(f"ranks {name} (n={len(text)})", flush=True)
    sa = expected_sa(text)
    want = oracle.sample_ranks_from_sa(sa, 32)
    got = gpu.gpu_rank_sample(text)
    assert np.array_equal(got, want), f"{name}: sample ranks differ"


def test_words(gpu):
    for word in [b"", b"a", b"banana", b"mississippi", b"ab", b"ba", b"aa", b"aab", b"abracadabra", b"yabbadabbado"]:
        check_case(gpu, repr(word), word)


def test_golden_banana(gpu):
    sa = gpu.gpu_suffix_array(b"banana")
    assert sa.tolist() == [5, 3, 1, 0, 4, 2]
    assert gpu.gpu_bwt(b"banana") == (b"annbaa", 4)
    assert gpu.gpu_bwt(b"a") == (b"a", 1)
    assert gpu.gpu_bwt(b"") == (b"", 0)


def test_structured_small(gpu):
    check_case(gpu, "run of a", b"a" * 1000)
    check_case(gpu, "zeros", b"\0" * 1000)
    check_case(gpu, "fibonacci", oracle.fibonacci_text(5000))
    check_case(gpu, "fibonacci 0/1", oracle.fibonacci_text(5000, 0, 1))
    check_case(gpu, "thue-morse", oracle.thue_morse_text(4096))
    check_case(gpu, "run then b", b"a" * 1999 + b"b")
    check_case(gpu, "b then run", b"b" + b"a" * 1999)


def test_periods(gpu):
    rng = oracle.Rng(3)
    for period in [1, 2, 3, 5, 7, 31, 32, 33, 63, 64, 65, 1000]:
        check_case(gpu, f"period {period}", rng.periodic_text(5000, period))


def test_all_lengths_to_80(gpu):
    rng = oracle.Rng(3)
    for n in range(0, 81):
        check_case(gpu, f"run of a n={n}", b"a" * n)
        check_case(gpu, f"random sigma=2 n={n}", rng.random_text(n, 2))


def test_sample_ranks_small(gpu):
    rng = oracle.Rng(3)
    for n in range(1, 81):
        check_ranks(gpu, f"run of a n={n}", b"a" * n)
        check_ranks(gpu, f"random sigma=2 n={n}", rng.random_text(n, 2))
    for period in [1, 2, 3, 7, 31, 32, 33, 64, 65, 1000]:
        check_ranks(gpu, f"period {period}", rng.periodic_text(5000, period))
    check_ranks(gpu, "banana", b"banana")
    check_ranks(gpu, "fibonacci 0/1", oracle.fibonacci_text(5000, 0, 1))
    check_ranks(gpu, "thue-morse", oracle.thue_morse_text(4096))
    check_ranks(gpu, "random sigma=256", rng.random_text(5000, 256))


def test_random_alphabets(gpu):
    rng = oracle.Rng(3)
    for sigma in [1, 2, 4, 26, 256]:
        for i in range(10):
            n = 1 + rng.next_below(20000)
            check_case(gpu, f"random sigma={sigma} #{i}", rng.random_text(n, sigma))


@pytest.mark.skipif(not oracle.ref_available(), reason="needs the reference build in oracle/_ref")
def test_large(gpu):
    rng = oracle.Rng(3)
    check_case(gpu, "large run of a", b"a" * LARGE)
    check_case(gpu, "large period 33", rng.periodic_text(LARGE, 33))
    check_case(gpu, "large fibonacci", oracle.fibonacci_text(LARGE))
    check_case(gpu, "large thue-morse", oracle.thue_morse_text(LARGE))
    check_case(gpu, "large random sigma=4", rng.random_text(LARGE, 4))
    check_case(gpu, "large random sigma=256", rng.random_text(LARGE, 256))


@pytest.mark.skipif(not oracle.ref_available(), reason="needs the reference build in oracle/_ref")
def test_large_sample_ranks(gpu):
    rng = oracle.Rng(5)
    check_ranks(gpu, "large run of a", b"a" * LARGE)
    check_ranks(gpu, "large period 33", rng.periodic_text(LARGE, 33))
    check_ranks(gpu, "large fibonacci", oracle.fibonacci_text(LARGE))
    check_ranks(gpu, "large random sigma=4", rng.random_text(LARGE, 4))


@pytest.mark.skipif(not oracle.ref_available(), reason="needs the reference build in oracle/_ref")
def test_context_reuse(gpu):
    """selftest.cpp:505-541: one context, largest input first, then smaller ones in dirty memory;
    an input over the capacity is refused and the context still works."""
    capacity = 20_000_000
    rng = oracle.Rng(3)
    with gpu.GpuContext(capacity) as ctx:
        check_case(gpu, "context random sigma=4 n=20M", rng.random_text(capacity, 4), ctx)
        check_case(gpu, "context large random sigma=256", rng.random_text(LARGE, 256), ctx)
        check_case(gpu, "context banana", b"banana", ctx)
        check_case(gpu, "context empty", b"", ctx)
        check_case(gpu, "context a", b"a", ctx)
        check_case(gpu, "context large fibonacci", oracle.fibonacci_text(LARGE), ctx)
        check_case(gpu, "context period 33", rng.periodic_text(5000, 33), ctx)
        for i in range(5):
            n = 1 + rng.next_below(20000)
            check_case(gpu, f"context random sigma=2 #{i}", rng.random_text(n, 2), ctx)
        too_large = np.full(capacity + 1, ord("a"), dtype=np.uint8)
        assert gpu.gpu_build_status(too_large, context=ctx) == gpu.GpuStatus.NotEnoughMemory
        check_case(gpu, "context after refusal", rng.random_text(5000, 26), ctx)


def test_fused_sa_bwt(gpu):
    """SA and BWT from one merge (in-place tile-major mode) equal the separate builds."""
    rng = oracle.Rng(11)
    for n in [2, 3, 100, 5000, 100_000, 1_000_000]:
        text = rng.random_text(n, 4)
        print(f"fused n={n}", flush=True)
        sa, b, primary = gpu.gpu_build(text, True, True)
        want = expected_sa(text)
        assert np.array_equal(sa, want)
        assert (b, primary) == oracle.bwt_from_sa(text, want)


def test_cover_128(gpu):
    """The same outputs with the difference cover modulo 128 (the CPU variant's default)."""
    rng = oracle.Rng(7)
    with gpu.GpuContext(1 << 21, cover=128) as ctx:
        for name, text in [("banana", b"banana"), ("run", b"a" * 3000), ("period 129", rng.periodic_text(20000, 129)),
                           ("random sigma=4", rng.random_text(1 << 20, 4)), ("random sigma=256", rng.random_text(300000, 256)),
                           ("fibonacci", oracle.fibonacci_text(1 << 20))]:
            check_case(gpu, f"cover 128 {name}", text, ctx)


def test_merge_parameters(gpu):
    """Small splitter spacing and tiles force several merge levels on modest inputs."""
    rng = oracle.Rng(9)
    text = rng.random_text(300_000, 3)
    for g, k in [(2, 1), (4, 3), (16, 2), (128, 4)]:
        with gpu.GpuContext(len(text), splitter_spacing=g, tile_splitters=k) as ctx:
            check_case(gpu, f"merge g={g} k={k}", text, ctx)


def test_batch_equals_single(gpu):
    rng = oracle.Rng(13)
    texts = [b"banana", b"a", b"ab" * 50, rng.random_text(3000, 4), rng.periodic_text(7000, 33),
             oracle.fibonacci_text(10000), rng.random_text(257, 256), b"\0" * 300, rng.random_text(100000, 26)]
    print(f"batch of {len(texts)}", flush=True)
    results = gpu.gpu_batch_build(texts)
    for i, text in enumerate(texts):
        sa, b, primary = results[i]
        want = expected_sa(text)
        assert np.array_equal(sa, want), f"batch text {i}: SA differs"
        assert (b, primary) == oracle.bwt_from_sa(text, want), f"batch text {i}: BWT differs"


def test_determinism(gpu):
    text = oracle.Rng(17).random_text(500_000, 4)
    a = gpu.gpu_build(text)
    b = gpu.gpu_build(text)
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]
