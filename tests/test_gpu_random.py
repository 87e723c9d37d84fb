"""GPU parity over seeded random cases: texts of random length and structure, both covers, and
merge parameters drawn from their valid range (splitter spacing, tile splitters, tile threads),
built alone and in batches; every suffix array and BWT equal to the oracle's.

The cases draw lengths from 1 to 500,000 bytes log-uniformly, so that the small tiles, the tail
ranking kernel and several merge levels all occur, on the alphabets and repeat structures the
reference's selftest generators cover (src/cli/selftest.cpp:432-606): uniform alphabets of 1 to
256 symbols, periodic texts, runs and Fibonacci words.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = 300


def expected_sa(text):
    if len(text) <= 20000:
        return oracle.naive_sa(text)
    return oracle.ref_suffix_array(text) if oracle.ref_available() else oracle.suffix_array(text)


def random_text(rng, np_rng, n):
    kind = int(np_rng.integers(0, 5))
    if kind == 0:
        return rng.random_text(n, int(np_rng.choice([1, 2, 3, 4, 20, 96, 256])))
    if kind == 1:
        return rng.periodic_text(n, int(np_rng.integers(1, 300)))
    if kind == 2:
        return oracle.fibonacci_text(n)
    if kind == 3:
        # Runs of one symbol among random bytes.
        t = bytearray(rng.random_text(n, 4))
        for _ in range(int(np_rng.integers(1, 6))):
            start = int(np_rng.integers(0, n))
            end = min(n, start + int(np_rng.integers(1, 5000)))
            t[start:end] = b"\x07" * (end - start)
        return bytes(t)
    # Copies of earlier pieces with point changes.
    t = bytearray(rng.random_text(n, 4))
    for _ in range(int(np_rng.integers(1, 8))):
        length = int(np_rng.integers(1, max(2, n // 3)))
        src = int(np_rng.integers(0, max(1, n - length)))
        dst = int(np_rng.integers(0, max(1, n - length)))
        t[dst:dst + length] = t[src:src + length]
        if length > 10:
            t[dst + length // 2] ^= 1
    return bytes(t[:n])


def random_params(np_rng):
    """(splitter spacing, tile splitters, tile threads); 0 = the cover's default."""
    g = int(np_rng.choice([0, 2, 3, 4, 8, 16]))
    k = int(np_rng.choice([0, 1, 2, 5, 17]))
    threads = int(np_rng.choice([0, 256, 512]))
    return g, k, threads


def draw_cases():
    np_rng = np.random.default_rng(2024)
    rng = oracle.Rng(2024)
    cases = []
    for c in range(CASES):
        n = int(np.exp(np_rng.uniform(0, np.log(500_000))))
        cover = int(np_rng.choice([32, 128]))
        cases.append((c, random_text(rng, np_rng, max(n, 1)), cover, random_params(np_rng)))
    return cases


def test_random_single_texts(gpu):
    for c, text, cover, (g, k, threads) in draw_cases():
        print(f"random case {c}: n={len(text)} cover {cover} g={g} k={k} threads={threads}", flush=True)
        with gpu.GpuContext(len(text), cover=cover, splitter_spacing=g, tile_splitters=k, tile_threads=threads) as ctx:
            sa, b, primary = gpu.gpu_build(text, True, True, context=ctx)
        want = expected_sa(text)
        assert np.array_equal(sa, want), f"random case {c}: SA differs"
        assert (b, primary) == oracle.bwt_from_sa(text, want), f"random case {c}: BWT differs"


def test_random_batches(gpu):
    cases = draw_cases()
    np_rng = np.random.default_rng(77)
    for cover in (32, 128):
        pool = [text for _, text, _, _ in cases]
        for b in range(8):
            picks = np_rng.choice(len(pool), size=int(np_rng.integers(2, 24)), replace=False)
            texts = [pool[i] for i in picks]
            g, k, threads = random_params(np_rng)
            print(f"random batch {b} cover {cover}: {len(texts)} texts, {sum(map(len, texts))} bytes, "
                  f"g={g} k={k} threads={threads}", flush=True)
            with gpu.GpuContext(0, cover=cover, splitter_spacing=g, tile_splitters=k, tile_threads=threads,
                                batch_lengths=[len(t) for t in texts]) as ctx:
                results = gpu.gpu_batch_build(texts, context=ctx)
            for i, text in enumerate(texts):
                sa, bwt, primary = results[i]
                want = expected_sa(text)
                assert np.array_equal(sa, want), f"random batch {b} cover {cover} text {i}: SA differs"
                assert (bwt, primary) == oracle.bwt_from_sa(text, want), f"random batch {b} cover {cover} text {i}: BWT differs"
