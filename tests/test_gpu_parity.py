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


def runs_among_repeats(n, seed):
    """Random ACGT text with planted repeats (many groups of two after the first pass) and a few
    long runs (one group of thousands of samples): passes after the first place the small groups
    in place and sort the run's group apart."""
    rng = np.random.default_rng(seed)
    text = rng.choice(np.frombuffer(b"ACGT", dtype=np.uint8), n)
    for _ in range(n // 400):
        length = int(rng.integers(200, 600))
        src, dst = (int(x) for x in rng.integers(0, n - length, 2))
        text[dst:dst + length] = text[src:src + length].copy()
    for length in (5000, 3000, 2000, 1500):
        at = int(rng.integers(0, n - length))
        text[at:at + length] = ord("N")
    at = int(rng.integers(0, n - 3000))
    text[at:at + 3000] = np.frombuffer(b"ab" * 1500, dtype=np.uint8)
    return text.tobytes()


def test_runs_among_repeats(gpu):
    for seed, n in [(1, 300_000), (2, 1 << 20)]:
        text = runs_among_repeats(n, seed)
        check_case(gpu, f"runs among repeats seed={seed}", text)
        check_ranks(gpu, f"runs among repeats seed={seed}", text)


def test_batch_runs_among_repeats(gpu):
    texts = [runs_among_repeats(n, 10 + i) for i, n in enumerate([200_000, 500_000, 350_000])]
    texts.append(oracle.Rng(5).random_text(100_000, 200))
    for i, (sa, bwt, primary) in enumerate(gpu.gpu_batch_build(texts)):
        want = expected_sa(texts[i])
        assert np.array_equal(sa, want), i
        assert (bwt, primary) == oracle.bwt_from_sa(texts[i], want), i


def test_batch_alphabets(gpu):
    """Batches whose texts' alphabets differ: with every byte value present somewhere (9-bit codes,
    no per-text codes) and without (each text's first pass packs its own codes), against the
    texts built alone."""
    from sadcs_amd import workloads
    rng = oracle.Rng(9)
    narrow = [workloads.dna_like(60_000, 71).tobytes(), b"ab" * 20_000, rng.random_text(50_000, 2),
              workloads.markov(80_000, 72).tobytes(), b"z" * 700]
    wide = narrow + [bytes(range(256)) * 40 + rng.random_text(5000, 256)]
    for name, texts in [("narrow", narrow), ("wide", wide)]:
        print(f"batch {name}: {len(texts)} texts", flush=True)
        for i, (sa, bwt, primary) in enumerate(gpu.gpu_batch_build(texts)):
            want = expected_sa(texts[i])
            assert np.array_equal(sa, want), (name, i)
            assert (bwt, primary) == oracle.bwt_from_sa(texts[i], want), (name, i)


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
    for g, k, threads in [(2, 1, 512), (4, 3, 512), (16, 2, 512), (128, 4, 512), (8, 0, 256), (16, 3, 256)]:
        with gpu.GpuContext(len(text), splitter_spacing=g, tile_splitters=k, tile_threads=threads) as ctx:
            check_case(gpu, f"merge g={g} k={k} threads={threads}", text, ctx)


def test_merge_parameters_cover_128(gpu):
    """Cover 128 (31 lists) with its default tiles (G = 8, 256 threads), the cover-32 defaults
    (G = 16, 512 threads) and small tiles that force several merge levels."""
    rng = oracle.Rng(19)
    text = rng.random_text(400_000, 4)
    for g, k, threads in [(0, 0, 0), (16, 0, 512), (2, 1, 256), (3, 2, 512), (4, 1, 256), (8, 0, 256)]:
        with gpu.GpuContext(len(text), cover=128, splitter_spacing=g, tile_splitters=k, tile_threads=threads) as ctx:
            check_case(gpu, f"cover 128 merge g={g} k={k} threads={threads}", text, ctx)
            sa, b, primary = gpu.gpu_build(text, True, True, context=ctx)  # the fused merge
            want = expected_sa(text)
            assert np.array_equal(sa, want)
            assert (b, primary) == oracle.bwt_from_sa(text, want)


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


def test_batch_cover_128(gpu):
    """A batch at cover 128 (31 lists with the text index in the keys), with its default tiles and
    with small ones, and the server at cover 128: every text equals the oracle's."""
    rng = oracle.Rng(23)
    texts = [b"banana", b"a", b"ab" * 70, rng.random_text(5000, 4), rng.periodic_text(9000, 129),
             oracle.fibonacci_text(20000), rng.random_text(300, 256), b"\0" * 500, rng.random_text(150000, 20)]
    lengths = [len(t) for t in texts]
    for g, k, threads in [(0, 0, 0), (2, 1, 512)]:
        print(f"cover 128 batch g={g} k={k} threads={threads}", flush=True)
        with gpu.GpuContext(0, cover=128, splitter_spacing=g, tile_splitters=k, batch_lengths=lengths, tile_threads=threads) as ctx:
            results = gpu.gpu_batch_build(texts, context=ctx)
        for i, text in enumerate(texts):
            sa, b, primary = results[i]
            want = expected_sa(text)
            assert np.array_equal(sa, want), f"cover 128 batch text {i}: SA differs"
            assert (b, primary) == oracle.bwt_from_sa(text, want), f"cover 128 batch text {i}: BWT differs"
    print("cover 128 server", flush=True)
    with gpu.Server(workers=2, cover=128) as server:
        handles = [server.submit(t) for t in texts]
        for i, (text, h) in enumerate(zip(texts, handles)):
            sa, b, primary = server.wait(h)[:3]
            want = expected_sa(text)
            assert np.array_equal(sa, want), f"cover 128 served text {i}: SA differs"
            assert (b, primary) == oracle.bwt_from_sa(text, want), f"cover 128 served text {i}: BWT differs"


def test_determinism(gpu):
    text = oracle.Rng(17).random_text(500_000, 4)
    a = gpu.gpu_build(text)
    b = gpu.gpu_build(text)
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


def test_repeated_fused_builds_cover_128(gpu):
    """The fused SA + BWT merge at cover 128, built four times over texts with long repeats (many
    comparisons past the records' 32 characters and down to the ranks): every build equals the
    reference CPU build."""
    from sadcs_amd import workloads
    for name, text in [("dna-like", workloads.dna_like(3 << 20, 41).tobytes()),
                       ("markov", workloads.markov(3 << 20, 42).tobytes())]:
        want = expected_sa(text)
        want_bwt = oracle.bwt_from_sa(text, want)
        with gpu.GpuContext(len(text), cover=128) as ctx:
            for r in range(4):
                print(f"cover 128 fused {name} build {r}", flush=True)
                sa, b, primary = gpu.gpu_build(text, True, True, context=ctx)
                assert np.array_equal(sa, want), f"{name} build {r}: SA differs"
                assert (b, primary) == want_bwt, f"{name} build {r}: BWT differs"


def test_serving_results_equal_alone_and_in_mix(gpu):
    """Requests submitted together are batched and overlapped; each result must equal the same text
    built alone (and the oracle)."""
    from sadcs_amd import workloads
    lengths, kinds = workloads.serving_requests(40, seed=21, median=20_000, sigma=1.2, cap=400_000)
    texts = [workloads.serving_text(n, k, 100 + i).tobytes() for i, (n, k) in enumerate(zip(lengths, kinds))]
    texts += [b"banana", b"a", b""]
    with gpu.Server(memory_budget=1 << 30, workers=2, max_batch_bytes=1 << 20) as server:
        ids = [server.submit(t) for t in texts]
        results = [server.wait(i) for i in ids]
        counters = server.counters()
    print(f"serving counters {counters}", flush=True)
    assert counters["batches"] < len(texts), "requests were not batched"
    for i, text in enumerate(texts):
        sa, b, primary, latency, queued, batch = results[i]
        alone = gpu.gpu_build(text, True, True)
        assert np.array_equal(sa, alone[0]), f"request {i}: SA differs from the single build"
        assert (b, primary) == (alone[1], alone[2]), f"request {i}: BWT differs from the single build"
        if len(text) <= 20000:
            assert np.array_equal(sa, oracle.naive_sa(text))
        assert latency >= queued >= 0


def test_serving_refuses_over_budget(gpu):
    with gpu.Server(memory_budget=64 << 20, workers=1) as server:
        big = server.submit(oracle.Rng(2).random_text(8_000_000, 4))
        small = server.submit(b"mississippi")
        with pytest.raises(gpu.GpuError) as info:
            server.wait(big)
        assert info.value.status == gpu.GpuStatus.NotEnoughMemory
        sa, b, primary, *_ = server.wait(small)
        assert sa.tolist() == oracle.naive_sa(b"mississippi").tolist()


def test_device_batch_equals_single(gpu):
    from sadcs_amd import workloads
    texts = [workloads.dna_like(n, 30 + i) for i, n in enumerate([1500, 70_000, 3000, 250_000])]
    texts += [workloads.markov(n, 40 + i) for i, n in enumerate([2048, 100_000, 9999])]
    lengths = np.array([t.size for t in texts], dtype=np.int64)
    offsets, span = gpu.gpu_batch_layout(lengths)
    host = np.zeros(span + 64, dtype=np.uint8)
    for t, o in zip(texts, offsets):
        host[o:o + t.size] = t
    d_text = gpu.DeviceBuffer.from_host(host)
    total = int(lengths.sum())
    d_sa = gpu.DeviceBuffer(4 * total)
    d_bwt = gpu.DeviceBuffer(total)
    primaries = gpu.gpu_batch_build_device(None, d_text.ptr, lengths, d_sa.ptr, d_bwt.ptr)
    sa = d_sa.to_host(np.int32)
    bwt = d_bwt.to_host(np.uint8)
    offset = 0
    for i, t in enumerate(texts):
        n = t.size
        want_sa, want_bwt, want_primary = gpu.gpu_build(t, True, True)
        assert np.array_equal(sa[offset:offset + n], want_sa), f"text {i}"
        assert bwt[offset:offset + n].tobytes() == want_bwt and primaries[i] == want_primary, f"text {i}"
        offset += n


def test_device_single_equals_host(gpu):
    from sadcs_amd import workloads
    text = workloads.markov(3_000_000, 77)
    d_text = gpu.DeviceBuffer.from_host(text, extra=64)
    d_sa = gpu.DeviceBuffer(4 * text.size)
    d_bwt = gpu.DeviceBuffer(text.size)
    with gpu.GpuContext(text.size) as ctx:
        stats = gpu.GpuStats()
        primary = gpu.gpu_build_device(ctx, d_text.ptr, text.size, d_sa.ptr, d_bwt.ptr, stats)
    want = oracle.ref_suffix_array(text.tobytes()) if oracle.ref_available() else oracle.suffix_array(text.tobytes())
    assert np.array_equal(d_sa.to_host(np.int32), want)
    assert (d_bwt.to_host(np.uint8).tobytes(), primary) == oracle.bwt_from_sa(text.tobytes(), want)
    assert stats.merge_kernel_seconds > 0 and stats.merge_kernel_elements == text.size


def test_batch_many_texts(gpu):
    """Enough texts that the text index takes most of a ranking key (12 bits): the key layout must
    still fit 64 bits. Each text's outputs equal the oracle's."""
    from sadcs_amd import workloads
    rng = np.random.default_rng(12)
    lengths = rng.integers(1, 300, 3000)
    texts = [workloads.dna_like(int(n), 500 + i) if i % 3 else workloads.markov(int(n), 900 + i) for i, n in enumerate(lengths)]
    texts[7] = np.frombuffer(b"a" * 250, dtype=np.uint8)  # one long run
    results = gpu.gpu_batch_build(texts, True, True)
    for i, (text, (sa, b, primary)) in enumerate(zip(texts, results)):
        raw = text.tobytes()
        want = oracle.naive_sa(raw)
        assert np.array_equal(sa, want), f"text {i}"
        assert (b, primary) == oracle.bwt_from_sa(raw, want), f"text {i}"


def test_tuner_picks_a_working_configuration(gpu):
    """The autotuner returns a configuration whose builds equal the defaults' (single and batch)."""
    from sadcs_amd import workloads
    text = workloads.markov(2_000_000, 31)
    d_text = gpu.DeviceBuffer.from_host(text, extra=64)
    best, seconds = gpu.gpu_tune(d_text.ptr, [text.size])
    assert best["cover"] == 32 and best["splitter_spacing"] in (8, 12, 16, 24, 32) and best["tile_splitters"] >= 1
    assert best["tile_threads"] in (256, 512)
    assert seconds > 0
    with gpu.GpuContext(text.size, splitter_spacing=best["splitter_spacing"], tile_splitters=best["tile_splitters"],
                        tile_threads=best["tile_threads"]) as ctx:
        tuned = gpu.gpu_build(text, True, True, context=ctx)
    plain = gpu.gpu_build(text, True, True)
    assert np.array_equal(tuned[0], plain[0]) and tuned[1:] == plain[1:]
    texts = [workloads.dna_like(n, 70 + i) for i, n in enumerate([5000, 12000, 700, 40000])]
    lengths = np.array([t.size for t in texts], dtype=np.int64)
    offsets, span = gpu.gpu_batch_layout(lengths)
    host = np.zeros(span + 64, dtype=np.uint8)
    for t, o in zip(texts, offsets):
        host[o:o + t.size] = t
    d_texts = gpu.DeviceBuffer.from_host(host)
    best_batch, _ = gpu.gpu_tune(d_texts.ptr, lengths)
    assert best_batch["splitter_spacing"] in (8, 12, 16, 24, 32)


def test_device_outputs_at_any_offset(gpu):
    """Device outputs may start at any byte offset of a larger buffer (as when many texts' outputs
    are packed into one allocation); before the merge the library uses them as scratch, which it
    must align itself. Offsets from a failure seen with packed outputs."""
    from sadcs_amd import workloads
    text = workloads.markov(151_701, 4 * 1_000_003 + 4)
    n = text.size
    want = gpu.gpu_build(text, True, True)
    d_text = gpu.DeviceBuffer.from_host(text, extra=64)
    for sa_offset, bwt_offset in [(12, 3), (8, 2), (4, 1), (0, 0)]:
        d_sa = gpu.DeviceBuffer(4 * n + 1024)
        d_bwt = gpu.DeviceBuffer(n + 1024)
        with gpu.GpuContext(n) as ctx:
            primary = gpu.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr + sa_offset, d_bwt.ptr + bwt_offset, gpu.GpuStats())
        sa = d_sa.to_host(np.uint8)[sa_offset:sa_offset + 4 * n].view(np.int32)
        bwt = d_bwt.to_host(np.uint8)[bwt_offset:bwt_offset + n].tobytes()
        assert np.array_equal(sa, want[0]), f"offsets {sa_offset}, {bwt_offset}: SA"
        assert (bwt, primary) == (want[1], want[2]), f"offsets {sa_offset}, {bwt_offset}: BWT"


def test_serving_submit_into_with_one_output(gpu):
    """submit_into with only the SA buffer on a server that makes both outputs: the SA arrives in
    the caller's buffer, the BWT is held and wait copies it into wait's own buffer (and wait leaves
    its SA buffer alone)."""
    import ctypes
    lib = gpu.lib
    texts = [oracle.Rng(31).random_text(n, 4) for n in (70_000, 3000, 1)]
    with gpu.Server(workers=1) as server:
        held = []
        for text in texts:
            t = np.frombuffer(text, dtype=np.uint8)
            sa = np.full(t.size, -7, dtype=np.int32)
            rid = lib.sadcs_hip_server_submit_into(server.handle, t.ctypes.data, t.size, sa.ctypes.data, None)
            assert rid >= 0
            held.append((rid, t, sa))
        for rid, t, sa in held:
            wait_sa = np.full(t.size, -9, dtype=np.int32)
            wait_bwt = np.zeros(t.size, dtype=np.uint8)
            primary = ctypes.c_int32(-1)
            status = lib.sadcs_hip_server_wait(server.handle, rid, wait_sa.ctypes.data, wait_bwt.ctypes.data,
                                               ctypes.byref(primary), None, None, None)
            assert status == 0, gpu.gpu_last_error()
            want = oracle.naive_sa(t.tobytes())
            assert np.array_equal(sa, want)
            assert (wait_bwt.tobytes(), primary.value) == oracle.bwt_from_sa(t.tobytes(), want)
            assert (wait_sa == -9).all(), "wait wrote an output that was given to submit_into"


def test_serving_texts_in_and_beyond_the_pinned_slabs(gpu):
    """Queued texts wait in 64 MiB pinned slabs that are reused once their texts are done; a text
    larger than a slab waits in pageable memory. Two waves of 3 MiB texts (more than one slab per
    wave, so slabs fill and are reused by the second wave) and one 65 MiB text: every result
    equals the reference build's."""
    from sadcs_amd import workloads
    waves = [[workloads.markov(3 << 20, seed=60 + 20 * w + i).tobytes() for i in range(20)] for w in range(2)]
    big = workloads.dna_like((65 << 20) + 17, seed=59).tobytes()
    texts = waves[0] + waves[1] + [big]
    want = oracle.ref_outputs(texts)
    results = []
    with gpu.Server(workers=2) as server:
        for wave in waves:
            ids = [server.submit(t) for t in wave]
            results += [server.wait(i) for i in ids]
        results.append(server.wait(server.submit(big)))
    for i, (got, (sa, b, primary)) in enumerate(zip(results, want)):
        print(f"request {i}: {len(texts[i])} bytes", flush=True)
        assert np.array_equal(got[0], sa), f"request {i}: SA differs from the reference build"
        assert (got[1], got[2]) == (b, primary), f"request {i}: BWT differs from the reference build"
