"""GPU parity at the benchmark configurations' full sizes, against the reference's own CPU build.

Each configuration's seeded text (sadcs_amd/workloads.py) is built by the HIP path at its stated
size and cover, and every suffix array and BWT must equal, byte for byte, the reference CPU build's
(oracle/_ref, src/cpu/dcsort.cpp: buildSuffixArray and buildBwt). The SA of a text is unique, so
the reference's cover 128 output is the expected output for either cover. This is the reference
selftest's GPU contract (src/cli/selftest.cpp:432-471: full SA and BWT against the CPU build) at the
configurations' sizes:
  C1  16 MiB i.i.d. ACGT, cover 128 (and 32)
  C2  256 MiB DNA-like, cover 128 (and 32)
  C3  1 GiB order-3 Markov text, cover 32 (the benchmark's) and 128
  C4  4096 texts of 1 KiB - 4 MiB in one batched construction, every text
  C5  a 300-request lognormal stream through the continuous-batching server, every request
"""
import threading
import time

import numpy as np
import pytest

import oracle

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not oracle.ref_available(), reason="needs oracle/_ref")]


def reference_single(text):
    """(sa, bwt, primary) of the reference build: buildSuffixArray, then buildBwt on its own."""
    t0 = time.time()
    sa = oracle.ref_suffix_array(text)
    bwt, primary = oracle.ref_bwt(text)
    print(f"  reference SA + BWT of {len(text)} bytes in {time.time() - t0:.1f}s", flush=True)
    return sa, bwt, primary


def assert_equal_outputs(name, got, want):
    sa, bwt, primary = got
    want_sa, want_bwt, want_primary = want
    assert sa.shape == want_sa.shape, name
    mismatch = np.flatnonzero(sa != want_sa)
    assert mismatch.size == 0, f"{name}: SA differs first at {mismatch[0]} of {sa.size}"
    assert primary == want_primary, f"{name}: primary {primary}, expected {want_primary}"
    if bwt != want_bwt:
        a = np.frombuffer(bwt, dtype=np.uint8)
        b = np.frombuffer(want_bwt, dtype=np.uint8)
        first = np.flatnonzero(a != b)[0] if a.size == b.size else -1
        raise AssertionError(f"{name}: BWT differs first at {first}")


def build_with_cover(gpu, text, cover):
    with gpu.GpuContext(len(text), cover=cover) as ctx:
        stats = gpu.GpuStats()
        t0 = time.time()
        out = gpu.gpu_build(text, True, True, context=ctx, stats=stats)
        print(f"  GPU cover {cover}: {time.time() - t0:.2f}s host call, device {stats.device_seconds * 1e3:.1f} ms, "
              f"{stats.character_passes} character passes, {stats.doubling_rounds} doubling rounds", flush=True)
        assert stats.cover == cover
    return out


@pytest.mark.parametrize("config,covers", [("c1", (128, 32)), ("c2", (128, 32)), ("c3", (32, 128))])
def test_single_text_config(gpu, config, covers):
    from sadcs_amd import workloads
    text = workloads.config_text(config)
    print(f"config {config}: {workloads.CONFIGS[config][0]}, n={text.size}", flush=True)
    want = reference_single(text)
    for cover in covers:
        print(f"case {config} cover {cover}", flush=True)
        assert_equal_outputs(f"{config} cover {cover}", build_with_cover(gpu, text, cover), want)


def test_c4_batch_of_4096_texts(gpu):
    """C4: the 4096 texts (2.02 GB, half DNA-like, half Markov) in one batched construction from host
    memory; every text's SA and BWT against the reference build of that text alone."""
    from sadcs_amd import workloads
    t0 = time.time()
    texts = workloads.batch_texts(4096)
    print(f"C4: {len(texts)} texts, {sum(t.size for t in texts)} bytes, generated in {time.time() - t0:.1f}s", flush=True)
    stats = gpu.GpuStats()
    t0 = time.time()
    results = gpu.gpu_batch_build(texts, True, True, stats=stats)
    print(f"  batch host call {time.time() - t0:.2f}s, device {stats.device_seconds * 1e3:.1f} ms", flush=True)
    t0 = time.time()
    want = oracle.ref_outputs(texts)
    print(f"  reference builds of every text in {time.time() - t0:.1f}s", flush=True)
    for i in range(len(texts)):
        assert_equal_outputs(f"C4 text {i} (n={texts[i].size})", results[i], want[i])


def test_c5_serving_stream(gpu):
    """C5: 300 requests (lognormal lengths, median 256 KiB, sigma 1.5, capped at 64 MiB; 45% ACGT,
    45% protein-like, 10% periodic) arriving as a Poisson stream while earlier ones run, through the
    continuous-batching server; every result against the reference build of that text alone."""
    from sadcs_amd import workloads
    lengths, kinds = workloads.serving_requests(300)
    texts = [workloads.serving_text(n, k, 1000 + i) for i, (n, k) in enumerate(zip(lengths, kinds))]
    total = int(lengths.sum())
    print(f"C5: {len(texts)} requests, {total} bytes, largest {int(lengths.max())}", flush=True)
    rng = np.random.default_rng(5)
    gaps = rng.exponential(0.0005, len(texts))
    ids = [None] * len(texts)
    with gpu.Server(workers=2) as server:
        def submitter():
            for i, text in enumerate(texts):
                time.sleep(gaps[i])
                ids[i] = server.submit(text)

        feeder = threading.Thread(target=submitter)
        t0 = time.time()
        feeder.start()
        feeder.join()
        results = [server.wait(rid) for rid in ids]
        seconds = time.time() - t0
        counters = server.counters()
    print(f"  served in {seconds:.2f}s ({total / seconds / 1e6:.0f} MB/s), counters {counters}", flush=True)
    assert counters["completed"] == len(texts)
    assert counters["batches"] < len(texts), "requests were not batched"
    want = oracle.ref_outputs(texts)
    for i, (sa, bwt, primary, latency, queued, batch) in enumerate(results):
        assert_equal_outputs(f"C5 request {i} (n={texts[i].size}, kind {int(kinds[i])})", (sa, bwt, primary), want[i])
        assert latency >= queued >= 0 and batch >= 1
