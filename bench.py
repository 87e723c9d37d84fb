"""Benchmark: suffix array + BWT construction of the 1 GiB configuration on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--no-cpu]

A step is one construction of the suffix array (int32) and the BWT of the whole text, in one
merge, with the text already resident in HBM and both outputs left in HBM (the device-resident
entry point of the C ABI). value = input bytes per second over the timed steps, in decimal MB/s.

Printed as one JSON line, with:
  roofline: the dominant kernel (the top-level tile merge) - its algorithmic bytes per launch
    (54 bytes per element for SA + BWT: list entry 4, record 44, SA 4, BWT 2) over its average
    duration measured with HIP events on the library's stream, against 8 TB/s of HBM; traffic is
    the kernel's measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy
    of known size, profiles/mergeTilesKernel_c3_1GiB_traffic.md);
  e2e_hbm_fraction: the whole step's algorithmic bytes (127.94 bytes per input byte for fused
    SA + BWT) over the step time, against 8 TB/s;
  host_call: the same 1 GiB text through the host-memory API (pageable text in, SA and BWT out,
    PCIe included; the download overlaps the end of the merge), best of 3 calls, outputs compared
    with the device-resident build's - a secondary figure, never value;
  cpu_baseline: the reference's CPU construction (oracle/_ref, OpenMP on all host threads, cover
    128) on the same whole text, SA and BWT built separately as the reference CLI does, best of 3
    each, MB/s = n / (SA seconds + BWT seconds), with the CPU model and core count;
  configs: secondary device-resident lines: the same text at the other cover (128: the cover the
    configurations state), then C1 (16 MiB ACGT) and C2 (256 MiB DNA-like) at covers 128 and 32,
    each checked on the device;
  batched: 4096 texts of log-uniform length 1 KiB - 4 MiB (even: DNA-like, odd: Markov), all
    resident in HBM, built in one batched construction, against the same texts built one at a
    time by the single-text device path; every text's outputs of both compared;
  served: requests (lognormal lengths, median 256 KiB, sigma 1.5, capped at 64 MiB; 45% ACGT,
    45% protein-like, 10% periodic) arriving as a Poisson stream at 0.5, 0.8 and 1.1 times the
    measured one-at-a-time capacity, through the continuous-batching server (host memory in and
    out); p50/p99 latency from submission to result and MB/s over the run. Further streams at 2, 3
    and 4 times that capacity and every request at once show where the server stops keeping up.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
BYTES_PER_INPUT_BYTE = 127.94   # fused SA + BWT, the reference design's logical traffic
# Top-level merge, per element: list entry 4 + record (16 characters and a 4-byte rank per sample
# of its window: 7 for cover 32, 13 for cover 128) + SA 4 + BWT 2.
MERGE_BYTES_PER_ELEMENT = {32: 54.0, 128: 78.0}
METRIC = "SA+BWT input MB/s at 1 GiB; batched/served MB/s + p99 vs sequential"


def measured_traffic(config, n):
    """HBM bytes per launch of the top-level merge: the rocprofv3 FETCH_SIZE + WRITE_SIZE figure per
    element, scaled by a copy of known size, from the committed summary for this configuration."""
    path = os.path.join(ROOT, "profiles", f"mergeTilesKernel_{config}_1GiB_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_element:"):
            return int(float(line.split(":")[1]) * n)
    return None


def host_cpu():
    """(model name, physical cores, logical CPUs this process may use) of the host."""
    model, cores = "unknown", set()
    physical = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, value = line.partition(":")
                key, value = key.strip(), value.strip()
                if key == "model name":
                    model = value
                elif key == "physical id":
                    physical = value
                elif key == "core id":
                    core = value
                elif not key and physical is not None:
                    cores.add((physical, core))
                    physical = core = None
        if physical is not None:
            cores.add((physical, core))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, len(cores) or None, usable


def cpu_baseline(text, runs=3):
    """The reference CPU construction (oracle/_ref, OpenMP on every thread it is given, cover 128)
    on the whole text: SA (buildSuffixArray) and BWT (buildBwt) built separately, as the reference
    CLI's build and build --bwt do, each the best of `runs`; value = n / (best SA + best BWT)."""
    import numpy as np

    import oracle
    if not oracle.ref_available():
        return None
    lib = oracle.ref_lib()
    threads = int(lib.ref_max_threads())
    model, physical, usable = host_cpu()
    n = int(text.size)
    padded = np.zeros(n + oracle.TEXT_PADDING, dtype=np.uint8)
    padded[:n] = text
    sa = np.empty(n, dtype=np.int32)
    out = np.empty(n, dtype=np.uint8)
    sa_times, bwt_times = [], []
    for r in range(runs):
        t0 = time.perf_counter()
        assert lib.ref_build_sa(padded.ctypes.data, n, sa.ctypes.data, 128) == 0
        t1 = time.perf_counter()
        primary = lib.ref_build_bwt(padded.ctypes.data, n, out.ctypes.data, sa.ctypes.data, 128)
        t2 = time.perf_counter()
        assert primary >= 0
        sa_times.append(t1 - t0)
        bwt_times.append(t2 - t1)
        print(f"# CPU reference run {r}: SA {t1 - t0:.2f}s, BWT {t2 - t1:.2f}s", file=sys.stderr, flush=True)
    sa_s, bwt_s = min(sa_times), min(bwt_times)
    return {"value": round(n / 1e6 / (sa_s + bwt_s), 2), "unit": "MB/s", "cores": threads, "kind": "reference",
            "sample": f"the whole {n >> 20} MiB text; reference CPU build (cover 128, {threads} OpenMP threads), "
                      f"SA and BWT built separately, best of {runs} each",
            "sa_seconds": round(sa_s, 3), "bwt_seconds": round(bwt_s, 3),
            "sa_mb_per_s": round(n / 1e6 / sa_s, 1), "bwt_mb_per_s": round(n / 1e6 / bwt_s, 1),
            "sa_plus_bwt_mib_per_s": round(n / 2**20 / (sa_s + bwt_s), 1),
            "cpu_model": model, "physical_cores": physical, "usable_logical_cpus": usable}


def cpu_texts_baseline(texts, what):
    """The reference CPU build (all host threads) over a list of texts one after another, SA then
    BWT of each: the sequential CPU path for the batched and served workloads."""
    import numpy as np

    import oracle
    if not oracle.ref_available():
        return None
    lib = oracle.ref_lib()
    total = 0
    t0 = time.perf_counter()
    for text in texts:
        n = int(text.size)
        padded = np.zeros(n + oracle.TEXT_PADDING, dtype=np.uint8)
        padded[:n] = text
        sa = np.empty(n, dtype=np.int32)
        assert lib.ref_build_sa(padded.ctypes.data, n, sa.ctypes.data, 128) == 0
        out = np.empty(n, dtype=np.uint8)
        lib.ref_build_bwt(padded.ctypes.data, n, out.ctypes.data, sa.ctypes.data, 128)
        total += n
    seconds = time.perf_counter() - t0
    return {"value": round(total / 1e6 / seconds, 2), "unit": "MB/s", "cores": int(lib.ref_max_threads()), "kind": "reference",
            "sample": f"{what}: {len(texts)} texts, {total} bytes, SA + BWT each, one after another, {seconds:.2f}s"}


def batched_measurement(count, steps, tune=True, cpu_texts=0):
    """Batched construction of `count` mixed texts against one-at-a-time device builds."""
    import numpy as np

    import sadcs_amd
    from sadcs_amd import workloads

    t0 = time.perf_counter()
    texts = workloads.batch_texts(count)
    lengths = np.array([t.size for t in texts], dtype=np.int64)
    total = int(lengths.sum())
    offsets, span = sadcs_amd.gpu_batch_layout(lengths)
    host = np.zeros(span + 64, dtype=np.uint8)
    for t, o in zip(texts, offsets):
        host[o:o + t.size] = t
    # The one-at-a-time path needs 64 zero bytes after each text.
    seqOffsets = np.zeros(count, dtype=np.int64)
    position = 0
    for i, n in enumerate(lengths):
        seqOffsets[i] = position
        position += (int(n) + 64 + 255) // 256 * 256
    seqHost = np.zeros(position + 64, dtype=np.uint8)
    for t, o in zip(texts, seqOffsets):
        seqHost[o:o + t.size] = t
    rows = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    print(f"# batched: {count} texts, {total} bytes, generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    d_texts = sadcs_amd.DeviceBuffer.from_host(host)
    d_seq = sadcs_amd.DeviceBuffer.from_host(seqHost)
    del host, seqHost
    d_sa = sadcs_amd.DeviceBuffer(4 * total)
    d_bwt = sadcs_amd.DeviceBuffer(total)
    d_sa2 = sadcs_amd.DeviceBuffer(4 * total)
    d_bwt2 = sadcs_amd.DeviceBuffer(total)
    # Autotune on the first 256 texts (their own batch layout), then build everything with it.
    tuned = {}
    if tune:
        sub_lengths = lengths[:256]
        sub_offsets, sub_span = sadcs_amd.gpu_batch_layout(sub_lengths)
        sub_host = np.zeros(sub_span + 64, dtype=np.uint8)
        for t, o in zip(texts[:256], sub_offsets):
            sub_host[o:o + t.size] = t
        d_sub = sadcs_amd.DeviceBuffer.from_host(sub_host)
        tuned, _ = sadcs_amd.gpu_tune(d_sub.ptr, sub_lengths)
        d_sub.free()
        print(f"# batched: tuned {tuned}", file=sys.stderr, flush=True)
    ctx = sadcs_amd.GpuContext(int(lengths.max()))
    batch_ctx = sadcs_amd.GpuContext(0, splitter_spacing=tuned.get("splitter_spacing", 0),
                                     tile_splitters=tuned.get("tile_splitters", 0), batch_lengths=lengths,
                                     tile_threads=tuned.get("tile_threads", 0))
    stats = sadcs_amd.GpuStats()

    batch_primaries = sadcs_amd.gpu_batch_build_device(batch_ctx, d_texts.ptr, lengths, d_sa.ptr, d_bwt.ptr)  # warmup
    sadcs_amd.gpu_synchronize()
    start = time.perf_counter()
    for _ in range(steps):
        batch_primaries = sadcs_amd.gpu_batch_build_device(batch_ctx, d_texts.ptr, lengths, d_sa.ptr, d_bwt.ptr, stats)
    sadcs_amd.gpu_synchronize()
    batch_seconds = (time.perf_counter() - start) / steps
    phases = {"ranking": round(stats.ranking_seconds * 1e3, 2), "classes": round(stats.class_sort_seconds * 1e3, 2),
              "merge": round(stats.merge_seconds * 1e3, 2), "merge_kernel": round(stats.merge_kernel_seconds * 1e3, 2)}
    print(f"# batched: {batch_seconds * 1e3:.1f} ms per batch (last one: {phases}; "
          f"{stats.character_passes} character passes, {stats.doubling_rounds} doubling rounds)", file=sys.stderr, flush=True)

    seq_primaries = np.empty(count, dtype=np.int64)
    sadcs_amd.gpu_synchronize()
    start = time.perf_counter()
    for i in range(count):
        seq_primaries[i] = sadcs_amd.gpu_build_device(ctx, d_seq.ptr + int(seqOffsets[i]), int(lengths[i]),
                                                      d_sa2.ptr + 4 * int(rows[i]), d_bwt2.ptr + int(rows[i]), stats)
    sadcs_amd.gpu_synchronize()
    seq_seconds = time.perf_counter() - start
    ctx.close()
    batch_ctx.close()

    # Every text: the batched outputs equal the one-at-a-time outputs.
    same = bool(np.array_equal(np.asarray(batch_primaries), seq_primaries))
    for a, b in ((d_sa, d_sa2), (d_bwt, d_bwt2)):
        same = same and np.array_equal(a.to_host(np.uint8), b.to_host(np.uint8))
    for buf in (d_texts, d_seq, d_sa, d_bwt, d_sa2, d_bwt2):
        buf.free()
    cpu = cpu_texts_baseline(texts[:cpu_texts], f"all {count} texts" if cpu_texts >= count else f"the first {cpu_texts} of the {count} texts") if cpu_texts else None
    return {"texts": count, "bytes": total, "mb_per_s": round(total / batch_seconds / 1e6, 1),
            "ms": round(batch_seconds * 1e3, 2), "sequential_mb_per_s": round(total / seq_seconds / 1e6, 1),
            "sequential_ms": round(seq_seconds * 1e3, 1), "speedup_vs_sequential": round(seq_seconds / batch_seconds, 2),
            "outputs_equal_sequential": same, "resident": "texts in HBM, outputs in HBM",
            "phases_ms": phases, "merge_params": tuned or "defaults", "cpu_baseline": cpu}


def served_measurement(requests, factors=(0.5, 0.8, 1.1, 2.0, 3.0, 4.0, None), cpu=False, workers=0, batch_mib=0):
    """Poisson request streams through the serving loop, against one-at-a-time host builds. A factor
    of None submits every request at once: the server's capacity with full batches."""
    import threading

    import numpy as np

    import sadcs_amd
    from sadcs_amd import workloads

    lengths, kinds = workloads.serving_requests(requests, seed=5)
    texts = [workloads.serving_text(int(n), int(k), 1000 + i) for i, (n, k) in enumerate(zip(lengths, kinds))]
    total = int(lengths.sum())

    # One at a time through the host API (the sequential capacity the rates are set against).
    with sadcs_amd.GpuContext(int(lengths.max())) as ctx:
        sadcs_amd.gpu_build(texts[0], True, True, context=ctx)
        start = time.perf_counter()
        for t in texts:
            sadcs_amd.gpu_build(t, True, True, context=ctx)
        seq_seconds = time.perf_counter() - start
    capacity = len(texts) / seq_seconds
    print(f"# served: {requests} requests, {total} bytes; one at a time {seq_seconds:.2f}s ({capacity:.0f} requests/s)",
          file=sys.stderr, flush=True)

    rates = []
    rng = np.random.default_rng(7)
    for factor in factors:
        rate = factor * capacity if factor else float("inf")
        gaps = rng.exponential(1.0 / rate, len(texts)) if factor else np.zeros(len(texts))
        results = [None] * len(texts)
        # 0: the library's defaults (5 workers, batches of at most 32 MiB)
        with sadcs_amd.Server(workers=workers, max_batch_bytes=batch_mib << 20) as server:
            ids = [None] * len(texts)

            def waiter():
                for i in range(len(texts)):
                    while ids[i] is None:
                        time.sleep(0.0002)
                    results[i] = server.wait(ids[i])

            thread = threading.Thread(target=waiter)
            thread.start()
            start = time.perf_counter()
            due = start
            for i, t in enumerate(texts):
                due += gaps[i]
                pause = due - time.perf_counter()
                if pause > 0:
                    time.sleep(pause)
                ids[i] = server.submit(t)
            thread.join()
            elapsed = time.perf_counter() - start
            counters = server.counters()
            phases = server.stats()
        latency = np.array([r[3] for r in results]) * 1e3
        rates.append({"offered_x_sequential": factor if factor else "all at once",
                      "offered_requests_per_s": round(rate, 1) if factor else None,
                      "mb_per_s": round(total / elapsed / 1e6, 1), "x_sequential": round(seq_seconds / elapsed, 2),
                      "p50_ms": round(float(np.percentile(latency, 50)), 2),
                      "p99_ms": round(float(np.percentile(latency, 99)), 2),
                      "mean_batch": round(float(np.mean([r[5] for r in results])), 2), "batches": counters["batches"],
                      "batch_phase_seconds": {k: round(phases[k], 4) for k in ("upload_seconds", "device_seconds",
                                                                               "download_seconds", "copy_out_seconds")}})
        print(f"# served at {factor}x: {rates[-1]}", file=sys.stderr, flush=True)
    return {"requests": requests, "bytes": total, "sequential_mb_per_s": round(total / seq_seconds / 1e6, 1),
            "sequential_requests_per_s": round(capacity, 1), "rates": rates, "io": "host memory in and out (PCIe)",
            "cpu_baseline": cpu_texts_baseline(texts, "every request") if cpu else None}


def host_call_measurement(ctx, text, d_sa, d_bwt, primary, runs):
    """The same text through the host-memory API (gpuSuffixArray + gpuBwt in one call): text uploaded
    from pageable memory, SA and BWT returned to fresh pageable arrays, PCIe included. Best of runs;
    the outputs must equal the device-resident build's."""
    import ctypes
    import numpy as np
    import sadcs_amd
    n = text.size
    best = None
    for r in range(runs):
        # The C ABI call alone, into fresh arrays (the Python wrapper's copy of the BWT into a
        # bytes object is not the library's).
        stats = sadcs_amd.GpuStats()
        sa = np.empty(n, dtype=np.int32)
        bwt = np.empty(n, dtype=np.uint8)
        got = ctypes.c_int32(-1)
        start = time.perf_counter()
        status = sadcs_amd.lib.sadcs_hip_build(ctx.handle, text.ctypes.data, n, sa.ctypes.data, bwt.ctypes.data,
                                               ctypes.byref(got), ctypes.byref(stats))
        wall = time.perf_counter() - start
        assert status == 0, sadcs_amd.gpu_last_error()
        got_primary = got.value
        print(f"# host call {r}: {wall * 1e3:.1f} ms (upload {stats.upload_seconds * 1e3:.1f}, device "
              f"{stats.device_seconds * 1e3:.1f}, download after the merge {stats.download_seconds * 1e3:.1f})",
              file=sys.stderr, flush=True)
        if best is None or wall < best[0]:
            best = (wall, stats.upload_seconds, stats.device_seconds, stats.download_seconds)
        if r + 1 < runs:
            del sa, bwt
    same = bool(got_primary == primary and np.array_equal(sa, d_sa.to_host(np.int32))
                and np.array_equal(bwt, d_bwt.to_host(np.uint8)))
    assert same, "host call outputs differ from the device-resident build's"
    wall, upload, device, download = best
    return {"value": round(n / wall / 1e6, 1), "unit": "MB/s", "ms": round(wall * 1e3, 1),
            "upload_ms": round(upload * 1e3, 1), "device_ms": round(device * 1e3, 1),
            "download_after_merge_ms": round(download * 1e3, 1), "runs": runs,
            "io": "pageable host memory in and out over PCIe; the download overlaps the top-level merge",
            "outputs_equal_device_build": same}


def device_line(sadcs_amd, d_text, n, d_sa, d_bwt, cover, steps, warmup, label):
    """Steps of one device-resident SA + BWT construction with the default merge configuration for
    `cover`, outputs checked on the device afterwards: a secondary line (the configurations' stated
    covers)."""
    stats = sadcs_amd.GpuStats()
    with sadcs_amd.GpuContext(n, cover=cover) as ctx:
        for _ in range(warmup):
            sadcs_amd.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, stats)
        sadcs_amd.gpu_synchronize()
        phases = []
        start = time.perf_counter()
        for _ in range(steps):
            primary = sadcs_amd.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, stats)
            phases.append((stats.ranking_seconds, stats.class_sort_seconds, stats.merge_seconds))
        sadcs_amd.gpu_synchronize()
        step = (time.perf_counter() - start) / steps
    verdict = sadcs_amd.gpu_check_device(d_text.ptr, n, d_sa.ptr, d_bwt.ptr, primary)
    assert verdict == (0, -1), f"{label} cover {cover}: output check failed: {verdict}"
    ranking, classes, merge = (sum(p[i] for p in phases) / len(phases) for i in range(3))
    line = {"workload": label, "n": n, "cover": cover, "ms_per_step": round(step * 1e3, 3),
            "mb_per_s": round(n / step / 1e6, 1), "e2e_hbm_fraction": round(BYTES_PER_INPUT_BYTE * n / step / HBM_PEAK, 4),
            "phases_ms": {"ranking": round(ranking * 1e3, 2), "classes": round(classes * 1e3, 2), "merge": round(merge * 1e3, 2)},
            "character_passes": stats.character_passes, "doubling_rounds": stats.doubling_rounds,
            "merge_kernel_ms": round(stats.merge_kernel_seconds * 1e3, 3),
            "peak_device_bytes_per_input_byte": round(stats.device_bytes / n, 3), "output_check": "pass (device, every row)"}
    print(f"# {label}, cover {cover}: {line}", file=sys.stderr, flush=True)
    return line


def config_lines(sadcs_amd, workloads, names, covers, steps=3, warmup=1):
    """device_line for each configuration text (generated here) at each cover."""
    lines = []
    for name in names:
        description, size, _, _ = workloads.CONFIGS[name]
        text = workloads.config_text(name, padding=64)
        d_text = sadcs_amd.DeviceBuffer.from_host(text)
        d_sa = sadcs_amd.DeviceBuffer(4 * size)
        d_bwt = sadcs_amd.DeviceBuffer(size)
        for cover in covers:
            lines.append(device_line(sadcs_amd, d_text, size, d_sa, d_bwt, cover, steps, warmup, f"{name}: {description}"))
        for buf in (d_text, d_sa, d_bwt):
            buf.free()
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--cover", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3, help="runs of the CPU reference's SA and BWT (best of)")
    ap.add_argument("--tune-mib", type=int, default=64,
                    help="autotune the merge configuration on a prefix of this many MiB before timing (0: defaults)")
    ap.add_argument("--batch-texts", type=int, default=4096, help="texts in the batched measurement (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests per offered rate (0: skip)")
    ap.add_argument("--serve-workers", type=int, default=0, help="server workers (0: the library default)")
    ap.add_argument("--serve-batch-mib", type=int, default=0, help="largest server batch in MiB (0: the library default)")
    ap.add_argument("--host-runs", type=int, default=3, help="host-memory calls of the same text (0: skip)")
    ap.add_argument("--config-lines", type=int, default=1, help="secondary lines: the other cover, C1 and C2 (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES device to device three times (the known-size kernel that scales HBM counters)")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="GATHERS",
                    help="first gather GATHERS random 128-byte lines of a 4 GiB array three times, whole and as the "
                         "merge reads its records (known-size random reads that scale HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(batches would shard by text across GPUs, one large text by a distributed sample sort; "
              "no multi-GPU path is built)", flush=True)
        return 2

    import sadcs_amd
    from sadcs_amd import workloads

    description, size, _, _ = workloads.CONFIGS[args.config]
    t0 = time.perf_counter()
    text = workloads.config_text(args.config, padding=64)
    n = size
    print(f"# generated {description}: {n} bytes in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    # Device buffers come from the library's own HIP runtime (one runtime in the process).
    d_text = sadcs_amd.DeviceBuffer.from_host(text)
    d_sa = sadcs_amd.DeviceBuffer(4 * n)
    d_bwt = sadcs_amd.DeviceBuffer(n)
    sadcs_amd.gpu_synchronize()

    if args.copy_calibration:
        source = sadcs_amd.DeviceBuffer(args.copy_calibration)
        target = sadcs_amd.DeviceBuffer(args.copy_calibration)
        source.zero()
        for _ in range(3):
            sadcs_amd.gpu_copy_device(target.ptr, source.ptr, args.copy_calibration)
        sadcs_amd.gpu_synchronize()
        source.free()
        target.free()
        print(f"# copied {args.copy_calibration} bytes device to device 3 times", file=sys.stderr, flush=True)

    if args.gather_calibration:
        lines = sadcs_amd.DeviceBuffer(4 << 30)
        lines.zero()
        for whole in (True, False):
            for _ in range(3):
                sadcs_amd.gpu_gather_lines(lines.ptr, (4 << 30) // 128, args.gather_calibration, whole)
        lines.free()
        print(f"# gathered {args.gather_calibration} random lines of a 4 GiB array 3 times whole, then 3 times as records",
              file=sys.stderr, flush=True)

    tuned, tune_seconds = {}, 0.0
    if args.tune_mib:
        start = time.perf_counter()
        sample = min(n, args.tune_mib << 20)
        d_sample = sadcs_amd.DeviceBuffer.from_host(text[:sample], extra=64)
        tuned, _ = sadcs_amd.gpu_tune(d_sample.ptr, [sample], cover=args.cover)
        d_sample.free()
        tune_seconds = time.perf_counter() - start
        print(f"# tuned on {sample >> 20} MiB in {tune_seconds:.2f}s: {tuned}", file=sys.stderr, flush=True)
    ctx = sadcs_amd.GpuContext(n, cover=args.cover, splitter_spacing=tuned.get("splitter_spacing", 0),
                               tile_splitters=tuned.get("tile_splitters", 0), tile_threads=tuned.get("tile_threads", 0))
    stats = sadcs_amd.GpuStats()
    for w in range(args.warmup):
        sadcs_amd.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, stats)
        print(f"# warmup {w}: device {stats.device_seconds * 1e3:.1f} ms (rank {stats.ranking_seconds * 1e3:.1f}, "
              f"class {stats.class_sort_seconds * 1e3:.1f}, merge {stats.merge_seconds * 1e3:.1f}; "
              f"{stats.character_passes} character passes, {stats.doubling_rounds} doubling rounds)", file=sys.stderr, flush=True)

    sadcs_amd.gpu_synchronize()
    merge_kernel = []
    phases = []
    primary = -1
    start = time.perf_counter()
    for _ in range(args.steps):
        primary = sadcs_amd.gpu_build_device(ctx, d_text.ptr, n, d_sa.ptr, d_bwt.ptr, stats)
        merge_kernel.append(stats.merge_kernel_seconds)
        phases.append((stats.ranking_seconds, stats.class_sort_seconds, stats.merge_seconds))
    sadcs_amd.gpu_synchronize()
    elapsed = time.perf_counter() - start
    peak_bytes = stats.device_bytes

    step = elapsed / args.steps
    value = n / step / 1e6
    kernel = sum(merge_kernel) / len(merge_kernel)
    achieved = MERGE_BYTES_PER_ELEMENT[args.cover] * n / kernel
    e2e = BYTES_PER_INPUT_BYTE * n / step / HBM_PEAK

    # The timed output checked whole on the device: a permutation, neighbours in order (the
    # reference's fast check), and the BWT and primary as the suffix array fixes them.
    start = time.perf_counter()
    verdict = sadcs_amd.gpu_check_device(d_text.ptr, n, d_sa.ptr, d_bwt.ptr, primary)
    check_seconds = time.perf_counter() - start
    assert verdict == (0, -1), f"output check failed: {sadcs_amd.CHECK_NAMES[verdict[0]]} at index {verdict[1]}"
    print(f"# device check of all {n} rows passed in {check_seconds:.2f}s", file=sys.stderr, flush=True)

    host_call = host_call_measurement(ctx, text[:n], d_sa, d_bwt, primary, args.host_runs) if args.host_runs else None
    ctx.close()

    # The other covers and configurations: cover 128 (the configurations' stated cover) on the same
    # resident text, then C1 and C2 at covers 128 and 32.
    configs = []
    if args.config_lines:
        other = 128 if args.cover == 32 else 32
        configs.append(device_line(sadcs_amd, d_text, n, d_sa, d_bwt, other, 3, 1, f"{args.config}: {description}"))

    baseline = None if args.no_cpu else cpu_baseline(text[:n], args.cpu_runs)

    for buf in (d_text, d_sa, d_bwt):
        buf.free()
    if args.config_lines:
        configs += config_lines(sadcs_amd, workloads, [c for c in ("c1", "c2") if c != args.config], (128, 32))
    batched = batched_measurement(args.batch_texts, 3, tune=args.tune_mib > 0,
                                  cpu_texts=0 if args.no_cpu else args.batch_texts) if args.batch_texts else None
    served = served_measurement(args.serve_requests, cpu=not args.no_cpu, workers=args.serve_workers,
                                batch_mib=args.serve_batch_mib) if args.serve_requests else None
    ranking, classes, merge = (sum(p[i] for p in phases) / len(phases) for i in range(3))
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8 text, u32 positions/ranks, u64 sort keys",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {description}", "n": n, "outputs": "SA int32 + BWT, one merge",
                   "cover": args.cover, "resident": "text in HBM before the timed region, outputs in HBM",
                   "merge_params": tuned or "defaults", "tuning_seconds": round(tune_seconds, 2)},
        "roofline": {"bound": "hbm", "kernel": "mergeTilesKernel (top level, SA+BWT in place)",
                     "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": measured_traffic(args.config, n) if args.cover == 32 else None,  # measured at cover 32
                     "alg_bytes_per_launch": int(MERGE_BYTES_PER_ELEMENT[args.cover] * n), "avg_launch_ms": round(kernel * 1e3, 3)},
        "e2e_hbm_fraction": round(e2e, 4),
        "phases_ms": {"ranking": round(ranking * 1e3, 2), "classes": round(classes * 1e3, 2), "merge": round(merge * 1e3, 2)},
        "peak_device_bytes_per_input_byte": round(peak_bytes / n, 3),
        "cpu_baseline": baseline,
        "output_check": {"kind": "device, every row: permutation, neighbour order, BWT and primary",
                         "result": "pass", "seconds": round(check_seconds, 3)},
        "host_call": host_call,
        "configs": configs,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
