"""Benchmark: suffix array + BWT construction of the 1 GiB configuration on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3] [--no-cpu]

A step is one construction of the suffix array (int32) and the BWT of the whole text, in one
merge, with the text already resident in HBM and both outputs left in HBM (the device-resident
entry point of the C ABI). value = input bytes per second over the timed steps, in decimal MB/s.

Printed as one JSON line, with:
  roofline: the dominant kernel (the top-level tile merge) - its algorithmic bytes per launch
    (54 bytes per element for SA + BWT: list entry 4, record 44, SA 4, BWT 2) over its average
    duration measured with HIP events on the library's stream, against 8 TB/s of HBM;
  e2e_hbm_fraction: the whole step's algorithmic bytes (127.94 bytes per input byte for fused
    SA + BWT) over the step time, against 8 TB/s;
  cpu_baseline: the reference's CPU construction (oracle/_ref, OpenMP on all host cores) timed
    on a 256 MiB prefix of the same generator's text, SA and BWT built separately as the reference
    does, MB/s = prefix bytes / (SA seconds + BWT seconds).
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
MERGE_BYTES_PER_ELEMENT = 54.0  # list entry 4 + record 44 + SA 4 + BWT 2
METRIC = "SA+BWT input MB/s at 1 GiB; batched/served MB/s + p99 vs sequential"


def cpu_baseline(text_prefix):
    import numpy as np

    import oracle
    if not oracle.ref_available():
        return None
    lib = oracle.ref_lib()
    threads = int(lib.ref_max_threads())
    n = int(text_prefix.size)
    padded = np.zeros(n + oracle.TEXT_PADDING, dtype=np.uint8)
    padded[:n] = text_prefix
    sa = np.empty(n, dtype=np.int32)
    t0 = time.perf_counter()
    assert lib.ref_build_sa(padded.ctypes.data, n, sa.ctypes.data, 128) == 0
    t1 = time.perf_counter()
    out = np.empty(n, dtype=np.uint8)
    lib.ref_build_bwt(padded.ctypes.data, n, out.ctypes.data, sa.ctypes.data, 128)
    t2 = time.perf_counter()
    return {"value": round(n / 1e6 / (t2 - t0), 2), "unit": "MB/s", "cores": threads, "kind": "reference",
            "sample": f"first {n >> 20} MiB of the same text; reference CPU build (cover 128), SA {t1 - t0:.2f}s + BWT {t2 - t1:.2f}s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--cover", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-mib", type=int, default=256)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first copy BYTES device to device three times (the known-size kernel that scales HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(texts would shard across GPUs as described in DESIGN.md, but no multi-GPU path is built)", flush=True)
        return 2

    import numpy as np

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

    ctx = sadcs_amd.GpuContext(n, cover=args.cover)
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
    achieved = MERGE_BYTES_PER_ELEMENT * n / kernel
    e2e = BYTES_PER_INPUT_BYTE * n / step / HBM_PEAK

    # A cheap independent check of the timed output: the BWT agrees with the suffix array, row by row.
    sa = d_sa.to_host(np.int32)
    bwt = d_bwt.to_host(np.uint8)
    check_rows = np.random.default_rng(0).integers(0, n, 1 << 16)
    row0 = int(np.flatnonzero(sa == 0)[0])
    assert primary == row0 + 1
    positions = sa[check_rows]
    expected = np.where(positions > 0, text[np.maximum(positions - 1, 0)], 0)
    index = np.where(check_rows < row0, check_rows + 1, check_rows)
    ok = (positions == 0) | (bwt[index] == expected)
    assert ok.all(), "BWT does not match the suffix array"
    del sa, bwt

    baseline = None
    if not args.no_cpu:
        sample = min(n, args.cpu_sample_mib << 20)
        baseline = cpu_baseline(text[:sample])

    ctx.close()
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
                   "cover": args.cover, "resident": "text in HBM before the timed region, outputs in HBM"},
        "roofline": {"bound": "hbm", "kernel": "mergeTilesKernel (top level, SA+BWT in place)",
                     "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "alg_bytes_per_launch": int(MERGE_BYTES_PER_ELEMENT * n), "avg_launch_ms": round(kernel * 1e3, 3)},
        "e2e_hbm_fraction": round(e2e, 4),
        "phases_ms": {"ranking": round(ranking * 1e3, 2), "classes": round(classes * 1e3, 2), "merge": round(merge * 1e3, 2)},
        "peak_device_bytes_per_input_byte": round(peak_bytes / n, 3),
        "cpu_baseline": baseline,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
