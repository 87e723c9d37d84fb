"""The command line (sadcs_amd/bin/sadcs) against the oracle, as the reference's CLI is used.

Host-only commands (check of .sa and .bwt files in all three modes, usage errors) run without a
GPU; their verdicts and first failing indices must equal the reference's own checker
(oracle/_ref, src/common/checker.cpp) or, where that is not built, the oracle's fast check.
Construction commands (build, build --bwt, build --both, test, selftest, check --device) are GPU
tests: their files must hold the oracle's suffix array and BWT in divsufsort's formats
(src/common/fileio.cpp:65-125).
"""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "sadcs_amd", "bin", "sadcs")


def run(*args, timeout=120):
    if not os.path.exists(BINARY):
        pytest.fail(f"{BINARY} is missing: build it with `python sadcs_amd/build.py`")
    print("sadcs " + " ".join(str(a) for a in args), flush=True)
    done = subprocess.run([BINARY, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return done.returncode, done.stdout + done.stderr


def write_sa(path, sa):
    np.ascontiguousarray(sa, dtype="<i4").tofile(path)


def write_bwt(path, bwt, primary):
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(bwt)))
        if bwt:
            f.write(struct.pack("<i", primary))
            f.write(bwt)


def read_bwt(path):
    data = open(path, "rb").read()
    n = struct.unpack_from("<i", data, 0)[0]
    if n == 0:
        return b"", 0
    return data[8:8 + n], struct.unpack_from("<i", data, 4)[0]


def reference_check(text, sa, mode):
    if oracle.ref_available():
        return oracle.ref_check_sa(text, sa, mode)
    return oracle.check_sa(text, sa) if mode == 0 else None


@pytest.fixture
def selftest_text(tmp_path):
    text = oracle.Rng(7).random_text(3000, 3)
    path = tmp_path / "in.txt"
    path.write_bytes(text)
    return text, path


def test_check_accepts_a_correct_suffix_array(selftest_text, tmp_path):
    text, path = selftest_text
    sa_path = tmp_path / "in.sa"
    write_sa(sa_path, oracle.naive_sa(text))
    for flags in ([], ["--lowmem"], ["--naive"], ["--threads", "2"]):
        rc, out = run("check", path, sa_path, *flags)
        assert rc == 0 and "suffix array OK (3000 entries" in out, out


@pytest.mark.parametrize("mode,flag", [(0, None), (1, "--lowmem"), (2, "--naive")])
def test_check_reports_the_first_failure(selftest_text, tmp_path, mode, flag):
    """The reference selftest's mutations (src/cli/selftest.cpp:321-327), reported at the index
    the reference's checker names in each mode."""
    text, path = selftest_text
    good = oracle.naive_sa(text)
    mutations = {
        "swapped neighbours": lambda sa: sa.__setitem__([1000, 1001], sa[[1001, 1000]]),
        "swapped far apart": lambda sa: sa.__setitem__([10, 2900], sa[[2900, 10]]),
        "duplicate entry": lambda sa: sa.__setitem__(500, sa[1500]),
        "out of range": lambda sa: sa.__setitem__(42, sa.size),
        "negative entry": lambda sa: sa.__setitem__(0, -1),
    }
    for name, apply in mutations.items():
        bad = good.copy()
        apply(bad)
        sa_path = tmp_path / "bad.sa"
        write_sa(sa_path, bad)
        rc, out = run("check", path, sa_path, *([flag] if flag else []))
        assert rc == 1, (name, out)
        match = re.search(r"check FAILED at SA\[(-?\d+)\]", out)
        assert match, (name, out)
        got = int(match.group(1))
        if mode == 0:
            assert got == oracle.check_sa(text, bad), (name, out)
        want = reference_check(text, bad, mode)
        if name == "duplicate entry" and mode == 0:
            # The reference's fast check lets OpenMP threads race for the inverse slot of a
            # duplicated value, so it names either copy; this library names the first.
            assert want in (500, 1500)
        elif want is not None:
            assert got == want, (name, out)


@pytest.mark.skipif(not oracle.ref_available(), reason="needs the reference build in oracle/_ref")
def test_random_mutations_reported_where_the_reference_reports_them(tmp_path):
    """Swaps, repeats and out-of-range entries at seeded random places, on texts with long runs and
    small alphabets: in every mode the first failing index equals the reference checker's
    (src/common/checker.cpp), the fast mode's repeat aside as above."""
    rng = np.random.default_rng(41)
    texts = [oracle.Rng(8).random_text(2000, 2), b"ab" * 700 + b"a" * 600, oracle.Rng(9).random_text(1500, 256)]
    for t, text in enumerate(texts):
        path = tmp_path / f"t{t}.txt"
        path.write_bytes(text)
        good = oracle.naive_sa(text)
        n = good.size
        for k in range(12):
            bad = good.copy()
            kind = k % 4
            i, j = (int(x) for x in rng.integers(0, n, 2))
            if kind == 0:
                bad[[i, j]] = bad[[j, i]]
            elif kind == 1:
                bad[[i, min(i + 1, n - 1)]] = bad[[min(i + 1, n - 1), i]]
            elif kind == 2:
                bad[i] = bad[j]
            else:
                bad[i] = n + j if k % 8 == 3 else -1 - j
            if np.array_equal(bad, good):
                continue
            sa_path = tmp_path / "bad.sa"
            write_sa(sa_path, bad)
            for mode, flag in [(0, None), (1, "--lowmem"), (2, "--naive")]:
                rc, out = run("check", path, sa_path, *([flag] if flag else []))
                assert rc == 1, (t, k, mode, out)
                got = int(re.search(r"check FAILED at SA\[(-?\d+)\]", out).group(1))
                want = oracle.ref_check_sa(text, bad, mode)
                if mode == 0 and kind == 2 and bad[i] != good[i]:
                    holders = np.flatnonzero(bad == bad[i])
                    assert want in holders and got == holders.min(), (t, k, out)
                else:
                    assert got == want, (t, k, mode, out)
        want_bwt, primary = oracle.bwt_from_sa(text, good)
        for k in range(6):
            changed = bytearray(want_bwt)
            at = int(rng.integers(0, n))
            changed[at] = (changed[at] + 1 + int(rng.integers(0, 255))) % 256
            bwt_path = tmp_path / "bad.bwt"
            write_bwt(bwt_path, bytes(changed), primary)
            rc, out = run("check", "--bwt", path, bwt_path)
            assert rc == 1, out
            got = int(re.search(r", index (-?\d+)$", out.strip()).group(1))
            assert got == oracle.ref_check_bwt(text, bytes(changed), primary), (t, k, out)


def test_check_bwt_files(selftest_text, tmp_path):
    text, path = selftest_text
    good, primary = oracle.bwt_from_sa(text, oracle.naive_sa(text))
    bwt_path = tmp_path / "in.bwt"
    write_bwt(bwt_path, good, primary)
    rc, out = run("check", "--bwt", path, bwt_path)
    assert rc == 0 and "BWT OK" in out, out
    changed = bytearray(good)
    changed[1234] ^= 1
    for bwt, p in [(bytes(changed), primary), (good, primary + 1 if primary < len(text) else primary - 1), (good, 0)]:
        write_bwt(bwt_path, bwt, p)
        rc, out = run("check", "--bwt", path, bwt_path)
        assert rc == 1 and "BWT check FAILED" in out, out
        want = oracle.ref_check_bwt(text, bwt, p) if oracle.ref_available() else oracle.check_bwt(text, bwt, p)
        got = int(re.search(r", index (-?\d+)$", out.strip()).group(1))
        # The reference reports a primary out of range at index -1; its test driver returns 0 for
        # it, keeping -1 for a pass.
        assert got == want or (got == -1 and want == 0 and "primary index not in 1..n" in out), out


def test_check_rejects_mismatched_sizes_and_bad_usage(selftest_text, tmp_path):
    text, path = selftest_text
    sa_path = tmp_path / "short.sa"
    write_sa(sa_path, oracle.naive_sa(text)[:-1])
    rc, out = run("check", path, sa_path)
    assert rc == 1 and "2999 entries" in out, out
    assert run("check", path)[0] == 1
    assert run("build", path, tmp_path / "x.sa", "--cover", "64")[0] == 1
    rc, out = run("frobnicate")
    assert rc == 1 and "usage" in out


# ---- construction on the GPU ------------------------------------------------------------------

@pytest.mark.gpu
def test_build_writes_divsufsort_files(gpu, tmp_path):
    from sadcs_amd import workloads
    cases = {"dna.txt": workloads.dna_like(200_000, 8).tobytes(), "words.txt": b"mississippi banana " * 500,
             "one.txt": b"x", "empty.txt": b""}
    for name, text in cases.items():
        path = tmp_path / name
        path.write_bytes(text)
        want = oracle.naive_sa(text) if len(text) <= 20000 else oracle.suffix_array(text)
        want_bwt = oracle.bwt_from_sa(text, want)
        for cover in (32, 128):
            rc, out = run("build", path, tmp_path / "out.sa", "--cover", cover)
            assert rc == 0, out
            assert np.array_equal(np.fromfile(tmp_path / "out.sa", dtype="<i4"), want), (name, cover)
            rc, out = run("build", "--bwt", path, tmp_path / "out.bwt", "--cover", cover)
            assert rc == 0, out
            assert read_bwt(tmp_path / "out.bwt") == want_bwt, (name, cover)
        rc, out = run("-b", "--both", path, tmp_path / "both.sa")
        assert rc == 0, out
        assert np.array_equal(np.fromfile(tmp_path / "both.sa", dtype="<i4"), want), name
        assert read_bwt(str(tmp_path / "both.sa") + ".bwt") == want_bwt, name
        rc, out = run("check", path, tmp_path / "both.sa")
        assert rc == 0, out


@pytest.mark.gpu
def test_device_check_matches_host_check(gpu, selftest_text, tmp_path):
    text, path = selftest_text
    good = oracle.naive_sa(text)
    sa_path = tmp_path / "in.sa"
    write_sa(sa_path, good)
    rc, out = run("check", path, sa_path, "--device")
    assert rc == 0 and "checked on the GPU" in out, out
    bad = good.copy()
    bad[[10, 2900]] = bad[[2900, 10]]
    write_sa(sa_path, bad)
    rc, out = run("check", path, sa_path, "--device")
    assert rc == 1 and f"check FAILED at SA[{oracle.check_sa(text, bad)}]" in out, out


@pytest.mark.gpu
def test_test_command_over_a_directory(gpu, tmp_path):
    from sadcs_amd import workloads
    corpus = tmp_path / "corpus"
    (corpus / "sub").mkdir(parents=True)
    (corpus / "a.txt").write_bytes(workloads.markov(100_000, 3).tobytes())
    (corpus / "b.txt").write_bytes(b"abracadabra" * 1000)
    (corpus / "sub" / "c.txt").write_bytes(workloads.dna_like(50_000, 4).tobytes())
    for flags, files in [([], 2), (["--recursive"], 3), (["--recursive", "--bwt"], 3), (["--recursive", "--device"], 3)]:
        rc, out = run("test", corpus, *flags)
        assert rc == 0 and f"{files}/{files} OK" in out, (flags, out)


@pytest.mark.gpu
def test_selftest_passes(gpu):
    rc, out = run("selftest", timeout=600)
    assert rc == 0, out[-2000:]


@pytest.mark.gpu
def test_bench_batch_and_serve_equal_one_at_a_time(gpu, tmp_path):
    from sadcs_amd import workloads
    corpus = tmp_path / "corpus"
    corpus.mkdir()
    rng = np.random.default_rng(3)
    for i, n in enumerate(rng.integers(1, 200_000, 12)):
        data = workloads.markov(int(n), 40 + i) if i % 2 else workloads.dna_like(int(n), 60 + i)
        (corpus / f"t{i:02d}.txt").write_bytes(data.tobytes())
    (corpus / "empty.txt").write_bytes(b"")
    (corpus / "one.txt").write_bytes(b"z")
    rc, out = run("bench", corpus, "--batch", "--serve", "--both", "--runs", "1", timeout=300)
    assert rc == 0, out
    assert "batch of 13 files" in out and "served 13 files" in out, out
    assert out.count("outputs equal") == 2, out


@pytest.mark.gpu
def test_build_from_pinned_buffers(gpu, tmp_path):
    from sadcs_amd import workloads
    text = workloads.markov(300_000, 17).tobytes()
    path = tmp_path / "in.txt"
    path.write_bytes(text)
    want = oracle.suffix_array(text)
    rc, out = run("build", "--both", path, tmp_path / "out.sa", "--pinned")
    assert rc == 0, out
    assert np.array_equal(np.fromfile(tmp_path / "out.sa", dtype="<i4"), want)
    assert read_bwt(str(tmp_path / "out.sa") + ".bwt") == oracle.bwt_from_sa(text, want)
