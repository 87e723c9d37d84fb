"""Every committed golden fixture (tests/golden/, written by the reference's CPU build through
tests/golden/make_golden.py) read by a test.

CPU: the word files (<name>.sa in divsufsort's mksary format, <name>.bwt as int32 n, int32 primary
when n > 0, then n bytes; src/common/fileio.cpp:65-125) equal the oracle's naive suffix array and
the BWT read off it; hashes.json's SHA-256 of the SA and BWT files of five seeded 4-64 KiB inputs
equal those of the oracle's construction (the C restatement, not the reference build).
GPU: the command line's `build` and `build --bwt` write files byte-identical to the fixtures.
"""
import hashlib
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BINARY = os.path.join(ROOT, "sadcs_amd", "bin", "sadcs")
WORDS = {"empty": b"", "a": b"a", "banana": b"banana", "mississippi": b"mississippi",
         "abracadabra": b"abracadabra", "yabbadabbado": b"yabbadabbado"}


def golden(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def hashed_inputs():
    """The seeded inputs of hashes.json, generated as make_golden.py generates them."""
    rng = oracle.Rng(3)
    return {"random4_64k": rng.random_text(1 << 16, 4), "random256_64k": rng.random_text(1 << 16, 256),
            "period33_64k": rng.periodic_text(1 << 16, 33), "fibonacci_64k": oracle.fibonacci_text(1 << 16),
            "run_4k": b"a" * 4096}


def sa_file(sa):
    return np.ascontiguousarray(sa, dtype="<i4").tobytes()


def bwt_file(n, primary, bwt):
    return struct.pack("<i", n) + (struct.pack("<i", primary) + bwt if n > 0 else b"")


def test_every_fixture_is_read():
    used = {f"{w}.{ext}" for w in WORDS for ext in ("sa", "bwt")} | {"hashes.json", "make_golden.py"}
    assert set(os.listdir(GOLDEN)) - {"__pycache__"} == used


@pytest.mark.parametrize("word", sorted(WORDS))
def test_word_files_equal_the_oracle(word):
    text = WORDS[word]
    sa = oracle.naive_sa(text)
    assert golden(f"{word}.sa") == sa_file(sa)
    b, primary = oracle.bwt_from_sa(text, sa) if text else (b"", 0)
    assert golden(f"{word}.bwt") == bwt_file(len(text), primary, b)


def test_hashes_equal_the_oracle_construction():
    table = json.loads(golden("hashes.json"))
    inputs = hashed_inputs()
    assert set(table) == set(inputs)
    for name, text in inputs.items():
        entry = table[name]
        assert entry["n"] == len(text), name
        sa = oracle.suffix_array(text)
        b, primary = oracle.bwt_from_sa(text, sa)
        assert primary == entry["primary"], name
        assert hashlib.sha256(sa_file(sa)).hexdigest() == entry["sa_sha256"], name
        assert hashlib.sha256(bwt_file(len(text), primary, b)).hexdigest() == entry["bwt_sha256"], name


def run_cli(*args):
    if not os.path.exists(BINARY):
        pytest.fail(f"{BINARY} is missing: build it with `python sadcs_amd/build.py`")
    print("sadcs " + " ".join(map(str, args)), flush=True)
    done = subprocess.run([BINARY, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert done.returncode == 0, done.stdout + done.stderr


@pytest.mark.gpu
def test_cli_files_equal_the_fixtures(gpu, tmp_path):
    inputs = {**WORDS, **hashed_inputs()}
    table = json.loads(golden("hashes.json"))
    for name, text in inputs.items():
        path = tmp_path / f"{name}.txt"
        path.write_bytes(text)
        for cover in (32, 128):
            run_cli("build", path, tmp_path / "out.sa", "--cover", cover)
            run_cli("build", "--bwt", path, tmp_path / "out.bwt", "--cover", cover)
            sa_bytes = (tmp_path / "out.sa").read_bytes()
            bwt_bytes = (tmp_path / "out.bwt").read_bytes()
            if name in WORDS:
                assert sa_bytes == golden(f"{name}.sa"), (name, cover)
                assert bwt_bytes == golden(f"{name}.bwt"), (name, cover)
            else:
                assert hashlib.sha256(sa_bytes).hexdigest() == table[name]["sa_sha256"], (name, cover)
                assert hashlib.sha256(bwt_bytes).hexdigest() == table[name]["bwt_sha256"], (name, cover)
