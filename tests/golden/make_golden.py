"""Writes the golden fixtures from the reference's CPU build (oracle/_ref, compiled from the
reference sources by oracle/Makefile). Run from the repository root:

    python tests/golden/make_golden.py

Files: <name>.sa (n little-endian int32, divsufsort's mksary format) and <name>.bwt (int32 n, int32
primary when n > 0, n bytes; src/common/fileio.cpp:108-125) for the selftest's words, and
hashes.json with the SHA-256 of the SA and BWT files of larger seeded inputs.
"""
import hashlib
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

WORDS = {"empty": b"", "a": b"a", "banana": b"banana", "mississippi": b"mississippi",
         "abracadabra": b"abracadabra", "yabbadabbado": b"yabbadabbado"}


def sa_file(sa):
    return b"".join(struct.pack("<i", int(x)) for x in sa)


def bwt_file(n, primary, bwt):
    return struct.pack("<i", n) + (struct.pack("<i", primary) + bwt if n > 0 else b"")


def inputs():
    rng = oracle.Rng(3)
    yield "random4_64k", rng.random_text(1 << 16, 4)
    yield "random256_64k", rng.random_text(1 << 16, 256)
    yield "period33_64k", rng.periodic_text(1 << 16, 33)
    yield "fibonacci_64k", oracle.fibonacci_text(1 << 16)
    yield "run_4k", b"a" * 4096


def main():
    assert oracle.ref_available(), "build oracle/_ref first (make -f oracle/Makefile)"
    for name, text in WORDS.items():
        sa = oracle.ref_suffix_array(text) if text else []
        bwt, primary = oracle.ref_bwt(text) if text else (b"", 0)
        with open(os.path.join(HERE, name + ".sa"), "wb") as f:
            f.write(sa_file(sa))
        with open(os.path.join(HERE, name + ".bwt"), "wb") as f:
            f.write(bwt_file(len(text), primary, bwt))
    hashes = {}
    for name, text in inputs():
        sa = oracle.ref_suffix_array(text)
        bwt, primary = oracle.ref_bwt(text)
        hashes[name] = {"n": len(text), "primary": primary,
                        "sa_sha256": hashlib.sha256(sa_file(sa)).hexdigest(),
                        "bwt_sha256": hashlib.sha256(bwt_file(len(text), primary, bwt)).hexdigest()}
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
