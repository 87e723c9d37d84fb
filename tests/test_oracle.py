"""The oracle pinned: golden vectors, brute force, closed forms and the reference's own CPU build.

Every function of oracle.c is checked here against something that does not share its formula:
  - cover tables: the definition of a difference cover (exhaustively) and brute-force tie offsets;
  - naive_sa: golden vectors from the reference's own outputs and brute force over all binary
    strings up to length 12;
  - bwt_from_sa / check_bwt: golden vectors, inversion round trips, the reference's buildBwt;
  - check_sa: mutations planted as src/cli/selftest.cpp:309-385 plants them;
  - rank_sample: the closed form (ranks read off the suffix array, gpusa.h:62-73);
  - class_lists: the closed form (the suffix array filtered to each class);
  - suffix_array (rank + classes + merge): the reference CPU build (oracle/_ref).
"""
import itertools
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


# ---- covers ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("v", [32, 64, 128, 256])
def test_cover_is_a_difference_cover(v):
    lib = oracle.oracle_lib()
    members = [r for r in range(v) if lib.oracle_is_sample(v, r)]
    assert len(members) == lib.oracle_cover_size(v)
    distances = {(b - a) % v for a in members for b in members}
    assert distances == set(range(v))


@pytest.mark.parametrize("v", [32, 128])
def test_tie_offsets_minimal(v):
    lib = oracle.oracle_lib()
    for a in range(v):
        for b in range(0, v, 3):
            k = lib.oracle_tie_offset(v, a, b)
            assert lib.oracle_is_sample(v, a + k) and lib.oracle_is_sample(v, b + k)
            assert all(not (lib.oracle_is_sample(v, a + j) and lib.oracle_is_sample(v, b + j)) for j in range(k))
            assert k < v


@pytest.mark.parametrize("v", [32, 64, 128, 256])
def test_pack_counts_samples(v):
    lib = oracle.oracle_lib()
    count = 0
    for p in range(3 * v):
        assert lib.oracle_pack(v, p) == count
        count += lib.oracle_is_sample(v, p)


def test_class_counts():
    lib = oracle.oracle_lib()
    # docs/gpu-design.md:27: mod 32, 7 classes; the CPU's mod 128 cover has a largest gap of 31.
    assert lib.oracle_cover_classes(32) == 7
    assert lib.oracle_cover_classes(128) == 31
    sizes = [sum(1 for r in range(32) if lib.oracle_class_of(32, r) == d) for d in range(7)]
    assert sizes == [7, 6, 5, 5, 4, 3, 2]


# ---- naive suffix array and BWT -----------------------------------------------------------------

def test_golden_banana():
    # Outputs of the reference binary (SA file and BWT file bytes), committed as fixtures.
    sa = np.frombuffer(load_golden("banana.sa"), dtype="<i4")
    assert oracle.naive_sa(b"banana").tolist() == sa.tolist() == [5, 3, 1, 0, 4, 2]
    raw = load_golden("banana.bwt")
    n, primary = np.frombuffer(raw[:8], dtype="<i4")
    assert (n, primary, raw[8:]) == (6, 4, b"annbaa")
    assert oracle.bwt_from_sa(b"banana", oracle.naive_sa(b"banana")) == (b"annbaa", 4)
    assert load_golden("a.bwt") == bytes([1, 0, 0, 0, 1, 0, 0, 0]) + b"a"
    assert oracle.bwt_from_sa(b"a", oracle.naive_sa(b"a")) == (b"a", 1)
    assert load_golden("empty.bwt") == bytes(4)


def brute_sa(text):
    return sorted(range(len(text)), key=lambda i: text[i:])


def test_naive_sa_brute_force():
    for n in range(0, 13):
        for bits in itertools.product(b"ab", repeat=n) if n <= 10 else [tuple(b"ab"[(i * 7 + n) % 2] for i in range(n))]:
            text = bytes(bits)
            assert oracle.naive_sa(text).tolist() == brute_sa(text)


def test_naive_sa_zero_bytes_are_characters():
    text = b"\0a\0\0a\0"
    assert oracle.naive_sa(text).tolist() == brute_sa(text)


def test_bwt_round_trip():
    rng = oracle.Rng(21)
    for sigma in [1, 2, 4, 256]:
        for n in [1, 2, 3, 17, 500]:
            text = rng.random_text(n, sigma)
            b, primary = oracle.bwt_from_sa(text, oracle.naive_sa(text))
            assert oracle.check_bwt(text, b, primary) == -1


# ---- checkers: planted mutations (src/cli/selftest.cpp:309-385) ---------------------------------

def test_check_sa_catches_mutations():
    text = oracle.Rng(7).random_text(3000, 4)
    sa = oracle.naive_sa(text)
    assert oracle.check_sa(text, sa) == -1
    swapped = sa.copy()
    swapped[[10, 11]] = swapped[[11, 10]]
    assert oracle.check_sa(text, swapped) >= 0
    far = sa.copy()
    far[[5, 2000]] = far[[2000, 5]]
    assert oracle.check_sa(text, far) >= 0
    duplicate = sa.copy()
    duplicate[100] = duplicate[101]
    assert oracle.check_sa(text, duplicate) >= 0
    for value in (len(text), -1):
        bad = sa.copy()
        bad[7] = value
        assert oracle.check_sa(text, bad) >= 0


def test_check_bwt_catches_mutations():
    text = oracle.Rng(7).random_text(2000, 4)
    b, primary = oracle.bwt_from_sa(text, oracle.naive_sa(text))
    assert oracle.check_bwt(text, b, primary) == -1
    arr = bytearray(b)
    i = next(k for k in range(1, len(arr)) if arr[k] != arr[k - 1])
    arr[i], arr[i - 1] = arr[i - 1], arr[i]
    assert oracle.check_bwt(text, bytes(arr), primary) != -1
    arr = bytearray(b)
    arr[50] = ord("z")
    assert oracle.check_bwt(text, bytes(arr), primary) != -1
    assert oracle.check_bwt(text, b, primary + 1 if primary < len(text) else primary - 1) != -1
    assert oracle.check_bwt(text, b, 0) != -1


# ---- the GPU algorithm's stages against closed forms ----------------------------------------------

def selftest_inputs():
    rng = oracle.Rng(1)
    yield "banana", b"banana"
    for n in [1, 2, 31, 32, 33, 100, 270]:
        yield f"run {n}", b"a" * n
    yield "zeros", b"\0" * 1000
    for period in [1, 2, 3, 5, 7, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000]:
        yield f"period {period}", rng.periodic_text(5000, period)
    yield "fibonacci", oracle.fibonacci_text(5000)
    yield "fibonacci 0/1", oracle.fibonacci_text(5000, 0, 1)
    yield "thue-morse", oracle.thue_morse_text(4096)
    for sigma in [1, 2, 3, 4, 26, 256]:
        for i in range(4):
            yield f"random {sigma} #{i}", rng.random_text(1 + rng.next_below(4000), sigma)


@pytest.mark.parametrize("v", [32, 128])
def test_rank_sample_matches_closed_form(v):
    for name, text in selftest_inputs():
        sa = oracle.naive_sa(text)
        ranks, _ = oracle.rank_sample(text, v)
        assert np.array_equal(ranks, oracle.sample_ranks_from_sa(sa, v)), name


@pytest.mark.parametrize("v", [32, 128])
def test_class_lists_match_closed_form(v):
    lib = oracle.oracle_lib()
    for name, text in selftest_inputs():
        sa = oracle.naive_sa(text)
        ranks, _ = oracle.rank_sample(text, v)
        lists, starts = oracle.class_lists(text, ranks, v)
        for d in range(len(starts) - 1):
            want = [p for p in sa if lib.oracle_class_of(v, int(p)) == d]
            assert lists[starts[d]:starts[d + 1]].tolist() == want, (name, d)


@pytest.mark.parametrize("v", [32, 128])
def test_suffix_array_matches_naive(v):
    for name, text in selftest_inputs():
        assert np.array_equal(oracle.suffix_array(text, v), oracle.naive_sa(text)), name


def test_planted_comparator_error_is_caught():
    """A comparator that drops the 'shorter suffix is smaller' rule must produce a wrong SA that the
    checkers reject: the tests above would notice a plausible mistake of this kind."""
    text = b"abcab" * 40 + b"abc"
    sa = oracle.naive_sa(text)
    wrong = sa.copy()
    i = int(np.flatnonzero(wrong == len(text) - 3)[0])
    wrong[[i, i + 1]] = wrong[[i + 1, i]]
    assert oracle.check_sa(text, wrong) >= 0


# ---- the reference CPU build ------------------------------------------------------------------------

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="the reference build oracle/_ref is absent")


@needs_ref
def test_reference_agrees_with_naive():
    for name, text in selftest_inputs():
        assert np.array_equal(oracle.ref_suffix_array(text), oracle.naive_sa(text)), name
        assert oracle.ref_bwt(text) == oracle.bwt_from_sa(text, oracle.naive_sa(text)), name


@needs_ref
def test_oracle_agrees_with_reference_large():
    rng = oracle.Rng(2)
    for name, text in [("random 4", rng.random_text(1 << 19, 4)), ("period 33", rng.periodic_text(1 << 18, 33)),
                       ("fibonacci", oracle.fibonacci_text(1 << 18)), ("random 256", rng.random_text(1 << 19, 256))]:
        sa = oracle.ref_suffix_array(text)
        assert oracle.ref_check_sa(text, sa) == -1
        assert np.array_equal(oracle.suffix_array(text), sa), name


@needs_ref
def test_checkers_agree_with_reference():
    text = oracle.Rng(4).random_text(4000, 3)
    sa = oracle.ref_suffix_array(text)
    bad = sa.copy()
    bad[[3, 4]] = bad[[4, 3]]
    assert (oracle.check_sa(text, sa) == -1) and (oracle.ref_check_sa(text, sa) == -1)
    assert (oracle.check_sa(text, bad) >= 0) and (oracle.ref_check_sa(text, bad) >= 0)
    b, primary = oracle.ref_bwt(text)
    assert oracle.check_bwt(text, b, primary) == -1 and oracle.ref_check_bwt(text, b, primary) == -1


def test_generators_match_reference_definitions():
    """splitmix64 (selftest.cpp:34-39): the first outputs for seed 1, computed from its definition."""
    mask = (1 << 64) - 1
    state = 1
    expected = []
    for _ in range(3):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        expected.append(z ^ (z >> 31))
    rng = oracle.Rng(1)
    assert [rng.next() for _ in range(3)] == expected
    assert oracle.fibonacci_text(8) == b"abaababa"
    assert oracle.thue_morse_text(8) == b"abbabaab"
