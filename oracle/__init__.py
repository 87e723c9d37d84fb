"""The oracle: CPU checkers of the HIP construction. TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package. It loads
  - liboracle.so: the C restatement of the reference's GPU algorithm (oracle.c), and
  - _ref/libsadcs_ref.so: the reference's own CPU construction and checker, compiled from its
    sources by oracle/Makefile where /root/reference exists (the built library travels with the
    tree, so it is also available where the reference sources are not).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsadcs_ref.so")
TEXT_PADDING = 64


def build():
    """Compiles liboracle.so and, where the reference sources exist, _ref/libsadcs_ref.so."""
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def _load_oracle():
    if not os.path.exists(ORACLE_PATH):
        build()
    lib = ctypes.CDLL(ORACLE_PATH)
    P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sig = {
        "oracle_cover_size": ([ctypes.c_int], ctypes.c_int),
        "oracle_cover_classes": ([ctypes.c_int], ctypes.c_int),
        "oracle_pack": ([ctypes.c_int, i64], i64),
        "oracle_is_sample": ([ctypes.c_int, i64], ctypes.c_int),
        "oracle_class_of": ([ctypes.c_int, i64], ctypes.c_int),
        "oracle_tie_offset": ([ctypes.c_int, i64, i64], ctypes.c_int),
        "oracle_naive_sa": ([P, i32, P], None),
        "oracle_bwt_from_sa": ([P, i32, P, P], i32),
        "oracle_sample_ranks_from_sa": ([ctypes.c_int, P, i32, P], None),
        "oracle_check_sa": ([P, i32, P], i64),
        "oracle_check_bwt": ([P, i32, P, i32], i64),
        "oracle_rank_sample": ([ctypes.c_int, P, i32, P, P], ctypes.c_int),
        "oracle_class_lists": ([ctypes.c_int, P, i32, P, P, P], None),
        "oracle_less": ([ctypes.c_int, P, i32, P, i64, i64], ctypes.c_int),
        "oracle_merge": ([ctypes.c_int, P, i32, P, P, P, P], None),
        "oracle_suffix_array": ([ctypes.c_int, P, i32, P], ctypes.c_int),
        "oracle_splitmix_next": ([P], ctypes.c_uint64),
        "oracle_random_text": ([P, i32, ctypes.c_uint32, P], None),
        "oracle_periodic_text": ([P, i32, i32, P], None),
        "oracle_fibonacci_text": ([i32, ctypes.c_uint8, ctypes.c_uint8, P], None),
        "oracle_thue_morse_text": ([i32, P], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        _oracle = _load_oracle()
    return _oracle


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    """The reference's CPU build, or None where it was never built."""
    global _ref
    if _ref is None and ref_available():
        lib = ctypes.CDLL(REF_PATH)
        P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        lib.ref_build_sa.argtypes = [P, i32, P, i32]
        lib.ref_build_sa.restype = ctypes.c_int
        lib.ref_build_bwt.argtypes = [P, i32, P, P, i32]
        lib.ref_build_bwt.restype = i32
        lib.ref_check_sa.argtypes = [P, i32, P, i32]
        lib.ref_check_sa.restype = i64
        lib.ref_check_bwt.argtypes = [P, i32, P, i32]
        lib.ref_check_bwt.restype = i64
        lib.ref_set_threads.argtypes = [i32]
        lib.ref_set_threads.restype = None
        lib.ref_max_threads.argtypes = []
        lib.ref_max_threads.restype = i32
        _ref = lib
    return _ref


def _padded(text):
    t = np.frombuffer(bytes(text), dtype=np.uint8) if not isinstance(text, np.ndarray) else text
    padded = np.zeros(t.size + TEXT_PADDING, dtype=np.uint8)
    padded[:t.size] = t
    return padded, t.size


# ---- the C restatement -------------------------------------------------------------------------

def naive_sa(text):
    p, n = _padded(text)
    sa = np.empty(max(n, 1), dtype=np.int32)
    oracle_lib().oracle_naive_sa(p.ctypes.data, n, sa.ctypes.data)
    return sa[:n]


def bwt_from_sa(text, sa):
    p, n = _padded(text)
    sa = np.ascontiguousarray(sa, dtype=np.int32)
    out = np.empty(max(n, 1), dtype=np.uint8)
    primary = oracle_lib().oracle_bwt_from_sa(p.ctypes.data, n, sa.ctypes.data, out.ctypes.data)
    return out[:n].tobytes(), primary


def sample_ranks_from_sa(sa, cover=32):
    sa = np.ascontiguousarray(sa, dtype=np.int32)
    n = sa.size
    m = oracle_lib().oracle_pack(cover, n)
    ranks = np.zeros(max(m, 1), dtype=np.uint32)
    oracle_lib().oracle_sample_ranks_from_sa(cover, sa.ctypes.data, n, ranks.ctypes.data)
    return ranks[:m]


def check_sa(text, sa):
    p, n = _padded(text)
    sa = np.ascontiguousarray(sa, dtype=np.int32)
    return oracle_lib().oracle_check_sa(p.ctypes.data, n, sa.ctypes.data)


def check_bwt(text, bwt, primary):
    p, n = _padded(text)
    b = np.frombuffer(bytes(bwt), dtype=np.uint8) if n else np.zeros(1, dtype=np.uint8)
    return oracle_lib().oracle_check_bwt(p.ctypes.data, n, b.ctypes.data, primary)


def rank_sample(text, cover=32):
    p, n = _padded(text)
    m = oracle_lib().oracle_pack(cover, n)
    ranks = np.zeros(max(m, 1), dtype=np.uint32)
    passes = np.zeros(2, dtype=np.int32)
    status = oracle_lib().oracle_rank_sample(cover, p.ctypes.data, n, ranks.ctypes.data, passes.ctypes.data)
    if status != 0:
        raise RuntimeError("oracle ranking failed")
    return ranks[:m], (int(passes[0]), int(passes[1]))


def class_lists(text, ranks, cover=32):
    p, n = _padded(text)
    ranks = np.ascontiguousarray(ranks, dtype=np.uint32)
    if ranks.size == 0:
        ranks = np.zeros(1, dtype=np.uint32)
    lists = np.zeros(max(n, 1), dtype=np.uint32)
    starts = np.zeros(64, dtype=np.int64)
    oracle_lib().oracle_class_lists(cover, p.ctypes.data, n, ranks.ctypes.data, lists.ctypes.data, starts.ctypes.data)
    c = oracle_lib().oracle_cover_classes(cover)
    return lists[:n], starts[:c + 1]


def suffix_array(text, cover=32):
    p, n = _padded(text)
    sa = np.empty(max(n, 1), dtype=np.int32)
    if oracle_lib().oracle_suffix_array(cover, p.ctypes.data, n, sa.ctypes.data) != 0:
        raise RuntimeError("oracle construction failed")
    return sa[:n]


# ---- generators of the reference selftest (src/cli/selftest.cpp:30-108) -------------------------

class Rng:
    """splitmix64 with a fixed seed, as the reference selftest's Rng."""

    def __init__(self, seed):
        self.state = np.array([seed], dtype=np.uint64)

    def next(self):
        return int(oracle_lib().oracle_splitmix_next(self.state.ctypes.data))

    def next_below(self, bound):
        return self.next() % bound

    def random_text(self, n, sigma):
        out = np.empty(max(n, 1), dtype=np.uint8)
        oracle_lib().oracle_random_text(self.state.ctypes.data, n, sigma, out.ctypes.data)
        return out[:n].tobytes()

    def periodic_text(self, n, period):
        out = np.empty(max(n, 1), dtype=np.uint8)
        oracle_lib().oracle_periodic_text(self.state.ctypes.data, n, period, out.ctypes.data)
        return out[:n].tobytes()


def fibonacci_text(n, a=ord("a"), b=ord("b")):
    out = np.empty(max(n, 1), dtype=np.uint8)
    oracle_lib().oracle_fibonacci_text(n, a, b, out.ctypes.data)
    return out[:n].tobytes()


def thue_morse_text(n):
    out = np.empty(max(n, 1), dtype=np.uint8)
    oracle_lib().oracle_thue_morse_text(n, out.ctypes.data)
    return out[:n].tobytes()


# ---- the reference build -------------------------------------------------------------------------

def ref_suffix_array(text, cover=128):
    lib = ref_lib()
    p, n = _padded(text)
    sa = np.empty(max(n, 1), dtype=np.int32)
    if lib.ref_build_sa(p.ctypes.data, n, sa.ctypes.data, cover) != 0:
        raise RuntimeError("reference build refused")
    return sa[:n]


def ref_bwt(text, cover=128):
    lib = ref_lib()
    p, n = _padded(text)
    out = np.empty(max(n, 1), dtype=np.uint8)
    sa = np.empty(max(n, 1), dtype=np.int32)
    primary = lib.ref_build_bwt(p.ctypes.data, n, out.ctypes.data, sa.ctypes.data, cover)
    return out[:n].tobytes(), primary


def ref_check_sa(text, sa, mode=0):
    p, n = _padded(text)
    sa = np.ascontiguousarray(sa, dtype=np.int32)
    return ref_lib().ref_check_sa(p.ctypes.data, n, sa.ctypes.data, mode)


def ref_check_bwt(text, bwt, primary):
    p, n = _padded(text)
    b = np.frombuffer(bytes(bwt), dtype=np.uint8) if n else np.zeros(1, dtype=np.uint8)
    return ref_lib().ref_check_bwt(p.ctypes.data, n, b.ctypes.data, primary)


def ref_outputs(texts, cover=128, workers=None, bwt=True):
    """The reference build's suffix array of every text, and the BWT read off it by bwt_from_sa,
    computed by a pool of threads that each run the reference single-threaded (texts are
    independent; the reference's output does not depend on its thread count). Returns a list of
    (sa, bwt bytes, primary) in the order of texts (bwt and primary None when bwt is False)."""
    import concurrent.futures

    lib = ref_lib()
    if lib is None:
        raise RuntimeError("the reference build (oracle/_ref) is missing")
    # omp_set_num_threads sets the calling thread's own team size, so each pool thread runs the
    # reference on one thread while other callers keep theirs.
    workers = workers or min(16, os.cpu_count() or 1)
    order = sorted(range(len(texts)), key=lambda i: -len(texts[i]))

    def one(i):
        lib.ref_set_threads(1)
        sa = ref_suffix_array(texts[i], cover)
        return i, (sa, *(bwt_from_sa(texts[i], sa) if bwt else (None, None)))

    out = [None] * len(texts)
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for i, result in pool.map(one, order):
            out[i] = result
    return out
