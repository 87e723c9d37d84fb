"""The benchmark configurations' synthetic texts (sadcs_amd/csrc/workloads.cpp, libsadcs_work.so).

Each text is a pure function of (seed, n). Configurations:
  C1  16 MiB i.i.d. uniform ACGT
  C2  256 MiB DNA-like: A/T 0.3, C/G 0.2, ~30% copies of earlier segments (1-10 KiB, 1% substitutions),
      'N' runs (100-5000) on ~0.1%
  C3  1 GiB order-3 Markov text over 96 printable bytes, Zipf(1.1) per context, ~5% exact repeats
      (64 B - 64 KiB)
  C4  4096 texts, lengths log-uniform in [1 KiB, 4 MiB]; even index C2-style, odd index C3-style
  C5  a request stream: lognormal lengths (median 256 KiB, sigma 1.5, capped at 64 MiB);
      45% ACGT, 45% protein-like, 10% periodic
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "libsadcs_work.so")
if not os.path.exists(_PATH):
    raise ImportError(f"{_PATH} is missing: build it with `python sadcs_amd/build.py`")
_lib = ctypes.CDLL(_PATH)
_u64, _i64, _P = ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p
for _name in ("sadcs_gen_dna_uniform", "sadcs_gen_markov", "sadcs_gen_protein", "sadcs_gen_periodic"):
    getattr(_lib, _name).argtypes = [_u64, _P, _i64]
    getattr(_lib, _name).restype = None
_lib.sadcs_gen_dna_like.argtypes = [_u64, _P, _i64, _P]
_lib.sadcs_gen_dna_like.restype = None
_lib.sadcs_gen_log_uniform_lengths.argtypes = [_u64, ctypes.c_int32, _i64, _i64, _P]
_lib.sadcs_gen_log_uniform_lengths.restype = None
_lib.sadcs_gen_serving_requests.argtypes = [_u64, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _i64, _P, _P]
_lib.sadcs_gen_serving_requests.restype = None
_lib.sadcs_gen_batch.argtypes = [_u64, ctypes.c_int32, _P, _P, _P]
_lib.sadcs_gen_batch.restype = None

MiB = 1 << 20


def _buffer(n, padding):
    out = np.zeros(n + padding, dtype=np.uint8)
    return out


def dna_uniform(n, seed=1, padding=0):
    out = _buffer(n, padding)
    _lib.sadcs_gen_dna_uniform(seed, out.ctypes.data, n)
    return out


def dna_like(n, seed=2, padding=0):
    out = _buffer(n, padding)
    stats = np.zeros(2, dtype=np.float64)
    _lib.sadcs_gen_dna_like(seed, out.ctypes.data, n, stats.ctypes.data)
    return out


def markov(n, seed=3, padding=0):
    out = _buffer(n, padding)
    _lib.sadcs_gen_markov(seed, out.ctypes.data, n)
    return out


def protein(n, seed=5, padding=0):
    out = _buffer(n, padding)
    _lib.sadcs_gen_protein(seed, out.ctypes.data, n)
    return out


def periodic(n, seed=5, padding=0):
    out = _buffer(n, padding)
    _lib.sadcs_gen_periodic(seed, out.ctypes.data, n)
    return out


CONFIGS = {
    "c1": ("16 MiB i.i.d. uniform ACGT", 16 * MiB, dna_uniform, 1),
    "c2": ("256 MiB DNA-like (30% repeats, N runs)", 256 * MiB, dna_like, 2),
    "c3": ("1 GiB order-3 Markov text over 96 printable bytes, Zipf(1.1), 5% repeats", 1024 * MiB, markov, 3),
}


def config_text(name, padding=0, n=None):
    """The text of configuration c1, c2 or c3 (optionally a prefix of n bytes of the same generator)."""
    _, size, gen, seed = CONFIGS[name]
    return gen(size if n is None else n, seed, padding)


def batch_lengths(count=4096, seed=4, lo=1024, hi=4 * MiB):
    lengths = np.zeros(count, dtype=np.int64)
    _lib.sadcs_gen_log_uniform_lengths(seed, count, lo, hi, lengths.ctypes.data)
    return lengths


def batch_buffer(count=4096, seed=4, lo=1024, hi=4 * MiB, offsets=None, padding=64):
    """C4 in one buffer: (buffer, offsets, lengths), text i at buffer[offsets[i]:offsets[i] + lengths[i]]
    (packed back to back unless offsets are given; the buffer ends with `padding` zero bytes).
    Texts are generated in parallel; each equals batch_texts' text i."""
    lengths = batch_lengths(count, seed, lo, hi)
    if offsets is None:
        offsets = np.concatenate(([0], np.cumsum(lengths)[:-1])).astype(np.int64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    span = int((offsets + lengths).max()) if count else 0
    buffer = np.zeros(span + padding, dtype=np.uint8)
    _lib.sadcs_gen_batch(seed, count, lengths.ctypes.data, offsets.ctypes.data, buffer.ctypes.data)
    return buffer, offsets, lengths


def batch_texts(count=4096, seed=4, lo=1024, hi=4 * MiB):
    """C4: texts of log-uniform lengths; even index DNA-like (seed * 1000003 + i), odd index Markov."""
    buffer, offsets, lengths = batch_buffer(count, seed, lo, hi)
    return [buffer[o:o + n] for o, n in zip(offsets.tolist(), lengths.tolist())]


def serving_requests(count, seed=5, median=256 * 1024, sigma=1.5, cap=64 * MiB):
    """C5: (lengths, kinds) with kind 0 ACGT, 1 protein-like, 2 periodic."""
    lengths = np.zeros(count, dtype=np.int64)
    kinds = np.zeros(count, dtype=np.int32)
    _lib.sadcs_gen_serving_requests(seed, count, median, sigma, cap, lengths.ctypes.data, kinds.ctypes.data)
    return lengths, kinds


def serving_text(length, kind, seed):
    return (dna_uniform, protein, periodic)[kind](int(length), seed)
