"""Python binding of the MI355X suffix array / BWT construction (the C ABI of include/sadcs_hip.h).

The names mirror the reference's GPU interface, src/cuda/gpusa.h: gpu_suffix_array, gpu_bwt,
gpu_rank_sample, GpuContext, GpuStatus, GpuStats, gpu_required_bytes, gpu_last_error. All work
runs in the HIP library built by sadcs_amd/build.py; importing this module without it raises.
"""
import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libsadcs_hip.so")


class GpuStatus(enum.IntEnum):
    """gpusa.h:22-27."""
    Ok = 0
    Unavailable = 1
    NotEnoughMemory = 2
    Failed = 3


class GpuDevice(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("compute_major", ctypes.c_int32), ("compute_minor", ctypes.c_int32),
                ("total_bytes", ctypes.c_int64), ("free_bytes", ctypes.c_int64)]


class GpuStats(ctypes.Structure):
    """gpusa.h:37-53, plus the cover used."""
    _fields_ = [("device_bytes", ctypes.c_int64), ("ranking_bytes", ctypes.c_int64),
                ("class_sort_bytes", ctypes.c_int64), ("merge_bytes", ctypes.c_int64),
                ("pinned", ctypes.c_int32), ("character_passes", ctypes.c_int32),
                ("doubling_rounds", ctypes.c_int32), ("cover", ctypes.c_int32),
                ("setup_seconds", ctypes.c_double), ("upload_seconds", ctypes.c_double),
                ("device_seconds", ctypes.c_double), ("ranking_seconds", ctypes.c_double),
                ("class_sort_seconds", ctypes.c_double), ("merge_seconds", ctypes.c_double),
                ("bwt_seconds", ctypes.c_double), ("download_seconds", ctypes.c_double),
                ("merge_kernel_seconds", ctypes.c_double), ("merge_kernel_elements", ctypes.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class GpuParams(ctypes.Structure):
    _fields_ = [("cover", ctypes.c_int32), ("splitter_spacing", ctypes.c_uint32), ("tile_splitters", ctypes.c_uint32),
                ("tile_threads", ctypes.c_uint32)]


class ServerParams(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_int64), ("max_batch_bytes", ctypes.c_int64), ("max_batch_texts", ctypes.c_int32),
                ("workers", ctypes.c_int32), ("cover", ctypes.c_int32), ("want_sa", ctypes.c_int32), ("want_bwt", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int64), ("texts", ctypes.c_int64), ("bytes", ctypes.c_int64),
                ("upload_seconds", ctypes.c_double), ("device_seconds", ctypes.c_double),
                ("download_seconds", ctypes.c_double), ("copy_out_seconds", ctypes.c_double),
                ("batch_seconds", ctypes.c_double)]


class GpuError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{GpuStatus(status).name}: {message}")
        self.status = GpuStatus(status)


def _load():
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(f"{LIBRARY_PATH} is missing: build it with `python sadcs_amd/build.py`")
    lib = ctypes.CDLL(LIBRARY_PATH)
    P, i32, i64, u8p = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    lib.sadcs_hip_query.argtypes = [ctypes.POINTER(GpuDevice)]
    lib.sadcs_hip_query.restype = ctypes.c_int
    lib.sadcs_hip_sample_size.argtypes = [i32, i32]
    lib.sadcs_hip_sample_size.restype = i64
    lib.sadcs_hip_required_bytes.argtypes = [i32]
    lib.sadcs_hip_required_bytes.restype = i64
    lib.sadcs_hip_ctx_create.argtypes = [i32, ctypes.POINTER(GpuParams), ctypes.POINTER(P)]
    lib.sadcs_hip_ctx_create.restype = ctypes.c_int
    lib.sadcs_hip_ctx_destroy.argtypes = [P]
    lib.sadcs_hip_ctx_destroy.restype = None
    lib.sadcs_hip_build.argtypes = [P, u8p, i32, P, P, ctypes.POINTER(i32), ctypes.POINTER(GpuStats)]
    lib.sadcs_hip_build.restype = ctypes.c_int
    lib.sadcs_hip_build_device.argtypes = [P, P, i32, P, P, ctypes.POINTER(i32), ctypes.POINTER(GpuStats)]
    lib.sadcs_hip_build_device.restype = ctypes.c_int
    lib.sadcs_hip_rank_sample.argtypes = [u8p, i32, P, ctypes.POINTER(GpuStats)]
    lib.sadcs_hip_rank_sample.restype = ctypes.c_int
    lib.sadcs_hip_alloc_pinned.argtypes = [ctypes.c_size_t]
    lib.sadcs_hip_alloc_pinned.restype = P
    lib.sadcs_hip_free_pinned.argtypes = [P]
    lib.sadcs_hip_free_pinned.restype = None
    lib.sadcs_hip_last_error.argtypes = []
    lib.sadcs_hip_last_error.restype = ctypes.c_char_p
    lib.sadcs_hip_batch_build.argtypes = [P, u8p, ctypes.POINTER(i64), i32, P, P, P, ctypes.POINTER(GpuStats)]
    lib.sadcs_hip_batch_build.restype = ctypes.c_int
    lib.sadcs_hip_batch_required_bytes.argtypes = [ctypes.POINTER(i64), i32]
    lib.sadcs_hip_batch_required_bytes.restype = i64
    lib.sadcs_hip_batch_layout.argtypes = [ctypes.POINTER(i64), i32, ctypes.POINTER(i64)]
    lib.sadcs_hip_batch_layout.restype = i64
    lib.sadcs_hip_batch_build_device.argtypes = [P, P, ctypes.POINTER(i64), i32, P, P, P, ctypes.POINTER(GpuStats)]
    lib.sadcs_hip_batch_build_device.restype = ctypes.c_int
    lib.sadcs_hip_server_create.argtypes = [ctypes.POINTER(ServerParams), ctypes.POINTER(P)]
    lib.sadcs_hip_server_create.restype = ctypes.c_int
    lib.sadcs_hip_server_submit.argtypes = [P, P, i64]
    lib.sadcs_hip_server_submit.restype = i64
    lib.sadcs_hip_server_submit_into.argtypes = [P, P, i64, P, P]
    lib.sadcs_hip_server_submit_into.restype = i64
    lib.sadcs_hip_server_wait.argtypes = [P, i64, P, P, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]
    lib.sadcs_hip_server_wait.restype = ctypes.c_int
    lib.sadcs_hip_server_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
    lib.sadcs_hip_server_stats.restype = ctypes.c_int
    lib.sadcs_hip_server_counters.argtypes = [P, ctypes.POINTER(i64)]
    lib.sadcs_hip_server_counters.restype = None
    lib.sadcs_hip_server_destroy.argtypes = [P]
    lib.sadcs_hip_server_destroy.restype = None
    lib.sadcs_hip_tune.argtypes = [P, ctypes.POINTER(i64), i32, ctypes.POINTER(GpuParams), ctypes.POINTER(GpuParams),
                                   ctypes.POINTER(ctypes.c_double)]
    lib.sadcs_hip_tune.restype = ctypes.c_int
    lib.sadcs_hip_device_required_bytes.argtypes = [i32, i32, i32]
    lib.sadcs_hip_device_required_bytes.restype = i64
    lib.sadcs_hip_ctx_create_batch.argtypes = [ctypes.POINTER(i64), i32, ctypes.POINTER(GpuParams), ctypes.POINTER(P)]
    lib.sadcs_hip_ctx_create_batch.restype = ctypes.c_int
    lib.sadcs_hip_device_alloc.argtypes = [ctypes.c_size_t]
    lib.sadcs_hip_device_alloc.restype = P
    lib.sadcs_hip_device_free.argtypes = [P]
    lib.sadcs_hip_device_free.restype = None
    lib.sadcs_hip_memcpy.argtypes = [P, P, ctypes.c_size_t, i32]
    lib.sadcs_hip_memcpy.restype = ctypes.c_int
    lib.sadcs_hip_memset.argtypes = [P, i32, ctypes.c_size_t]
    lib.sadcs_hip_memset.restype = ctypes.c_int
    lib.sadcs_hip_synchronize.argtypes = []
    lib.sadcs_hip_synchronize.restype = ctypes.c_int
    lib.sadcs_hip_check_device.argtypes = [P, i32, P, P, i32, ctypes.POINTER(i64)]
    lib.sadcs_hip_check_device.restype = ctypes.c_int
    lib.sadcs_hip_check_device_noalloc.argtypes = [P, i32, P, P, i32, ctypes.POINTER(i64)]
    lib.sadcs_hip_check_device_noalloc.restype = ctypes.c_int
    lib.sadcs_hip_gather_lines.argtypes = [P, i64, i64, i32]
    lib.sadcs_hip_gather_lines.restype = ctypes.c_int
    return lib


lib = _load()

# The symbols include/sadcs_hip.h declares (checked by the CPU tests).
EXPORTED_SYMBOLS = (
    "sadcs_hip_query", "sadcs_hip_sample_size", "sadcs_hip_required_bytes", "sadcs_hip_ctx_create",
    "sadcs_hip_ctx_destroy", "sadcs_hip_build", "sadcs_hip_build_device", "sadcs_hip_rank_sample",
    "sadcs_hip_alloc_pinned", "sadcs_hip_free_pinned", "sadcs_hip_last_error", "sadcs_hip_batch_build",
    "sadcs_hip_batch_required_bytes", "sadcs_hip_batch_layout", "sadcs_hip_batch_build_device",
    "sadcs_hip_server_create", "sadcs_hip_server_submit", "sadcs_hip_server_submit_into", "sadcs_hip_server_wait", "sadcs_hip_server_counters",
    "sadcs_hip_server_destroy", "sadcs_hip_ctx_create_batch", "sadcs_hip_device_required_bytes", "sadcs_hip_tune", "sadcs_hip_device_alloc", "sadcs_hip_device_free", "sadcs_hip_memcpy",
    "sadcs_hip_memset", "sadcs_hip_synchronize", "sadcs_hip_check_device", "sadcs_hip_check_device_noalloc",
    "sadcs_hip_server_stats", "sadcs_hip_gather_lines",
)


def gpu_last_error():
    """gpusa.h:122-123."""
    return lib.sadcs_hip_last_error().decode(errors="replace")


def _check(status):
    if status != GpuStatus.Ok:
        raise GpuError(status, gpu_last_error())


def gpu_query_device():
    """gpusa.h:76: (status, GpuDevice)."""
    device = GpuDevice()
    status = lib.sadcs_hip_query(ctypes.byref(device))
    return GpuStatus(status), device


def gpu_sample_size(n, cover=32):
    """gpusa.h:58-60."""
    return lib.sadcs_hip_sample_size(n, cover)


def gpu_required_bytes(n):
    """gpusa.h:79."""
    return lib.sadcs_hip_required_bytes(n)


def _as_bytes(text):
    if isinstance(text, np.ndarray):
        if text.dtype != np.uint8:
            raise TypeError("text must be uint8")
        return np.ascontiguousarray(text)
    return np.frombuffer(bytes(text), dtype=np.uint8)


class GpuContext:
    """gpusa.h:84-95: device memory for inputs up to capacity bytes, kept across calls."""

    def __init__(self, capacity, cover=32, splitter_spacing=0, tile_splitters=0, batch_lengths=None, tile_threads=0):
        """capacity: the longest single text; batch_lengths: plan for a batch of these lengths instead
        (sadcs_hip_ctx_create_batch)."""
        params = GpuParams(cover, splitter_spacing, tile_splitters, tile_threads)
        handle = ctypes.c_void_p()
        if batch_lengths is not None:
            lengths = np.ascontiguousarray(batch_lengths, dtype=np.int64)
            _check(lib.sadcs_hip_ctx_create_batch(lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), lengths.size,
                                                  ctypes.byref(params), ctypes.byref(handle)))
            capacity = int(lengths.max()) if lengths.size else 0
        else:
            _check(lib.sadcs_hip_ctx_create(int(capacity), ctypes.byref(params), ctypes.byref(handle)))
        self.handle = handle
        self.capacity = capacity
        self.cover = cover

    def close(self):
        if self.handle:
            lib.sadcs_hip_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _handle(context):
    return context.handle if context is not None else None


def gpu_build(text, suffix_array=True, bwt=True, context=None, stats=None):
    """Suffix array and/or BWT of text in one construction.

    Returns (sa or None, bwt bytes or None, primary). Raises GpuError on failure.
    """
    t = _as_bytes(text)
    n = t.size
    sa_out = np.empty(n, dtype=np.int32) if suffix_array else None
    bwt_out = np.empty(n, dtype=np.uint8) if bwt else None
    primary = ctypes.c_int32(-1)
    st = stats if stats is not None else GpuStats()
    status = lib.sadcs_hip_build(_handle(context), t.ctypes.data if n else None, n,
                                 sa_out.ctypes.data if sa_out is not None and n else None,
                                 bwt_out.ctypes.data if bwt_out is not None and n else None,
                                 ctypes.byref(primary), ctypes.byref(st))
    _check(status)
    return sa_out, (bwt_out.tobytes() if bwt_out is not None else None), primary.value


def gpu_suffix_array(text, context=None, stats=None):
    """gpusa.h:101-102: the suffix array (int32 numpy array)."""
    return gpu_build(text, True, False, context, stats)[0]


def gpu_bwt(text, context=None, stats=None):
    """gpusa.h:107-108: (bwt bytes, primary index) in buildBwt()'s convention."""
    _, b, primary = gpu_build(text, False, True, context, stats)
    return b, primary


def gpu_build_status(text, n=None, context=None, suffix_array=True, bwt=False):
    """The raw status of a build (for refusal tests); n may exceed len(text) only for n <= 0."""
    t = _as_bytes(text)
    n = t.size if n is None else n
    sa_out = np.empty(max(n, 1), dtype=np.int32) if suffix_array else None
    bwt_out = np.empty(max(n, 1), dtype=np.uint8) if bwt else None
    primary = ctypes.c_int32(-1)
    st = GpuStats()
    status = lib.sadcs_hip_build(_handle(context), t.ctypes.data if t.size else None, n,
                                 sa_out.ctypes.data if sa_out is not None else None,
                                 bwt_out.ctypes.data if bwt_out is not None else None,
                                 ctypes.byref(primary), ctypes.byref(st))
    return GpuStatus(status)


def gpu_rank_sample(text, stats=None):
    """gpusa.h:110-114: ranks[pack(p)] of the cover-32 sample positions p."""
    t = _as_bytes(text)
    n = t.size
    ranks = np.empty(max(gpu_sample_size(n), 1), dtype=np.uint32)
    st = stats if stats is not None else GpuStats()
    _check(lib.sadcs_hip_rank_sample(t.ctypes.data if n else None, n, ranks.ctypes.data, ctypes.byref(st)))
    return ranks[:gpu_sample_size(n)]


def gpu_build_device(context, d_text, n, d_sa=None, d_bwt=None, stats=None):
    """Device-resident build. d_text: device pointer to n bytes plus 64 zero bytes (8-byte aligned);
    d_sa (4n bytes) and d_bwt (n bytes): device pointers or None. Returns the primary index."""
    primary = ctypes.c_int32(-1)
    st = stats if stats is not None else GpuStats()
    _check(lib.sadcs_hip_build_device(_handle(context), ctypes.c_void_p(d_text), int(n),
                                      ctypes.c_void_p(d_sa) if d_sa else None,
                                      ctypes.c_void_p(d_bwt) if d_bwt else None,
                                      ctypes.byref(primary), ctypes.byref(st)))
    return primary.value


def gpu_batch_build(texts, suffix_array=True, bwt=True, context=None, stats=None):
    """Many texts in one construction. Returns a list of (sa, bwt bytes, primary) per text."""
    arrays = [_as_bytes(t) for t in texts]
    lengths = np.array([a.size for a in arrays], dtype=np.int64)
    if (lengths < 1).any():
        raise ValueError("batched texts need at least one byte each")
    joined = np.concatenate(arrays) if arrays else np.empty(0, dtype=np.uint8)
    total = int(lengths.sum())
    sa_out = np.empty(max(total, 1), dtype=np.int32) if suffix_array else None
    bwt_out = np.empty(max(total, 1), dtype=np.uint8) if bwt else None
    primaries = np.empty(max(len(arrays), 1), dtype=np.int32)
    st = stats if stats is not None else GpuStats()
    _check(lib.sadcs_hip_batch_build(_handle(context), joined.ctypes.data, lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     len(arrays), sa_out.ctypes.data if sa_out is not None else None,
                                     bwt_out.ctypes.data if bwt_out is not None else None,
                                     primaries.ctypes.data, ctypes.byref(st)))
    results = []
    offset = 0
    for i, n in enumerate(lengths):
        n = int(n)
        results.append((sa_out[offset:offset + n].copy() if sa_out is not None else None,
                        bwt_out[offset:offset + n].tobytes() if bwt_out is not None else None,
                        int(primaries[i]) if bwt_out is not None else None))
        offset += n
    return results


def gpu_allocate_pinned(size):
    """gpusa.h:116-120: (pointer, numpy view) of page-locked host memory, or (None, None)."""
    pointer = lib.sadcs_hip_alloc_pinned(size)
    if not pointer:
        return None, None
    view = np.ctypeslib.as_array(ctypes.cast(pointer, ctypes.POINTER(ctypes.c_uint8)), shape=(max(size, 1),))
    return pointer, view


def gpu_free_pinned(pointer):
    lib.sadcs_hip_free_pinned(pointer)


def gpu_batch_layout(lengths):
    """(offsets, span) of a device batch: text i at offsets[i]; the buffer needs span + 64 zero bytes."""
    lengths = np.ascontiguousarray(lengths, dtype=np.int64)
    offsets = np.zeros(max(lengths.size, 1), dtype=np.int64)
    span = lib.sadcs_hip_batch_layout(lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), lengths.size,
                                      offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return offsets[:lengths.size], int(span)


def gpu_batch_build_device(context, d_texts, lengths, d_sa=None, d_bwt=None, stats=None):
    """A device-resident batch (texts at gpu_batch_layout offsets). Returns the primaries (numpy)."""
    lengths = np.ascontiguousarray(lengths, dtype=np.int64)
    primaries = np.full(max(lengths.size, 1), -1, dtype=np.int32)
    st = stats if stats is not None else GpuStats()
    _check(lib.sadcs_hip_batch_build_device(_handle(context), ctypes.c_void_p(d_texts),
                                            lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), lengths.size,
                                            ctypes.c_void_p(d_sa) if d_sa else None,
                                            ctypes.c_void_p(d_bwt) if d_bwt else None,
                                            primaries.ctypes.data, ctypes.byref(st)))
    return primaries[:lengths.size]


class Server:
    """A continuous-batching construction server (sadcs_hip_server_*): submit() queues a text and
    returns an id; wait(id) blocks until its suffix array / BWT are ready."""

    def __init__(self, memory_budget=0, workers=0, max_batch_bytes=0, max_batch_texts=0, cover=32,
                 suffix_array=True, bwt=True):
        params = ServerParams(memory_budget, max_batch_bytes, max_batch_texts, workers, cover, int(suffix_array), int(bwt))
        handle = ctypes.c_void_p()
        _check(lib.sadcs_hip_server_create(ctypes.byref(params), ctypes.byref(handle)))
        self.handle = handle
        self.suffix_array = suffix_array
        self.bwt = bwt
        self._outputs = {}  # id -> (n, sa, bwt): arrays the server writes into, alive until wait

    def submit(self, text):
        """Queues a text; its outputs go straight into arrays allocated here (submit_into)."""
        t = _as_bytes(text)
        n = t.size
        sa = np.empty(max(n, 1), dtype=np.int32) if self.suffix_array else None
        b = np.empty(max(n, 1), dtype=np.uint8) if self.bwt else None
        rid = lib.sadcs_hip_server_submit_into(self.handle, t.ctypes.data if n else None, n,
                                               sa.ctypes.data if sa is not None else None,
                                               b.ctypes.data if b is not None else None)
        if rid < 0:
            _check(-rid)
        self._outputs[rid] = (n, sa, b)
        return rid

    def wait(self, rid):
        """Returns (sa, bwt bytes, primary, latency seconds, queued seconds, batch size)."""
        n, sa, b = self._outputs[rid]
        primary = ctypes.c_int32(-1)
        latency = ctypes.c_double(0)
        queued = ctypes.c_double(0)
        batch = ctypes.c_int32(0)
        try:
            _check(lib.sadcs_hip_server_wait(self.handle, rid, None, None, ctypes.byref(primary),
                                             ctypes.byref(latency), ctypes.byref(queued), ctypes.byref(batch)))
        finally:
            del self._outputs[rid]
        return (sa[:n] if sa is not None else None, b[:n].tobytes() if b is not None else None, primary.value,
                latency.value, queued.value, batch.value)

    def counters(self):
        c = (ctypes.c_int64 * 5)()
        lib.sadcs_hip_server_counters(self.handle, c)
        return {"completed": c[0], "batches": c[1], "peak_pool_bytes": c[2], "budget": c[3], "max_concurrent_batches": c[4]}

    def stats(self):
        """Per-phase seconds summed over the batches so far (sadcs_hip_server_stats)."""
        st = ServerStats()
        _check(lib.sadcs_hip_server_stats(self.handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def close(self):
        """Finishes the queued requests (they write into this object's arrays), then stops."""
        if self.handle:
            lib.sadcs_hip_server_destroy(self.handle)
            self.handle = None
        self._outputs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "handle", None):
            self.close()


class DeviceBuffer:
    """Device memory from the library's own HIP runtime (sadcs_hip_device_alloc)."""

    def __init__(self, size):
        self.size = int(size)
        self.ptr = lib.sadcs_hip_device_alloc(self.size)
        if not self.ptr:
            raise GpuError(GpuStatus.NotEnoughMemory, f"device allocation of {self.size} bytes failed")

    @classmethod
    def from_host(cls, array, extra=0, zero=True):
        a = np.ascontiguousarray(array)
        buf = cls(a.nbytes + extra)
        if zero and extra:
            _check(lib.sadcs_hip_memset(buf.ptr, 0, buf.size))
        if a.nbytes:
            _check(lib.sadcs_hip_memcpy(buf.ptr, a.ctypes.data, a.nbytes, 1))
        return buf

    def zero(self):
        _check(lib.sadcs_hip_memset(self.ptr, 0, self.size))

    def to_host(self, dtype=np.uint8, count=None):
        dtype = np.dtype(dtype)
        count = self.size // dtype.itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _check(lib.sadcs_hip_memcpy(out.ctypes.data, self.ptr, out.nbytes, 2))
        return out

    def free(self):
        if self.ptr:
            lib.sadcs_hip_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gpu_synchronize():
    _check(lib.sadcs_hip_synchronize())


def gpu_copy_device(dst, src, size):
    _check(lib.sadcs_hip_memcpy(dst, src, int(size), 3))


def gpu_gather_lines(d_lines, line_count, gathers, whole_lines=True):
    """sadcs_hip_gather_lines: `gathers` reads of random 128-byte lines (whole, or 96 bytes of each
    as the top-level merge reads records), the known-size kernel for scaling HBM counters."""
    _check(lib.sadcs_hip_gather_lines(ctypes.c_void_p(d_lines), int(line_count), int(gathers), int(bool(whole_lines))))


CHECK_NAMES = {0: "pass", 1: "entry out of range", 2: "duplicate entry", 3: "SA[index] and SA[index + 1] out of order",
               4: "primary index differs", 5: "BWT byte differs"}


def gpu_check_device(d_text, n, d_sa, d_bwt=None, primary=-1):
    """sadcs_hip_check_device: checks a device-resident suffix array (and BWT with its primary)
    on the GPU, as the reference's fast checker does (src/common/checker.cpp:53-109).
    Returns (category, first failing index): (0, -1) on a pass; 1 out of range, 2 duplicate,
    3 out of order (the index the reference reports), 4 primary, 5 BWT byte (its index)."""
    result = (ctypes.c_int64 * 2)()
    _check(lib.sadcs_hip_check_device(ctypes.c_void_p(d_text), int(n), ctypes.c_void_p(d_sa),
                                      ctypes.c_void_p(d_bwt) if d_bwt else None, int(primary), result))
    return int(result[0]), int(result[1])


def gpu_tune(d_texts, lengths, cover=32):
    """sadcs_hip_tune: the fastest merge configuration for a device-resident sample (one text with
    64 zero bytes after it, or a batch in the gpu_batch_layout layout). Returns
    (dict(cover, splitter_spacing, tile_splitters, tile_threads), device seconds of the best build)."""
    lengths = np.ascontiguousarray(np.atleast_1d(lengths), dtype=np.int64)
    base = GpuParams(cover, 0, 0, 0)
    best = GpuParams()
    seconds = ctypes.c_double(0)
    _check(lib.sadcs_hip_tune(ctypes.c_void_p(d_texts), lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), lengths.size,
                              ctypes.byref(base), ctypes.byref(best), ctypes.byref(seconds)))
    return {"cover": best.cover, "splitter_spacing": best.splitter_spacing, "tile_splitters": best.tile_splitters,
            "tile_threads": best.tile_threads}, seconds.value
