/*
 * sadcs_hip.h - C ABI of the MI355X (gfx950) suffix array and BWT construction.
 *
 * Plain C: pointers, sizes and status codes, no HIP or C++ types. Each entry point replaces one
 * call of the reference's GPU interface, src/cuda/gpusa.h (cited per function), with the same
 * meaning of arguments, the same refusal rules and the same status order:
 *   SADCS_OK, SADCS_UNAVAILABLE, SADCS_NOT_ENOUGH_MEMORY, SADCS_FAILED  (gpusa.h:22-27).
 * The C++ API of gpusa.h itself is provided on top of this by gpusa.cpp.
 *
 * Output conventions (divsufsort's, as the reference): the suffix array holds n int32 positions;
 * the BWT is the BWT of text plus a sentinel without the sentinel, n bytes, with the primary
 * index (the sentinel's row) returned separately (gpusa.h:104-108).
 */
#ifndef SADCS_HIP_H
#define SADCS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SADCS_OK = 0,
    SADCS_UNAVAILABLE = 1,         /* no device or no usable driver */
    SADCS_NOT_ENOUGH_MEMORY = 2,   /* the input does not fit in free device memory or the context */
    SADCS_FAILED = 3               /* a HIP call or an internal check failed; see sadcs_hip_last_error() */
};

/* gpusa.h:29-35 (GpuDevice). */
typedef struct sadcs_device {
    char name[256];
    int32_t compute_major;
    int32_t compute_minor;
    int64_t total_bytes;
    int64_t free_bytes;
} sadcs_device;

/* gpusa.h:37-53 (GpuStats); zeroed at the start of every call. Seconds are host wall clock
 * around device-synchronised phases, as in the reference; device_seconds is text resident on the
 * device to outputs complete on the device. */
typedef struct sadcs_stats {
    int64_t device_bytes;        /* peak device memory of the call, including caller-owned device buffers */
    int64_t ranking_bytes;
    int64_t class_sort_bytes;
    int64_t merge_bytes;
    int32_t pinned;              /* transfers went through pinned staging or pinned caller buffers */
    int32_t character_passes;    /* sample ranking: character passes */
    int32_t doubling_rounds;     /* sample ranking: prefix-doubling rounds */
    int32_t cover;               /* the difference cover modulus used */
    double setup_seconds;
    double upload_seconds;
    double device_seconds;
    double ranking_seconds;
    double class_sort_seconds;
    double merge_seconds;
    double bwt_seconds;
    double download_seconds;
    double merge_kernel_seconds; /* HIP-event time of the top-level tile-merge kernel on the context's stream */
    int64_t merge_kernel_elements; /* elements that kernel merged (the whole text) */
} sadcs_stats;

/* Construction parameters; zero fields take the defaults (of the cover, for G and tile threads). */
typedef struct sadcs_params {
    int32_t cover;               /* 32 (default) or 128 */
    uint32_t splitter_spacing;   /* G: every G-th element of a list is a splitter; 16 (cover 32), 8 (cover 128) */
    uint32_t tile_splitters;     /* K: sorted splitters per merge tile; the most that fit the tiles in LDS */
    uint32_t tile_threads;       /* threads per merge tile: 512 (two tiles per CU; cover 32) or 256 (four; cover 128) */
} sadcs_params;

typedef struct sadcs_ctx sadcs_ctx;

/* gpusa.h:76 (gpuQueryDevice): the first device. */
int sadcs_hip_query(sadcs_device* device);

/* gpusa.h:58-60 (gpuSampleSize): number of sample positions of a text of n bytes for the cover. */
int64_t sadcs_hip_sample_size(int32_t n, int32_t cover);

/* gpusa.h:79 (gpuRequiredBytes): peak device bytes of a suffix array + BWT build of n bytes,
 * -1 if unknown. */
int64_t sadcs_hip_required_bytes(int32_t n);

/* gpusa.h:84-95 (gpuCreateContext / gpuDestroyContext): device memory reserved for inputs up
 * to capacity bytes, kept across calls. Refuses with SADCS_NOT_ENOUGH_MEMORY, allocating nothing,
 * if the planned peak plus 256 MiB exceeds the free device memory. Not thread-safe: one context per
 * thread. params may be null. Destroy ignores null. */
int sadcs_hip_ctx_create(int32_t capacity, const sadcs_params* params, sadcs_ctx** context);
void sadcs_hip_ctx_destroy(sadcs_ctx* context);

/* gpusa.h:97-108 (gpuSuffixArray and gpuBwt): builds from host memory. sa (n int32) and/or
 * bwt (n bytes, may alias text) select the outputs; both non-null builds both in one merge.
 * primary receives the BWT's primary index (-1 until success; 0 for n = 0). A null context
 * uses a temporary one. n over the context's capacity gives SADCS_NOT_ENOUGH_MEMORY. n <= 0 is Ok. */
int sadcs_hip_build(sadcs_ctx* context, const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt,
                    int32_t* primary, sadcs_stats* stats);

/* The same with the text already on the device and the outputs left there (no host transfers).
 * d_text: n bytes followed by at least 64 zero bytes, 16-byte aligned. d_sa (4n bytes, 4-byte
 * aligned) and d_bwt (n bytes, any alignment) are device buffers or null; until the merge writes
 * them the library uses them as scratch. No counterpart in the reference: the device-resident form
 * of gpuSuffixArray/gpuBwt for callers that keep texts in HBM. */
int sadcs_hip_build_device(sadcs_ctx* context, const void* d_text, int32_t n, void* d_sa, void* d_bwt,
                           int32_t* primary, sadcs_stats* stats);

/* Device bytes the library itself needs for sadcs_hip_build_device on n bytes (cover 32): the
 * caller's text and output buffers are not included, and before the merge writes them the
 * outputs serve as scratch. No counterpart in the reference. */
int64_t sadcs_hip_device_required_bytes(int32_t n, int32_t want_sa, int32_t want_bwt);

/* gpusa.h:110-114 (gpuRankSample): uploads text and ranks the sample only;
 * ranks receives sadcs_hip_sample_size(n, 32) entries: ranks[pack(p)] for each sample position p. */
int sadcs_hip_rank_sample(const uint8_t* text, int32_t n, uint32_t* ranks, sadcs_stats* stats);

/* gpusa.h:116-120 (gpuAllocatePinned / gpuFreePinned). */
void* sadcs_hip_alloc_pinned(size_t bytes);
void sadcs_hip_free_pinned(void* pointer);

/* gpusa.h:122-123 (gpuLastError): "" or a description of the last failure of this thread. */
const char* sadcs_hip_last_error(void);

/* ---- batching: many texts in one construction (no counterpart in the reference) ----
 * texts: the concatenated texts; lengths[i] bytes each (>= 1). Outputs are concatenated in the
 * same order: sa receives lengths[i] int32 per text (local positions), bwt lengths[i] bytes per
 * text, primaries one int32 per text. Either output may be null. The batch runs as one pipeline:
 * every sort and merge is segmented by a text-index key. */
int sadcs_hip_batch_build(sadcs_ctx* context, const uint8_t* texts, const int64_t* lengths, int32_t count,
                          int32_t* sa, uint8_t* bwt, int32_t* primaries, sadcs_stats* stats);

/* Device bytes a batch of these lengths needs. */
int64_t sadcs_hip_batch_required_bytes(const int64_t* lengths, int32_t count);

/* A context whose device memory is planned for a batch of these lengths, so that repeated batch
 * builds (of this layout or any whose plan fits) allocate nothing. Same refusal rule as
 * sadcs_hip_ctx_create; single-text builds up to the longest length also fit. */
int sadcs_hip_ctx_create_batch(const int64_t* lengths, int32_t count, const sadcs_params* params, sadcs_ctx** context);

/* The device layout of a batch: text i starts at offsets[i] (a multiple of 256). Returns the span;
 * a device buffer for the batch holds span bytes plus 64 zero bytes. */
int64_t sadcs_hip_batch_layout(const int64_t* lengths, int32_t count, int64_t* offsets);

/* A batch already on the device in that layout; outputs left on the device, concatenated per text
 * (d_sa: sum of lengths int32, local positions; d_bwt: sum of lengths bytes); primaries on the host. */
int sadcs_hip_batch_build_device(sadcs_ctx* context, const void* d_texts, const int64_t* lengths, int32_t count,
                                 void* d_sa, void* d_bwt, int32_t* primaries, sadcs_stats* stats);

/* ---- autotuner (no counterpart in the reference) ----
 * Builds SA + BWT of a device-resident sample with each candidate merge configuration - tile
 * threads 512 or 256, splitter spacing G in {8, 12, 16, 24, 32} (cover 128: {4, 6, 8, 12}) and tile splitters K at 1, 3/4 and
 * 1/2 of the most that keep the tiles a CU runs at once in its LDS - and writes the fastest (by device time, best of two builds) to *best; the
 * cover comes from base (null: 32). count 1: d_texts is one text of lengths[0] bytes followed by
 * 64 zero bytes; count > 1: a batch in the sadcs_hip_batch_layout layout. best_seconds may be null. */
int sadcs_hip_tune(const void* d_texts, const int64_t* lengths, int32_t count, const sadcs_params* base,
                   sadcs_params* best, double* best_seconds);

/* ---- checking device-resident outputs (no counterpart in the reference's GPU API) ----
 * The reference's fast suffix-array check (src/common/checker.cpp:53-109) on the device: every
 * entry in [0, n), no duplicates, neighbours in order by first character and then by the rank of
 * the following suffix. With d_bwt (may be null), the BWT and primary are checked against the
 * suffix array in divsufsort's convention; a valid suffix array fixes them, so this passes exactly
 * when the reference's inversion check (src/common/checker.cpp:196-245) would. result[0] is the
 * first failing check in that order (SADCS_CHECK_*), result[1] the first failing index of the
 * suffix array (range, duplicate, order: the index the reference reports) or of the BWT, else -1.
 * Allocates 4n bytes of device memory for the inverse; synchronous. */
enum {
    SADCS_CHECK_PASS = 0,
    SADCS_CHECK_RANGE = 1,      /* an entry outside [0, n) */
    SADCS_CHECK_DUPLICATE = 2,  /* an entry that appears twice */
    SADCS_CHECK_ORDER = 3,      /* SA[i] and SA[i + 1] out of order */
    SADCS_CHECK_PRIMARY = 4,    /* primary differs from the row of suffix 0, plus one */
    SADCS_CHECK_BWT = 5         /* a BWT byte differs */
};
int sadcs_hip_check_device(const void* d_text, int32_t n, const void* d_sa, const void* d_bwt, int32_t primary,
                           int64_t* result);

/* sadcs_hip_check_device without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The same checks and results, for n up to 8388608 (2^23) entries; larger n
 * gives SADCS_NOT_ENOUGH_MEMORY and checks nothing. The inverse lives in a device array of the
 * library (32 MiB + 256 B, held by every device the library runs on until the process ends),
 * shared by every call on that device: calls that overlap on one device race and must not be
 * made. Each call clears what it uses; nothing carries over from the last call. */
int sadcs_hip_check_device_noalloc(const void* d_text, int32_t n, const void* d_sa, const void* d_bwt, int32_t primary,
                                   int64_t* result);

/* ---- device memory for callers without their own HIP setup (no counterpart in the reference) ----
 * The library's HIP runtime allocates and copies, so a caller that keeps texts and outputs on the
 * device needs no second runtime in the process. kind: 1 host to device, 2 device to host,
 * 3 device to device. */
void* sadcs_hip_device_alloc(size_t bytes);
void sadcs_hip_device_free(void* pointer);
int sadcs_hip_memcpy(void* dst, const void* src, size_t bytes, int32_t kind);
int sadcs_hip_memset(void* dst, int32_t value, size_t bytes);
int sadcs_hip_synchronize(void);

/* A known-size random gather for scaling the HBM counters of a profiler (no counterpart in the
 * reference): `gathers` reads of random 128-byte lines of d_lines (line_count lines, at most
 * 2^32 - 1): whole lines, or (whole_lines = 0) 96 bytes of each as the top-level merge reads its
 * records. Synchronous; kernel `gatherLinesKernel`. */
int sadcs_hip_gather_lines(const void* d_lines, int64_t line_count, int64_t gathers, int32_t whole_lines);

/* ---- serving: a request loop with continuous batching (no counterpart in the reference) ----
 * Submitted texts queue; whenever one of `workers` device streams is idle, a batch is formed from
 * the front of the queue (at most max_batch_bytes of text and max_batch_texts texts; from a deep
 * queue, about half of the queued bytes per worker but at least 4 MiB, so that every worker has a
 * share and outputs are copied out while other batches run) and admitted
 * if its planned device memory (about 10 bytes per input byte) fits the free part of a pool of
 * memory_budget bytes (0: 80% of the free device memory). Batches on different streams overlap.
 * A text whose plan alone exceeds the budget completes with SADCS_NOT_ENOUGH_MEMORY. */
typedef struct sadcs_server_params {
    int64_t memory_budget;
    int64_t max_batch_bytes;     /* 0: 32 MiB */
    int32_t max_batch_texts;     /* 0: 4096 */
    int32_t workers;             /* 0: 5 */
    int32_t cover;               /* 0: 32 */
    int32_t want_sa;
    int32_t want_bwt;
} sadcs_server_params;

typedef struct sadcs_server sadcs_server;

int sadcs_hip_server_create(const sadcs_server_params* params, sadcs_server** server);

/* Copies the text (into the server's pinned host slabs while they have room, 1 GiB at most) and
 * queues it; returns the request id (>= 0) or minus a status. */
int64_t sadcs_hip_server_submit(sadcs_server* server, const uint8_t* text, int64_t n);

/* As sadcs_hip_server_submit, with the outputs written straight into the caller's buffers (sa: n
 * int32, bwt: n bytes, either may be null) when the request completes, instead of being held for
 * wait to copy; the buffers must stay valid until wait returns for this id. An output the server
 * makes (want_sa, want_bwt) whose buffer is null here is held as after sadcs_hip_server_submit,
 * and wait copies it into its own sa or bwt argument; wait writes nothing for an output given here.
 * No counterpart in the reference: the server is this library's. */
int64_t sadcs_hip_server_submit_into(sadcs_server* server, const uint8_t* text, int64_t n, int32_t* sa, uint8_t* bwt);

/* Blocks until the request is done and copies its outputs (sa: n int32, bwt: n bytes, either may be
 * null). latency: submission to completion; queued: submission to admission; batch_size: texts in its batch. */
int sadcs_hip_server_wait(sadcs_server* server, int64_t id, int32_t* sa, uint8_t* bwt, int32_t* primary,
                          double* latency_seconds, double* queued_seconds, int32_t* batch_size);

/* counters[0..4]: completed requests, batches, peak pool bytes in use, budget, most batches at once. */
void sadcs_hip_server_counters(sadcs_server* server, int64_t* counters);

/* Where the server's batches spent their time, summed over the batches so far (batches run on
 * `workers` streams at once, so the sums can exceed the wall time). No counterpart in the reference. */
typedef struct sadcs_server_stats {
    int64_t batches;
    int64_t texts;
    int64_t bytes;               /* input bytes */
    double upload_seconds;       /* texts to the device */
    double device_seconds;       /* construction on the device */
    double download_seconds;     /* outputs to the worker's host buffer after the construction */
    double copy_out_seconds;     /* outputs from that buffer into the requests' or callers' arrays */
    double batch_seconds;        /* the whole of each batch on its worker */
} sadcs_server_stats;

int sadcs_hip_server_stats(sadcs_server* server, sadcs_server_stats* stats);

/* Finishes the queued requests, then stops. */
void sadcs_hip_server_destroy(sadcs_server* server);

#ifdef __cplusplus
}
#endif

#endif /* SADCS_HIP_H */
