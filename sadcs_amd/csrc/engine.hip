/*
 * engine.hip - The construction on one gfx950 device, behind the C ABI of include/sadcs_hip.h.
 *
 * One translation unit: the phases (rank.h, classes.h, merge.h) share the __constant__ cover
 * tables defined here. The host side plans memory by running the allocation sequence of a build
 * against a dry-run arena (so the plan is exactly what the build allocates, fragmentation
 * included), reserves that much once per context, and sequences the phases on one stream.
 *
 * Reference counterparts: src/cuda/gpusa.cu (host driver: planMemory 171-201, checkMemory 203-218,
 * uploadAndRank 226-267, GpuContext 321-396, buildOnDevice 406-466, gpuBwt 544-605,
 * gpuRankSample 632-676).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/sadcs_hip.h"
#include "check.h"
#include "classes.h"
#include "cover.h"
#include "dev_common.h"
#include "merge.h"
#include "prim.h"
#include "rank.h"

namespace sadcs {

// ==== device cover tables =======================================================================

__constant__ CoverTables<32> gCover32;
__constant__ CoverTables<128> gCover128;

template <>
__device__ __forceinline__ const CoverTables<32>& deviceCover<32>() {
    return gCover32;
}

template <>
__device__ __forceinline__ const CoverTables<128>& deviceCover<128>() {
    return gCover128;
}

// Cover 32 answers residue questions by bit arithmetic on the cover mask (no table load, which
// would add a dependent memory round trip in front of the gathers that use the answer).
template <int V>
__device__ __forceinline__ uint32_t coverBelow(uint32_t r) {
    if constexpr (V == 32) {
        return __builtin_popcount(coverMask<32>() & ((1u << r) - 1));
    } else {
        return deviceCover<V>().below[r];
    }
}

template <int V>
__device__ __forceinline__ uint32_t packPosition(uint32_t p) {
    return (p >> DifferenceCover<V>::kLog2) * CoverSet<V>::kSize + coverBelow<V>(p & (V - 1));
}

// Position of compact index t. The residue comes from a chain of selects over the cover's
// elements (compile-time constants), not a table in memory: a load here would stall each store.
template <int V>
__device__ __forceinline__ uint32_t unpackIndex(uint32_t t) {
    const uint32_t e = t % CoverSet<V>::kSize;
    uint32_t residue = 0;
#pragma unroll
    for (int k = 1; k < CoverSet<V>::kSize; k++) {
        residue = e == static_cast<uint32_t>(k) ? static_cast<uint32_t>(CoverSet<V>::kElements[k]) : residue;
    }
    return (t / CoverSet<V>::kSize) * V + residue;
}

template <int V>
__device__ __forceinline__ uint32_t classOf(uint32_t p) {
    if constexpr (V == 32) {
        return __builtin_ctz(windowOf32(coverMask<32>(), p));
    } else {
        return deviceCover<V>().distance[p & (V - 1)];
    }
}

// ==== small kernels ==============================================================================

__global__ void fillWordsKernel(uint32_t* __restrict__ out, uint64_t count, uint32_t value) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
        out[i] = value;
    }
}

__global__ void fillBytesKernel(uint8_t* __restrict__ out, uint64_t count, uint8_t value) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
        out[i] = value;
    }
}

// The bytes that occur in the texts, as a 256-bit mask (bit b mod 32 of word b / 32): the dense
// codes need only which bytes occur, not how often. Each thread reads 16-byte chunks (the text
// buffer is 16-byte aligned and padded); gap bytes between the texts of a batch are masked off (a
// chunk never spans a text's start, which is 256-aligned). Every thread ORs each byte's bit into
// its own 8 words of LDS, one LDS atomic per byte with no dependent read; 9 words apart, the
// lanes' words fall in distinct banks. (Selecting the word in registers cost 8 compares and
// selects per byte: the kernel was VALU-bound at a third of the text's read rate.)
template <class Texts>
__global__ __launch_bounds__(256) void bytePresenceKernel(const uint4* __restrict__ chunks, uint32_t span, Texts texts, uint32_t* __restrict__ present) {
    constexpr uint32_t kStride = 9;
    __shared__ uint32_t bits[256 * kStride];
    uint32_t* mine = bits + threadIdx.x * kStride;
    for (uint32_t x = 0; x < 8; x++) {
        mine[x] = 0;
    }
    const uint32_t count = static_cast<uint32_t>((static_cast<uint64_t>(span) + 15) / 16);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < count; c += gridDim.x * blockDim.x) {
        const uint32_t p0 = c * 16;
        const uint32_t end = min(texts.endOf(p0), span);
        const uint32_t valid = p0 < end ? min(16u, end - p0) : 0u;
        const uint4 q = chunks[c];
        const uint32_t words[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            const uint32_t b = (words[k >> 2] >> ((k & 3) * 8)) & 255u;
            if (k < valid) {
                atomicOr(&mine[b >> 5], 1u << (b & 31));
            }
        }
    }
    __syncthreads();
    // Thread t ORs word t % 8 of every 32nd thread from t / 8, then the wave's lanes with the same
    // word (lane bits 3-5) combine by shuffles; lanes 0-7 publish the wave's words.
    const uint32_t x = threadIdx.x & 7;
    uint32_t mask = 0;
    for (uint32_t l = threadIdx.x >> 3; l < 256; l += 32) {
        mask |= bits[l * kStride + x];
    }
    for (int offset = 8; offset < kWave; offset *= 2) {
        mask |= __shfl_xor(mask, offset);
    }
    if (threadIdx.x % kWave < 8 && mask) {
        atomicOr(&present[x], mask);
    }
}

// Rows of a merge level that is a single tile: zeros, then the lengths and their sum.
__global__ void singleTileRowsKernel(ListSet lists, uint32_t* __restrict__ rows) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int lc = lists.count;
        uint32_t total = 0;
        for (int l = 0; l < lc; l++) {
            rows[l] = 0;
            rows[lc + 1 + l] = lists.length[l];
            total += lists.length[l];
        }
        rows[lc] = 0;
        rows[2 * lc + 1] = total;
    }
}

// BWT byte 0 of every text: T[end - 1], the row of the sentinel.
__global__ void firstBwtBytesKernel(const uint8_t* __restrict__ bytes, const Segment* __restrict__ segments, uint32_t count, uint8_t* __restrict__ bwt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < count) {
        bwt[segments[s].rowBase] = bytes[segments[s].end - 1];
    }
}

// The same from the lines (T[x] is byte x mod 32 + 1 of line x / 32), for a BWT that overwrites
// the text.
__global__ void firstBwtBytesFromLinesKernel(const uint4* __restrict__ lines, const Segment* __restrict__ segments, uint32_t count, uint8_t* __restrict__ bwt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < count) {
        const uint64_t x = segments[s].end - 1;
        bwt[segments[s].rowBase] = reinterpret_cast<const uint8_t*>(lines)[(x / kLinePositions) * kLineBytes + x % kLinePositions + 1];
    }
}

// ==== errors =====================================================================================

thread_local char tLastError[512] = "";

void setError(const char* format, ...) __attribute__((format(printf, 1, 2)));

void setError(const char* format, ...) {
    va_list args;
    va_start(args, format);
    vsnprintf(tLastError, sizeof tLastError, format, args);
    va_end(args);
}

int statusOf(hipError_t e) {
    return e == hipErrorOutOfMemory ? SADCS_NOT_ENOUGH_MEMORY : SADCS_FAILED;
}

#define SADCS_HIP(call)                                                                   \
    do {                                                                                  \
        const hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                           \
            setError("%s: %s", #call, hipGetErrorString(e_));                             \
            return statusOf(e_);                                                          \
        }                                                                                 \
    } while (0)

#define SADCS_TRY(call)                 \
    do {                                \
        const int s_ = (call);          \
        if (s_ != SADCS_OK) {           \
            return s_;                  \
        }                               \
    } while (0)

double secondsSince(std::chrono::steady_clock::time_point start) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
}

int bitWidth(uint64_t x) {
    int b = 0;
    while (x) {
        b++;
        x >>= 1;
    }
    return b;
}

// ==== arena ======================================================================================

// First-fit allocator over one device block. In dry-run mode it has no memory and only records
// the high-water mark of the same allocation sequence, which becomes the plan.
class Arena {
 public:
    static constexpr size_t kAlign = 256;

    void reset(uint8_t* base, size_t capacity, bool dry) {
        m_base = base;
        m_capacity = capacity;
        m_dry = dry;
        m_free.assign(1, Block{0, capacity});
        m_used.clear();
        m_highWater = 0;
        m_live = 0;
    }

    bool dry() const {
        return m_dry;
    }

    size_t highWater() const {
        return m_highWater;
    }

    void resetHighWater() {
        size_t top = 0;
        for (const Block& b : m_used) {
            top = std::max(top, b.offset + b.size);
        }
        m_highWater = top;
    }

    template <class T>
    int allocate(T*& pointer, size_t count) {
        if (tryAllocate(pointer, count)) {
            return SADCS_OK;
        }
        const size_t bytes = std::max<size_t>(kAlign, (count * sizeof(T) + kAlign - 1) / kAlign * kAlign);
        setError("device arena exhausted: %zu bytes requested, capacity %zu", bytes, m_capacity);
        return SADCS_NOT_ENOUGH_MEMORY;
    }

    // allocate() without an error when there is no room: for optional buffers.
    template <class T>
    bool tryAllocate(T*& pointer, size_t count) {
        const size_t bytes = std::max<size_t>(kAlign, (count * sizeof(T) + kAlign - 1) / kAlign * kAlign);
        for (size_t i = 0; i < m_free.size(); i++) {
            if (m_free[i].size >= bytes) {
                const size_t offset = m_free[i].offset;
                m_free[i].offset += bytes;
                m_free[i].size -= bytes;
                if (m_free[i].size == 0) {
                    m_free.erase(m_free.begin() + static_cast<std::ptrdiff_t>(i));
                }
                m_used.push_back(Block{offset, bytes});
                m_highWater = std::max(m_highWater, offset + bytes);
                m_live += bytes;
                pointer = reinterpret_cast<T*>((m_dry ? reinterpret_cast<uint8_t*>(kAlign) : m_base) + offset);
                return true;
            }
        }
        return false;
    }

    template <class T>
    void release(T*& pointer) {
        if (!pointer) {
            return;
        }
        const size_t offset = static_cast<size_t>(reinterpret_cast<uint8_t*>(pointer) -
                                                  (m_dry ? reinterpret_cast<uint8_t*>(kAlign) : m_base));
        for (size_t i = 0; i < m_used.size(); i++) {
            if (m_used[i].offset == offset) {
                insertFree(m_used[i]);
                m_live -= m_used[i].size;
                m_used.erase(m_used.begin() + static_cast<std::ptrdiff_t>(i));
                break;
            }
        }
        pointer = nullptr;
    }

 private:
    struct Block {
        size_t offset;
        size_t size;
    };

    void insertFree(Block b) {
        size_t i = 0;
        while (i < m_free.size() && m_free[i].offset < b.offset) {
            i++;
        }
        m_free.insert(m_free.begin() + static_cast<std::ptrdiff_t>(i), b);
        if (i + 1 < m_free.size() && m_free[i].offset + m_free[i].size == m_free[i + 1].offset) {
            m_free[i].size += m_free[i + 1].size;
            m_free.erase(m_free.begin() + static_cast<std::ptrdiff_t>(i + 1));
        }
        if (i > 0 && m_free[i - 1].offset + m_free[i - 1].size == m_free[i].offset) {
            m_free[i - 1].size += m_free[i].size;
            m_free.erase(m_free.begin() + static_cast<std::ptrdiff_t>(i));
        }
    }

    uint8_t* m_base = nullptr;
    size_t m_capacity = 0;
    bool m_dry = true;
    std::vector<Block> m_free;
    std::vector<Block> m_used;
    size_t m_highWater = 0;
    size_t m_live = 0;
};

// Caller-owned device memory the pipeline may use as scratch until the merge writes it: the SA
// buffer of a device call, and the BWT buffer of a fused one. A stack: the top moves down when its
// last entry is released (entries may be released in any order). Allocations that do not fit go
// to the arena instead. In a dry run the base is a placeholder that is never dereferenced.
// The kernels assume the alignment of arena allocations, and a caller's buffer may start anywhere:
// the base is rounded up to Arena::kAlign and the usable size is the same whatever the offset,
// (size - (kAlign - 1)) rounded down to kAlign, so a dry run plans exactly what a build borrows.
class ScratchStack {
 public:
    void reset(void* base, size_t size) {
        const uintptr_t b = reinterpret_cast<uintptr_t>(base);
        const uintptr_t aligned = (b + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign;
        m_size = base && size >= Arena::kAlign ? (size - (Arena::kAlign - 1)) / Arena::kAlign * Arena::kAlign : 0;
        m_base = m_size ? reinterpret_cast<uint8_t*>(aligned) : nullptr;
        m_top = 0;
        m_entries.clear();
    }

    template <class T>
    bool take(T*& pointer, uint64_t count) {
        const size_t bytes = (std::max<size_t>(count, 1) * sizeof(T) + Arena::kAlign - 1) / Arena::kAlign * Arena::kAlign;
        if (!m_base || m_top + bytes > m_size) {
            return false;
        }
        pointer = reinterpret_cast<T*>(m_base + m_top);
        m_entries.push_back(Entry{m_top, bytes, false});
        m_top += bytes;
        return true;
    }

    // False if the pointer is not from this stack.
    template <class T>
    bool give(T*& pointer) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(pointer);
        if (!pointer || !m_base || p < m_base || p >= m_base + m_size) {
            return false;
        }
        const size_t offset = static_cast<size_t>(p - m_base);
        for (Entry& entry : m_entries) {
            if (entry.offset == offset) {
                entry.freed = true;
            }
        }
        while (!m_entries.empty() && m_entries.back().freed) {
            m_top = m_entries.back().offset;
            m_entries.pop_back();
        }
        pointer = nullptr;
        return true;
    }

 private:
    struct Entry {
        size_t offset, bytes;
        bool freed;
    };
    uint8_t* m_base = nullptr;
    size_t m_size = 0, m_top = 0;
    std::vector<Entry> m_entries;
};

// ==== jobs =======================================================================================

struct Params {
    int cover = 32;
    uint32_t g = defaultSplitterSpacing(32);  // splitter spacing
    uint32_t k = 0;                           // sorted splitters per tile; 0: the most that keep the CU's tiles in its LDS
    int threads = defaultTileThreads(32);     // threads per merge tile (512 or 256)

    // Zero fields take the defaults of the cover.
    void apply(const sadcs_params* p) {
        if (p && p->cover) {
            cover = p->cover;
        }
        threads = p && p->tile_threads ? static_cast<int>(p->tile_threads) : defaultTileThreads(cover);
        g = p && p->splitter_spacing ? p->splitter_spacing : defaultSplitterSpacing(cover);
        if (p && p->tile_splitters) {
            k = p->tile_splitters;
        }
    }
};

// One construction: texts on the device (segments of one position space) and what to produce.
struct Job {
    const uint8_t* text = nullptr;          // device, span + 64 bytes, zero padded
    uint32_t span = 0;                      // positions, gaps between texts included
    std::vector<Segment> segments;          // host copy
    const Segment* deviceSegments = nullptr;
    const uint32_t* blockSegment = nullptr; // batch only
    bool batch = false;
    bool rankOnly = false;
    uint32_t* sa = nullptr;                 // device outputs, null = not wanted / allocate
    uint8_t* bwt = nullptr;
    bool wantSA = false;
    bool wantBWT = false;
    uint32_t* firstRows = nullptr;          // device, one per text (BWT)
    bool textScratch = false;               // the text is this call's upload: free to overwrite late in the merge
    uint32_t maxLength = 0;
    uint64_t totalLength = 0;               // positions in texts (output rows)
    int32_t* hostSA = nullptr;              // host calls: where the outputs go; the merge
    uint8_t* hostBWT = nullptr;             //   downloads them as it finishes slices of rows
    bool downloaded = false;                //   and sets this when it has
};

// The outputs of a top-level merge in slices, for downloadSlices: rows[0] = 0 < ... <
// rows[slices] = totalRows (host copies of the slices' first rows).
struct OutputSlices {
    uint32_t slices = 0;
    const uint32_t* rows = nullptr;
    uint64_t totalRows = 0;
    uint8_t* hostSA = nullptr;
    const uint8_t* deviceSA = nullptr;
    uint8_t* hostBWT = nullptr;
    const uint8_t* deviceBWT = nullptr;
};

int downloadSlices(sadcs_ctx& ctx, const OutputSlices& o);

struct PhaseTimes {
    double ranking = 0, classes = 0, merge = 0, bwt = 0;
    double download = 0;  // host seconds of a download that overlapped the merge, past its end
    double mergeKernel = 0;
    uint64_t mergeElements = 0;
    int characterPasses = 0, doublingRounds = 0;
    size_t rankingBytes = 0, classBytes = 0, mergeBytes = 0;
};

}  // namespace sadcs

// ==== the context ================================================================================

struct sadcs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t capacity = 0;
    sadcs::Params params;
    uint8_t* arenaBase = nullptr;
    size_t arenaBytes = 0;
    sadcs::Arena arena;
    uint32_t* hostWords = nullptr;     // pinned, for readbacks
    uint8_t* staging[2] = {};          // pinned staging chunks
    hipEvent_t stagingDone[2] = {};
    hipEvent_t kernelTimer[2] = {};    // around the top-level merge kernel
    int sms = 0;
    // Downloads that overlap the top-level merge: its launches over consecutive row slices each
    // record sliceDone; copyStream copies finished slices to the host while later ones merge.
    static constexpr int kSlices = 16;
    hipStream_t copyStream = nullptr;  // null: no overlap (download after the merge)
    hipEvent_t sliceDone[kSlices] = {};
    hipEvent_t mergeBegin = nullptr;
};

namespace sadcs {

constexpr size_t kStagingChunk = size_t(32) << 20;
// Pinned readback words: [0, 264) the text's byte presence (at the start of ranking), [0, 16)
// counters, [16, 1016) the first rows of a batch (after the build), [kSliceRowWords, +kSlices + 1)
// the row bounds of the top-level merge's slices.
constexpr size_t kHostWords = 2048;
constexpr size_t kSliceRowWords = 1024;
// The top-level merge of a host call is launched in slices of about this many output rows (at
// most sadcs_ctx::kSlices), so the download of finished rows overlaps the rest of the merge.
constexpr uint64_t kSliceRows = uint64_t(1) << 24;
constexpr int64_t kMemoryReserve = int64_t(256) << 20;
constexpr size_t kTextPadding = 64;

bool gTablesReady[2] = {false, false};

int ensureTables() {
    if (!gTablesReady[0]) {
        static CoverTables<32> t32;
        t32.fill();
        SADCS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gCover32), &t32, sizeof t32));
        gTablesReady[0] = true;
    }
    if (!gTablesReady[1]) {
        static CoverTables<128> t128;
        t128.fill();
        SADCS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(gCover128), &t128, sizeof t128));
        gTablesReady[1] = true;
    }
    return SADCS_OK;
}

template <int V>
uint64_t classSizeOf(uint64_t n, int d) {
    uint64_t size = 0;
    for (int r = 0; r < V; r++) {
        if (DifferenceCover<V>::nextSampleDistance(r) == d) {
            size += n / V + (static_cast<uint64_t>(r) < n % V ? 1 : 0);
        }
    }
    return size;
}

// A pass after the first sorts within groups only while its groups hold at most this many
// elements on average (u / groups); larger groups make the per-element scans cost more than the
// radix passes.
constexpr uint64_t kGroupRankMeanMax = 8;

// The first ranking pass of a large sample leaves its resolved rank stores to
// scatterResolvedRanksKernel, in launches over consecutive rank ranges of at most this many bytes
// (what the 256 MB memory-side cache holds: at 1 GiB, 4 launches measured 1.2 ms faster than 8,
// and 2 launches, of 470 MB each, 2 ms slower than 4).
constexpr uint32_t kDeferredRankMin = 1u << 22;
constexpr uint64_t kDeferredRankWindow = uint64_t(240) << 20;

template <int V>
uint32_t mergeCapacity(uint32_t g, uint32_t k) {
    return k * g + DifferenceCover<V>::kClasses * (g - 1);
}

// ==== the pipeline ===============================================================================

template <int V, class Texts>
class Pipeline {
    using Cover = DifferenceCover<V>;
    static constexpr int kClasses = Cover::kClasses;
    static constexpr bool kBatch = Texts::kBatch;

 public:
    Pipeline(sadcs_ctx& ctx, Job& job, Texts texts) : m_ctx(ctx), m_arena(ctx.arena), m_job(job), m_texts(texts) {
        m_stream = ctx.stream;
        m_dry = m_arena.dry();
        m_segments = static_cast<uint32_t>(job.segments.size());
        m_samples = 0;
        for (const Segment& s : job.segments) {
            m_samples += static_cast<uint32_t>(Cover::pack(s.end - s.start));
        }
        m_ctxMerge.text = reinterpret_cast<const uint64_t*>(job.text);
        m_ctxMerge.segments = job.deviceSegments;
        m_ctxMerge.n = job.span;
        m_g = ctx.params.g;
        m_threads = ctx.params.threads;
        m_k = ctx.params.k ? ctx.params.k : defaultTileSplitters<V, kBatch>(m_g, m_threads);
        // Caller-owned outputs are idle until the merge writes them. The BWT buffer can hold the
        // ranks only when the merge reads records from the lines (fused SA + BWT).
        m_saScratch.reset(job.sa, job.sa ? static_cast<size_t>(job.totalLength) * sizeof(uint32_t) : 0);
        m_bwtScratch.reset(job.sa && job.bwt && !job.rankOnly ? job.bwt : nullptr, job.bwt ? static_cast<size_t>(job.totalLength) : 0);
    }

    ~Pipeline() {
        releaseAll();
    }

    PhaseTimes times;

    int run() {
        std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();
        SADCS_TRY(rankSample());
        SADCS_TRY(sync());
        times.ranking = secondsSince(start);
        times.rankingBytes = m_arena.highWater();
        if (m_job.rankOnly) {
            return SADCS_OK;  // the caller reads ranks() before the pipeline goes
        }
        start = std::chrono::steady_clock::now();
        SADCS_TRY(sortClasses());
        SADCS_TRY(sync());
        times.classes = secondsSince(start);
        times.classBytes = m_arena.highWater();
        start = std::chrono::steady_clock::now();
        SADCS_TRY(merge());
        SADCS_TRY(sync());
        times.merge = secondsSince(start);
        times.mergeBytes = m_arena.highWater();
        if (m_job.downloaded) {
            // The merge phase ends with its last slice; the host time past that was the download's.
            float ms = 0;
            SADCS_HIP(hipEventElapsedTime(&ms, m_ctx.mergeBegin, m_ctx.sliceDone[m_slices - 1]));
            times.download = std::max(0.0, times.merge - ms * 1e-3);
            times.merge = ms * 1e-3;
        }
        return SADCS_OK;
    }

    const uint32_t* ranks() const {
        return m_ranks;
    }

    uint32_t samples() const {
        return m_samples;
    }

 private:
    sadcs_ctx& m_ctx;
    Arena& m_arena;
    Job& m_job;
    Texts m_texts;
    hipStream_t m_stream;
    bool m_dry;
    uint32_t m_segments;
    uint32_t m_samples;
    MergeCtx m_ctxMerge{};
    uint32_t m_g = 0, m_k = 0;  // splitter spacing, sorted splitters per tile
    int m_threads = kMergeThreads;  // threads per merge tile
    ScratchStack m_saScratch;   // the caller's SA buffer before the merge
    ScratchStack m_bwtScratch;  // the caller's BWT buffer before the merge (fused builds)

    template <class T>
    int scratch(T*& pointer, uint64_t count) {
        if (m_saScratch.take(pointer, count)) {
            return SADCS_OK;
        }
        return m_arena.allocate(pointer, count);
    }

    template <class T>
    void unscratch(T*& pointer) {
        if (!m_saScratch.give(pointer) && !m_bwtScratch.give(pointer)) {
            m_arena.release(pointer);
        }
    }

    uint32_t* m_ranks = nullptr;
    uint32_t* m_lists[kClasses] = {};
    uint64_t m_listLength[kClasses] = {};

    int sync() {
        if (!m_dry) {
            SADCS_HIP(hipStreamSynchronize(m_stream));
        }
        return SADCS_OK;
    }

    void releaseAll() {
        for (int d = 0; d < kClasses; d++) {
            m_arena.release(m_lists[d]);
        }
        unscratch(m_ranks);
    }

    static unsigned blocksFor(uint64_t count, unsigned threads = 256) {
        return static_cast<unsigned>(std::max<uint64_t>(1, (count + threads - 1) / threads));
    }

    void fillWords(uint32_t* out, uint64_t count, uint32_t value) {
        if (!m_dry && count) {
            fillWordsKernel<<<std::min<unsigned>(blocksFor(count), 4096), 256, 0, m_stream>>>(out, count, value);
        }
    }

    // ---- phase (a) ------------------------------------------------------------------------------

    // The elements of large groups (flags from groupRankSortKernel, `large` of them): compacted,
    // sorted by radix and placed at their list positions in keys/values[1 - y]. sorted = false
    // (and nothing done) when the arena has no room for them.
    int sortLargeGroups(uint64_t* const* keys, uint32_t* const* values, int y, uint32_t u, uint32_t large, int keyBits,
                        uint8_t* flags, uint32_t* hist, Sum1* sortAggregates, Sum1* scanAggregates, bool& sorted) {
        uint64_t* subKeys[2] = {};
        uint32_t* subValues[2] = {};
        uint32_t* subPos = nullptr;
        bool room = true;
        for (int b = 0; b < 2 && room; b++) {
            room = m_arena.tryAllocate(subKeys[b], large) && m_arena.tryAllocate(subValues[b], large);
        }
        room = room && m_arena.tryAllocate(subPos, large);
        if (room) {
            scanPipeline<Sum1, kSumItems>(m_stream, u, LargeFlagSource{flags},
                                          LargeGroupCompact{flags, keys[y], values[y], subKeys[0], subValues[0], subPos}, scanAggregates);
            SortBuffers<uint64_t> sub;
            for (int b = 0; b < 2; b++) {
                sub.keys[b] = subKeys[b];
                sub.values[b] = subValues[b];
            }
            sub.hist = hist;
            sub.aggregates = sortAggregates;
            sub.digits = flags;  // read by the compaction above, before the sort writes digits
            radixSort<uint64_t>(m_stream, sub, large, 0, keyBits);
            placeLargeGroupsKernel<<<blocksFor(large), 256, 0, m_stream>>>(sub.keys[sub.current], sub.values[sub.current], subPos, large,
                                                                          keys[1 - y], values[1 - y]);
            SADCS_HIP(hipGetLastError());
            sorted = true;
        }
        // Stream order keeps these kernels' uses ahead of any reuse of the memory.
        m_arena.release(subPos);
        for (int b = 1; b >= 0; b--) {
            m_arena.release(subValues[b]);
            m_arena.release(subKeys[b]);
        }
        return SADCS_OK;
    }

    int rankSample() {
        const uint64_t rankWords = static_cast<uint64_t>(Cover::pack(m_job.span)) + CoverSet<V>::kSize + 8;
        if (!(V == kLinePositions && m_bwtScratch.take(m_ranks, rankWords))) {
            SADCS_TRY(m_arena.allocate(m_ranks, rankWords));
        }
        SADCS_TRY(m_arena.allocate(m_lists[0], std::max<uint32_t>(m_samples, 1)));
        m_listLength[0] = m_samples;
        fillWords(m_ranks, rankWords, 0);
        fillWords(m_lists[0], std::max<uint32_t>(m_samples, 1), ~0u);  // unwritten sorted-sample entries

        const uint32_t M = m_samples;
        const uint32_t groupCapacity = M / 2 + m_segments + 1;
        uint64_t* keys[2] = {};
        uint32_t* values[2] = {};
        uint32_t* groupIds[2] = {};
        uint32_t* hist = nullptr;
        Sum1* sortAggregates = nullptr;
        MarkVec* markAggregates = nullptr;
        uint32_t* counts = nullptr;
        uint32_t* byteCounts = nullptr;
        uint8_t* digits = nullptr;  // the radix passes' next digits, one byte per sample
        uint32_t* textMasks = nullptr;  // a batch's per-text alphabets, for its first pass
        uint8_t* textCodes = nullptr;
        uint8_t* textShape = nullptr;

        int status = SADCS_OK;
        auto cleanup = [&]() {
            m_arena.release(textShape);
            m_arena.release(textCodes);
            m_arena.release(textMasks);
            unscratch(digits);
            m_arena.release(byteCounts);
            m_arena.release(counts);
            m_arena.release(markAggregates);
            m_arena.release(sortAggregates);
            m_arena.release(hist);
            for (int i = 1; i >= 0; i--) {
                m_arena.release(groupIds[i]);
                m_arena.release(values[i]);
                unscratch(keys[i]);
            }
        };

        for (int i = 0; i < 2 && status == SADCS_OK; i++) {
            status = scratch(keys[i], M);
            if (status == SADCS_OK) status = m_arena.allocate(values[i], M);
            if (status == SADCS_OK) status = m_arena.allocate(groupIds[i], groupCapacity);
        }
        if (status == SADCS_OK) status = m_arena.allocate(hist, radixHistWords(M));
        if (status == SADCS_OK) status = m_arena.allocate(sortAggregates, radixAggregateBytes(M) / sizeof(Sum1) + 1);
        if (status == SADCS_OK) status = m_arena.allocate(markAggregates, scanScratchBytes<MarkVec, kMarkItems>(M) / sizeof(MarkVec) + 1);
        if (status == SADCS_OK) status = m_arena.allocate(counts, 4);
        if (status == SADCS_OK) status = m_arena.allocate(byteCounts, 8);
        if (status == SADCS_OK) status = scratch(digits, std::max<uint32_t>(M, 1));
        if (kBatch && m_segments > 1) {
            if (status == SADCS_OK) status = m_arena.allocate(textMasks, static_cast<uint64_t>(m_segments) * 8);
            if (status == SADCS_OK) status = m_arena.allocate(textCodes, static_cast<uint64_t>(m_segments) * 256);
            if (status == SADCS_OK) status = m_arena.allocate(textShape, static_cast<uint64_t>(m_segments) * 2);
        }
        if (status != SADCS_OK || m_dry || M == 0) {
            cleanup();
            return status;
        }

        // The alphabet: dense codes 1..sigma for the bytes present.
        fillWords(byteCounts, 8, 0);
        bytePresenceKernel<Texts><<<std::min<unsigned>(blocksFor((m_job.span + 15) / 16), 2048), 256, 0, m_stream>>>(
            reinterpret_cast<const uint4*>(m_job.text), m_job.span, m_texts, byteCounts);
        SADCS_HIP(hipMemcpyAsync(m_ctx.hostWords + 256, byteCounts, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, m_stream));
        SADCS_HIP(hipStreamSynchronize(m_stream));
        for (int b = 0; b < 256; b++) {
            m_ctx.hostWords[b] = (m_ctx.hostWords[256 + b / 32] >> (b % 32)) & 1u;
        }
        CodeTable codes{};
        int sigma = 0;
        for (int b = 0; b < 256; b++) {
            if (m_ctx.hostWords[b]) {
                codes.code[b] = static_cast<uint8_t>(++sigma);  // 256 wraps to 0 only if sigma = 256: see below
            }
        }
        const int charBits = std::max(1, bitWidth(static_cast<uint64_t>(sigma)));
        if (sigma == 256) {
            // 9-bit codes do not fit a byte table; use the byte plus one computed in the key kernel.
            for (int b = 0; b < 256; b++) {
                codes.code[b] = static_cast<uint8_t>(b);
            }
        }

        RankParams rp{};
        rp.text = reinterpret_cast<const uint64_t*>(m_job.text);
        rp.ranks = m_ranks;
        rp.sortedSample = m_lists[0];
        rp.samples = M;
        rp.segments = m_segments;
        rp.segmentTable = m_job.deviceSegments;
        // A batch whose texts have different alphabets: the first pass packs each text's own
        // dense codes, as many characters as fit (the later passes use the common codes).
        const int firstBits = std::min(56, 64 - (m_segments > 1 ? bitWidth(m_segments - 1) : 0));
        if constexpr (kBatch) {
            if (textCodes && sigma < 256) {
                fillWords(textMasks, m_segments * 8, 0);
                textPresenceKernel<<<static_cast<unsigned>((static_cast<uint64_t>(m_job.span) + 65535) / 65536), 256, 0, m_stream>>>(
                    reinterpret_cast<const uint4*>(m_job.text), m_job.span, m_texts, textMasks);
                textCodesKernel<<<blocksFor(m_segments), 256, 0, m_stream>>>(textMasks, m_segments, firstBits, textCodes, textShape);
                SADCS_HIP(hipGetLastError());
                rp.textCodes = textCodes;
                rp.textShape = textShape;
            }
        }

        initGroupIdsKernel<<<blocksFor(m_segments), 256, 0, m_stream>>>(m_job.deviceSegments, m_segments, groupIds[0]);

        SortBuffers<uint64_t> sb;
        sb.keys[0] = keys[0];
        sb.keys[1] = keys[1];
        sb.values[0] = values[0];
        sb.values[1] = values[1];
        sb.hist = hist;
        sb.aggregates = sortAggregates;
        sb.digits = digits;

        uint32_t u = M;
        uint32_t groups = m_segments;
        uint64_t depth = 0;
        bool first = true;
        int listBuffer = 1;  // where the list's t and ordinals are (values[x], keys[x] as words)
        int gcur = 0;
        const int rankBits = bitWidth(M);

        while (u > 0) {
            if (!first && u <= kTailMax) {
                // The remaining passes in one workgroup, with one wait for all of them.
                TailParams tp{};
                tp.rp = rp;
                tp.codes = codes;
                tp.charBits = sigma == 256 ? 9 : charBits;
                tp.rankBits = rankBits;
                tp.maxLength = m_job.maxLength;
                tp.depth = depth;
                tp.listT = values[listBuffer];
                tp.listOrdinal = reinterpret_cast<const uint32_t*>(keys[listBuffer]);
                tp.groupIds = groupIds[gcur];
                tp.u = u;
                tp.groups = groups;
                tp.status = counts;
                const size_t tailShared = rankTailSharedBytes();
                SADCS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(rankTailKernel<V, Texts>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                              static_cast<int>(tailShared)));
                rankTailKernel<V, Texts><<<1, kTailThreads, tailShared, m_stream>>>(tp, m_texts);
                SADCS_HIP(hipGetLastError());
                SADCS_HIP(hipMemcpyAsync(m_ctx.hostWords, counts, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, m_stream));
                SADCS_HIP(hipStreamSynchronize(m_stream));
                times.characterPasses += static_cast<int>(m_ctx.hostWords[0]);
                times.doublingRounds += static_cast<int>(m_ctx.hostWords[1]);
                if (m_ctx.hostWords[2] != 0) {
                    setError("sample ranking: groups left after doubling past the longest text (internal error)");
                    cleanup();
                    return SADCS_FAILED;
                }
                break;
            }
            const int ordinalBits = groups > 1 ? bitWidth(groups - 1) : 0;
            PassSpec spec{};
            if (depth < static_cast<uint64_t>(V)) {
                spec.doubling = false;
                spec.offset = static_cast<uint32_t>(depth);
                spec.charBits = sigma == 256 ? 9 : charBits;
                // The first pass sorts every sample: 56 key bits (7 digit passes) instead of up to 63
                // (8) resolves nearly as many, and the later passes sort only the few left.
                spec.chars = std::max(1, (first ? std::min(56, 64 - ordinalBits) : 64 - ordinalBits) / spec.charBits);
                spec.subkeyBits = spec.chars * spec.charBits;
                if (first && rp.textCodes) {
                    // Per-text codes fill the first pass's bits; the depth still advances by the
                    // common codes' characters, which every text's key covers.
                    spec.subkeyBits = firstBits;
                }
            } else {
                const uint64_t h = depth / V * V;
                if (h >= m_job.maxLength) {
                    setError("sample ranking: groups left at depth %llu (internal error)", static_cast<unsigned long long>(depth));
                    cleanup();
                    return SADCS_FAILED;
                }
                spec.doubling = true;
                spec.offset = static_cast<uint32_t>(h);
                spec.subkeyBits = rankBits;
            }

            const int y = 1 - listBuffer;
            // The first pass of a single text sorts the identity as its values (t = index) without
            // storing them, and its key kernel writes the first digit for the radix histograms.
            const bool identity = first && !kBatch && u > 1;
            uint32_t* valuesOut = identity ? nullptr : values[y];
            uint8_t* digitsOut = first ? digits : nullptr;
            if (sigma == 256 && !spec.doubling) {
                makePassKeysKernel<V, Texts, true><<<blocksFor(u), 256, 0, m_stream>>>(
                    rp, m_texts, spec, codes, first, values[listBuffer], reinterpret_cast<const uint32_t*>(keys[listBuffer]), u, keys[y], valuesOut, digitsOut);
            } else {
                makePassKeysKernel<V, Texts, false><<<blocksFor(u), 256, 0, m_stream>>>(
                    rp, m_texts, spec, codes, first, values[listBuffer], reinterpret_cast<const uint32_t*>(keys[listBuffer]), u, keys[y], valuesOut, digitsOut);
            }
            // After the first pass the unresolved groups are mostly pairs: sort within groups. The
            // few elements of groups too large for that (long runs) are sorted apart; when they are
            // many, or the groups are large on average (the scans would cost more than the radix
            // passes), the pass is sorted by radix.
            bool sorted = false;
            if (!first && ordinalBits > 0 && spec.subkeyBits < 64 && u <= kGroupRankMeanMax * static_cast<uint64_t>(groups)) {
                SADCS_HIP(hipMemsetAsync(counts + 2, 0, sizeof(uint32_t), m_stream));
                groupRankSortKernel<<<blocksFor(u), 256, 0, m_stream>>>(keys[y], values[y], u, spec.subkeyBits, keys[1 - y], values[1 - y], digits, counts + 2);
                SADCS_HIP(hipMemcpyAsync(m_ctx.hostWords + 2, counts + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, m_stream));
                SADCS_HIP(hipStreamSynchronize(m_stream));
                const uint32_t large = m_ctx.hostWords[2];
                if (large == 0) {
                    sorted = true;
                } else if (large <= u / 4) {
                    SADCS_TRY(sortLargeGroups(keys, values, y, u, large, ordinalBits + spec.subkeyBits, digits, hist, sortAggregates,
                                              reinterpret_cast<Sum1*>(markAggregates), sorted));
                }
                if (sorted) {
                    sb.current = 1 - y;
                }
            }
            if (!sorted) {
                sb.current = y;
                sb.identityValues = identity;
                sb.digitsReady = first;
                radixSort<uint64_t>(m_stream, sb, u, 0, ordinalBits + spec.subkeyBits);
                sb.identityValues = false;
                sb.digitsReady = false;
            }
            const int c = sb.current;

            MarkState ms{};
            ms.keys = keys[c];
            ms.values = values[c];
            ms.u = u;
            ms.subkeyBits = spec.subkeyBits;
            ms.groupIds = groupIds[gcur];
            ms.ranks = m_ranks;
            ms.sortedSample = m_lists[0];
            ms.nextT = values[1 - c];
            ms.nextOrdinal = reinterpret_cast<uint32_t*>(keys[1 - c]);
            ms.nextGroupIds = groupIds[1 - gcur];
            ms.counts = counts;
            ms.samples = M;
            ms.deferResolved = first && u >= kDeferredRankMin;
            scanPipeline<MarkVec, kMarkItems>(m_stream, u, MarkProducer{ms}, MarkConsumer<V>{ms}, markAggregates);
            SADCS_HIP(hipGetLastError());
            if (ms.deferResolved) {
                const uint32_t rankCount = static_cast<uint32_t>(Cover::pack(m_job.span));
                const uint64_t rounds = std::max<uint64_t>(1, (uint64_t(rankCount) * sizeof(uint32_t) + kDeferredRankWindow - 1) / kDeferredRankWindow);
                const uint32_t step = static_cast<uint32_t>((rankCount + rounds - 1) / rounds);
                const unsigned grid = std::min<unsigned>(blocksFor(M), 8192);
                for (uint32_t lo = 0; lo < rankCount; lo += step) {
                    scatterResolvedRanksKernel<V><<<grid, 256, 0, m_stream>>>(m_lists[0], M, lo, std::min(rankCount, lo + step), m_ranks);
                }
                SADCS_HIP(hipGetLastError());
            }
            SADCS_HIP(hipMemcpyAsync(m_ctx.hostWords, counts, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, m_stream));
            SADCS_HIP(hipStreamSynchronize(m_stream));

            u = m_ctx.hostWords[0];
            groups = m_ctx.hostWords[1];
            gcur = 1 - gcur;
            listBuffer = 1 - c;
            first = false;
            if (spec.doubling) {
                depth += spec.offset;
                times.doublingRounds++;
            } else {
                depth += static_cast<uint64_t>(spec.chars);
                times.characterPasses++;
            }
        }
        cleanup();
        return SADCS_OK;
    }

    // ---- phase (b) ------------------------------------------------------------------------------

    int sortClasses() {
        uint64_t largest = 0;
        for (int d = 1; d < kClasses; d++) {
            uint64_t size = 0;
            for (const Segment& s : m_job.segments) {
                size += classSizeOf<V>(s.end - s.start, d);
            }
            m_listLength[d] = size;
            largest = std::max(largest, size);
            SADCS_TRY(m_arena.allocate(m_lists[d], std::max<uint64_t>(size, 1)));
        }

        using Key = typename std::conditional<kBatch, uint64_t, uint32_t>::type;
        Key* keys[2] = {};
        Key* tmpKeys = nullptr;
        uint32_t* tmpValues = nullptr;
        uint32_t* hist = nullptr;
        Sum1* aggregates = nullptr;
        uint32_t* walkKeys = nullptr;  // the walks of classes that read the text, evaluated once
        uint32_t* walkValues = nullptr;
        uint8_t* classDigits = nullptr;  // a batch's digit bytes between its passes over the text index
        // A batch sorts each class by (text, character) in one pass when the per-text digit
        // counts (3 words per text and digit) are small beside the class lists; else it adds
        // digit passes over the text index.
        const uint64_t textCells = static_cast<uint64_t>(m_segments) * kRadixBins;
        const bool byText = kBatch && m_segments > 1 && textCells <= largest;
        TextSortScratch<Key> ts;
        const uint64_t walk = largest + m_segments + static_cast<uint64_t>(m_listLength[0]);
        uint64_t textWalk = 0;
        for (int d = 1; d < kClasses; d++) {
            if ((d - 1) % 4 == 0) {
                textWalk = std::max<uint64_t>(textWalk, m_listLength[d - 1] + m_segments);
            }
        }
        int status = SADCS_OK;
        auto cleanup = [&]() {
            unscratch(ts.endDigits);
            unscratch(ts.endValues);
            unscratch(ts.endKeys);
            unscratch(ts.delta);
            unscratch(ts.byDigit);
            unscratch(ts.byText);
            unscratch(classDigits);
            unscratch(walkValues);
            unscratch(walkKeys);
            unscratch(aggregates);
            unscratch(hist);
            unscratch(tmpValues);
            unscratch(tmpKeys);
            unscratch(keys[1]);
            unscratch(keys[0]);
        };
        status = scratch(keys[0], std::max<uint64_t>(largest, 1));
        if (status == SADCS_OK) status = scratch(keys[1], std::max<uint64_t>(largest, 1));
        if (kBatch && status == SADCS_OK) status = scratch(tmpKeys, std::max<uint64_t>(largest, 1));
        if (kBatch && status == SADCS_OK) status = scratch(tmpValues, std::max<uint64_t>(largest, 1));
        // The histograms too go to the caller's buffer when it has room: every class list is live
        // now, the peak of a device-resident build.
        if (status == SADCS_OK) status = scratch(hist, radixHistWords(walk));
        const size_t aggregateBytes = std::max(radixAggregateBytes(walk), byText ? scanScratchBytes<Sum1, kSumItems>(textCells) : size_t(0));
        if (status == SADCS_OK) status = scratch(aggregates, aggregateBytes / sizeof(Sum1) + 1);
        if (status == SADCS_OK) status = scratch(walkKeys, std::max<uint64_t>(textWalk, 1));
        if (status == SADCS_OK) status = scratch(walkValues, std::max<uint64_t>(textWalk, 1));
        if (kBatch && !byText && status == SADCS_OK) status = scratch(classDigits, std::max<uint64_t>(largest, 1));
        if (byText) {
            if (status == SADCS_OK) status = scratch(ts.byText, textCells);
            if (status == SADCS_OK) status = scratch(ts.byDigit, textCells);
            if (status == SADCS_OK) status = scratch(ts.delta, textCells);
            if (status == SADCS_OK) status = scratch(ts.endKeys, m_segments);
            if (status == SADCS_OK) status = scratch(ts.endValues, m_segments);
            if (status == SADCS_OK) status = scratch(ts.endDigits, m_segments);
        }
        if (status != SADCS_OK || m_dry) {
            cleanup();
            return status;
        }

        const int segmentBits = m_segments > 1 ? bitWidth(m_segments - 1) : 0;
        int prev = 0;  // keys[prev]: the previous class's keys
        for (int d = 1; d < kClasses; d++) {
            if (m_listLength[d] == 0) {
                continue;
            }
            const bool fromText = (d - 1) % 4 == 0;
            ClassWalkSource<V, Texts, Key> source{};
            source.bytes = m_job.text;
            source.previous = m_lists[d - 1];
            source.previousKeys = fromText ? nullptr : keys[prev];
            source.previousCount = static_cast<uint32_t>(m_listLength[d - 1]);
            source.segmentTable = m_job.deviceSegments;
            source.segments = m_segments;
            source.n = m_job.span;
            source.texts = m_texts;
            source.d = d;
            const uint32_t count = static_cast<uint32_t>(m_listLength[d - 1]) + m_segments;
            const int next = 1 - prev;

            SortBuffers<Key> sb;
            sb.hist = hist;
            sb.aggregates = aggregates;
            const int extraPasses = kBatch && !byText ? (segmentBits + kRadixBits - 1) / kRadixBits : 0;
            const int passes = 1 + extraPasses;
            // Buffer 1 is the destination (next keys, class d); buffer 0 the temporaries.
            sb.keys[0] = kBatch ? tmpKeys : keys[prev];
            sb.values[0] = tmpValues;
            sb.keys[1] = keys[next];
            sb.values[1] = m_lists[d];
            sb.current = passes % 2 == 1 ? 0 : 1;
            // A batch's passes over the text index (bits 32 and up) count digit bytes that the pass
            // before wrote, instead of reading the 8-byte keys.
            const uint32_t kept = static_cast<uint32_t>(m_listLength[d]);
            auto segmentWidth = [&](int e) { return std::min(kRadixBits, segmentBits - e * kRadixBits); };
            sb.digits = classDigits;
            const int firstNext = extraPasses > 0 && classDigits ? 32 : -1;
            if (byText) {
                if (fromText) {
                    materializeWalkKernel<decltype(source), Key><<<blocksFor(count), 256, 0, m_stream>>>(source, count, walkKeys, walkValues);
                    sortClassByText<Key>(m_stream, sb, MaterializedWalk<Key, Texts>{walkKeys, walkValues, m_texts}, count, kept, m_segments, 24, ts);
                } else {
                    sortClassByText<Key>(m_stream, sb, source, count, kept, m_segments, 24, ts);
                }
            } else if (fromText) {
                materializeWalkKernel<decltype(source), Key><<<blocksFor(count), 256, 0, m_stream>>>(source, count, walkKeys, walkValues);
                radixPassDigits<Key>(m_stream, sb, MaterializedWalk<Key, Texts>{walkKeys, walkValues, m_texts}, count, 24, 8, kept,
                                     false, firstNext, firstNext >= 0 ? segmentWidth(0) : 0);
            } else {
                radixPassDigits<Key>(m_stream, sb, source, count, 24, 8, kept, false, firstNext, firstNext >= 0 ? segmentWidth(0) : 0);
            }
            for (int e = 0; e < extraPasses; e++) {
                const int shift = 32 + e * kRadixBits;
                const int next = e + 1 < extraPasses && classDigits ? shift + kRadixBits : -1;
                radixPassDigits<Key>(m_stream, sb, ArraySource<Key>{sb.keys[sb.current], sb.values[sb.current]}, kept, shift, segmentWidth(e),
                                     0xFFFFFFFFu, classDigits != nullptr, next, next >= 0 ? segmentWidth(e + 1) : 0);
            }
            SADCS_HIP(hipGetLastError());
            prev = next;
        }
        cleanup();
        return SADCS_OK;
    }

    // ---- phase (c) ------------------------------------------------------------------------------

    ListSet topLists() const {
        ListSet l{};
        l.count = kClasses;
        uint32_t base = 0;
        for (int d = 0; d < kClasses; d++) {
            l.ptr[d] = m_lists[d];
            l.length[d] = static_cast<uint32_t>(m_listLength[d]);
            l.indexBase[d] = base;
            base += l.length[d];
        }
        l.indexBase[kClasses] = base;
        return l;
    }

    template <int kFlags>
    int launchTiles(const ListSet& lists, RowTable rows, uint32_t tiles, const TileOut& out, uint32_t* errors, uint32_t slices = 1) {
        const uint32_t capacity = mergeCapacity<V>(m_g, m_k);
        const size_t shared = mergeSharedBytes<V, kBatch>(capacity);
        if (m_dry) {
            return SADCS_OK;
        }
        constexpr bool kTop = (kFlags & kOutIndex) == 0;
        if (kTop && m_ctx.kernelTimer[0]) {
            SADCS_HIP(hipEventRecord(m_ctx.kernelTimer[0], m_stream));
        }
        auto kernel = m_threads == 256 ? mergeTilesKernel<V, Texts, kFlags, 256> : mergeTilesKernel<V, Texts, kFlags, 512>;
        SADCS_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shared)));
        // Slices (a host call's top level): consecutive tile ranges, each followed by its event.
        for (uint32_t s = 0; s < slices; s++) {
            const uint32_t first = static_cast<uint32_t>(static_cast<uint64_t>(tiles) * s / slices);
            const uint32_t last = static_cast<uint32_t>(static_cast<uint64_t>(tiles) * (s + 1) / slices);
            kernel<<<last - first, m_threads, shared, m_stream>>>(m_ctxMerge, m_texts, lists, rows, first, capacity, out, errors);
            SADCS_HIP(hipGetLastError());
            if (slices > 1) {
                SADCS_HIP(hipEventRecord(m_ctx.sliceDone[s], m_stream));
            }
        }
        if (kTop && m_ctx.kernelTimer[0]) {
            SADCS_HIP(hipEventRecord(m_ctx.kernelTimer[1], m_stream));
            m_timedMerge = true;
            times.mergeElements = lists.indexBase[lists.count];
        }
        return SADCS_OK;
    }

    bool m_timedMerge = false;
    uint32_t m_slices = 1;  // launches of the top level (more than one: the download overlapped them)
    uint4* m_lines = nullptr;
    bool m_bwtInText = false;

    // Cover 32: the lines of the whole span, then the ranks are no longer needed (the lines hold
    // them). Other covers read records from the text and the rank array.
    int buildLines() {
        if constexpr (V != kLinePositions) {
            return SADCS_OK;
        } else {
            const uint64_t count = lineCount(m_job.span);
            SADCS_TRY(m_arena.allocate(m_lines, count * (kLineBytes / sizeof(uint4))));
            if (!m_dry) {
                const uint64_t threads = count * (kLineBytes / sizeof(uint4));
                buildLinesKernel<<<static_cast<unsigned>((threads + 255) / 256), 256, 0, m_stream>>>(
                    reinterpret_cast<const uint4*>(m_job.text), m_ranks, count, m_lines,
                    kBatch ? m_job.blockSegment : nullptr, kBatch ? m_job.deviceSegments : nullptr);
                SADCS_HIP(hipGetLastError());
            }
            // Stream order keeps the build's reads of the ranks ahead of any reuse of their memory
            // (the BWT writes, when the ranks were in the caller's BWT buffer).
            unscratch(m_ranks);
            m_ctxMerge.ranks = nullptr;
            m_ctxMerge.lines = m_lines;
            return SADCS_OK;
        }
    }

    // Rows of the tiles of one level: (tiles + 1) rows of lists.count + 1 words, allocated here.
    // The top level of a fused single-text build (cover 32) borrows them from the free tail of the
    // caller's BWT buffer (after the ranks) when they fit the lines' spare words, where they move
    // before the merge writes the BWT: the table then costs no memory of the library's own while
    // every class list is live, the build's peak.
    bool m_topRowsInBwt = false;

    int allocateRows(uint32_t*& rows, uint64_t words, bool top) {
        if (top) {
            m_topRowsInBwt = false;
            if (!kBatch && V == kLinePositions && m_job.wantSA && m_job.wantBWT && words <= 2 * lineCount(m_job.span) &&
                m_bwtScratch.take(rows, words)) {
                m_topRowsInBwt = true;
                return SADCS_OK;
            }
        } else if (m_topRowsInBwt && m_bwtScratch.take(rows, words)) {
            return SADCS_OK;  // a splitter level's table, released before the top level's tiles run
        }
        return m_arena.allocate(rows, words);
    }

    int levelRows(const ListSet& lists, uint32_t*& rows, uint32_t& tiles, uint32_t* errors, bool top = false) {
        const uint32_t g = m_g, k = m_k;
        const uint32_t capacity = mergeCapacity<V>(g, k);
        const int lc = lists.count;
        const uint64_t total = lists.indexBase[lc];

        if (total <= capacity) {
            tiles = 1;
            SADCS_TRY(allocateRows(rows, 2 * (lc + 1), top));
            if (!m_dry) {
                singleTileRowsKernel<<<1, 1, 0, m_stream>>>(lists, rows);
            }
            return SADCS_OK;
        }

        // Splitters: every g-th element of every list, sorted by the level below.
        ListSet splitters{};
        splitters.count = lc;
        uint32_t base = 0;
        for (int l = 0; l < lc; l++) {
            splitters.length[l] = (lists.length[l] + g - 1) / g;
            splitters.indexBase[l] = base;
            base += splitters.length[l];
        }
        splitters.indexBase[lc] = base;
        const uint32_t count = base;
        const uint32_t boundaries = (count + k - 1) / k;
        tiles = boundaries;

        SADCS_TRY(allocateRows(rows, static_cast<uint64_t>(boundaries + 1) * (lc + 1), top));
        uint32_t* positions = nullptr;
        uint32_t* sorted = nullptr;
        uint32_t* counts = nullptr;
        Sum1* aggregates = nullptr;
        // Scratch (the caller's SA buffer when there is one): released before the tiles run.
        int status = scratch(positions, count);
        if (status == SADCS_OK) status = scratch(sorted, count);
        if (status == SADCS_OK) status = scratch(counts, static_cast<uint64_t>(boundaries) * lc);
        if (status == SADCS_OK) status = scratch(aggregates, scanScratchBytes<Sum1, kSumItems>(static_cast<uint64_t>(boundaries) * lc) / sizeof(Sum1) + 1);
        for (int l = 0; l < lc && status == SADCS_OK; l++) {
            splitters.ptr[l] = positions + splitters.indexBase[l];
        }
        if (status == SADCS_OK && !m_dry) {
            gatherSplittersKernel<<<blocksFor(count), 256, 0, m_stream>>>(lists, splitters, g, count, positions);
        }
        if (status == SADCS_OK) {
            TileOut childOut{};
            childOut.index = sorted;
            status = mergeLevel<kOutIndex>(splitters, childOut, errors);
        }
        if (status == SADCS_OK && !m_dry) {
            countBoundarySplittersKernel<<<blocksFor(static_cast<uint64_t>(boundaries) * kWave), 256, 0, m_stream>>>(splitters, sorted, count, k, boundaries, counts);
            exclusiveSum(m_stream, counts, counts, boundaries * static_cast<uint32_t>(lc), aggregates);
            boundaryRowsKernel<V, Texts><<<blocksFor(boundaries + 1ull), 256, 0, m_stream>>>(
                m_ctxMerge, m_texts, lists, splitters, sorted, counts, g, k, boundaries, rows);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                setError("merge boundaries: %s", hipGetErrorString(e));
                status = SADCS_FAILED;
            }
        }
        unscratch(aggregates);
        unscratch(counts);
        unscratch(sorted);
        unscratch(positions);
        if (status != SADCS_OK) {
            unscratch(rows);
        }
        return status;
    }

    template <int kFlags>
    int mergeLevel(const ListSet& lists, const TileOut& out, uint32_t* errors) {
        uint32_t* rows = nullptr;
        uint32_t tiles = 0;
        SADCS_TRY(levelRows(lists, rows, tiles, errors));
        const int status = launchTiles<kFlags>(lists, RowTable{rows, 0}, tiles, out, errors);
        unscratch(rows);
        return status;
    }

    int merge() {
        m_ctxMerge.ranks = m_ranks;
        const ListSet lists = topLists();
        uint32_t* errors = nullptr;
        uint32_t* firstRows = m_job.firstRows;
        bool ownFirstRows = false;
        // errors[0]: the tiles' error flags; errors[1..]: the slices' first rows.
        SADCS_TRY(m_arena.allocate(errors, 2 + sadcs_ctx::kSlices));
        fillWords(errors, 1, 0);
        if (m_ctx.mergeBegin && !m_dry) {
            SADCS_HIP(hipEventRecord(m_ctx.mergeBegin, m_stream));
        }

        if (m_job.wantBWT && !firstRows) {
            SADCS_TRY(m_arena.allocate(firstRows, m_segments));
            ownFirstRows = true;
        }
        if (m_job.wantBWT && !m_dry) {
            firstSuffixRowsKernel<V, Texts><<<blocksFor(m_segments), 256, 0, m_stream>>>(
                m_ctxMerge, m_texts, lists, m_segments, firstRows, nullptr);
        }

        ListSet mergeLists = lists;
        int status = SADCS_OK;
        uint32_t* rows = nullptr;
        uint32_t tiles = 0;
        status = levelRows(lists, rows, tiles, errors, true);
        RowTable rowTable{rows, 0};

        // A host call downloads the top level's output rows slice by slice while later slices
        // merge (a host-side choice: the kernels and their results are those of one launch).
        uint32_t slices = 1;
        if (status == SADCS_OK && m_ctx.copyStream && (m_job.hostSA || m_job.hostBWT)) {
            slices = static_cast<uint32_t>(std::min<uint64_t>({static_cast<uint64_t>(sadcs_ctx::kSlices),
                                                               (m_job.totalLength + kSliceRows - 1) / kSliceRows, tiles}));
        }
        auto launchSliceRows = [&]() -> int {
            if (slices > 1 && !m_dry) {
                sliceRowsKernel<<<1, 64, 0, m_stream>>>(rowTable, static_cast<uint32_t>(lists.count), tiles, slices, errors + 1);
                SADCS_HIP(hipMemcpyAsync(m_ctx.hostWords + kSliceRowWords, errors + 1, (slices + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, m_stream));
            }
            return SADCS_OK;
        };

        TileOut out{};
        out.firstSuffixRow = firstRows;
        out.bwtLength = static_cast<uint32_t>(m_job.totalLength);
        uint32_t* ownSA = nullptr;
        uint8_t* ownBWT = nullptr;

        if (status == SADCS_OK && m_job.wantSA) {
            // Tile-major copy of the lists into the SA buffer, free the lists; cover 32 then builds
            // the lines in their place and frees the ranks. Every tile reads its positions from its
            // own range of the SA buffer and writes its rows there.
            out.sa = m_job.sa;
            if (!out.sa) {
                status = m_arena.allocate(ownSA, m_job.span);
                out.sa = ownSA;
            }
            if (status == SADCS_OK && !m_dry) {
                tileMajorCopyKernel<<<(tiles + 3) / 4, 256, 0, m_stream>>>(lists, rows, tiles, out.sa);
            }
            // Stream order keeps the copy's reads of the lists ahead of any write to their memory.
            for (int d = 0; d < kClasses; d++) {
                m_arena.release(m_lists[d]);
            }
            if (status == SADCS_OK) {
                status = buildLines();
            }
            if (status == SADCS_OK && m_topRowsInBwt) {
                // The row table leaves the BWT buffer for the lines' spare words before any BWT write.
                const uint64_t words = static_cast<uint64_t>(tiles + 1) * (lists.count + 1);
                if (!m_dry) {
                    relocateRowsKernel<<<std::min<unsigned>(blocksFor(words), 4096), 256, 0, m_stream>>>(
                        rows, words, reinterpret_cast<uint32_t*>(m_lines));
                }
                rowTable = RowTable{reinterpret_cast<const uint32_t*>(m_lines), 1};
                unscratch(rows);
                m_topRowsInBwt = false;
            }
            if (status == SADCS_OK && m_job.wantBWT) {
                out.bwt = m_job.bwt;
                if (!out.bwt && m_lines && m_job.textScratch) {
                    out.bwt = const_cast<uint8_t*>(m_job.text);  // the merge reads only the lines now
                    m_bwtInText = true;
                }
                if (!out.bwt) {
                    status = m_arena.allocate(ownBWT, m_job.span);
                    out.bwt = ownBWT;
                }
            }
            if (status == SADCS_OK && m_job.wantBWT && !m_dry) {
                if (m_lines) {
                    firstBwtBytesFromLinesKernel<<<blocksFor(m_segments), 256, 0, m_stream>>>(m_lines, m_job.deviceSegments, m_segments, out.bwt);
                } else {
                    firstBwtBytesKernel<<<blocksFor(m_segments), 256, 0, m_stream>>>(m_job.text, m_job.deviceSegments, m_segments, out.bwt);
                }
            }
            if (status == SADCS_OK) {
                status = launchSliceRows();
            }
            if (status == SADCS_OK) {
                status = m_job.wantBWT ? launchTiles<kOutSA | kOutBWT | kInPlace>(mergeLists, rowTable, tiles, out, errors, slices)
                                       : launchTiles<kOutSA | kInPlace>(mergeLists, rowTable, tiles, out, errors, slices);
            }
        } else if (status == SADCS_OK && m_job.wantBWT) {
            out.bwt = m_job.bwt;
            if (!out.bwt) {
                status = m_arena.allocate(ownBWT, m_job.span);
                out.bwt = ownBWT;
            }
            if (status == SADCS_OK && !m_dry) {
                firstBwtBytesKernel<<<blocksFor(m_segments), 256, 0, m_stream>>>(m_job.text, m_job.deviceSegments, m_segments, out.bwt);
            }
            if (status == SADCS_OK) {
                status = launchSliceRows();
            }
            if (status == SADCS_OK) {
                status = launchTiles<kOutBWT>(mergeLists, rowTable, tiles, out, errors, slices);
            }
        }

        if (status == SADCS_OK && slices > 1 && !m_dry) {
            // The slices' rows were read back before the first slice ran.
            SADCS_HIP(hipEventSynchronize(m_ctx.sliceDone[0]));
            OutputSlices o;
            o.slices = slices;
            o.rows = m_ctx.hostWords + kSliceRowWords;
            o.totalRows = m_job.totalLength;
            if (m_job.wantSA && m_job.hostSA) {
                o.hostSA = reinterpret_cast<uint8_t*>(m_job.hostSA);
                o.deviceSA = reinterpret_cast<const uint8_t*>(out.sa);
            }
            if (m_job.wantBWT && m_job.hostBWT) {
                o.hostBWT = m_job.hostBWT;
                o.deviceBWT = out.bwt;
            }
            status = downloadSlices(m_ctx, o);
            m_job.downloaded = status == SADCS_OK;
            m_slices = slices;
        }

        if (status == SADCS_OK && !m_dry) {
            SADCS_HIP(hipMemcpyAsync(m_ctx.hostWords + 8, errors, sizeof(uint32_t), hipMemcpyDeviceToHost, m_stream));
            SADCS_HIP(hipStreamSynchronize(m_stream));
            if (m_ctx.hostWords[8] != 0) {
                setError("merge: inconsistent tile layout (internal error, code %u)", m_ctx.hostWords[8]);
                status = SADCS_FAILED;
            }
            if (m_timedMerge) {
                float ms = 0;
                SADCS_HIP(hipEventElapsedTime(&ms, m_ctx.kernelTimer[0], m_ctx.kernelTimer[1]));
                times.mergeKernel = ms * 1e-3;
            }
        }
        m_outSA = ownSA;
        m_outBWT = m_bwtInText ? const_cast<uint8_t*>(m_job.text) : ownBWT;
        m_ownBWT = ownBWT;
        m_ctxMerge.lines = nullptr;
        m_arena.release(m_lines);
        unscratch(rows);
        if (ownFirstRows) {
            m_firstRowsOwned = firstRows;
        }
        m_arena.release(errors);
        return status;
    }

 public:
    // Outputs the pipeline allocated itself (host-memory calls), valid until destruction.
    uint32_t* m_outSA = nullptr;
    uint8_t* m_outBWT = nullptr;      // the BWT, possibly in the text's buffer
    uint8_t* m_ownBWT = nullptr;      // the BWT's own allocation, if any
    uint32_t* m_firstRowsOwned = nullptr;

    void releaseOutputs() {
        m_arena.release(m_outSA);
        m_arena.release(m_ownBWT);
        m_arena.release(m_firstRowsOwned);
    }
};

}  // namespace sadcs

// ==== host side ==================================================================================

namespace sadcs {

enum class Mode { SA, BWT, Both, RankOnly };

// What a job needs besides the pipeline: the text buffer and the segment tables.
struct Layout {
    uint32_t span = 0;
    std::vector<Segment> segments;
    std::vector<uint32_t> blockSegment;  // batch only
    uint32_t maxLength = 0;
    uint64_t totalLength = 0;
};

template <int V>
uint32_t samplesOf(uint64_t n) {
    return static_cast<uint32_t>(DifferenceCover<V>::pack(static_cast<int64_t>(n)));
}

// Texts at aligned offsets; lengths >= 1.
template <int V>
Layout makeLayout(const int64_t* lengths, int32_t count, bool batch) {
    Layout layout;
    uint64_t start = 0, rows = 0, ranks = 0;
    for (int32_t i = 0; i < count; i++) {
        const uint64_t n = static_cast<uint64_t>(lengths[i]);
        Segment s{};
        s.start = static_cast<uint32_t>(start);
        s.end = static_cast<uint32_t>(start + n);
        s.rankBase = static_cast<uint32_t>(ranks);
        s.rowBase = static_cast<uint32_t>(rows);
        layout.segments.push_back(s);
        layout.maxLength = std::max<uint32_t>(layout.maxLength, static_cast<uint32_t>(n));
        rows += n;
        ranks += samplesOf<V>(n);
        start = batch ? (start + n + kSegmentAlign - 1) / kSegmentAlign * kSegmentAlign : start + n;
    }
    layout.totalLength = rows;
    layout.span = count > 0 ? layout.segments.back().end : 0;
    if (batch) {
        const uint64_t blocks = (static_cast<uint64_t>(layout.span) >> kSegmentBlockLog) + 2;
        layout.blockSegment.assign(blocks, 0);
        for (int32_t i = 0; i < count; i++) {
            const uint64_t first = layout.segments[i].start >> kSegmentBlockLog;
            const uint64_t last = i + 1 < count ? (layout.segments[i + 1].start >> kSegmentBlockLog) : blocks;
            for (uint64_t b = first; b < last; b++) {
                layout.blockSegment[b] = static_cast<uint32_t>(i);
            }
        }
    }
    return layout;
}

// Device buffers of the job's inputs, allocated from the arena (the text only if not given).
struct JobInputs {
    uint8_t* text = nullptr;
    Segment* segments = nullptr;
    uint32_t* blockSegment = nullptr;
    uint32_t* firstRows = nullptr;
};

int allocateInputs(Arena& arena, const Layout& layout, bool ownText, JobInputs& in) {
    if (ownText) {
        SADCS_TRY(arena.allocate(in.text, static_cast<uint64_t>(layout.span) + kTextPadding + 8));
    }
    SADCS_TRY(arena.allocate(in.segments, layout.segments.size()));
    SADCS_TRY(arena.allocate(in.firstRows, layout.segments.size()));
    if (!layout.blockSegment.empty()) {
        SADCS_TRY(arena.allocate(in.blockSegment, layout.blockSegment.size()));
    }
    return SADCS_OK;
}

// Runs the pipeline for one cover and text policy. In a dry-run arena only the allocation
// sequence runs; arena.highWater() is then the plan.
template <int V, class Texts>
int runPipeline(sadcs_ctx& ctx, Job& job, Texts texts, PhaseTimes& times, uint32_t** saOut, uint8_t** bwtOut, std::unique_ptr<Pipeline<V, Texts>>& keep) {
    keep.reset(new Pipeline<V, Texts>(ctx, job, texts));
    const int status = keep->run();
    times = keep->times;
    if (saOut) {
        *saOut = keep->m_outSA;
    }
    if (bwtOut) {
        *bwtOut = keep->m_outBWT;
    }
    return status;
}

// The planned peak of a job of this layout and mode, in bytes (a dry run of the allocation sequence).
template <int V>
size_t planBytes(const sadcs_ctx& base, const Layout& layout, Mode mode, bool ownText, bool ownOutputs) {
    sadcs_ctx ctx;
    ctx.params = base.params;
    ctx.arena.reset(nullptr, SIZE_MAX / 4, true);
    JobInputs in;
    if (allocateInputs(ctx.arena, layout, ownText, in) != SADCS_OK) {
        return SIZE_MAX;
    }
    Job job;
    job.text = in.text;
    job.span = layout.span;
    job.segments = layout.segments;
    job.deviceSegments = in.segments;
    job.blockSegment = in.blockSegment;
    job.batch = !layout.blockSegment.empty();
    job.rankOnly = mode == Mode::RankOnly;
    job.wantSA = mode == Mode::SA || mode == Mode::Both;
    job.wantBWT = mode == Mode::BWT || mode == Mode::Both;
    job.firstRows = in.firstRows;
    job.textScratch = ownText;
    job.maxLength = layout.maxLength;
    job.totalLength = layout.totalLength;
    uint32_t* dummySA = ownOutputs ? nullptr : reinterpret_cast<uint32_t*>(uintptr_t(1) << 40);
    uint8_t* dummyBWT = ownOutputs ? nullptr : reinterpret_cast<uint8_t*>(uintptr_t(1) << 41);
    job.sa = job.wantSA ? dummySA : nullptr;
    job.bwt = job.wantBWT ? dummyBWT : nullptr;
    PhaseTimes times;
    int status;
    if (job.batch) {
        std::unique_ptr<Pipeline<V, TextBatch>> keep;
        status = runPipeline<V, TextBatch>(ctx, job, TextBatch{}, times, nullptr, nullptr, keep);
    } else {
        std::unique_ptr<Pipeline<V, SingleText>> keep;
        status = runPipeline<V, SingleText>(ctx, job, SingleText{layout.span}, times, nullptr, nullptr, keep);
    }
    return status == SADCS_OK ? ctx.arena.highWater() : SIZE_MAX;
}

size_t planFor(const sadcs_ctx& base, const Layout& layout, Mode mode, bool ownText, bool ownOutputs) {
    return base.params.cover == 128 ? planBytes<128>(base, layout, mode, ownText, ownOutputs)
                                    : planBytes<32>(base, layout, mode, ownText, ownOutputs);
}

// The largest plan over the modes a context serves: SA, BWT, both.
size_t contextPlan(const sadcs_ctx& base, int64_t capacity) {
    if (capacity <= 1) {
        return 1 << 20;
    }
    const int64_t lengths[1] = {capacity};
    Layout layout = base.params.cover == 128 ? makeLayout<128>(lengths, 1, false) : makeLayout<32>(lengths, 1, false);
    size_t peak = 0;
    for (Mode mode : {Mode::SA, Mode::BWT, Mode::Both}) {
        peak = std::max(peak, planFor(base, layout, mode, true, true));
    }
    return peak;
}

int queryDevice(sadcs_device& device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        setError("no HIP device%s%s", e != hipSuccess ? ": " : "", e != hipSuccess ? hipGetErrorString(e) : "");
        return SADCS_UNAVAILABLE;
    }
    hipDeviceProp_t properties;
    if ((e = hipGetDeviceProperties(&properties, 0)) != hipSuccess) {
        setError("hipGetDeviceProperties: %s", hipGetErrorString(e));
        return SADCS_UNAVAILABLE;
    }
    size_t freeBytes = 0, totalBytes = 0;
    if ((e = hipMemGetInfo(&freeBytes, &totalBytes)) != hipSuccess) {
        setError("hipMemGetInfo: %s", hipGetErrorString(e));
        return SADCS_UNAVAILABLE;
    }
    // The marketing name can be empty where the runtime lacks the device-id table.
    snprintf(device.name, sizeof device.name, "%s (%s)", properties.name[0] ? properties.name : "AMD GPU", properties.gcnArchName);
    device.compute_major = properties.major;
    device.compute_minor = properties.minor;
    device.total_bytes = static_cast<int64_t>(totalBytes);
    device.free_bytes = static_cast<int64_t>(freeBytes);
    return SADCS_OK;
}

bool isPinned(const void* host) {
    hipPointerAttribute_t attributes;
    if (hipPointerGetAttributes(&attributes, host) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attributes.type == hipMemoryTypeHost;
}

// ---- host copies ---------------------------------------------------------------------------------
//
// Between pageable caller memory and the pinned staging chunks the host copies every byte, and one
// thread's memcpy (about 13 GB/s measured, first-touch page faults of fresh result arrays included)
// is well below what PCIe 5 x16 moves: returning a suffix array to pageable memory was bound by it.
// A process-wide pool of threads copies each chunk in slices. Calls from several contexts (server
// workers) take turns: they share the host's memory bandwidth anyway.
class HostCopier {
public:
    static HostCopier& instance() {
        static HostCopier pool;
        return pool;
    }

    struct Piece {
        void* dst;
        const void* src;
        size_t bytes;
    };

    void copy(void* dst, const void* src, size_t bytes) {
        const Piece piece{dst, src, bytes};
        copyPieces(&piece, 1);
    }

    // Copies every piece, the whole list split evenly over the pool by bytes (pieces are cut where
    // the slices fall), so many small copies cost one hand-off instead of one each.
    void copyPieces(const Piece* pieces, size_t count) {
        size_t bytes = 0;
        for (size_t i = 0; i < count; i++) {
            bytes += pieces[i].bytes;
        }
        if (bytes < kParallelMin || m_threads.empty()) {
            for (size_t i = 0; i < count; i++) {
                memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
            }
            return;
        }
        std::lock_guard<std::mutex> turn(m_turn);
        const size_t parts = m_threads.size() + 1;
        const size_t slice = ((bytes + parts - 1) / parts + 63) / 64 * 64;
        {
            std::lock_guard<std::mutex> lock(m_mutex);
            m_pieces = pieces;
            m_count = count;
            m_bytes = bytes;
            m_slice = slice;
            m_pending = static_cast<int>(m_threads.size());
            m_generation++;
        }
        m_start.notify_all();
        copyRange(pieces, count, 0, std::min(slice, bytes));
        std::unique_lock<std::mutex> lock(m_mutex);
        m_done.wait(lock, [&] { return m_pending == 0; });
    }

private:
    static constexpr size_t kParallelMin = size_t(1) << 20;

    HostCopier() {
        const unsigned cores = std::thread::hardware_concurrency();
        const unsigned workers = std::min(7u, cores > 1 ? cores - 1 : 0u);
        for (unsigned w = 0; w < workers; w++) {
            m_threads.emplace_back(&HostCopier::work, this, w + 1);
        }
    }

    ~HostCopier() {
        {
            std::lock_guard<std::mutex> lock(m_mutex);
            m_stop = true;
        }
        m_start.notify_all();
        for (std::thread& t : m_threads) {
            t.join();
        }
    }

    // Bytes [begin, end) of the pieces taken one after another.
    static void copyRange(const Piece* pieces, size_t count, size_t begin, size_t end) {
        size_t at = 0;
        for (size_t i = 0; i < count && at < end; i++) {
            const size_t b = std::max(begin, at), e = std::min(end, at + pieces[i].bytes);
            if (b < e) {
                memcpy(static_cast<uint8_t*>(pieces[i].dst) + (b - at), static_cast<const uint8_t*>(pieces[i].src) + (b - at), e - b);
            }
            at += pieces[i].bytes;
        }
    }

    void work(size_t part) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lock(m_mutex);
        for (;;) {
            m_start.wait(lock, [&] { return m_stop || m_generation != seen; });
            if (m_stop) {
                return;
            }
            seen = m_generation;
            const size_t begin = std::min(part * m_slice, m_bytes), end = std::min(begin + m_slice, m_bytes);
            const Piece* pieces = m_pieces;
            const size_t count = m_count;
            lock.unlock();
            if (end > begin) {
                copyRange(pieces, count, begin, end);
            }
            lock.lock();
            if (--m_pending == 0) {
                m_done.notify_one();
            }
        }
    }

    std::vector<std::thread> m_threads;
    std::mutex m_turn, m_mutex;
    std::condition_variable m_start, m_done;
    uint64_t m_generation = 0;
    int m_pending = 0;
    bool m_stop = false;
    const Piece* m_pieces = nullptr;
    size_t m_count = 0;
    size_t m_bytes = 0, m_slice = 0;
};

// Host to device through two pinned chunks, or directly when the host buffer is pinned.
int upload(sadcs_ctx& ctx, void* device, const void* host, size_t bytes, bool& pinned) {
    pinned = true;
    if (bytes == 0) {
        return SADCS_OK;
    }
    if (isPinned(host) || !ctx.staging[0]) {
        pinned = ctx.staging[0] != nullptr || isPinned(host);
        SADCS_HIP(hipMemcpyAsync(device, host, bytes, hipMemcpyHostToDevice, ctx.stream));
        return SADCS_OK;
    }
    for (size_t k = 0, offset = 0; offset < bytes; k++, offset += kStagingChunk) {
        const size_t length = std::min(kStagingChunk, bytes - offset);
        const int b = static_cast<int>(k & 1);
        SADCS_HIP(hipEventSynchronize(ctx.stagingDone[b]));
        HostCopier::instance().copy(ctx.staging[b], static_cast<const uint8_t*>(host) + offset, length);
        SADCS_HIP(hipMemcpyAsync(static_cast<uint8_t*>(device) + offset, ctx.staging[b], length, hipMemcpyHostToDevice, ctx.stream));
        SADCS_HIP(hipEventRecord(ctx.stagingDone[b], ctx.stream));
    }
    return SADCS_OK;
}

// Device to host after the stream's earlier work; returns with the data on the host.
int download(sadcs_ctx& ctx, void* host, const void* device, size_t bytes) {
    if (bytes == 0) {
        return SADCS_OK;
    }
    if (isPinned(host) || !ctx.staging[0]) {
        SADCS_HIP(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, ctx.stream));
        SADCS_HIP(hipStreamSynchronize(ctx.stream));
        return SADCS_OK;
    }
    const size_t chunks = (bytes + kStagingChunk - 1) / kStagingChunk;
    auto issue = [&](size_t k) -> int {
        const size_t offset = k * kStagingChunk;
        const int b = static_cast<int>(k & 1);
        SADCS_HIP(hipMemcpyAsync(ctx.staging[b], static_cast<const uint8_t*>(device) + offset,
                                 std::min(kStagingChunk, bytes - offset), hipMemcpyDeviceToHost, ctx.stream));
        SADCS_HIP(hipEventRecord(ctx.stagingDone[b], ctx.stream));
        return SADCS_OK;
    };
    SADCS_TRY(issue(0));
    for (size_t k = 0; k < chunks; k++) {
        if (k + 1 < chunks) {
            SADCS_TRY(issue(k + 1));
        }
        const int b = static_cast<int>(k & 1);
        const size_t offset = k * kStagingChunk;
        SADCS_HIP(hipEventSynchronize(ctx.stagingDone[b]));
        HostCopier::instance().copy(static_cast<uint8_t*>(host) + offset, ctx.staging[b], std::min(kStagingChunk, bytes - offset));
    }
    return SADCS_OK;
}

void releaseOverlap(sadcs_ctx& ctx) {
    if (ctx.copyStream) {
        (void)hipStreamSynchronize(ctx.copyStream);
        (void)hipStreamDestroy(ctx.copyStream);
        ctx.copyStream = nullptr;
    }
    for (hipEvent_t& e : ctx.sliceDone) {
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
    if (ctx.mergeBegin) {
        (void)hipEventDestroy(ctx.mergeBegin);
        ctx.mergeBegin = nullptr;
    }
}

// Downloads the outputs of a top-level merge launched in slices (OutputSlices) while it runs: after
// slice s the SA rows [rows[s], rows[s + 1]) and the BWT bytes [rows[s], rows[s + 1]) are final (BWT
// byte j comes from row j - 1 or row j, byte 0 of each text from before the merge), so the copy
// stream waits for that slice's event and copies them, through the staging chunks for pageable
// memory and directly into pinned memory. Returns with the data on the host.
int downloadSlices(sadcs_ctx& ctx, const OutputSlices& o) {
    struct Chunk {
        uint8_t* host;
        const uint8_t* device;
        size_t bytes;
        uint32_t slice;
        int buffer;  // staging chunk, or -1: a direct copy into pinned memory
    };
    std::vector<Chunk> chunks;
    const bool pinnedSA = o.hostSA && isPinned(o.hostSA);
    const bool pinnedBWT = o.hostBWT && isPinned(o.hostBWT);
    uint64_t staged = 0;
    auto add = [&](uint8_t* host, const uint8_t* device, size_t bytes, uint32_t slice, bool pinned) {
        for (size_t offset = 0; offset < bytes; offset += kStagingChunk) {
            const size_t length = pinned ? bytes : std::min(kStagingChunk, bytes - offset);
            chunks.push_back(Chunk{host + offset, device + offset, length, slice, pinned ? -1 : static_cast<int>(staged++ & 1)});
            if (pinned) {
                break;
            }
        }
    };
    for (uint32_t s = 0; s < o.slices; s++) {
        const uint64_t a = o.rows[s], b = o.rows[s + 1];
        if (a > b || b > o.totalRows || (s + 1 == o.slices && b != o.totalRows)) {
            setError("merge slices: rows %llu..%llu of %llu (internal error)", static_cast<unsigned long long>(a),
                     static_cast<unsigned long long>(b), static_cast<unsigned long long>(o.totalRows));
            return SADCS_FAILED;
        }
        if (o.hostSA) {
            add(o.hostSA + 4 * a, o.deviceSA + 4 * a, 4 * (b - a), s, pinnedSA);
        }
        if (o.hostBWT) {
            add(o.hostBWT + a, o.deviceBWT + a, b - a, s, pinnedBWT);
        }
    }
    uint32_t waited = UINT32_MAX;
    auto issue = [&](size_t k) -> int {
        const Chunk& c = chunks[k];
        if (c.slice != waited) {
            SADCS_HIP(hipStreamWaitEvent(ctx.copyStream, ctx.sliceDone[c.slice], 0));
            waited = c.slice;
        }
        uint8_t* target = c.buffer < 0 ? c.host : ctx.staging[c.buffer];
        SADCS_HIP(hipMemcpyAsync(target, c.device, c.bytes, hipMemcpyDeviceToHost, ctx.copyStream));
        if (c.buffer >= 0) {
            SADCS_HIP(hipEventRecord(ctx.stagingDone[c.buffer], ctx.copyStream));
        }
        return SADCS_OK;
    };
    // Staged chunks alternate between the two buffers; a buffer is refilled (chunk k + 1) only
    // after the host has copied out everything before chunk k.
    if (!chunks.empty()) {
        SADCS_TRY(issue(0));
    }
    for (size_t k = 0; k < chunks.size(); k++) {
        if (k + 1 < chunks.size()) {
            SADCS_TRY(issue(k + 1));
        }
        const Chunk& c = chunks[k];
        if (c.buffer >= 0) {
            SADCS_HIP(hipEventSynchronize(ctx.stagingDone[c.buffer]));
            HostCopier::instance().copy(c.host, ctx.staging[c.buffer], c.bytes);
        }
    }
    SADCS_HIP(hipStreamSynchronize(ctx.copyStream));
    return SADCS_OK;
}

int createContext(size_t arenaBytes, const Params& params, int32_t capacity, sadcs_ctx*& out) {
    out = nullptr;
    sadcs_device device;
    SADCS_TRY(queryDevice(device));
    if (static_cast<int64_t>(arenaBytes) + kMemoryReserve > device.free_bytes) {
        setError("needs %lld MiB of device memory plus %lld MiB reserve, %lld MiB free",
                 static_cast<long long>(arenaBytes >> 20), static_cast<long long>(kMemoryReserve >> 20),
                 static_cast<long long>(device.free_bytes >> 20));
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    std::unique_ptr<sadcs_ctx> ctx(new (std::nothrow) sadcs_ctx);
    if (!ctx) {
        setError("out of host memory");
        return SADCS_FAILED;
    }
    ctx->params = params;
    ctx->capacity = capacity;
    SADCS_HIP(hipSetDevice(0));
    SADCS_TRY(ensureTables());
    SADCS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    SADCS_HIP(hipDeviceGetAttribute(&ctx->sms, hipDeviceAttributeMultiprocessorCount, 0));
    ctx->arenaBytes = std::max<size_t>(arenaBytes, Arena::kAlign);
    {
        const hipError_t e = hipMalloc(&ctx->arenaBase, ctx->arenaBytes);
        if (e != hipSuccess) {
            setError("hipMalloc(%zu): %s", ctx->arenaBytes, hipGetErrorString(e));
            (void)hipStreamDestroy(ctx->stream);
            return statusOf(e);
        }
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&ctx->hostWords), kHostWords * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        setError("hipHostMalloc of readback words failed");
        (void)hipFree(ctx->arenaBase);
        (void)hipStreamDestroy(ctx->stream);
        return SADCS_FAILED;
    }
    for (int i = 0; i < 2; i++) {
        if (hipEventCreate(&ctx->kernelTimer[i]) != hipSuccess) {
            (void)hipGetLastError();
            ctx->kernelTimer[i] = nullptr;
        }
    }
    if (!ctx->kernelTimer[0] || !ctx->kernelTimer[1]) {
        ctx->kernelTimer[0] = nullptr;
    }
    for (int i = 0; i < 2; i++) {
        if (hipHostMalloc(reinterpret_cast<void**>(&ctx->staging[i]), kStagingChunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&ctx->stagingDone[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (int j = 0; j <= i; j++) {
                if (ctx->staging[j]) {
                    (void)hipHostFree(ctx->staging[j]);
                    ctx->staging[j] = nullptr;
                }
            }
            break;
        }
    }
    // Without its stream or any of its events the context downloads after the merge.
    bool overlap = ctx->staging[0] != nullptr && hipStreamCreateWithFlags(&ctx->copyStream, hipStreamNonBlocking) == hipSuccess;
    overlap = overlap && hipEventCreate(&ctx->mergeBegin) == hipSuccess;
    for (int i = 0; i < sadcs_ctx::kSlices && overlap; i++) {
        overlap = hipEventCreate(&ctx->sliceDone[i]) == hipSuccess;
    }
    if (!overlap) {
        (void)hipGetLastError();
        releaseOverlap(*ctx);
    }
    out = ctx.release();
    return SADCS_OK;
}

void destroyContext(sadcs_ctx* ctx) {
    if (!ctx) {
        return;
    }
    (void)hipStreamSynchronize(ctx->stream);
    releaseOverlap(*ctx);
    for (int i = 0; i < 2; i++) {
        if (ctx->kernelTimer[i]) {
            (void)hipEventDestroy(ctx->kernelTimer[i]);
        }
        if (ctx->stagingDone[i]) {
            (void)hipEventDestroy(ctx->stagingDone[i]);
        }
        if (ctx->staging[i]) {
            (void)hipHostFree(ctx->staging[i]);
        }
    }
    if (ctx->hostWords) {
        (void)hipHostFree(ctx->hostWords);
    }
    if (ctx->arenaBase) {
        (void)hipFree(ctx->arenaBase);
    }
    if (ctx->stream) {
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

void fillStats(sadcs_stats* stats, const PhaseTimes& t, size_t deviceBytes, int cover) {
    if (!stats) {
        return;
    }
    stats->device_bytes = static_cast<int64_t>(deviceBytes);
    stats->ranking_bytes = static_cast<int64_t>(t.rankingBytes);
    stats->class_sort_bytes = static_cast<int64_t>(t.classBytes);
    stats->merge_bytes = static_cast<int64_t>(t.mergeBytes);
    stats->ranking_seconds = t.ranking;
    stats->class_sort_seconds = t.classes;
    stats->merge_seconds = t.merge;
    stats->bwt_seconds = t.bwt;
    stats->device_seconds = t.ranking + t.classes + t.merge + t.bwt;
    stats->merge_kernel_seconds = t.mergeKernel;
    stats->merge_kernel_elements = static_cast<int64_t>(t.mergeElements);
    stats->character_passes = t.characterPasses;
    stats->doubling_rounds = t.doublingRounds;
    stats->cover = cover;
}

// A whole build of one layout on a context: texts either uploaded from the host (hostText) or
// already on the device (deviceText, laid out as the layout says, zero padded).
struct BuildRequest {
    Mode mode = Mode::Both;
    const uint8_t* hostText = nullptr;       // concatenated without gaps, lengths from the layout
    const uint8_t* const* hostTexts = nullptr;  // or one pointer per text (a batch)
    const int64_t* hostLengths = nullptr;
    const uint8_t* deviceText = nullptr;
    uint32_t* deviceSA = nullptr;            // device outputs (device calls) or null
    uint8_t* deviceBWT = nullptr;
    int32_t* hostSA = nullptr;               // host outputs (host calls) or null
    uint8_t* hostBWT = nullptr;
    int32_t* hostPrimaries = nullptr;        // one per text
    uint32_t* hostRanks = nullptr;           // rank-only
};

template <int V, class Texts>
int buildWith(sadcs_ctx& ctx, const Layout& layout, const BuildRequest& rq, Texts texts, sadcs_stats* stats) {
    const auto wallStart = std::chrono::steady_clock::now();
    Arena& arena = ctx.arena;
    arena.reset(ctx.arenaBase, ctx.arenaBytes, false);
    const bool deviceCall = rq.deviceText != nullptr;
    JobInputs in;
    SADCS_TRY(allocateInputs(arena, layout, !deviceCall, in));
    const uint8_t* text = deviceCall ? rq.deviceText : in.text;
    const bool batch = !layout.blockSegment.empty();
    if constexpr (Texts::kBatch) {
        texts.blockSegment = in.blockSegment;
        texts.segments = in.segments;
    }

    // Segment tables through the pinned readback words when they fit, else synchronously.
    const size_t segmentBytes = layout.segments.size() * sizeof(Segment);
    SADCS_HIP(hipMemcpy(in.segments, layout.segments.data(), segmentBytes, hipMemcpyHostToDevice));
    if (batch) {
        SADCS_HIP(hipMemcpy(in.blockSegment, layout.blockSegment.data(), layout.blockSegment.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    double uploadSeconds = 0;
    bool pinned = false;
    if (!deviceCall) {
        const auto start = std::chrono::steady_clock::now();
        SADCS_HIP(hipMemsetAsync(in.text, 0, static_cast<size_t>(layout.span) + kTextPadding + 8, ctx.stream));
        if (!batch) {
            SADCS_TRY(upload(ctx, in.text, rq.hostText, layout.span, pinned));
        } else {
            const uint8_t* src = rq.hostText;
            for (size_t i = 0; i < layout.segments.size(); i++) {
                const size_t n = layout.segments[i].end - layout.segments[i].start;
                SADCS_TRY(upload(ctx, in.text + layout.segments[i].start, rq.hostTexts ? rq.hostTexts[i] : src, n, pinned));
                src += n;
            }
        }
        SADCS_HIP(hipStreamSynchronize(ctx.stream));
        uploadSeconds = secondsSince(start);
    }

    Job job;
    job.text = text;
    job.span = layout.span;
    job.segments = layout.segments;
    job.deviceSegments = in.segments;
    job.blockSegment = in.blockSegment;
    job.batch = batch;
    job.rankOnly = rq.mode == Mode::RankOnly;
    job.wantSA = rq.mode == Mode::SA || rq.mode == Mode::Both;
    job.wantBWT = rq.mode == Mode::BWT || rq.mode == Mode::Both;
    job.sa = job.wantSA ? rq.deviceSA : nullptr;
    job.bwt = job.wantBWT ? rq.deviceBWT : nullptr;
    job.firstRows = in.firstRows;
    job.textScratch = !deviceCall;
    job.maxLength = layout.maxLength;
    job.totalLength = layout.totalLength;
    if (!deviceCall) {
        job.hostSA = job.wantSA ? rq.hostSA : nullptr;
        job.hostBWT = job.wantBWT ? rq.hostBWT : nullptr;
    }

    PhaseTimes times;
    uint32_t* outSA = nullptr;
    uint8_t* outBWT = nullptr;
    std::unique_ptr<Pipeline<V, Texts>> keep;
    const int status = runPipeline<V, Texts>(ctx, job, texts, times, &outSA, &outBWT, keep);
    size_t deviceBytes = arena.highWater();
    if (deviceCall) {
        deviceBytes += static_cast<size_t>(layout.span) + kTextPadding;
        deviceBytes += job.wantSA ? static_cast<size_t>(layout.totalLength) * 4 : 0;
        deviceBytes += job.wantBWT ? static_cast<size_t>(layout.totalLength) : 0;
    }
    fillStats(stats, times, deviceBytes, V);
    if (stats) {
        stats->upload_seconds = uploadSeconds;
        stats->pinned = pinned ? 1 : 0;
    }
    if (status != SADCS_OK) {
        return status;
    }

    const auto downloadStart = std::chrono::steady_clock::now();
    const size_t count = layout.segments.size();
    if (job.wantBWT && rq.hostPrimaries) {
        SADCS_TRY(download(ctx, ctx.hostWords + 16, in.firstRows, count <= 1000 ? count * sizeof(uint32_t) : 0));
        std::vector<uint32_t> rows;
        const uint32_t* firstRows = ctx.hostWords + 16;
        if (count > 1000) {
            rows.resize(count);
            SADCS_HIP(hipMemcpy(rows.data(), in.firstRows, count * sizeof(uint32_t), hipMemcpyDeviceToHost));
            firstRows = rows.data();
        }
        for (size_t i = 0; i < count; i++) {
            rq.hostPrimaries[i] = static_cast<int32_t>(firstRows[i]) + 1;
        }
    }
    if (job.rankOnly && rq.hostRanks) {
        SADCS_TRY(download(ctx, rq.hostRanks, keep->ranks(), sizeof(uint32_t) * samplesOf<V>(layout.span)));
    }
    if (!deviceCall && !job.downloaded) {
        if (job.wantSA && rq.hostSA) {
            SADCS_TRY(download(ctx, rq.hostSA, outSA, static_cast<size_t>(layout.totalLength) * 4));
        }
        if (job.wantBWT && rq.hostBWT) {
            SADCS_TRY(download(ctx, rq.hostBWT, outBWT, static_cast<size_t>(layout.totalLength)));
        }
    }
    SADCS_HIP(hipStreamSynchronize(ctx.stream));
    if (stats) {
        stats->download_seconds = secondsSince(downloadStart) + times.download;
        (void)wallStart;
    }
    return SADCS_OK;
}

int build(sadcs_ctx& ctx, const Layout& layout, const BuildRequest& rq, sadcs_stats* stats) {
    (void)hipGetLastError();  // errors of earlier, unrelated calls are not this construction's
    const bool batch = !layout.blockSegment.empty();
    if (ctx.params.cover == 128) {
        return batch ? buildWith<128, TextBatch>(ctx, layout, rq, TextBatch{}, stats)
                     : buildWith<128, SingleText>(ctx, layout, rq, SingleText{layout.span}, stats);
    }
    return batch ? buildWith<32, TextBatch>(ctx, layout, rq, TextBatch{}, stats)
                 : buildWith<32, SingleText>(ctx, layout, rq, SingleText{layout.span}, stats);
}

Layout layoutFor(const Params& params, const int64_t* lengths, int32_t count, bool batch) {
    return params.cover == 128 ? makeLayout<128>(lengths, count, batch) : makeLayout<32>(lengths, count, batch);
}

bool validParams(const Params& p) {
    if (p.cover != 32 && p.cover != 128) {
        setError("cover must be 32 or 128, not %d", p.cover);
        return false;
    }
    if (p.g < 2 || p.g > 4096 || p.k > 4096 || (p.threads != 256 && p.threads != 512)) {
        setError("invalid merge parameters (splitter spacing %u, tile splitters %u, tile threads %d)", p.g, p.k, p.threads);
        return false;
    }
    const uint32_t k = p.k ? p.k : p.cover == 128 ? defaultTileSplitters<128, true>(p.g, p.threads) : defaultTileSplitters<32, true>(p.g, p.threads);
    const uint32_t capacity = p.cover == 128 ? mergeCapacity<128>(p.g, k) : mergeCapacity<32>(p.g, k);
    const size_t shared = p.cover == 128 ? mergeSharedBytes<128, true>(capacity) : mergeSharedBytes<32, true>(capacity);
    if (capacity > 65535 || shared > 160 * 1024 - 1024) {
        setError("merge tiles of up to %u elements need %zu bytes of LDS, more than a workgroup has", capacity, shared);
        return false;
    }
    return true;
}

// Builds with the given context, or a temporary one sized for this layout.
int buildOn(sadcs_ctx* context, const Params& params, const Layout& layout, const BuildRequest& rq, sadcs_stats* stats) {
    if (context) {
        const bool deviceCall = rq.deviceText != nullptr;
        // A device call brings its text and its outputs; a host call has the library own both.
        const size_t need = planFor(*context, layout, rq.mode, !deviceCall, !deviceCall);
        if (need > context->arenaBytes) {
            setError("input needs %zu bytes of the context's %zu", need, context->arenaBytes);
            return SADCS_NOT_ENOUGH_MEMORY;
        }
        return build(*context, layout, rq, stats);
    }
    const auto start = std::chrono::steady_clock::now();
    sadcs_ctx planner;
    planner.params = params;
    const size_t need = planFor(planner, layout, rq.mode, rq.deviceText == nullptr, rq.deviceText == nullptr);
    if (need == SIZE_MAX) {
        setError("planning the device memory failed");
        return SADCS_FAILED;
    }
    sadcs_ctx* ctx = nullptr;
    SADCS_TRY(createContext(need, params, 0, ctx));
    const double setup = secondsSince(start);
    const int status = build(*ctx, layout, rq, stats);
    if (stats) {
        stats->setup_seconds += setup;
    }
    destroyContext(ctx);
    return status;
}

}  // namespace sadcs

// ==== C ABI ======================================================================================

using namespace sadcs;

namespace sadcs {

// Gathers of random 128-byte lines of a lines array: whole lines (eight lanes, 16 bytes each), or
// as the top-level merge reads its records (a lane pair per line: one lane three quads from the
// first half, the other three from the second). Each thread writes the XOR of what it loaded.
__global__ __launch_bounds__(256) void gatherLinesKernel(const uint4* __restrict__ lines, uint32_t lineCount, uint64_t gathers,
                                                         int wholeLines, uint32_t* __restrict__ out) {
    const uint64_t thread = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t lanes = wholeLines ? 8 : 2;
    const uint32_t part = static_cast<uint32_t>(thread % lanes);
    const uint64_t slots = static_cast<uint64_t>(gridDim.x) * blockDim.x / lanes;
    uint32_t acc = 0;
    for (uint64_t g = thread / lanes; g < gathers; g += slots) {
        uint32_t h = static_cast<uint32_t>(g) * 0x9E3779B9u ^ static_cast<uint32_t>(g >> 32);
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        h *= 0xC2B2AE35u;
        h ^= h >> 16;
        const uint4* line = lines + (static_cast<uint64_t>(h) * lineCount >> 32) * 8;
        if (wholeLines) {
            const uint4 v = line[part];
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        } else {
            const uint4* from = line + (part ? 4 : 0) + (h & 1);
            const uint4 a = from[0], b = from[1], c = from[2];
            acc ^= a.x ^ a.w ^ b.y ^ b.z ^ c.x ^ c.w;
        }
    }
    out[thread] = acc;
}

}  // namespace sadcs

extern "C" {

int sadcs_hip_query(sadcs_device* device) {
    tLastError[0] = '\0';
    sadcs_device local;
    return queryDevice(device ? *device : local);
}

int64_t sadcs_hip_sample_size(int32_t n, int32_t cover) {
    if (n <= 0) {
        return 0;
    }
    return cover == 128 ? DifferenceCover<128>::pack(n) : DifferenceCover<32>::pack(n);
}

int64_t sadcs_hip_required_bytes(int32_t n) {
    if (n <= 0) {
        return 0;
    }
    sadcs_ctx planner;
    const size_t plan = contextPlan(planner, n);
    return plan == SIZE_MAX ? -1 : static_cast<int64_t>(plan);
}

int sadcs_hip_tune(const void* d_texts, const int64_t* lengths, int32_t count, const sadcs_params* base,
                   sadcs_params* best, double* best_seconds) {
    tLastError[0] = '\0';
    if (!d_texts || !lengths || count < 1 || !best) {
        setError("invalid arguments");
        return SADCS_FAILED;
    }
    int64_t total = 0;
    for (int32_t i = 0; i < count; i++) {
        total += lengths[i];
    }
    if (total < 2 || total > INT32_MAX) {
        setError("a tuning sample needs 2 to 2^31 - 1 bytes");
        return SADCS_FAILED;
    }
    const int32_t cover = base && base->cover ? base->cover : 32;
    const bool batch = count > 1;
    void* d_sa = nullptr;
    void* d_bwt = nullptr;
    SADCS_HIP(hipMalloc(&d_sa, static_cast<size_t>(total) * sizeof(uint32_t)));
    if (hipMalloc(&d_bwt, static_cast<size_t>(total)) != hipSuccess) {
        (void)hipFree(d_sa);
        setError("hipMalloc failed");
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    std::vector<int32_t> primaries(static_cast<size_t>(count));
    double bestTime = 1e30;
    int status = SADCS_OK;
    // One candidate: a context planned for it, two builds, the faster device time kept.
    auto tryConfig = [&](const sadcs_params& p) -> int {
        sadcs_ctx* ctx = nullptr;
        SADCS_TRY(sadcs_hip_ctx_create_batch(lengths, count, &p, &ctx));
        double fastest = 1e30;
        int s = SADCS_OK;
        for (int run = 0; run < 2 && s == SADCS_OK; run++) {
            sadcs_stats stats{};
            s = batch ? sadcs_hip_batch_build_device(ctx, d_texts, lengths, count, d_sa, d_bwt, primaries.data(), &stats)
                      : sadcs_hip_build_device(ctx, d_texts, static_cast<int32_t>(lengths[0]), d_sa, d_bwt, primaries.data(), &stats);
            fastest = std::min(fastest, stats.device_seconds);
        }
        sadcs_hip_ctx_destroy(ctx);
        if (s == SADCS_OK && fastest < bestTime) {
            bestTime = fastest;
            *best = p;
        }
        return s;
    };
    // Cover 128's 31 lists want smaller spacings than cover 32's 7 (the spill room of a tile).
    const std::vector<uint32_t> spacings = cover == 128 ? std::vector<uint32_t>{4, 6, 8, 12} : std::vector<uint32_t>{8, 12, 16, 24, 32};
    for (int threads : {512, 256}) {
        for (uint32_t g : spacings) {
            const uint32_t most = cover == 128 ? (batch ? defaultTileSplitters<128, true>(g, threads) : defaultTileSplitters<128, false>(g, threads))
                                               : (batch ? defaultTileSplitters<32, true>(g, threads) : defaultTileSplitters<32, false>(g, threads));
            for (uint32_t k : {most, std::max(1u, 3 * most / 4), std::max(1u, most / 2)}) {
                sadcs_params p{};
                p.cover = cover;
                p.splitter_spacing = g;
                p.tile_splitters = k;
                p.tile_threads = static_cast<uint32_t>(threads);
                if (status == SADCS_OK) {
                    status = tryConfig(p);
                }
            }
        }
    }
    (void)hipFree(d_bwt);
    (void)hipFree(d_sa);
    if (status == SADCS_OK && best_seconds) {
        *best_seconds = bestTime;
    }
    return status;
}

int64_t sadcs_hip_device_required_bytes(int32_t n, int32_t want_sa, int32_t want_bwt) {
    if (n <= 0 || (!want_sa && !want_bwt)) {
        return 0;
    }
    sadcs_ctx planner;
    const int64_t lengths[1] = {n};
    const Layout layout = makeLayout<32>(lengths, 1, false);
    const Mode mode = want_sa && want_bwt ? Mode::Both : want_sa ? Mode::SA : Mode::BWT;
    const size_t plan = planFor(planner, layout, mode, false, false);
    return plan == SIZE_MAX ? -1 : static_cast<int64_t>(plan);
}

int sadcs_hip_ctx_create_batch(const int64_t* lengths, int32_t count, const sadcs_params* params, sadcs_ctx** context) {
    tLastError[0] = '\0';
    if (!context || (!lengths && count > 0) || count < 0) {
        setError("invalid arguments");
        return SADCS_FAILED;
    }
    *context = nullptr;
    Params p;
    p.apply(params);
    if (!validParams(p)) {
        return SADCS_FAILED;
    }
    int64_t total = 0, longest = 0;
    for (int32_t i = 0; i < count; i++) {
        if (lengths[i] < 1 || lengths[i] > INT32_MAX) {
            setError("text %d has length %lld", i, static_cast<long long>(lengths[i]));
            return SADCS_FAILED;
        }
        total += lengths[i];
        longest = std::max(longest, lengths[i]);
    }
    if (total + int64_t(count) * kSegmentAlign > INT32_MAX) {
        setError("a batch holds at most 2^31 - 1 positions");
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    sadcs_ctx planner;
    planner.params = p;
    size_t plan = contextPlan(planner, longest);
    if (count > 0) {
        const Layout layout = layoutFor(p, lengths, count, true);
        for (Mode mode : {Mode::SA, Mode::BWT, Mode::Both}) {
            plan = std::max(plan, planFor(planner, layout, mode, true, true));
        }
    }
    if (plan == SIZE_MAX) {
        setError("planning the device memory failed");
        return SADCS_FAILED;
    }
    return createContext(plan, p, static_cast<int32_t>(longest), *context);
}

int sadcs_hip_ctx_create(int32_t capacity, const sadcs_params* params, sadcs_ctx** context) {
    tLastError[0] = '\0';
    if (!context) {
        setError("null context pointer");
        return SADCS_FAILED;
    }
    *context = nullptr;
    Params p;
    p.apply(params);
    if (!validParams(p)) {
        return SADCS_FAILED;
    }
    sadcs_device device;
    SADCS_TRY(queryDevice(device));
    sadcs_ctx planner;
    planner.params = p;
    const size_t plan = contextPlan(planner, std::max(capacity, 0));
    if (plan == SIZE_MAX) {
        setError("planning the device memory failed");
        return SADCS_FAILED;
    }
    return createContext(plan, p, std::max(capacity, 0), *context);
}

void sadcs_hip_ctx_destroy(sadcs_ctx* context) {
    destroyContext(context);
}

static int smallCase(const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt, int32_t* primary) {
    if (n == 1) {
        if (sa) {
            sa[0] = 0;
        }
        if (bwt) {
            bwt[0] = text[0];
        }
        if (primary) {
            *primary = 1;
        }
    } else if (primary) {
        *primary = 0;
    }
    return SADCS_OK;
}

int sadcs_hip_build(sadcs_ctx* context, const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt, int32_t* primary, sadcs_stats* stats) {
    tLastError[0] = '\0';
    if (stats) {
        *stats = sadcs_stats{};
    }
    if (primary) {
        *primary = n > 0 ? -1 : 0;
    }
    if (n <= 1) {
        sadcs_device device;
        const int status = context ? SADCS_OK : queryDevice(device);
        return status == SADCS_OK ? smallCase(text, n, sa, bwt, primary) : status;
    }
    if (context && n > context->capacity) {
        setError("input of %d bytes exceeds the context's capacity of %d bytes", n, context->capacity);
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    if (!sa && !bwt) {
        return SADCS_OK;
    }
    Params params = context ? context->params : Params{};
    const int64_t lengths[1] = {n};
    const Layout layout = layoutFor(params, lengths, 1, false);
    BuildRequest rq;
    rq.mode = sa && bwt ? Mode::Both : sa ? Mode::SA : Mode::BWT;
    rq.hostText = text;
    rq.hostSA = sa;
    rq.hostBWT = bwt;
    int32_t first = -1;
    rq.hostPrimaries = &first;
    const int status = buildOn(context, params, layout, rq, stats);
    if (status == SADCS_OK && primary && bwt) {
        *primary = first;
    }
    return status;
}

int sadcs_hip_build_device(sadcs_ctx* context, const void* d_text, int32_t n, void* d_sa, void* d_bwt, int32_t* primary, sadcs_stats* stats) {
    tLastError[0] = '\0';
    if (stats) {
        *stats = sadcs_stats{};
    }
    if (primary) {
        *primary = n > 0 ? -1 : 0;
    }
    if (n <= 0 || (!d_sa && !d_bwt)) {
        return SADCS_OK;
    }
    if (!d_text || (reinterpret_cast<uintptr_t>(d_text) & 15) != 0) {
        setError("the device text must be 16-byte aligned");
        return SADCS_FAILED;
    }
    if ((reinterpret_cast<uintptr_t>(d_sa) & 3) != 0) {
        setError("the device suffix array must be 4-byte aligned");
        return SADCS_FAILED;
    }
    if (context && n > context->capacity) {
        setError("input of %d bytes exceeds the context's capacity of %d bytes", n, context->capacity);
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    if (n == 1) {
        uint8_t c = 0;
        SADCS_HIP(hipMemcpy(&c, d_text, 1, hipMemcpyDeviceToHost));
        int32_t zero = 0;
        if (d_sa) {
            SADCS_HIP(hipMemcpy(d_sa, &zero, 4, hipMemcpyHostToDevice));
        }
        if (d_bwt) {
            SADCS_HIP(hipMemcpy(d_bwt, &c, 1, hipMemcpyHostToDevice));
        }
        if (primary) {
            *primary = 1;
        }
        return SADCS_OK;
    }
    Params params = context ? context->params : Params{};
    const int64_t lengths[1] = {n};
    const Layout layout = layoutFor(params, lengths, 1, false);
    BuildRequest rq;
    rq.mode = d_sa && d_bwt ? Mode::Both : d_sa ? Mode::SA : Mode::BWT;
    rq.deviceText = static_cast<const uint8_t*>(d_text);
    rq.deviceSA = static_cast<uint32_t*>(d_sa);
    rq.deviceBWT = static_cast<uint8_t*>(d_bwt);
    int32_t first = -1;
    rq.hostPrimaries = &first;
    const int status = buildOn(context, params, layout, rq, stats);
    if (status == SADCS_OK && primary && d_bwt) {
        *primary = first;
    }
    return status;
}

int sadcs_hip_rank_sample(const uint8_t* text, int32_t n, uint32_t* ranks, sadcs_stats* stats) {
    tLastError[0] = '\0';
    if (stats) {
        *stats = sadcs_stats{};
    }
    if (n <= 0) {
        sadcs_device device;
        return queryDevice(device);
    }
    Params params;
    const int64_t lengths[1] = {n};
    const Layout layout = layoutFor(params, lengths, 1, false);
    BuildRequest rq;
    rq.mode = Mode::RankOnly;
    rq.hostText = text;
    rq.hostRanks = ranks;
    return buildOn(nullptr, params, layout, rq, stats);
}

void* sadcs_hip_device_alloc(size_t bytes) {
    void* pointer = nullptr;
    if (hipMalloc(&pointer, std::max<size_t>(bytes, 1)) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return pointer;
}

void sadcs_hip_device_free(void* pointer) {
    if (pointer) {
        (void)hipFree(pointer);
    }
}

int sadcs_hip_memcpy(void* dst, const void* src, size_t bytes, int32_t kind) {
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    SADCS_HIP(hipMemcpy(dst, src, bytes, k));
    return SADCS_OK;
}

int sadcs_hip_memset(void* dst, int32_t value, size_t bytes) {
    SADCS_HIP(hipMemset(dst, value, bytes));
    return SADCS_OK;
}

int sadcs_hip_synchronize(void) {
    SADCS_HIP(hipDeviceSynchronize());
    return SADCS_OK;
}

int sadcs_hip_gather_lines(const void* d_lines, int64_t line_count, int64_t gathers, int32_t whole_lines) {
    if (!d_lines || line_count <= 0 || line_count > INT64_C(0xFFFFFFFF) || gathers < 0) {
        setError("gather_lines: bad arguments");
        return SADCS_FAILED;
    }
    constexpr unsigned kBlocks = 2048, kThreads = 256;
    uint32_t* out = nullptr;
    SADCS_HIP(hipMalloc(&out, sizeof(uint32_t) * kBlocks * kThreads));
    sadcs::gatherLinesKernel<<<kBlocks, kThreads>>>(static_cast<const uint4*>(d_lines), static_cast<uint32_t>(line_count),
                                                    static_cast<uint64_t>(gathers), whole_lines, out);
    const hipError_t launched = hipGetLastError();
    const hipError_t done = hipDeviceSynchronize();
    (void)hipFree(out);
    SADCS_HIP(launched);
    SADCS_HIP(done);
    return SADCS_OK;
}

void* sadcs_hip_alloc_pinned(size_t bytes) {
    void* pointer = nullptr;
    if (hipHostMalloc(&pointer, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return pointer;
}

void sadcs_hip_free_pinned(void* pointer) {
    if (pointer) {
        (void)hipHostFree(pointer);
    }
}

const char* sadcs_hip_last_error(void) {
    return tLastError;
}

int sadcs_hip_batch_build(sadcs_ctx* context, const uint8_t* texts, const int64_t* lengths, int32_t count, int32_t* sa, uint8_t* bwt, int32_t* primaries, sadcs_stats* stats) {
    tLastError[0] = '\0';
    if (stats) {
        *stats = sadcs_stats{};
    }
    if (count <= 0 || (!sa && !bwt)) {
        return SADCS_OK;
    }
    uint64_t total = 0;
    for (int32_t i = 0; i < count; i++) {
        if (lengths[i] < 1 || lengths[i] > INT32_MAX) {
            setError("text %d has length %lld; batched texts need 1 to 2^31-1 bytes", i, static_cast<long long>(lengths[i]));
            return SADCS_FAILED;
        }
        total += static_cast<uint64_t>(lengths[i]) + kSegmentAlign;
    }
    if (total >= (uint64_t(1) << 31)) {
        setError("a batch holds less than 2 GiB of text");
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    Params params = context ? context->params : Params{};
    const Layout layout = layoutFor(params, lengths, count, true);
    BuildRequest rq;
    rq.mode = sa && bwt ? Mode::Both : sa ? Mode::SA : Mode::BWT;
    rq.hostText = texts;
    rq.hostLengths = lengths;
    rq.hostSA = sa;
    rq.hostBWT = bwt;
    std::vector<int32_t> first(static_cast<size_t>(count), -1);
    rq.hostPrimaries = first.data();
    const int status = buildOn(context, params, layout, rq, stats);
    if (status == SADCS_OK && primaries && bwt) {
        std::copy(first.begin(), first.end(), primaries);
    }
    return status;
}

int64_t sadcs_hip_batch_layout(const int64_t* lengths, int32_t count, int64_t* offsets) {
    int64_t start = 0, span = 0;
    for (int32_t i = 0; i < count; i++) {
        if (offsets) {
            offsets[i] = start;
        }
        span = start + lengths[i];
        start = (span + kSegmentAlign - 1) / kSegmentAlign * kSegmentAlign;
    }
    return span;
}

int sadcs_hip_batch_build_device(sadcs_ctx* context, const void* d_texts, const int64_t* lengths, int32_t count, void* d_sa, void* d_bwt, int32_t* primaries, sadcs_stats* stats) {
    tLastError[0] = '\0';
    if (stats) {
        *stats = sadcs_stats{};
    }
    if (count <= 0 || (!d_sa && !d_bwt)) {
        return SADCS_OK;
    }
    if (!d_texts || (reinterpret_cast<uintptr_t>(d_texts) & 15) != 0) {
        setError("the device texts must be 16-byte aligned");
        return SADCS_FAILED;
    }
    if ((reinterpret_cast<uintptr_t>(d_sa) & 3) != 0) {
        setError("the device suffix array must be 4-byte aligned");
        return SADCS_FAILED;
    }
    uint64_t total = 0;
    for (int32_t i = 0; i < count; i++) {
        if (lengths[i] < 1 || lengths[i] > INT32_MAX) {
            setError("text %d has length %lld; batched texts need 1 to 2^31-1 bytes", i, static_cast<long long>(lengths[i]));
            return SADCS_FAILED;
        }
        total += static_cast<uint64_t>(lengths[i]) + kSegmentAlign;
    }
    if (total >= (uint64_t(1) << 31)) {
        setError("a batch holds less than 2 GiB of text");
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    Params params = context ? context->params : Params{};
    const Layout layout = layoutFor(params, lengths, count, true);
    BuildRequest rq;
    rq.mode = d_sa && d_bwt ? Mode::Both : d_sa ? Mode::SA : Mode::BWT;
    rq.deviceText = static_cast<const uint8_t*>(d_texts);
    rq.deviceSA = static_cast<uint32_t*>(d_sa);
    rq.deviceBWT = static_cast<uint8_t*>(d_bwt);
    std::vector<int32_t> first(static_cast<size_t>(count), -1);
    rq.hostPrimaries = first.data();
    const int status = buildOn(context, params, layout, rq, stats);
    if (status == SADCS_OK && primaries && d_bwt) {
        std::copy(first.begin(), first.end(), primaries);
    }
    return status;
}

namespace {
// The device check's three passes with a caller-provided inverse (n words) and result words
// (kCheckCount), on the null stream, then the verdict. The inverse and result words are cleared
// here with a fill kernel.
int runDeviceCheck(const void* d_text, int32_t n, const void* d_sa, const void* d_bwt, int32_t primary,
                   int32_t* inverse, uint32_t* first, int64_t* result) {
    const auto* text = static_cast<const uint8_t*>(d_text);
    const auto* sa = static_cast<const int32_t*>(d_sa);
    const uint32_t blocks = std::min<uint32_t>(kCheckBlocks, (static_cast<uint32_t>(n) + kCheckThreads - 1) / kCheckThreads);
    fillWordsKernel<<<blocks, kCheckThreads>>>(reinterpret_cast<uint32_t*>(inverse), static_cast<uint64_t>(n), ~0u);
    fillWordsKernel<<<1, kCheckThreads>>>(first, kCheckCount, ~0u);
    checkRangeKernel<<<blocks, kCheckThreads>>>(sa, static_cast<uint32_t>(n), inverse, first);
    checkOrderKernel<<<blocks, kCheckThreads>>>(text, sa, static_cast<uint32_t>(n), inverse, first);
    if (d_bwt) {
        checkBwtKernel<<<blocks, kCheckThreads>>>(text, sa, static_cast<uint32_t>(n), inverse, static_cast<const uint8_t*>(d_bwt), first);
    }
    SADCS_HIP(hipGetLastError());
    uint32_t found[kCheckCount];
    SADCS_HIP(hipMemcpy(found, first, sizeof found, hipMemcpyDeviceToHost));
    // The first failing check in the reference's order: range, duplicates, order; then the BWT.
    const int32_t categories[] = {SADCS_CHECK_RANGE, SADCS_CHECK_DUPLICATE, SADCS_CHECK_ORDER};
    for (int c = 0; c < 3; c++) {
        if (found[c] != UINT32_MAX) {
            result[0] = categories[c];
            result[1] = found[c];
            return SADCS_OK;
        }
    }
    if (d_bwt) {
        if (static_cast<int64_t>(primary) != static_cast<int64_t>(found[kCheckRow0]) + 1) {
            result[0] = SADCS_CHECK_PRIMARY;
        } else if (found[kCheckBwt] != UINT32_MAX) {
            result[0] = SADCS_CHECK_BWT;
            result[1] = found[kCheckBwt];
        }
    }
    return SADCS_OK;
}

// Argument checks shared by both entry points; returns -1 to go on, else the status to return.
int checkArguments(const void* d_text, int32_t n, const void* d_sa, int64_t* result) {
    tLastError[0] = '\0';
    if (!result) {
        setError("invalid arguments");
        return SADCS_FAILED;
    }
    result[0] = SADCS_CHECK_PASS;
    result[1] = -1;
    if (n <= 0) {
        return SADCS_OK;
    }
    if (!d_text || !d_sa || (reinterpret_cast<uintptr_t>(d_sa) & 3) != 0) {
        setError("the device text and a 4-byte aligned device suffix array are needed");
        return SADCS_FAILED;
    }
    return -1;
}

// The allocation-free check's scratch: the inverse of up to kNoAllocCheckMax entries, then the
// result words. A file-scope device array: every device the library runs on holds its own copy
// (32 MiB + 256 B) from module load until the process ends. Host code never passes the array's
// name to a kernel (in host code it is a stand-in, not a device address): each device's address
// comes from hipGetSymbolAddress, taken once per device with that device current.
constexpr uint32_t kNoAllocCheckMax = 1u << 23;
constexpr uint32_t kNoAllocScratchWords = kNoAllocCheckMax + 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
uint32_t* gScratchAddress[kMaxDevices] = {};

// This device's address of `scratch`, looked up on the first call with the device current.
int noAllocScratch(uint32_t*& out) {
    int device = 0;
    SADCS_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) {
        setError("device %d is beyond the %d devices the allocation-free check tracks", device, kMaxDevices);
        return SADCS_FAILED;
    }
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchAddress[device]) {
        void* address = nullptr;
        const hipError_t e = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !address) {
            setError("the allocation-free check's scratch on device %d: %s", device, hipGetErrorString(e));
            return SADCS_FAILED;
        }
        gScratchAddress[device] = static_cast<uint32_t*>(address);
    }
    out = gScratchAddress[device];
    return SADCS_OK;
}
}  // namespace

int sadcs_hip_check_device(const void* d_text, int32_t n, const void* d_sa, const void* d_bwt, int32_t primary, int64_t* result) {
    const int early = checkArguments(d_text, n, d_sa, result);
    if (early >= 0) {
        return early;
    }
    // inverse: n int32, then the first failing index of each check.
    const size_t inverseBytes = (static_cast<size_t>(n) * sizeof(int32_t) + 255) / 256 * 256;
    void* memory = nullptr;
    SADCS_HIP(hipMalloc(&memory, inverseBytes + 256));
    std::unique_ptr<void, void (*)(void*)> owner(memory, [](void* m) { (void)hipFree(m); });
    return runDeviceCheck(d_text, n, d_sa, d_bwt, primary, static_cast<int32_t*>(memory),
                          reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(memory) + inverseBytes), result);
}

int sadcs_hip_check_device_noalloc(const void* d_text, int32_t n, const void* d_sa, const void* d_bwt, int32_t primary, int64_t* result) {
    const int early = checkArguments(d_text, n, d_sa, result);
    if (early >= 0) {
        return early;
    }
    if (static_cast<uint32_t>(n) > kNoAllocCheckMax) {
        setError("the allocation-free check takes at most %u entries (got %d); use sadcs_hip_check_device", kNoAllocCheckMax, n);
        return SADCS_NOT_ENOUGH_MEMORY;
    }
    uint32_t* base = nullptr;
    SADCS_TRY(noAllocScratch(base));
    return runDeviceCheck(d_text, n, d_sa, d_bwt, primary, reinterpret_cast<int32_t*>(base), base + kNoAllocCheckMax, result);
}

int64_t sadcs_hip_batch_required_bytes(const int64_t* lengths, int32_t count) {
    if (count <= 0) {
        return 0;
    }
    Params params;
    const Layout layout = layoutFor(params, lengths, count, true);
    sadcs_ctx planner;
    const size_t plan = planFor(planner, layout, Mode::Both, true, true);
    return plan == SIZE_MAX ? -1 : static_cast<int64_t>(plan);
}

}  // extern "C"

// ==== serving ====================================================================================
//
// A request loop with continuous batching: submitted texts wait in a FIFO; a dispatcher forms a
// batch from the front of the queue whenever a worker (one stream each) is idle and the batch's
// planned device memory (the dry-run plan, about 10 bytes per input byte) fits in the free part of
// the server's HBM pool. Batches of different workers run concurrently on the device; a batch's
// texts and sample ranks stay resident from upload to the end of its merge. Requests that arrive
// while others run are admitted as soon as a worker and memory free up.

#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace sadcs {

// A host array left uninitialised (results are copied over it whole).
template <class T>
struct HostArray {
    std::unique_ptr<T[]> data;
    size_t size = 0;

    void allocate(size_t n) {
        data.reset(n ? new T[n] : nullptr);
        size = n;
    }
};

// Queued texts wait in pinned host memory, so that a batch uploads each with one DMA: no copy into
// the staging chunks and no wait for a chunk between texts. Slabs of kSlab bytes are carved in
// arrival order and reused once every text in them is done; a text larger than a slab, or one
// arriving while kMaxSlabs slabs are all in use, waits in pageable memory instead.
class TextPool {
 public:
    struct Block {
        uint8_t* data = nullptr;
        int slab = -1;
    };

    ~TextPool() {
        for (Slab& s : m_slabs) {
            (void)hipHostFree(s.base);
        }
    }

    // Pins `count` slabs now (pinning is slow: better at start than while texts arrive).
    void reserve(size_t count) {
        std::lock_guard<std::mutex> lock(m_mutex);
        while (m_slabs.size() < std::min(count, kMaxSlabs)) {
            uint8_t* base = nullptr;
            if (hipHostMalloc(reinterpret_cast<void**>(&base), kSlab, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                return;
            }
            m_slabs.push_back(Slab{base, 0, 0});
        }
    }

    Block allocate(size_t n) {
        const size_t need = (n + 63) / 64 * 64;
        if (n == 0 || need > kSlab) {
            return {};
        }
        std::lock_guard<std::mutex> lock(m_mutex);
        if (m_current < 0 || m_slabs[m_current].used + need > kSlab) {
            m_current = -1;
            for (size_t i = 0; i < m_slabs.size(); i++) {
                if (m_slabs[i].live == 0) {
                    m_slabs[i].used = 0;
                    m_current = static_cast<int>(i);
                    break;
                }
            }
            if (m_current < 0) {
                uint8_t* base = nullptr;
                if (m_slabs.size() >= kMaxSlabs || hipHostMalloc(reinterpret_cast<void**>(&base), kSlab, hipHostMallocDefault) != hipSuccess) {
                    (void)hipGetLastError();
                    return {};
                }
                m_slabs.push_back(Slab{base, 0, 0});
                m_current = static_cast<int>(m_slabs.size()) - 1;
            }
        }
        Slab& s = m_slabs[m_current];
        Block b{s.base + s.used, m_current};
        s.used += need;
        s.live++;
        return b;
    }

    void release(Block& b) {
        if (!b.data) {
            return;
        }
        std::lock_guard<std::mutex> lock(m_mutex);
        m_slabs[b.slab].live--;
        b = {};
    }

 private:
    static constexpr size_t kSlab = size_t(64) << 20;
    static constexpr size_t kMaxSlabs = 16;
    struct Slab {
        uint8_t* base;
        size_t used;
        int live;
    };
    std::mutex m_mutex;
    std::vector<Slab> m_slabs;
    int m_current = -1;
};

// A queued text: in a pinned block of the server's pool, or in pageable memory.
struct RequestText {
    HostArray<uint8_t> pageable;
    TextPool::Block pinned;
    size_t size = 0;

    const uint8_t* ptr() const {
        return pinned.data ? pinned.data : pageable.data.get();
    }
};

struct ServerRequest {
    int64_t id = 0;
    RequestText text;
    HostArray<int32_t> sa;
    HostArray<uint8_t> bwt;
    int32_t* outSA = nullptr;   // the caller's buffers, written at completion (submit_into);
    uint8_t* outBWT = nullptr;  //   otherwise the outputs are held here until wait copies them
    int32_t primary = -1;
    int status = SADCS_OK;
    std::string error;
    std::chrono::steady_clock::time_point arrival, admitted, done;
    bool finished = false;
    int64_t batchId = -1;
    int32_t batchSize = 0;
};

class Server {
 public:
    Server(const sadcs_server_params& p) {
        m_params.cover = p.cover ? p.cover : 32;
        Params tmp;
        sadcs_params sp{};
        sp.cover = m_params.cover;
        tmp.apply(&sp);
        m_params = tmp;
        m_budget = static_cast<size_t>(p.memory_budget);
        m_workerCount = p.workers > 0 ? p.workers : 5;
        m_maxBatchBytes = p.max_batch_bytes > 0 ? static_cast<uint64_t>(p.max_batch_bytes) : (uint64_t(32) << 20);
        m_maxBatchTexts = p.max_batch_texts > 0 ? p.max_batch_texts : 4096;
        m_wantSA = p.want_sa != 0;
        m_wantBWT = p.want_bwt != 0;
    }

    ~Server() {
        stop();
    }

    int start() {
        sadcs_device device;
        SADCS_TRY(queryDevice(device));
        if (m_budget == 0) {
            m_budget = static_cast<size_t>(std::max<int64_t>(0, device.free_bytes - kMemoryReserve) / 5 * 4);
        }
        if (static_cast<int64_t>(m_budget) + kMemoryReserve > device.free_bytes) {
            setError("server budget of %zu MiB plus the reserve exceeds the free device memory", m_budget >> 20);
            return SADCS_NOT_ENOUGH_MEMORY;
        }
        SADCS_TRY(ensureTables());
        SADCS_HIP(hipMalloc(&m_poolBase, m_budget));
        m_pool.reset(static_cast<uint8_t*>(m_poolBase), m_budget, false);
        for (int w = 0; w < m_workerCount; w++) {
            sadcs_ctx* ctx = nullptr;
            const int s = createContext(0, m_params, INT32_MAX, ctx);
            if (s != SADCS_OK) {
                return s;
            }
            // Worker contexts borrow their arena from the pool for each batch.
            (void)hipFree(ctx->arenaBase);
            ctx->arenaBase = nullptr;
            ctx->arenaBytes = 0;
            m_workers.push_back(std::make_unique<Worker>());
            m_workers.back()->ctx = ctx;
        }
        // Host memory is pinned now rather than while requests arrive: text slabs for the queue, and
        // each worker's output buffer for a full batch (grown later only for a larger lone text).
        m_textPool.reserve(kReservedTextSlabs);
        const size_t outputBytes = static_cast<size_t>(m_maxBatchBytes) * ((m_wantSA ? sizeof(int32_t) : 0) + (m_wantBWT ? 1 : 0));
        for (auto& w : m_workers) {
            if (outputBytes && hipHostMalloc(reinterpret_cast<void**>(&w->pinnedOut), outputBytes, hipHostMallocDefault) == hipSuccess) {
                w->pinnedBytes = outputBytes;
            } else {
                (void)hipGetLastError();
                w->pinnedOut = nullptr;
            }
        }
        for (int w = 0; w < m_workerCount; w++) {
            m_workers[w]->thread = std::thread(&Server::workerLoop, this, w);
        }
        m_dispatcher = std::thread(&Server::dispatchLoop, this);
        return SADCS_OK;
    }

    void stop() {
        {
            std::lock_guard<std::mutex> lock(m_mutex);
            if (m_stopped) {
                return;
            }
            m_stopping = true;
        }
        m_wake.notify_all();
        if (m_dispatcher.joinable()) {
            m_dispatcher.join();
        }
        for (auto& w : m_workers) {
            if (w->thread.joinable()) {
                w->thread.join();
            }
            w->ctx->arenaBase = nullptr;  // never free a pool chunk as an arena
            destroyContext(w->ctx);
            if (w->pinnedOut) {
                (void)hipHostFree(w->pinnedOut);
                w->pinnedOut = nullptr;
            }
        }
        m_workers.clear();
        if (m_poolBase) {
            (void)hipFree(m_poolBase);
            m_poolBase = nullptr;
        }
        std::lock_guard<std::mutex> lock(m_mutex);
        m_stopped = true;
    }

    int64_t submit(const uint8_t* text, int64_t n, int32_t* outSA = nullptr, uint8_t* outBWT = nullptr) {
        auto r = std::make_shared<ServerRequest>();
        r->text.size = static_cast<size_t>(n);
        if (n > 1) {
            r->text.pinned = m_textPool.allocate(static_cast<size_t>(n));
        }
        if (!r->text.pinned.data) {
            r->text.pageable.allocate(static_cast<size_t>(n));
        }
        // A text below a few MiB is copied by this thread: the copy pool is shared with the workers'
        // output copies, and waiting for its turn would hold up the submitting thread.
        if (n >= kPooledTextCopy) {
            HostCopier::instance().copy(const_cast<uint8_t*>(r->text.ptr()), text, static_cast<size_t>(n));
        } else if (n > 0) {
            memcpy(const_cast<uint8_t*>(r->text.ptr()), text, static_cast<size_t>(n));
        }
        r->outSA = m_wantSA ? outSA : nullptr;
        r->outBWT = m_wantBWT ? outBWT : nullptr;
        r->arrival = std::chrono::steady_clock::now();
        std::lock_guard<std::mutex> lock(m_mutex);
        r->id = m_nextId++;
        m_requests[r->id] = r;
        m_pending.push_back(r);
        m_wake.notify_all();
        return r->id;
    }

    int wait(int64_t id, int32_t* sa, uint8_t* bwt, int32_t* primary, double* latency, double* queued, int32_t* batchSize) {
        std::shared_ptr<ServerRequest> r;
        {
            std::unique_lock<std::mutex> lock(m_mutex);
            auto it = m_requests.find(id);
            if (it == m_requests.end()) {
                setError("unknown request %lld", static_cast<long long>(id));
                return SADCS_FAILED;
            }
            r = it->second;
            m_done.wait(lock, [&] { return r->finished; });
            m_requests.erase(id);
        }
        if (r->status != SADCS_OK) {
            setError("%s", r->error.c_str());
            return r->status;
        }
        if (sa && r->sa.size) {
            HostCopier::instance().copy(sa, r->sa.data.get(), r->sa.size * sizeof(int32_t));
        }
        if (bwt && r->bwt.size) {
            HostCopier::instance().copy(bwt, r->bwt.data.get(), r->bwt.size);
        }
        if (primary) {
            *primary = r->primary;
        }
        if (latency) {
            *latency = std::chrono::duration<double>(r->done - r->arrival).count();
        }
        if (queued) {
            *queued = std::chrono::duration<double>(r->admitted - r->arrival).count();
        }
        if (batchSize) {
            *batchSize = r->batchSize;
        }
        return SADCS_OK;
    }

    void counters(int64_t* out) {
        std::lock_guard<std::mutex> lock(m_mutex);
        out[0] = m_completed;
        out[1] = m_batches;
        out[2] = static_cast<int64_t>(m_peakPool);
        out[3] = static_cast<int64_t>(m_budget);
        out[4] = m_maxConcurrent;
    }

 private:
    struct Worker {
        sadcs_ctx* ctx = nullptr;
        uint8_t* pinnedOut = nullptr;  // batch outputs, downloaded without staging
        size_t pinnedBytes = 0;
        std::thread thread;
        std::vector<std::shared_ptr<ServerRequest>> batch;
        uint8_t* chunk = nullptr;
        size_t bytes = 0;
        bool busy = false;
        int64_t batchId = 0;
    };

    Mode mode() const {
        return m_wantSA && m_wantBWT ? Mode::Both : m_wantSA ? Mode::SA : Mode::BWT;
    }

    size_t planOf(const std::vector<std::shared_ptr<ServerRequest>>& batch) {
        std::vector<int64_t> lengths;
        for (auto& r : batch) {
            lengths.push_back(static_cast<int64_t>(r->text.size));
        }
        const Layout layout = layoutFor(m_params, lengths.data(), static_cast<int32_t>(lengths.size()), batch.size() > 1);
        sadcs_ctx planner;
        planner.params = m_params;
        return planFor(planner, layout, mode(), true, true);
    }

    void releaseText(ServerRequest& r) {
        m_textPool.release(r.text.pinned);
        r.text.pageable.allocate(0);
    }

    void finish(const std::shared_ptr<ServerRequest>& r, int status, const char* error) {
        releaseText(*r);
        r->status = status;
        r->error = error ? error : "";
        r->done = std::chrono::steady_clock::now();
        r->finished = true;
        m_completed++;
    }

    // With the lock held: the batch to run next, or empty.
    // A batch from the front of the queue with its pool chunk, or an empty batch (nothing admissible
    // now). tooLarge is set when the front request alone can never fit the budget.
    std::vector<std::shared_ptr<ServerRequest>> formBatch(uint8_t*& chunk, size_t& bytes, bool& tooLarge) {
        std::vector<std::shared_ptr<ServerRequest>> batch;
        tooLarge = false;
        // A deep queue is split so that every worker gets a share (about half of pending / workers,
        // at least kMinBatchBytes): one worker's outputs are copied to their callers while the
        // others' batches run on the device, instead of a few large batches running one after
        // another with all their copies at the end.
        uint64_t pendingBytes = 0;
        for (auto& r : m_pending) {
            pendingBytes += r->text.size;
        }
        const uint64_t share = pendingBytes / (2 * static_cast<uint64_t>(m_workerCount));
        const uint64_t limit = std::min(m_maxBatchBytes, std::max(kMinBatchBytes, share));
        uint64_t total = 0;
        for (auto& r : m_pending) {
            if (r->text.size <= 1) {
                continue;  // answered by the dispatcher directly
            }
            if (!batch.empty() && (total + r->text.size > limit || static_cast<int>(batch.size()) >= m_maxBatchTexts)) {
                break;
            }
            batch.push_back(r);
            total += r->text.size + kSegmentAlign;
        }
        while (!batch.empty()) {
            const size_t plan = planOf(batch);
            if (plan == SIZE_MAX) {
                batch.clear();
                break;
            }
            if (plan > m_budget && batch.size() == 1) {
                tooLarge = true;  // can never fit: refused by the dispatcher
                batch.clear();
                break;
            }
            uint8_t* p = nullptr;
            if (plan <= m_budget && m_pool.allocate(p, plan) == SADCS_OK) {
                chunk = p;
                bytes = plan;
                return batch;
            }
            if (batch.size() == 1) {
                batch.clear();  // waits for memory
                break;
            }
            batch.resize(batch.size() / 2);  // admit fewer
        }
        return batch;
    }

    void dispatchLoop() {
        std::unique_lock<std::mutex> lock(m_mutex);
        while (true) {
            // Texts of length 0 or 1 need no device work.
            for (auto it = m_pending.begin(); it != m_pending.end();) {
                auto& r = *it;
                if (r->text.size <= 1) {
                    r->admitted = std::chrono::steady_clock::now();
                    if (r->text.size == 1) {
                        r->sa.allocate(m_wantSA && !r->outSA ? 1 : 0);
                        r->bwt.allocate(m_wantBWT && !r->outBWT ? 1 : 0);
                        if (m_wantSA) {
                            (r->outSA ? r->outSA : r->sa.data.get())[0] = 0;
                        }
                        if (m_wantBWT) {
                            (r->outBWT ? r->outBWT : r->bwt.data.get())[0] = r->text.ptr()[0];
                        }
                        r->primary = 1;
                    } else {
                        r->primary = 0;
                    }
                    finish(r, SADCS_OK, nullptr);
                    it = m_pending.erase(it);
                    m_done.notify_all();
                } else {
                    ++it;
                }
            }
            if (m_stopping && m_pending.empty()) {
                break;
            }
            int idle = -1;
            for (int w = 0; w < m_workerCount; w++) {
                if (!m_workers[w]->busy) {
                    idle = w;
                    break;
                }
            }
            if (idle >= 0 && !m_pending.empty()) {
                uint8_t* chunk = nullptr;
                size_t bytes = 0;
                bool tooLarge = false;
                auto batch = formBatch(chunk, bytes, tooLarge);
                if (!batch.empty()) {
                    for (auto& r : batch) {
                        r->admitted = std::chrono::steady_clock::now();
                        r->batchId = m_batches;
                        r->batchSize = static_cast<int32_t>(batch.size());
                    }
                    m_pending.erase(m_pending.begin(), m_pending.begin() + static_cast<std::ptrdiff_t>(batch.size()));
                    Worker& w = *m_workers[idle];
                    w.batch = std::move(batch);
                    w.chunk = chunk;
                    w.bytes = bytes;
                    w.busy = true;
                    w.batchId = m_batches++;
                    m_poolInUse += bytes;
                    m_peakPool = std::max(m_peakPool, m_poolInUse);
                    int busy = 0;
                    for (auto& x : m_workers) {
                        busy += x->busy ? 1 : 0;
                    }
                    m_maxConcurrent = std::max<int64_t>(m_maxConcurrent, busy);
                    m_wake.notify_all();
                    continue;
                }
                // Nothing admissible: the head request never fits, or memory is held by running batches.
                bool anyBusy = false;
                for (auto& x : m_workers) {
                    anyBusy = anyBusy || x->busy;
                }
                if (tooLarge || !anyBusy) {
                    auto r = m_pending.front();
                    m_pending.pop_front();
                    r->admitted = std::chrono::steady_clock::now();
                    finish(r, SADCS_NOT_ENOUGH_MEMORY, "the text's planned device memory exceeds the server's budget");
                    m_done.notify_all();
                    continue;
                }
            }
            m_wake.wait(lock);
        }
        for (auto& w : m_workers) {
            w->busy = w->busy;  // workers exit once idle and stopping
        }
        m_shutdown = true;
        m_wake.notify_all();
    }

    void workerLoop(int index) {
        Worker& w = *m_workers[index];
        std::unique_lock<std::mutex> lock(m_mutex);
        while (true) {
            m_wake.wait(lock, [&] { return (w.busy && !w.batch.empty()) || m_shutdown; });
            if (!(w.busy && !w.batch.empty())) {
                break;
            }
            auto batch = w.batch;
            lock.unlock();
            runBatch(w, batch);
            lock.lock();
            for (auto& r : batch) {
                r->done = std::chrono::steady_clock::now();
                r->finished = true;
                m_completed++;
            }
            m_pool.release(w.chunk);
            m_poolInUse -= w.bytes;
            w.batch.clear();
            w.busy = false;
            m_done.notify_all();
            m_wake.notify_all();
        }
    }

    void runBatch(Worker& w, std::vector<std::shared_ptr<ServerRequest>>& batch) {
        std::vector<int64_t> lengths;
        std::vector<const uint8_t*> texts;
        uint64_t total = 0;
        for (auto& r : batch) {
            lengths.push_back(static_cast<int64_t>(r->text.size));
            texts.push_back(r->text.ptr());
            total += r->text.size;
        }
        const Layout layout = layoutFor(m_params, lengths.data(), static_cast<int32_t>(lengths.size()), batch.size() > 1);
        // Outputs land in the worker's pinned buffer (one DMA each, no staging copies), grown to
        // the largest batch; without it, in an uninitialised pageable buffer.
        const size_t saBytes = m_wantSA ? total * sizeof(int32_t) : 0, bwtBytes = m_wantBWT ? total : 0;
        if (w.pinnedBytes < saBytes + bwtBytes) {
            if (w.pinnedOut) {
                (void)hipHostFree(w.pinnedOut);
            }
            w.pinnedOut = nullptr;
            w.pinnedBytes = 0;
            if (hipHostMalloc(reinterpret_cast<void**>(&w.pinnedOut), saBytes + bwtBytes, hipHostMallocDefault) == hipSuccess) {
                w.pinnedBytes = saBytes + bwtBytes;
            } else {
                (void)hipGetLastError();
                w.pinnedOut = nullptr;
            }
        }
        HostArray<uint8_t> pageable;
        uint8_t* out = w.pinnedOut;
        if (!out) {
            pageable.allocate(saBytes + bwtBytes);
            out = pageable.data.get();
        }
        int32_t* sa = reinterpret_cast<int32_t*>(out);
        uint8_t* bwt = out + saBytes;
        std::vector<int32_t> primaries(batch.size(), -1);
        BuildRequest rq;
        rq.mode = mode();
        rq.hostText = texts[0];  // a batch of one is built as a single text
        rq.hostTexts = texts.data();
        rq.hostLengths = lengths.data();
        rq.hostSA = m_wantSA ? sa : nullptr;
        rq.hostBWT = m_wantBWT ? bwt : nullptr;
        rq.hostPrimaries = primaries.data();
        w.ctx->arenaBase = w.chunk;
        w.ctx->arenaBytes = w.bytes;
        w.ctx->capacity = INT32_MAX;
        sadcs_stats stats{};
        const auto batchStart = std::chrono::steady_clock::now();
        const int status = build(*w.ctx, layout, rq, &stats);
        w.ctx->arenaBase = nullptr;  // the chunk belongs to the pool
        w.ctx->arenaBytes = 0;
        const std::string error = status == SADCS_OK ? "" : tLastError;
        const auto copyStart = std::chrono::steady_clock::now();
        // Every request's outputs leave the batch buffer in one parallel copy.
        std::vector<HostCopier::Piece> pieces;
        pieces.reserve(2 * batch.size());
        uint64_t offset = 0;
        for (size_t i = 0; i < batch.size(); i++) {
            auto& r = batch[i];
            const size_t n = r->text.size;
            r->status = status;
            r->error = error;
            if (status == SADCS_OK) {
                if (m_wantSA) {
                    if (!r->outSA) {
                        r->sa.allocate(n);
                    }
                    pieces.push_back({r->outSA ? r->outSA : r->sa.data.get(), sa + offset, n * sizeof(int32_t)});
                }
                if (m_wantBWT) {
                    if (!r->outBWT) {
                        r->bwt.allocate(n);
                    }
                    pieces.push_back({r->outBWT ? r->outBWT : r->bwt.data.get(), bwt + offset, n});
                    r->primary = primaries[i];
                }
            }
            offset += n;
        }
        HostCopier::instance().copyPieces(pieces.data(), pieces.size());
        for (auto& r : batch) {
            releaseText(*r);
        }
        std::lock_guard<std::mutex> lock(m_statsMutex);
        m_stats.batches++;
        m_stats.texts += static_cast<int64_t>(batch.size());
        m_stats.bytes += static_cast<int64_t>(total);
        m_stats.upload_seconds += stats.upload_seconds;
        m_stats.device_seconds += stats.device_seconds;
        m_stats.download_seconds += stats.download_seconds;
        m_stats.copy_out_seconds += secondsSince(copyStart);
        m_stats.batch_seconds += secondsSince(batchStart);
    }

 public:
    void statsInto(sadcs_server_stats* out) {
        std::lock_guard<std::mutex> lock(m_statsMutex);
        *out = m_stats;
    }

 private:

    Params m_params;
    size_t m_budget = 0;
    int m_workerCount = 3;
    uint64_t m_maxBatchBytes = 0;
    static constexpr uint64_t kMinBatchBytes = uint64_t(4) << 20;
    int m_maxBatchTexts = 4096;
    bool m_wantSA = true, m_wantBWT = true;
    void* m_poolBase = nullptr;
    Arena m_pool;
    size_t m_poolInUse = 0, m_peakPool = 0;
    std::mutex m_mutex;
    std::condition_variable m_wake, m_done;
    std::deque<std::shared_ptr<ServerRequest>> m_pending;
    std::unordered_map<int64_t, std::shared_ptr<ServerRequest>> m_requests;
    TextPool m_textPool;  // queued texts, pinned
    static constexpr int64_t kPooledTextCopy = int64_t(8) << 20;
    static constexpr size_t kReservedTextSlabs = 4;
    std::vector<std::unique_ptr<Worker>> m_workers;
    std::thread m_dispatcher;
    int64_t m_nextId = 0, m_completed = 0, m_batches = 0, m_maxConcurrent = 0;
    bool m_stopping = false, m_shutdown = false, m_stopped = false;
    std::mutex m_statsMutex;
    sadcs_server_stats m_stats{};
};

}  // namespace sadcs

struct sadcs_server {
    std::unique_ptr<sadcs::Server> server;
};

extern "C" {

int sadcs_hip_server_create(const sadcs_server_params* params, sadcs_server** out) {
    tLastError[0] = '\0';
    if (!out) {
        setError("null server pointer");
        return SADCS_FAILED;
    }
    *out = nullptr;
    sadcs_server_params p{};
    if (params) {
        p = *params;
    } else {
        p.want_sa = 1;
        p.want_bwt = 1;
    }
    if (p.cover != 0 && p.cover != 32 && p.cover != 128) {
        setError("cover must be 32 or 128, not %d", p.cover);
        return SADCS_FAILED;
    }
    std::unique_ptr<sadcs_server> s(new sadcs_server);
    s->server.reset(new Server(p));
    const int status = s->server->start();
    if (status != SADCS_OK) {
        return status;
    }
    *out = s.release();
    return SADCS_OK;
}

int64_t sadcs_hip_server_submit(sadcs_server* server, const uint8_t* text, int64_t n) {
    if (!server || n < 0 || n > INT32_MAX - 1024 || (n > 0 && !text)) {
        setError("invalid submission");
        return -SADCS_FAILED;
    }
    return server->server->submit(text, n);
}

int64_t sadcs_hip_server_submit_into(sadcs_server* server, const uint8_t* text, int64_t n, int32_t* sa, uint8_t* bwt) {
    if (!server || n < 0 || n > INT32_MAX - 1024 || (n > 0 && !text)) {
        setError("invalid submission");
        return -SADCS_FAILED;
    }
    return server->server->submit(text, n, sa, bwt);
}

int sadcs_hip_server_wait(sadcs_server* server, int64_t id, int32_t* sa, uint8_t* bwt, int32_t* primary, double* latency_seconds,
                          double* queued_seconds, int32_t* batch_size) {
    if (!server) {
        setError("null server");
        return SADCS_FAILED;
    }
    return server->server->wait(id, sa, bwt, primary, latency_seconds, queued_seconds, batch_size);
}

int sadcs_hip_server_stats(sadcs_server* server, sadcs_server_stats* stats) {
    if (!server || !stats) {
        setError("invalid arguments");
        return SADCS_FAILED;
    }
    server->server->statsInto(stats);
    return SADCS_OK;
}

void sadcs_hip_server_counters(sadcs_server* server, int64_t* counters) {
    if (server && counters) {
        server->server->counters(counters);
    }
}

void sadcs_hip_server_destroy(sadcs_server* server) {
    if (server) {
        server->server->stop();
        delete server;
    }
}

}  // extern "C"
