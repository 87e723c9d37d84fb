/*
 * merge.h - Phase (c): one multiway merge of the sorted class lists into the suffix array, with
 * the BWT emitted by the same kernel.
 *
 * Suffixes i and j compare by k = the smallest offset with i + k and j + k both sampled: the
 * first k characters, then the sample ranks at i + k and j + k; if one suffix ends first it is
 * the smaller (src/cuda/merge.cuh:53-71, :102-128). The output is cut into tiles at splitters
 * (every G-th element of every list, ordered by this same merge one level down;
 * src/cuda/merge.cuh:1-9, :133-172, :718-789), and each tile is merged in LDS.
 *
 * The design for gfx950:
 *  - A tile spans K consecutive sorted splitters, so its size is at most K*G + C*(G-1)
 *    elements for C lists: one fixed LDS capacity always suffices and there is no overflow path.
 *    Exact list positions are searched only at tile boundaries, not for every splitter.
 *  - Inside a tile the C runs are merged pairwise in ceil(log2 C) rounds of merge paths over
 *    16-bit index arrays: about log2(C) comparisons per element instead of C - 1, and no atomics.
 *    Each record (16 characters, the V-window of sample ranks, the position) is loaded once.
 *  - In-place mode: the lists are first copied tile-major into the suffix-array buffer, so every
 *    tile reads and writes the same range. The class lists are then free before the merge and
 *    the BWT's n bytes fit in their place: SA plus BWT in one merge within about 10n bytes.
 *  - The row of each text's first suffix is found before the merge (C binary searches), so the
 *    BWT bytes land at their final index directly: byte 0 = T[n-1], no hole to close afterwards.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "cover.h"
#include "dev_common.h"

namespace sadcs {

constexpr int kMaxLists = 32;
constexpr int kRecordWords = 2;

template <int V>
__device__ __forceinline__ const CoverTables<V>& deviceCover();

template <int V>
__device__ __forceinline__ uint32_t packPosition(uint32_t p);

struct MergeCtx {
    const uint64_t* text;
    const uint32_t* ranks;
    const Segment* segments;
    uint32_t n;  // single text: its length
};

// The sorted lists of one merge level.
struct ListSet {
    const uint32_t* ptr[kMaxLists];
    uint32_t length[kMaxLists];
    uint32_t indexBase[kMaxLists + 1];  // prefix sums of the lengths
    int count;
};

template <int V, bool kBatch>
struct MergeRecord {
    uint64_t words[kRecordWords];
    uint32_t ranks[CoverSet<V>::kSize];
    uint32_t position;
    uint32_t tag;  // batch: segment << 9 | min(remaining, V)
};

template <int V>
struct MergeRecord<V, false> {
    uint64_t words[kRecordWords];
    uint32_t ranks[CoverSet<V>::kSize];
    uint32_t position;
};

template <int V, class Texts>
__device__ __forceinline__ void loadRecord(const MergeCtx& c, const Texts& texts, uint32_t p, MergeRecord<V, Texts::kBatch>& r) {
#pragma unroll
    for (int w = 0; w < kRecordWords; w++) {
        r.words[w] = load8(c.text, p + 8 * w);
    }
    const uint32_t base = packPosition<V>(p);
#pragma unroll
    for (int s = 0; s < CoverSet<V>::kSize; s++) {
        r.ranks[s] = c.ranks[base + s];
    }
    r.position = p;
    if constexpr (Texts::kBatch) {
        const uint32_t segment = texts.segmentOf(p);
        const uint32_t remaining = c.segments[segment].end - p;
        r.tag = (segment << 9) | min(remaining, static_cast<uint32_t>(V));
    }
}

template <int V, class Texts>
__device__ __forceinline__ uint32_t remainingOf(const MergeCtx& c, const MergeRecord<V, Texts::kBatch>& r) {
    if constexpr (Texts::kBatch) {
        return r.tag & 0x1FF;
    } else {
        return min(c.n - r.position, static_cast<uint32_t>(V));
    }
}

// Record a < record b.
template <int V, class Texts>
__device__ __forceinline__ bool lessRecord(const MergeCtx& c, const MergeRecord<V, Texts::kBatch>& a, const MergeRecord<V, Texts::kBatch>& b) {
    if constexpr (Texts::kBatch) {
        const uint32_t sa = a.tag >> 9, sb = b.tag >> 9;
        if (sa != sb) {
            return sa < sb;
        }
    }
    const CoverTables<V>& cv = deviceCover<V>();
    const uint32_t i = a.position, j = b.position;
    const uint32_t k = cv.tie[i & (V - 1)][j & (V - 1)];
    const uint32_t ra = remainingOf<V, Texts>(c, a), rb = remainingOf<V, Texts>(c, b);
    const uint32_t common = min(k, min(ra, rb));

#pragma unroll
    for (uint32_t w = 0; w < V / 8; w++) {
        if (8 * w >= common) {
            break;
        }
        const uint64_t mask = prefixMask(common - 8 * w);
        const uint64_t x = (w < kRecordWords ? a.words[w] : load8(c.text, i + 8 * w)) & mask;
        const uint64_t y = (w < kRecordWords ? b.words[w] : load8(c.text, j + 8 * w)) & mask;
        if (x != y) {
            return x < y;
        }
    }
    if (common < k || ra == k || rb == k) {
        return i > j;  // one is a prefix of the other: the shorter suffix is smaller
    }
    return a.ranks[cv.slot[i & (V - 1)][k]] < b.ranks[cv.slot[j & (V - 1)][k]];
}

// Suffix i (read from global memory) < the suffix of record b.
template <int V, class Texts>
__device__ __forceinline__ bool lessThanRecord(const MergeCtx& c, const Texts& texts, uint32_t i, const MergeRecord<V, Texts::kBatch>& b) {
    uint32_t endI;
    if constexpr (Texts::kBatch) {
        const uint32_t si = texts.segmentOf(i), sb = b.tag >> 9;
        if (si != sb) {
            return si < sb;
        }
        endI = c.segments[si].end;
    } else {
        endI = c.n;
    }
    const CoverTables<V>& cv = deviceCover<V>();
    const uint32_t j = b.position;
    const uint32_t k = cv.tie[i & (V - 1)][j & (V - 1)];
    const uint32_t ra = min(endI - i, static_cast<uint32_t>(V)), rb = remainingOf<V, Texts>(c, b);
    const uint32_t common = min(k, min(ra, rb));

#pragma unroll
    for (uint32_t w = 0; w < V / 8; w++) {
        if (8 * w >= common) {
            break;
        }
        const uint64_t mask = prefixMask(common - 8 * w);
        const uint64_t x = load8(c.text, i + 8 * w) & mask;
        const uint64_t y = (w < kRecordWords ? b.words[w] : load8(c.text, j + 8 * w)) & mask;
        if (x != y) {
            return x < y;
        }
    }
    if (common < k || ra == k || rb == k) {
        return i > j;
    }
    // The record lives in registers: select its rank without indexing it dynamically.
    const uint32_t slot = cv.slot[j & (V - 1)][k];
    uint32_t rankB = 0;
#pragma unroll
    for (int s = 0; s < CoverSet<V>::kSize; s++) {
        rankB = s == static_cast<int>(slot) ? b.ranks[s] : rankB;
    }
    return c.ranks[packPosition<V>(i + k)] < rankB;
}

__device__ __forceinline__ int listOfIndex(const ListSet& lists, uint32_t index) {
    int l = 0;
    while (l + 1 < lists.count && index >= lists.indexBase[l + 1]) {
        l++;
    }
    return l;
}

// ---- splitters ------------------------------------------------------------------------------

// out[x] = element k*G of list l, for the x-th splitter (splitter lists concatenated).
__global__ void gatherSplittersKernel(ListSet lists, ListSet splitters, uint32_t g, uint32_t total, uint32_t* __restrict__ out) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < total) {
        const int l = listOfIndex(splitters, x);
        out[x] = lists.ptr[l][static_cast<uint64_t>(x - splitters.indexBase[l]) * g];
    }
}

// counts[l * boundaries + b] = splitters of list l among sorted splitters [b*K, (b+1)*K).
__global__ void countBoundarySplittersKernel(ListSet splitters, const uint32_t* __restrict__ sorted, uint32_t total, uint32_t perTile, uint32_t boundaries, uint32_t* __restrict__ counts) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= boundaries) {
        return;
    }
    uint32_t local[kMaxLists];
    for (int l = 0; l < splitters.count; l++) {
        local[l] = 0;
    }
    const uint64_t first = static_cast<uint64_t>(b) * perTile;
    const uint64_t last = min(first + perTile, static_cast<uint64_t>(total));
    for (uint64_t x = first; x < last; x++) {
        local[listOfIndex(splitters, sorted[x])]++;
    }
    for (int l = 0; l < splitters.count; l++) {
        counts[static_cast<uint64_t>(l) * boundaries + b] = local[l];
    }
}

// Row b: for every list, the number of its elements smaller than sorted splitter b*K, then their
// sum (the tile's first output row). before[l * boundaries + b] + indexBase[l] = the l-splitters
// before it (an exclusive scan of the counts above). Row `boundaries` is the end: the lengths.
template <int V, class Texts>
__global__ void boundaryRowsKernel(MergeCtx c, Texts texts, ListSet lists, ListSet splitters, const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ before, uint32_t g, uint32_t perTile, uint32_t boundaries, uint32_t* __restrict__ rows) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    const int lc = lists.count;
    const uint32_t stride = lc + 1;

    if (b == boundaries) {
        uint32_t total = 0;
        for (int l = 0; l < lc; l++) {
            rows[static_cast<uint64_t>(b) * stride + l] = lists.length[l];
            total += lists.length[l];
        }
        rows[static_cast<uint64_t>(b) * stride + lc] = total;
        return;
    }
    if (b > boundaries) {
        return;
    }

    uint32_t s = sorted[static_cast<uint64_t>(b) * perTile];
    if (s >= splitters.indexBase[lc]) {
        s = 0;  // an invalid child merge; the tile checks then reject the rows
    }
    const int own = listOfIndex(splitters, s);
    uint32_t ownIndex = (s - splitters.indexBase[own]) * g;
    if (ownIndex >= lists.length[own]) {
        ownIndex = 0;
    }
    MergeRecord<V, Texts::kBatch> record;
    loadRecord<V, Texts>(c, texts, lists.ptr[own][ownIndex], record);
    uint32_t total = 0;

    for (int l = 0; l < lc; l++) {
        uint32_t r;
        if (l == own) {
            r = ownIndex;
        } else {
            const uint32_t count = before[static_cast<uint64_t>(l) * boundaries + b] - splitters.indexBase[l];
            if (count == 0) {
                r = 0;
            } else {
                uint32_t lo = (count - 1) * g + 1;
                uint32_t hi = min(count * g, lists.length[l]);
                const uint32_t* list = lists.ptr[l];
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (lessThanRecord<V, Texts>(c, texts, list[mid], record)) {
                        lo = mid + 1;
                    } else {
                        hi = mid;
                    }
                }
                r = lo;
            }
        }
        rows[static_cast<uint64_t>(b) * stride + l] = r;
        total += r;
    }
    rows[static_cast<uint64_t>(b) * stride + lc] = total;
}

// ---- tile layout ----------------------------------------------------------------------------

// Reads tile `tile`'s rows: every list's part must lie inside the list, rows must not go backwards
// and the parts must add up to the difference of the output bases. False (nothing to do, and an
// internal error if flagged by the caller) otherwise.
__device__ __forceinline__ bool tileLayout(const ListSet& lists, const uint32_t* __restrict__ rows, uint64_t tile, uint32_t* segStart, uint32_t* rowStart, uint32_t& outBase) {
    const int lc = lists.count;
    const uint32_t stride = lc + 1;
    uint32_t sum = 0;
    bool ok = true;
    for (int l = 0; l < lc; l++) {
        const uint32_t a = rows[tile * stride + l], b = rows[(tile + 1) * stride + l];
        ok = ok && a <= b && b <= lists.length[l];
        segStart[l] = sum;
        rowStart[l] = a;
        sum += ok ? b - a : 0;
    }
    segStart[lc] = sum;
    outBase = rows[tile * stride + lc];
    const uint32_t nextBase = rows[(tile + 1) * stride + lc];
    ok = ok && outBase <= nextBase && nextBase - outBase == sum && nextBase <= lists.indexBase[lc];
    return ok;
}

// ---- tile-major copy ----------------------------------------------------------------------

// Copies every tile's parts of the lists to its output range, list after list.
__global__ void tileMajorCopyKernel(ListSet lists, const uint32_t* __restrict__ rows, uint32_t* __restrict__ out) {
    __shared__ uint32_t segStart[kMaxLists + 1];
    __shared__ uint32_t rowStart[kMaxLists];
    __shared__ uint32_t outBase;
    const int lc = lists.count;
    const uint64_t tile = blockIdx.x;

    __shared__ bool valid;
    if (threadIdx.x == 0) {
        valid = tileLayout(lists, rows, tile, segStart, rowStart, outBase);
    }
    __syncthreads();
    if (!valid) {
        return;
    }

    int l = 0;
    for (uint32_t e = threadIdx.x; e < segStart[lc]; e += blockDim.x) {
        while (l + 1 < lc && e >= segStart[l + 1]) {
            l++;
        }
        out[outBase + e] = lists.ptr[l][rowStart[l] + (e - segStart[l])];
    }
}

// ---- the tile merge ---------------------------------------------------------------------------

enum : int {
    kOutSA = 1,       // text positions (local to their text)
    kOutBWT = 2,      // T[p - 1] at its BWT index
    kOutIndex = 4,    // index into the concatenated input lists (splitter levels)
    kInPlace = 8,     // input read from the SA buffer, tile-major
};

struct TileOut {
    uint32_t* sa;
    uint8_t* bwt;
    uint32_t bwtLength;  // bytes of the BWT buffer
    uint32_t* index;
    const uint32_t* firstSuffixRow;  // per text: local row of its suffix 0 (BWT)
};

constexpr int kMergeThreads = 512;

template <int V, bool kBatch>
inline size_t mergeSharedBytes(uint32_t capacity) {
    return static_cast<size_t>(capacity) * (sizeof(MergeRecord<V, kBatch>) + 2 * sizeof(uint16_t) + 1) + 16;
}

template <int V, class Texts>
__device__ __forceinline__ void mergeChunk(const MergeCtx& c, const MergeRecord<V, Texts::kBatch>* recs, const uint16_t* src, uint16_t* dst, const uint32_t* bounds, uint32_t runs, uint32_t o, uint32_t o1) {
    const uint32_t pairs = (runs + 1) / 2;
    uint32_t j = 0;
    while (o < o1) {
        while (j + 1 < pairs && bounds[2 * (j + 1)] <= o) {
            j++;
        }
        const uint32_t aStart = bounds[2 * j];
        const uint32_t aEnd = bounds[min(2 * j + 1, runs)];
        const uint32_t bEnd = bounds[min(2 * j + 2, runs)];
        const uint32_t lengthA = aEnd - aStart, lengthB = bEnd - aEnd;
        const uint32_t d = o - aStart;
        uint32_t lo = d > lengthB ? d - lengthB : 0;
        uint32_t hi = min(d, lengthA);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (lessRecord<V, Texts>(c, recs[src[aStart + mid]], recs[src[aEnd + d - 1 - mid]])) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        uint32_t ia = lo, ib = d - lo;
        const uint32_t stop = min(o1, bEnd);
        if (stop <= o) {
            return;  // inconsistent bounds; the output check reports it
        }
        while (o < stop) {
            bool takeA;
            if (ib >= lengthB) {
                takeA = true;
            } else if (ia >= lengthA) {
                takeA = false;
            } else {
                takeA = lessRecord<V, Texts>(c, recs[src[aStart + ia]], recs[src[aEnd + ib]]);
            }
            dst[o++] = takeA ? src[aStart + ia++] : src[aEnd + ib++];
        }
    }
}

template <int V, class Texts, int kFlags>
__global__ __launch_bounds__(kMergeThreads) void mergeTilesKernel(MergeCtx c, Texts texts, ListSet lists, const uint32_t* __restrict__ rows, uint32_t capacity, TileOut out, uint32_t* __restrict__ errors) {
    using Record = MergeRecord<V, Texts::kBatch>;
    extern __shared__ __align__(16) unsigned char shared[];
    Record* recs = reinterpret_cast<Record*>(shared);
    uint16_t* indexA = reinterpret_cast<uint16_t*>(recs + capacity);
    uint16_t* indexB = indexA + capacity;
    uint8_t* before = reinterpret_cast<uint8_t*>(indexB + capacity);
    __shared__ uint32_t bounds[kMaxLists + 1];
    __shared__ uint32_t segStart[kMaxLists + 1];
    __shared__ uint32_t rowStart[kMaxLists];
    __shared__ uint32_t outBase;

    const int lc = lists.count;
    const uint64_t tile = blockIdx.x;

    __shared__ bool valid;
    if (threadIdx.x == 0) {
        valid = tileLayout(lists, rows, tile, segStart, rowStart, outBase) && segStart[lc] <= capacity;
        for (int l = 0; l <= lc; l++) {
            bounds[l] = segStart[l];
        }
        if (!valid) {
            atomicOr(errors, 1u);
        }
    }
    __syncthreads();
    if (!valid) {
        return;
    }
    const uint32_t size = segStart[lc];

    // Load the records; the BWT character before each.
    {
        int l = 0;
        for (uint32_t e = threadIdx.x; e < size; e += kMergeThreads) {
            while (l + 1 < lc && e >= segStart[l + 1]) {
                l++;
            }
            uint32_t p = (kFlags & kInPlace) ? out.sa[outBase + e] : lists.ptr[l][rowStart[l] + (e - segStart[l])];
            if (p >= c.n) {
                atomicOr(errors, 2u);
                p = 0;
            }
            loadRecord<V, Texts>(c, texts, p, recs[e]);
            if constexpr ((kFlags & kOutBWT) != 0) {
                before[e] = p > 0 ? reinterpret_cast<const uint8_t*>(c.text)[p - 1] : 0;
            }
            indexA[e] = static_cast<uint16_t>(e);
        }
    }
    __syncthreads();

    uint16_t* src = indexA;
    uint16_t* dst = indexB;
    uint32_t runs = static_cast<uint32_t>(lc);
    const uint32_t chunk = max(4u, (size + kMergeThreads - 1) / kMergeThreads);

    while (runs > 1) {
        for (uint64_t o = static_cast<uint64_t>(threadIdx.x) * chunk; o < size; o += static_cast<uint64_t>(kMergeThreads) * chunk) {
            mergeChunk<V, Texts>(c, recs, src, dst, bounds, runs, static_cast<uint32_t>(o), static_cast<uint32_t>(min(o + chunk, static_cast<uint64_t>(size))));
        }
        __syncthreads();
        const uint32_t pairs = (runs + 1) / 2;
        if (threadIdx.x == 0) {
            for (uint32_t j = 0; j < pairs; j++) {
                bounds[j] = bounds[2 * j];
            }
            bounds[pairs] = size;
        }
        __syncthreads();
        uint16_t* t = src;
        src = dst;
        dst = t;
        runs = pairs;
    }

    for (uint32_t r = threadIdx.x; r < size; r += kMergeThreads) {
        const uint32_t e = src[r];
        const uint32_t p = recs[e].position;
        const uint32_t row = outBase + r;
        uint32_t start = 0, rowBase = 0, segment = 0;
        if constexpr (Texts::kBatch) {
            segment = recs[e].tag >> 9;
            start = c.segments[segment].start;
            rowBase = c.segments[segment].rowBase;
        }
        if constexpr ((kFlags & kOutSA) != 0) {
            out.sa[row] = p - start;
        }
        if constexpr ((kFlags & kOutIndex) != 0) {
            int l = 0;
            while (l + 1 < lc && e >= segStart[l + 1]) {
                l++;
            }
            out.index[row] = lists.indexBase[l] + rowStart[l] + (e - segStart[l]);
        }
        if constexpr ((kFlags & kOutBWT) != 0) {
            if (p != start) {
                const uint32_t local = row - rowBase;
                const uint32_t first = out.firstSuffixRow[segment];
                const uint32_t target = rowBase + (local < first ? local + 1 : local);
                if (target < out.bwtLength) {
                    out.bwt[target] = before[e];
                } else {
                    atomicOr(errors, 4u);
                }
            }
        }
    }
}

// Per text: the row of its first suffix (the number of its suffixes smaller than it), and the
// BWT's byte 0 for that text, T[end - 1].
template <int V, class Texts>
__global__ void firstSuffixRowsKernel(MergeCtx c, Texts texts, ListSet lists, uint32_t segments, uint32_t* __restrict__ rowsOut, uint8_t* __restrict__ bwt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= segments) {
        return;
    }
    uint32_t start = 0, end = c.n, rowBase = 0;
    if constexpr (Texts::kBatch) {
        start = c.segments[s].start;
        end = c.segments[s].end;
        rowBase = c.segments[s].rowBase;
    }
    MergeRecord<V, Texts::kBatch> record;
    loadRecord<V, Texts>(c, texts, start, record);
    uint32_t total = 0;
    for (int l = 0; l < lists.count; l++) {
        uint32_t lo = 0, hi = lists.length[l];
        const uint32_t* list = lists.ptr[l];
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (list[mid] != start && lessThanRecord<V, Texts>(c, texts, list[mid], record)) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        total += lo;
    }
    rowsOut[s] = total - rowBase;
    if (bwt) {
        bwt[rowBase] = reinterpret_cast<const uint8_t*>(c.text)[end - 1];
    }
}

}  // namespace sadcs
