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
 *    Each record (32 characters, the V-window of sample ranks, the position) is loaded once.
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
// Characters kept in a record: 32, so with the mod-32 cover (tie offsets below 32) no comparison
// inside a tile reads the text; with mod 128 the rest comes from global memory.
constexpr int kRecordWords = 4;

// The tables a tile comparison reads for covers other than 32 (cover 32 computes tie offsets by
// bit arithmetic): the tie offsets, flattened [a * V + b], in __constant__ memory, and per
// difference d an anchor, a cover element x with x + d in the cover (cover 128: a copy in LDS).
struct TieTables {
    const uint8_t* tie;
    const uint8_t* anchor;
};

// Cover 32 as a bit mask: bit k of rotr(mask, r) says whether r + k is a sample position (mod 32).
// The tie offset of residues a and b is then the lowest set bit of rotr(mask, a) & rotr(mask, b),
// and the slot of a + k in a's rank window (samples in [a, a + k)) a population count.
template <int V>
constexpr uint32_t coverMask() {
    static_assert(V == 32, "a 32-bit mask");
    uint32_t mask = 0;
    for (int e = 0; e < CoverSet<V>::kSize; e++) {
        mask |= 1u << CoverSet<V>::kElements[e];
    }
    return mask;
}

__device__ __forceinline__ uint32_t windowOf32(uint32_t mask, uint32_t p) {
    return __builtin_rotateright32(mask, p & 31);
}

template <int V>
__device__ __forceinline__ const CoverTables<V>& deviceCover();

template <int V>
__device__ __forceinline__ uint32_t packPosition(uint32_t p);

template <int V>
__device__ __forceinline__ uint32_t coverBelow(uint32_t r);

struct MergeCtx {
    const uint64_t* text;
    const uint32_t* ranks;
    const Segment* segments;
    uint32_t n;          // single text: its length
    const uint4* lines;  // cover 32: one line per 32 positions (buildLinesKernel), or null
};

// ---- lines: one 128-byte line per 32 text positions (cover 32) ------------------------------
//
// Line b holds T[32b - 1 .. 32b + 62] (bytes 0..63, T[-1] = 0) and the ranks of the 14 sample
// positions in [32b, 32b + 64), pack indices 7b .. 7b + 13 (bytes 64..119). A record of any p in
// [32b, 32b + 32) - T[p - 1 .. p + 31] and the 7 ranks from the first sample at or after p - lies
// inside line b, so a record costs one line of HBM traffic instead of the 2 to 3 lines its text
// window and rank window straddle when read from the text and the rank array.
constexpr int kLineBytes = 128;
constexpr int kLinePositions = 32;

inline uint64_t lineCount(uint64_t span) {
    return (span + kLinePositions - 1) / kLinePositions;
}

// 8 threads per line, one 16-byte word each: words 0..3 text, 4..7 ranks. The text is 16-byte
// aligned and readable to span + 63; the rank array holds 15 words past its last sample.
__global__ void buildLinesKernel(const uint4* __restrict__ text, const uint32_t* __restrict__ ranks, uint64_t lines, uint4* __restrict__ out,
                                 const uint32_t* __restrict__ blockSegment, const Segment* __restrict__ segments) {
    const uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (x >= lines * 8) {
        return;
    }
    const uint64_t b = x >> 3;
    const uint32_t t = static_cast<uint32_t>(x & 7);
    uint4 v;
    if (t < 4) {
        // Bytes 32b - 1 + 16t .. 32b + 14 + 16t: the last byte of the aligned chunk before, then 15.
        const uint64_t chunk = 2 * b + t;  // the aligned chunk at 32b + 16t
        const uint4 cur = text[chunk];
        const uint32_t prevLast = chunk > 0 ? text[chunk - 1].w >> 24 : 0;
        v.x = (cur.x << 8) | prevLast;
        v.y = (cur.y << 8) | (cur.x >> 24);
        v.z = (cur.z << 8) | (cur.y >> 24);
        v.w = (cur.w << 8) | (cur.z >> 24);
    } else {
        const uint64_t first = 7 * b + 4 * (t - 4);
        v.x = ranks[first];
        v.y = ranks[first + 1];
        v.z = t < 7 ? ranks[first + 2] : 0;
        v.w = t < 7 ? ranks[first + 3] : 0;
        if (t == 7 && blockSegment) {
            // A batch: the line's text and that text's end in the last two words (a line's 32
            // positions lie in one text, which starts on a block boundary).
            const uint32_t segment = blockSegment[(b * kLinePositions) >> kSegmentBlockLog];
            v.z = segment;
            v.w = segments[segment].end;
        }
    }
    out[x] = v;
}

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
    uint32_t ranks[CoverSet<V>::kSize];  // the rank window, where recordRanks<V>()
    uint32_t position;
    uint32_t tag;  // batch: segment << 9 | min(remaining, V)
};

template <int V>
struct MergeRecord<V, false> {
    uint64_t words[kRecordWords];
    uint32_t ranks[CoverSet<V>::kSize];  // the rank window, where recordRanks<V>()
    uint32_t position;
};

// Whether a record holds its rank window. Cover 32 does: its window (7 ranks) lies in the record's
// line. Cover 128 does not: 13 ranks would more than double a record's LDS (88 bytes instead of
// 36) and halve the records a tile holds, while its comparisons rarely reach the ranks (only when
// the first k characters tie, k averaging 59); those that do read the two ranks from the rank
// array.
template <int V>
constexpr bool recordRanks() {
    return V == 32;
}

// A record with few, wide gathers, in two steps so that a caller can issue the loads early and use
// them late: fetchRecord issues three 16-byte loads holding T[p - 1 .. p + 31] and kQuads 16-byte
// loads holding the rank window (none without one); decodeRecord shifts them into place. From the
// lines (cover 32) all of them lie in one 128-byte line; otherwise they come from the text and the
// rank array (16-byte aligned and padded past their ends). `below` is the number of cover elements
// below p mod V (read by the caller from whichever copy of the table is closest).
template <int V>
struct RawRecord {
    static constexpr int kQuads = recordRanks<V>() ? (CoverSet<V>::kSize + 3 + 3) / 4 : 0;  // 16-byte loads covering kSize words from any offset
    uint4 chunk[3];
    uint4 quad[kQuads > 0 ? kQuads : 1];
};

template <int V>
__device__ __forceinline__ bool recordFromLines(const MergeCtx& c) {
    if constexpr (V == kLinePositions) {
        return c.lines != nullptr;
    } else {
        return false;
    }
}

template <int V>
__device__ __forceinline__ void fetchRecord(const MergeCtx& c, uint32_t p, uint32_t below, RawRecord<V>& raw) {
    const uint4* chunks;
    const uint4* quads;
    if (recordFromLines<V>(c)) {
        const uint4* line = c.lines + (static_cast<uint64_t>(p / kLinePositions) << 3);
        chunks = line + ((p & (kLinePositions - 1)) >> 4);
        quads = line + 4 + (below >> 2);
    } else {
        const uint32_t first = p > 0 ? p - 1 : 0;
        chunks = reinterpret_cast<const uint4*>(c.text) + (first >> 4);
        const uint32_t index = (p / V) * CoverSet<V>::kSize + below;
        quads = reinterpret_cast<const uint4*>(c.ranks) + (index >> 2);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        raw.chunk[q] = chunks[q];
    }
#pragma unroll
    for (int q = 0; q < RawRecord<V>::kQuads; q++) {
        raw.quad[q] = quads[q];
    }
}

template <int V, class Texts>
__device__ __forceinline__ void decodeRecord(const MergeCtx& c, const Texts& texts, uint32_t p, uint32_t below, const RawRecord<V>& raw, MergeRecord<V, Texts::kBatch>& r, uint8_t* before) {
    static_assert(kRecordWords == 4, "three 16-byte loads cover 32 characters and the one before");
    constexpr int kSize = CoverSet<V>::kSize;
    constexpr int kQuads = RawRecord<V>::kQuads;
    // T[p] sits at byte `offset` of the 48 loaded bytes; the first rank at word `o` of the quads.
    uint32_t offset, o;
    if (recordFromLines<V>(c)) {
        const uint32_t q = p & (kLinePositions - 1);
        offset = q + 1 - (q & 16);  // 1 .. 16
        o = below & 3;
    } else {
        const uint32_t first = p > 0 ? p - 1 : 0;
        offset = p - (first & ~15u);  // 0 .. 16
        o = ((p / V) * kSize + below) & 3;
    }
    const uint4 q0 = raw.chunk[0], q1 = raw.chunk[1], q2 = raw.chunk[2];
    uint64_t w[6] = {q0.x | (uint64_t(q0.y) << 32), q0.z | (uint64_t(q0.w) << 32),
                     q1.x | (uint64_t(q1.y) << 32), q1.z | (uint64_t(q1.w) << 32),
                     q2.x | (uint64_t(q2.y) << 32), q2.z | (uint64_t(q2.w) << 32)};
    if (before) {
        const uint32_t b = offset - 1;  // only meaningful for p > 0
        *before = p > 0 ? static_cast<uint8_t>(w[(b >> 3) & 1] >> ((b & 7) * 8)) : 0;
    }
    // Shift by whole words (0, 1 or 2), then by bytes.
    const uint32_t wordShift = offset >> 3;
    uint64_t v[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const uint64_t a0 = w[k];
        const uint64_t a1 = k + 1 < 6 ? w[k + 1] : 0;
        const uint64_t a2 = k + 2 < 6 ? w[k + 2] : 0;
        v[k] = wordShift == 0 ? a0 : wordShift == 1 ? a1 : a2;
    }
    const uint32_t s = (offset & 7) * 8;
#pragma unroll
    for (int k = 0; k < kRecordWords; k++) {
        r.words[k] = __builtin_bswap64(s ? (v[k] >> s) | (v[k + 1] << (64 - s)) : v[k]);
    }
    if constexpr (kQuads > 0) {
        uint32_t x[4 * kQuads];
#pragma unroll
        for (int q = 0; q < kQuads; q++) {
            x[4 * q] = raw.quad[q].x;
            x[4 * q + 1] = raw.quad[q].y;
            x[4 * q + 2] = raw.quad[q].z;
            x[4 * q + 3] = raw.quad[q].w;
        }
#pragma unroll
        for (int s2 = 0; s2 < kSize; s2++) {
            r.ranks[s2] = o == 0 ? x[s2] : o == 1 ? x[s2 + 1] : o == 2 ? x[s2 + 2] : x[s2 + 3];
        }
    }
    r.position = p;
    if constexpr (Texts::kBatch) {
        const uint32_t segment = texts.segmentOf(p);
        const uint32_t remaining = c.segments[segment].end - p;
        r.tag = (segment << 9) | min(remaining, static_cast<uint32_t>(V));
    }
}

template <int V, class Texts>
__device__ __forceinline__ void loadRecord(const MergeCtx& c, const Texts& texts, uint32_t p, MergeRecord<V, Texts::kBatch>& r, uint8_t* before = nullptr) {
    const uint32_t below = coverBelow<V>(p & (V - 1));
    RawRecord<V> raw;
    fetchRecord<V>(c, p, below, raw);
    decodeRecord<V, Texts>(c, texts, p, below, raw, r, before);
}

template <int V, class Texts>
__device__ __forceinline__ uint32_t remainingOf(const MergeCtx& c, const MergeRecord<V, Texts::kBatch>& r) {
    if constexpr (Texts::kBatch) {
        return r.tag & 0x1FF;
    } else {
        return min(c.n - r.position, static_cast<uint32_t>(V));
    }
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
    if constexpr (!recordRanks<V>()) {
        return c.ranks[packPosition<V>(i + k)] < c.ranks[packPosition<V>(j + k)];
    } else {
        // The record lives in registers: select its rank without indexing it dynamically.
        const uint32_t slot = cv.slot[j & (V - 1)][k];
        uint32_t rankB = 0;
#pragma unroll
        for (int s = 0; s < CoverSet<V>::kSize; s++) {
            rankB = s == static_cast<int>(slot) ? b.ranks[s] : rankB;
        }
        return c.ranks[packPosition<V>(i + k)] < rankB;
    }
}

// The list holding concatenated index `index`: the number of list starts after the first at or
// below it. Branch-free, against bases that are the same for every lane (kernel arguments).
__device__ __forceinline__ int listOfIndex(const ListSet& lists, uint32_t index) {
    int l = 0;
#pragma unroll
    for (int k = 1; k < kMaxLists; k++) {
        l += k < lists.count && index >= lists.indexBase[k] ? 1 : 0;
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

// counts[l * boundaries + b] = splitters of list l among sorted splitters [b*K, (b+1)*K). One
// wave per boundary: the lanes classify 64 splitters at a time into the wave's LDS histogram.
__global__ __launch_bounds__(256) void countBoundarySplittersKernel(ListSet splitters, const uint32_t* __restrict__ sorted, uint32_t total, uint32_t perTile, uint32_t boundaries, uint32_t* __restrict__ counts) {
    __shared__ uint32_t histogram[256 / kWave][kMaxLists];
    const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (b >= boundaries) {
        return;  // whole waves leave together, and no wave shares its histogram
    }
    const int lc = splitters.count;
    uint32_t* h = histogram[wave];
    if (lane < static_cast<uint32_t>(kMaxLists)) {
        h[lane] = 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint64_t first = static_cast<uint64_t>(b) * perTile;
    const uint64_t last = min(first + perTile, static_cast<uint64_t>(total));
    for (uint64_t x0 = first; x0 < last; x0 += kWave) {
        const uint64_t x = x0 + lane;
        if (x < last) {
            atomicAdd(&h[listOfIndex(splitters, sorted[x])], 1u);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < static_cast<uint32_t>(lc)) {
        counts[static_cast<uint64_t>(lane) * boundaries + b] = h[lane];
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

// A level's tile-row table, (tiles + 1) rows of lists + 1 words: plain words, or the top level of a
// fused single-text build, whose table moves into the lines before the merge (relocateRowsKernel):
// word w is then the spare dword 30 + (w mod 2) of line w / 2.
struct RowTable {
    const uint32_t* words;
    uint32_t inLines;

    __device__ __forceinline__ uint32_t operator[](uint64_t w) const {
        return inLines ? words[(w >> 1) * (kLineBytes / 4) + (kLineBytes / 4 - 2) + (w & 1)] : words[w];
    }
};

// Moves a row table of `count` words into the two spare dwords of consecutive lines (count is at
// most twice the line count; the lines' builder leaves those dwords zero for a single text).
__global__ void relocateRowsKernel(const uint32_t* __restrict__ rows, uint64_t count, uint32_t* __restrict__ lines) {
    for (uint64_t w = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; w < count; w += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        lines[(w >> 1) * (kLineBytes / 4) + (kLineBytes / 4 - 2) + (w & 1)] = rows[w];
    }
}

// A tile's layout is read by wave 0 of the workgroup, one list per lane: tileRows loads the row
// entries (a = the tile's, b = the next tile's), tilePublish checks them - every list's part must
// lie inside the list, rows must not go backwards, the parts must add up to the difference of the
// output bases and fit `capacity` - and writes the run starts segStart[0..lc], the list rows
// rowStart[0..lc) (if wanted), the output base and the verdict. The caller synchronises before
// reading them. Both are no-ops outside wave 0.
__device__ __forceinline__ void tileRows(const ListSet& lists, const RowTable& rows, uint64_t tile, uint32_t& a, uint32_t& b) {
    const uint32_t lane = threadIdx.x;
    const uint32_t lc = static_cast<uint32_t>(lists.count);
    a = 0;
    b = 0;
    if (lane <= lc) {
        a = rows[tile * (lc + 1) + lane];
        b = rows[(tile + 1) * (lc + 1) + lane];
    }
}

__device__ __forceinline__ void tilePublish(const ListSet& lists, uint32_t a, uint32_t b, uint32_t capacity, uint32_t* segStart, uint32_t* rowStart, uint32_t* outBase, uint32_t* valid) {
    if (threadIdx.x >= kWave) {
        return;
    }
    const uint32_t lane = threadIdx.x;
    const uint32_t lc = static_cast<uint32_t>(lists.count);
    const bool isList = lane < lc;
    const bool ok = !isList || (a <= b && b <= lists.length[lane]);
    const uint32_t length = isList && ok ? b - a : 0;
    const uint32_t inclusive = waveInclusive<OpSum>(length);
    const uint32_t base = readLane(a, lc), end = readLane(b, lc), total = readLane(inclusive, lc);
    const bool allOk = __all(ok);
    if (lane <= lc) {
        segStart[lane] = inclusive - length;
    }
    if (isList && rowStart) {
        rowStart[lane] = a;
    }
    if (lane == 0) {
        *outBase = base;
        *valid = allOk && base <= end && end - base == total && end <= lists.indexBase[lc] && total <= capacity;
    }
}

__device__ __forceinline__ void tileLayout(const ListSet& lists, const RowTable& rows, uint64_t tile, uint32_t capacity, uint32_t* segStart, uint32_t* rowStart, uint32_t* outBase, uint32_t* valid) {
    if (threadIdx.x < kWave) {
        uint32_t a, b;
        tileRows(lists, rows, tile, a, b);
        tilePublish(lists, a, b, capacity, segStart, rowStart, outBase, valid);
    }
}

// The first output row of every slice of a level's tiles, and the total after them: slice s starts
// at tile s * tiles / slices, and row tile * (lists + 1) + lists of the row table is a tile's first
// output row (tile = tiles: the total). One workgroup; slices < 64.
__global__ void sliceRowsKernel(RowTable rows, uint32_t lists, uint32_t tiles, uint32_t slices, uint32_t* __restrict__ out) {
    const uint32_t s = threadIdx.x;
    if (s <= slices) {
        const uint64_t tile = static_cast<uint64_t>(tiles) * s / slices;
        out[s] = rows[tile * (lists + 1) + lists];
    }
}

// ---- tile-major copy ----------------------------------------------------------------------

// Copies every tile's parts of the lists to its output range, list after list.
__global__ __launch_bounds__(256) void tileMajorCopyKernel(ListSet lists, const uint32_t* __restrict__ rows, uint32_t tiles, uint32_t* __restrict__ out) {
    // One wave per tile, four tiles per workgroup: a tile is ~1200 elements, and a workgroup per
    // tile left the copy bound by workgroup dispatch rather than by HBM.
    constexpr uint32_t kTileWaves = 4;
    constexpr int kCopyItems = 19;  // a default tile (at most 1216 elements) in one batch of loads
    __shared__ uint32_t segStart[kTileWaves][kMaxLists + 1];
    __shared__ uint32_t rowStart[kTileWaves][kMaxLists];
    __shared__ const uint32_t* listPtr[kMaxLists];  // indexed per lane: LDS, not kernel-argument loads
    const uint32_t lc = static_cast<uint32_t>(lists.count);
    const uint32_t wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kTileWaves + wave;
    if (threadIdx.x < lc) {
        listPtr[threadIdx.x] = lists.ptr[threadIdx.x];
    }
    __syncthreads();
    if (tile >= tiles) {
        return;
    }
    // The tile's rows of every list (lane l: list l), as tilePublish checks them.
    uint32_t a = 0, b = 0;
    if (lane <= lc) {
        a = rows[tile * (lc + 1) + lane];
        b = rows[(tile + 1) * (lc + 1) + lane];
    }
    const bool isList = lane < lc;
    const bool ok = !isList || (a <= b && b <= lists.length[lane]);
    const uint32_t length = isList && ok ? b - a : 0;
    const uint32_t inclusive = waveInclusive<OpSum>(length);
    const uint32_t base = readLane(a, lc), end = readLane(b, lc), size = readLane(inclusive, lc);
    if (!__all(ok) || base > end || end - base != size || end > lists.indexBase[lc]) {
        return;  // inconsistent rows: the merge kernel reports them
    }
    if (lane <= lc) {
        segStart[wave][lane] = inclusive - length;
    }
    if (isList) {
        rowStart[wave][lane] = a;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t* seg = segStart[wave];
    const uint32_t* row = rowStart[wave];
    // All loads first, then the stores (which the compiler must assume could alias the lists).
    for (uint32_t e0 = lane; e0 < size; e0 += kCopyItems * kWave) {
        uint32_t v[kCopyItems];
        uint32_t l = 0;
#pragma unroll
        for (int j = 0; j < kCopyItems; j++) {
            const uint32_t e = min(e0 + j * kWave, size - 1);
            while (l + 1 < lc && e >= seg[l + 1]) {
                l++;
            }
            v[j] = listPtr[l][row[l] + (e - seg[l])];
        }
#pragma unroll
        for (int j = 0; j < kCopyItems; j++) {
            const uint32_t e = e0 + j * kWave;
            if (e < size) {
                out[base + e] = v[j];
            }
        }
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

// Threads per tile: 512 (two tiles per CU in the LDS) or 256 (four tiles per CU, each with half the
// records), a tuned parameter. kMergeThreads is the default.
constexpr int kMergeThreads = 512;

// The defaults per cover, measured on one MI355X (16 MiB to 1 GiB texts). A tile holds up to
// K*G + C*(G-1) records for C lists, C*(G-1) of them the room its unaligned run ends need, so the
// 31 lists of cover 128 want a smaller G than the 7 of cover 32: at G = 16 that room would be 465
// of the 973 records of a 256-thread tile. Cover 128 merges in 256-thread tiles (four per CU) at
// G = 8; cover 32 in 512-thread tiles (two per CU) at G = 16.
constexpr uint32_t defaultSplitterSpacing(int cover) {
    return cover == 128 ? 8 : 16;
}
constexpr int defaultTileThreads(int cover) {
    return cover == 128 ? 256 : kMergeThreads;
}
constexpr int kLoadSlots = 3;  // records each thread gathers at once (default tiles hold at most 3 * threads)

// A tile's records in LDS as a structure of arrays: a wave reading one field of 64 random records
// then hits random banks. (As 64-byte structs, the same field of different records would map to
// 2 of the 32 banks and every comparison would serialise on bank conflicts.)
template <int V, bool kBatch>
struct LdsRecords {
    static constexpr int kSize = CoverSet<V>::kSize;
    // A batch keeps each record's position and tag side by side, so that a merge step reads both
    // with one 8-byte LDS load.
    static constexpr uint32_t kPositionStride = kBatch ? 2 : 1;
    uint32_t capacity;
    uint64_t* words;     // [kRecordWords][capacity]
    uint32_t* position;  // [capacity], or (position, tag) pairs [capacity][2] for a batch
    uint32_t* ranks;     // [kSize][capacity]

    static constexpr int kStoredRanks = recordRanks<V>() ? kSize : 0;
    static constexpr size_t bytesPerRecord() {
        return 8 * kRecordWords + 4 + 4 * kStoredRanks + (kBatch ? 4 : 0);
    }

    __device__ __forceinline__ void carve(unsigned char* base, uint32_t cap) {
        capacity = cap;
        words = reinterpret_cast<uint64_t*>(base);
        position = reinterpret_cast<uint32_t*>(words + static_cast<size_t>(kRecordWords) * cap);
        ranks = position + static_cast<size_t>(kPositionStride) * cap;
    }

    __device__ __forceinline__ uint32_t positionOf(uint32_t e) const {
        return position[e * kPositionStride];
    }
    __device__ __forceinline__ uint32_t tagOf(uint32_t e) const {  // batch only
        return position[e * 2 + 1];
    }
    __device__ __forceinline__ void setPosition(uint32_t e, uint32_t p) const {
        position[e * kPositionStride] = p;
    }
    __device__ __forceinline__ void setTag(uint32_t e, uint32_t t) const {  // batch only
        position[e * 2 + 1] = t;
    }

    __device__ __forceinline__ void store(uint32_t e, const MergeRecord<V, kBatch>& r) const {
#pragma unroll
        for (int w = 0; w < kRecordWords; w++) {
            words[w * capacity + e] = r.words[w];
        }
        if constexpr (kBatch) {
            reinterpret_cast<uint2*>(position)[e] = uint2{r.position, r.tag};
        } else {
            position[e] = r.position;
        }
#pragma unroll
        for (int s = 0; s < kStoredRanks; s++) {
            ranks[s * capacity + e] = r.ranks[s];
        }
    }
};

template <int V, bool kBatch>
inline size_t mergeSharedBytes(uint32_t capacity) {
    return static_cast<size_t>(capacity) * (LdsRecords<V, kBatch>::bytesPerRecord() + 2 * sizeof(uint16_t) + 1) + 16;
}

// Tiles are latency-bound, so tiles per CU matter more than tile size: the default K is the largest
// whose tiles (at most K*G + C*(G-1) records) leave room in the 160 KB of LDS for as many tiles as
// the CU runs at once (VGPRs allow 16 waves per CU: two 512-thread tiles or four 256-thread ones).
inline size_t tileLdsBudget(int threads) {
    return (threads == 256 ? 40 : 80) * 1024 - 1024;
}

template <int V, bool kBatch>
inline uint32_t defaultTileSplitters(uint32_t g, int threads = kMergeThreads) {
    const size_t perRecord = LdsRecords<V, kBatch>::bytesPerRecord() + 2 * sizeof(uint16_t) + 1;
    const size_t records = (tileLdsBudget(threads) - 16) / perRecord;
    const size_t spill = static_cast<size_t>(DifferenceCover<V>::kClasses) * (g - 1);
    return records > spill + g ? static_cast<uint32_t>((records - spill) / g) : 1;
}

// The fields of an LDS record that decide most comparisons (cover 32): a merge keeps them in
// registers for its two heads, so a merge step loads one record's key instead of two.
template <bool kBatch>
struct HeadKey {
    uint32_t position;
    uint32_t tag;  // batch only
    uint64_t word;
};

template <int V, bool kBatch>
__device__ __forceinline__ HeadKey<kBatch> headKey(const LdsRecords<V, kBatch>& R, uint32_t x) {
    HeadKey<kBatch> k;
    if constexpr (kBatch) {
        const uint2 pt = reinterpret_cast<const uint2*>(R.position)[x];
        k.position = pt.x;
        k.tag = pt.y;
    } else {
        k.position = R.position[x];
        k.tag = 0;
    }
    k.word = R.words[x];
    return k;
}

// Record x < record y (cover 32) given their keys; the other words and the ranks come from LDS
// only if the first words tie. Cover 32 keeps all the characters a comparison can need in the
// record and finds the tie offset and rank slots by bit arithmetic.
template <int V, class Texts>
__device__ __forceinline__ bool lessKeyed(const MergeCtx& c, const LdsRecords<V, Texts::kBatch>& R, uint32_t x, const HeadKey<Texts::kBatch>& kx, uint32_t y, const HeadKey<Texts::kBatch>& ky) {
    static_assert(V == 32, "keys hold what a cover-32 comparison reads first");
    const uint32_t cap = R.capacity;
    const uint32_t i = kx.position, j = ky.position;
    const uint64_t a0 = kx.word, b0 = ky.word;
    uint32_t ra, rb, sameText = 1, textLess = 0;
    if constexpr (Texts::kBatch) {
        const uint32_t ta = kx.tag, tb = ky.tag;
        sameText = (ta >> 9) == (tb >> 9);
        textLess = (ta >> 9) < (tb >> 9);
        ra = ta & 0x1FF;
        rb = tb & 0x1FF;
    } else {
        ra = min(c.n - i, static_cast<uint32_t>(V));
        rb = min(c.n - j, static_cast<uint32_t>(V));
    }
    const uint32_t wi = windowOf32(coverMask<V>(), i), wj = windowOf32(coverMask<V>(), j);
    const uint32_t k = __builtin_ctz(wi & wj);
    const uint32_t common = min(k, min(ra, rb));
    const uint64_t m0 = common == 0 ? 0 : prefixMask(common);
    bool decided = (a0 & m0) != (b0 & m0);
    bool less = (a0 & m0) < (b0 & m0);
    if (!decided) {
        const bool rankTie = common == k && ra != k && rb != k;
        uint32_t rankA = 0, rankB = 0;
        if (rankTie) {
            const uint32_t before = (1u << k) - 1;
            rankA = R.ranks[__builtin_popcount(wi & before) * cap + x];
            rankB = R.ranks[__builtin_popcount(wj & before) * cap + y];
        }
        if (common > 8) {
#pragma unroll
            for (int w = 1; w < kRecordWords; w++) {
                const uint32_t rest = common > 8u * w ? common - 8u * w : 0u;
                const uint64_t mask = rest == 0 ? 0 : prefixMask(rest);
                const uint64_t av = R.words[w * cap + x] & mask, bv = R.words[w * cap + y] & mask;
                less = decided ? less : av < bv;
                decided = decided || av != bv;
            }
        }
        if (!decided) {
            less = rankTie ? rankA < rankB : i > j;  // i > j: one is a prefix of the other, the shorter is smaller
        }
    }
    return sameText ? less : textLess;
}

// Record x < record y, both in LDS: at most two dependent LDS round trips for cover 32 (the keys,
// then - only if those tie - the other words and the two ranks together).
template <int V, class Texts>
__device__ __forceinline__ bool lessLds(const MergeCtx& c, const TieTables& t, const LdsRecords<V, Texts::kBatch>& R, uint32_t x, uint32_t y) {
    if constexpr (V == 32) {
        return lessKeyed<V, Texts>(c, R, x, headKey<V, Texts::kBatch>(R, x), y, headKey<V, Texts::kBatch>(R, y));
    } else {
        uint32_t ra, rb;
        if constexpr (Texts::kBatch) {
            const uint32_t ta = R.tagOf(x), tb = R.tagOf(y);
            if ((ta >> 9) != (tb >> 9)) {
                return (ta >> 9) < (tb >> 9);
            }
            ra = ta & 0x1FF;
            rb = tb & 0x1FF;
        }
        const uint32_t i = R.positionOf(x), j = R.positionOf(y);
        if constexpr (!Texts::kBatch) {
            ra = min(c.n - i, static_cast<uint32_t>(V));
            rb = min(c.n - j, static_cast<uint32_t>(V));
        }
        uint32_t k;
        if constexpr (V == 128) {
            // i + k = x and j + k = x + d (mod 128) are both samples, for the anchor x of the
            // difference d = j - i: one LDS byte instead of the tie table in global memory. Not
            // always the smallest such k, but any k with both sampled orders the two the same.
            k = (t.anchor[(j - i) & (V - 1)] - i) & (V - 1);
        } else {
            k = t.tie[(i & (V - 1)) * V + (j & (V - 1))];
        }
        const uint32_t common = min(k, min(ra, rb));
        const uint32_t cap = R.capacity;

#pragma unroll
        for (uint32_t w = 0; w < V / 8; w++) {
            if (8 * w >= common) {
                break;
            }
            const uint64_t mask = prefixMask(common - 8 * w);
            const uint64_t a = (w < kRecordWords ? R.words[w * cap + x] : load8(c.text, i + 8 * w)) & mask;
            const uint64_t b = (w < kRecordWords ? R.words[w * cap + y] : load8(c.text, j + 8 * w)) & mask;
            if (a != b) {
                return a < b;
            }
        }
        if (common < k || ra == k || rb == k) {
            return i > j;  // one is a prefix of the other: the shorter suffix is smaller
        }
        // Both suffixes run past i + k and j + k, sample positions: their ranks decide.
        return c.ranks[packPosition<V>(i + k)] < c.ranks[packPosition<V>(j + k)];
    }
}

// Outputs [o, o1) of one round of pairwise merges: run 2p and run 2p + 1 of src (bounds[0..runs])
// merge into dst at the same place. Each output range starts with a merge-path search for its
// first output, then merges sequentially with the two head indices in registers and the indices
// after them loaded together with each comparison. src holds `capacity` entries; kIdentity: src
// is the identity (the first round, runs still in list order) and is not read.
template <int V, class Texts, bool kIdentity>
__device__ __forceinline__ void mergeChunk(const MergeCtx& c, const TieTables& t, const LdsRecords<V, Texts::kBatch>& recs, const uint16_t* src, uint16_t* dst, const uint32_t* bounds, uint32_t runs, uint32_t o, uint32_t o1) {
    const uint32_t pairs = (runs + 1) / 2;
    const uint32_t last = recs.capacity - 1;
    auto at = [&](uint32_t x) -> uint16_t { return kIdentity ? static_cast<uint16_t>(x) : src[x]; };
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
        // The number of A elements among the first d outputs: the first m in [lo, hi] for which
        // A[m] < B[d - 1 - m] is false (hi when there is none).
        uint32_t lo = d > lengthB ? d - lengthB : 0;
        uint32_t hi = min(d, lengthA);
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (lessLds<V, Texts>(c, t, recs, at(aStart + mid), at(aEnd + d - 1 - mid))) {
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
        uint16_t headA = at(min(aStart + ia, last)), headB = at(min(aEnd + ib, last));
        if constexpr (V == 32) {
            // Heads' keys in registers; each step loads the key of the new head and the index
            // after it (one LDS round trip on the dependent path). A head past the end of its run
            // can be an unwritten slot: its key is loaded clamped and never used.
            using Key = HeadKey<Texts::kBatch>;
            Key keyA = headKey<V, Texts::kBatch>(recs, min<uint32_t>(headA, last));
            Key keyB = headKey<V, Texts::kBatch>(recs, min<uint32_t>(headB, last));
            uint16_t nextA = at(min(aStart + ia + 1, last)), nextB = at(min(aEnd + ib + 1, last));
            while (o < stop) {
                bool takeA;
                if (ib >= lengthB) {
                    takeA = true;
                } else if (ia >= lengthA) {
                    takeA = false;
                } else {
                    takeA = lessKeyed<V, Texts>(c, recs, headA, keyA, headB, keyB);
                }
                dst[o++] = takeA ? headA : headB;
                const uint16_t head = takeA ? nextA : nextB;
                const Key key = headKey<V, Texts::kBatch>(recs, min<uint32_t>(head, last));
                const uint16_t next = at(min(takeA ? aStart + ia + 2 : aEnd + ib + 2, last));
                if (takeA) {
                    ia++;
                    headA = head;
                    keyA = key;
                    nextA = next;
                } else {
                    ib++;
                    headB = head;
                    keyB = key;
                    nextB = next;
                }
            }
        } else {
            while (o < stop) {
                const uint16_t nextA = at(min(aStart + ia + 1, last));
                const uint16_t nextB = at(min(aEnd + ib + 1, last));
                bool takeA;
                if (ib >= lengthB) {
                    takeA = true;
                } else if (ia >= lengthA) {
                    takeA = false;
                } else {
                    takeA = lessLds<V, Texts>(c, t, recs, headA, headB);
                }
                dst[o++] = takeA ? headA : headB;
                if (takeA) {
                    ia++;
                    headA = nextA;
                } else {
                    ib++;
                    headB = nextB;
                }
            }
        }
    }
}

// The comparison tables of a tile merge for covers other than 32 (cover 32 computes them).
template <int V>
__device__ __forceinline__ TieTables setupTables() {
    return TieTables{&deviceCover<V>().tie[0][0], &deviceCover<V>().anchor[0]};
}

// Merges the runs of indices in src (run starts bounds[0..runs), bounds[runs] = size) in LDS by
// rounds of pairwise merges; returns the buffer holding the merged order. afterRound() runs after
// every round, between barriers, on every thread. identityFirst: src[i] = i, so the first round
// need not read it. Each thread takes a chunk of at least two consecutive outputs: a tile with
// fewer records than 2 * kThreads leaves the rest of its threads idle, and in tiles of about 340
// records in 256 threads (cover 128 with G = 4 and rank windows in the records) chunks of 2
// merged 5% faster than chunks of 3.
template <int V, class Texts, int kThreads, class AfterRound>
__device__ __forceinline__ uint16_t* mergeRuns(const MergeCtx& c, const TieTables& tables, const LdsRecords<V, Texts::kBatch>& recs, uint16_t* src, uint16_t* dst, uint32_t* bounds, uint32_t runs, uint32_t size, bool identityFirst, AfterRound afterRound) {
    const uint32_t chunk = max(2u, (size + kThreads - 1) / kThreads);
    bool first = true;
    while (runs > 1) {
        for (uint64_t o = static_cast<uint64_t>(threadIdx.x) * chunk; o < size; o += static_cast<uint64_t>(kThreads) * chunk) {
            const uint32_t o0 = static_cast<uint32_t>(o), o1 = static_cast<uint32_t>(min(o + chunk, static_cast<uint64_t>(size)));
            if (first && identityFirst) {
                mergeChunk<V, Texts, true>(c, tables, recs, src, dst, bounds, runs, o0, o1);
            } else {
                mergeChunk<V, Texts, false>(c, tables, recs, src, dst, bounds, runs, o0, o1);
            }
        }
        first = false;
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
        afterRound();
    }
    return src;
}

// Writes a merged tile: order[r] is the record of output row outBase + r.
template <int V, class Texts, int kFlags, int kThreads>
__device__ __forceinline__ void writeTile(const MergeCtx& c, const ListSet& lists, const LdsRecords<V, Texts::kBatch>& recs, const uint16_t* order, const uint8_t* before, const uint32_t* segStart, const uint32_t* rowStart, uint32_t outBase, uint32_t size, const TileOut& out, uint32_t* errors) {
    const int lc = lists.count;
    for (uint32_t r = threadIdx.x; r < size; r += kThreads) {
        const uint32_t e = min<uint32_t>(order[r], recs.capacity - 1);
        const uint32_t p = recs.positionOf(e);
        const uint32_t row = outBase + r;
        uint32_t start = 0, rowBase = 0, segment = 0;
        if constexpr (Texts::kBatch) {
            segment = recs.tagOf(e) >> 9;
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

// The records of a cover-32 tile from the lines, two lanes per record: random gathers cost per
// request, and lanes of one instruction that hit the same line share one. The text lane loads
// the three quads holding T[p - 1 .. p + 31] (bytes q .. q + 32 of the line, q = p mod 32) and
// writes the four text words and the character before; the rank lane loads the quads holding
// the record's seven ranks (dwords 16 + below(q) .. 22 + below(q) of the line, in three quads
// when below(q) mod 4 = 3) and writes them, the position
// and the index. Each lane decodes only what it loaded, in registers. Positions are the tile's
// rows of the SA buffer.
constexpr int kPairSlots = 5;  // records per lane pair in flight: a default tile (<= 1280) in one batch

template <class Texts, int kFlags, int kThreads>
__device__ __forceinline__ void loadRecordsFromLines(const MergeCtx& c, const Texts& texts, const LdsRecords<32, Texts::kBatch>& recs, const uint32_t* __restrict__ rowsIn, uint32_t size, uint8_t* before, uint16_t* indexA, uint32_t* errors) {
    constexpr uint32_t kPairs = kThreads / 2;
    const bool textLane = (threadIdx.x & 1) == 0;
    const uint32_t cap = recs.capacity;
    for (uint32_t e0 = threadIdx.x / 2; e0 < size; e0 += kPairs * kPairSlots) {
        uint32_t pos[kPairSlots];
#pragma unroll
        for (int h = 0; h < kPairSlots; h++) {
            pos[h] = rowsIn[min(e0 + h * kPairs, size - 1)];
        }
        // A batch's rank lane loads all four rank quads: the last holds the line's text and its end.
        uint4 quad[kPairSlots][Texts::kBatch ? 4 : 3];
#pragma unroll
        for (int h = 0; h < kPairSlots; h++) {
            uint32_t p = pos[h];
            if (p >= c.n) {
                if (e0 + h * kPairs < size && textLane) {
                    atomicOr(errors, 2u);
                }
                p = 0;
                pos[h] = 0;
            }
            const uint32_t q = p & (kLinePositions - 1);
            const uint4* line = c.lines + (static_cast<uint64_t>(p / kLinePositions) << 3);
            const uint4* from = textLane ? line + (q >> 4) : line + 4 + (Texts::kBatch ? 0 : coverBelow<32>(q) >> 2);
            quad[h][0] = from[0];
            quad[h][1] = from[1];
            quad[h][2] = from[2];
            if constexpr (Texts::kBatch) {
                quad[h][3] = textLane ? uint4{0, 0, 0, 0} : from[3];
            }
        }
#pragma unroll
        for (int h = 0; h < kPairSlots; h++) {
            const uint32_t e = e0 + h * kPairs;
            if (e >= size) {
                continue;
            }
            const uint32_t p = pos[h];
            const uint32_t q = p & (kLinePositions - 1);
            if (textLane) {
                // T[p] is byte (q mod 16) + 1 of the 48 loaded bytes.
                const uint4 q0 = quad[h][0], q1 = quad[h][1], q2 = quad[h][2];
                const uint64_t w[6] = {q0.x | (uint64_t(q0.y) << 32), q0.z | (uint64_t(q0.w) << 32),
                                       q1.x | (uint64_t(q1.y) << 32), q1.z | (uint64_t(q1.w) << 32),
                                       q2.x | (uint64_t(q2.y) << 32), q2.z | (uint64_t(q2.w) << 32)};
                const uint32_t offset = (q & 15) + 1;
                const uint32_t wordShift = offset >> 3, shift = (offset & 7) * 8;
                uint64_t v[5];
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const uint64_t a1 = k + 1 < 6 ? w[k + 1] : 0;
                    const uint64_t a2 = k + 2 < 6 ? w[k + 2] : 0;
                    v[k] = wordShift == 0 ? w[k] : wordShift == 1 ? a1 : a2;
                }
#pragma unroll
                for (int k = 0; k < kRecordWords; k++) {
                    recs.words[k * cap + e] = __builtin_bswap64(shift ? (v[k] >> shift) | (v[k + 1] << (64 - shift)) : v[k]);
                }
                if constexpr ((kFlags & kOutBWT) != 0) {
                    const uint32_t b = offset - 1;
                    before[e] = p > 0 ? static_cast<uint8_t>(w[(b >> 3) & 1] >> ((b & 7) * 8)) : 0;
                }
            } else if constexpr (Texts::kBatch) {
                // Ranks at words below(q) .. below(q) + 6 of the four quads (below(q) is 0 .. 7);
                // words 14 and 15 are the text and its end.
                const uint32_t o = coverBelow<32>(q);
                uint32_t x[16];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    x[4 * k] = quad[h][k].x;
                    x[4 * k + 1] = quad[h][k].y;
                    x[4 * k + 2] = quad[h][k].z;
                    x[4 * k + 3] = quad[h][k].w;
                }
#pragma unroll
                for (int k = 0; k < CoverSet<32>::kSize; k++) {
                    uint32_t r = x[k];
#pragma unroll
                    for (uint32_t j = 1; j <= CoverSet<32>::kSize; j++) {
                        r = o == j ? x[k + j] : r;
                    }
                    recs.ranks[k * cap + e] = r;
                }
                recs.setPosition(e, p);
                recs.setTag(e, (x[14] << 9) | min(x[15] - p, 32u));
                indexA[e] = static_cast<uint16_t>(e);
            } else {
                const uint32_t o = coverBelow<32>(q) & 3;
                const uint4 r0 = quad[h][0], r1 = quad[h][1], r2 = quad[h][2];
                const uint32_t x[12] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z, r2.w};
#pragma unroll
                for (int k = 0; k < CoverSet<32>::kSize; k++) {
                    recs.ranks[k * cap + e] = o == 0 ? x[k] : o == 1 ? x[k + 1] : o == 2 ? x[k + 2] : x[k + 3];
                }
                recs.setPosition(e, p);
                indexA[e] = static_cast<uint16_t>(e);
            }
        }
    }
}

// One tile per workgroup: load its records into LDS, merge, write.
template <int V, class Texts, int kFlags, int kThreads>
__global__ __launch_bounds__(kThreads) void mergeTilesKernel(MergeCtx c, Texts texts, ListSet lists, RowTable rows, uint32_t firstTile, uint32_t capacity, TileOut out, uint32_t* __restrict__ errors) {
    using Record = MergeRecord<V, Texts::kBatch>;
    extern __shared__ __align__(16) unsigned char shared[];
    LdsRecords<V, Texts::kBatch> recs;
    recs.carve(shared, capacity);
    uint16_t* indexA = reinterpret_cast<uint16_t*>(shared + static_cast<size_t>(capacity) * LdsRecords<V, Texts::kBatch>::bytesPerRecord());
    uint16_t* indexB = indexA + capacity;
    uint8_t* before = reinterpret_cast<uint8_t*>(indexB + capacity);
    __shared__ uint32_t bounds[kMaxLists + 1];
    __shared__ uint32_t segStart[kMaxLists + 1];
    __shared__ uint32_t rowStart[kMaxLists];
    __shared__ uint32_t outBase, valid;
    __shared__ const uint32_t* listPtr[kMaxLists];  // indexed per lane: LDS, not kernel-argument loads

    const int lc = lists.count;
    const uint64_t tile = static_cast<uint64_t>(firstTile) + blockIdx.x;
    TieTables tables = setupTables<V>();
    if constexpr (V == 128) {
        __shared__ uint8_t anchorLds[V];
        if (threadIdx.x < static_cast<uint32_t>(V)) {
            anchorLds[threadIdx.x] = tables.anchor[threadIdx.x];
        }
        tables.anchor = anchorLds;  // read after the barrier below
    }

    if (threadIdx.x < static_cast<uint32_t>(lc)) {
        listPtr[threadIdx.x] = lists.ptr[threadIdx.x];
    }
    tileLayout(lists, rows, tile, capacity, segStart, rowStart, &outBase, &valid);
    __syncthreads();
    if (!valid) {
        if (threadIdx.x == 0) {
            atomicOr(errors, 1u);
        }
        return;
    }
    if (threadIdx.x <= static_cast<uint32_t>(lc)) {
        bounds[threadIdx.x] = segStart[threadIdx.x];
    }
    const uint32_t size = segStart[lc];

    bool fromLines = false;
    if constexpr (V == kLinePositions && (kFlags & kInPlace) != 0) {
        fromLines = c.lines != nullptr;
        if (fromLines) {
            loadRecordsFromLines<Texts, kFlags, kThreads>(c, texts, recs, out.sa + outBase, size, before, indexA, errors);
        }
    }
    // Otherwise kLoadSlots records per thread at a time: every slot's position first, then every
    // slot's record gathers, so a tile of up to kLoadSlots * kThreads records costs two dependent
    // round trips (straight-line code, clamped indices); the BWT character before each record too.
    if (!fromLines) {
        for (uint32_t e0 = threadIdx.x; e0 < size; e0 += kLoadSlots * kThreads) {
            uint32_t pos[kLoadSlots];
            int l = 0;
#pragma unroll
            for (int h = 0; h < kLoadSlots; h++) {
                const uint32_t e = min(e0 + h * kThreads, size - 1);
                if constexpr ((kFlags & kInPlace) != 0) {
                    pos[h] = out.sa[outBase + e];
                } else {
                    while (l + 1 < lc && e >= segStart[l + 1]) {
                        l++;
                    }
                    pos[h] = listPtr[l][rowStart[l] + (e - segStart[l])];
                }
            }
            RawRecord<V> raw[kLoadSlots];
#pragma unroll
            for (int h = 0; h < kLoadSlots; h++) {
                if (pos[h] >= c.n) {
                    if (e0 + h * kThreads < size) {
                        atomicOr(errors, 2u);
                    }
                    pos[h] = 0;
                }
                fetchRecord<V>(c, pos[h], coverBelow<V>(pos[h] & (V - 1)), raw[h]);
            }
#pragma unroll
            for (int h = 0; h < kLoadSlots; h++) {
                const uint32_t e = e0 + h * kThreads;
                if (e < size) {
                    Record r;
                    uint8_t characterBefore;
                    decodeRecord<V, Texts>(c, texts, pos[h], coverBelow<V>(pos[h] & (V - 1)), raw[h], r, &characterBefore);
                    recs.store(e, r);
                    if constexpr ((kFlags & kOutBWT) != 0) {
                        before[e] = characterBefore;
                    }
                    indexA[e] = static_cast<uint16_t>(e);
                }
            }
        }
    }
    __syncthreads();

    const uint16_t* order = mergeRuns<V, Texts, kThreads>(c, tables, recs, indexA, indexB, bounds, static_cast<uint32_t>(lc), size, true, [] {});
    writeTile<V, Texts, kFlags, kThreads>(c, lists, recs, order, before, segStart, rowStart, outBase, size, out, errors);
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
