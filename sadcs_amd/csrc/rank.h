/*
 * rank.h - Phase (a): ranking the sample suffixes.
 *
 * Follows the reference's scheme of passes over a shrinking list of unresolved sample suffixes
 * (src/cuda/samplerank.cuh:1-20, :314-355): every suffix carries a group id, the number of sample
 * suffixes known to be smaller than every member of its group; a pass splits groups by a subkey;
 * a new group's id is the old id plus its offset in the old group; groups of one are final and
 * leave the list. Character passes run until every group shares at least V characters, then
 * prefix-doubling rounds use the group id at p + h (Larsson-Sadakane without in-round refinement,
 * src/cuda/samplerank.cuh:15-18) until nothing is left.
 *
 * What differs, and why it suits gfx950:
 *  - One sort per pass on a composite 64-bit key (dense group ordinal, subkey) instead of a
 *    subkey sort, a gather of group ids and a second sort. The ordinal is the group's index in
 *    the list, so it needs only bit_width(groups - 1) bits.
 *  - Characters are packed by a dense code: code 0 is "past the end", the bytes present get
 *    codes 1..sigma in byte order. Shorter suffixes then sort first with no length tag, and a
 *    64-bit key holds 21 DNA characters, 9 of 96 printable ones, or 7 arbitrary bytes.
 *  - Marking, id assignment, rank stores, the sorted-sample output and the compaction are one
 *    scan pipeline (four fields: last head, group start, unresolved count, unresolved groups).
 *  - The sorted sample positions are written as suffixes resolve (mostly coalesced, since
 *    resolved ids rise along the list). That array is class 0 of the merge input.
 *  - Segments: a batch ranks all its texts at once; pass 0 starts with one group per text.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "cover.h"
#include "dev_common.h"
#include "prim.h"

namespace sadcs {

template <int V>
struct DeviceCover;  // __constant__ tables, defined by the engine

// Byte -> dense character code (0 = past the end).
struct CodeTable {
    uint8_t code[256];
};

struct RankParams {
    const uint64_t* text;    // padded with zero words
    uint32_t* ranks;         // compact, indexed by pack(p)
    uint32_t* sortedSample;  // ranks -> positions, filled as suffixes resolve
    uint32_t samples;        // valid sample suffixes (all texts)
    uint32_t segments;
    const Segment* segmentTable;
    // A batch's first pass: per text, dense codes of its own bytes (256 per text) and its code
    // width and characters per key (2 per text); null otherwise.
    const uint8_t* textCodes;
    const uint8_t* textShape;
};

// Subkey of one pass: characters at depth `depth`, or the group id at p + h plus one.
struct PassSpec {
    bool doubling;
    uint32_t offset;    // depth (characters) or h (doubling)
    int chars;          // characters in a character subkey
    int charBits;       // bits per character code
    int subkeyBits;
};

template <int V>
__device__ __forceinline__ uint32_t packPosition(uint32_t p);

template <int V>
__device__ __forceinline__ uint32_t unpackIndex(uint32_t t);

// Pass 0: the i-th valid sample suffix of the batch, and its text's index.
// The text of sample rank i in a batch: the last with rankBase <= i.
__device__ __forceinline__ uint32_t textOfRank(const RankParams& rp, uint32_t i) {
    uint32_t lo = 0, hi = rp.segments;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (rp.segmentTable[mid].rankBase <= i) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    return lo;
}

// The first pass's element i: its compact index t and its text. blockText/blockTextEnd: the text
// of the workgroup's first element and the rank where the next text starts; elements before that
// skip the search.
template <int V, class Texts>
__device__ __forceinline__ void firstPassEntry(const RankParams& rp, uint32_t i, uint32_t& t, uint32_t& segment, uint32_t blockText, uint32_t blockTextEnd) {
    if constexpr (!Texts::kBatch) {
        t = i;
        segment = 0;
    } else {
        segment = i < blockTextEnd ? blockText : textOfRank(rp, i);
        const Segment s = rp.segmentTable[segment];
        t = packPosition<V>(s.start) + (i - s.rankBase);
    }
}

// Keys of one pass: key = (ordinal << subkeyBits) | subkey, value = compact index t (not written
// when valuesOut is null: the first pass of a single text, whose t is the index). digitsOut, if
// not null, gets the key's low byte, the first digit of the radix sort.
// kWideCodes: all 256 byte values occur, the code is the byte plus one (9 bits).
template <int V, class Texts, bool kWideCodes>
__global__ __launch_bounds__(256) void makePassKeysKernel(RankParams rp, Texts texts, PassSpec spec, CodeTable codes, bool first, const uint32_t* __restrict__ listT, const uint32_t* __restrict__ listOrdinal, uint32_t u, uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valuesOut, uint8_t* __restrict__ digitsOut) {
    __shared__ uint8_t code[256];
    __shared__ uint8_t blockCodes[256];  // a batch's first pass: the codes of the workgroup's first text
    __shared__ uint32_t blockText, blockTextEnd;
    for (uint32_t x = threadIdx.x; x < 256; x += blockDim.x) {
        code[x] = codes.code[x];
    }
    if constexpr (Texts::kBatch) {
        if (first) {
            // The workgroup's elements are consecutive ranks, mostly of one text: one search.
            if (threadIdx.x == 0) {
                const uint32_t text = textOfRank(rp, min(blockIdx.x * blockDim.x, u - 1));
                blockText = text;
                blockTextEnd = text + 1 < rp.segments ? rp.segmentTable[text + 1].rankBase : ~0u;
            }
            __syncthreads();
            if (rp.textCodes) {
                for (uint32_t x = threadIdx.x; x < 256; x += blockDim.x) {
                    blockCodes[x] = rp.textCodes[static_cast<size_t>(blockText) * 256 + x];
                }
            }
        }
    }
    __syncthreads();

    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= u) {
        return;
    }

    uint32_t t, ordinal;
    if (first) {
        firstPassEntry<V, Texts>(rp, i, t, ordinal, Texts::kBatch ? blockText : 0u, Texts::kBatch ? blockTextEnd : 0u);
    } else {
        t = listT[i];
        ordinal = listOrdinal[i];
    }
    const uint32_t p = unpackIndex<V>(t);
    const uint32_t end = texts.endOf(p);
    uint64_t subkey = 0;

    if (spec.doubling) {
        const uint32_t q = p + spec.offset;
        subkey = q < end ? uint64_t(rp.ranks[packPosition<V>(q)]) + 1 : 0;
    } else if (first && rp.textCodes) {
        // Texts of a batch never compare, so each packs its own codes as densely as they allow
        // (a DNA text 3 bits per character among texts of 96 symbols), left-aligned in the
        // subkey. The ordinal of a first-pass element is its text.
        // The workgroup's first text reads its codes from LDS, others from the global table.
        const uint8_t* table = ordinal == blockText ? blockCodes : rp.textCodes + static_cast<size_t>(ordinal) * 256;
        const int bits = rp.textShape[2 * ordinal], chars = rp.textShape[2 * ordinal + 1];
        const uint32_t q = p + spec.offset;
        const uint32_t remaining = q < end ? end - q : 0;
        // Eight characters per text word, their code lookups independent of each other.
        for (int c0 = 0; c0 < chars; c0 += 8) {
            uint64_t w = static_cast<uint32_t>(c0) < remaining ? load8(rp.text, q + c0) : 0;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = c0 + j;
                const uint32_t byte = static_cast<uint32_t>(w >> 56);
                w <<= 8;
                const uint32_t ch = static_cast<uint32_t>(c) < remaining ? table[byte] : 0u;
                subkey = c < chars ? (subkey << bits) | ch : subkey;
            }
        }
        subkey <<= spec.subkeyBits - chars * bits;
    } else {
        const uint32_t q = p + spec.offset;
        const uint32_t remaining = q < end ? end - q : 0;
        const int chars = spec.chars;
        uint64_t w = 0;
        for (int c = 0; c < chars; c++) {
            if ((c & 7) == 0) {
                w = static_cast<uint32_t>(c) < remaining ? load8(rp.text, q + c) : 0;
            }
            const uint32_t byte = static_cast<uint32_t>(w >> 56);
            w <<= 8;
            const uint32_t ch = static_cast<uint32_t>(c) < remaining ? (kWideCodes ? byte + 1 : code[byte]) : 0;
            subkey = (subkey << spec.charBits) | ch;
        }
    }

    const uint64_t key = (spec.subkeyBits < 64 ? (uint64_t(ordinal) << spec.subkeyBits) : 0) | subkey;
    keysOut[i] = key;
    if (valuesOut) {
        valuesOut[i] = t;
    }
    if (digitsOut) {
        digitsOut[i] = static_cast<uint8_t>(key);
    }
}

// Per text of a batch: the bytes that occur in it (8 words per text, zeroed by the caller), for
// codes dense within each text. Thread t of a workgroup takes 256-byte block t of the workgroup's
// 64 KiB (a text starts on a block boundary, so the block is one text's or a gap's) into its own
// LDS bitmap; the first thread of each run of threads on one text combines the run's bitmaps and
// publishes them.
__global__ __launch_bounds__(256) void textPresenceKernel(const uint4* __restrict__ chunks, uint32_t span, TextBatch texts, uint32_t* __restrict__ textMasks) {
    constexpr uint32_t kStride = 9;
    __shared__ uint32_t bits[256 * kStride];
    __shared__ uint32_t runText[256];
    uint32_t* mine = bits + threadIdx.x * kStride;
    for (uint32_t x = 0; x < 8; x++) {
        mine[x] = 0;
    }
    const uint64_t p0 = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) << kSegmentBlockLog;
    uint32_t text = ~0u;
    if (p0 < span) {
        const uint32_t segment = texts.segmentOf(static_cast<uint32_t>(p0));
        const Segment s = texts.segments[segment];
        const uint32_t end = min(s.end, span);
        if (p0 >= s.start && p0 < end) {
            text = segment;
            const uint32_t valid = min(1u << kSegmentBlockLog, end - static_cast<uint32_t>(p0));
            for (uint32_t k = 0; k * 16 < valid; k++) {
                const uint4 q = chunks[(p0 >> 4) + k];
                const uint32_t words[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (uint32_t j = 0; j < 16; j++) {
                    const uint32_t b = (words[j >> 2] >> ((j & 3) * 8)) & 255u;
                    if (k * 16 + j < valid) {
                        atomicOr(&mine[b >> 5], 1u << (b & 31));
                    }
                }
            }
        }
    }
    runText[threadIdx.x] = text;
    __syncthreads();
    if (text != ~0u && (threadIdx.x == 0 || runText[threadIdx.x - 1] != text)) {
        uint32_t m[8];
        for (uint32_t x = 0; x < 8; x++) {
            m[x] = mine[x];
        }
        for (uint32_t t = threadIdx.x + 1; t < 256 && runText[t] == text; t++) {
            for (uint32_t x = 0; x < 8; x++) {
                m[x] |= bits[t * kStride + x];
            }
        }
        for (uint32_t x = 0; x < 8; x++) {
            if (m[x]) {
                atomicOr(&textMasks[static_cast<size_t>(text) * 8 + x], m[x]);
            }
        }
    }
}

// Per text: codes 1..sigma of its bytes in byte order (0 for absent bytes; sigma < 256), the code
// width and the characters that fit subkeyBits.
__global__ void textCodesKernel(const uint32_t* __restrict__ textMasks, uint32_t texts, int subkeyBits, uint8_t* __restrict__ textCodes, uint8_t* __restrict__ textShape) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= texts) {
        return;
    }
    uint32_t count = 0;
    for (uint32_t b = 0; b < 256; b++) {
        const uint32_t present = (textMasks[static_cast<size_t>(s) * 8 + b / 32] >> (b % 32)) & 1u;
        count += present;
        textCodes[static_cast<size_t>(s) * 256 + b] = static_cast<uint8_t>(present ? count : 0);
    }
    const int bits = count > 1 ? 32 - __clz(count) : 1;
    textShape[2 * s] = static_cast<uint8_t>(bits);
    textShape[2 * s + 1] = static_cast<uint8_t>(max(1, subkeyBits / bits));
}

// The sort of a pass after the first, when its groups are small. The list arrives grouped: keys
// are (ordinal << subkeyBits) | subkey with ordinals non-decreasing along the list (the previous
// pass compacted its unresolved groups in sorted order). The stable sort of such a list puts
// element i at its group's start plus the members with a smaller key, plus those with an equal key
// before i: the same order as the stable radix sort, with each element reading only its group.
// Unresolved groups are mostly pairs (a repeat and its source), so this replaces the several
// digit passes of a 50- to 64-bit radix sort. An element of a group of more than kGroupRankMax
// elements is not placed: it sets largeFlags[i] (0 otherwise) and counts itself in *largeCount,
// and the caller sorts those elements apart (or the whole pass by radix). Ordinals do not decrease
// along the list, so one read kGroupRankMax away shows whether the group is larger than that:
// elements of a huge group (a long run) stop at once instead of scanning.
constexpr uint32_t kGroupRankMax = 256;

__global__ __launch_bounds__(256) void groupRankSortKernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ values, uint32_t u, int subkeyBits, uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valuesOut, uint8_t* __restrict__ largeFlags, uint32_t* __restrict__ largeCount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool large = false;
    if (i < u) {
        const uint64_t key = keys[i];
        const uint64_t ordinal = key >> subkeyBits;
        large = (i >= kGroupRankMax && (keys[i - kGroupRankMax] >> subkeyBits) == ordinal) ||
                (i + kGroupRankMax < u && (keys[i + kGroupRankMax] >> subkeyBits) == ordinal);
        if (!large) {
            uint32_t start = i;
            for (uint32_t step = 0; step < kGroupRankMax && start > 0 && (keys[start - 1] >> subkeyBits) == ordinal; step++) {
                start--;
            }
            uint32_t end = i + 1;
            for (uint32_t step = 0; step < kGroupRankMax && end < u && (keys[end] >> subkeyBits) == ordinal; step++) {
                end++;
            }
            const bool leftOpen = start > 0 && (keys[start - 1] >> subkeyBits) == ordinal;
            const bool rightOpen = end < u && (keys[end] >> subkeyBits) == ordinal;
            large = leftOpen || rightOpen || end - start > kGroupRankMax;
            if (!large) {
                uint32_t rank = start;
                for (uint32_t j = start; j < end; j++) {
                    const uint64_t other = keys[j];
                    rank += (other < key || (other == key && j < i)) ? 1u : 0u;
                }
                keysOut[rank] = key;
                valuesOut[rank] = values[i];
            }
        }
        largeFlags[i] = large ? 1 : 0;
    }
    const uint64_t wave = ballot(large);
    if (wave != 0 && laneId() == static_cast<uint32_t>(__builtin_ctzll(wave))) {
        atomicAdd(largeCount, static_cast<uint32_t>(__builtin_popcountll(wave)));
    }
}

// The elements of large groups, in list order: subKeys/subValues to be sorted, and subPos, the
// list positions they came from (ascending). Sorted stably by their keys, the k-th of them belongs
// at subPos[k]: the large groups keep their order (ordinals ascend along the list) and each
// occupies a contiguous range of positions.
struct LargeFlagSource {
    const uint8_t* flags;
    __device__ Vec<1, 0u> operator()(uint32_t i) const {
        Vec<1, 0u> r;
        r.v[0] = flags[i];
        return r;
    }
};

struct LargeGroupCompact {
    const uint8_t* flags;
    const uint64_t* keys;
    const uint32_t* values;
    uint64_t* subKeys;
    uint32_t* subValues;
    uint32_t* subPos;
    struct Prepared {};
    __device__ Prepared prepare(uint32_t) const {
        return {};
    }
    __device__ void complete(Prepared&) const {}
    __device__ void operator()(uint32_t i, const Prepared&, const Vec<1, 0u>& item, const Vec<1, 0u>& exclusive, bool valid) const {
        if (valid && item.v[0] != 0) {
            const uint32_t k = exclusive.v[0];
            subKeys[k] = keys[i];
            subValues[k] = values[i];
            subPos[k] = i;
        }
    }
};

__global__ void placeLargeGroupsKernel(const uint64_t* __restrict__ subKeys, const uint32_t* __restrict__ subValues, const uint32_t* __restrict__ subPos, uint32_t count, uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valuesOut) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        const uint32_t at = subPos[k];
        keysOut[at] = subKeys[k];
        valuesOut[at] = subValues[k];
    }
}

// Four fields: last head (max), group start (max), unresolved before (sum), unresolved groups before (sum).
using MarkVec = Vec<4, 0x3u>;
constexpr int kMarkItems = 8;

struct MarkState {
    const uint64_t* keys;      // sorted composite keys
    const uint32_t* values;    // compact indices in the same order
    uint32_t u;
    int subkeyBits;
    const uint32_t* groupIds;  // ordinal -> group id
    // outputs
    uint32_t* ranks;
    uint32_t* sortedSample;
    uint32_t* nextT;
    uint32_t* nextOrdinal;
    uint32_t* nextGroupIds;
    uint32_t* counts;          // [0] unresolved left, [1] groups left
    uint32_t samples;          // entries of sortedSample
    bool deferResolved;        // resolved elements' rank stores are left to scatterResolvedRanksKernel
};

struct MarkFlags {
    bool head, groupStart, resolved;
};

// Branch-free: the three loads of an item are issued together with those of the other items.
__device__ __forceinline__ MarkFlags markFlags(const MarkState& s, uint32_t i) {
    const uint64_t key = s.keys[i];
    const uint64_t previous = s.keys[i > 0 ? i - 1 : 0];
    const uint64_t next = s.keys[i + 1 < s.u ? i + 1 : i];
    // Bitwise operators: a short-circuit would let the compiler move a load under a branch.
    const bool first = i == 0, last = i + 1 == s.u;
    const int shift = s.subkeyBits < 64 ? s.subkeyBits : 0;
    const uint64_t groupMask = s.subkeyBits < 64 ? ~uint64_t(0) : 0;
    MarkFlags f;
    f.head = first | (previous != key);
    f.groupStart = first | ((((previous >> shift) ^ (key >> shift)) & groupMask) != 0);
    f.resolved = f.head & (last | (next != key));
    return f;
}

struct MarkProducer {
    MarkState s;
    __device__ MarkVec operator()(uint32_t i) const {
        const MarkFlags f = markFlags(s, i);
        MarkVec r;
        r.v[0] = f.head ? i : 0;
        r.v[1] = f.groupStart ? i : 0;
        r.v[2] = f.resolved ? 0 : 1;
        r.v[3] = (!f.resolved && f.head) ? 1 : 0;
        return r;
    }
};

template <int V>
struct MarkConsumer {
    MarkState s;
    struct Prepared {
        uint64_t key;
        uint32_t t, groupId;
    };
    // The loads of item i (any index in range): its key and value, then its group's id.
    __device__ Prepared prepare(uint32_t i) const {
        return Prepared{s.keys[i], s.values[i], 0};
    }
    __device__ void complete(Prepared& p) const {
        p.groupId = s.groupIds[s.subkeyBits >= 64 ? 0 : static_cast<uint32_t>(p.key >> s.subkeyBits)];
    }
    // The stores of item i, for valid items only.
    __device__ void operator()(uint32_t i, const Prepared& p, const MarkVec& item, const MarkVec& exclusive, bool valid) const {
        const uint32_t lastHead = max(exclusive.v[0], item.v[0]);
        const uint32_t groupStart = max(exclusive.v[1], item.v[1]);
        const uint32_t id = p.groupId + (lastHead - groupStart);
        const uint32_t t = p.t;
        if (!valid) {
            return;
        }

        if (!(s.deferResolved && item.v[2] == 0)) {
            s.ranks[t] = id;
        }
        if (item.v[2] == 0) {
            if (id < s.samples) {
                s.sortedSample[id] = unpackIndex<V>(t);
            }
        } else {
            const uint32_t slot = exclusive.v[2];
            s.nextT[slot] = t;
            s.nextOrdinal[slot] = exclusive.v[3] + item.v[3] - 1;
            if (item.v[3]) {
                s.nextGroupIds[exclusive.v[3]] = id;
            }
        }
        if (i + 1 == s.u) {
            s.counts[0] = exclusive.v[2] + item.v[2];
            s.counts[1] = exclusive.v[3] + item.v[3];
        }
    }
};

// The deferred rank stores of one pass: ranks[pack(sortedSample[i])] = i for every entry written
// in that pass (entries still holding ~0 are skipped). The mark scan meets resolved elements in
// rank order, so their rank stores would scatter 4-byte writes over the whole rank array, each a
// read-modify-write of a memory line; here one launch writes only the ranks in [first, last), a
// range small enough for the partial lines to merge in the memory-side cache, and the launches
// over consecutive ranges read the (coalesced) sorted sample again each time. Each thread loads
// kItems entries before storing any: one load and store at a time left the launch latency-bound.
template <int V>
__global__ __launch_bounds__(256) void scatterResolvedRanksKernel(const uint32_t* __restrict__ sortedSample, uint32_t samples, uint32_t first, uint32_t last, uint32_t* __restrict__ ranks) {
    constexpr uint32_t kItems = 16;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i0 < samples; i0 += stride * kItems) {
        uint32_t p[kItems];
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            const uint64_t i = i0 + k * stride;
            p[k] = i < samples ? sortedSample[i] : ~0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kItems; k++) {
            if (p[k] != ~0u) {
                const uint32_t t = packPosition<V>(p[k]);
                if (t >= first && t < last) {
                    ranks[t] = static_cast<uint32_t>(i0 + k * stride);
                }
            }
        }
    }
}

// ---- the last passes in one workgroup ----------------------------------------------------------
//
// Once few suffixes are unresolved, a pass costs a dozen small launches and a host round trip to
// learn the next list's size. rankTailKernel runs every remaining pass in one workgroup instead:
// the list, the group ids, the keys and the scans live in LDS, the workgroup loops until the list
// is empty, and the host waits once. A pass is the same as above - the keys of makePassKeysKernel,
// a sort by key, the mark scan of MarkConsumer - so ids, ranks and the sorted sample come out the
// same. The sort is a bitonic network on (key, list index); ties among equal keys only reorder
// members of one unresolved group, whose id they share.
constexpr uint32_t kTailMax = 2048;  // list entries the workgroup takes
constexpr int kTailThreads = 1024;
constexpr int kTailItems = kTailMax / kTailThreads;  // consecutive sorted entries per thread in the scan

struct TailParams {
    RankParams rp;
    CodeTable codes;
    int charBits;         // bits per character code; 9: all 256 bytes occur (code = byte + 1)
    int rankBits;         // bits of a doubling subkey (group id + 1)
    uint32_t maxLength;   // longest text: a doubling step as long is an internal error
    uint64_t depth;       // characters every group shares
    const uint32_t* listT;
    const uint32_t* listOrdinal;
    const uint32_t* groupIds;
    uint32_t u, groups;
    uint32_t* status;     // out: character passes, doubling rounds, error (1: doubling beyond maxLength)
};

inline size_t rankTailSharedBytes() {
    return static_cast<size_t>(kTailMax) * (sizeof(uint64_t) + 6 * sizeof(uint32_t) + sizeof(uint16_t));
}

template <int V, class Texts>
__global__ __launch_bounds__(kTailThreads) void rankTailKernel(TailParams tp, Texts texts) {
    extern __shared__ __align__(16) unsigned char shared[];
    uint64_t* key = reinterpret_cast<uint64_t*>(shared);
    uint32_t* listT[2] = {reinterpret_cast<uint32_t*>(key + kTailMax), nullptr};
    listT[1] = listT[0] + kTailMax;
    uint32_t* listOrdinal[2] = {listT[1] + kTailMax, listT[1] + 2 * kTailMax};
    uint32_t* groupId[2] = {listOrdinal[1] + kTailMax, listOrdinal[1] + 2 * kTailMax};
    uint16_t* index = reinterpret_cast<uint16_t*>(groupId[1] + kTailMax);
    __shared__ uint8_t code[256];
    __shared__ MarkVec waveTotal[kTailThreads / kWave];

    const uint32_t tid = threadIdx.x;
    const RankParams& rp = tp.rp;
    for (uint32_t x = tid; x < 256; x += kTailThreads) {
        code[x] = tp.codes.code[x];
    }
    uint32_t u = min(tp.u, kTailMax), groups = min(tp.groups, kTailMax);
    for (uint32_t i = tid; i < u; i += kTailThreads) {
        listT[0][i] = tp.listT[i];
        listOrdinal[0][i] = tp.listOrdinal[i];
    }
    for (uint32_t g = tid; g < groups; g += kTailThreads) {
        groupId[0][g] = tp.groupIds[g];
    }
    __syncthreads();

    uint64_t depth = tp.depth;
    uint32_t characterPasses = 0, doublingRounds = 0, error = 0;
    int cur = 0;
    while (u > 0) {
        // The pass, as the host would choose it.
        const int ordinalBits = groups > 1 ? 32 - __clz(groups - 1) : 0;
        bool doubling;
        uint32_t offset;
        int chars = 0, subkeyBits;
        if (depth < static_cast<uint64_t>(V)) {
            doubling = false;
            offset = static_cast<uint32_t>(depth);
            chars = max(1, (64 - ordinalBits) / tp.charBits);
            subkeyBits = chars * tp.charBits;
        } else {
            const uint64_t h = depth / V * V;
            if (h >= tp.maxLength) {
                error = 1;
                break;
            }
            doubling = true;
            offset = static_cast<uint32_t>(h);
            subkeyBits = tp.rankBits;
        }

        // Keys (makePassKeysKernel after the first pass); entries past u sort last.
        uint32_t width = 1;
        while (width < u) {
            width <<= 1;
        }
        for (uint32_t i = tid; i < width; i += kTailThreads) {
            if (i < u) {
                const uint32_t t = listT[cur][i];
                const uint32_t p = unpackIndex<V>(t);
                const uint32_t end = texts.endOf(p);
                const uint32_t q = p + offset;
                uint64_t subkey = 0;
                if (doubling) {
                    subkey = q < end ? uint64_t(rp.ranks[packPosition<V>(q)]) + 1 : 0;
                } else {
                    const uint32_t remaining = q < end ? end - q : 0;
                    uint64_t w = 0;
                    for (int c = 0; c < chars; c++) {
                        if ((c & 7) == 0) {
                            w = static_cast<uint32_t>(c) < remaining ? load8(rp.text, q + c) : 0;
                        }
                        const uint32_t byte = static_cast<uint32_t>(w >> 56);
                        w <<= 8;
                        const uint32_t ch = static_cast<uint32_t>(c) < remaining ? (tp.charBits == 9 ? byte + 1 : code[byte]) : 0;
                        subkey = (subkey << tp.charBits) | ch;
                    }
                }
                key[i] = (subkeyBits < 64 ? (uint64_t(listOrdinal[cur][i]) << subkeyBits) : 0) | subkey;
                index[i] = static_cast<uint16_t>(i);
            } else {
                key[i] = ~uint64_t(0);
                index[i] = 0xFFFF;
            }
        }
        __syncthreads();

        // Bitonic sort of (key, index), ascending.
        for (uint32_t k = 2; k <= width; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = tid; i < width; i += kTailThreads) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const uint64_t a = key[i], b = key[l];
                        const uint16_t ia = index[i], ib = index[l];
                        const bool greater = a > b || (a == b && ia > ib);
                        if (greater == ((i & k) == 0)) {
                            key[i] = b;
                            key[l] = a;
                            index[i] = ib;
                            index[l] = ia;
                        }
                    }
                }
                __syncthreads();
            }
        }

        // The mark scan (MarkProducer/MarkConsumer) over sorted entries tid * kTailItems + 0..3.
        const int shift = subkeyBits < 64 ? subkeyBits : 0;
        const uint64_t groupMask = subkeyBits < 64 ? ~uint64_t(0) : 0;
        MarkVec item[kTailItems];
        MarkVec running = MarkVec::zero();
#pragma unroll
        for (int j = 0; j < kTailItems; j++) {
            const uint32_t i = tid * kTailItems + j;
            MarkVec r = MarkVec::zero();
            if (i < u) {
                const uint64_t k0 = key[i];
                const uint64_t previous = key[i > 0 ? i - 1 : 0];
                const uint64_t next = key[i + 1 < u ? i + 1 : i];
                const bool first = i == 0, last = i + 1 == u;
                const bool head = first | (previous != k0);
                const bool groupStart = first | ((((previous >> shift) ^ (k0 >> shift)) & groupMask) != 0);
                const bool resolved = head & (last | (next != k0));
                r.v[0] = head ? i : 0;
                r.v[1] = groupStart ? i : 0;
                r.v[2] = resolved ? 0 : 1;
                r.v[3] = (!resolved && head) ? 1 : 0;
            }
            item[j] = r;
            running = MarkVec::combine(running, r);
        }
        // Exclusive prefix of this thread's total over the workgroup.
        const MarkVec inclusive = MarkVec::waveInclusive(running);
        const uint32_t lane = tid % kWave, wave = tid / kWave;
        if (lane == kWave - 1) {
            waveTotal[wave] = inclusive;
        }
        __syncthreads();
        MarkVec before = MarkVec::shiftUp(inclusive);
        MarkVec all = MarkVec::zero();
        for (uint32_t w = 0; w < kTailThreads / kWave; w++) {
            if (w < wave) {
                before = MarkVec::combine(waveTotal[w], before);
            }
            all = MarkVec::combine(all, waveTotal[w]);
        }
        const int next = 1 - cur;
#pragma unroll
        for (int j = 0; j < kTailItems; j++) {
            const uint32_t i = tid * kTailItems + j;
            const MarkVec it = item[j];
            if (i < u) {
                const uint64_t k0 = key[i];
                const uint32_t ordinal = subkeyBits >= 64 ? 0u : static_cast<uint32_t>(k0 >> subkeyBits);
                const uint32_t lastHead = max(before.v[0], it.v[0]);
                const uint32_t groupStart = max(before.v[1], it.v[1]);
                const uint32_t id = groupId[cur][ordinal] + (lastHead - groupStart);
                const uint32_t t = listT[cur][index[i]];
                rp.ranks[t] = id;
                if (it.v[2] == 0) {
                    if (id < rp.samples) {
                        rp.sortedSample[id] = unpackIndex<V>(t);
                    }
                } else {
                    const uint32_t slot = before.v[2];
                    listT[next][slot] = t;
                    listOrdinal[next][slot] = before.v[3] + it.v[3] - 1;
                    if (it.v[3]) {
                        groupId[next][before.v[3]] = id;
                    }
                }
            }
            before = MarkVec::combine(before, it);
        }
        u = all.v[2];
        groups = all.v[3];
        cur = next;
        if (doubling) {
            depth += offset;
            doublingRounds++;
        } else {
            depth += static_cast<uint64_t>(chars);
            characterPasses++;
        }
        __syncthreads();
    }
    if (tid == 0) {
        tp.status[0] = characterPasses;
        tp.status[1] = doublingRounds;
        tp.status[2] = error;
    }
}

__global__ void initGroupIdsKernel(const Segment* __restrict__ segments, uint32_t count, uint32_t* __restrict__ groupIds) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < count) {
        groupIds[s] = segments[s].rankBase;
    }
}

}  // namespace sadcs
