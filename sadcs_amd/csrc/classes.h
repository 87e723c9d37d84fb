/*
 * classes.h - Phase (b): the sorted list of every class of non-sample suffixes.
 *
 * Class d holds the positions whose next sample position is d characters ahead. For i in class
 * d >= 1, i + 1 is in class d - 1, so suffix i is ordered by (T[i], suffix i + 1) and class d is
 * class d - 1 walked in order, restricted to positions whose predecessor is in class d, stably
 * sorted by that predecessor's character (src/cuda/classsort.cuh:1-20, :184-244). The walk starts
 * with the text's end n, standing for the empty suffix: position n - 1 then precedes every other
 * suffix starting with T[n - 1] (src/cuda/classsort.cuh:14-17, :77-80). Class 0 is the sample in
 * rank order, written by phase (a).
 *
 * Each element carries four characters T[q], T[q-1], T[q-2], T[q-3] from the top byte down; the
 * next class shifts them left a byte, so only every fourth class reads the text
 * (src/cuda/classsort.cuh:8-12).
 *
 * Here the walk, the filter and the key are a Source functor of the radix sort, so one class is
 * one filtered stable pass: histogram, scan, scatter. In a batch the text index sits above the
 * characters in a 64-bit key and adds passes on its bits; the ends of all texts lead the walk,
 * which keeps each end ahead of its own text's elements with the same character.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "dev_common.h"
#include "prim.h"

namespace sadcs {

template <int V>
__device__ __forceinline__ uint32_t classOf(uint32_t p);

// T[q], T[q-1], T[q-2], T[q-3] from the top byte down; bytes before `start` read as 0. That is
// T[q-3 .. q] as a little-endian word: one 8-byte load from the 4-byte boundary at or below q - 3
// holds it (a random gather costs per request, so one load, not four byte loads). The text buffer
// is 16-byte aligned and padded past its end.
__device__ __forceinline__ uint32_t characterKey(const uint8_t* __restrict__ bytes, uint32_t q, uint32_t start) {
    if (q >= start + 3) {
        const uint32_t a = (q - 3) & ~3u;
        const uint2 v = *reinterpret_cast<const uint2*>(__builtin_assume_aligned(bytes + a, 4));
        return __builtin_amdgcn_alignbyte(v.y, v.x, q - 3 - a);  // bytes q - 3 - a .. q - a of v
    }
    uint32_t key = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
        key = (key << 8) | (q >= start + i ? bytes[q - i] : 0u);
    }
    return key;
}

// The walk over class d - 1 with every text's end in front, filtered to the elements whose
// predecessor is in class d. Key: text index above the four characters (batch) or the characters.
template <int V, class Texts, class K>
struct ClassWalkSource {
    const uint8_t* bytes;
    const uint32_t* previous;      // class d - 1, sorted
    const K* previousKeys;         // its four-character keys (low 32 bits), or null to read the text
    uint32_t previousCount;
    const Segment* segmentTable;
    uint32_t segments;
    uint32_t n;                    // positions in the text buffer (a batch: all texts and gaps)
    Texts texts;
    int d;

    __device__ __forceinline__ bool operator()(uint32_t k, K& key, uint32_t& value) const {
        uint32_t w, start, segment;
        bool fromEnd = k < segments;
        // Both list loads first and unconditionally (index clamped for the ends; the lists hold at
        // least one entry): the sort kernels issue every item's loads before using any, which a
        // load behind the class test would serialise into a second round trip per item.
        const uint32_t index = fromEnd ? 0 : k - segments;
        const uint32_t listed = previous[index];
        const uint32_t listedChars = previousKeys ? static_cast<uint32_t>(previousKeys[index]) : 0u;
        if (fromEnd) {
            if constexpr (Texts::kBatch) {
                const Segment s = segmentTable[k];
                w = s.end;
                start = s.start;
            } else {
                w = n;
                start = 0;
            }
            segment = k;
        } else {
            w = listed;
            if (w >= n) {
                return false;  // not a position: drop it (the class then comes out short and fails its check)
            }
            if constexpr (Texts::kBatch) {
                segment = texts.segmentOf(w);
                start = segmentTable[segment].start;
            } else {
                segment = 0;
                start = 0;
            }
        }
        if (w <= start || classOf<V>(w - 1) != static_cast<uint32_t>(d)) {
            return false;
        }
        const uint32_t q = w - 1;
        const uint32_t chars = (previousKeys && !fromEnd) ? listedChars << 8 : characterKey(bytes, q, start);
        if constexpr (sizeof(K) == 8) {
            key = (static_cast<K>(segment) << 32) | chars;
        } else {
            key = chars;
        }
        value = q;
        return true;
    }
};

// A walk evaluated once into arrays (walk order; value ~0 marks a dropped element), so that a
// class whose characters come from the text pays for its random text reads once, not once in the
// histogram pass and again in the scatter pass. Only the four characters are kept: a batch's text
// index (the key's high word) follows from the position again.
template <class Source, class K>
__global__ void materializeWalkKernel(Source source, uint32_t count, uint32_t* __restrict__ chars, uint32_t* __restrict__ values) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        K key = 0;
        uint32_t value = 0;
        const bool kept = source(i, key, value);
        chars[i] = static_cast<uint32_t>(key);
        values[i] = kept ? value : ~0u;
    }
}

template <class K, class Texts>
struct MaterializedWalk {
    const uint32_t* chars;
    const uint32_t* values;
    Texts texts;

    __device__ __forceinline__ bool operator()(uint32_t i, K& key, uint32_t& value) const {
        value = values[i];
        key = chars[i];
        if constexpr (sizeof(K) == 8) {
            if (value != ~0u) {
                key |= static_cast<K>(texts.segmentOf(value)) << 32;
            }
        }
        return value != ~0u;
    }
};

// ---- a batch's class sort in one digit pass ---------------------------------------------------
//
// A batch sorts each class by (text, character) with key = text << 32 | characters. The walk is
// every text's end first, then class d - 1 in (text, suffix) order. Over the list part alone the
// items come in ascending text order, so one pass on the character digit can place them: a plain
// pass puts an item at (kept list items of smaller digits) + (those of its digit before it), and
// the second term exceeds its rank among its own text's items of its digit by exactly the items
// of its digit in earlier texts. Its place in the class is its text's start plus its text's items
// of smaller digits, plus its text's end if that has the same digit (the end comes first in the
// walk), plus that rank. With per-text digit counts, the difference is one table lookup per
// item (textDelta). The ends, one per text, go straight to the first place of their
// (text, digit). Same order as the passes over the text index that it replaces.

// The list part of a walk: walk index i + ends.
template <class Source>
struct ListPart {
    Source source;
    uint32_t ends;

    template <class K>
    __device__ __forceinline__ bool operator()(uint32_t i, K& key, uint32_t& value) const {
        return source(i + ends, key, value);
    }
};

// Per text s: its walk end (walk index s), kept or not; a kept end adds one to its text's count
// of its digit (counts, text-major: the list items' counts already there).
template <class Source, class K>
__global__ void textEndsKernel(Source source, uint32_t texts, int shift, K* __restrict__ endKeys, uint32_t* __restrict__ endValues,
                               uint32_t* __restrict__ endDigits, uint32_t* __restrict__ countsByText) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= texts) {
        return;
    }
    K key = 0;
    uint32_t value = 0;
    uint32_t digit = ~0u;
    if (source(s, key, value)) {
        digit = static_cast<uint32_t>(key >> shift) & (kRadixBins - 1);
        countsByText[static_cast<size_t>(s) * kRadixBins + digit] += 1;
    }
    endKeys[s] = key;
    endValues[s] = value;
    endDigits[s] = digit;
}

// byDigit[c * texts + s] = the list items of (s, c): all of them less the end.
__global__ void listCountsByDigitKernel(const uint32_t* __restrict__ byText, const uint32_t* __restrict__ endDigits, uint32_t texts, uint32_t* __restrict__ byDigit) {
    const uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (x < static_cast<uint64_t>(texts) * kRadixBins) {
        const uint32_t s = static_cast<uint32_t>(x / kRadixBins), c = static_cast<uint32_t>(x % kRadixBins);
        byDigit[static_cast<size_t>(c) * texts + s] = byText[x] - (endDigits[s] == c ? 1u : 0u);
    }
}

// After the exclusive sums: byText[s][c] = first place of (s, c) in the class; byDigit[c][s] = a
// plain pass's first place for the list items of (s, c). delta = the difference, past the end.
__global__ void textDigitDeltaKernel(const uint32_t* __restrict__ byText, const uint32_t* __restrict__ byDigit, const uint32_t* __restrict__ endDigits,
                                     uint32_t texts, uint32_t* __restrict__ delta) {
    const uint64_t x = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (x < static_cast<uint64_t>(texts) * kRadixBins) {
        const uint32_t s = static_cast<uint32_t>(x / kRadixBins), c = static_cast<uint32_t>(x % kRadixBins);
        delta[x] = byText[x] + (endDigits[s] == c ? 1u : 0u) - byDigit[static_cast<size_t>(c) * texts + s];
    }
}

template <class K>
__global__ void placeTextEndsKernel(const K* __restrict__ endKeys, const uint32_t* __restrict__ endValues, const uint32_t* __restrict__ endDigits,
                                    const uint32_t* __restrict__ byText, uint32_t texts, uint32_t limit, K* __restrict__ keysOut, uint32_t* __restrict__ valuesOut) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < texts && endDigits[s] != ~0u) {
        const uint32_t at = byText[static_cast<size_t>(s) * kRadixBins + endDigits[s]];
        if (at < limit) {
            keysOut[at] = endKeys[s];
            valuesOut[at] = endValues[s];
        }
    }
}

// Scratch of sortClassByText for `texts` texts.
template <class K>
struct TextSortScratch {
    uint32_t* byText = nullptr;   // texts * 256
    uint32_t* byDigit = nullptr;  // texts * 256
    uint32_t* delta = nullptr;    // texts * 256
    K* endKeys = nullptr;         // texts
    uint32_t* endValues = nullptr;
    uint32_t* endDigits = nullptr;
};

// One class of a batch: the walk (ends + list, `count` items, `kept` of them kept) sorted by
// (text, digit at shift) into buffer 1 - b.current. b.aggregates scans texts * 256 words.
template <class K, class Source>
inline void sortClassByText(hipStream_t stream, SortBuffers<K>& b, Source source, uint32_t count, uint32_t kept, uint32_t texts, int shift,
                            const TextSortScratch<K>& t) {
    const uint32_t cells = texts * kRadixBins;
    const uint32_t listCount = count - texts;
    const uint32_t tiles = std::max<uint32_t>(1, ceilDiv(listCount, kSortTile));
    const uint32_t mask = kRadixBins - 1;
    const ListPart<Source> list{source, texts};
    fillKernelWords(stream, t.byText, cells);
    radixUpsweepKernel<K, ListPart<Source>, true><<<tiles, kSortThreads, 0, stream>>>(list, listCount, shift, mask, tiles, b.hist, t.byText);
    exclusiveSum(stream, b.hist, b.hist, tiles * kRadixBins, b.aggregates);
    textEndsKernel<Source, K><<<ceilDiv(texts, 256), 256, 0, stream>>>(source, texts, shift, t.endKeys, t.endValues, t.endDigits, t.byText);
    listCountsByDigitKernel<<<ceilDiv(cells, 256), 256, 0, stream>>>(t.byText, t.endDigits, texts, t.byDigit);
    exclusiveSum(stream, t.byText, t.byText, cells, b.aggregates);
    exclusiveSum(stream, t.byDigit, t.byDigit, cells, b.aggregates);
    textDigitDeltaKernel<<<ceilDiv(cells, 256), 256, 0, stream>>>(t.byText, t.byDigit, t.endDigits, texts, t.delta);
    radixDownsweepKernel<K, ListPart<Source>><<<tiles, kSortThreads, 0, stream>>>(
        list, listCount, shift, mask, tiles, b.hist, b.keys[1 - b.current], b.values[1 - b.current], kept, nullptr, 0, 0, t.delta);
    placeTextEndsKernel<K><<<ceilDiv(texts, 256), 256, 0, stream>>>(t.endKeys, t.endValues, t.endDigits, t.byText, texts, kept,
                                                                  b.keys[1 - b.current], b.values[1 - b.current]);
    b.current = 1 - b.current;
}

// Class 0 of a batch needs no work: phase (a) ranks by (text, suffix), and the sorted sample
// array is indexed by those global ranks.

}  // namespace sadcs
