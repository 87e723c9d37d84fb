/*
 * prim.h - Device-wide primitives written for gfx950: a scan pipeline and an LSD radix sort.
 *
 * Both are reduce-then-scan: a tile kernel reduces, one block scans the tile aggregates, a second
 * tile kernel scans with the tile's prefix. Nothing waits on another block, so there is no
 * forward-progress assumption between workgroups.
 *
 * Scan pipeline. A Producer maps item i to a Vec of 32-bit fields (sum or max per field); a
 * Consumer receives the item and its exclusive prefix. Fusing the per-item work into the
 * consumer is how the phases mark groups, store ranks and compact lists in one pass.
 *
 * Radix sort. Stable LSD over a bit range with RB-bit digits. A tile of THREADS * ITEMS keys
 * is ranked with ballots: the lanes holding the same digit are found with RB ballots, the
 * rank among them is a masked popcount (mbcnt), and one leader lane per digit bumps the wave's
 * counter. Keys are then placed in LDS in tile order and written out in runs per digit.
 * The first pass may read from a Source functor that filters and computes keys, so a
 * stable partition of a derived sequence (the class sorts) is one sort pass.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include "dev_common.h"

namespace sadcs {

inline uint32_t ceilDiv(uint64_t a, uint64_t b) {
    return static_cast<uint32_t>((a + b - 1) / b);
}

// ==== scan pipeline ==============================================================================

// A compiler-only fence (no instruction, no wait): memory operations are not moved across it, so
// loads written before it are all issued before those written after it.
__device__ __forceinline__ void orderLoads() {
    __asm__ volatile("" ::: "memory");
}

// An item index for loads that are issued for every item, in range or not: past the end it reads
// the last item (or item 0 of an empty range, which every buffer here has room for).
__device__ __forceinline__ uint32_t clampIndex(uint64_t i, uint32_t count) {
    return i < count ? static_cast<uint32_t>(i) : (count ? count - 1 : 0);
}

constexpr int kScanThreads = 256;
constexpr int kScanWaves = kScanThreads / kWave;

template <class VecT, int ITEMS, class Producer>
__global__ __launch_bounds__(kScanThreads) void scanReduceKernel(Producer produce, uint32_t count, VecT* __restrict__ aggregates) {
    __shared__ VecT waveTotals[kScanWaves];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanThreads * ITEMS + wave * ITEMS * kWave + lane;
    VecT sum = VecT::zero();

    // Every item's loads first, in straight-line code (indices past the end read the last item and
    // are dropped), so they are in flight together rather than one round trip per item.
    VecT items[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        items[j] = produce(clampIndex(i, count));
    }
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        if (i < count) {
            sum = VecT::combine(sum, items[j]);
        }
    }

    sum = VecT::lane(VecT::waveInclusive(sum), kWave - 1);
    if (lane == 0) {
        waveTotals[wave] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        VecT total = VecT::zero();
        for (int w = 0; w < kScanWaves; w++) {
            total = VecT::combine(total, waveTotals[w]);
        }
        aggregates[blockIdx.x] = total;
    }
}

// Exclusive scan of count aggregates in place by one block, looping; the total goes to aggregates[count].
// Each thread takes kAggregateItems consecutive aggregates per step, so a step covers 4096 (the
// first ranking pass has about 115K tiles: 28 steps, where one aggregate per thread took 450).
constexpr uint32_t kAggregateItems = 16;

template <class VecT>
__global__ __launch_bounds__(kScanThreads) void scanAggregatesKernel(VecT* __restrict__ aggregates, uint32_t count) {
    __shared__ VecT waveTotals[kScanWaves];
    __shared__ VecT carryShared;
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;

    if (threadIdx.x == 0) {
        carryShared = VecT::zero();
    }
    __syncthreads();

    for (uint32_t chunk = 0; chunk < count; chunk += kScanThreads * kAggregateItems) {
        const uint32_t first = chunk + threadIdx.x * kAggregateItems;
        VecT x[kAggregateItems];
        VecT total = VecT::zero();
#pragma unroll
        for (uint32_t k = 0; k < kAggregateItems; k++) {
            x[k] = first + k < count ? aggregates[first + k] : VecT::zero();
        }
#pragma unroll
        for (uint32_t k = 0; k < kAggregateItems; k++) {
            total = VecT::combine(total, x[k]);
        }
        const VecT inclusive = VecT::waveInclusive(total);
        if (lane == kWave - 1) {
            waveTotals[wave] = inclusive;
        }
        __syncthreads();
        VecT prefix = carryShared;
        for (uint32_t w = 0; w < wave; w++) {
            prefix = VecT::combine(prefix, waveTotals[w]);
        }
        prefix = VecT::combine(prefix, VecT::shiftUp(inclusive));
#pragma unroll
        for (uint32_t k = 0; k < kAggregateItems; k++) {
            if (first + k < count) {
                aggregates[first + k] = prefix;
            }
            prefix = VecT::combine(prefix, x[k]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            VecT c = carryShared;
            for (int w = 0; w < kScanWaves; w++) {
                c = VecT::combine(c, waveTotals[w]);
            }
            carryShared = c;
        }
        __syncthreads();
    }

    if (threadIdx.x == 0) {
        aggregates[count] = carryShared;
    }
}

template <class VecT, int ITEMS, class Producer, class Consumer>
__global__ __launch_bounds__(kScanThreads) void scanDownsweepKernel(Producer produce, Consumer consume, uint32_t count, const VecT* __restrict__ aggregates) {
    __shared__ VecT waveTotals[kScanWaves];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanThreads * ITEMS + wave * ITEMS * kWave + lane;
    VecT items[ITEMS];
    VecT exclusive[ITEMS];
    VecT carry = VecT::zero();

#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        items[j] = produce(clampIndex(i, count));
    }
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        if (i >= count) {
            items[j] = VecT::zero();
        }
        const VecT inclusive = VecT::waveInclusive(items[j]);
        exclusive[j] = VecT::combine(carry, VecT::shiftUp(inclusive));
        carry = VecT::combine(carry, VecT::lane(inclusive, kWave - 1));
    }

    if (lane == 0) {
        waveTotals[wave] = carry;
    }
    __syncthreads();

    VecT prefix = aggregates[blockIdx.x];
    for (uint32_t w = 0; w < wave; w++) {
        prefix = VecT::combine(prefix, waveTotals[w]);
    }

    // The consumer's loads for every item (clamped indices), then its stores (valid items only):
    // stores could alias later loads, so loads interleaved with them would run one at a time.
    // (Two stages, for loads that depend on other loads of the same item.)
    typename Consumer::Prepared prepared[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        prepared[j] = consume.prepare(clampIndex(base + static_cast<uint64_t>(j) * kWave, count));
    }
    orderLoads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        consume.complete(prepared[j]);
    }
    orderLoads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        consume(clampIndex(i, count), prepared[j], items[j], VecT::combine(prefix, exclusive[j]), i < count);
    }
}

// Scratch for a scan over up to `count` items: tile aggregates plus the total.
template <class VecT, int ITEMS>
inline size_t scanScratchBytes(uint64_t count) {
    const uint64_t tiles = std::max<uint64_t>(1, ceilDiv(count, uint64_t(kScanThreads) * ITEMS));
    return (tiles + 1) * sizeof(VecT);
}

// Runs produce/consume over [0, count). aggregates holds scanScratchBytes; afterwards
// aggregates[tiles] is the total.
template <class VecT, int ITEMS, class Producer, class Consumer>
inline void scanPipeline(hipStream_t stream, uint32_t count, Producer produce, Consumer consume, VecT* aggregates) {
    const uint32_t tiles = std::max<uint32_t>(1, ceilDiv(count, uint64_t(kScanThreads) * ITEMS));
    scanReduceKernel<VecT, ITEMS><<<tiles, kScanThreads, 0, stream>>>(produce, count, aggregates);
    scanAggregatesKernel<VecT><<<1, kScanThreads, 0, stream>>>(aggregates, tiles);
    scanDownsweepKernel<VecT, ITEMS><<<tiles, kScanThreads, 0, stream>>>(produce, consume, count, aggregates);
}

using Sum1 = Vec<1, 0u>;

struct LoadU32 {
    const uint32_t* in;  // may be the output (in-place scans)
    __device__ Sum1 operator()(uint32_t i) const {
        Sum1 r;
        r.v[0] = in[i];
        return r;
    }
};

struct StoreExclusiveU32 {
    uint32_t* out;
    struct Prepared {};
    __device__ Prepared prepare(uint32_t) const {
        return {};
    }
    __device__ void complete(Prepared&) const {}
    __device__ void operator()(uint32_t i, const Prepared&, const Sum1&, const Sum1& exclusive, bool valid) const {
        if (valid) {
            out[i] = exclusive.v[0];
        }
    }
};

constexpr int kSumItems = 16;

// Exclusive sum of count words, in place allowed.
inline void exclusiveSum(hipStream_t stream, const uint32_t* in, uint32_t* out, uint32_t count, Sum1* aggregates) {
    scanPipeline<Sum1, kSumItems>(stream, count, LoadU32{in}, StoreExclusiveU32{out}, aggregates);
}

__global__ void fillZeroKernel(uint32_t* __restrict__ out, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        out[i] = 0;
    }
}

inline void fillKernelWords(hipStream_t stream, uint32_t* out, uint32_t count) {
    fillZeroKernel<<<std::max<uint32_t>(1, std::min<uint32_t>(ceilDiv(count, 256), 1024)), 256, 0, stream>>>(out, count);
}

// ==== radix sort =================================================================================

constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kSortThreads = 512;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortWaves = kSortThreads / kWave;

template <class K>
struct ArraySource {
    const K* __restrict__ keys;
    const uint32_t* __restrict__ values;  // null: value i (the identity)

    __device__ __forceinline__ bool operator()(uint32_t i, K& key, uint32_t& value) const {
        key = keys[i];
        value = values ? values[i] : i;
        return true;
    }
};

template <class K>
__device__ __forceinline__ uint32_t digitOf(K key, int shift, uint32_t mask) {
    return static_cast<uint32_t>(key >> shift) & mask;
}

// The text of a batch item's key (text << 32 | characters).
template <class K>
__device__ __forceinline__ uint32_t textOf(K key) {
    return static_cast<uint32_t>(static_cast<uint64_t>(key) >> 32);
}

constexpr uint32_t kTextSlots = 4;

// hist[d * tiles + tile] = number of kept items of the tile with digit d. kTexts (items in
// ascending text order, key = text << 32 | ...): also textCounts[text * 256 + d] += the tile's
// items of that text and digit, through LDS for the tile's first kTextSlots texts.
template <class K, class Source, bool kTexts = false>
__global__ __launch_bounds__(kSortThreads) void radixUpsweepKernel(Source source, uint32_t count, int shift, uint32_t mask, uint32_t tiles, uint32_t* __restrict__ hist,
                                                                   uint32_t* __restrict__ textCounts = nullptr) {
    __shared__ uint32_t counts[kSortWaves][kRadixBins];
    __shared__ uint32_t textHist[kTexts ? kTextSlots : 1][kRadixBins];
    __shared__ uint32_t textBase;
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;

    for (uint32_t x = threadIdx.x; x < kSortWaves * kRadixBins; x += kSortThreads) {
        (&counts[0][0])[x] = 0;
    }
    if constexpr (kTexts) {
        for (uint32_t x = threadIdx.x; x < kTextSlots * kRadixBins; x += kSortThreads) {
            (&textHist[0][0])[x] = 0;
        }
        if (threadIdx.x == 0) {
            textBase = ~0u;
        }
    }
    __syncthreads();

    const uint32_t tile = xcdTile(blockIdx.x, tiles);
    const uint64_t base = static_cast<uint64_t>(tile) * kSortTile + wave * kSortItems * kWave + lane;

    // Every item's loads first (straight-line, clamped indices), then the counting.
    K keys[kSortItems];
    bool valids[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        uint32_t value = 0;
        keys[j] = 0;
        valids[j] = source(clampIndex(i, count), keys[j], value) && i < count;
    }
    if constexpr (kTexts) {
        // The tile's first text: the least over its kept items.
        uint32_t lo = ~0u;
#pragma unroll
        for (int j = 0; j < kSortItems; j++) {
            lo = valids[j] ? min(lo, textOf(keys[j])) : lo;
        }
        for (int offset = kWave / 2; offset > 0; offset /= 2) {
            lo = min(lo, static_cast<uint32_t>(__shfl_xor(static_cast<int>(lo), offset)));
        }
        if (lane == 0 && lo != ~0u) {
            atomicMin(&textBase, lo);
        }
        __syncthreads();
    }
    // Counting needs no ranks: one LDS atomic add per item into the wave's own histogram (lanes
    // with equal digits serialise in the LDS; matching them first cost more VALU time).
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        if (valids[j]) {
            const uint32_t digit = digitOf(keys[j], shift, mask);
            atomicAdd(&counts[wave][digit], 1u);
            if constexpr (kTexts) {
                const uint32_t text = textOf(keys[j]);
                const uint32_t slot = text - textBase;
                if (slot < kTextSlots) {
                    atomicAdd(&textHist[slot][digit], 1u);
                } else {
                    atomicAdd(&textCounts[static_cast<size_t>(text) * kRadixBins + digit], 1u);
                }
            }
        }
    }
    __syncthreads();

    for (uint32_t d = threadIdx.x; d < kRadixBins; d += kSortThreads) {
        uint32_t total = 0;
        for (int w = 0; w < kSortWaves; w++) {
            total += counts[w][d];
        }
        hist[static_cast<uint64_t>(d) * tiles + tile] = total;
    }
    if constexpr (kTexts) {
        for (uint32_t x = threadIdx.x; x < kTextSlots * kRadixBins; x += kSortThreads) {
            const uint32_t c = (&textHist[0][0])[x];
            if (c != 0) {
                atomicAdd(&textCounts[(static_cast<size_t>(textBase) + x / kRadixBins) * kRadixBins + x % kRadixBins], c);
            }
        }
    }
}

// The histograms of a pass whose digits the previous pass wrote out, one byte per item in the
// items' order (radixDownsweepKernel with kDigits): a tile reads 1 byte per item instead of its
// keys. hist as radixUpsweepKernel writes it.
__global__ __launch_bounds__(kSortThreads) void radixDigitUpsweepKernel(const uint8_t* __restrict__ digits, uint32_t count, uint32_t tiles, uint32_t* __restrict__ hist) {
    // Two histograms per wave (even and odd lanes): a skewed digit's lanes collide half as often.
    constexpr int kCopies = 2 * kSortWaves;
    __shared__ uint32_t counts[kCopies][kRadixBins];
    const uint32_t copy = 2 * (threadIdx.x / kWave) + (threadIdx.x & 1);

    for (uint32_t x = threadIdx.x; x < kCopies * kRadixBins; x += kSortThreads) {
        (&counts[0][0])[x] = 0;
    }
    __syncthreads();

    // Thread x counts items [first, first + 8) of the tile: one 8-byte load (8-aligned: the buffer
    // is, and so are tile and thread offsets) unless the tile ends inside them.
    const uint32_t tile = xcdTile(blockIdx.x, tiles);
    const uint64_t first = static_cast<uint64_t>(tile) * kSortTile + static_cast<uint64_t>(threadIdx.x) * kSortItems;
    static_assert(kSortItems == 8, "one 8-byte load per thread");
    uint64_t packed = 0;
    if (first + kSortItems <= count) {
        packed = *reinterpret_cast<const uint64_t*>(digits + first);
    } else {
        for (int j = 0; j < kSortItems; j++) {
            if (first + j < count) {
                packed |= static_cast<uint64_t>(digits[first + j]) << (8 * j);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        if (first + j < count) {
            atomicAdd(&counts[copy][static_cast<uint32_t>(packed >> (8 * j)) & 0xffu], 1u);
        }
    }
    __syncthreads();

    for (uint32_t d = threadIdx.x; d < kRadixBins; d += kSortThreads) {
        uint32_t total = 0;
        for (int w = 0; w < kCopies; w++) {
            total += counts[w][d];
        }
        hist[static_cast<uint64_t>(d) * tiles + tile] = total;
    }
}

// offsets = exclusive sum of hist (digit-major): the first output index of each digit of each tile.
// Writes beyond `limit` outputs are dropped (a source that keeps more items than planned). With
// kDigits, every output item's next digit ((key >> nextShift) & nextMask) goes to digitsOut at its
// output index, for radixDigitUpsweepKernel.
// textDelta (a batch class sort by text and digit, sortClassByText): each item's output index moves by
// textDelta[text * 256 + digit].
template <class K, class Source, bool kDigits = false>
__global__ __launch_bounds__(kSortThreads) void radixDownsweepKernel(Source source, uint32_t count, int shift, uint32_t mask, uint32_t tiles, const uint32_t* __restrict__ offsets, K* __restrict__ keysOut, uint32_t* __restrict__ valuesOut, uint32_t limit,
                                                                     uint8_t* __restrict__ digitsOut = nullptr, int nextShift = 0, uint32_t nextMask = 0,
                                                                     const uint32_t* __restrict__ textDelta = nullptr) {
    __shared__ uint32_t waveCounts[kSortWaves][kRadixBins];
    __shared__ uint32_t digitStart[kRadixBins];
    __shared__ uint32_t globalStart[kRadixBins];
    __shared__ uint32_t waveSums[kRadixBins / kWave];
    __shared__ K sortedKeys[kSortTile];
    __shared__ uint32_t sortedValues[kSortTile];

    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = threadIdx.x % kWave;

    for (uint32_t x = threadIdx.x; x < kSortWaves * kRadixBins; x += kSortThreads) {
        (&waveCounts[0][0])[x] = 0;
    }
    const uint32_t tile = xcdTile(blockIdx.x, tiles);
    for (uint32_t d = threadIdx.x; d < kRadixBins; d += kSortThreads) {
        globalStart[d] = offsets[static_cast<uint64_t>(d) * tiles + tile];
    }
    __syncthreads();

    const uint64_t base = static_cast<uint64_t>(tile) * kSortTile + wave * kSortItems * kWave + lane;
    K keys[kSortItems];
    uint32_t values[kSortItems];
    uint32_t ranks[kSortItems];
    bool valids[kSortItems];

    // Every item's loads first (straight-line, clamped indices).
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint64_t i = base + static_cast<uint64_t>(j) * kWave;
        keys[j] = 0;
        values[j] = 0;
        valids[j] = source(clampIndex(i, count), keys[j], values[j]) && i < count;
    }
    // Rank each item among the wave's earlier items with the same digit.
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        const uint32_t digit = digitOf(keys[j], shift, mask);
        const uint64_t peers = matchDigit<kRadixBits>(digit, valids[j]);
        const uint32_t before = valids[j] ? waveCounts[wave][digit] : 0;
        ranks[j] = before + countBelow(peers);
        if (valids[j] && lane == static_cast<uint32_t>(__builtin_ctzll(peers))) {
            waveCounts[wave][digit] = before + static_cast<uint32_t>(__builtin_popcountll(peers));
        }
    }
    __syncthreads();

    // Per digit: the waves' prefixes, the tile's count; then the digits' starts in the tile.
    uint32_t total = 0;
    if (threadIdx.x < kRadixBins) {
        const uint32_t d = threadIdx.x;
        for (int w = 0; w < kSortWaves; w++) {
            const uint32_t c = waveCounts[w][d];
            waveCounts[w][d] = total;
            total += c;
        }
        const uint32_t inclusive = waveInclusive<OpSum>(total);
        if (lane == kWave - 1) {
            waveSums[wave] = inclusive;
        }
        digitStart[d] = inclusive - total;
    }
    __syncthreads();
    if (threadIdx.x < kRadixBins) {
        uint32_t prefix = 0;
        for (uint32_t w = 0; w < wave; w++) {
            prefix += waveSums[w];
        }
        digitStart[threadIdx.x] += prefix;
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        if (valids[j]) {
            const uint32_t digit = digitOf(keys[j], shift, mask);
            const uint32_t local = digitStart[digit] + waveCounts[wave][digit] + ranks[j];
            sortedKeys[local] = keys[j];
            sortedValues[local] = values[j];
        }
    }
    __syncthreads();

    uint32_t kept = 0;
    for (int w = 0; w < kRadixBins / kWave; w++) {
        kept += waveSums[w];
    }
    for (uint32_t x = threadIdx.x; x < kept; x += kSortThreads) {
        const K key = sortedKeys[x];
        const uint32_t digit = digitOf(key, shift, mask);
        uint32_t target = globalStart[digit] + (x - digitStart[digit]);
        if (textDelta) {
            target += textDelta[static_cast<size_t>(textOf(key)) * kRadixBins + digit];
        }
        if (target < limit) {
            keysOut[target] = key;
            valuesOut[target] = sortedValues[x];
            if constexpr (kDigits) {
                digitsOut[target] = static_cast<uint8_t>(digitOf(key, nextShift, nextMask));
            }
        }
    }
}

// Buffers and scratch of a sort over up to `capacity` items (reduce-then-scan per digit pass:
// tile histograms, one scan, scatter).
template <class K>
struct SortBuffers {
    K* keys[2] = {};
    uint32_t* values[2] = {};
    int current = 0;        // which buffer holds the data
    uint32_t* hist = nullptr;
    Sum1* aggregates = nullptr;
    uint8_t* digits = nullptr;  // optional, one byte per item: radixSort's passes hand on their next digit
    bool digitsReady = false;   // digits already holds the first pass's digit (written with the keys)
    bool identityValues = false;  // the values are 0, 1, 2, ... and not stored: the first pass makes them
};

inline size_t radixHistWords(uint64_t capacity) {
    return static_cast<size_t>(std::max<uint32_t>(1, ceilDiv(capacity, kSortTile))) * kRadixBins;
}

inline size_t radixAggregateBytes(uint64_t capacity) {
    return scanScratchBytes<Sum1, kSumItems>(radixHistWords(capacity));
}

// One pass on the digit at shift (width bits), reading the source, writing buffer 1 - current.
template <class K, class Source>
inline void radixPass(hipStream_t stream, SortBuffers<K>& b, Source source, uint32_t count, int shift, int width, uint32_t limit = 0xFFFFFFFFu) {
    const uint32_t tiles = std::max<uint32_t>(1, ceilDiv(count, kSortTile));
    const uint32_t mask = (1u << width) - 1;
    const uint32_t histWords = tiles * kRadixBins;
    radixUpsweepKernel<K, Source><<<tiles, kSortThreads, 0, stream>>>(source, count, shift, mask, tiles, b.hist);
    exclusiveSum(stream, b.hist, b.hist, histWords, b.aggregates);
    radixDownsweepKernel<K, Source><<<tiles, kSortThreads, 0, stream>>>(
        source, count, shift, mask, tiles, b.hist, b.keys[1 - b.current], b.values[1 - b.current], limit == 0xFFFFFFFFu ? count : limit);
    b.current = 1 - b.current;
}

// radixPass with digit bytes (b.digits): fromDigits, the histograms count the bytes the pass
// before wrote instead of the source's keys; nextShift >= 0, the scatter writes each output
// item's digit at nextShift (nextWidth bits) for the pass after.
template <class K, class Source>
inline void radixPassDigits(hipStream_t stream, SortBuffers<K>& b, Source source, uint32_t count, int shift, int width, uint32_t limit,
                            bool fromDigits, int nextShift, int nextWidth) {
    const uint32_t tiles = std::max<uint32_t>(1, ceilDiv(count, kSortTile));
    const uint32_t mask = (1u << width) - 1;
    if (fromDigits) {
        radixDigitUpsweepKernel<<<tiles, kSortThreads, 0, stream>>>(b.digits, count, tiles, b.hist);
    } else {
        radixUpsweepKernel<K, Source><<<tiles, kSortThreads, 0, stream>>>(source, count, shift, mask, tiles, b.hist);
    }
    exclusiveSum(stream, b.hist, b.hist, tiles * kRadixBins, b.aggregates);
    const uint32_t kept = limit == 0xFFFFFFFFu ? count : limit;
    if (nextShift >= 0) {
        radixDownsweepKernel<K, Source, true><<<tiles, kSortThreads, 0, stream>>>(
            source, count, shift, mask, tiles, b.hist, b.keys[1 - b.current], b.values[1 - b.current], kept, b.digits, nextShift, (1u << nextWidth) - 1);
    } else {
        radixDownsweepKernel<K, Source><<<tiles, kSortThreads, 0, stream>>>(
            source, count, shift, mask, tiles, b.hist, b.keys[1 - b.current], b.values[1 - b.current], kept);
    }
    b.current = 1 - b.current;
}

// Stable sort of b.keys[current], b.values[current] (count items) by bits [beginBit, endBit).
template <class K>
inline void radixSort(hipStream_t stream, SortBuffers<K>& b, uint32_t count, int beginBit, int endBit) {
    if (count <= 1 || endBit <= beginBit) {
        return;
    }
    const uint32_t tiles = std::max<uint32_t>(1, ceilDiv(count, kSortTile));
    for (int shift = beginBit; shift < endBit; shift += kRadixBits) {
        const int width = std::min(kRadixBits, endBit - shift);
        const bool firstPass = shift == beginBit;
        const ArraySource<K> source{b.keys[b.current], firstPass && b.identityValues ? nullptr : b.values[b.current]};
        const int nextShift = shift + kRadixBits;
        if (!b.digits || (firstPass && nextShift >= endBit && !b.digitsReady)) {
            radixPass<K>(stream, b, source, count, shift, width);
            continue;
        }
        // With a digit buffer: every pass after the first counts the bytes the one before wrote
        // (the first too when the keys' writer wrote its digits). The buffer is read by this pass's
        // histograms before its scatter overwrites it.
        const uint32_t mask = (1u << width) - 1;
        if (firstPass && !b.digitsReady) {
            radixUpsweepKernel<K, ArraySource<K>><<<tiles, kSortThreads, 0, stream>>>(source, count, shift, mask, tiles, b.hist);
        } else {
            radixDigitUpsweepKernel<<<tiles, kSortThreads, 0, stream>>>(b.digits, count, tiles, b.hist);
        }
        exclusiveSum(stream, b.hist, b.hist, tiles * kRadixBins, b.aggregates);
        if (nextShift < endBit) {
            const uint32_t nextMask = (1u << std::min(kRadixBits, endBit - nextShift)) - 1;
            radixDownsweepKernel<K, ArraySource<K>, true><<<tiles, kSortThreads, 0, stream>>>(
                source, count, shift, mask, tiles, b.hist, b.keys[1 - b.current], b.values[1 - b.current], count, b.digits, nextShift, nextMask);
        } else {
            radixDownsweepKernel<K, ArraySource<K>><<<tiles, kSortThreads, 0, stream>>>(
                source, count, shift, mask, tiles, b.hist, b.keys[1 - b.current], b.values[1 - b.current], count);
        }
        b.current = 1 - b.current;
    }
}

}  // namespace sadcs
