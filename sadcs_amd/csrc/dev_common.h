/*
 * dev_common.h - Device helpers shared by every phase: text windows, wave64 scans on DPP,
 * ballot-based peer matching, and the segment policies that let one pipeline build a single
 * text or a batch of texts.
 *
 * Wavefronts are 64 lanes on gfx950. Scans use DPP row shifts plus the GFX9 row broadcasts
 * (row_bcast:15 / row_bcast:31), so a 64-lane inclusive scan is six DPP steps without LDS.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sadcs {

constexpr int kWave = 64;

// ---- text ---------------------------------------------------------------------------------

// The 8 text bytes from p as a big-endian word (byte p in the top byte). The text buffer is
// padded with zero words, so the second word read is always inside it.
__device__ __forceinline__ uint64_t load8(const uint64_t* __restrict__ words, uint64_t p) {
    const uint64_t lo = words[p >> 3];
    const uint64_t hi = words[(p >> 3) + 1];
    const uint32_t s = static_cast<uint32_t>(p & 7) * 8;
    const uint64_t x = s ? (lo >> s) | (hi << (64 - s)) : lo;
    return __builtin_bswap64(x);
}

// Mask keeping the first `remaining` bytes of a big-endian word (all of it for 8 or more).
__device__ __forceinline__ uint64_t prefixMask(uint32_t remaining) {
    return remaining >= 8 ? ~uint64_t(0) : ~uint64_t(0) << (64 - 8 * remaining);
}

// ---- XCD-aware tile order -------------------------------------------------------------------------

// MI355X hands workgroups to its 8 XCDs round robin by index, and each XCD has its own L2. This
// maps workgroup `block` of `tiles` to a tile so that each XCD takes one contiguous range of tiles:
// neighbouring tiles, whose outputs share cache lines, then run on the same XCD at about the same
// time and their partial lines merge in its L2 instead of in memory. A bijection on [0, tiles).
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcdTile(uint32_t block, uint32_t tiles) {
    const uint32_t q = tiles / kXcds, r = tiles % kXcds;
    const uint32_t xcd = block % kXcds, slot = block / kXcds;
    return xcd * q + min(xcd, r) + slot;
}

// ---- lanes ------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t laneId() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0));
}

// Number of set bits of mask below this lane.
__device__ __forceinline__ uint32_t countBelow(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0));
}

__device__ __forceinline__ uint64_t ballot(bool predicate) {
    return __ballot(predicate);
}

// Lanes of the wave that are valid and hold the same BITS-bit digit as this lane.
template <int BITS>
__device__ __forceinline__ uint64_t matchDigit(uint32_t digit, bool valid) {
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
        const bool bit = (digit >> b) & 1;
        const uint64_t set = ballot(bit);
        peers &= bit ? set : ~set;
    }
    return peers;
}

// ---- wave scans on DPP --------------------------------------------------------------------------

// DPP controls (GFX9): row_shr:n = 0x110 + n, wave_shr:1 = 0x138, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143. Lanes without a source read 0, the identity of every operator used here.
template <int kCtrl, int kRowMask = 0xF, int kBankMask = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(x), kCtrl, kRowMask, kBankMask, false));
}

struct OpSum {
    __device__ static uint32_t apply(uint32_t a, uint32_t b) {
        return a + b;
    }
};

struct OpMax {
    __device__ static uint32_t apply(uint32_t a, uint32_t b) {
        return a > b ? a : b;
    }
};

// Inclusive scan over the 64 lanes; identity 0.
template <class Op>
__device__ __forceinline__ uint32_t waveInclusive(uint32_t x) {
    x = Op::apply(x, dpp<0x111>(x));
    x = Op::apply(x, dpp<0x112>(x));
    x = Op::apply(x, dpp<0x114>(x));
    x = Op::apply(x, dpp<0x118>(x));
    x = Op::apply(x, dpp<0x142, 0xA>(x));
    x = Op::apply(x, dpp<0x143, 0xC>(x));
    return x;
}

// The value of the lane before (0 in lane 0).
__device__ __forceinline__ uint32_t waveShiftUp(uint32_t x) {
    return dpp<0x138>(x);
}

__device__ __forceinline__ uint32_t readLane(uint32_t x, int lane) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), lane));
}

// A vector of N 32-bit fields scanned field by field; bit f of kMaxMask selects max for field f,
// else sum. Both have identity 0.
template <int N, uint32_t kMaxMask>
struct Vec {
    uint32_t v[N];

    __device__ static Vec zero() {
        Vec r;
#pragma unroll
        for (int f = 0; f < N; f++) {
            r.v[f] = 0;
        }
        return r;
    }

    __device__ static Vec combine(const Vec& a, const Vec& b) {
        Vec r;
#pragma unroll
        for (int f = 0; f < N; f++) {
            r.v[f] = ((kMaxMask >> f) & 1) ? OpMax::apply(a.v[f], b.v[f]) : OpSum::apply(a.v[f], b.v[f]);
        }
        return r;
    }

    __device__ static Vec waveInclusive(const Vec& x) {
        Vec r;
#pragma unroll
        for (int f = 0; f < N; f++) {
            r.v[f] = ((kMaxMask >> f) & 1) ? sadcs::waveInclusive<OpMax>(x.v[f]) : sadcs::waveInclusive<OpSum>(x.v[f]);
        }
        return r;
    }

    __device__ static Vec shiftUp(const Vec& x) {
        Vec r;
#pragma unroll
        for (int f = 0; f < N; f++) {
            r.v[f] = waveShiftUp(x.v[f]);
        }
        return r;
    }

    __device__ static Vec lane(const Vec& x, int l) {
        Vec r;
#pragma unroll
        for (int f = 0; f < N; f++) {
            r.v[f] = readLane(x.v[f], l);
        }
        return r;
    }
};

// ---- segments -----------------------------------------------------------------------------------

// One text of a batch. Texts sit in one position space, each starting at a multiple of
// kSegmentAlign (a multiple of every cover period used), so residues modulo V are the same in
// local and global positions.
struct Segment {
    uint32_t start;        // first position
    uint32_t end;          // one past the last position
    uint32_t rankBase;     // sample suffixes of earlier texts: the first sample rank of this text
    uint32_t rowBase;      // positions of earlier texts: the first output row of this text
};

constexpr uint32_t kSegmentAlign = 256;
constexpr int kSegmentBlockLog = 8;  // positions per entry of the block-to-segment table (= align)

// Single text: one segment [0, n).
struct SingleText {
    uint32_t n;

    __device__ __forceinline__ uint32_t segmentOf(uint32_t) const {
        return 0;
    }
    __device__ __forceinline__ uint32_t endOf(uint32_t) const {
        return n;
    }
    __device__ __forceinline__ uint32_t startOf(uint32_t) const {
        return 0;
    }
    static constexpr bool kBatch = false;
};

// A batch: the segment of position p comes from a table with one entry per aligned block.
struct TextBatch {
    const uint32_t* __restrict__ blockSegment;
    const Segment* __restrict__ segments;

    __device__ __forceinline__ uint32_t segmentOf(uint32_t p) const {
        return blockSegment[p >> kSegmentBlockLog];
    }
    __device__ __forceinline__ uint32_t endOf(uint32_t p) const {
        return segments[segmentOf(p)].end;
    }
    __device__ __forceinline__ uint32_t startOf(uint32_t p) const {
        return segments[segmentOf(p)].start;
    }
    static constexpr bool kBatch = true;
};

}  // namespace sadcs
