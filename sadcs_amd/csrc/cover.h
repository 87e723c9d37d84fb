/*
 * cover.h - Difference covers modulo V and the tables derived from them, for host and device.
 *
 * A difference cover D modulo V is a set of residues such that every d in [0, V) is (b - a) mod V
 * for some a, b in D. The sample is every position p with p mod V in D. For any two positions
 * i and j there is then a k < V with both i + k and j + k sampled, so two suffixes agreeing on
 * their first k characters are ordered by the ranks of the sample suffixes at i + k and j + k.
 *
 * The covers are the minimum-size covers the reference uses (src/common/differencecover.h:27-47):
 * V = 32 is its GPU cover (src/cuda/gpusa.h:55), V = 128 its CPU default. The sample ranks returned
 * by the rank-sample entry point are indexed by pack(p), so the V = 32 cover must be the same set
 * for those ranks to match the reference's.
 *
 * Everything here is constexpr; the device copies of the two-dimensional tables are filled once
 * per device by the engine (see DeviceCover in engine code).
 */
#pragma once

#include <cstdint>

namespace sadcs {

template <int V>
struct CoverSet;

template <>
struct CoverSet<32> {
    static constexpr int kSize = 7;
    static constexpr int kElements[kSize] = {0, 1, 3, 8, 12, 19, 25};
};

template <>
struct CoverSet<64> {
    static constexpr int kSize = 9;
    static constexpr int kElements[kSize] = {0, 1, 2, 5, 14, 16, 34, 42, 59};
};

template <>
struct CoverSet<128> {
    static constexpr int kSize = 13;
    static constexpr int kElements[kSize] = {0, 6, 11, 26, 30, 34, 42, 43, 44, 75, 82, 104, 107};
};

template <>
struct CoverSet<256> {
    static constexpr int kSize = 20;
    static constexpr int kElements[kSize] = {
        0, 24, 31, 34, 36, 54, 58, 123, 139, 147, 154, 165, 168, 186, 203, 219, 228, 237, 243, 244};
};

constexpr int log2Exact(int v) {
    int k = 0;
    while ((1 << k) < v) {
        k++;
    }
    return k;
}

// The one-dimensional tables of a cover: membership, elements below a residue, distance to the
// next element. Small enough to evaluate at compile time for every V.
template <int V>
struct CoverBasics {
    bool member[V];
    uint16_t below[V];
    uint16_t distance[V];
    int classes;

    constexpr CoverBasics() : member{}, below{}, distance{}, classes(0) {
        for (int e = 0; e < CoverSet<V>::kSize; e++) {
            member[CoverSet<V>::kElements[e]] = true;
        }
        int count = 0;
        for (int r = 0; r < V; r++) {
            below[r] = static_cast<uint16_t>(count);
            count += member[r] ? 1 : 0;
        }
        for (int r = 0; r < V; r++) {
            int d = 0;
            while (!member[(r + d) & (V - 1)]) {
                d++;
            }
            distance[r] = static_cast<uint16_t>(d);
            classes = d + 1 > classes ? d + 1 : classes;
        }
    }
};

// The cover's arithmetic: isSample, pack (the compact index of a sample position, or the number
// of sample positions below p), nextSampleDistance (the class of p), tieOffset.
template <int V>
struct DifferenceCover {
    static constexpr int kPeriod = V;
    static constexpr int kMask = V - 1;
    static constexpr int kLog2 = log2Exact(V);
    static constexpr int kSize = CoverSet<V>::kSize;
    static constexpr CoverBasics<V> kBasics{};
    static constexpr int kClasses = kBasics.classes;

    static_assert((1 << kLog2) == V, "the period must be a power of two");

    static constexpr bool isSample(int64_t p) {
        return kBasics.member[p & kMask];
    }

    // Number of sample positions below p; for a sample position its index in the compact arrays.
    static constexpr int64_t pack(int64_t p) {
        return (p >> kLog2) * kSize + kBasics.below[p & kMask];
    }

    // Position of the sample with compact index t.
    static constexpr int64_t unpack(int64_t t) {
        return (t / kSize) * V + CoverSet<V>::kElements[t % kSize];
    }

    static constexpr int nextSampleDistance(int64_t p) {
        return kBasics.distance[p & kMask];
    }

    // Smallest k >= 0 with a + k and b + k both sampled.
    static constexpr int tieOffset(int64_t a, int64_t b) {
        int best = V;
        for (int e = 0; e < kSize; e++) {
            const int k = static_cast<int>((CoverSet<V>::kElements[e] - a) & kMask);
            if (k < best && isSample(b + k)) {
                best = k;
            }
        }
        return best;
    }
};

// The tables the kernels index with per-lane residues. Filled on the host and copied to
// __constant__ memory of each device once.
template <int V>
struct CoverTables {
    static constexpr int kSize = CoverSet<V>::kSize;
    uint8_t tie[V][V];         // smallest k with a + k and b + k sampled
    uint8_t slot[V][V];        // [a][k]: index of sample a + k among the kSize samples in [a, a + V)
    uint16_t below[V];         // samples in [0, r)
    uint16_t residue[kSize];   // the cover, ascending
    uint8_t distance[V];       // class of residue r
    uint8_t anchor[V];         // [d]: the smallest cover element x with x + d (mod V) in the cover

    void fill() {
        using C = DifferenceCover<V>;
        for (int a = 0; a < V; a++) {
            below[a] = C::kBasics.below[a];
            distance[a] = static_cast<uint8_t>(C::kBasics.distance[a]);
            for (int b = 0; b < V; b++) {
                tie[a][b] = static_cast<uint8_t>(C::tieOffset(a, b));
            }
            for (int k = 0; k < V; k++) {
                slot[a][k] = static_cast<uint8_t>(C::pack(a + k) - C::pack(a));
            }
        }
        for (int e = 0; e < kSize; e++) {
            residue[e] = static_cast<uint16_t>(CoverSet<V>::kElements[e]);
        }
        for (int d = 0; d < V; d++) {
            anchor[d] = 0;
            for (int e = kSize - 1; e >= 0; e--) {
                if (C::isSample(CoverSet<V>::kElements[e] + d)) {
                    anchor[d] = static_cast<uint8_t>(CoverSet<V>::kElements[e]);
                }
            }
        }
    }
};

}  // namespace sadcs
