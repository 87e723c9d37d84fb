/*
 * checker.cpp - The checks of checker.h, written from the published methods:
 *
 *   Fast       Burkhardt and Kärkkäinen's check (CPM 2003, "Fast lightweight suffix array
 *              construction and checking"): SA is the suffix array of T exactly when it is a
 *              permutation of [0, n) and, for every neighbour pair (a, b) = (SA[i], SA[i + 1]),
 *              (T[a], rank(a + 1)) < (T[b], rank(b + 1)), with the empty suffix ranked -1.
 *              Needs the inverse array (4n bytes). OpenMP-parallel.
 *   LowMemory  The induced order of suffixes: scanning the rows in order, suffix j > 0 places
 *              suffix j - 1 = T[j - 1] . suffix(j) in the bucket of T[j - 1], and the suffixes a
 *              bucket receives this way arrive in sorted order. So suffix j - 1 must occupy the
 *              next unclaimed slot of its bucket when row j is scanned. n bits plus 4 KiB.
 *   Naive      Every neighbour pair compared character by character.
 *   BWT        Inversion by the LF mapping (Burrows and Wheeler 1994) over the n + 1 rows of
 *              the sorted rotations of T plus a sentinel.
 *
 * The index a failed check reports is the interface of checker.h: the first SA index in scan
 * order (for the fast check, the smallest failing index of its first failing test), and for the
 * BWT the text position at which the inversion departs from the text.
 */
#include "checker.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <vector>

namespace {
    constexpr int64_t kNothing = INT64_MAX;

    const char* const kOutOfRange = "SA[index] is not a position of the text";
    const char* const kRepeated = "SA[index] repeats another entry";
    const char* const kMisordered = "suffix SA[index] is not smaller than suffix SA[index + 1]";
    const char* const kFirstCharacters = "SA[index] starts with a larger character than SA[index + 1]";
    const char* const kNotInduced = "the suffix one longer than SA[index] is not at its bucket's next slot";
    const char* const kPrimaryRange = "primary index not in 1..n";
    const char* const kEarlySentinel = "the inverted BWT reaches the sentinel early, at text position index";
    const char* const kWrongCharacter = "the inverted BWT differs from the text at position index";
    const char* const kNotClosed = "the inverted BWT does not end at the primary row";

    CheckResult passed() {
        return CheckResult{true, -1, nullptr};
    }

    CheckResult failed(int64_t index, const char* what) {
        return CheckResult{false, index, what};
    }

    // A scan in order: the first entry outside [0, n) or equal to an earlier entry.
    CheckResult scanPermutation(int32_t n, const int32_t* sa) {
        std::vector<uint64_t> marked((static_cast<size_t>(n) + 63) / 64, 0);
        for (int64_t i = 0; i < n; i++) {
            const uint32_t p = static_cast<uint32_t>(sa[i]);  // a negative entry becomes >= n
            if (p >= static_cast<uint32_t>(n)) {
                return failed(i, kOutOfRange);
            }
            uint64_t& word = marked[p / 64];
            const uint64_t bit = uint64_t(1) << (p % 64);
            if (word & bit) {
                return failed(i, kRepeated);
            }
            word |= bit;
        }
        return passed();
    }

    CheckResult fastCheck(const uint8_t* text, int32_t n, const int32_t* sa) {
        std::vector<int32_t> rank(static_cast<size_t>(n));
        int64_t outOfRange = kNothing;
        // rank[SA[i]] = i. Where an entry repeats, any of its indices may take the slot: the pass
        // below finds the repeat whichever did.
#pragma omp parallel for reduction(min : outOfRange)
        for (int64_t i = 0; i < n; i++) {
            const uint32_t p = static_cast<uint32_t>(sa[i]);
            if (p >= static_cast<uint32_t>(n)) {
                outOfRange = std::min(outOfRange, i);
            } else {
                std::atomic_ref<int32_t>(rank[p]).store(static_cast<int32_t>(i), std::memory_order_relaxed);
            }
        }
        if (outOfRange != kNothing) {
            return failed(outOfRange, kOutOfRange);
        }
        // An index that lost its slot holds a repeated value, and so does the index that took it;
        // the smallest index holding a repeated value is the same whichever took the slot.
        int64_t repeated = kNothing;
#pragma omp parallel for reduction(min : repeated)
        for (int64_t i = 0; i < n; i++) {
            const int64_t owner = rank[static_cast<uint32_t>(sa[i])];
            if (owner != i) {
                repeated = std::min(repeated, std::min(owner, i));
            }
        }
        if (repeated != kNothing) {
            return failed(repeated, kRepeated);
        }
        // (first character, rank of the suffix one on + 1) as one 64-bit key; 0 is the empty suffix.
        const auto key = [&](int32_t p) {
            const uint64_t on = p + 1 < n ? static_cast<uint64_t>(rank[p + 1]) + 1 : 0;
            return (static_cast<uint64_t>(text[p]) << 32) | on;
        };
        int64_t misordered = kNothing;
#pragma omp parallel for reduction(min : misordered)
        for (int64_t i = 0; i < static_cast<int64_t>(n) - 1; i++) {
            if (key(sa[i]) >= key(sa[i + 1])) {
                misordered = std::min(misordered, i);
            }
        }
        return misordered == kNothing ? passed() : failed(misordered, kMisordered);
    }

    CheckResult inducedOrderCheck(const uint8_t* text, int32_t n, const int32_t* sa) {
        const CheckResult permutation = scanPermutation(n, sa);
        if (!permutation.ok) {
            return permutation;
        }
        for (int64_t i = 0; i + 1 < n; i++) {
            if (text[sa[i]] > text[sa[i + 1]]) {
                return failed(i, kFirstCharacters);
            }
        }
        // Bucket c holds rows [edge[c], edge[c + 1]).
        std::array<int64_t, 257> edge{};
        for (int64_t p = 0; p < n; p++) {
            edge[text[p] + 1u]++;
        }
        for (int c = 0; c < 256; c++) {
            edge[c + 1] += edge[c];
        }
        std::array<int64_t, 256> unclaimed;
        std::copy(edge.begin(), edge.begin() + 256, unclaimed.begin());
        // The empty suffix precedes every row and places suffix n - 1 first in its bucket. Suffix 0
        // places nothing; its row is where the walk confirms that slot.
        const int64_t slotOfLast = unclaimed[text[n - 1]]++;
        for (int64_t i = 0; i < n; i++) {
            const int32_t j = sa[i];
            int64_t slot = slotOfLast;
            int32_t placed = n - 1;
            if (j > 0) {
                const uint8_t c = text[j - 1];
                placed = j - 1;
                slot = unclaimed[c] < edge[c + 1u] ? unclaimed[c]++ : -1;
            }
            if (slot < 0 || sa[slot] != placed) {
                return failed(i, kNotInduced);
            }
        }
        return passed();
    }

    CheckResult pairwiseCheck(const uint8_t* text, int32_t n, const int32_t* sa) {
        const CheckResult permutation = scanPermutation(n, sa);
        if (!permutation.ok) {
            return permutation;
        }
        int64_t misordered = kNothing;
#pragma omp parallel for schedule(dynamic, 4096) reduction(min : misordered)
        for (int64_t i = 0; i < static_cast<int64_t>(n) - 1; i++) {
            const int32_t a = sa[i], b = sa[i + 1];
            const int order = memcmp(text + a, text + b, static_cast<size_t>(std::min(n - a, n - b)));
            // On a common prefix as long as the shorter suffix, the shorter (later) suffix is smaller.
            if (order > 0 || (order == 0 && a < b)) {
                misordered = std::min(misordered, i);
            }
        }
        return misordered == kNothing ? passed() : failed(misordered, kMisordered);
    }
}

CheckResult checkSuffixArray(const uint8_t* text, int32_t n, const int32_t* sa, CheckMode mode) {
    if (n <= 0) {
        return passed();
    }
    if (mode == CheckMode::LowMemory) {
        return inducedOrderCheck(text, n, sa);
    }
    if (mode == CheckMode::Naive) {
        return pairwiseCheck(text, n, sa);
    }
    return fastCheck(text, n, sa);
}

// Rows 0..n of the sorted rotations of T$ ($ the sentinel, smallest): row 0 starts with $, and
// its last symbol is T[n - 1]. The last column holds $ at row `primary` and bwt, in order, in the
// other rows. LF(row) is the row of the rotation one step to the left: 1 + the number of smaller
// last-column symbols + the earlier rows with the same one; LF(primary) = 0. From row 0, the
// symbols met along LF spell T backwards and the walk must meet $ exactly after T[0].
CheckResult checkBwt(const uint8_t* text, int32_t n, const uint8_t* bwt, int32_t primary) {
    if (n <= 0) {
        return primary == 0 ? passed() : failed(-1, kPrimaryRange);
    }
    if (primary < 1 || primary > n) {
        return failed(-1, kPrimaryRange);
    }
    const auto lastSymbol = [&](int64_t row) { return row == primary ? -1 : static_cast<int>(bwt[row - (row > primary)]); };
    std::array<int64_t, 256> next{};
    for (int64_t u = 0; u < n; u++) {
        next[bwt[u]]++;
    }
    for (int64_t c = 0, rows = 1; c < 256; c++) {
        const int64_t count = next[c];
        next[c] = rows;
        rows += count;
    }
    std::vector<int32_t> lf(static_cast<size_t>(n) + 1);
    for (int64_t row = 0; row <= n; row++) {
        const int s = lastSymbol(row);
        lf[row] = s < 0 ? 0 : static_cast<int32_t>(next[s]++);
    }
    int64_t row = 0;
    for (int64_t i = static_cast<int64_t>(n) - 1; i >= 0; i--) {
        const int s = lastSymbol(row);
        if (s < 0) {
            return failed(i, kEarlySentinel);
        }
        if (s != text[i]) {
            return failed(i, kWrongCharacter);
        }
        row = lf[row];
    }
    return row == primary ? passed() : failed(0, kNotClosed);
}
