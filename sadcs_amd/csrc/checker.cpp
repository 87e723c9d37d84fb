/*
 * checker.cpp - The checks of checker.h. OpenMP-parallel where the check is a pure scan.
 */
#include "checker.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

namespace {
    constexpr int64_t kNone = INT64_MAX;

    CheckResult ok() {
        return CheckResult{true, -1, nullptr};
    }

    CheckResult failAt(int64_t index, const char* reason) {
        return CheckResult{false, index, reason};
    }

    CheckResult permutation(int32_t n, const int32_t* sa) {
        std::vector<uint64_t> seen((static_cast<size_t>(n) + 63) / 64, 0);
        for (int32_t i = 0; i < n; i++) {
            const int32_t p = sa[i];
            if (p < 0 || p >= n) {
                return failAt(i, "entry out of range");
            }
            uint64_t& word = seen[static_cast<size_t>(p) >> 6];
            const uint64_t bit = uint64_t(1) << (p & 63);
            if (word & bit) {
                return failAt(i, "duplicate entry");
            }
            word |= bit;
        }
        return ok();
    }

    CheckResult fast(const uint8_t* text, int32_t n, const int32_t* sa) {
        std::vector<int32_t> inverse(static_cast<size_t>(n), -1);
        int64_t bad = kNone;
#pragma omp parallel for reduction(min : bad)
        for (int32_t i = 0; i < n; i++) {
            const int32_t p = sa[i];
            if (p < 0 || p >= n) {
                bad = std::min<int64_t>(bad, i);
            } else {
                // The last of equal entries owns the slot, as in a sequential pass, so the first
                // duplicate reported does not depend on the thread schedule.
                std::atomic_ref<int32_t> slot(inverse[p]);
                int32_t seen = slot.load(std::memory_order_relaxed);
                while (seen < i && !slot.compare_exchange_weak(seen, i, std::memory_order_relaxed)) {
                }
            }
        }
        if (bad != kNone) {
            return failAt(bad, "entry out of range");
        }
#pragma omp parallel for reduction(min : bad)
        for (int32_t i = 0; i < n; i++) {
            if (inverse[sa[i]] != i) {
                bad = std::min<int64_t>(bad, i);
            }
        }
        if (bad != kNone) {
            return failAt(bad, "duplicate entry");
        }
        // With equal first characters, the order of the suffixes one position on decides (the empty suffix is -1).
#pragma omp parallel for reduction(min : bad)
        for (int32_t i = 0; i < n - 1; i++) {
            const int32_t a = sa[i], b = sa[i + 1];
            bool ordered = text[a] < text[b];
            if (text[a] == text[b]) {
                ordered = (a + 1 < n ? inverse[a + 1] : -1) < (b + 1 < n ? inverse[b + 1] : -1);
            }
            if (!ordered) {
                bad = std::min<int64_t>(bad, i);
            }
        }
        return bad == kNone ? ok() : failAt(bad, "SA[index] and SA[index + 1] out of order");
    }

    // In SA order, the suffix one position before each suffix must be the next unvisited member of its
    // first-character bucket; suffix n-1 is first of its bucket and is visited from the empty suffix.
    CheckResult lowMemory(const uint8_t* text, int32_t n, const int32_t* sa) {
        const CheckResult p = permutation(n, sa);
        if (!p.ok) {
            return p;
        }
        for (int32_t i = 1; i < n; i++) {
            if (text[sa[i - 1]] > text[sa[i]]) {
                return failAt(i - 1, "first characters out of order");
            }
        }
        int64_t next[256] = {};
        for (int32_t i = 0; i < n; i++) {
            next[text[i]]++;
        }
        for (int c = 0, offset = 0; c < 256; c++) {
            const int64_t count = next[c];
            next[c] = offset;
            offset += static_cast<int>(count);
        }
        const int64_t lastSlot = next[text[n - 1]]++;
        for (int32_t i = 0; i < n; i++) {
            const bool fromEnd = sa[i] == 0;
            const int32_t before = fromEnd ? n - 1 : sa[i] - 1;
            const int64_t slot = fromEnd ? lastSlot : next[text[before]];
            if (slot < 0 || slot >= n || sa[slot] != before) {
                return failAt(i, "predecessor of SA[index] not where its bucket expects it");
            }
            if (!fromEnd) {
                const uint8_t bucket = text[before];
                int64_t& s = ++next[bucket];
                if (s >= n || text[sa[s]] != bucket) {
                    s = -1;
                }
            }
        }
        return ok();
    }

    CheckResult naive(const uint8_t* text, int32_t n, const int32_t* sa) {
        const CheckResult p = permutation(n, sa);
        if (!p.ok) {
            return p;
        }
        int64_t bad = kNone;
#pragma omp parallel for schedule(dynamic, 4096) reduction(min : bad)
        for (int32_t i = 0; i < n - 1; i++) {
            const int32_t a = sa[i], b = sa[i + 1];
            const int32_t comparable = std::min(n - a, n - b);
            int32_t common = 0;
            while (common < comparable && text[a + common] == text[b + common]) {
                common++;
            }
            const bool ordered = common < comparable ? text[a + common] < text[b + common] : a > b;
            if (!ordered) {
                bad = std::min<int64_t>(bad, i);
            }
        }
        return bad == kNone ? ok() : failAt(bad, "SA[index] and SA[index + 1] out of order");
    }
}

CheckResult checkSuffixArray(const uint8_t* text, int32_t n, const int32_t* sa, CheckMode mode) {
    if (n <= 0) {
        return ok();
    }
    switch (mode) {
        case CheckMode::LowMemory:
            return lowMemory(text, n, sa);
        case CheckMode::Naive:
            return naive(text, n, sa);
        default:
            return fast(text, n, sa);
    }
}

// The matrix of the text plus a sentinel has n + 1 rows: row 0 is the sentinel's suffix, row
// `primary` is suffix 0's (whose BWT character is the sentinel). bwt holds the other rows in order.
CheckResult checkBwt(const uint8_t* text, int32_t n, const uint8_t* bwt, int32_t primary) {
    if (n <= 0) {
        return primary == 0 ? ok() : failAt(-1, "primary index out of range");
    }
    if (primary < 1 || primary > n) {
        return failAt(-1, "primary index out of range");
    }
    std::vector<int32_t> lf(static_cast<size_t>(n));
    int64_t start[256] = {};
    for (int32_t u = 0; u < n; u++) {
        start[bwt[u]]++;
    }
    int64_t sum = 1;  // row 0 is the sentinel's
    for (int c = 0; c < 256; c++) {
        const int64_t count = start[c];
        start[c] = sum;
        sum += count;
    }
    for (int32_t u = 0; u < n; u++) {
        lf[u] = static_cast<int32_t>(start[bwt[u]]++);
    }
    int32_t row = 0;
    for (int32_t i = n - 1; i >= 0; i--) {
        if (row == primary) {
            return failAt(i, "inverse BWT ends early at text position index");
        }
        const int32_t u = row < primary ? row : row - 1;
        if (bwt[u] != text[i]) {
            return failAt(i, "inverse BWT differs from the text at position index");
        }
        row = lf[u];
    }
    return row == primary ? ok() : failAt(0, "inverse BWT doesn't end at the primary index");
}
