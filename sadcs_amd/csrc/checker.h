/*
 * checker.h - Independent verification of a suffix array or BWT (semantics of the reference's
 * src/common/checker.h:14-28). Shares no code with the construction.
 */
#pragma once

#include <cstdint>

enum class CheckMode { Fast, LowMemory, Naive };

struct CheckResult {
    bool ok;
    int64_t index;       // SA index (or text position, for the BWT) of the first failure, -1 if ok
    const char* reason;  // null if ok
};

// Fast: inverse suffix array plus one comparison per neighbour pair (checker.cpp:53-109).
// LowMemory: permutation bitmap plus a walk of the first-character buckets (:112-168).
// Naive: every neighbour pair compared character by character (:170-193).
CheckResult checkSuffixArray(const uint8_t* text, int32_t n, const int32_t* sa, CheckMode mode);

// Inverts the BWT (buildBwt()'s convention) and compares it with the text (:196-245).
CheckResult checkBwt(const uint8_t* text, int32_t n, const uint8_t* bwt, int32_t primary);
