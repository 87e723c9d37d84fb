/*
 * check.h - Checking a suffix array (and BWT) on the device.
 *
 * The checks of the reference's fast checker (src/common/checker.cpp:53-109): every entry in
 * [0, n); no two entries equal (the entry that owns the inverse slot of its position is the only
 * one); neighbours in order by their first character and then by the rank of the suffix that
 * follows (the empty suffix ranks lowest). Together they hold exactly for the suffix array. The
 * BWT check adds what the suffix array fixes for the BWT (divsufsort's convention, as
 * src/cuda/gpusa.cu:579-604 writes it): byte 0 is T[n-1], the row of suffix 0 is skipped, every
 * other row r holds T[SA[r] - 1] at index r + 1 before that row and r after it, and the primary
 * index is that row plus one. Each check records the first failing index with an atomic minimum,
 * as the reference's OpenMP loops reduce with min.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sadcs {

enum : uint32_t {
    kCheckRange = 0,      // entry out of range
    kCheckDuplicate = 1,  // duplicate entry
    kCheckOrder = 2,      // SA[i] and SA[i + 1] out of order
    kCheckBwt = 3,        // BWT byte differs
    kCheckRow0 = 4,       // the row of suffix 0 (written by checkBwtKernel)
    kCheckCount = 5
};

// Grid of the check kernels: grid-stride loops over at most 2^31 - 1 entries.
constexpr uint32_t kCheckThreads = 256;
constexpr uint32_t kCheckBlocks = 4096;

// inverse[p] = the largest i with sa[i] = p (inverse filled with -1 beforehand): of equal entries
// the last owns the slot, as in a sequential pass, so the first duplicate reported is determined.
__global__ void checkRangeKernel(const int32_t* __restrict__ sa, uint32_t n, int32_t* __restrict__ inverse, uint32_t* __restrict__ first) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t p = sa[i];
        if (p < 0 || static_cast<uint32_t>(p) >= n) {
            atomicMin(&first[kCheckRange], i);
        } else {
            atomicMax(&inverse[p], static_cast<int32_t>(i));
        }
    }
}

// After checkRangeKernel in stream order; does nothing if an entry was out of range (the inverse
// is then incomplete and entries may index past it). The order check counts only if no entry is
// duplicated: the host reports the first failing category.
__global__ void checkOrderKernel(const uint8_t* __restrict__ text, const int32_t* __restrict__ sa, uint32_t n, const int32_t* __restrict__ inverse, uint32_t* __restrict__ first) {
    if (first[kCheckRange] != UINT32_MAX) {
        return;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t left = sa[i];
        if (inverse[left] != static_cast<int32_t>(i)) {
            atomicMin(&first[kCheckDuplicate], i);
        }
        if (i + 1 < n) {
            const int32_t right = sa[i + 1];
            bool inOrder = text[left] < text[right];
            if (text[left] == text[right]) {
                const int32_t leftNext = static_cast<uint32_t>(left) + 1 < n ? inverse[left + 1] : -1;
                const int32_t rightNext = static_cast<uint32_t>(right) + 1 < n ? inverse[right + 1] : -1;
                inOrder = leftNext < rightNext;
            }
            if (!inOrder) {
                atomicMin(&first[kCheckOrder], i);
            }
        }
    }
}

// The BWT against the suffix array, after checkOrderKernel; does nothing unless the array is a
// permutation (row0 = inverse[0], the row of suffix 0, is then defined).
__global__ void checkBwtKernel(const uint8_t* __restrict__ text, const int32_t* __restrict__ sa, uint32_t n, const int32_t* __restrict__ inverse, const uint8_t* __restrict__ bwt, uint32_t* __restrict__ first) {
    if (first[kCheckRange] != UINT32_MAX || first[kCheckDuplicate] != UINT32_MAX) {
        return;
    }
    const uint32_t row0 = static_cast<uint32_t>(inverse[0]);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        first[kCheckRow0] = row0;
    }
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const int32_t p = sa[r];
        uint32_t index;
        uint8_t expected;
        if (p == 0) {
            index = 0;
            expected = text[n - 1];
        } else {
            index = r < row0 ? r + 1 : r;
            expected = text[p - 1];
        }
        if (bwt[index] != expected) {
            atomicMin(&first[kCheckBwt], index);
        }
    }
}

}  // namespace sadcs
