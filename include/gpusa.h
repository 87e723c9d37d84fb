/*
 * gpusa.h - The C++ interface of the GPU construction, as the reference declares it
 * (src/cuda/gpusa.h:22-123): the same names, argument meanings, refusal rules and status values,
 * so callers written against the reference compile and behave the same. Implemented in
 * sadcs_amd/csrc/gpusa.cpp over the C ABI of sadcs_hip.h; plain C++, no HIP types.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include "../sadcs_amd/csrc/cover.h"

enum class GpuStatus {
    Ok,
    Unavailable,      // no device or no usable driver
    NotEnoughMemory,  // the input doesn't fit in free device memory or the context
    Failed,           // a HIP call failed; gpuLastError() has the details
};

struct GpuDevice {
    char name[256];
    int computeMajor;
    int computeMinor;
    int64_t totalBytes;
    int64_t freeBytes;
};

struct GpuStats {
    int64_t deviceBytes = 0;
    int64_t rankingBytes = 0;
    int64_t classSortBytes = 0;
    int64_t mergeBytes = 0;
    bool pinned = false;
    double setupSeconds = 0.0;
    double uploadSeconds = 0.0;
    double deviceSeconds = 0.0;
    double rankingSeconds = 0.0;
    double classSortSeconds = 0.0;
    double mergeSeconds = 0.0;
    double bwtSeconds = 0.0;
    double downloadSeconds = 0.0;
    int32_t characterPasses = 0;
    int32_t doublingRounds = 0;
};

// The cover of the rank-sample entry point (the GPU variant's, gpusa.h:55).
using GpuCover = sadcs::DifferenceCover<32>;

inline int64_t gpuSampleSize(int32_t n) {
    return n > 0 ? GpuCover::pack(n) : 0;
}

// ranks[GpuCover::pack(p)] = the number of sample suffixes smaller than the one at sample position p.
inline void gpuSampleRanksFromSuffixArray(const int32_t* sa, int32_t n, uint32_t* ranks) {
    uint32_t rank = 0;
    for (int32_t i = 0; i < n; i++) {
        if (GpuCover::isSample(sa[i])) {
            ranks[GpuCover::pack(sa[i])] = rank++;
        }
    }
}

GpuStatus gpuQueryDevice(GpuDevice& device);

int64_t gpuRequiredBytes(int32_t n);

constexpr int64_t kGpuMemoryReserve = int64_t(256) << 20;

struct GpuContext;

// Cover 32 by default; gpuCreateContextWithCover selects 128.
GpuStatus gpuCreateContext(int32_t capacity, GpuContext*& context);
GpuStatus gpuCreateContextWithCover(int32_t capacity, int cover, GpuContext*& context);
void gpuDestroyContext(GpuContext* context);

GpuStatus gpuSuffixArray(GpuContext* context, const uint8_t* text, int32_t n, int32_t* sa, GpuStats* stats = nullptr);
GpuStatus gpuSuffixArray(const uint8_t* text, int32_t n, int32_t* sa, GpuStats* stats = nullptr);

GpuStatus gpuBwt(GpuContext* context, const uint8_t* text, int32_t n, uint8_t* bwt, int32_t& primary, GpuStats* stats = nullptr);
GpuStatus gpuBwt(const uint8_t* text, int32_t n, uint8_t* bwt, int32_t& primary, GpuStats* stats = nullptr);

// Both outputs from one merge (no counterpart in the reference, which builds one or the other).
GpuStatus gpuSuffixArrayAndBwt(GpuContext* context, const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt,
                               int32_t& primary, GpuStats* stats = nullptr);

GpuStatus gpuRankSample(const uint8_t* text, int32_t n, uint32_t* ranks, GpuStats* stats = nullptr);

void* gpuAllocatePinned(size_t bytes);
void gpuFreePinned(void* pointer);

const char* gpuLastError();
