/*
 * gpusa.cpp - The reference's C++ GPU interface (include/gpusa.h, after src/cuda/gpusa.h) on top of
 * the C ABI of include/sadcs_hip.h. Status values map one to one; stats are converted field by field.
 */
#include "gpusa.h"

#include "sadcs_hip.h"

struct GpuContext {
    sadcs_ctx* handle = nullptr;
};

namespace {
    GpuStatus statusOf(int s) {
        switch (s) {
            case SADCS_OK:
                return GpuStatus::Ok;
            case SADCS_UNAVAILABLE:
                return GpuStatus::Unavailable;
            case SADCS_NOT_ENOUGH_MEMORY:
                return GpuStatus::NotEnoughMemory;
            default:
                return GpuStatus::Failed;
        }
    }

    void convert(const sadcs_stats& in, GpuStats* out) {
        if (!out) {
            return;
        }
        out->deviceBytes = in.device_bytes;
        out->rankingBytes = in.ranking_bytes;
        out->classSortBytes = in.class_sort_bytes;
        out->mergeBytes = in.merge_bytes;
        out->pinned = in.pinned != 0;
        out->setupSeconds = in.setup_seconds;
        out->uploadSeconds = in.upload_seconds;
        out->deviceSeconds = in.device_seconds;
        out->rankingSeconds = in.ranking_seconds;
        out->classSortSeconds = in.class_sort_seconds;
        out->mergeSeconds = in.merge_seconds;
        out->bwtSeconds = in.bwt_seconds;
        out->downloadSeconds = in.download_seconds;
        out->characterPasses = in.character_passes;
        out->doublingRounds = in.doubling_rounds;
    }

    GpuStatus run(GpuContext* context, const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt, int32_t* primary, GpuStats* stats) {
        sadcs_stats s{};
        const int status = sadcs_hip_build(context ? context->handle : nullptr, text, n, sa, bwt, primary, &s);
        convert(s, stats);
        return statusOf(status);
    }
}

GpuStatus gpuQueryDevice(GpuDevice& device) {
    sadcs_device d{};
    const int status = sadcs_hip_query(&d);
    for (size_t i = 0; i < sizeof device.name; i++) {
        device.name[i] = d.name[i];
    }
    device.computeMajor = d.compute_major;
    device.computeMinor = d.compute_minor;
    device.totalBytes = d.total_bytes;
    device.freeBytes = d.free_bytes;
    return statusOf(status);
}

int64_t gpuRequiredBytes(int32_t n) {
    return sadcs_hip_required_bytes(n);
}

GpuStatus gpuCreateContextWithCover(int32_t capacity, int cover, GpuContext*& context) {
    context = nullptr;
    sadcs_params params{};
    params.cover = cover;
    sadcs_ctx* handle = nullptr;
    const int status = sadcs_hip_ctx_create(capacity, &params, &handle);
    if (status == SADCS_OK) {
        context = new GpuContext;
        context->handle = handle;
    }
    return statusOf(status);
}

GpuStatus gpuCreateContext(int32_t capacity, GpuContext*& context) {
    return gpuCreateContextWithCover(capacity, 32, context);
}

void gpuDestroyContext(GpuContext* context) {
    if (context) {
        sadcs_hip_ctx_destroy(context->handle);
        delete context;
    }
}

GpuStatus gpuSuffixArray(GpuContext* context, const uint8_t* text, int32_t n, int32_t* sa, GpuStats* stats) {
    return run(context, text, n, sa, nullptr, nullptr, stats);
}

GpuStatus gpuSuffixArray(const uint8_t* text, int32_t n, int32_t* sa, GpuStats* stats) {
    return run(nullptr, text, n, sa, nullptr, nullptr, stats);
}

GpuStatus gpuBwt(GpuContext* context, const uint8_t* text, int32_t n, uint8_t* bwt, int32_t& primary, GpuStats* stats) {
    return run(context, text, n, nullptr, bwt, &primary, stats);
}

GpuStatus gpuBwt(const uint8_t* text, int32_t n, uint8_t* bwt, int32_t& primary, GpuStats* stats) {
    return run(nullptr, text, n, nullptr, bwt, &primary, stats);
}

GpuStatus gpuSuffixArrayAndBwt(GpuContext* context, const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt, int32_t& primary, GpuStats* stats) {
    return run(context, text, n, sa, bwt, &primary, stats);
}

GpuStatus gpuRankSample(const uint8_t* text, int32_t n, uint32_t* ranks, GpuStats* stats) {
    sadcs_stats s{};
    const int status = sadcs_hip_rank_sample(text, n, ranks, &s);
    convert(s, stats);
    return statusOf(status);
}

void* gpuAllocatePinned(size_t bytes) {
    return sadcs_hip_alloc_pinned(bytes);
}

void gpuFreePinned(void* pointer) {
    sadcs_hip_free_pinned(pointer);
}

const char* gpuLastError() {
    return sadcs_hip_last_error();
}
