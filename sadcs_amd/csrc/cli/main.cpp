/*
 * main.cpp - The sadcs command line on the MI355X construction (the reference's src/cli/main.cpp:
 * commands build|-b, check|-c, test|-t, bench, selftest with --bwt, --gpu, --cover, --threads,
 * --runs, --context, --pinned, --recursive, --lowmem, --naive, --verbose).
 *
 * bench --batch builds every file as one batch (sadcs_hip_batch_build) and compares it with one
 * file at a time; bench --serve submits every file at once to the continuous-batching server
 * (sadcs_hip_server_*). build --pinned writes the output files straight from pinned host buffers.
 * None of the three has a counterpart in the reference.
 *
 * --device (check and test, suffix arrays) runs the fast check on the GPU instead of the host
 * (sadcs_hip_check_device: the same checks, the same first failing index); it has no counterpart
 * in the reference.
 *
 * Every construction runs on the GPU: --gpu is accepted and implied (the reference's CPU variant
 * is not part of this library). --cover selects the GPU's difference cover, 32 (default) or 128.
 * --both builds the suffix array and the BWT in one merge and writes <out> and <out>.bwt.
 * Sizes are MiB and rates MiB/s, as the reference prints them (main.cpp:175-177).
 */
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <memory>
#include <string>
#include <vector>

#include "../checker.h"
#include "../fileio.h"
#include "gpusa.h"
#include "sadcs_hip.h"
#include "selftest.h"

namespace {
    constexpr int kOk = 0;
    constexpr int kError = 1;
    constexpr size_t kPadding = 64;
    constexpr int64_t kMaxText = INT32_MAX - 1024;

    struct Options {
        const char* paths[2] = {};
        int pathCount = 0;
        int cover = 32;
        CheckMode checkMode = CheckMode::Fast;
        bool bwt = false;
        bool both = false;
        bool verbose = false;
        bool recursive = false;
        bool context = false;
        bool pinned = false;
        bool deviceCheck = false;
        bool batch = false;
        bool serve = false;
        int runs = 5;
    };

    void usage() {
        printf(
            "sadcs - difference-cover suffix array and BWT construction on an MI355X GPU\n"
            "\n"
            "usage:\n"
            "  sadcs build|-b <in> <out.sa> [--cover 32|128] [--gpu]\n"
            "  sadcs build|-b --bwt <in> <out.bwt> [--cover 32|128] [--gpu]\n"
            "  sadcs build|-b --both <in> <out.sa>      (also writes <out.sa>.bwt, one merge)\n"
            "  sadcs build|-b ... --pinned              (outputs in pinned host memory, written from there)\n"
            "  sadcs check|-c <in> <in.sa> [--lowmem | --naive | --device] [--threads N]\n"
            "  sadcs check|-c --bwt <in> <in.bwt>\n"
            "  sadcs test|-t <in|dir> [--bwt] [--cover 32|128] [--threads N] [--recursive] [--device]\n"
            "  sadcs bench <in|dir> [--bwt | --both] [--runs N] [--context] [--pinned] [--recursive]\n"
            "  sadcs bench <dir> --batch | --serve [--bwt | --both] [--runs N] [--recursive]\n"
            "  sadcs selftest [--verbose]\n"
            "\n"
            "The .sa file is n little-endian int32 entries (divsufsort's mksary format).\n"
            "The .bwt file is divsufsort's bwt format as one block: n, the primary index, the n bytes.\n"
            "test builds on the GPU and checks the result with the independent checker.\n"
            "--device runs the suffix-array check on the GPU (text and array uploaded; 4n bytes of device memory).\n"
            "bench times the build host to host: a warmup, then the best of --runs, with each phase.\n"
            "--context reuses one GPU context per file; --pinned uses page-locked host buffers.\n"
            "--batch builds all files in one batched construction; --serve submits them all at once to the\n"
            "serving loop; both compare with one file at a time and check that the outputs are equal.\n");
    }

    bool parse(int argc, char** argv, Options& o) {
        for (int i = 2; i < argc; i++) {
            const char* a = argv[i];
            if (!strcmp(a, "--cover") && i + 1 < argc) {
                o.cover = atoi(argv[++i]);
                if (o.cover != 32 && o.cover != 128) {
                    printf("sadcs: --cover must be 32 or 128 on the GPU\n");
                    return false;
                }
            } else if (!strcmp(a, "--threads") && i + 1 < argc) {
                const int t = atoi(argv[++i]);
                if (t < 1) {
                    printf("sadcs: --threads must be at least 1\n");
                    return false;
                }
                omp_set_num_threads(t);
            } else if (!strcmp(a, "--lowmem")) {
                o.checkMode = CheckMode::LowMemory;
            } else if (!strcmp(a, "--naive")) {
                o.checkMode = CheckMode::Naive;
            } else if (!strcmp(a, "--device")) {
                o.deviceCheck = true;
            } else if (!strcmp(a, "--batch")) {
                o.batch = true;
            } else if (!strcmp(a, "--serve")) {
                o.serve = true;
            } else if (!strcmp(a, "--bwt")) {
                o.bwt = true;
            } else if (!strcmp(a, "--both")) {
                o.both = true;
            } else if (!strcmp(a, "--gpu")) {
                // implied
            } else if (!strcmp(a, "--runs") && i + 1 < argc) {
                o.runs = atoi(argv[++i]);
                if (o.runs < 1) {
                    printf("sadcs: --runs must be at least 1\n");
                    return false;
                }
            } else if (!strcmp(a, "--context")) {
                o.context = true;
            } else if (!strcmp(a, "--pinned")) {
                o.pinned = true;
            } else if (!strcmp(a, "--recursive")) {
                o.recursive = true;
            } else if (!strcmp(a, "--verbose")) {
                o.verbose = true;
            } else if (!strncmp(a, "--", 2)) {
                printf("sadcs: unknown option '%s'\n", a);
                return false;
            } else if (o.pathCount < 2) {
                o.paths[o.pathCount++] = a;
            } else {
                printf("sadcs: unexpected argument '%s'\n", a);
                return false;
            }
        }
        return true;
    }

    bool loadText(const char* path, std::vector<uint8_t>& text, int32_t& n) {
        if (!readWholeFile(path, text, kPadding)) {
            printf("sadcs: cannot read '%s'\n", path);
            return false;
        }
        const int64_t size = static_cast<int64_t>(text.size()) - static_cast<int64_t>(kPadding);
        if (size > kMaxText) {
            printf("sadcs: '%s' is too large (%lld bytes, at most %lld supported)\n", path, static_cast<long long>(size),
                   static_cast<long long>(kMaxText));
            return false;
        }
        n = static_cast<int32_t>(size);
        return true;
    }

    double mibPerSecond(int64_t bytes, double seconds) {
        return seconds > 0 ? static_cast<double>(bytes) / (1024.0 * 1024.0) / seconds : 0.0;
    }

    bool report(GpuStatus s) {
        if (s != GpuStatus::Ok) {
            printf("sadcs: GPU %s: %s\n",
                   s == GpuStatus::Unavailable ? "unavailable" : s == GpuStatus::NotEnoughMemory ? "out of memory" : "failed",
                   gpuLastError());
            return false;
        }
        return true;
    }

    // A construction with a temporary context of the requested cover.
    GpuStatus construct(const Options& o, GpuContext* context, const uint8_t* text, int32_t n, int32_t* sa, uint8_t* bwt,
                        int32_t& primary, GpuStats& stats) {
        std::unique_ptr<GpuContext, void (*)(GpuContext*)> own(nullptr, gpuDestroyContext);
        if (!context && o.cover != 32 && n > 1) {
            GpuContext* c = nullptr;
            const GpuStatus s = gpuCreateContextWithCover(n, o.cover, c);
            if (s != GpuStatus::Ok) {
                return s;
            }
            own.reset(c);
            context = c;
        }
        if (sa && bwt) {
            return gpuSuffixArrayAndBwt(context, text, n, sa, bwt, primary, &stats);
        }
        if (bwt) {
            return gpuBwt(context, text, n, bwt, primary, &stats);
        }
        return gpuSuffixArray(context, text, n, sa, &stats);
    }

    void printFailure(const uint8_t* text, int32_t n, const int32_t* sa, const CheckResult& r) {
        printf("sadcs: check FAILED at SA[%lld]: %s\n", static_cast<long long>(r.index), r.reason);
        for (int64_t i = r.index; i >= 0 && i < n && i <= r.index + 1; i++) {
            printf("  SA[%lld] = %d:", static_cast<long long>(i), sa[i]);
            for (int32_t p = sa[i]; p >= 0 && p < n && p < sa[i] + 24; p++) {
                const uint8_t c = text[p];
                if (c >= 0x20 && c < 0x7F) {
                    printf("%c", c);
                } else {
                    printf("\\x%02X", c);
                }
            }
            printf("\n");
        }
    }

    int cmdBuild(const Options& o) {
        if (o.pathCount != 2) {
            usage();
            return kError;
        }
        std::vector<uint8_t> text;
        int32_t n;
        if (!loadText(o.paths[0], text, n)) {
            return kError;
        }
        const double start = omp_get_wtime();
        GpuStats stats;
        int32_t primary = 0;
        const bool wantSA = !o.bwt || o.both;
        // Outputs left uninitialised (every entry is written); with --pinned in page-locked memory,
        // so the device writes them with one DMA each and the files are written from there.
        using Host = std::unique_ptr<void, void (*)(void*)>;
        auto hostBuffer = [&](size_t bytes) {
            if (o.pinned) {
                return Host(bytes ? gpuAllocatePinned(bytes) : nullptr, gpuFreePinned);
            }
            return Host(bytes ? ::operator new(bytes) : nullptr, [](void* q) { ::operator delete(q); });
        };
        Host saBuffer = hostBuffer(wantSA ? static_cast<size_t>(n) * sizeof(int32_t) : 0);
        Host bwtBuffer = hostBuffer(o.both ? static_cast<size_t>(n) : 0);
        if ((wantSA && n > 0 && !saBuffer) || (o.both && n > 0 && !bwtBuffer)) {
            printf("sadcs: cannot allocate %shost memory for the outputs\n", o.pinned ? "pinned " : "");
            return kError;
        }
        int32_t* sa = static_cast<int32_t*>(saBuffer.get());
        // --bwt alone writes the BWT in place of the host text: n bytes in all (main.cpp:351-361).
        uint8_t* bwtOut = o.both ? static_cast<uint8_t*>(bwtBuffer.get()) : o.bwt ? text.data() : nullptr;
        if (!report(construct(o, nullptr, text.data(), n, wantSA ? sa : nullptr, bwtOut, primary, stats))) {
            return kError;
        }
        const double seconds = omp_get_wtime() - start;
        bool written;
        if (o.bwt && !o.both) {
            written = writeBwtFile(o.paths[1], text.data(), n, primary);
        } else {
            written = writeWholeFile(o.paths[1], sa, static_cast<size_t>(n) * sizeof(int32_t));
            if (o.both) {
                const std::string bwtPath = std::string(o.paths[1]) + ".bwt";
                written = written && writeBwtFile(bwtPath.c_str(), bwtOut, n, primary);
            }
        }
        if (!written) {
            printf("sadcs: cannot write '%s'\n", o.paths[1]);
            return kError;
        }
        printf("%s -> %s: %d bytes, GPU, cover %d, %.3fs (%.1f MB/s), peak device memory %.1f MiB (%.2f bytes per input byte)\n",
               o.paths[0], o.paths[1], n, o.cover, seconds, mibPerSecond(n, seconds), stats.deviceBytes / (1024.0 * 1024.0),
               n > 0 ? static_cast<double>(stats.deviceBytes) / n : 0.0);
        return kOk;
    }

    // The fast check on the GPU: uploads the text and the array, returns what checkSuffixArray's
    // fast mode returns (the first failing index and the reason).
    CheckResult deviceCheckSuffixArray(const uint8_t* text, int32_t n, const int32_t* sa, bool& failedToRun) {
        failedToRun = false;
        if (n <= 0) {
            return {true, -1, nullptr};
        }
        using Device = std::unique_ptr<void, void (*)(void*)>;
        Device dText(sadcs_hip_device_alloc(static_cast<size_t>(n)), sadcs_hip_device_free);
        Device dSA(sadcs_hip_device_alloc(static_cast<size_t>(n) * 4), sadcs_hip_device_free);
        int64_t result[2] = {0, -1};
        if (!dText || !dSA || sadcs_hip_memcpy(dText.get(), text, static_cast<size_t>(n), 1) != SADCS_OK ||
            sadcs_hip_memcpy(dSA.get(), sa, static_cast<size_t>(n) * 4, 1) != SADCS_OK ||
            sadcs_hip_check_device(dText.get(), n, dSA.get(), nullptr, 0, result) != SADCS_OK) {
            printf("sadcs: device check could not run: %s\n", dText && dSA ? sadcs_hip_last_error() : "device allocation failed");
            failedToRun = true;
            return {false, -1, "device check could not run"};
        }
        switch (result[0]) {
            case SADCS_CHECK_PASS:
                return {true, -1, nullptr};
            case SADCS_CHECK_RANGE:
                return {false, result[1], "SA[index] is not a position of the text"};
            case SADCS_CHECK_DUPLICATE:
                return {false, result[1], "SA[index] repeats another entry"};
            default:
                return {false, result[1], "suffix SA[index] is not smaller than suffix SA[index + 1]"};
        }
    }

    CheckResult checkArray(const Options& o, const uint8_t* text, int32_t n, const int32_t* sa, bool& failedToRun) {
        failedToRun = false;
        return o.deviceCheck ? deviceCheckSuffixArray(text, n, sa, failedToRun) : checkSuffixArray(text, n, sa, o.checkMode);
    }

    int cmdCheck(const Options& o) {
        if (o.pathCount != 2) {
            usage();
            return kError;
        }
        std::vector<uint8_t> text;
        int32_t n;
        if (!loadText(o.paths[0], text, n)) {
            return kError;
        }
        const double start = omp_get_wtime();
        if (o.bwt) {
            std::vector<uint8_t> bwt;
            int32_t primary;
            if (!readBwtFile(o.paths[1], bwt, primary)) {
                printf("sadcs: cannot read '%s' as a single-block BWT file\n", o.paths[1]);
                return kError;
            }
            if (static_cast<int64_t>(bwt.size()) != n) {
                printf("sadcs: '%s' has %zu bytes of BWT, but '%s' has %d bytes\n", o.paths[1], bwt.size(), o.paths[0], n);
                return kError;
            }
            const CheckResult r = checkBwt(text.data(), n, bwt.data(), primary);
            if (!r.ok) {
                printf("sadcs: BWT check FAILED (primary index %d): %s, index %lld\n", primary, r.reason,
                       static_cast<long long>(r.index));
                return kError;
            }
            printf("%s: BWT OK (%d bytes, primary index %d, %.3fs)\n", o.paths[1], n, primary, omp_get_wtime() - start);
            return kOk;
        }
        std::vector<int32_t> sa;
        if (!readSuffixArray(o.paths[1], sa)) {
            printf("sadcs: cannot read '%s' as a suffix array\n", o.paths[1]);
            return kError;
        }
        if (static_cast<int64_t>(sa.size()) != n) {
            printf("sadcs: '%s' has %zu entries, but '%s' has %d bytes\n", o.paths[1], sa.size(), o.paths[0], n);
            return kError;
        }
        bool failedToRun;
        const CheckResult r = checkArray(o, text.data(), n, sa.data(), failedToRun);
        if (failedToRun) {
            return kError;
        }
        if (!r.ok) {
            printFailure(text.data(), n, sa.data(), r);
            return kError;
        }
        printf("%s: suffix array OK (%d entries%s, %.3fs)\n", o.paths[1], n, o.deviceCheck ? ", checked on the GPU" : "",
               omp_get_wtime() - start);
        return kOk;
    }

    struct TestRun {
        int64_t bytes = 0;
        double gpuSeconds = 0, checkSeconds = 0, peak = 0;
        GpuStats gpu;
        bool ok = true;
    };

    TestRun runTest(std::vector<uint8_t>& text, int32_t n, const Options& o) {
        TestRun run;
        run.bytes = n;
        const double start = omp_get_wtime();
        int32_t primary = 0;
        if (o.bwt) {
            std::vector<uint8_t> bwt(static_cast<size_t>(n));
            if (!report(construct(o, nullptr, text.data(), n, nullptr, bwt.data(), primary, run.gpu))) {
                run.ok = false;
                return run;
            }
            run.gpuSeconds = omp_get_wtime() - start;
            const double checkStart = omp_get_wtime();
            const CheckResult r = checkBwt(text.data(), n, bwt.data(), primary);
            run.checkSeconds = omp_get_wtime() - checkStart;
            if (!r.ok) {
                printf("sadcs: BWT check FAILED (primary index %d): %s, index %lld\n", primary, r.reason, static_cast<long long>(r.index));
            }
            run.ok = r.ok;
        } else {
            std::vector<int32_t> sa(static_cast<size_t>(n));
            if (!report(construct(o, nullptr, text.data(), n, sa.data(), nullptr, primary, run.gpu))) {
                run.ok = false;
                return run;
            }
            run.gpuSeconds = omp_get_wtime() - start;
            const double checkStart = omp_get_wtime();
            bool failedToRun;
            const CheckResult r = checkArray(o, text.data(), n, sa.data(), failedToRun);
            run.checkSeconds = omp_get_wtime() - checkStart;
            if (!r.ok && !failedToRun) {
                printFailure(text.data(), n, sa.data(), r);
            }
            run.ok = r.ok;
        }
        run.peak = n > 0 ? static_cast<double>(run.gpu.deviceBytes) / n : 0.0;
        return run;
    }

    bool listFiles(const std::filesystem::path& dir, bool recursive, std::vector<std::filesystem::path>& files) {
        std::error_code error;
        auto collect = [&](auto it) {
            for (const decltype(it) end{}; !error && it != end; it.increment(error)) {
                std::error_code ignored;
                if (it->is_regular_file(ignored)) {
                    files.push_back(it->path());
                }
            }
        };
        if (recursive) {
            collect(std::filesystem::recursive_directory_iterator(dir, error));
        } else {
            collect(std::filesystem::directory_iterator(dir, error));
        }
        std::sort(files.begin(), files.end());
        return !error;
    }

    int cmdTest(const Options& o) {
        if (o.pathCount != 1) {
            usage();
            return kError;
        }
        std::error_code error;
        std::vector<std::filesystem::path> files;
        const bool directory = std::filesystem::is_directory(o.paths[0], error);
        if (directory) {
            if (!listFiles(o.paths[0], o.recursive, files)) {
                printf("sadcs: cannot list directory '%s'\n", o.paths[0]);
                return kError;
            }
        } else {
            files.push_back(o.paths[0]);
        }
        printf("%s: %zu file%s, %s, GPU cover %d, checker on %d threads (times in seconds)\n", o.paths[0], files.size(),
               files.size() == 1 ? "" : "s", o.bwt ? "BWT" : "suffix array", o.cover, omp_get_max_threads());
        printf("%-32s %12s %8s %8s %8s %8s %6s %8s\n", "file", "bytes", "rank", "class", "merge", "gpu", "peak/n", "check");
        size_t passed = 0;
        for (const auto& file : files) {
            std::vector<uint8_t> text;
            int32_t n;
            if (!loadText(file.string().c_str(), text, n)) {
                continue;
            }
            const TestRun run = runTest(text, n, o);
            const std::string name = directory ? file.lexically_relative(o.paths[0]).generic_string() : file.filename().string();
            printf("%-32s %12lld %8.3f %8.3f %8.3f %8.3f %6.2f %8.3f  %s\n", name.c_str(), static_cast<long long>(run.bytes),
                   run.gpu.rankingSeconds, run.gpu.classSortSeconds, run.gpu.mergeSeconds, run.gpuSeconds, run.peak,
                   run.checkSeconds, run.ok ? "OK" : "FAILED");
            passed += run.ok ? 1 : 0;
        }
        printf("%zu/%zu OK\n", passed, files.size());
        return passed == files.size() ? kOk : kError;
    }

    struct BenchRun {
        int64_t bytes = 0;
        double seconds = 0;
        GpuStats stats;
    };

    bool benchOne(const uint8_t* text, int32_t n, const Options& o, BenchRun& best) {
        GpuContext* context = nullptr;
        if (o.context && !report(gpuCreateContextWithCover(n, o.cover, context))) {
            return false;
        }
        std::unique_ptr<GpuContext, void (*)(GpuContext*)> ownContext(context, gpuDestroyContext);
        const bool wantSA = !o.bwt || o.both;
        const bool wantBWT = o.bwt || o.both;
        using Pinned = std::unique_ptr<void, void (*)(void*)>;
        Pinned pinnedText(o.pinned ? gpuAllocatePinned(static_cast<size_t>(n)) : nullptr, gpuFreePinned);
        Pinned pinnedSA(o.pinned && wantSA ? gpuAllocatePinned(static_cast<size_t>(n) * 4) : nullptr, gpuFreePinned);
        Pinned pinnedBWT(o.pinned && wantBWT ? gpuAllocatePinned(static_cast<size_t>(n)) : nullptr, gpuFreePinned);
        std::vector<int32_t> sa(!o.pinned && wantSA ? static_cast<size_t>(n) : 0);
        std::vector<uint8_t> bwt(!o.pinned && wantBWT ? static_cast<size_t>(n) : 0);
        const uint8_t* input = text;
        int32_t* saOut = wantSA ? (o.pinned ? static_cast<int32_t*>(pinnedSA.get()) : sa.data()) : nullptr;
        uint8_t* bwtOut = wantBWT ? (o.pinned ? static_cast<uint8_t*>(pinnedBWT.get()) : bwt.data()) : nullptr;
        if (o.pinned) {
            if (!pinnedText || (wantSA && !saOut) || (wantBWT && !bwtOut)) {
                printf("sadcs: cannot allocate pinned host memory\n");
                return false;
            }
            memcpy(pinnedText.get(), text, static_cast<size_t>(n));
            input = static_cast<const uint8_t*>(pinnedText.get());
        }
        best = BenchRun{};
        best.bytes = n;
        for (int run = 0; run <= o.runs; run++) {
            GpuStats stats;
            int32_t primary = 0;
            const double start = omp_get_wtime();
            const GpuStatus s = construct(o, context, input, n, saOut, bwtOut, primary, stats);
            const double seconds = omp_get_wtime() - start;
            if (!report(s)) {
                return false;
            }
            if (run > 0 && (run == 1 || seconds < best.seconds)) {
                best.seconds = seconds;
                best.stats = stats;
            }
        }
        return true;
    }

    struct Corpus {
        std::vector<std::string> names;
        std::vector<uint8_t> joined;      // the texts back to back
        std::vector<int64_t> lengths;
        std::vector<int64_t> offsets;
        int64_t total = 0;
    };

    bool loadCorpus(const std::vector<std::filesystem::path>& files, Corpus& c) {
        for (const auto& file : files) {
            std::vector<uint8_t> text;
            int32_t n;
            if (!loadText(file.string().c_str(), text, n)) {
                return false;
            }
            if (n < 1) {
                continue;  // empty files have nothing to build
            }
            c.names.push_back(file.string());
            c.offsets.push_back(c.total);
            c.lengths.push_back(n);
            c.joined.insert(c.joined.end(), text.begin(), text.begin() + n);
            c.total += n;
        }
        return true;
    }

    // One file at a time through the reference API, best of the runs per file; outputs kept for
    // the comparison.
    bool oneAtATime(const Options& o, const Corpus& c, std::vector<int32_t>& sa, std::vector<uint8_t>& bwt, std::vector<int32_t>& primaries, double& seconds) {
        const bool wantSA = !o.bwt || o.both, wantBWT = o.bwt || o.both;
        sa.resize(wantSA ? c.total : 0);
        bwt.resize(wantBWT ? c.total : 0);
        primaries.assign(c.lengths.size(), -1);
        seconds = 0;
        for (size_t i = 0; i < c.lengths.size(); i++) {
            const int32_t n = static_cast<int32_t>(c.lengths[i]);
            double best = 1e30;
            for (int run = 0; run <= o.runs; run++) {
                GpuStats stats;
                int32_t primary = 0;
                const double start = omp_get_wtime();
                if (!report(construct(o, nullptr, c.joined.data() + c.offsets[i], n, wantSA ? sa.data() + c.offsets[i] : nullptr,
                                      wantBWT ? bwt.data() + c.offsets[i] : nullptr, primary, stats))) {
                    return false;
                }
                if (run > 0) {
                    best = std::min(best, omp_get_wtime() - start);
                }
                primaries[i] = primary;
            }
            seconds += best;
        }
        return true;
    }

    bool sameOutputs(const std::vector<int32_t>& sa, const std::vector<uint8_t>& bwt, const std::vector<int32_t>& primaries,
                     const int32_t* sa2, const uint8_t* bwt2, const int32_t* primaries2) {
        return (sa.empty() || memcmp(sa.data(), sa2, sa.size() * sizeof(int32_t)) == 0) &&
               (bwt.empty() || (memcmp(bwt.data(), bwt2, bwt.size()) == 0 &&
                                memcmp(primaries.data(), primaries2, primaries.size() * sizeof(int32_t)) == 0));
    }

    int benchBatch(const Options& o, const Corpus& c, const std::vector<int32_t>& seqSA, const std::vector<uint8_t>& seqBWT,
                   const std::vector<int32_t>& seqPrimaries, double seqSeconds) {
        const bool wantSA = !o.bwt || o.both, wantBWT = o.bwt || o.both;
        std::vector<int32_t> sa(wantSA ? c.total : 0);
        std::vector<uint8_t> bwt(wantBWT ? c.total : 0);
        std::vector<int32_t> primaries(c.lengths.size(), -1);
        double best = 1e30;
        for (int run = 0; run <= o.runs; run++) {
            sadcs_stats stats{};
            const double start = omp_get_wtime();
            if (sadcs_hip_batch_build(nullptr, c.joined.data(), c.lengths.data(), static_cast<int32_t>(c.lengths.size()),
                                      wantSA ? sa.data() : nullptr, wantBWT ? bwt.data() : nullptr, primaries.data(), &stats) != SADCS_OK) {
                printf("sadcs: batch build failed: %s\n", sadcs_hip_last_error());
                return kError;
            }
            if (run > 0) {
                best = std::min(best, omp_get_wtime() - start);
            }
        }
        const bool equal = sameOutputs(seqSA, seqBWT, seqPrimaries, sa.data(), bwt.data(), primaries.data());
        printf("batch of %zu files, %lld bytes: %.1f ms (%.1f MB/s); one at a time %.1f ms (%.1f MB/s): %.2fx; outputs %s\n",
               c.lengths.size(), static_cast<long long>(c.total), best * 1e3, mibPerSecond(c.total, best), seqSeconds * 1e3,
               mibPerSecond(c.total, seqSeconds), seqSeconds / best, equal ? "equal" : "DIFFER");
        return equal ? kOk : kError;
    }

    int benchServe(const Options& o, const Corpus& c, const std::vector<int32_t>& seqSA, const std::vector<uint8_t>& seqBWT,
                   const std::vector<int32_t>& seqPrimaries, double seqSeconds) {
        const bool wantSA = !o.bwt || o.both, wantBWT = o.bwt || o.both;
        sadcs_server_params params{};
        params.cover = o.cover;
        params.want_sa = wantSA;
        params.want_bwt = wantBWT;
        sadcs_server* server = nullptr;
        if (sadcs_hip_server_create(&params, &server) != SADCS_OK) {
            printf("sadcs: cannot start the server: %s\n", sadcs_hip_last_error());
            return kError;
        }
        std::vector<int32_t> sa(wantSA ? c.total : 0);
        std::vector<uint8_t> bwt(wantBWT ? c.total : 0);
        std::vector<int32_t> primaries(c.lengths.size(), -1);
        std::vector<double> latency(c.lengths.size());
        std::vector<int64_t> ids;
        const double start = omp_get_wtime();
        for (size_t i = 0; i < c.lengths.size(); i++) {
            ids.push_back(sadcs_hip_server_submit(server, c.joined.data() + c.offsets[i], c.lengths[i]));
        }
        bool ok = true;
        for (size_t i = 0; i < ids.size(); i++) {
            double queued = 0;
            int32_t batchSize = 0;
            ok = ok && ids[i] >= 0 &&
                 sadcs_hip_server_wait(server, ids[i], wantSA ? sa.data() + c.offsets[i] : nullptr, wantBWT ? bwt.data() + c.offsets[i] : nullptr,
                                       &primaries[i], &latency[i], &queued, &batchSize) == SADCS_OK;
        }
        const double seconds = omp_get_wtime() - start;
        int64_t counters[5] = {};
        sadcs_hip_server_counters(server, counters);
        sadcs_hip_server_destroy(server);
        if (!ok) {
            printf("sadcs: a request failed: %s\n", sadcs_hip_last_error());
            return kError;
        }
        std::sort(latency.begin(), latency.end());
        auto percentile = [&](double q) { return latency[std::min(latency.size() - 1, static_cast<size_t>(q * latency.size()))] * 1e3; };
        const bool equal = sameOutputs(seqSA, seqBWT, seqPrimaries, sa.data(), bwt.data(), primaries.data());
        printf("served %zu files at once, %lld bytes: %.1f ms (%.1f MB/s) in %lld batches; latency p50 %.1f ms, p99 %.1f ms; "
               "one at a time %.1f ms (%.1f MB/s): %.2fx; outputs %s\n",
               c.lengths.size(), static_cast<long long>(c.total), seconds * 1e3, mibPerSecond(c.total, seconds), static_cast<long long>(counters[1]),
               percentile(0.5), percentile(0.99), seqSeconds * 1e3, mibPerSecond(c.total, seqSeconds), seqSeconds / seconds, equal ? "equal" : "DIFFER");
        return equal ? kOk : kError;
    }

    int cmdBench(const Options& o) {
        if (o.pathCount != 1) {
            usage();
            return kError;
        }
        if (o.batch || o.serve) {
            std::error_code error;
            std::vector<std::filesystem::path> files;
            if (std::filesystem::is_directory(o.paths[0], error)) {
                if (!listFiles(o.paths[0], o.recursive, files)) {
                    printf("sadcs: cannot list directory '%s'\n", o.paths[0]);
                    return kError;
                }
            } else {
                files.push_back(o.paths[0]);
            }
            Corpus corpus;
            if (!loadCorpus(files, corpus)) {
                return kError;
            }
            if (corpus.lengths.empty()) {
                printf("sadcs: no non-empty files in '%s'\n", o.paths[0]);
                return kError;
            }
            std::vector<int32_t> sa, primaries;
            std::vector<uint8_t> bwt;
            double seqSeconds = 0;
            if (!oneAtATime(o, corpus, sa, bwt, primaries, seqSeconds)) {
                return kError;
            }
            int status = kOk;
            if (o.batch) {
                status = benchBatch(o, corpus, sa, bwt, primaries, seqSeconds);
            }
            if (o.serve && status == kOk) {
                status = benchServe(o, corpus, sa, bwt, primaries, seqSeconds);
            }
            return status;
        }
        std::error_code error;
        std::vector<std::filesystem::path> files;
        const bool directory = std::filesystem::is_directory(o.paths[0], error);
        if (directory) {
            if (!listFiles(o.paths[0], o.recursive, files)) {
                printf("sadcs: cannot list directory '%s'\n", o.paths[0]);
                return kError;
            }
        } else {
            files.push_back(o.paths[0]);
        }
        GpuDevice device;
        if (!report(gpuQueryDevice(device))) {
            return kError;
        }
        printf("%s: %zu file%s, GPU %s, %s, cover %d, best of %d runs%s%s (times in ms)\n", o.paths[0], files.size(),
               files.size() == 1 ? "" : "s", device.name, o.both ? "suffix array + BWT" : o.bwt ? "BWT" : "suffix array", o.cover,
               o.runs, o.context ? ", context reused" : "", o.pinned ? ", pinned host buffers" : "");
        printf("%-24s %12s %7s %7s %7s %7s %7s %7s %7s %8s %8s %6s\n", "file", "bytes", "setup", "upload", "rank", "class", "merge",
               "bwt", "down", "total", "MB/s", "peak/n");
        for (const auto& file : files) {
            std::vector<uint8_t> text;
            int32_t n;
            BenchRun run;
            if (!loadText(file.string().c_str(), text, n) || !benchOne(text.data(), n, o, run)) {
                return kError;
            }
            const GpuStats& s = run.stats;
            const std::string name = directory ? file.lexically_relative(o.paths[0]).generic_string() : file.filename().string();
            printf("%-24s %12lld %7.1f %7.1f %7.1f %7.1f %7.1f %7.1f %7.1f %8.1f %8.1f %6.2f\n", name.c_str(),
                   static_cast<long long>(run.bytes), s.setupSeconds * 1e3, s.uploadSeconds * 1e3, s.rankingSeconds * 1e3,
                   s.classSortSeconds * 1e3, s.mergeSeconds * 1e3, s.bwtSeconds * 1e3, s.downloadSeconds * 1e3, run.seconds * 1e3,
                   mibPerSecond(run.bytes, run.seconds), n > 0 ? static_cast<double>(s.deviceBytes) / n : 0.0);
        }
        return kOk;
    }
}

int main(int argc, char** argv) {
    if (argc < 2) {
        usage();
        return kError;
    }
    const char* command = argv[1];
    Options o;
    if (!parse(argc, argv, o)) {
        return kError;
    }
    if (!strcmp(command, "build") || !strcmp(command, "-b")) {
        return cmdBuild(o);
    }
    if (!strcmp(command, "check") || !strcmp(command, "-c")) {
        return cmdCheck(o);
    }
    if (!strcmp(command, "test") || !strcmp(command, "-t")) {
        return cmdTest(o);
    }
    if (!strcmp(command, "bench")) {
        return cmdBench(o);
    }
    if (!strcmp(command, "selftest")) {
        return runSelfTest(o.verbose) == 0 ? kOk : kError;
    }
    usage();
    return kError;
}
