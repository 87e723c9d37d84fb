/*
 * selftest.cpp - The GPU cases of the reference's selftest (src/cli/selftest.cpp:432-606), with its
 * generators and seeds (:30-108), and its planted checker errors (:309-385). Without the reference's
 * CPU construction, expected suffix arrays come from a naive sort up to 20000 bytes and from the
 * independent checker above that; every BWT must equal the one read off the checked suffix array.
 */
#include "selftest.h"

#include <algorithm>
#include <bit>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../checker.h"
#include "gpusa.h"

namespace {
    constexpr int32_t kLarge = 1 << 22;
    constexpr int32_t kNaiveLimit = 20000;

    struct Rng {
        uint64_t state;
        uint64_t next() {
            uint64_t z = (state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        }
        uint32_t below(uint32_t bound) {
            return static_cast<uint32_t>(next() % bound);
        }
    };

    using Text = std::vector<uint8_t>;

    Text randomText(Rng& rng, int32_t n, uint32_t sigma) {
        Text t(static_cast<size_t>(n));
        for (uint8_t& c : t) {
            c = static_cast<uint8_t>(sigma == 256 ? rng.below(256) : 'a' + rng.below(sigma));
        }
        return t;
    }

    Text periodicText(Rng& rng, int32_t n, int32_t period) {
        const Text pattern = randomText(rng, period, 4);
        Text t(static_cast<size_t>(n));
        for (int32_t i = 0; i < n; i++) {
            t[i] = pattern[i % period];
        }
        return t;
    }

    Text fibonacciText(int32_t n, uint8_t a, uint8_t b) {
        Text shorter{a}, longer{a, b};
        while (static_cast<int32_t>(longer.size()) < n) {
            Text next(longer);
            next.insert(next.end(), shorter.begin(), shorter.end());
            shorter = std::move(longer);
            longer = std::move(next);
        }
        longer.resize(static_cast<size_t>(n));
        return longer;
    }

    Text thueMorseText(int32_t n) {
        Text t(static_cast<size_t>(n));
        for (int32_t i = 0; i < n; i++) {
            t[i] = static_cast<uint8_t>('a' + (std::popcount(static_cast<uint32_t>(i)) & 1));
        }
        return t;
    }

    Text textOf(const char* s) {
        return Text(s, s + strlen(s));
    }

    Text padded(const Text& t) {
        Text p(t);
        p.resize(t.size() + 64, 0);
        return p;
    }

    std::vector<int32_t> naiveSuffixArray(const uint8_t* t, int32_t n) {
        std::vector<int32_t> sa(static_cast<size_t>(n));
        std::iota(sa.begin(), sa.end(), 0);
        std::sort(sa.begin(), sa.end(), [&](int32_t p, int32_t q) {
            const int32_t comparable = std::min(n - p, n - q);
            const int c = memcmp(t + p, t + q, static_cast<size_t>(comparable));
            return c != 0 ? c < 0 : p > q;
        });
        return sa;
    }

    // selftest.cpp:141-159.
    std::vector<uint8_t> bwtOf(const uint8_t* t, int32_t n, const std::vector<int32_t>& sa, int32_t& primary) {
        std::vector<uint8_t> b;
        primary = 0;
        if (n == 0) {
            return b;
        }
        b.push_back(t[n - 1]);
        for (int32_t row = 0; row < n; row++) {
            if (sa[row] == 0) {
                primary = row + 1;
            } else {
                b.push_back(t[sa[row] - 1]);
            }
        }
        return b;
    }

    struct Suite {
        bool verbose;
        int cases = 0;
        int failures = 0;
        void fail(const std::string& name, const char* what, int64_t index) {
            failures++;
            printf("  FAIL %s: %s (index %lld)\n", name.c_str(), what, static_cast<long long>(index));
        }
    };

    void testGpu(Suite& suite, const std::string& name, const Text& input, GpuContext* context = nullptr) {
        const Text text = padded(input);
        const int32_t n = static_cast<int32_t>(input.size());
        std::vector<int32_t> sa(static_cast<size_t>(n), -1);
        suite.cases++;
        if (gpuSuffixArray(context, text.data(), n, sa.data()) != GpuStatus::Ok) {
            suite.fail(name + " GPU", gpuLastError(), -1);
            return;
        }
        if (n <= kNaiveLimit) {
            const std::vector<int32_t> expected = naiveSuffixArray(text.data(), n);
            const auto m = std::mismatch(expected.begin(), expected.end(), sa.begin());
            if (m.first != expected.end()) {
                suite.fail(name + " GPU", "differs from the naive suffix array", m.first - expected.begin());
                return;
            }
        } else {
            const CheckResult r = checkSuffixArray(text.data(), n, sa.data(), CheckMode::Fast);
            if (!r.ok) {
                suite.fail(name + " GPU", r.reason, r.index);
                return;
            }
        }
        // The BWT in place of a copy of the text, as build --bwt does.
        Text copy(text);
        int32_t primary = -1;
        suite.cases++;
        if (gpuBwt(context, copy.data(), n, copy.data(), primary) != GpuStatus::Ok) {
            suite.fail(name + " GPU BWT", gpuLastError(), -1);
            return;
        }
        int32_t expectedPrimary;
        const std::vector<uint8_t> expected = bwtOf(text.data(), n, sa, expectedPrimary);
        if (primary != expectedPrimary) {
            suite.fail(name + " GPU BWT", "wrong primary index", primary);
            return;
        }
        if (!std::equal(expected.begin(), expected.end(), copy.begin())) {
            suite.fail(name + " GPU BWT", "BWT differs", -1);
            return;
        }
        if (suite.verbose) {
            printf("  ok   %s GPU, suffix array and BWT (n=%d)\n", name.c_str(), n);
        }
    }

    void testRanks(Suite& suite, const std::string& name, const Text& input) {
        const Text text = padded(input);
        const int32_t n = static_cast<int32_t>(input.size());
        std::vector<int32_t> sa(static_cast<size_t>(n));
        suite.cases++;
        if (n > 0 && gpuSuffixArray(text.data(), n, sa.data()) != GpuStatus::Ok) {
            suite.fail(name + " GPU sample", gpuLastError(), -1);
            return;
        }
        if (n > 0 && !checkSuffixArray(text.data(), n, sa.data(), CheckMode::Fast).ok) {
            suite.fail(name + " GPU sample", "suffix array fails the check", -1);
            return;
        }
        std::vector<uint32_t> expected(static_cast<size_t>(gpuSampleSize(n)));
        std::vector<uint32_t> actual(expected.size() + 1);
        gpuSampleRanksFromSuffixArray(sa.data(), n, expected.data());
        GpuStats stats;
        if (gpuRankSample(text.data(), n, actual.data(), &stats) != GpuStatus::Ok) {
            suite.fail(name + " GPU sample", gpuLastError(), -1);
            return;
        }
        const auto m = std::mismatch(expected.begin(), expected.end(), actual.begin());
        if (m.first != expected.end()) {
            suite.fail(name + " GPU sample", "ranks differ from those the suffix array implies", m.first - expected.begin());
            return;
        }
        if (suite.verbose) {
            printf("  ok   %s GPU sample (n=%d, %d character passes, %d doubling rounds)\n", name.c_str(), n,
                   stats.characterPasses, stats.doublingRounds);
        }
    }

    // selftest.cpp:505-541.
    void contextCases(Suite& suite, Rng& rng) {
        constexpr int32_t kCapacity = 20000000;
        GpuContext* context = nullptr;
        suite.cases++;
        if (gpuCreateContext(kCapacity, context) != GpuStatus::Ok) {
            suite.fail("GPU context", gpuLastError(), -1);
            return;
        }
        testGpu(suite, "context, random sigma=4, n=20M", randomText(rng, kCapacity, 4), context);
        testGpu(suite, "context, large random sigma=256", randomText(rng, kLarge, 256), context);
        testGpu(suite, "context, \"banana\"", textOf("banana"), context);
        testGpu(suite, "context, \"\"", {}, context);
        testGpu(suite, "context, \"a\"", textOf("a"), context);
        testGpu(suite, "context, large fibonacci", fibonacciText(kLarge, 'a', 'b'), context);
        testGpu(suite, "context, period 33", periodicText(rng, 5000, 33), context);
        for (int i = 0; i < 5; i++) {
            const int32_t n = 1 + static_cast<int32_t>(rng.below(20000));
            testGpu(suite, "context, random sigma=2 #" + std::to_string(i), randomText(rng, n, 2), context);
        }
        const Text tooLarge = padded(Text(static_cast<size_t>(kCapacity) + 1, 'a'));
        std::vector<int32_t> sa(static_cast<size_t>(kCapacity) + 1);
        suite.cases++;
        if (gpuSuffixArray(context, tooLarge.data(), kCapacity + 1, sa.data()) != GpuStatus::NotEnoughMemory) {
            suite.fail("GPU context, over capacity", "not refused with NotEnoughMemory", -1);
        }
        testGpu(suite, "context, after refusal", randomText(rng, 5000, 26), context);
        gpuDestroyContext(context);
    }

    // selftest.cpp:309-385: the checker must reject planted errors.
    void checkerCases(Suite& suite) {
        Rng rng{7};
        const Text text = padded(randomText(rng, 5000, 4));
        const int32_t n = 5000;
        const std::vector<int32_t> sa = naiveSuffixArray(text.data(), n);
        struct Mutation {
            const char* name;
            int32_t i, j, value;  // swap i and j, or set i to value (j < 0)
        };
        const Mutation mutations[] = {{"adjacent swap", 100, 101, 0}, {"distant swap", 5, 4000, 0},
                                      {"duplicate", 10, -1, -2}, {"out of range", 20, -1, n}, {"negative", 30, -1, -1}};
        for (const Mutation& m : mutations) {
            std::vector<int32_t> bad(sa);
            if (m.j >= 0) {
                std::swap(bad[m.i], bad[m.j]);
            } else {
                bad[m.i] = m.value == -2 ? bad[m.i + 1] : m.value;
            }
            for (const CheckMode mode : {CheckMode::Fast, CheckMode::LowMemory, CheckMode::Naive}) {
                suite.cases++;
                if (checkSuffixArray(text.data(), n, bad.data(), mode).ok) {
                    suite.fail("checker", m.name, -1);
                }
            }
        }
        int32_t primary;
        const std::vector<uint8_t> bwt = bwtOf(text.data(), n, sa, primary);
        std::vector<uint8_t> changed(bwt);
        changed[77] = 'z';
        suite.cases += 3;
        if (checkBwt(text.data(), n, changed.data(), primary).ok) {
            suite.fail("BWT checker", "changed character", -1);
        }
        if (checkBwt(text.data(), n, bwt.data(), primary == n ? primary - 1 : primary + 1).ok) {
            suite.fail("BWT checker", "primary off by one", -1);
        }
        if (checkBwt(text.data(), n, bwt.data(), 0).ok) {
            suite.fail("BWT checker", "primary 0", -1);
        }
    }
}

int runSelfTest(bool verbose) {
    Suite suite{verbose};
    printf("selftest: checker catches planted errors...\n");
    checkerCases(suite);

    GpuDevice device;
    if (gpuQueryDevice(device) != GpuStatus::Ok) {
        printf("selftest: GPU cases skipped (%s)\n", gpuLastError());
    } else {
        printf("selftest: GPU cases on %s (%lld MiB free)...\n", device.name, static_cast<long long>(device.freeBytes >> 20));
        Rng rng{3};
        testGpu(suite, "\"\"", {});
        testGpu(suite, "\"a\"", textOf("a"));
        testGpu(suite, "\"banana\"", textOf("banana"));
        testGpu(suite, "\"mississippi\"", textOf("mississippi"));
        testGpu(suite, "run of a", Text(1000, 'a'));
        testGpu(suite, "fibonacci", fibonacciText(5000, 'a', 'b'));
        testGpu(suite, "fibonacci 0/1", fibonacciText(5000, 0, 1));
        testGpu(suite, "thue-morse", thueMorseText(4096));
        for (const int32_t period : {1, 2, 3, 5, 7, 31, 32, 33, 63, 64, 65, 1000}) {
            testGpu(suite, "period " + std::to_string(period), periodicText(rng, 5000, period));
        }
        for (int32_t n = 0; n <= 80; n++) {
            testGpu(suite, "run of a, n=" + std::to_string(n), Text(static_cast<size_t>(n), 'a'));
            testGpu(suite, "random sigma=2, n=" + std::to_string(n), randomText(rng, n, 2));
            testRanks(suite, "run of a, n=" + std::to_string(n), Text(static_cast<size_t>(n), 'a'));
            testRanks(suite, "random sigma=2, n=" + std::to_string(n), randomText(rng, n, 2));
        }
        for (const uint32_t sigma : {1u, 2u, 4u, 26u, 256u}) {
            for (int i = 0; i < 10; i++) {
                const int32_t n = 1 + static_cast<int32_t>(rng.below(20000));
                testGpu(suite, "random sigma=" + std::to_string(sigma) + " #" + std::to_string(i), randomText(rng, n, sigma));
            }
        }
        testGpu(suite, "large run of a", Text(kLarge, 'a'));
        testGpu(suite, "large period 33", periodicText(rng, kLarge, 33));
        testGpu(suite, "large fibonacci", fibonacciText(kLarge, 'a', 'b'));
        testGpu(suite, "large thue-morse", thueMorseText(kLarge));
        testGpu(suite, "large random sigma=4", randomText(rng, kLarge, 4));
        testGpu(suite, "large random sigma=256", randomText(rng, kLarge, 256));
        contextCases(suite, rng);
        for (const int32_t period : {1, 2, 3, 7, 31, 32, 33, 64, 65, 1000}) {
            testRanks(suite, "period " + std::to_string(period), periodicText(rng, 5000, period));
        }
        testRanks(suite, "\"banana\"", textOf("banana"));
        testRanks(suite, "fibonacci 0/1", fibonacciText(5000, 0, 1));
        testRanks(suite, "thue-morse", thueMorseText(4096));
        testRanks(suite, "large fibonacci", fibonacciText(kLarge, 'a', 'b'));
        testRanks(suite, "large random sigma=4", randomText(rng, kLarge, 4));
    }
    printf("selftest: %d of %d cases passed\n", suite.cases - suite.failures, suite.cases);
    return suite.failures;
}
