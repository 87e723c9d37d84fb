/*
 * selftest.cpp - `sadcs selftest`.
 *
 * Two parts:
 *   1. The host checker (checker.h) must reject planted errors in a correct suffix array and BWT
 *      of a seeded text; the correct suffix array comes from prefix doubling (Manber and Myers).
 *   2. The GPU construction over the inputs of the reference's GPU selftest. The inputs are the
 *      reference's: its splitmix64 generator, seeds and text shapes, drawn in the order its cases
 *      draw them (src/cli/selftest.cpp:30-108, 455-607). Everything else here is this file's own.
 *      A GPU suffix array must pass the checker: the pairwise comparison of neighbours up to
 *      20000 bytes, the fast check above. A sorted permutation of the suffixes is unique, so a
 *      passing array is the suffix array. Each BWT, built in place of a copy of the text as
 *      `build --bwt` does, must equal the one the suffix array determines; sample ranks must equal
 *      gpuSampleRanksFromSuffixArray's.
 */
#include "selftest.h"

#include <algorithm>
#include <bit>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

#include "../checker.h"
#include "gpusa.h"

namespace {
    constexpr int32_t kBig = 1 << 22;            // the large inputs' size
    constexpr int32_t kPairwiseLimit = 20000;    // longest input checked neighbour by neighbour
    constexpr int32_t kContextCapacity = 20000000;
    constexpr size_t kPadding = 64;              // zero bytes after every text

    using Bytes = std::vector<uint8_t>;

    // splitmix64 (Steele, Lea and Flood 2014).
    class SplitMix {
     public:
        explicit SplitMix(uint64_t seed) : m_state(seed) {}

        uint64_t operator()() {
            m_state += 0x9E3779B97F4A7C15ull;
            uint64_t z = m_state;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        }

        uint32_t below(uint32_t bound) {
            return static_cast<uint32_t>((*this)() % bound);
        }

     private:
        uint64_t m_state;
    };

    // The reference selftest's text shapes.
    namespace shape {
        Bytes uniform(SplitMix& g, int32_t n, uint32_t sigma) {
            Bytes t(static_cast<size_t>(n));
            for (uint8_t& c : t) {
                c = static_cast<uint8_t>(sigma == 256 ? g.below(256) : 'a' + g.below(sigma));
            }
            return t;
        }

        Bytes periodic(SplitMix& g, int32_t n, int32_t period) {
            const Bytes unit = uniform(g, period, 4);
            Bytes t(static_cast<size_t>(n));
            for (size_t i = 0; i < t.size(); i++) {
                t[i] = unit[i % unit.size()];
            }
            return t;
        }

        // The Fibonacci word over {a, b}: w1 = a, w2 = ab, w(k+1) = w(k) w(k-1), cut to n.
        Bytes fibonacci(int32_t n, uint8_t a, uint8_t b) {
            Bytes older{a}, newer{a, b};
            while (newer.size() < static_cast<size_t>(n)) {
                Bytes joined = newer;
                joined.insert(joined.end(), older.begin(), older.end());
                older.swap(newer);
                newer.swap(joined);
            }
            newer.resize(static_cast<size_t>(n));
            return newer;
        }

        // Thue-Morse: the parity of the set bits of i.
        Bytes thueMorse(int32_t n) {
            Bytes t(static_cast<size_t>(n));
            for (int32_t i = 0; i < n; i++) {
                t[i] = static_cast<uint8_t>(std::popcount(static_cast<uint32_t>(i)) % 2 ? 'b' : 'a');
            }
            return t;
        }

        Bytes run(int32_t n, char c = 'a') {
            return Bytes(static_cast<size_t>(n), static_cast<uint8_t>(c));
        }

        Bytes word(const char* s) {
            return Bytes(s, s + strlen(s));
        }
    }

    enum class Verify { Outputs, SampleRanks };

    struct Probe {
        std::string label;
        Bytes text;
        Verify verify;
    };

    // The GPU cases, their inputs drawn from seed 3 in the reference's order. `shared` runs in one
    // context of capacity 20,000,000; `afterRefusal` runs in it after an over-capacity call.
    struct GpuPlan {
        std::vector<Probe> alone;
        std::vector<Probe> shared;
        Probe afterRefusal;
        std::vector<Probe> ranks;
    };

    GpuPlan planGpuCases() {
        SplitMix g(3);
        GpuPlan plan;
        auto& a = plan.alone;
        const auto both = Verify::Outputs;
        for (const char* w : {"", "a", "banana", "mississippi"}) {
            a.push_back({std::string("word '") + w + "'", shape::word(w), both});
        }
        a.push_back({"1000 x a", shape::run(1000), both});
        a.push_back({"fibonacci a/b, 5000", shape::fibonacci(5000, 'a', 'b'), both});
        a.push_back({"fibonacci 0/1, 5000", shape::fibonacci(5000, 0, 1), both});
        a.push_back({"thue-morse, 4096", shape::thueMorse(4096), both});
        for (const int32_t p : {1, 2, 3, 5, 7, 31, 32, 33, 63, 64, 65, 1000}) {
            a.push_back({"period " + std::to_string(p) + ", 5000", shape::periodic(g, 5000, p), both});
        }
        for (int32_t n = 0; n <= 80; n++) {
            const std::string len = ", n=" + std::to_string(n);
            a.push_back({"run" + len, shape::run(n), both});
            a.push_back({"binary" + len, shape::uniform(g, n, 2), both});
            plan.ranks.push_back({"run" + len, shape::run(n), Verify::SampleRanks});
            plan.ranks.push_back({"binary" + len, shape::uniform(g, n, 2), Verify::SampleRanks});
        }
        for (const uint32_t sigma : {1u, 2u, 4u, 26u, 256u}) {
            for (int k = 0; k < 10; k++) {
                const int32_t n = 1 + static_cast<int32_t>(g.below(20000));
                a.push_back({"sigma " + std::to_string(sigma) + " sample " + std::to_string(k), shape::uniform(g, n, sigma), both});
            }
        }
        a.push_back({"4M run", shape::run(kBig), both});
        a.push_back({"4M period 33", shape::periodic(g, kBig, 33), both});
        a.push_back({"4M fibonacci", shape::fibonacci(kBig, 'a', 'b'), both});
        a.push_back({"4M thue-morse", shape::thueMorse(kBig), both});
        a.push_back({"4M sigma 4", shape::uniform(g, kBig, 4), both});
        a.push_back({"4M sigma 256", shape::uniform(g, kBig, 256), both});

        auto& s = plan.shared;
        s.push_back({"shared: 20M sigma 4", shape::uniform(g, kContextCapacity, 4), both});
        s.push_back({"shared: 4M sigma 256", shape::uniform(g, kBig, 256), both});
        s.push_back({"shared: word 'banana'", shape::word("banana"), both});
        s.push_back({"shared: empty", Bytes{}, both});
        s.push_back({"shared: word 'a'", shape::word("a"), both});
        s.push_back({"shared: 4M fibonacci", shape::fibonacci(kBig, 'a', 'b'), both});
        s.push_back({"shared: period 33, 5000", shape::periodic(g, 5000, 33), both});
        for (int k = 0; k < 5; k++) {
            const int32_t n = 1 + static_cast<int32_t>(g.below(20000));
            s.push_back({"shared: binary sample " + std::to_string(k), shape::uniform(g, n, 2), both});
        }
        plan.afterRefusal = {"shared, after a refusal: sigma 26, 5000", shape::uniform(g, 5000, 26), both};

        auto& r = plan.ranks;
        const auto ranks = Verify::SampleRanks;
        for (const int32_t p : {1, 2, 3, 7, 31, 32, 33, 64, 65, 1000}) {
            r.push_back({"period " + std::to_string(p) + ", 5000", shape::periodic(g, 5000, p), ranks});
        }
        r.push_back({"word 'banana'", shape::word("banana"), ranks});
        r.push_back({"fibonacci 0/1, 5000", shape::fibonacci(5000, 0, 1), ranks});
        r.push_back({"thue-morse, 4096", shape::thueMorse(4096), ranks});
        r.push_back({"sigma 256, 5000", shape::uniform(g, 5000, 256), ranks});
        r.push_back({"4M run", shape::run(kBig), ranks});
        r.push_back({"4M period 33", shape::periodic(g, kBig, 33), ranks});
        r.push_back({"4M fibonacci", shape::fibonacci(kBig, 'a', 'b'), ranks});
        r.push_back({"4M thue-morse", shape::thueMorse(kBig), ranks});
        r.push_back({"4M sigma 4", shape::uniform(g, kBig, 4), ranks});
        r.push_back({"4M sigma 256", shape::uniform(g, kBig, 256), ranks});
        return plan;
    }

    // Counts checks and prints each miss as it happens (and each pass when verbose).
    class Tally {
     public:
        explicit Tally(bool verbose) : m_verbose(verbose) {}

        void ok(const std::string& label, const std::string& note) {
            m_checks++;
            if (m_verbose) {
                printf("  ok      %s: %s\n", label.c_str(), note.c_str());
            }
        }

        void miss(const std::string& label, const std::string& detail, int64_t at = -1) {
            m_checks++;
            m_misses++;
            if (at >= 0) {
                printf("  MISSED  %s: %s [at %lld]\n", label.c_str(), detail.c_str(), static_cast<long long>(at));
            } else {
                printf("  MISSED  %s: %s\n", label.c_str(), detail.c_str());
            }
        }

        int checks() const { return m_checks; }
        int misses() const { return m_misses; }

     private:
        bool m_verbose;
        int m_checks = 0;
        int m_misses = 0;
    };

    Bytes padded(const Bytes& t) {
        Bytes p(t);
        p.resize(t.size() + kPadding, 0);
        return p;
    }

    template <class T>
    int64_t firstDifference(const T* a, const T* b, size_t count) {
        const auto at = std::mismatch(a, a + count, b);
        return at.first == a + count ? -1 : at.first - a;
    }

    // The BWT a suffix array determines (buildBwt()'s convention): byte 0 is T[n - 1] (the
    // sentinel's row); row r holds T[SA[r] - 1] at r + 1 before the row of suffix 0 and at r after
    // it; the primary index is that row plus one.
    int32_t bwtOfSuffixArray(const uint8_t* t, int32_t n, const int32_t* sa, uint8_t* out) {
        if (n == 0) {
            return 0;
        }
        const int32_t zeroRow = static_cast<int32_t>(std::find(sa, sa + n, 0) - sa);
        out[0] = t[n - 1];
        for (int32_t r = 0; r < n; r++) {
            if (r != zeroRow) {
                out[r < zeroRow ? r + 1 : r] = t[sa[r] - 1];
            }
        }
        return zeroRow + 1;
    }

    CheckResult verifySuffixArray(const Bytes& text, int32_t n, const int32_t* sa) {
        return checkSuffixArray(text.data(), n, sa, n <= kPairwiseLimit ? CheckMode::Naive : CheckMode::Fast);
    }

    void runOutputs(Tally& tally, const Probe& probe, GpuContext* context) {
        const Bytes text = padded(probe.text);
        const int32_t n = static_cast<int32_t>(probe.text.size());
        std::vector<int32_t> sa(static_cast<size_t>(n), -1);
        if (gpuSuffixArray(context, text.data(), n, sa.data()) != GpuStatus::Ok) {
            tally.miss(probe.label, std::string("suffix array build failed: ") + gpuLastError());
            return;
        }
        const CheckResult r = verifySuffixArray(text, n, sa.data());
        if (!r.ok) {
            tally.miss(probe.label, std::string("suffix array: ") + r.reason, r.index);
            return;
        }
        tally.ok(probe.label, "suffix array (n=" + std::to_string(n) + ")");

        Bytes inPlace(text);
        int32_t primary = -1;
        if (gpuBwt(context, inPlace.data(), n, inPlace.data(), primary) != GpuStatus::Ok) {
            tally.miss(probe.label, std::string("BWT build failed: ") + gpuLastError());
            return;
        }
        Bytes expected(static_cast<size_t>(n));
        const int32_t expectedPrimary = bwtOfSuffixArray(text.data(), n, sa.data(), expected.data());
        if (primary != expectedPrimary) {
            tally.miss(probe.label, "BWT primary " + std::to_string(primary) + ", expected " + std::to_string(expectedPrimary));
            return;
        }
        const int64_t at = firstDifference(expected.data(), inPlace.data(), expected.size());
        if (at >= 0) {
            tally.miss(probe.label, "BWT byte differs", at);
            return;
        }
        tally.ok(probe.label, "BWT in place");
    }

    void runSampleRanks(Tally& tally, const Probe& probe) {
        const Bytes text = padded(probe.text);
        const int32_t n = static_cast<int32_t>(probe.text.size());
        std::vector<int32_t> sa(static_cast<size_t>(n));
        if (n > 0 && gpuSuffixArray(text.data(), n, sa.data()) != GpuStatus::Ok) {
            tally.miss(probe.label, std::string("suffix array build failed: ") + gpuLastError());
            return;
        }
        if (n > 0 && !verifySuffixArray(text, n, sa.data()).ok) {
            tally.miss(probe.label, "suffix array for the expected ranks fails the checker");
            return;
        }
        const size_t m = static_cast<size_t>(gpuSampleSize(n));
        std::vector<uint32_t> implied(m), ranked(m + 1);
        gpuSampleRanksFromSuffixArray(sa.data(), n, implied.data());
        GpuStats stats;
        if (gpuRankSample(text.data(), n, ranked.data(), &stats) != GpuStatus::Ok) {
            tally.miss(probe.label, std::string("sample ranking failed: ") + gpuLastError());
            return;
        }
        const int64_t at = firstDifference(implied.data(), ranked.data(), m);
        if (at >= 0) {
            tally.miss(probe.label, "sample rank differs from the suffix array's", at);
            return;
        }
        tally.ok(probe.label, "sample ranks (" + std::to_string(stats.characterPasses) + " character passes, " +
                                  std::to_string(stats.doublingRounds) + " doubling rounds)");
    }

    void runGpuCases(Tally& tally) {
        const GpuPlan plan = planGpuCases();
        for (const Probe& p : plan.alone) {
            runOutputs(tally, p, nullptr);
        }
        GpuContext* context = nullptr;
        if (gpuCreateContext(kContextCapacity, context) != GpuStatus::Ok) {
            tally.miss("shared context", std::string("creation failed: ") + gpuLastError());
        } else {
            for (const Probe& p : plan.shared) {
                runOutputs(tally, p, context);
            }
            const Bytes over = padded(shape::run(kContextCapacity + 1));
            std::vector<int32_t> sa(over.size());
            const GpuStatus s = gpuSuffixArray(context, over.data(), kContextCapacity + 1, sa.data());
            if (s == GpuStatus::NotEnoughMemory) {
                tally.ok("shared context", "an input over its capacity is refused");
            } else {
                tally.miss("shared context", "an input over its capacity was not refused with NotEnoughMemory");
            }
            runOutputs(tally, plan.afterRefusal, context);
            gpuDestroyContext(context);
        }
        for (const Probe& p : plan.ranks) {
            runSampleRanks(tally, p);
        }
    }

    // Prefix doubling (Manber and Myers 1993) with comparison sorts, O(n log^2 n): the correct
    // suffix array of the planted-error cases, independent of the GPU.
    std::vector<int32_t> doublingSuffixArray(const uint8_t* t, int32_t n) {
        std::vector<int32_t> sa(static_cast<size_t>(n)), rank(static_cast<size_t>(n)), next(static_cast<size_t>(n));
        std::iota(sa.begin(), sa.end(), 0);
        for (int32_t i = 0; i < n; i++) {
            rank[i] = t[i];
        }
        for (int64_t h = 1; n > 1; h *= 2) {
            // Past the end ranks -1, so a suffix that ends within h characters sorts first.
            const auto key = [&](int32_t i) { return std::make_pair(rank[i], i + h < n ? rank[i + h] : -1); };
            std::sort(sa.begin(), sa.end(), [&](int32_t x, int32_t y) { return key(x) < key(y); });
            next[sa[0]] = 0;
            for (int32_t i = 1; i < n; i++) {
                next[sa[i]] = next[sa[i - 1]] + (key(sa[i - 1]) < key(sa[i]) ? 1 : 0);
            }
            rank.swap(next);
            if (rank[sa[n - 1]] == n - 1) {
                break;
            }
        }
        return sa;
    }

    void runCheckerCases(Tally& tally) {
        SplitMix g(7);
        const int32_t n = 5000;
        const Bytes text = padded(shape::uniform(g, n, 4));
        const std::vector<int32_t> sa = doublingSuffixArray(text.data(), n);
        for (const CheckMode mode : {CheckMode::Fast, CheckMode::LowMemory, CheckMode::Naive}) {
            const CheckResult r = checkSuffixArray(text.data(), n, sa.data(), mode);
            if (r.ok) {
                tally.ok("checker", "accepts the suffix array");
            } else {
                tally.miss("checker", std::string("rejects the suffix array: ") + r.reason, r.index);
            }
        }
        using Edit = std::function<void(std::vector<int32_t>&)>;
        const std::pair<const char*, Edit> edits[] = {
            {"neighbours 2500 and 2501 swapped", [](auto& a) { std::swap(a[2500], a[2501]); }},
            {"entries 17 and 4444 swapped", [](auto& a) { std::swap(a[17], a[4444]); }},
            {"entry 900 repeated at 901", [](auto& a) { a[901] = a[900]; }},
            {"entry 3000 set to n", [n](auto& a) { a[3000] = n; }},
            {"entry 0 set to -5", [](auto& a) { a[0] = -5; }},
        };
        const char* modeNames[] = {"fast", "low-memory", "pairwise"};
        for (const auto& [what, edit] : edits) {
            std::vector<int32_t> bad(sa);
            edit(bad);
            for (const CheckMode mode : {CheckMode::Fast, CheckMode::LowMemory, CheckMode::Naive}) {
                const std::string label = std::string("checker (") + modeNames[static_cast<int>(mode)] + ")";
                if (checkSuffixArray(text.data(), n, bad.data(), mode).ok) {
                    tally.miss(label, std::string("accepted ") + what);
                } else {
                    tally.ok(label, std::string("rejects ") + what);
                }
            }
        }
        Bytes bwt(static_cast<size_t>(n));
        const int32_t primary = bwtOfSuffixArray(text.data(), n, sa.data(), bwt.data());
        const CheckResult good = checkBwt(text.data(), n, bwt.data(), primary);
        if (good.ok) {
            tally.ok("BWT checker", "accepts the BWT");
        } else {
            tally.miss("BWT checker", std::string("rejects the BWT: ") + good.reason, good.index);
        }
        Bytes flipped(bwt);
        flipped[1234] ^= 0x40;
        const std::pair<const char*, CheckResult> bwtEdits[] = {
            {"a changed byte", checkBwt(text.data(), n, flipped.data(), primary)},
            {"the primary index moved by one", checkBwt(text.data(), n, bwt.data(), primary == n ? primary - 1 : primary + 1)},
            {"primary index 0", checkBwt(text.data(), n, bwt.data(), 0)},
        };
        for (const auto& [what, r] : bwtEdits) {
            if (r.ok) {
                tally.miss("BWT checker", std::string("accepted ") + what);
            } else {
                tally.ok("BWT checker", std::string("rejects ") + what);
            }
        }
    }
}

int runSelfTest(bool verbose) {
    Tally tally(verbose);
    printf("selftest: host checker against planted errors\n");
    runCheckerCases(tally);
    GpuDevice device;
    if (gpuQueryDevice(device) != GpuStatus::Ok) {
        printf("selftest: no GPU, construction cases skipped (%s)\n", gpuLastError());
    } else {
        printf("selftest: construction on %s, %lld MiB free\n", device.name, static_cast<long long>(device.freeBytes >> 20));
        runGpuCases(tally);
    }
    printf("selftest: %d checks, %d missed\n", tally.checks(), tally.misses());
    return tally.misses();
}
