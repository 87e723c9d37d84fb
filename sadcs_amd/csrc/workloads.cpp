/*
 * workloads.cpp - Seeded synthetic texts of the benchmark configurations, as a small C library
 * (libsadcs_work.so) for bench.py, the tests and the command line. The bytes are a pure function of
 * (seed, n): large texts are generated in independent chunks with seeds derived from the chunk
 * index, so the result does not depend on the number of threads.
 *
 *   dna_uniform  i.i.d. uniform over ACGT
 *   dna_like     i.i.d. background (A/T 0.3, C/G 0.2), ~30% of positions covered by copies of earlier
 *                segments (length U[1 KiB, 10 KiB], 1% point substitutions), runs of 'N' (length
 *                U[100, 5000]) covering ~0.1%
 *   markov       order-3 Markov chain over the 96 bytes 0x20..0x7F: every context has its own seeded
 *                permutation of the symbols, and the symbol of rank r is drawn with the Zipf(1.1)
 *                weight of r; then ~5% of positions overwritten by exact copies of earlier segments
 *                (length log-uniform in [64 B, 64 KiB]). A context's permutation is
 *                rank -> outer[(a * inner[rank] + b) mod 96], with two seeded shuffles inner and
 *                outer and (a, b) drawn from a hash of the context (a a unit modulo 96), so no
 *                per-context table is needed.
 *   protein      i.i.d. over the 20 amino-acid letters with seeded frequencies in [1%, 10%]
 *   periodic     a random pattern (period U[2, 64]) over ACGT repeated, 0.1% substitutions
 */
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {
    uint64_t mix(uint64_t& state) {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }

    struct Rng {
        uint64_t state;
        explicit Rng(uint64_t seed) : state(seed) {}
        uint64_t next() {
            return mix(state);
        }
        double uniform() {
            return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0);
        }
        uint64_t below(uint64_t bound) {
            return next() % bound;
        }
    };

    uint64_t derive(uint64_t seed, uint64_t index) {
        uint64_t s = seed * 0x100000001B3ull + index * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
        return mix(s);
    }

    constexpr int64_t kChunk = int64_t(1) << 20;

    // Walker alias table over k outcomes.
    struct Alias {
        std::vector<double> probability;
        std::vector<uint32_t> alias;

        explicit Alias(const std::vector<double>& weights) {
            const size_t k = weights.size();
            probability.resize(k);
            alias.resize(k);
            double total = 0;
            for (double w : weights) {
                total += w;
            }
            std::vector<double> scaled(k);
            std::vector<uint32_t> small, large;
            for (size_t i = 0; i < k; i++) {
                scaled[i] = weights[i] * static_cast<double>(k) / total;
                (scaled[i] < 1.0 ? small : large).push_back(static_cast<uint32_t>(i));
            }
            while (!small.empty() && !large.empty()) {
                const uint32_t s = small.back(), l = large.back();
                small.pop_back();
                probability[s] = scaled[s];
                alias[s] = l;
                scaled[l] -= 1.0 - scaled[s];
                if (scaled[l] < 1.0) {
                    large.pop_back();
                    small.push_back(l);
                }
            }
            for (uint32_t i : large) {
                probability[i] = 1.0;
                alias[i] = i;
            }
            for (uint32_t i : small) {
                probability[i] = 1.0;
                alias[i] = i;
            }
        }

        uint32_t sample(Rng& rng) const {
            const uint64_t r = rng.next();
            const uint32_t i = static_cast<uint32_t>((r >> 32) % probability.size());
            const double u = static_cast<double>(r & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
            return u < probability[i] ? i : alias[i];
        }
    };

    // Copies random earlier segments over the text until `fraction` of it is covered.
    // Lengths: uniform in [lo, hi] (logUniform false) or log-uniform. Substitutions with probability sub,
    // by a symbol drawn from `alphabet` that differs from the copied one.
    void plantRepeats(Rng& rng, uint8_t* t, int64_t n, double fraction, int64_t lo, int64_t hi, bool logUniform, double sub,
                      const char* alphabet) {
        const int64_t target = static_cast<int64_t>(fraction * static_cast<double>(n));
        const int alphabetSize = static_cast<int>(strlen(alphabet));
        int64_t covered = 0;
        while (covered < target) {
            int64_t length;
            if (logUniform) {
                const double l = std::log(static_cast<double>(lo)) + rng.uniform() * (std::log(static_cast<double>(hi)) - std::log(static_cast<double>(lo)));
                length = static_cast<int64_t>(std::exp(l));
            } else {
                length = lo + static_cast<int64_t>(rng.below(static_cast<uint64_t>(hi - lo + 1)));
            }
            if (length >= n / 2) {
                length = std::max<int64_t>(1, n / 4);
            }
            const int64_t destination = 1 + static_cast<int64_t>(rng.below(static_cast<uint64_t>(n - length)));
            const int64_t source = static_cast<int64_t>(rng.below(static_cast<uint64_t>(destination)));
            for (int64_t i = 0; i < length; i++) {
                uint8_t c = t[source + i];
                if (sub > 0 && rng.uniform() < sub) {
                    uint8_t d;
                    do {
                        d = static_cast<uint8_t>(alphabet[rng.below(static_cast<uint64_t>(alphabetSize))]);
                    } while (d == c && alphabetSize > 1);
                    c = d;
                }
                t[destination + i] = c;
            }
            covered += length;
        }
    }

    template <class Fill>
    void chunked(uint64_t seed, uint8_t* out, int64_t n, Fill fill) {
        const int64_t chunks = (n + kChunk - 1) / kChunk;
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t c = 0; c < chunks; c++) {
            Rng rng(derive(seed, static_cast<uint64_t>(c)));
            const int64_t begin = c * kChunk;
            fill(rng, out + begin, std::min(kChunk, n - begin));
        }
    }
}

extern "C" {

void sadcs_gen_dna_uniform(uint64_t seed, uint8_t* out, int64_t n) {
    static const char kACGT[] = "ACGT";
    chunked(seed, out, n, [](Rng& rng, uint8_t* t, int64_t m) {
        for (int64_t i = 0; i < m; i++) {
            t[i] = static_cast<uint8_t>(kACGT[rng.next() >> 62]);
        }
    });
}

// stats[0] = fraction covered by repeats, stats[1] = fraction covered by N runs (may be null).
void sadcs_gen_dna_like(uint64_t seed, uint8_t* out, int64_t n, double* stats) {
    chunked(seed, out, n, [](Rng& rng, uint8_t* t, int64_t m) {
        for (int64_t i = 0; i < m; i++) {
            const double u = rng.uniform();
            t[i] = u < 0.3 ? 'A' : u < 0.5 ? 'C' : u < 0.7 ? 'G' : 'T';
        }
    });
    if (n < 4) {
        return;
    }
    Rng rng(derive(seed, ~0ull));
    const int64_t lo = std::min<int64_t>(1024, std::max<int64_t>(1, n / 8)), hi = std::min<int64_t>(10240, std::max<int64_t>(1, n / 4));
    plantRepeats(rng, out, n, 0.30, lo, std::max(lo, hi), false, 0.01, "ACGT");
    const int64_t target = n / 1000;
    int64_t covered = 0;
    while (covered < target) {
        const int64_t length = std::min<int64_t>(100 + static_cast<int64_t>(rng.below(4901)), n / 2);
        const int64_t at = static_cast<int64_t>(rng.below(static_cast<uint64_t>(n - length)));
        memset(out + at, 'N', static_cast<size_t>(length));
        covered += length;
    }
    if (stats) {
        int64_t ns = 0;
        for (int64_t i = 0; i < n; i++) {
            ns += out[i] == 'N';
        }
        stats[0] = 0.30;
        stats[1] = static_cast<double>(ns) / static_cast<double>(n);
    }
}

void sadcs_gen_markov(uint64_t seed, uint8_t* out, int64_t n) {
    constexpr int kSymbols = 96;
    constexpr int kContexts = kSymbols * kSymbols * kSymbols;
    std::vector<double> weights(kSymbols);
    for (int r = 0; r < kSymbols; r++) {
        weights[r] = 1.0 / std::pow(static_cast<double>(r + 1), 1.1);
    }
    const Alias zipf(weights);
    uint8_t inner[kSymbols], outer[kSymbols];
    uint8_t units[kSymbols];
    int unitCount = 0;
    {
        Rng rng(derive(seed, ~6ull));
        for (int i = 0; i < kSymbols; i++) {
            inner[i] = outer[i] = static_cast<uint8_t>(i);
            if (std::gcd(i, kSymbols) == 1) {
                units[unitCount++] = static_cast<uint8_t>(i);
            }
        }
        for (int i = kSymbols - 1; i > 0; i--) {
            std::swap(inner[i], inner[rng.below(static_cast<uint64_t>(i + 1))]);
            std::swap(outer[i], outer[rng.below(static_cast<uint64_t>(i + 1))]);
        }
    }
    static_assert(kContexts == 96 * 96 * 96, "order-3 contexts");
    chunked(seed, out, n, [&](Rng& rng, uint8_t* t, int64_t m) {
        uint32_t a = static_cast<uint32_t>(rng.below(kSymbols)), b = static_cast<uint32_t>(rng.below(kSymbols)),
                 c = static_cast<uint32_t>(rng.below(kSymbols));
        for (int64_t i = 0; i < m; i++) {
            const uint32_t context = (a * kSymbols + b) * kSymbols + c;
            uint64_t h = seed ^ (static_cast<uint64_t>(context) * 0xD6E8FEB86659FD93ull);
            h = mix(h);
            const uint32_t unit = units[h % static_cast<uint64_t>(unitCount)];
            const uint32_t shift = static_cast<uint32_t>((h >> 32) % kSymbols);
            const uint32_t symbol = outer[(unit * inner[zipf.sample(rng)] + shift) % kSymbols];
            t[i] = static_cast<uint8_t>(0x20 + symbol);
            a = b;
            b = c;
            c = symbol;
        }
    });
    if (n >= 256) {
        Rng rng(derive(seed, ~1ull));
        plantRepeats(rng, out, n, 0.05, 64, std::min<int64_t>(65536, n / 4), true, 0.0, "");
    }
}

void sadcs_gen_protein(uint64_t seed, uint8_t* out, int64_t n) {
    static const char kAmino[] = "ACDEFGHIKLMNPQRSTVWY";
    Rng setup(derive(seed, ~2ull));
    std::vector<double> weights(20);
    for (double& w : weights) {
        w = 0.01 + 0.09 * setup.uniform();
    }
    const Alias table(weights);
    chunked(seed, out, n, [&](Rng& rng, uint8_t* t, int64_t m) {
        for (int64_t i = 0; i < m; i++) {
            t[i] = static_cast<uint8_t>(kAmino[table.sample(rng)]);
        }
    });
}

void sadcs_gen_periodic(uint64_t seed, uint8_t* out, int64_t n) {
    static const char kACGT[] = "ACGT";
    Rng rng(derive(seed, ~3ull));
    const int period = 2 + static_cast<int>(rng.below(63));
    uint8_t pattern[64];
    for (int i = 0; i < period; i++) {
        pattern[i] = static_cast<uint8_t>(kACGT[rng.below(4)]);
    }
    for (int64_t i = 0; i < n; i++) {
        out[i] = pattern[i % period];
    }
    const int64_t substitutions = n / 1000;
    for (int64_t s = 0; s < substitutions; s++) {
        const int64_t at = static_cast<int64_t>(rng.below(static_cast<uint64_t>(n)));
        uint8_t d;
        do {
            d = static_cast<uint8_t>(kACGT[rng.below(4)]);
        } while (d == out[at]);
        out[at] = d;
    }
}

// The batch configuration's texts, text i (of lengths[i] bytes, seed base * 1000003 + i) at
// out + offsets[i]: DNA-like for even i, Markov for odd i. Texts are generated in parallel, each on
// one thread (the nested chunk loops run serially), and every text is the same bytes as the
// single-text generator gives for its seed and length.
void sadcs_gen_batch(uint64_t base, int32_t count, const int64_t* lengths, const int64_t* offsets, uint8_t* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < count; i++) {
        const uint64_t seed = base * 1000003ull + static_cast<uint64_t>(i);
        if (i % 2 == 0) {
            sadcs_gen_dna_like(seed, out + offsets[i], lengths[i], nullptr);
        } else {
            sadcs_gen_markov(seed, out + offsets[i], lengths[i]);
        }
    }
}

// Lengths floor(exp(U[ln lo, ln hi])).
void sadcs_gen_log_uniform_lengths(uint64_t seed, int32_t count, int64_t lo, int64_t hi, int64_t* lengths) {
    Rng rng(derive(seed, ~4ull));
    for (int32_t i = 0; i < count; i++) {
        const double l = std::log(static_cast<double>(lo)) + rng.uniform() * (std::log(static_cast<double>(hi)) - std::log(static_cast<double>(lo)));
        lengths[i] = std::max<int64_t>(lo, std::min<int64_t>(hi, static_cast<int64_t>(std::exp(l))));
    }
}

// Lengths lognormal(ln median, sigma), capped; kinds: 0 ACGT (45%), 1 protein (45%), 2 periodic (10%).
void sadcs_gen_serving_requests(uint64_t seed, int32_t count, double median, double sigma, int64_t cap, int64_t* lengths, int32_t* kinds) {
    Rng rng(derive(seed, ~5ull));
    for (int32_t i = 0; i < count; i++) {
        const double u1 = std::max(rng.uniform(), 1e-300), u2 = rng.uniform();
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        lengths[i] = std::max<int64_t>(1, std::min<int64_t>(cap, static_cast<int64_t>(median * std::exp(sigma * z))));
        const double k = rng.uniform();
        kinds[i] = k < 0.45 ? 0 : k < 0.90 ? 1 : 2;
    }
}

}
