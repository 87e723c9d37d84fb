/*
 * oracle.c - CPU restatement of the reference's GPU construction. TEST INFRASTRUCTURE ONLY
 * (see oracle.h): loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline only.
 *
 * How each function is pinned (tests/test_oracle.py):
 *  - cover arithmetic: the definition of a difference cover checked exhaustively (every distance
 *    covered, tie offsets minimal by brute force), and the reference's own cover sets;
 *  - oracle_naive_sa: the reference's golden vectors (banana: 5 3 1 0 4 2) and brute force over
 *    all binary strings up to length 12; against oracle/_ref (the reference CPU build) on the
 *    selftest inputs;
 *  - oracle_bwt_from_sa / oracle_check_bwt: golden vectors (banana: "annbaa", primary 4; "a":
 *    primary 1), the reference build's buildBwt, and inversion round trips;
 *  - oracle_check_sa: planted mutations (swaps, duplicates, out of range) as the reference's
 *    selftest plants them (src/cli/selftest.cpp:309-385), and agreement with ref_check_sa;
 *  - oracle_rank_sample: the closed form oracle_sample_ranks_from_sa(SA) for the reference's SA;
 *  - oracle_class_lists: the closed form "SA filtered to class d" for the reference's SA;
 *  - oracle_merge / oracle_suffix_array: equality with the reference build's SA.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---- covers: src/common/differencecover.h:27-47 ---------------------------------------------- */

static const int kCover32[] = {0, 1, 3, 8, 12, 19, 25};
static const int kCover64[] = {0, 1, 2, 5, 14, 16, 34, 42, 59};
static const int kCover128[] = {0, 6, 11, 26, 30, 34, 42, 43, 44, 75, 82, 104, 107};
static const int kCover256[] = {0, 24, 31, 34, 36, 54, 58, 123, 139, 147, 154, 165, 168, 186, 203, 219, 228, 237, 243, 244};

static const int* cover_elements(int v, int* size) {
    switch (v) {
        case 32: *size = 7; return kCover32;
        case 64: *size = 9; return kCover64;
        case 128: *size = 13; return kCover128;
        case 256: *size = 20; return kCover256;
        default: *size = 0; return NULL;
    }
}

int oracle_cover_size(int v) {
    int size;
    cover_elements(v, &size);
    return size;
}

int oracle_is_sample(int v, int64_t p) {
    int size;
    const int* e = cover_elements(v, &size);
    const int r = (int)(p & (v - 1));
    for (int i = 0; i < size; i++) {
        if (e[i] == r) {
            return 1;
        }
    }
    return 0;
}

/* differencecover.h:161-164: samples below p. */
int64_t oracle_pack(int v, int64_t p) {
    int size;
    const int* e = cover_elements(v, &size);
    const int r = (int)(p & (v - 1));
    int below = 0;
    for (int i = 0; i < size; i++) {
        below += e[i] < r;
    }
    return (p / v) * size + below;
}

/* differencecover.h:171-173 (nextSampleDistance): the class of p. */
int oracle_class_of(int v, int64_t p) {
    int d = 0;
    while (!oracle_is_sample(v, p + d)) {
        d++;
    }
    return d;
}

int oracle_cover_classes(int v) {
    int largest = 0;
    for (int r = 0; r < v; r++) {
        const int d = oracle_class_of(v, r);
        largest = d > largest ? d : largest;
    }
    return largest + 1;
}

/* src/cuda/devicecommon.cuh:76-85 (cTieOffset): the smallest k with a + k and b + k sampled. */
int oracle_tie_offset(int v, int64_t a, int64_t b) {
    int k = 0;
    while (!oracle_is_sample(v, a + k) || !oracle_is_sample(v, b + k)) {
        k++;
    }
    return k;
}

/* ---- conventions ---------------------------------------------------------------------------- */

static const uint8_t* g_text;
static int32_t g_n;

static int compare_suffixes(const void* x, const void* y) {
    const int32_t p = *(const int32_t*)x, q = *(const int32_t*)y;
    const int32_t lp = g_n - p, lq = g_n - q;
    const int32_t common = lp < lq ? lp : lq;
    const int c = memcmp(g_text + p, g_text + q, (size_t)common);
    if (c != 0) {
        return c;
    }
    return p > q ? -1 : (p < q ? 1 : 0); /* the shorter suffix first */
}

/* src/cli/selftest.cpp:114-125 (naiveSuffixArray). */
void oracle_naive_sa(const uint8_t* text, int32_t n, int32_t* sa) {
    for (int32_t i = 0; i < n; i++) {
        sa[i] = i;
    }
    g_text = text;
    g_n = n;
    qsort(sa, (size_t)n, sizeof(int32_t), compare_suffixes);
}

/* src/cli/selftest.cpp:141-159 (bwtFromSuffixArray): byte 0 is T[n-1], then one byte per row
 * except suffix 0's; returns that row plus one (the primary index). */
int32_t oracle_bwt_from_sa(const uint8_t* text, int32_t n, const int32_t* sa, uint8_t* bwt) {
    if (n == 0) {
        return 0;
    }
    int32_t out = 0, primary = 0;
    bwt[out++] = text[n - 1];
    for (int32_t row = 0; row < n; row++) {
        if (sa[row] == 0) {
            primary = row + 1;
        } else {
            bwt[out++] = text[sa[row] - 1];
        }
    }
    return primary;
}

/* src/cuda/gpusa.h:62-73 (gpuSampleRanksFromSuffixArray), generic in the cover. */
void oracle_sample_ranks_from_sa(int v, const int32_t* sa, int32_t n, uint32_t* ranks) {
    uint32_t rank = 0;
    for (int32_t i = 0; i < n; i++) {
        if (oracle_is_sample(v, sa[i])) {
            ranks[oracle_pack(v, sa[i])] = rank++;
        }
    }
}

/* src/common/checker.cpp:53-109 (checkFast): permutation via the inverse, then each neighbour pair
 * ordered by first character and the inverse of the next position (-1 for the empty suffix).
 * Returns -1 if the array is the suffix array, else the first failing index. */
int64_t oracle_check_sa(const uint8_t* text, int32_t n, const int32_t* sa) {
    if (n == 0) {
        return -1;
    }
    int32_t* isa = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int64_t failure = -1;
    for (int32_t i = 0; i < n && failure < 0; i++) {
        if (sa[i] < 0 || sa[i] >= n) {
            failure = i;
        } else {
            isa[sa[i]] = i;
        }
    }
    for (int32_t i = 0; i < n && failure < 0; i++) {
        if (isa[sa[i]] != i) {
            failure = i;
        }
    }
    for (int32_t i = 0; i + 1 < n && failure < 0; i++) {
        const int32_t a = sa[i], b = sa[i + 1];
        int ordered = text[a] < text[b];
        if (text[a] == text[b]) {
            const int32_t ra = a + 1 < n ? isa[a + 1] : -1;
            const int32_t rb = b + 1 < n ? isa[b + 1] : -1;
            ordered = ra < rb;
        }
        if (!ordered) {
            failure = i;
        }
    }
    free(isa);
    return failure;
}

/* src/common/checker.cpp:196-245 (checkBwt): inverts the BWT through the LF mapping, with the
 * sentinel's row 0 and the primary index as suffix 0's row. Returns -1 if it spells the text. */
int64_t oracle_check_bwt(const uint8_t* text, int32_t n, const uint8_t* bwt, int32_t primary) {
    if (n == 0) {
        return primary == 0 ? -1 : 0;
    }
    if (primary < 1 || primary > n) {
        return 0;
    }
    int32_t* lf = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t next[256] = {0};
    for (int32_t u = 0; u < n; u++) {
        next[bwt[u]]++;
    }
    for (int c = 0, sum = 1; c < 256; c++) {
        const int32_t count = next[c];
        next[c] = sum;
        sum += count;
    }
    for (int32_t u = 0; u < n; u++) {
        lf[u] = next[bwt[u]]++;
    }
    int32_t row = 0;
    int64_t failure = -1;
    for (int32_t i = n - 1; i >= 0; i--) {
        if (row == primary) {
            failure = i;
            break;
        }
        const int32_t u = row < primary ? row : row - 1;
        if (bwt[u] != text[i]) {
            failure = i;
            break;
        }
        row = lf[u];
    }
    if (failure < 0 && row != primary) {
        failure = 0;
    }
    free(lf);
    return failure;
}

/* ---- (a) sample ranking: src/cuda/samplerank.cuh -------------------------------------------- */

/* load8 (devicecommon.cuh:146-150): the 8 bytes from p big-endian, zero past n. */
static uint64_t load8(const uint8_t* text, int64_t n, int64_t p) {
    uint64_t w = 0;
    for (int i = 0; i < 8; i++) {
        w = (w << 8) | (p + i < n ? text[p + i] : 0);
    }
    return w;
}

/* suffixKey (samplerank.cuh:50-63): 7 characters in the top 56 bits, min(remaining, 7) in the low
 * byte; 0 at or past the end. */
static uint64_t suffix_key(const uint8_t* text, int64_t n, int64_t p) {
    if (p >= n) {
        return 0;
    }
    const int64_t remaining = n - p;
    const uint64_t characters = load8(text, n, p) & ~(uint64_t)0xFF;
    if (remaining >= 7) {
        return characters | 7;
    }
    return (characters & (~(uint64_t)0 << (64 - 8 * remaining))) | (uint64_t)remaining;
}

/* wideHeadKey / wideTailKey (samplerank.cuh:66-86): characters 0-7 (zero past the end), then
 * characters 8-14 with min(remaining, 15) in 4 bits. */
static uint64_t wide_head_key(const uint8_t* text, int64_t n, int64_t p) {
    return load8(text, n, p);
}

static uint64_t wide_tail_key(const uint8_t* text, int64_t n, int64_t p) {
    const int64_t remaining = n - p;
    uint64_t characters = 0;
    if (remaining > 8) {
        characters = load8(text, n, p + 8) >> 8;
    }
    return (characters << 4) | (uint64_t)(remaining < 15 ? remaining : 15);
}

typedef struct {
    uint32_t group;
    uint64_t key1;
    uint64_t key2;
    int64_t position;
} PassEntry;

static int compare_entries(const void* x, const void* y) {
    const PassEntry* a = (const PassEntry*)x;
    const PassEntry* b = (const PassEntry*)y;
    if (a->group != b->group) {
        return a->group < b->group ? -1 : 1;
    }
    if (a->key1 != b->key1) {
        return a->key1 < b->key1 ? -1 : 1;
    }
    if (a->key2 != b->key2) {
        return a->key2 < b->key2 ? -1 : 1;
    }
    return a->position < b->position ? -1 : (a->position > b->position ? 1 : 0);
}

/* SampleRanking::run (samplerank.cuh:314-355) and runPass (:406-491) without the device
 * mechanics: every pass sorts the unresolved list by (group id, subkey), a new group's id is its
 * old id plus its offset in the old group, and singletons leave the list. Pass 0 is the wide
 * 15-character pass, then 7-character passes at depths 15, 22, ... while the depth is below V
 * (the reference's three passes reach 36 >= 32), then doubling rounds h = V, 2V, ... reading
 * 1 + the id at p + h (0 past the end). Returns 0, or -1 if h reaches n with groups left.
 * passes[0] = character passes, passes[1] = doubling rounds. */
int oracle_rank_sample(int v, const uint8_t* text, int32_t n, uint32_t* ranks, int32_t* passes) {
    const int64_t m = oracle_pack(v, n);
    PassEntry* list = (PassEntry*)malloc(sizeof(PassEntry) * (size_t)(m > 0 ? m : 1));
    int64_t u = 0;
    for (int64_t p = 0; p < n; p++) {
        if (oracle_is_sample(v, p)) {
            ranks[oracle_pack(v, p)] = 0;
            list[u++].position = p;
        }
    }
    int64_t depth = 0;
    int character = 0, doubling = 0;
    int status = 0;
    while (u > 0) {
        int isDoubling = depth >= v;
        int64_t h = 0;
        if (isDoubling) {
            h = (int64_t)v << doubling;
            if (h >= n) {
                status = -1;
                break;
            }
        }
        for (int64_t i = 0; i < u; i++) {
            const int64_t p = list[i].position;
            list[i].group = ranks[oracle_pack(v, p)];
            if (isDoubling) {
                list[i].key1 = p + h < n ? (uint64_t)ranks[oracle_pack(v, p + h)] + 1 : 0;
                list[i].key2 = 0;
            } else if (character == 0) {
                list[i].key1 = wide_head_key(text, n, p);
                list[i].key2 = wide_tail_key(text, n, p);
            } else {
                list[i].key1 = suffix_key(text, n, p + depth);
                list[i].key2 = 0;
            }
        }
        qsort(list, (size_t)u, sizeof(PassEntry), compare_entries);
        /* All subkeys were read above, before any id of this pass is stored. */
        int64_t kept = 0, groupStart = 0, head = 0;
        uint32_t previousGroup = 0;
        uint64_t previousKey1 = 0, previousKey2 = 0;
        for (int64_t i = 0; i < u; i++) {
            const uint32_t group = list[i].group;
            if (i == 0 || group != previousGroup) {
                groupStart = i;
            }
            if (i == 0 || group != previousGroup || list[i].key1 != previousKey1 || list[i].key2 != previousKey2) {
                head = i;
            }
            previousGroup = group;
            previousKey1 = list[i].key1;
            previousKey2 = list[i].key2;
            list[i].group = group + (uint32_t)(head - groupStart); /* the new id, kept for the tie test below */
        }
        for (int64_t i = 0; i < u; i++) {
            ranks[oracle_pack(v, list[i].position)] = list[i].group;
        }
        for (int64_t i = 0; i < u; i++) {
            const int tied = (i > 0 && list[i - 1].group == list[i].group) || (i + 1 < u && list[i + 1].group == list[i].group);
            if (tied) {
                list[kept++] = list[i];
            }
        }
        u = kept;
        if (isDoubling) {
            depth += h;
            doubling++;
        } else {
            depth += character == 0 ? 15 : 7;
            character++;
        }
    }
    free(list);
    if (passes) {
        passes[0] = character;
        passes[1] = doubling;
    }
    return status;
}

/* ---- (b) class lists: src/cuda/classsort.cuh -------------------------------------------------- */

typedef struct {
    uint8_t character;
    int64_t order;
    uint32_t position;
} ClassEntry;

static int compare_class_entries(const void* x, const void* y) {
    const ClassEntry* a = (const ClassEntry*)x;
    const ClassEntry* b = (const ClassEntry*)y;
    if (a->character != b->character) {
        return a->character < b->character ? -1 : 1;
    }
    return a->order < b->order ? -1 : (a->order > b->order ? 1 : 0);
}

/* ClassSort::sortAll (classsort.cuh:184-244): class 0 is the sample scattered by rank
 * (scatterSample :51-57); class d walks class d - 1 with n prepended (previousElement :77-80),
 * keeps w whose predecessor w - 1 is in class d (predecessorInClass :82-85), and sorts those
 * predecessors stably by T[w - 1]. lists holds n entries, class after class; starts[d] is where
 * class d begins, starts[classes] = n. */
void oracle_class_lists(int v, const uint8_t* text, int32_t n, const uint32_t* ranks, uint32_t* lists, int64_t* starts) {
    const int classes = oracle_cover_classes(v);
    starts[0] = 0;
    for (int d = 0; d < classes; d++) {
        int64_t size = 0;
        for (int64_t p = 0; p < n; p++) {
            size += oracle_class_of(v, p) == d;
        }
        starts[d + 1] = starts[d] + size;
    }
    for (int64_t p = 0; p < n; p++) {
        if (oracle_is_sample(v, p)) {
            lists[ranks[oracle_pack(v, p)]] = (uint32_t)p;
        }
    }
    ClassEntry* entries = (ClassEntry*)malloc(sizeof(ClassEntry) * (size_t)(n + 1));
    for (int d = 1; d < classes; d++) {
        const int64_t count = starts[d] - starts[d - 1];
        int64_t kept = 0;
        for (int64_t k = 0; k <= count; k++) {
            const int64_t w = k == 0 ? n : lists[starts[d - 1] + k - 1];
            if (w > 0 && oracle_class_of(v, w - 1) == d) {
                entries[kept].character = text[w - 1];
                entries[kept].order = k;
                entries[kept].position = (uint32_t)(w - 1);
                kept++;
            }
        }
        qsort(entries, (size_t)kept, sizeof(ClassEntry), compare_class_entries);
        for (int64_t i = 0; i < kept; i++) {
            lists[starts[d] + i] = entries[i].position;
        }
    }
    free(entries);
}

/* ---- (c) merge: src/cuda/merge.cuh ------------------------------------------------------------ */

/* lessGlobal (merge.cuh:53-71): k = the tie offset; the first min(k, remaining) characters; if one
 * suffix ends within them, the shorter (larger position) is smaller; else the ranks at i + k, j + k. */
int oracle_less(int v, const uint8_t* text, int32_t n, const uint32_t* ranks, int64_t i, int64_t j) {
    const int64_t k = oracle_tie_offset(v, i, j);
    int64_t common = k;
    if (n - i < common) {
        common = n - i;
    }
    if (n - j < common) {
        common = n - j;
    }
    for (int64_t o = 0; o < common; o++) {
        if (text[i + o] != text[j + o]) {
            return text[i + o] < text[j + o];
        }
    }
    if (common < k || i + k == n || j + k == n) {
        return i > j;
    }
    return ranks[oracle_pack(v, i + k)] < ranks[oracle_pack(v, j + k)];
}

/* Merger (merge.cuh:622-830) computes, for every element, the number of smaller elements in every
 * other list; restated as a sequential multiway merge with the same comparator. */
void oracle_merge(int v, const uint8_t* text, int32_t n, const uint32_t* ranks, const uint32_t* lists,
                  const int64_t* starts, int32_t* sa) {
    const int classes = oracle_cover_classes(v);
    int64_t* head = (int64_t*)malloc(sizeof(int64_t) * (size_t)classes);
    for (int d = 0; d < classes; d++) {
        head[d] = starts[d];
    }
    for (int32_t r = 0; r < n; r++) {
        int best = -1;
        for (int d = 0; d < classes; d++) {
            if (head[d] < starts[d + 1] &&
                (best < 0 || oracle_less(v, text, n, ranks, lists[head[d]], lists[head[best]]))) {
                best = d;
            }
        }
        sa[r] = (int32_t)lists[head[best]++];
    }
    free(head);
}

/* The whole construction (gpusa.cu:406-466 buildOnDevice without the device). */
int oracle_suffix_array(int v, const uint8_t* text, int32_t n, int32_t* sa) {
    if (n <= 0) {
        return 0;
    }
    const int64_t m = oracle_pack(v, n);
    uint32_t* ranks = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(m + 1));
    uint32_t* lists = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    int64_t starts[64];
    int status = oracle_rank_sample(v, text, n, ranks, NULL);
    if (status == 0) {
        oracle_class_lists(v, text, n, ranks, lists, starts);
        oracle_merge(v, text, n, ranks, lists, starts, sa);
    }
    free(lists);
    free(ranks);
    return status;
}

/* ---- generators: src/cli/selftest.cpp:30-108 --------------------------------------------------- */

/* Rng::next (selftest.cpp:34-39), splitmix64. */
uint64_t oracle_splitmix_next(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* randomText (selftest.cpp:63-71). */
void oracle_random_text(uint64_t* state, int32_t n, uint32_t sigma, uint8_t* out) {
    for (int32_t i = 0; i < n; i++) {
        const uint32_t r = (uint32_t)(oracle_splitmix_next(state) % (sigma == 256 ? 256 : sigma));
        out[i] = (uint8_t)(sigma == 256 ? r : 'a' + r);
    }
}

/* periodicText (selftest.cpp:73-82). */
void oracle_periodic_text(uint64_t* state, int32_t n, int32_t period, uint8_t* out) {
    uint8_t* pattern = (uint8_t*)malloc((size_t)period);
    oracle_random_text(state, period, 4, pattern);
    for (int32_t i = 0; i < n; i++) {
        out[i] = pattern[i % period];
    }
    free(pattern);
}

/* fibonacciText (selftest.cpp:84-98): the prefix of the Fibonacci word over {a, b}. */
void oracle_fibonacci_text(int32_t n, uint8_t a, uint8_t b, uint8_t* out) {
    /* The word f_k = f_{k-1} f_{k-2} with f_1 = a, f_2 = a b: build it in place, longest prefix first. */
    if (n <= 0) {
        return;
    }
    int64_t shorter = 1, longer = 2;
    uint8_t* buffer = (uint8_t*)malloc((size_t)n + 2 * (size_t)n + 4);
    buffer[0] = a;
    buffer[1] = b;
    /* buffer holds the longer word; the shorter one is always its prefix. */
    while (longer < n) {
        const int64_t next = longer + shorter;
        memcpy(buffer + longer, buffer, (size_t)shorter);
        shorter = longer;
        longer = next;
    }
    memcpy(out, buffer, (size_t)n);
    free(buffer);
}

/* thueMorseText (selftest.cpp:100-108). */
void oracle_thue_morse_text(int32_t n, uint8_t* out) {
    for (int32_t i = 0; i < n; i++) {
        out[i] = (uint8_t)('a' + (__builtin_popcount((unsigned)i) & 1));
    }
}
