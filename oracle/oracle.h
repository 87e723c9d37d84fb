/*
 * oracle.h - CPU restatement of the reference's GPU suffix array / BWT construction.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker of the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and never as the thing
 * measured or shipped. The product (sadcs_amd) neither links nor calls it.
 *
 * Every function cites the reference file:line it restates. How each one is pinned (golden
 * vectors, the reference build in oracle/_ref, brute force, closed forms) is listed in the
 * header comment of oracle.c.
 */
#ifndef SADCS_ORACLE_H
#define SADCS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cover arithmetic (src/common/differencecover.h:27-47, :154-194; src/cuda/devicecommon.cuh:76-85). */
int oracle_cover_size(int v);
int oracle_cover_classes(int v);
int64_t oracle_pack(int v, int64_t p);
int oracle_is_sample(int v, int64_t p);
int oracle_class_of(int v, int64_t p);
int oracle_tie_offset(int v, int64_t a, int64_t b);

/* Conventions (src/cli/selftest.cpp:114-159, src/cuda/gpusa.h:62-73, src/common/checker.cpp). */
void oracle_naive_sa(const uint8_t* text, int32_t n, int32_t* sa);
int32_t oracle_bwt_from_sa(const uint8_t* text, int32_t n, const int32_t* sa, uint8_t* bwt);
void oracle_sample_ranks_from_sa(int v, const int32_t* sa, int32_t n, uint32_t* ranks);
int64_t oracle_check_sa(const uint8_t* text, int32_t n, const int32_t* sa);
int64_t oracle_check_bwt(const uint8_t* text, int32_t n, const uint8_t* bwt, int32_t primary);

/* The GPU algorithm, stage by stage (src/cuda/samplerank.cuh, classsort.cuh, merge.cuh). */
int oracle_rank_sample(int v, const uint8_t* text, int32_t n, uint32_t* ranks, int32_t* passes);
void oracle_class_lists(int v, const uint8_t* text, int32_t n, const uint32_t* ranks, uint32_t* lists, int64_t* starts);
int oracle_less(int v, const uint8_t* text, int32_t n, const uint32_t* ranks, int64_t i, int64_t j);
void oracle_merge(int v, const uint8_t* text, int32_t n, const uint32_t* ranks, const uint32_t* lists,
                  const int64_t* starts, int32_t* sa);
int oracle_suffix_array(int v, const uint8_t* text, int32_t n, int32_t* sa);

/* The generators of the reference's selftest (src/cli/selftest.cpp:30-108). */
uint64_t oracle_splitmix_next(uint64_t* state);
void oracle_random_text(uint64_t* state, int32_t n, uint32_t sigma, uint8_t* out);
void oracle_periodic_text(uint64_t* state, int32_t n, int32_t period, uint8_t* out);
void oracle_fibonacci_text(int32_t n, uint8_t a, uint8_t b, uint8_t* out);
void oracle_thue_morse_text(int32_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif

#endif
