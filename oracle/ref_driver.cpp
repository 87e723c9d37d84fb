// ref_driver.cpp - C entry points over the reference's CPU construction and checker, compiled
// together with the reference sources (src/cpu/dcsort.cpp, src/cpu/doubling.cpp,
// src/common/checker.cpp) where they lie under /root/reference. Test infrastructure only: the
// parity tests and the bench's CPU baseline load oracle/_ref/libsadcs_ref.so through oracle/ref.py.
//
// Entry points follow src/cpu/dcsort.h:37-44 (buildSuffixArray, buildBwt) and
// src/common/checker.h:23-28 (checkSuffixArray, checkBwt). Texts must carry 64 zero bytes of
// padding after n (src/common/simd.h:20).
#include <cstdint>
#include <omp.h>
#include "checker.h"
#include "dcsort.h"

extern "C" {

int ref_build_sa(const uint8_t* text, int32_t n, int32_t* sa, int32_t cover) {
    return buildSuffixArray(text, n, sa, cover) ? 0 : -1;
}

// bwt may be text; sa is n int32 of working space. Returns the primary index or -1.
int32_t ref_build_bwt(const uint8_t* text, int32_t n, uint8_t* bwt, int32_t* sa, int32_t cover) {
    return buildBwt(text, n, bwt, sa, cover);
}

// mode: 0 fast, 1 low memory, 2 naive. Returns -1 if ok, else the failing index (>= 0).
int64_t ref_check_sa(const uint8_t* text, int32_t n, const int32_t* sa, int32_t mode) {
    const CheckResult r = checkSuffixArray(text, n, sa, mode == 0 ? CheckMode::Fast : mode == 1 ? CheckMode::LowMemory : CheckMode::Naive);
    return r.ok ? -1 : (r.index >= 0 ? r.index : 0);
}

int64_t ref_check_bwt(const uint8_t* text, int32_t n, const uint8_t* bwt, int32_t primary) {
    const CheckResult r = checkBwt(text, n, bwt, primary);
    return r.ok ? -1 : (r.index >= 0 ? r.index : 0);
}

void ref_set_threads(int32_t threads) {
    omp_set_num_threads(threads);
}

int32_t ref_max_threads(void) {
    return omp_get_max_threads();
}

}
