// Device layout of the chained-BLA (bilinear approximation) table and the
// level search shared by the perturbation engines.
//
// Entry (k, d) advances the perturbation delta from iteration k by 2^d
// iterations: dz' = A dz + B dc, valid while |dz| < r (reference
// math_core/high_precision.py:1100-1238).  The reference stores
// float64[levels][ref_len][5]; this table stores, for each k, the levels of
// one quantity contiguously, so the level search of one lane reads one
// 128-byte row instead of `levels` scattered words:
//   r2   double [ref_len][levels]      r^2 (0 = invalid)
//   r1   double [ref_len][levels]      sqrt(r^2), the value the reference
//                                      computes at lookup when |dz| is tiny
//   coef double4 [ref_len][levels]     (A_re, A_im, B_re, B_im)
//   lr   float  [ref_len][levels]      log2 r for the FP32-mantissa engine
//                                      (-inf = invalid)
//   c32  Coef32 [ref_len][levels]      A and B as float32 mantissa pairs with
//                                      an int32 exponent each
// The radii of one k never grow with d (r_d = min(r_{d-1}, ...), and an
// invalid entry makes every higher level invalid), so "the highest valid
// level" is found by binary search; it equals the reference's top-down scan.
#pragma once

#include "ap_device.hpp"

namespace ap {

struct Coef32 {
    float ar, ai, br, bi;
    int ae, be;
};

struct BlaView {
    const double* r2;
    const double* r1;
    const double4* coef;
    const float* lr;
    const Coef32* c32;
    int levels;   // 0 = no table
    int ref_len;
};

__host__ __device__ inline size_t bla_entries(int ref_len, int levels) {
    return (size_t)(ref_len > 0 ? ref_len : 0) * (size_t)(levels > 0 ? levels : 0);
}

__host__ __device__ inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the regions inside one table allocation.
struct BlaLayout {
    size_t r2, r1, coef, lr, c32, total;
};

__host__ __device__ inline BlaLayout bla_layout(int ref_len, int levels) {
    size_t n = bla_entries(ref_len, levels);
    BlaLayout L;
    L.r2 = 0;
    L.r1 = L.r2 + align256(n * sizeof(double));
    L.coef = L.r1 + align256(n * sizeof(double));
    L.lr = L.coef + align256(n * sizeof(double4));
    L.c32 = L.lr + align256(n * sizeof(float));
    L.total = L.c32 + align256(n * sizeof(Coef32));
    return L;
}

inline BlaView bla_view(const void* table, int ref_len, int levels) {
    BlaView b;
    if (!table || levels < 2 || ref_len < 4) {
        b.r2 = b.r1 = nullptr;
        b.coef = nullptr;
        b.lr = nullptr;
        b.c32 = nullptr;
        b.levels = 0;
        b.ref_len = ref_len;
        return b;
    }
    BlaLayout L = bla_layout(ref_len, levels);
    const char* p = (const char*)table;
    b.r2 = (const double*)(p + L.r2);
    b.r1 = (const double*)(p + L.r1);
    b.coef = (const double4*)(p + L.coef);
    b.lr = (const float*)(p + L.lr);
    b.c32 = (const Coef32*)(p + L.c32);
    b.levels = levels;
    b.ref_len = ref_len;
    return b;
}

// Highest level d in [1, hi] for which valid(d) holds, or 0; valid() must be
// monotone (true up to some level, false above).
template <class Pred>
__device__ inline int highest_valid_level(int hi, Pred valid) {
    int lo = 0;  // valid(lo) known true (level 0 checked by the caller)
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (valid(mid))
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Largest d with k + 2^d < ref_len (k < ref_len - 1 holds at the call).
__device__ inline int level_limit_by_length(int k, int ref_len) {
    unsigned rem = (unsigned)(ref_len - k - 1);  // >= 1
    return 31 - __builtin_clz(rem);
}

}  // namespace ap
