// Series approximation of the first iterations of the perturbation loop
// (BBSA): host code, no kernels.
//
// For the standard FP64-delta perturbation (no FloatExp scale), the reference
// replaces the first `skip` iterations of every pixel by a polynomial in
// u = dc / r, r = 0.75 w (math_core/engine.py:1533-1546):
//   dz_skip = A u + B u^2 + ... (+ H u^8)
// whose coefficients it iterates along the reference orbit while the
// truncation error stays below tol relative to the linear term
// (_jit_compute_bbsa_coeffs, math_core/high_precision.py:51-536).  The
// recurrence is the composition with z -> 2 Z dz + dz^2 + dc (types 0/2) or
// z -> (3 Z^2 + k) dz + 3 Z dz^2 + dz^3 + dc (type 4, n = 3), truncated to
// the order: with a_j the coefficients after k steps and
//   q_j = sum_{i+l=j} a_i a_l          (the square, pairs doubled)
//   w_j = sum_{m=2}^{j-1} q_m a_(j-m)  (the cube, type 4)
// the next coefficients are
//   types 0/2: a_1' = 2 Z a_1 + c,  a_j' = 2 Z a_j + q_j
//   type 4:    a_1' = C1 a_1 + c,   a_2' = C1 a_2 + C2 q_2,
//              a_j' = (C1 a_j + C2 q_j) + w_j,  C1 = 3 Z^2 + k, C2 = 3 Z,
// c = r for types 0 and 4 (dc = r u), 0 for Julia (whose a_1 starts at r).
// Each sum is accumulated in ascending index, and doubled before the square
// term is added, as the reference writes it; with contraction off every
// coefficient rounds as its float64 formula does.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "apeiron_hip.h"

namespace {

struct Cx {
    double r, i;
};

inline Cx cmul(Cx a, Cx b) { return Cx{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
inline Cx cadd(Cx a, Cx b) { return Cx{a.r + b.r, a.i + b.i}; }
inline Cx cscale(double s, Cx a) { return Cx{s * a.r, s * a.i}; }
inline double cabs_(Cx a) { return sqrt(a.r * a.r + a.i * a.i); }
inline bool is_finite(double x) { return !(isnan(x) || isinf(x)); }

// q_j from a[1..]: the pairs (i, j-i), i < j-i, summed in ascending i and
// doubled, then the square a_(j/2)^2 for even j (q_2 is the square alone).
inline Cx square_term(const Cx* a, int j) {
    Cx s{0.0, 0.0};
    bool any = false;
    for (int i = 1; 2 * i < j; ++i) {
        const Cx t = cmul(a[i], a[j - i]);
        s = any ? cadd(s, t) : t;
        any = true;
    }
    if (any) s = cscale(2.0, s);
    if (j % 2 == 0) {
        const Cx sq = cmul(a[j / 2], a[j / 2]);
        s = any ? cadd(s, sq) : sq;
    }
    return s;
}

// w_j = q_2 a_(j-2) + q_3 a_(j-3) + ... + q_(j-1) a_1, left to right.
inline Cx cube_term(const Cx* q, const Cx* a, int j) {
    Cx s = cmul(q[2], a[j - 2]);
    for (int m = 3; m <= j - 1; ++m) s = cadd(s, cmul(q[m], a[j - m]));
    return s;
}

}  // namespace

extern "C" int ap_bbsa_coeffs(const double* ref_re, const double* ref_im, int32_t ref_len,
                              double max_r, int32_t fractal_type, double tol, int32_t order,
                              int32_t gen_n, double gen_kr, double gen_ki, int32_t e_scale,
                              double* out) {
    if (!out) return AP_ERR_ARG;
    memset(out, 0, 17 * sizeof(double));
    const int type = fractal_type;
    if (!(type == 0 || type == 2 || type == 4) || !(tol > 0.0) || order <= 0 || ref_len < 2 ||
        !(max_r > 0.0))
        return AP_OK;
    if (!ref_re || !ref_im) return AP_ERR_ARG;
    // Chained BLA replaces the series; a FloatExp frame runs the scaled
    // recurrence from k = 0 (math_core/high_precision.py:61-77).
    if (order == 32 || !is_finite(max_r) || e_scale > 0) return AP_OK;
    if (type == 4 && gen_n != 3) return AP_OK;
    const int N = order == 4 ? 4 : 8;
    const double r = max_r;
    const Cx cconst{type == 2 ? 0.0 : r, 0.0};
    Cx a[10], q[10], nx[10], best[10];
    for (int j = 0; j < 10; ++j) a[j] = q[j] = nx[j] = best[j] = Cx{0.0, 0.0};
    a[1] = Cx{type == 2 ? r : 0.0, 0.0};
    best[1] = a[1];
    int best_skip = 0;
    for (int k = 0; k < ref_len; ++k) {
        const Cx Z{ref_re[k], ref_im[k]};
        // q_2 .. q_(N+1): q_(N+1) is the first dropped term (types 0/2).
        for (int j = 2; j <= N + 1; ++j) q[j] = square_term(a, j);
        double err;
        if (type == 4) {
            const Cx z2 = cmul(Z, Z);
            const Cx c1{3.0 * z2.r + gen_kr, 3.0 * z2.i + gen_ki};
            const Cx c2{3.0 * Z.r, 3.0 * Z.i};
            nx[1] = cadd(cmul(c1, a[1]), cconst);
            nx[2] = cadd(cmul(c1, a[2]), cmul(c2, q[2]));
            for (int j = 3; j <= N; ++j)
                nx[j] = cadd(cadd(cmul(c1, a[j]), cmul(c2, q[j])), cube_term(q, a, j));
            err = cabs_(nx[N]) * 0.1;
        } else {
            const Cx c1{2.0 * Z.r, 2.0 * Z.i};
            nx[1] = cadd(cmul(c1, a[1]), cconst);
            for (int j = 2; j <= N; ++j) nx[j] = cadd(cmul(c1, a[j]), q[j]);
            err = cabs_(q[N + 1]) + cabs_(nx[N]) * 0.1;
        }
        const double lin = cabs_(nx[1]);
        const double top = cabs_(nx[N]);
        if (!is_finite(lin) || !is_finite(top)) break;
        if (lin > 0.0 && ((err / lin) > tol || lin > 0.5)) break;
        for (int j = 1; j <= N; ++j) a[j] = best[j] = nx[j];
        best_skip = k + 1;
    }
    const int cap = ref_len - 2 > 0 ? ref_len - 2 : 0;
    out[0] = (double)(best_skip < cap ? best_skip : cap);
    for (int j = 1; j <= N; ++j) {
        out[2 * j - 1] = best[j].r;
        out[2 * j] = best[j].i;
    }
    return AP_OK;
}
