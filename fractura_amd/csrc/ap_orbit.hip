// Reference orbit in binary fixed point: host code, no kernels.
//
// The reference iterates Z_(k+1) = f(Z_k) in MPFR (gmpy2) or Decimal at
// compute_required_prec_bits(w) bits and stores float(Z_k) until
// |float Z_k|^2 > 65536 (compute_single_highprec_orbit,
// math_core/high_precision.py:706-875).  Here Z = X / 2^P with X a signed
// integer of fixed width (P = precision bits + guard bits), every product
// truncated towards minus infinity by a right shift, and float(Z) the
// correctly rounded value of X / 2^P (round half to even, subnormals
// included).  The storage rule, the escape test on the stored doubles, the
// period detection of types 0/2 (exact equality of the double and of the
// fixed-point value at doubling checkpoints, math_core/high_precision.py:
// 749-761) and the store of Z_ref_len after the loop follow the reference.
//
// Numbers are sign + magnitude over L 64-bit limbs, little endian; products
// are formed in 2L + 1 limbs (Karatsuba above 40 limbs, below it schoolbook
// by columns with a 192-bit accumulator; a square forms each cross product
// once).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "apeiron_hip.h"

namespace {

typedef unsigned __int128 u128;

struct Num {
    std::vector<uint64_t> m;  // magnitude
    bool neg = false;
    explicit Num(int n = 0) : m(n, 0) {}
};

int top_limb(const std::vector<uint64_t>& a) {
    for (int i = (int)a.size() - 1; i >= 0; --i)
        if (a[i]) return i;
    return -1;
}

bool is_zero(const std::vector<uint64_t>& a) { return top_limb(a) < 0; }

int cmp_mag(const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
    const int n = (int)(a.size() > b.size() ? a.size() : b.size());
    for (int i = n - 1; i >= 0; --i) {
        const uint64_t x = i < (int)a.size() ? a[i] : 0, y = i < (int)b.size() ? b[i] : 0;
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

// r = a + b (magnitudes, r sized by the caller, no overflow past r).
void add_mag(std::vector<uint64_t>& r, const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
    uint64_t carry = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        const u128 s = (u128)(i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0) + carry;
        r[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
}

// r = a - b with |a| >= |b|.
void sub_mag(std::vector<uint64_t>& r, const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
    uint64_t borrow = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        const uint64_t x = i < a.size() ? a[i] : 0, y = i < b.size() ? b[i] : 0;
        const uint64_t d = x - y - borrow;
        borrow = (x < y || (x == y && borrow)) ? 1 : 0;
        r[i] = d;
    }
}

// r = a + (negate_b ? -b : b), signed, exact (r may alias neither).
void add_signed(Num& r, const Num& a, const Num& b, bool negate_b = false) {
    const bool bneg = negate_b ? !b.neg : b.neg;
    if (a.neg == bneg) {
        add_mag(r.m, a.m, b.m);
        r.neg = a.neg;
    } else if (cmp_mag(a.m, b.m) >= 0) {
        sub_mag(r.m, a.m, b.m);
        r.neg = a.neg;
    } else {
        sub_mag(r.m, b.m, a.m);
        r.neg = bneg;
    }
    if (is_zero(r.m)) r.neg = false;
}

void sub_signed(Num& r, const Num& a, const Num& b) { add_signed(r, a, b, true); }

// 192-bit column accumulator of the product scan: lo (128 bits) + hi.
struct Acc {
    unsigned long long w0 = 0, w1 = 0, w2 = 0;
    void add(u128 p) {
        unsigned long long c;
        w0 = __builtin_addcll(w0, (unsigned long long)p, 0, &c);
        w1 = __builtin_addcll(w1, (unsigned long long)(p >> 64), c, &c);
        w2 += c;
    }
    // Emit the low word, shift right by 64.
    uint64_t take() {
        const uint64_t w = w0;
        w0 = w1;
        w1 = w2;
        w2 = 0;
        return w;
    }
};

// r[0, na + nb) = a[0, na) * b[0, nb) by columns (product scanning): the
// multiplies of a column are independent of each other, so they issue back
// to back instead of waiting on a row's carry chain.  sq: a == b (each cross
// product of a column is formed once and the sum doubled).
void comba(uint64_t* r, const uint64_t* a, int na, const uint64_t* b, int nb, bool sq) {
    Acc acc;
    const int ncol = na + nb - 1;
    if (sq) {
        for (int k = 0; k < ncol; ++k) {
            const int i0 = k - na + 1 > 0 ? k - na + 1 : 0;
            Acc cross;
            for (int i = i0; 2 * i < k; ++i) cross.add((u128)a[i] * a[k - i]);
            // acc += 2 * cross (+ the diagonal square for even k)
            acc.add(((u128)cross.w1 << 65) | ((u128)cross.w0 << 1));
            acc.w2 += (cross.w2 << 1) | (cross.w1 >> 63);
            if ((k & 1) == 0) acc.add((u128)a[k >> 1] * a[k >> 1]);
            r[k] = acc.take();
        }
    } else {
        for (int k = 0; k < ncol; ++k) {
            const int i0 = k - nb + 1 > 0 ? k - nb + 1 : 0;
            const int i1 = k < na - 1 ? k : na - 1;
            for (int i = i0; i <= i1; ++i) acc.add((u128)a[i] * b[k - i]);
            r[k] = acc.take();
        }
    }
    r[ncol] = acc.take();
}

// r[0, n) += a[0, m) (m <= n), carries propagated to r's end; returns the
// carry out of r.
uint64_t add_into(uint64_t* r, int n, const uint64_t* a, int m) {
    unsigned long long c = 0;
    int i = 0;
    for (; i < m; ++i) r[i] = __builtin_addcll(r[i], a[i], c, &c);
    for (; i < n && c; ++i) r[i] = __builtin_addcll(r[i], 0, c, &c);
    return c;
}

// r[0, n) -= a[0, m) (m <= n, the result non-negative).
void sub_from(uint64_t* r, int n, const uint64_t* a, int m) {
    unsigned long long c = 0;
    int i = 0;
    for (; i < m; ++i) r[i] = __builtin_subcll(r[i], a[i], c, &c);
    for (; i < n && c; ++i) r[i] = __builtin_subcll(r[i], 0, c, &c);
}

// Below this many limbs the column product is faster than a Karatsuba split.
constexpr int kKaratsuba = 40;

// r[0, 2n) = a[0, n) * b[0, n) (sq: a == b), Karatsuba: with a = a1 B^h + a0,
// a b = a1 b1 B^2h + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) B^h + a0 b0, every
// part exact.  t: scratch of at least 6 (n + 2) limbs per level.
void kmul(uint64_t* r, const uint64_t* a, const uint64_t* b, int n, bool sq, uint64_t* t) {
    if (n < kKaratsuba) {
        comba(r, a, n, b, n, sq);
        return;
    }
    const int h = n / 2, m = n - h;  // low h limbs, high m >= h limbs
    kmul(r, a, b, h, sq, t);                  // a0 b0 -> r[0, 2h)
    kmul(r + 2 * h, a + h, b + h, m, sq, t);  // a1 b1 -> r[2h, 2n)
    const int s = m + 1;
    uint64_t* sa = t;
    uint64_t* sb = t + s;
    uint64_t* mid = t + 2 * s;  // 2s limbs
    memcpy(sa, a + h, m * sizeof(uint64_t));
    sa[m] = 0;
    add_into(sa, s, a, h);
    if (!sq) {
        memcpy(sb, b + h, m * sizeof(uint64_t));
        sb[m] = 0;
        add_into(sb, s, b, h);
    }
    kmul(mid, sa, sq ? sa : sb, s, sq, t + 4 * s);
    sub_from(mid, 2 * s, r, 2 * h);
    sub_from(mid, 2 * s, r + 2 * h, 2 * m);
    // mid < 2^(64 (n + 2)); add it at limb h (the total fits in 2n limbs).
    int len = 2 * s;
    while (len > 0 && mid[len - 1] == 0) --len;
    add_into(r + h, 2 * n - h, mid, len);
}

// r (size >= |a| + |b| limbs) = |a| * |b|.  Operands of similar length above
// the Karatsuba threshold (the orbit's squares and products) split; others go
// by columns.
void mul_mag(std::vector<uint64_t>& r, const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
    const int na = top_limb(a) + 1, nb = top_limb(b) + 1;
    if (na == 0 || nb == 0) {
        std::fill(r.begin(), r.end(), 0);
        return;
    }
    const bool sq = &a == &b;
    const int n = na > nb ? na : nb;
    int done;
    const bool balanced = 2 * na > n && 2 * nb > n;  // a short factor (type 4's k) goes by columns
    if (n >= kKaratsuba && balanced && (size_t)n <= a.size() && (size_t)n <= b.size() &&
        2 * (size_t)n <= r.size()) {
        // Both operands as n limbs (the shorter one's high limbs are zero).
        static thread_local std::vector<uint64_t> scratch;
        const size_t need = 16 * ((size_t)n + 8);
        if (scratch.size() < need) scratch.resize(need);
        kmul(r.data(), a.data(), b.data(), n, sq, scratch.data());
        done = 2 * n;
    } else {
        comba(r.data(), a.data(), na, b.data(), nb, sq);
        done = na + nb;
    }
    std::fill(r.begin() + done, r.end(), 0);
}

// Signed product of a and b into a wide number.
void mul_signed(Num& r, const Num& a, const Num& b) {
    mul_mag(r.m, a.m, b.m);
    r.neg = (a.neg != b.neg) && !is_zero(r.m);
}

void square(Num& r, const Num& a) {
    mul_mag(r.m, a.m, a.m);
    r.neg = false;
}

// r (L limbs) = floor(w / 2^s): the magnitude shifted, plus one for a
// negative value whose dropped bits are not all zero.
void shift_floor(Num& r, const Num& w, int s) {
    const int q = s >> 6, b = s & 63;
    bool dropped = false;  // only a negative value needs it
    if (w.neg) {
        for (int i = 0; i < q && i < (int)w.m.size() && !dropped; ++i) dropped = w.m[i] != 0;
        if (b && q < (int)w.m.size()) dropped |= (w.m[q] & ((uint64_t(1) << b) - 1)) != 0;
    }
    for (size_t i = 0; i < r.m.size(); ++i) {
        const size_t k = i + q;
        const uint64_t lo = k < w.m.size() ? w.m[k] : 0;
        const uint64_t hi = k + 1 < w.m.size() ? w.m[k + 1] : 0;
        r.m[i] = b ? (lo >> b) | (hi << (64 - b)) : lo;
    }
    r.neg = w.neg;
    if (w.neg && dropped) {
        uint64_t carry = 1;
        for (size_t i = 0; i < r.m.size() && carry; ++i) {
            r.m[i] += carry;
            carry = r.m[i] == 0;
        }
    }
    if (is_zero(r.m)) r.neg = false;
}

int bit_length(const std::vector<uint64_t>& a) {
    const int t = top_limb(a);
    return t < 0 ? 0 : 64 * t + (64 - __builtin_clzll(a[t]));
}

bool bit_at(const std::vector<uint64_t>& a, long i) {
    if (i < 0 || (size_t)(i >> 6) >= a.size()) return false;
    return (a[i >> 6] >> (i & 63)) & 1;
}

// Bits [pos, pos + 64) of a (pos >= 0).
uint64_t window(const std::vector<uint64_t>& a, long pos) {
    const long q = pos >> 6;
    const int b = (int)(pos & 63);
    const uint64_t lo = q < (long)a.size() ? a[q] : 0;
    const uint64_t hi = q + 1 < (long)a.size() ? a[q + 1] : 0;
    return b ? (lo >> b) | (hi << (64 - b)) : lo;
}

bool any_below(const std::vector<uint64_t>& a, long i) {  // bits [0, i)
    if (i <= 0) return false;
    const long q = i >> 6;
    for (long k = 0; k < q && k < (long)a.size(); ++k)
        if (a[k]) return true;
    if ((i & 63) && q < (long)a.size()) return (a[q] & ((uint64_t(1) << (i & 63)) - 1)) != 0;
    return false;
}

// Correctly rounded double of x / 2^P (round half to even; the quantum is
// 2^(e-52) for a value in [2^e, 2^(e+1)), and 2^-1074 below 2^-1022).
double to_double(const Num& x, int P) {
    const int nb = bit_length(x.m);
    if (nb == 0) return 0.0;
    const long e = (long)nb - 1 - P;
    const long q = e - 52 > -1074 ? e - 52 : -1074;  // exponent of the last kept bit
    const long drop = q + P;                          // low bits of x below it
    uint64_t m = 0;
    if (drop <= 0) {
        m = x.m[0] << -drop;  // exact: the whole value has at most 53 bits
    } else {
        const long kept = (long)nb - drop;  // <= 53; <= 0 below half the quantum
        m = kept > 0 ? window(x.m, drop) & ((uint64_t(1) << kept) - 1) : 0;
        const bool half = bit_at(x.m, drop - 1);
        const bool rest = any_below(x.m, drop - 1);
        if (half && (rest || (m & 1))) ++m;
    }
    const double v = ldexp((double)m, (int)q);
    return x.neg ? -v : v;
}

bool equal_num(const Num& a, const Num& b) { return a.neg == b.neg && cmp_mag(a.m, b.m) == 0; }

}  // namespace

extern "C" int ap_reference_orbit(int32_t fractal_type, int32_t max_iter, int32_t frac_bits,
                                  int32_t limbs, const uint64_t* init, const int32_t* signs,
                                  int32_t gen_n, double* re, double* im, int32_t* ref_len) {
    if (!init || !signs || !re || !im || !ref_len) return AP_ERR_ARG;
    if (max_iter < 0 || frac_bits < 2 || limbs < 1 || limbs > AP_MAX_ORBIT_LIMBS) return AP_ERR_ARG;
    const int type = fractal_type;
    if (type < 0 || type > 4 || (type == 4 && (gen_n < 2 || gen_n > 16))) return AP_ERR_ARG;
    const int P = frac_bits;
    // Values stay below 2^(P + 8 n + 24) (|Z| <= 256 when multiplied, n
    // factors for type 4); products below twice that.
    const int n_pow = type == 4 ? gen_n : 2;
    // Inputs wider than that escape before they are multiplied.
    const int L = std::max((P + 8 * n_pow + 24) / 64 + 2, limbs + 2);
    auto load = [&](int i) {
        Num x(L);
        for (int k = 0; k < limbs; ++k) x.m[k] = init[(size_t)i * limbs + k];
        x.neg = signs[i] != 0 && !is_zero(x.m);
        return x;
    };
    // init: X0, Y0, CX, CY, KR, KI
    Num X = load(0), Y = load(1);
    const Num CX = load(2), CY = load(3), KR = load(4), KI = load(5);
    for (int k = 0; k <= max_iter; ++k) re[k] = im[k] = 0.0;
    Num X2(L), Y2(L), t(L), u(L), v(L), sX(L), sY(L);
    Num w1(2 * L + 1), w2(2 * L + 1), w3(2 * L + 1);
    int len = max_iter;
    if (type == 0 || type == 2 || type == 1 || type == 3) {
        const bool periodic = type == 0 || type == 2;
        int check_step = 8, saved_k = 0;
        double sfx = 0.0, sfy = 0.0;
        sX = X;
        sY = Y;
        int k = 0;
        for (; k < max_iter; ++k) {
            const double rx = to_double(X, P), ry = to_double(Y, P);
            re[k] = rx;
            im[k] = ry;
            if (rx * rx + ry * ry > 65536.0) {
                len = k;
                break;
            }
            if (periodic) {
                if (k == check_step) {
                    check_step <<= 1;
                    saved_k = k;
                    sX = X;
                    sY = Y;
                    sfx = rx;
                    sfy = ry;
                } else if (k > saved_k && rx == sfx && ry == sfy && equal_num(X, sX) &&
                           equal_num(Y, sY)) {
                    const int period = k - saved_k;
                    for (int j = k; j <= max_iter; ++j) {
                        re[j] = re[j - period];
                        im[j] = im[j - period];
                    }
                    *ref_len = max_iter;
                    return AP_OK;
                }
            }
            square(w1, X);
            shift_floor(X2, w1, P);
            square(w2, Y);
            shift_floor(Y2, w2, P);
            if (periodic) {
                mul_signed(w3, X, Y);
                shift_floor(t, w3, P - 1);
                add_signed(Y, t, CY);
            } else {
                mul_mag(w3.m, X.m, Y.m);  // |X| |Y|
                w3.neg = false;
                shift_floor(t, w3, P - 1);
                sub_signed(Y, t, CY);
            }
            sub_signed(t, X2, Y2);
            add_signed(X, t, CX);
        }
    } else {
        Num pr(L), pi(L), npr(L), npi(L), kzr(L), kzi(L);
        int k = 0;
        for (; k < max_iter; ++k) {
            const double rx = to_double(X, P), ry = to_double(Y, P);
            re[k] = rx;
            im[k] = ry;
            if (rx * rx + ry * ry > 65536.0) {
                len = k;
                break;
            }
            pr = X;
            pi = Y;
            for (int e = 0; e < gen_n - 1; ++e) {
                mul_signed(w1, pr, X);
                mul_signed(w2, pi, Y);
                sub_signed(w3, w1, w2);
                shift_floor(npr, w3, P);
                mul_signed(w1, pr, Y);
                mul_signed(w2, pi, X);
                add_signed(w3, w1, w2);
                shift_floor(npi, w3, P);
                pr = npr;
                pi = npi;
            }
            mul_signed(w1, KR, X);
            mul_signed(w2, KI, Y);
            sub_signed(w3, w1, w2);
            shift_floor(kzr, w3, P);
            mul_signed(w1, KR, Y);
            mul_signed(w2, KI, X);
            add_signed(w3, w1, w2);
            shift_floor(kzi, w3, P);
            add_signed(t, pr, kzr);
            add_signed(X, t, CX);
            add_signed(u, pi, kzi);
            add_signed(Y, u, CY);
        }
    }
    re[len] = to_double(X, P);
    im[len] = to_double(Y, P);
    *ref_len = len;
    return AP_OK;
}
