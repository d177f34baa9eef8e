// Multi-limb fixed-point escape-time evaluator: the Strict-Metric exhaustive
// fallback for pixels still glitched after rebasing.
//
// Numbers are signed magnitudes of L base-2^32 limbs, limb 0 the integer part
// and limbs 1..L-1 the fraction (math_core/cuda_kernels.py:2237-2364).  One
// wave64 evaluates one pixel: limb q = lane + 64 m sits in register m of
// lane `lane`, so a multi-limb operation is a handful of vector instructions
// instead of a serial loop per lane, and the few residual pixels of a frame
// still spread over every SIMD.
//
//   product   column sums over L steps: step t broadcasts a_t (readlane) and
//             multiplies it into b shifted t limbs up (DPP wave_shr:1), with
//             a 96-bit accumulator per column (v_mad_u64_u32 + carry word);
//             the reference's truncation (the high words of the first
//             dropped column kept, its low words dropped) is one reversed
//             product per lane and a wave sum;
//   add / sub carries by lookahead: generate and propagate bits of every
//             limb are two ballots, and one 64-bit add over the bit-reversed
//             masks resolves the carry chain of the whole number;
//   compare   the most significant differing limb is the first set bit of
//             one ballot.
//
// Semantics 0 restates compute_ml_sample (math_core/cuda_kernels.py:
// 2367-2467) at L = 34, bit-for-bit: integer-limb escape test, smooth value
// from the top two limbs of |z|^2, float32 map value, -1 for the interior,
// type 4 not evaluated.  Semantics 1 restates compute_highprec_escape_scalar
// (math_core/high_precision.py:581-656), which the reference runs on the CPU
// in MPFR at prec_bits, in fixed point at L = 1 + ceil(prec_bits / 32) limbs:
// escape when |z|^2 > 65536 exactly, the period check of types 0/2, a
// float64 smooth value, max_iter for the interior.  Fixed point cannot hold
// |z| >= 2^16 squared or the powers of type 4 beyond 2^32; where that would
// happen the escape is already certain (|z| >= 2^15 before a square,
// |z^m| >= 2^22 inside the power), and |z|^2 is taken in float64 from the
// leading limbs.
#include "ap_device.hpp"
#include "ap_launch.hpp"

namespace ap {
namespace {

constexpr int kLimbWaves = 4;
constexpr int kLimbBlock = 64 * kLimbWaves;

__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }

// Value of register x in lane l (l wave-uniform).
__device__ inline uint32_t lane_value(uint32_t x, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, l);
}

// Limb q takes limb q-1 (limb 0 takes 0): DPP wave_shr:1, lane 0 of each
// register fed from lane 63 of the register before.
template <int P>
__device__ inline void shift_up(uint32_t (&x)[P]) {
    uint32_t in[P];
#pragma unroll
    for (int m = 1; m < P; ++m) in[m] = lane_value(x[m - 1], 63);
    // bound_ctrl: the lane without a source reads 0
    x[0] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[0], 0x138, 0xF, 0xF, true);
#pragma unroll
    for (int m = 1; m < P; ++m)
        x[m] = (uint32_t)__builtin_amdgcn_update_dpp((int)in[m], (int)x[m], 0x138, 0xF, 0xF, false);
}

// Limb q takes limb q+1 (the last takes 0): DPP wave_shl:1.
template <int P>
__device__ inline void shift_down(uint32_t (&x)[P]) {
    uint32_t in[P];
#pragma unroll
    for (int m = 0; m + 1 < P; ++m) in[m] = lane_value(x[m + 1], 0);
#pragma unroll
    for (int m = 0; m + 1 < P; ++m)
        x[m] = (uint32_t)__builtin_amdgcn_update_dpp((int)in[m], (int)x[m], 0x130, 0xF, 0xF, false);
    x[P - 1] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[P - 1], 0x130, 0xF, 0xF, true);
}

// Carry (or borrow) chain: g[q] says limb q sends a carry to limb q-1; a limb
// whose digit is all ones (all zeros for a borrow) passes on a carry it
// receives.  With the limb order reversed (least significant limb first) the
// carries received are (X + Y) ^ X ^ Y for X = generate | propagate and
// Y = generate, one 64-bit add per register.
template <int P, bool kBorrow>
__device__ inline void resolve(uint32_t (&x)[P], const bool (&g)[P], int L) {
    const int ln = lane_id();
    unsigned long long G[P], Pm[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const bool live = ln + 64 * m < L;
        G[m] = __ballot(g[m]);
        Pm[m] = __ballot(live && (kBorrow ? x[m] == 0u : x[m] == 0xFFFFFFFFu));
    }
    unsigned long long C[P];
    unsigned long long cin = 0;
#pragma unroll
    for (int w = 0; w < P; ++w) {
        const int m = P - 1 - w;
        const unsigned long long gw = __builtin_bitreverse64(G[m]);
        const unsigned long long X = gw | __builtin_bitreverse64(Pm[m]);
        const unsigned long long s1 = X + gw;
        const unsigned long long c1 = s1 < X ? 1ull : 0ull;
        const unsigned long long s2 = s1 + cin;
        const unsigned long long c2 = s2 < s1 ? 1ull : 0ull;
        cin = c1 | c2;
        C[m] = __builtin_bitreverse64(s2 ^ X ^ gw);
    }
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const uint32_t c = (uint32_t)(C[m] >> ln) & 1u;
        x[m] = kBorrow ? x[m] - c : x[m] + c;
    }
}

// ml_add_mag (:2299-2306); the carry out of limb 0 is dropped, as every
// caller of the reference drops it.
template <int P>
__device__ inline void add_mag(const uint32_t (&a)[P], const uint32_t (&b)[P], uint32_t (&out)[P],
                               int L) {
    bool g[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const uint32_t s = a[m] + b[m];
        g[m] = s < a[m];
        out[m] = s;
    }
    resolve<P, false>(out, g, L);
}

// ml_sub_mag (:2309-2320), a >= b.
template <int P>
__device__ inline void sub_mag(const uint32_t (&a)[P], const uint32_t (&b)[P], uint32_t (&out)[P],
                               int L) {
    bool g[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        g[m] = a[m] < b[m];
        out[m] = a[m] - b[m];
    }
    resolve<P, true>(out, g, L);
}

// ml_cmp_mag (:2289-2296).
template <int P>
__device__ inline int cmp_mag(const uint32_t (&a)[P], const uint32_t (&b)[P]) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const unsigned long long ne = __ballot(a[m] != b[m]);
        if (ne) {
            const int l = __builtin_ctzll(ne);
            return lane_value(a[m], l) > lane_value(b[m], l) ? 1 : -1;
        }
    }
    return 0;
}

template <int P>
__device__ inline bool is_zero(const uint32_t (&a)[P]) {
    bool z = true;
#pragma unroll
    for (int m = 0; m < P; ++m) z = z && __ballot(a[m] != 0u) == 0ull;
    return z;
}

__device__ inline unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// ml_mul_mag (:2343-2364): the product truncated to L limbs, where column k
// sums the low words of a_i b_j over i+j = k and the high words over
// i+j = k+1, and the last column adds the high words of the first dropped
// column (i+j = L).  As an integer that is
//   sum_{q<L} F_q 2^(32(L-1-q)) + H  (mod 2^(32L)),
//   F_q = sum_{i+j=q} a_i b_j,   H = sum_{i+j=L} hi32(a_i b_j),
// which is what is computed: F_q in 96 bits per lane (the column sums of the
// loop over t below), H by one reversed product per lane and a wave sum, then
// two carry passes.
template <int P>
struct Cols {
    unsigned long long acc[P];  // low 64 bits of F_q
    uint32_t top[P];            // bits 64..95
};

template <int P>
__device__ inline void cols_zero(Cols<P>& c) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
        c.acc[m] = 0ull;
        c.top[m] = 0u;
    }
}

template <int P>
__device__ inline void cols_add(Cols<P>& c, uint32_t at, const uint32_t (&bs)[P]) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const unsigned long long s = c.acc[m] + (unsigned long long)at * bs[m];
        c.top[m] += s < c.acc[m] ? 1u : 0u;
        c.acc[m] = s;
    }
}

// c += d (96-bit column sums).
template <int P>
__device__ inline void cols_merge(Cols<P>& c, const Cols<P>& d) {
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const unsigned long long s = c.acc[m] + d.acc[m];
        c.top[m] += d.top[m] + (s < c.acc[m] ? 1u : 0u);
        c.acc[m] = s;
    }
}

// The product a b from its column sums.
template <int P>
__device__ inline void cols_finish(const Cols<P>& c, const uint32_t (&a)[P], const uint32_t (&b)[P],
                                   uint32_t (&out)[P], int L) {
    const int ln = lane_id();
    // H: lane q (1 <= q <= L-1) multiplies b_q by a_(L-q).
    unsigned long long h = 0ull;
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int q = ln + 64 * m;
        const int src = L - q;
        const int addr = (src & 63) << 2;
        uint32_t av = 0u;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            const uint32_t v = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)a[r]);
            if ((src >> 6) == r) av = v;
        }
        if (q >= 1 && q <= L - 1) h += (unsigned long long)__umulhi(av, b[m]);
    }
    const unsigned long long H = wave_sum(h);
    // Limb q gathers lo32(F_q) + hi32(lo64(F_(q+1))) + top(F_(q+2)) (+ H at
    // q = L-1); columns q >= L are dropped.
    uint32_t w0[P], w1[P], w2[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const bool live = ln + 64 * m < L;
        w0[m] = live ? (uint32_t)c.acc[m] : 0u;
        w1[m] = live ? (uint32_t)(c.acc[m] >> 32) : 0u;
        w2[m] = live ? c.top[m] : 0u;
    }
    shift_down(w1);
    shift_down(w2);
    shift_down(w2);
    uint32_t lo[P], hi[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        unsigned long long v = (unsigned long long)w0[m] + w1[m] + w2[m];
        if (ln + 64 * m == L - 1) v += H;
        lo[m] = (uint32_t)v;
        hi[m] = (uint32_t)(v >> 32);
    }
    shift_down(hi);
    bool g[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const unsigned long long u = (unsigned long long)lo[m] + hi[m];
        out[m] = (uint32_t)u;
        g[m] = (u >> 32) != 0ull;
    }
    resolve<P, false>(out, g, L);
}

template <int P>
__device__ inline void mul_mag(const uint32_t (&a)[P], const uint32_t (&b)[P], uint32_t (&out)[P],
                               int L) {
    Cols<P> c;
    cols_zero(c);
    uint32_t bs[P];
#pragma unroll
    for (int m = 0; m < P; ++m) bs[m] = b[m];
#pragma unroll
    for (int ma = 0; ma < P; ++ma) {
        const int tend = min(64, L - 64 * ma);
        for (int t = 0; t < tend; ++t) {
            cols_add(c, lane_value(a[ma], t), bs);
            shift_up(bs);
        }
    }
    cols_finish(c, a, b, out, L);
}

// Products of two numbers by two others in one pass over t: p00 = a0 b0,
// p11 = a1 b1, p01 = a0 b1 and, when kCross, p10 = a1 b0.  The three or four
// multiply-accumulate chains are independent, so one wave overlaps them;
// each result equals mul_mag's.
template <int P, bool kCross>
__device__ inline void mul22(const uint32_t (&a0)[P], const uint32_t (&a1)[P],
                             const uint32_t (&b0)[P], const uint32_t (&b1)[P], uint32_t (&p00)[P],
                             uint32_t (&p11)[P], uint32_t (&p01)[P], uint32_t (&p10)[P], int L) {
    // Even and odd t accumulate into separate column sums: twice as many
    // independent chains in flight, merged once at the end.
    Cols<P> c00, c11, c01, c10, e00, e11, e01, e10;
    cols_zero(c00);
    cols_zero(c11);
    cols_zero(c01);
    cols_zero(e00);
    cols_zero(e11);
    cols_zero(e01);
    if (kCross) {
        cols_zero(c10);
        cols_zero(e10);
    }
    uint32_t s0[P], s1[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        s0[m] = b0[m];
        s1[m] = b1[m];
    }
#pragma unroll
    for (int ma = 0; ma < P; ++ma) {
        const int tend = min(64, L - 64 * ma);
        int t = 0;
        for (; t + 1 < tend; t += 2) {
            const uint32_t x = lane_value(a0[ma], t), y = lane_value(a1[ma], t);
            const uint32_t x1 = lane_value(a0[ma], t + 1), y1 = lane_value(a1[ma], t + 1);
            cols_add(c00, x, s0);
            cols_add(c11, y, s1);
            cols_add(c01, x, s1);
            if (kCross) cols_add(c10, y, s0);
            shift_up(s0);
            shift_up(s1);
            cols_add(e00, x1, s0);
            cols_add(e11, y1, s1);
            cols_add(e01, x1, s1);
            if (kCross) cols_add(e10, y1, s0);
            shift_up(s0);
            shift_up(s1);
        }
        if (t < tend) {
            const uint32_t x = lane_value(a0[ma], t), y = lane_value(a1[ma], t);
            cols_add(c00, x, s0);
            cols_add(c11, y, s1);
            cols_add(c01, x, s1);
            if (kCross) cols_add(c10, y, s0);
            shift_up(s0);
            shift_up(s1);
        }
    }
    cols_merge(c00, e00);
    cols_merge(c11, e11);
    cols_merge(c01, e01);
    if (kCross) cols_merge(c10, e10);
    cols_finish(c00, a0, b0, p00, L);
    cols_finish(c11, a1, b1, p11, L);
    cols_finish(c01, a0, b1, p01, L);
    if (kCross) cols_finish(c10, a1, b0, p10, L);
}

template <int P>
struct Num {
    uint32_t l[P];
    int sign;  // 1 = negative (wave-uniform)
};

template <int P>
__device__ inline void copy(Num<P>& d, const Num<P>& s) {
#pragma unroll
    for (int m = 0; m < P; ++m) d.l[m] = s.l[m];
    d.sign = s.sign;
}

template <int P>
__device__ inline void set_zero(Num<P>& d) {
#pragma unroll
    for (int m = 0; m < P; ++m) d.l[m] = 0u;
    d.sign = 0;
}

// ml_add_signed (:2323-2334) and ml_sub_signed (:2337-2340).
template <int P>
__device__ inline void add_signed(const Num<P>& a, int bsign, const uint32_t (&b)[P], Num<P>& out,
                                  int L) {
    if (a.sign == bsign) {
        add_mag(a.l, b, out.l, L);
        out.sign = a.sign;
    } else if (cmp_mag(a.l, b) >= 0) {
        sub_mag(a.l, b, out.l, L);
        out.sign = a.sign;
    } else {
        sub_mag(b, a.l, out.l, L);
        out.sign = bsign;
    }
}
template <int P>
__device__ inline void add_s(const Num<P>& a, const Num<P>& b, Num<P>& out, int L) {
    add_signed(a, b.sign, b.l, out, L);
}
template <int P>
__device__ inline void sub_s(const Num<P>& a, const Num<P>& b, Num<P>& out, int L) {
    add_signed(a, b.sign == 0 ? 1 : 0, b.l, out, L);
}
template <int P>
__device__ inline void mul_s(const Num<P>& a, const Num<P>& b, Num<P>& out, int L) {
    mul_mag(a.l, b.l, out.l, L);
    out.sign = a.sign != b.sign ? 1 : 0;
}

// Leading limbs as a double: the top 64 bits rounded once, plus the third limb.
template <int P>
__device__ inline double mag_double(const uint32_t (&a)[P]) {
    const unsigned long long top =
        ((unsigned long long)lane_value(a[0], 0) << 32) | lane_value(a[0], 1);
    return (double)top * 2.3283064365386963e-10 + (double)lane_value(a[0], 2) * 5.421010862427522e-20;
}
template <int P>
__device__ inline double signed_double(const Num<P>& a) {
    const double d = mag_double(a.l);
    return a.sign ? -d : d;
}

template <int P>
__device__ inline bool equal_value(const Num<P>& a, const Num<P>& b) {
    if (cmp_mag(a.l, b.l) != 0) return false;
    return a.sign == b.sign || is_zero(a.l);
}

template <int P>
__device__ inline void load(Num<P>& d, const uint32_t* limbs, int sign, int L) {
    const int ln = lane_id();
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int q = ln + 64 * m;
        d.l[m] = q < L ? limbs[q] : 0u;
    }
    d.sign = sign;
}

struct LimbResult {
    int count;
    float s;        // float32 map value
    double smooth;  // float64 smooth value, -1 interior
    double surv;    // value the arbitrary-precision tier colours from
};

__device__ inline LimbResult limb_escape(const ap_view& v, int k, double mag2, bool gpu34) {
    LimbResult r;
    r.count = k;
    double inv_ref, inv_exact;
    inv_ln_constants(v, inv_ref, inv_exact);
    const double lm = log(mag2);
    if (lm > 1.0) {
        const double ll = log(0.5 * lm);
        r.smooth = (double)(k + 1) - (ll - kLnLn256) * inv_exact;
        if (gpu34) {
            r.s = (float)(k + 1.0) - (float)((ll - v.ln_bailout_log) * v.inv_ln2);
            r.surv = (double)r.s;
        } else {
            // float64 throughout, 1/ln n unrounded for type 4
            // (math_core/high_precision.py:614, 637, 645).
            const double inv = (v.fractal_type == AP_GENERAL_MANDELBROT && v.gen_n > 1)
                                   ? 1.0 / log((double)v.gen_n)
                                   : v.inv_ln2;
            r.surv = (k + 1.0) - (ll - v.ln_bailout_log) * inv;
            r.s = (float)r.surv;
        }
    } else {
        r.smooth = (double)k;
        r.surv = (double)k;
        r.s = (float)k;
    }
    return r;
}

__device__ inline LimbResult limb_interior(const ap_view& v, bool gpu34) {
    LimbResult r;
    r.count = v.max_iter;
    r.smooth = -1.0;
    // The GPU tier returns -1; the arbitrary-precision tier float(max_iter)
    // (math_core/engine.py:250-251).
    r.surv = gpu34 ? -1.0 : (double)v.max_iter;
    r.s = (float)r.surv;
    return r;
}

struct LimbArgs {
    const int32_t* queue;
    int n;
    int L;
    const uint32_t* coords;  // n x 2 x L
    const int32_t* signs;    // n x 2
    const uint32_t* params;  // 4 x L: julia_cx, julia_cy, gen_kr, gen_ki
    const int32_t* psigns;   // 4
};

template <int P, bool kGpu34>
__device__ LimbResult limb_sample(const ap_view& v, const LimbArgs& A, int px) {
    const int L = A.L;
    const int type = v.fractal_type;
    if (kGpu34 && type == AP_GENERAL_MANDELBROT) {
        LimbResult r;
        r.count = -1;
        r.s = -1.0f;
        r.smooth = -1.0;
        r.surv = -1.0;
        return r;
    }
    Num<P> zx, zy, cx, cy, zx2, zy2, zxy, sub, sx, sy;
    {
        Num<P> x, y;
        load(x, A.coords + (size_t)(2 * px) * L, A.signs[2 * px], L);
        load(y, A.coords + (size_t)(2 * px + 1) * L, A.signs[2 * px + 1], L);
        if (type == AP_JULIA_MANDELBROT || type == AP_JULIA_BURNING_SHIP) {
            copy(zx, x);
            copy(zy, y);
            load(cx, A.params, A.psigns[0], L);
            load(cy, A.params + L, A.psigns[1], L);
        } else {
            set_zero(zx);
            set_zero(zy);
            copy(cx, x);
            copy(cy, y);
        }
    }
    const bool period = !kGpu34 && (type == AP_MANDELBROT || type == AP_JULIA_MANDELBROT);
    int check_step = 8, saved_k = 0;
    copy(sx, zx);
    copy(sy, zy);
    for (int k = 0; k < v.max_iter; ++k) {
        if (kGpu34 && (type == AP_BURNING_SHIP || type == AP_JULIA_BURNING_SHIP)) {
            zx.sign = 0;
            zy.sign = 0;
        }
        if (!kGpu34) {
            if (lane_value(zx.l[0], 0) >= (1u << 15) || lane_value(zy.l[0], 0) >= (1u << 15)) {
                const double dx = mag_double(zx.l), dy = mag_double(zy.l);
                return limb_escape(v, k, dx * dx + dy * dy, false);
            }
        }
        // zx^2, zy^2 and |zx||zy| in one pass (the cross product is used by
        // the update of types 0-3; its sign is set there).
        {
            uint32_t unused[P];
            mul22<P, false>(zx.l, zy.l, zx.l, zy.l, zx2.l, zy2.l, zxy.l, unused, L);
        }
        zx2.sign = 0;
        zy2.sign = 0;
        if (kGpu34) {
            const uint32_t x0 = lane_value(zx2.l[0], 0), y0 = lane_value(zy2.l[0], 0);
            if (x0 >= 65536u || y0 >= 65536u || (unsigned long long)x0 + y0 >= 65536ull) {
                const double mag2 = (double)x0 + (double)y0 +
                                    ((double)lane_value(zx2.l[0], 1) + (double)lane_value(zy2.l[0], 1)) *
                                        2.3283064365386963e-10;
                return limb_escape(v, k, mag2, true);
            }
        } else {
            Num<P> m2;
            add_mag(zx2.l, zy2.l, m2.l, L);
            const uint32_t m0 = lane_value(m2.l[0], 0);
            bool over = m0 > 65536u;
            if (m0 == 65536u) {
                bool frac = false;
#pragma unroll
                for (int m = 0; m < P; ++m)
                    frac = frac || __ballot(m2.l[m] != 0u && lane_id() + 64 * m != 0) != 0ull;
                over = frac;
            }
            if (over) return limb_escape(v, k, mag_double(m2.l), false);
            if (period) {
                if (k == check_step) {
                    check_step <<= 1;
                    saved_k = k;
                    copy(sx, zx);
                    copy(sy, zy);
                } else if (k > saved_k && equal_value(zx, sx) && equal_value(zy, sy)) {
                    return limb_interior(v, false);
                }
            }
        }
        if (type == AP_MANDELBROT || type == AP_JULIA_MANDELBROT) {
            // zy = 2 zx zy + cy; zx = (zx^2 - zy^2) + cx
            zxy.sign = zx.sign != zy.sign ? 1 : 0;
            Num<P> two;
            add_mag(zxy.l, zxy.l, two.l, L);
            two.sign = zxy.sign;
            sub_s(zx2, zy2, sub, L);
            add_s(sub, cx, zx, L);
            add_s(two, cy, zy, L);
        } else if (type == AP_BURNING_SHIP || type == AP_JULIA_BURNING_SHIP) {
            // zy = 2 |zx| |zy| - cy; zx = (zx^2 - zy^2) + cx
            Num<P> two;
            add_mag(zxy.l, zxy.l, two.l, L);
            two.sign = 0;
            sub_s(zx2, zy2, sub, L);
            add_s(sub, cx, zx, L);
            sub_s(two, cy, zy, L);
        } else {
            // z^n by repeated multiplication, + k z + c
            Num<P> zr, zi, t0, t1, kr, ki;
            copy(zr, zx);
            copy(zi, zy);
            bool bail = false;
            for (int e = 0; e < v.gen_n - 1; ++e) {
                if (lane_value(zr.l[0], 0) >= (1u << 22) || lane_value(zi.l[0], 0) >= (1u << 22)) {
                    bail = true;
                    break;
                }
                // zr zx, zi zy, zr zy, zi zx in one pass
                Num<P> p2, p3;
                mul22<P, true>(zr.l, zi.l, zx.l, zy.l, t0.l, t1.l, p2.l, p3.l, L);
                t0.sign = zr.sign != zx.sign ? 1 : 0;
                t1.sign = zi.sign != zy.sign ? 1 : 0;
                p2.sign = zr.sign != zy.sign ? 1 : 0;
                p3.sign = zi.sign != zx.sign ? 1 : 0;
                sub_s(t0, t1, zr, L);
                add_s(p2, p3, zi, L);
            }
            load(kr, A.params + 2 * L, A.psigns[2], L);
            load(ki, A.params + 3 * L, A.psigns[3], L);
            if (bail) {
                // |z^n| >= 2^22 - |k z| - |c| > 256: escaped at k + 1.
                const double dx = signed_double(zx), dy = signed_double(zy);
                double pr = dx, pi = dy;
                for (int e = 0; e < v.gen_n - 1; ++e) {
                    const double t = pr * dx - pi * dy;
                    pi = pr * dy + pi * dx;
                    pr = t;
                }
                const double dkr = signed_double(kr), dki = signed_double(ki);
                const double nx = pr + (dkr * dx - dki * dy) + signed_double(cx);
                const double ny = pi + (dkr * dy + dki * dx) + signed_double(cy);
                if (k + 1 < v.max_iter) return limb_escape(v, k + 1, nx * nx + ny * ny, false);
                return limb_interior(v, false);
            }
            // kr zx, ki zy, kr zy, ki zx in one pass
            Num<P> kz, p2, p3;
            mul22<P, true>(kr.l, ki.l, zx.l, zy.l, t0.l, t1.l, p2.l, p3.l, L);
            t0.sign = kr.sign != zx.sign ? 1 : 0;
            t1.sign = ki.sign != zy.sign ? 1 : 0;
            p2.sign = kr.sign != zy.sign ? 1 : 0;
            p3.sign = ki.sign != zx.sign ? 1 : 0;
            sub_s(t0, t1, kz, L);   // kz_r
            add_s(zr, kz, t0, L);   // zn_r + kz_r
            add_s(p2, p3, kz, L);   // kz_i
            add_s(zi, kz, t1, L);   // zn_i + kz_i
            add_s(t0, cx, zx, L);
            add_s(t1, cy, zy, L);
        }
    }
    return limb_interior(v, kGpu34);
}

template <int P, bool kGpu34>
__global__ __launch_bounds__(kLimbBlock) void limb_pass(ap_view v, ap_outputs o, LimbArgs A) {
    const int px = blockIdx.x * kLimbWaves + (int)(threadIdx.x >> 6);
    if (px >= A.n) return;  // whole waves
    const LimbResult r = limb_sample<P, kGpu34>(v, A, px);
    if (lane_id() != 0) return;
    const long long idx = A.queue[px];
    if (o.counts) o.counts[idx] = r.count;
    if (o.smooth) o.smooth[idx] = r.smooth;
    if (o.iter_map) o.iter_map[idx] = r.s;
    if (o.glitch) o.glitch[idx] = 0;
    if (o.rgb) {
        double cr = 0.0, cg = 0.0, cb = 0.0;
        if (kGpu34) {
            if (r.s >= 0.0f) sample_colour(v, o, r.s, cr, cg, cb);
        } else if (!(r.surv >= (double)v.max_iter || r.surv < 0.0)) {
            sample_colour(v, o, r.surv, cr, cg, cb);
        }
        store_rgb(o.rgb, v.is_hdr, idx, cr, cg, cb);
    }
}

template <int P>
hipError_t launch_p(bool gpu34, const ap_view& v, const ap_outputs& o, const LimbArgs& A,
                    hipStream_t s) {
    const dim3 grid((A.n + kLimbWaves - 1) / kLimbWaves);
    if (gpu34)
        hipLaunchKernelGGL((limb_pass<P, true>), grid, dim3(kLimbBlock), 0, s, v, o, A);
    else
        hipLaunchKernelGGL((limb_pass<P, false>), grid, dim3(kLimbBlock), 0, s, v, o, A);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_limbs(int semantics, int limbs, const ap_view& v, const int32_t* queue, int n,
                        const uint32_t* coords, const int32_t* signs, const uint32_t* params,
                        const int32_t* psigns, const ap_outputs& o, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const LimbArgs A{queue, n, limbs, coords, signs, params, psigns};
    const bool gpu34 = semantics == 0;
    if (limbs <= 64) return launch_p<1>(gpu34, v, o, A, s);
    if (limbs <= 128) return launch_p<2>(gpu34, v, o, A, s);
    if (limbs <= AP_MAX_LIMBS) return launch_p<4>(gpu34, v, o, A, s);
    return hipErrorInvalidValue;
}

}  // namespace ap
