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
// which is what is computed: F_q in 96 bits per lane, H by one reversed
// product per lane and a wave sum, then two carry passes.
template <int P>
__device__ inline void mul_mag(const uint32_t (&a)[P], const uint32_t (&b)[P], uint32_t (&out)[P],
                               int L) {
    const int ln = lane_id();
    unsigned long long acc[P];
    uint32_t top[P], bs[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        acc[m] = 0ull;
        top[m] = 0u;
        bs[m] = b[m];
    }
#pragma unroll
    for (int ma = 0; ma < P; ++ma) {
        const int tend = min(64, L - 64 * ma);
        for (int t = 0; t < tend; ++t) {
            const uint32_t at = lane_value(a[ma], t);
#pragma unroll
            for (int m = 0; m < P; ++m) {
                const unsigned long long s = acc[m] + (unsigned long long)at * bs[m];
                top[m] += s < acc[m] ? 1u : 0u;
                acc[m] = s;
            }
            shift_up(bs);
        }
    }
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
        w0[m] = live ? (uint32_t)acc[m] : 0u;
        w1[m] = live ? (uint32_t)(acc[m] >> 32) : 0u;
        w2[m] = live ? top[m] : 0u;
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
        mul_mag(zx.l, zx.l, zx2.l, L);
        mul_mag(zy.l, zy.l, zy2.l, L);
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
            mul_s(zx, zy, zxy, L);
            Num<P> two;
            add_mag(zxy.l, zxy.l, two.l, L);
            two.sign = zxy.sign;
            sub_s(zx2, zy2, sub, L);
            add_s(sub, cx, zx, L);
            add_s(two, cy, zy, L);
        } else if (type == AP_BURNING_SHIP || type == AP_JULIA_BURNING_SHIP) {
            // zy = 2 |zx| |zy| - cy; zx = (zx^2 - zy^2) + cx
            mul_mag(zx.l, zy.l, zxy.l, L);
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
                mul_s(zr, zx, t0, L);
                mul_s(zi, zy, t1, L);
                Num<P> re;
                sub_s(t0, t1, re, L);
                mul_s(zr, zy, t0, L);
                mul_s(zi, zx, t1, L);
                add_s(t0, t1, zi, L);
                copy(zr, re);
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
            Num<P> kz;
            mul_s(kr, zx, t0, L);
            mul_s(ki, zy, t1, L);
            sub_s(t0, t1, kz, L);   // kz_r
            add_s(zr, kz, t0, L);   // zn_r + kz_r
            mul_s(kr, zy, t1, L);
            mul_s(ki, zx, kz, L);
            add_s(t1, kz, zr, L);   // kz_i (zr is free)
            add_s(zi, zr, t1, L);   // zn_i + kz_i
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
