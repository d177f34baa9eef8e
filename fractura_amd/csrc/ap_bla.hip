// Chained-BLA table, built on the device as a level-by-level dyadic merge.
//
// Level 0 and the merge rule restate _jit_build_bla_table
// (math_core/high_precision.py:1100-1238) operation by operation in FP64
// (contraction off), so A, B and r^2 are bit-identical to the CPU
// restatement.  Level d only reads level d-1, so each level is one launch
// over all k; the reference builds the same table serially on the host and
// copies 16 x ref_len x 40 bytes over PCIe (math_core/engine.py:416-423).
#include "ap_bla.hpp"
#include "ap_launch.hpp"

namespace ap {

struct BlaMut {
    double4* coef;  // [levels][ref_len]
    double* r2m;    // [levels][ref_len]
    double* r2;     // [ref_len][levels]
    double* r1;     // [ref_len][levels]
    Coef32* c32;    // [ref_len][levels]
    int levels;
    int ref_len;
};

static BlaMut bla_mut(void* table, int ref_len, int levels) {
    BlaLayout L = bla_layout(ref_len, levels);
    char* p = (char*)table;
    BlaMut b;
    b.coef = (double4*)(p + L.coef);
    b.r2m = (double*)(p + L.r2m);
    b.r2 = (double*)(p + L.r2);
    b.r1 = (double*)(p + L.r1);
    b.c32 = (Coef32*)(p + L.c32);
    b.levels = levels;
    b.ref_len = ref_len;
    return b;
}

__global__ __launch_bounds__(256) void bla_level0_kernel(BlaMut t, const double* __restrict__ zr,
                                                         const double* __restrict__ zi, int type,
                                                         int n, double kr, double ki, double tol) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= t.ref_len) return;
    if (k >= t.ref_len - 1) {
        t.coef[k] = make_double4(0.0, 0.0, 0.0, 0.0);
        t.r2m[k] = 0.0;
        return;
    }
    double rx = zr[k], ry = zi[k];
    double r_mag2 = rx * rx + ry * ry;
    double ar, ai, r_sq;
    if (type == AP_GENERAL_MANDELBROT) {
        double pr = 1.0, pi = 0.0;
        for (int j = 0; j < n - 1; ++j) {
            double t2 = pr * rx - pi * ry;
            pi = pr * ry + pi * rx;
            pr = t2;
        }
        ar = (double)n * pr + kr;
        ai = (double)n * pi + ki;
        double a_mag = sqrt(ar * ar + ai * ai);
        double tol_a = tol * a_mag;
        if (!(tol_a > 0.0) || isinf(tol_a)) {
            r_sq = 0.0;
        } else {
            double r_mag = sqrt(r_mag2);
            double r_pow = n == 2 ? 1.0 : (n == 3 ? r_mag : (n == 4 ? r_mag2 : pow(r_mag, (double)(n - 2))));
            double b = 0.25 * (double)(n * (n - 1)) * r_pow;
            double denom = b + sqrt(b * b + tol_a);
            double x = denom > 1e-300 ? tol_a / denom : 0.0;
            if (n > 3) {
                double x_max = pow(tol_a, 1.0 / (double)(n - 1));
                if (x > x_max) x = x_max;
            }
            if (x > 0.5) x = 0.5;
            r_sq = x * x;
            if (k == 0 && r_mag2 < 1e-14) r_sq = 0.0;
        }
    } else {
        ar = 2.0 * rx;
        ai = 2.0 * ry;
        double tol_sq_4 = 4.0 * tol * tol;
        r_sq = tol_sq_4 * r_mag2;
        if (r_sq > 0.25) r_sq = 0.25;
        if (k == 0 && r_mag2 < 1e-14) r_sq = 0.0;
    }
    t.coef[k] = make_double4(ar, ai, 1.0, 0.0);
    t.r2m[k] = r_sq;
}

// Level d from level d-1, both level-major: lanes k and k+1 read adjacent
// words, so every load and store is coalesced.
__global__ __launch_bounds__(256) void bla_merge_kernel(BlaMut t, int d) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= t.ref_len) return;
    const size_t N = (size_t)t.ref_len;
    const int step = 1 << (d - 1);
    double4 out = make_double4(0.0, 0.0, 0.0, 0.0);
    double r_out = 0.0;
    if (k < t.ref_len - 2 * step) {
        size_t e1 = (size_t)(d - 1) * N + k;
        size_t e2 = e1 + step;
        double r1_sq = t.r2m[e1], r2_sq = t.r2m[e2];
        if (r1_sq > 0.0 && r2_sq > 0.0) {
            double4 c1 = t.coef[e1], c2 = t.coef[e2];
            double ar = c2.x * c1.x - c2.y * c1.y;
            double ai = c2.x * c1.y + c2.y * c1.x;
            double br = (c2.x * c1.z - c2.y * c1.w) + c2.z;
            double bi = (c2.x * c1.w + c2.y * c1.z) + c2.w;
            double a_mag2 = ar * ar + ai * ai;
            double b_mag2 = br * br + bi * bi;
            bool bad = isnan(a_mag2) || isinf(a_mag2) || isnan(b_mag2) || isinf(b_mag2) ||
                       a_mag2 > 1e150 || b_mag2 > 1e150;
            if (!bad) {
                out = make_double4(ar, ai, br, bi);
                double a1_mag2 = c1.x * c1.x + c1.y * c1.y;
                double cand;
                if (a1_mag2 > 1e-300) {
                    double r2s = (0.25 * r2_sq) / a1_mag2;
                    cand = r1_sq < r2s ? r1_sq : r2s;
                } else {
                    cand = r1_sq;
                }
                if (isnan(cand) || isinf(cand) || cand <= 0.0)
                    r_out = 0.0;
                else
                    r_out = cand > 0.25 ? 0.25 : cand;
            }
        }
    }
    size_t e = (size_t)d * N + k;
    t.coef[e] = out;
    t.r2m[e] = r_out;
}

// Derived regions.  A block takes 16 consecutive k for every level: reads the
// level-major build state coalesced, transposes it through LDS and writes the
// per-k rows (r^2 and sqrt(r^2) for the FP64 engines, the keyed FP32 entry
// for the FP32-mantissa engine).
constexpr int kDeriveK = 16;

__device__ inline Coef32 make_coef32(double4 c, double r2) {
    Coef32 q;
    double am = fmax(fabs(c.x), fabs(c.y));
    double bm = fmax(fabs(c.z), fabs(c.w));
    q.ae = am > 0.0 ? ilogb(am) : kZeroExp;
    q.be = bm > 0.0 ? ilogb(bm) : kZeroExp;
    q.ar = am > 0.0 ? (float)ldexp(c.x, -q.ae) : 0.0f;
    q.ai = am > 0.0 ? (float)ldexp(c.y, -q.ae) : 0.0f;
    q.br = bm > 0.0 ? (float)ldexp(c.z, -q.be) : 0.0f;
    q.bi = bm > 0.0 ? (float)ldexp(c.w, -q.be) : 0.0f;
    q.key = r2_key(r2);
    q.pad = 0;
    return q;
}

__global__ __launch_bounds__(256) void bla_derive_kernel(BlaMut t) {
    __shared__ double rt[32][kDeriveK + 1];
    __shared__ Coef32 ct[32][kDeriveK + 1];
    const int L = t.levels;
    const size_t N = (size_t)t.ref_len;
    const int k0 = blockIdx.x * kDeriveK;
    for (int i = threadIdx.x; i < L * kDeriveK; i += 256) {
        int d = i / kDeriveK, kl = i % kDeriveK;
        int k = k0 + kl;
        if (k < t.ref_len) {
            size_t e = (size_t)d * N + k;
            double r2 = t.r2m[e];
            rt[d][kl] = r2;
            ct[d][kl] = make_coef32(t.coef[e], r2);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < L * kDeriveK; i += 256) {
        int kl = i / L, d = i % L;
        int k = k0 + kl;
        if (k < t.ref_len) {
            double r2 = rt[d][kl];
            size_t row = (size_t)k * L + d;
            t.r2[row] = r2;
            t.r1[row] = sqrt(r2);
            t.c32[row] = ct[d][kl];
        }
    }
}

__global__ __launch_bounds__(256) void bla_export_kernel(BlaView t, const double* __restrict__ r2m,
                                                         double* __restrict__ out) {
    size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;  // level-major index
    size_t n = (size_t)t.ref_len * t.levels;
    if (e >= n) return;
    int d = (int)(e / t.ref_len), k = (int)(e % t.ref_len);
    double4 c = t.coef[e];
    double* o = out + ((size_t)d * t.ref_len + k) * 5;
    o[0] = c.x;
    o[1] = c.y;
    o[2] = c.z;
    o[3] = c.w;
    o[4] = r2m[e];
}

size_t bla_bytes(int ref_len, int levels) { return bla_layout(ref_len, levels).total; }

hipError_t launch_bla_build(const ap_orbit& orb, int type, int n, double kr, double ki,
                            double tol, int levels, void* table, hipStream_t s) {
    const int ref_len = orb.len;
    BlaMut t = bla_mut(table, ref_len, levels);
    int blocks = (ref_len + 255) / 256;
    hipLaunchKernelGGL(bla_level0_kernel, dim3(blocks), dim3(256), 0, s, t, orb.re, orb.im, type,
                       n, kr, ki, tol);
    for (int d = 1; d < levels; ++d)
        hipLaunchKernelGGL(bla_merge_kernel, dim3(blocks), dim3(256), 0, s, t, d);
    hipLaunchKernelGGL(bla_derive_kernel, dim3((ref_len + kDeriveK - 1) / kDeriveK), dim3(256), 0,
                       s, t);
    return hipGetLastError();
}

hipError_t launch_bla_export(const void* table, int ref_len, int levels, double* out,
                             hipStream_t s) {
    BlaView v = bla_view(table, ref_len, levels);
    BlaLayout L = bla_layout(ref_len, levels);
    const double* r2m = (const double*)((const char*)table + L.r2m);
    size_t ne = bla_entries(ref_len, levels);
    hipLaunchKernelGGL(bla_export_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, v,
                       r2m, out);
    return hipGetLastError();
}

}  // namespace ap
