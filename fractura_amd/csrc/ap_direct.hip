// Direct escape-time engines: plain FP64 and double-double.
//
// FP64 restates cpu_compute_fp64_sample (math_core/cpu_kernels.py:227-266)
// operation by operation; with contraction off the counts are bit-identical
// to the CPU restatement.  DD restates cpu_compute_dd_sample
// (math_core/cpu_kernels.py:270-369) with the fma exact product of the GPU
// twin (math_core/cuda_kernels.py:51-55).
#include "ap_pass.hpp"
#include "ap_launch.hpp"

namespace ap {

__device__ inline void inv_ln_pair(const ap_view& v, double& ref, double& exact) {
    ref = exact = v.inv_ln2;
    if (v.fractal_type == AP_GENERAL_MANDELBROT && v.gen_n > 1) {
        exact = 1.0 / log((double)v.gen_n);
        ref = (double)(float)exact;  // the reference rounds 1/ln n to float32
    }
}

// Escape test every iteration; the update is the reference's, written so the
// compiler sees exactly its multiplies and adds.
template <int T>
__device__ inline Sample fp64_sample(const ap_view& v, double x, double y) {
    if (T == AP_MANDELBROT && in_main_components(x, y)) return interior(-1, 0);
    double zx, zy, cx, cy;
    if (T == AP_JULIA_MANDELBROT || T == AP_JULIA_BURNING_SHIP) {
        zx = x; zy = y; cx = v.julia_cx; cy = v.julia_cy;
    } else {
        zx = 0.0; zy = 0.0; cx = x; cy = y;
    }
    double inv_ref, inv_exact;
    inv_ln_pair(v, inv_ref, inv_exact);
    const int max_iter = v.max_iter;
    const double kr = v.gen_kr, ki = v.gen_ki;
    const int n = v.gen_n;
    // One exit for both ends (escape at k < max_iter, or k = max_iter): the
    // escape is re-tested after the loop from the same z, so the iterations
    // carry nothing for the exit paths.  At k = max_iter |z|^2 is formed and
    // not tested, as the reference returns there without testing it.
    int k = 0;
    for (;; ++k) {
        double zx2 = zx * zx;
        double zy2 = zy * zy;
        double mag2 = zx2 + zy2;
        if (k >= max_iter || mag2 > kBailout2) break;
        if (T == AP_MANDELBROT || T == AP_JULIA_MANDELBROT) {
            zy = 2.0 * zx * zy + cy;
            zx = zx2 - zy2 + cx;
        } else if (T == AP_BURNING_SHIP || T == AP_JULIA_BURNING_SHIP) {
            zy = 2.0 * fabs(zx) * fabs(zy) - cy;
            zx = zx2 - zy2 + cx;
        } else {
            double pr = zx, pi = zy;
            for (int e = 0; e < n - 1; ++e) {
                double t = pr * zx - pi * zy;
                pi = pr * zy + pi * zx;
                pr = t;
            }
            double kzr = kr * zx - ki * zy;
            double kzi = kr * zy + ki * zx;
            zx = pr + kzr + cx;
            zy = pi + kzi + cy;
        }
    }
    if (k < max_iter) return escaped(k, zx * zx + zy * zy, inv_ref, inv_exact, v.ln_bailout_log);
    return interior(max_iter, max_iter);
}

template <int T>
struct EngFP64 {
    struct Data {};
    static constexpr bool kIterInSubsamples = false;
    __device__ static Sample sample(const ap_view& v, const Data&, double si, double sj) {
        double x = v.x_min + si * v.dx;
        double y = v.y_max - sj * v.dy;
        return fp64_sample<T>(v, x, y);
    }
};

template <int T>
__device__ inline Sample dd_sample(const ap_view& v, dd cx_in, dd cy_in) {
    if (T == AP_MANDELBROT && in_main_components(cx_in.hi, cy_in.hi)) return interior(-1, 0);
    dd zx, zy, cx, cy;
    if (T == AP_JULIA_MANDELBROT || T == AP_JULIA_BURNING_SHIP) {
        zx = cx_in; zy = cy_in;
        cx = dd{v.julia_cx, v.julia_cx_lo};
        cy = dd{v.julia_cy, v.julia_cy_lo};
    } else {
        zx = dd{0.0, 0.0}; zy = dd{0.0, 0.0};
        cx = cx_in; cy = cy_in;
    }
    double inv_ref, inv_exact;
    inv_ln_pair(v, inv_ref, inv_exact);
    const int max_iter = v.max_iter;
    const int n = v.gen_n;
    // One exit for both ends, as in the FP64 loop (the escape test is on the
    // high words only, math_core/cpu_kernels.py:292-295).
    int k = 0;
    double mag2;
    for (;; ++k) {
        dd zx2 = dd_sqr(zx);
        dd zy2 = dd_sqr(zy);
        mag2 = zx2.hi + zy2.hi;
        if (k >= max_iter || mag2 > kBailout2) break;
        if (T == AP_MANDELBROT || T == AP_JULIA_MANDELBROT) {
            dd sub = dd_sub(zx2, zy2);
            dd zxy = dd_mul(zx, zy);
            zx = dd_add(sub, cx);
            zy = dd_add(dd{zxy.hi * 2.0, zxy.lo * 2.0}, cy);
        } else if (T == AP_BURNING_SHIP || T == AP_JULIA_BURNING_SHIP) {
            dd sub = dd_sub(zx2, zy2);
            dd zxy = dd_mul(dd_abs(zx), dd_abs(zy));
            zx = dd_add(sub, cx);
            zy = dd_sub(dd{zxy.hi * 2.0, zxy.lo * 2.0}, cy);
        } else {
            dd pr = zx, pi = zy;
            for (int e = 0; e < n - 1; ++e) {
                dd re = dd_sub(dd_mul(pr, zx), dd_mul(pi, zy));
                dd im = dd_add(dd_mul(pr, zy), dd_mul(pi, zx));
                pr = re;
                pi = im;
            }
            dd kzr = dd_sub(dd_mul_f64(zx, v.gen_kr), dd_mul_f64(zy, v.gen_ki));
            dd kzi = dd_add(dd_mul_f64(zy, v.gen_kr), dd_mul_f64(zx, v.gen_ki));
            zx = dd_add(dd_add(pr, kzr), cx);
            zy = dd_add(dd_add(pi, kzi), cy);
        }
    }
    if (k < max_iter) return escaped(k, mag2, inv_ref, inv_exact, v.ln_bailout_log);
    return interior(max_iter, max_iter);
}

template <int T>
struct EngDD {
    struct Data {};
    static constexpr bool kIterInSubsamples = false;
    __device__ static Sample sample(const ap_view& v, const Data&, double si, double sj) {
        // cx = x_min + dx*sub_i, cy = y_max - dy*sub_j in double-double
        // (math_core/cpu_kernels.py:1230-1234)
        dd cx = dd_add(dd{v.x_min, v.x_min_lo}, dd_mul_f64(dd{v.dx, v.dx_lo}, si));
        dd cy = dd_sub(dd{v.y_max, v.y_max_lo}, dd_mul_f64(dd{v.dy, v.dy_lo}, sj));
        return dd_sample<T>(v, cx, cy);
    }
};

template <template <int> class E>
static hipError_t dispatch_type(const ap_view& v, const ap_outputs& o, hipStream_t s) {
    switch (v.fractal_type) {
    case 0: return launch_pass<E<0>>(v, o, {}, s);
    case 1: return launch_pass<E<1>>(v, o, {}, s);
    case 2: return launch_pass<E<2>>(v, o, {}, s);
    case 3: return launch_pass<E<3>>(v, o, {}, s);
    case 4: return launch_pass<E<4>>(v, o, {}, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_fp64(const ap_view& v, const ap_outputs& o, hipStream_t s) {
    return dispatch_type<EngFP64>(v, o, s);
}

hipError_t launch_dd(const ap_view& v, const ap_outputs& o, hipStream_t s) {
    return dispatch_type<EngDD>(v, o, s);
}

}  // namespace ap
