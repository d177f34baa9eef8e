// Perturbation engines over a host-computed reference orbit Z_k.
//
//  PERT64  FP64 delta (w >= 1e-300): restates cpu_compute_perturbation_sample
//          (math_core/cpu_kernels.py:373-602) with the chained-BLA lookup.
//  FEXP64  FloatExp, FP64 mantissa + int32 exponent (w < 1e-300): restates
//          cpu_compute_perturbation_sample_floatexp (:605-1021).
//  FEXP32  FloatExp with an FP32 mantissa (no reference counterpart).  FP64
//          stays where stability needs it: the near-bailout test and smooth
//          value, Stage 2, and the BLA table build.
//
// The FP64 engines follow the CPU restatement operation by operation
// (contraction off).  Two lookups are reorganised without changing results:
// the BLA level search is a binary search over radii that never grow with
// the level (same answer as the reference's top-down scan), and sqrt(r^2)
// is read from the table instead of being recomputed.
#include "ap_bla.hpp"
#include "ap_launch.hpp"
#include "ap_pass.hpp"

namespace ap {

constexpr double kInv2p16 = 1.52587890625e-05;  // 2^-16

// pow2_f64 of the reference (math_core/cpu_kernels.py:161-167): 2^e with
// e >= 1023 -> DBL_MAX and e <= -1022 -> 0.
__device__ inline double pow2_f64(int e) {
    if (e >= 1023) return 1.7976931348623157e+308;
    if (e <= -1022) return 0.0;
    return ldexp(1.0, e);
}

// Python max(a, b) (returns a unless b > a; matters only for NaN).
__device__ inline double pymax(double a, double b) { return b > a ? b : a; }

// int(log(x) / ln 2) for x > 0, computed from the binary exponent: truncation
// toward zero of log2 x.  Equal to the reference's log-based value except
// when x lies within rounding distance of a power of two.
__device__ inline int trunc_log2(double x) {
    int e;
    double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    if (x >= 1.0) return e - 1;
    return m == 0.5 ? e - 1 : e;
}

// Pauldelbrot test in rescaled exponent space
// (check_glitch_cancellation_floatexp, math_core/cpu_kernels.py:170-201).
__device__ inline int glitch_floatexp(double rx, double ry, double dzx, double dzy, int eff) {
    double r_max = pymax(fabs(rx), fabs(ry));
    if (r_max < 1e-4) return 0;
    double dz_max = pymax(fabs(dzx), fabs(dzy));
    if (dz_max <= 0.0) return 0;
    int r_exp = trunc_log2(r_max);
    int dz_exp = trunc_log2(dz_max);
    int diff = r_exp - (eff + dz_exp);
    if (diff > 10 || diff < -10) return 0;
    double sr = pow2_f64(-r_exp), sd = pow2_f64(eff - r_exp);
    double rxs = rx * sr, rys = ry * sr;
    double zx = rxs + dzx * sd, zy = rys + dzy * sd;
    double m2 = zx * zx + zy * zy;
    double rm2 = rxs * rxs + rys * rys;
    return m2 < 1e-6 * rm2 ? 1 : 0;
}

// Keep max(|dzx|, |dzy|) inside [2^-16, 2^16] in steps of 2^16.
__device__ inline void renorm64(double& dzx, double& dzy, int& z_exp) {
    double mag = pymax(fabs(dzx), fabs(dzy));
    while (mag > 65536.0 && isfinite(mag)) {
        dzx *= kInv2p16;
        dzy *= kInv2p16;
        z_exp += 16;
        mag = pymax(fabs(dzx), fabs(dzy));
    }
    while (mag < kInv2p16 && mag > 0.0 && isfinite(mag)) {
        dzx *= 65536.0;
        dzy *= 65536.0;
        z_exp -= 16;
        mag = pymax(fabs(dzx), fabs(dzy));
    }
}

__device__ inline Sample nan_exit(int k) {
    Sample r;
    r.count = k;
    r.state = kNanExit;
    r.mag2 = 0.0;
    r.glitch = 1;
    r.plain = 0;
    r.bla = 0;
    return r;
}

// Step 0 of a Mandelbrot-type orbit (types 0 and 4, Z_0 = 0, dz_0 = 0) gives
// dz_1 = dc exactly: every engine's step formula reduces to it, and neither
// the escape test (|Z_0|^2 = 0) nor the glitch test (k > 15) fires at k = 0.
// The engines start such pixels at k = 1; the skipped step counts as plain.
template <int T>
__device__ inline bool mandel_start(const ap_view& v, const double* zr, const double* zi) {
    return (T == 0 || T == 4) && v.ref_len >= 1 && v.max_iter >= 1 && zr[0] == 0.0 && zi[0] == 0.0;
}

struct PertConsts {
    double inv_ref, inv_exact;
};

__device__ inline PertConsts pert_consts(const ap_view& v) {
    PertConsts c;
    c.inv_ref = c.inv_exact = v.inv_ln2;
    if (v.fractal_type == AP_GENERAL_MANDELBROT && v.gen_n > 1) {
        c.inv_exact = 1.0 / log((double)v.gen_n);
        c.inv_ref = (double)(float)c.inv_exact;
    }
    return c;
}

// Stage 2 of every perturbation engine: the reference escaped (or ran out),
// so the pixel continues as a direct FP64 iteration of Z_ref_len + dz.
template <int T>
__device__ inline Sample stage2(const ap_view& v, const PertConsts& pc, double zx, double zy,
                                double cx, double cy, int k, int glitch, int plain, int bla) {
    const int max_iter = v.max_iter;
    const int n = v.gen_n;
    while (k < max_iter) {
        double zx2 = zx * zx, zy2 = zy * zy;
        double mag2 = zx2 + zy2;
        if (isnan(mag2) || isinf(mag2)) {
            Sample r = nan_exit(k);
            r.plain = plain;
            r.bla = bla;
            return r;
        }
        if (mag2 > kBailout2) {
            Sample r = escaped(k, mag2, pc.inv_ref, pc.inv_exact, v.ln_bailout_log);
            r.glitch = glitch;
            r.plain = plain;
            r.bla = bla;
            return r;
        }
        if (T == 0 || T == 2) {
            zy = 2.0 * zx * zy + cy;
            zx = zx2 - zy2 + cx;
        } else if (T == 1 || T == 3) {
            zy = 2.0 * fabs(zx) * fabs(zy) - cy;
            zx = zx2 - zy2 + cx;
        } else {
            double pr = zx, pi = zy;
            for (int e = 0; e < n - 1; ++e) {
                double t = pr * zx - pi * zy;
                pi = pr * zy + pi * zx;
                pr = t;
            }
            double kzr = v.gen_kr * zx - v.gen_ki * zy;
            double kzi = v.gen_kr * zy + v.gen_ki * zx;
            zx = pr + kzr + cx;
            zy = pi + kzi + cy;
        }
        ++k;
        ++plain;
    }
    Sample r = interior(max_iter, max_iter);
    r.glitch = glitch;
    r.plain = plain;
    r.bla = bla;
    return r;
}

// =================================================================== PERT64
struct PertData {
    const double* zr;
    const double* zi;
    BlaView bla;
};

template <int T>
__device__ Sample pert64_sample(const ap_view& v, const PertData& D, double si, double sj) {
    const bool julia = (T == 2 || T == 3);
    const double dcy = (v.ref_py - sj) * v.dy;
    const double dcx = (si - v.ref_px) * v.dx;
    double dzx = julia ? dcx : 0.0, dzy = julia ? dcy : 0.0;
    const double dpx = julia ? 0.0 : dcx, dpy = julia ? 0.0 : dcy;
    const int ref_len = v.ref_len;
    const bool strict = v.strict_glitch > 0;
    const bool has_bla = D.bla.levels > 1 && (T == 0 || T == 2 || T == 4);
    const PertConsts pc = pert_consts(v);
    int k = 0, glitch = 0, plain = 0, nbla = 0;
    const int n = v.gen_n;
    const double kr = v.gen_kr, ki = v.gen_ki;
    if (mandel_start<T>(v, D.zr, D.zi)) {
        dzx = dpx;
        dzy = dpy;
        k = plain = 1;
    }

    while (k < ref_len) {
        const double rx = D.zr[k], ry = D.zi[k];
        const double zx = rx + dzx, zy = ry + dzy;
        const double mag2 = zx * zx + zy * zy;
        if (mag2 > kBailout2) {
            if (strict && k > 15 && (rx * rx + ry * ry) < 4.0) glitch = 1;
            Sample r = escaped(k, mag2, pc.inv_ref, pc.inv_exact, v.ln_bailout_log);
            r.glitch = glitch;
            r.plain = plain;
            r.bla = nbla;
            return r;
        }
        if (strict && k > 15) {
            double r_mag2 = rx * rx + ry * ry;
            if (mag2 < 1e-6 * r_mag2 && r_mag2 > 1e-8) glitch = 1;
        }
        if (has_bla && k < ref_len - 1) {
            const double dz_mag2 = dzx * dzx + dzy * dzy;
            int max_d = D.bla.levels - 1;
            if (dz_mag2 == 0.0 && (dzx != 0.0 || dzy != 0.0)) max_d = max_d < 8 ? max_d : 8;
            const int hi = min(max_d, level_limit_by_length(k, ref_len));
            // Level 0 and the top level are read together; the search runs
            // only when the top level is not valid.
            // The top level's coefficients are fetched with the radii: in deep
            // views the top level is the one taken, and the leap then costs a
            // single round trip to the cache.
            const unsigned itop = D.bla.at(k, hi);
            const double r0 = entry(D.bla.r2, (unsigned)k), rtop = entry(D.bla.r2, itop);
            const double4 ctop = entry(D.bla.coef, itop);
            if (dz_mag2 < r0) {
                int best = dz_mag2 < rtop ? hi
                                          : highest_valid_level(hi - 1, [&](int d) {
                                                return dz_mag2 < entry(D.bla.r2, D.bla.at(k, d));
                                            });
                // A value copy, not a conditional lvalue: selecting between a
                // local and a table reference would put the local in scratch.
                double4 c = ctop;
                if (best != hi) c = entry(D.bla.coef, D.bla.at(k, best));
                double ndx = (c.x * dzx - c.y * dzy) + (c.z * dpx - c.w * dpy);
                dzy = (c.x * dzy + c.y * dzx) + (c.z * dpy + c.w * dpx);
                dzx = ndx;
                ++nbla;
                if (isnan(dzx) || isnan(dzy) || isinf(dzx) || isinf(dzy)) {
                    Sample r = nan_exit(k);
                    r.plain = plain;
                    r.bla = nbla;
                    return r;
                }
                k += 1 << best;
                continue;
            }
        }
        if (T == 0 || T == 2) {
            double ndx = (2.0 * rx + dzx) * dzx - (2.0 * ry + dzy) * dzy + dpx;
            dzy = 2.0 * ((rx + dzx) * dzy + ry * dzx) + dpy;
            dzx = ndx;
        } else if (T == 1 || T == 3) {
            double ndx = (2.0 * rx + dzx) * dzx - (2.0 * ry + dzy) * dzy + dpx;
            double sxs = rx >= 0.0 ? 1.0 : -1.0;
            double sys = ry >= 0.0 ? 1.0 : -1.0;
            double ddx = (rx + dzx) * sxs >= 0.0 ? sxs * dzx : -sxs * (2.0 * rx + dzx);
            double ddy = (ry + dzy) * sys >= 0.0 ? sys * dzy : -sys * (2.0 * ry + dzy);
            dzy = 2.0 * (fabs(rx) * ddy + fabs(ry) * ddx + ddx * ddy) - dpy;
            dzx = ndx;
        } else {
            // (Z+dz)^n - Z^n by the binomial recurrence of the reference.
            double pr = dzx, pi = dzy, qr = rx, qi = ry;
            for (int e = 0; e < n - 1; ++e) {
                double t1r = pr * rx - pi * ry, t1i = pr * ry + pi * rx;
                double sr = pr + qr, sim = pi + qi;
                double t2r = sr * dzx - sim * dzy, t2i = sr * dzy + sim * dzx;
                pr = t1r + t2r;
                pi = t1i + t2i;
                double nq = qr * rx - qi * ry;
                qi = qr * ry + qi * rx;
                qr = nq;
            }
            double kdr = kr * dzx - ki * dzy, kdi = kr * dzy + ki * dzx;
            dzx = pr + kdr + dpx;
            dzy = pi + kdi + dpy;
        }
        ++k;
        ++plain;
    }
    const double zx = D.zr[ref_len] + dzx, zy = D.zi[ref_len] + dzy;
    const double cx = julia ? v.julia_cx : v.ref_cx + dcx;
    const double cy = julia ? v.julia_cy : v.ref_cy + dcy;
    if (ref_len < v.max_iter) glitch = 1;
    return stage2<T>(v, pc, zx, zy, cx, cy, k, glitch, plain, nbla);
}

template <int T>
struct EngPert64 {
    using Data = PertData;
    static constexpr bool kIterInSubsamples = true;
    __device__ static Sample sample(const ap_view& v, const Data& d, double si, double sj) {
        return pert64_sample<T>(v, d, si, sj);
    }
};

// =================================================================== FEXP64
// One BLA leap of the FP64 FloatExp engine from iteration k, if level 0 is
// valid (math_core/cpu_kernels.py:677-733): returns the number of iterations
// advanced (0 = no leap).  The radius test is the reference's: |dz|^2 < r^2
// when the shift is >= -510, else |dz_m| 2^max(-1020, shift) < r with r read
// from the table.  Level 0 and the top level are tested with one round trip,
// the top level's coefficients fetched alongside.
__device__ inline int bla_leap64(const PertData& D, double& dzx, double& dzy, int& z_exp,
                                 double dpx, double dpy, int k, int E, int ref_len) {
    const int shift = z_exp - E;
    const int hi = min(D.bla.levels - 1, level_limit_by_length(k, ref_len));
    const unsigned i0 = (unsigned)k, itop = D.bla.at(k, hi);
    bool can0, cantop;
    double dz_mag2 = 0.0, tiny = 0.0;
    if (shift >= -510) {
        double sc = pow2_f64(shift);
        dz_mag2 = (dzx * dzx + dzy * dzy) * (sc * sc);
        const double r0 = entry(D.bla.r2, i0), rtop = entry(D.bla.r2, itop);
        can0 = dz_mag2 < r0;
        cantop = dz_mag2 < rtop;
    } else {
        double dz_mag = sqrt(dzx * dzx + dzy * dzy);
        tiny = dz_mag * pow2_f64(shift > -1020 ? shift : -1020);
        const double r0 = entry(D.bla.r2, i0), rtop = entry(D.bla.r2, itop);
        const double s0 = entry(D.bla.r1, i0), stop = entry(D.bla.r1, itop);
        can0 = r0 > 0.0 && tiny < s0;
        cantop = rtop > 0.0 && tiny < stop;
    }
    const double4 ctop = entry(D.bla.coef, itop);
    if (!can0) return 0;
    int best;
    if (cantop)
        best = hi;
    else if (shift >= -510)
        best = highest_valid_level(hi - 1, [&](int d) { return dz_mag2 < entry(D.bla.r2, D.bla.at(k, d)); });
    else
        best = highest_valid_level(hi - 1, [&](int d) {
            const unsigned i = D.bla.at(k, d);
            return entry(D.bla.r2, i) > 0.0 && tiny < entry(D.bla.r1, i);
        });
    double4 c = ctop;  // value copy (a conditional lvalue would go to scratch)
    if (best != hi) c = entry(D.bla.coef, D.bla.at(k, best));
    double scale_c = (z_exp >= -1020 && z_exp <= 1020) ? pow2_f64(-z_exp) : 0.0;
    double dxs = dpx * scale_c, dys = dpy * scale_c;
    double ndx = (c.x * dzx - c.y * dzy) + (c.z * dxs - c.w * dys);
    dzy = (c.x * dzy + c.y * dzx) + (c.z * dys + c.w * dxs);
    dzx = ndx;
    return 1 << best;
}

template <int T>
__device__ Sample fexp64_sample(const ap_view& v, const PertData& D, double si, double sj) {
    const bool julia = (T == 2 || T == 3);
    const double dcy = (v.ref_py - sj) * v.dy;  // normalised: dc * 2^E
    const double dcx = (si - v.ref_px) * v.dx;
    double dzx = julia ? dcx : 0.0, dzy = julia ? dcy : 0.0;
    const double dpx = julia ? 0.0 : dcx, dpy = julia ? 0.0 : dcy;
    int z_exp = 0;
    const int E = v.e_scale;
    const int ref_len = v.ref_len;
    const bool strict = v.strict_glitch > 0;
    const bool has_bla = D.bla.levels > 1 && (T == 0 || T == 2 || T == 4);
    const PertConsts pc = pert_consts(v);
    int k = 0, glitch = 0, plain = 0, nbla = 0;
    const int n = v.gen_n;
    const double kr = v.gen_kr, ki = v.gen_ki;
    if (mandel_start<T>(v, D.zr, D.zi)) {
        dzx = dpx;
        dzy = dpy;
        renorm64(dzx, dzy, z_exp);
        k = plain = 1;
    }

    while (k < ref_len) {
        // Deep regime (eff < -50): the reference's escape test (eff >= -50) and
        // the Pauldelbrot test (returns 0 below eff = -40) are both skipped, so
        // consecutive BLA leaps run in a loop of their own.
        bool no_leap = false;  // level 0 already failed at this k and state
        if ((T == 0 || T == 2 || T == 4) && has_bla) {
            while (z_exp - E < -50 && k < ref_len - 1) {
                int step = bla_leap64(D, dzx, dzy, z_exp, dpx, dpy, k, E, ref_len);
                if (step == 0) {
                    no_leap = true;
                    break;
                }
                ++nbla;
                if (isnan(dzx) || isnan(dzy) || isinf(dzx) || isinf(dzy)) {
                    Sample r = nan_exit(k);
                    r.plain = plain;
                    r.bla = nbla;
                    return r;
                }
                renorm64(dzx, dzy, z_exp);
                k += step;
            }
            if (k >= ref_len) break;
        }
        const double rx = D.zr[k], ry = D.zi[k];
        const int eff = z_exp - E;
        if (eff >= -50) {
            double sc = pow2_f64(eff);
            double zx = rx + dzx * sc, zy = ry + dzy * sc;
            double mag2 = zx * zx + zy * zy;
            if (mag2 > kBailout2) {
                Sample r = escaped(k, mag2, pc.inv_ref, pc.inv_exact, v.ln_bailout_log);
                r.glitch = glitch;
                r.plain = plain;
                r.bla = nbla;
                return r;
            }
        }
        // With max(|dz_m|) <= 2^16 and r_max >= 1e-4 the test can only fire
        // when eff + log2|dz_m| >= -24: below eff = -40 it returns 0.
        if (strict && k > 15 && eff >= -40) {
            if (glitch_floatexp(rx, ry, dzx, dzy, eff)) glitch = 1;
        }
        if (T == 0 || T == 2 || T == 4) {
            if (has_bla && !no_leap && k < ref_len - 1) {
                int step = bla_leap64(D, dzx, dzy, z_exp, dpx, dpy, k, E, ref_len);
                if (step > 0) {
                    ++nbla;
                    if (isnan(dzx) || isnan(dzy) || isinf(dzx) || isinf(dzy)) {
                        Sample r = nan_exit(k);
                        r.plain = plain;
                        r.bla = nbla;
                        return r;
                    }
                    renorm64(dzx, dzy, z_exp);
                    k += step;
                    continue;
                }
            }
        }
        if (T == 0 || T == 2) {
            double lx = 2.0 * (rx * dzx - ry * dzy);
            double ly = 2.0 * (rx * dzy + ry * dzx);
            const int snl = z_exp - E;
            if (snl >= -1020) {
                double s = pow2_f64(snl);
                lx += (dzx * dzx - dzy * dzy) * s;
                ly += (2.0 * dzx * dzy) * s;
            }
            if (z_exp >= -1020 && z_exp <= 1020) {
                double s = pow2_f64(-z_exp);
                lx += dpx * s;
                ly += dpy * s;
            }
            dzx = lx;
            dzy = ly;
            renorm64(dzx, dzy, z_exp);
        } else if (T == 1 || T == 3) {
            const int snl = z_exp - E;
            double s_nl = snl >= -1020 ? pow2_f64(snl) : 0.0;
            double s_c = (z_exp >= -1020 && z_exp <= 1020) ? pow2_f64(-z_exp) : 0.0;
            double sxs = rx >= 0.0 ? 1.0 : -1.0;
            double sys = ry >= 0.0 ? 1.0 : -1.0;
            double s_esc = eff >= -1020 ? pow2_f64(eff) : 0.0;
            double adx = dzx * s_esc, ady = dzy * s_esc;
            double fx, ex, fy, ey;
            if ((rx + adx) * sxs >= 0.0) {
                fx = sxs;
                ex = 0.0;
            } else {
                fx = -sxs;
                ex = eff <= 1020 ? -sxs * 2.0 * rx * pow2_f64(-eff) : 0.0;
            }
            if ((ry + ady) * sys >= 0.0) {
                fy = sys;
                ey = 0.0;
            } else {
                fy = -sys;
                ey = eff <= 1020 ? -sys * 2.0 * ry * pow2_f64(-eff) : 0.0;
            }
            double lx = (2.0 * rx + dzx * s_nl) * dzx - (2.0 * ry + dzy * s_nl) * dzy + dpx * s_c;
            double dxn = fx * dzx + ex;
            double dyn = fy * dzy + ey;
            double ly = 2.0 * (fabs(rx) * dyn + fabs(ry) * dxn + dxn * dyn * s_nl) - dpy * s_c;
            dzx = lx;
            dzy = ly;
            // The Burning Ship branch of the reference renormalises without a
            // finiteness guard (math_core/cpu_kernels.py:868-878), which never
            // ends on an infinite delta; the guard changes nothing else.
            renorm64(dzx, dzy, z_exp);
        } else {
            double pr = dzx, pi = dzy, qr = rx, qi = ry;
            const int snl = z_exp - E;
            double s_nl = snl >= -1020 ? pow2_f64(snl) : 0.0;
            for (int e = 0; e < n - 1; ++e) {
                double t1r = pr * rx - pi * ry, t1i = pr * ry + pi * rx;
                double t2r = qr * dzx - qi * dzy, t2i = qr * dzy + qi * dzx;
                double t3r = pr * dzx - pi * dzy, t3i = pr * dzy + pi * dzx;
                pr = t1r + t2r + t3r * s_nl;
                pi = t1i + t2i + t3i * s_nl;
                double nq = qr * rx - qi * ry;
                qi = qr * ry + qi * rx;
                qr = nq;
            }
            double kdr = kr * dzx - ki * dzy, kdi = kr * dzy + ki * dzx;
            double s_c = (z_exp >= -1020 && z_exp <= 1020) ? pow2_f64(-z_exp) : 0.0;
            dzx = pr + kdr + dpx * s_c;
            dzy = pi + kdi + dpy * s_c;
            renorm64(dzx, dzy, z_exp);
        }
        ++k;
        ++plain;
    }
    const int eff = z_exp - E;
    const double sfb = eff >= -60 ? pow2_f64(eff) : 0.0;
    const double zx = D.zr[ref_len] + dzx * sfb, zy = D.zi[ref_len] + dzy * sfb;
    const double scfb = E <= 1020 ? pow2_f64(-E) : 0.0;
    const double cx = julia ? v.julia_cx : v.ref_cx + dpx * scfb;
    const double cy = julia ? v.julia_cy : v.ref_cy + dpy * scfb;
    const double mag2 = zx * zx + zy * zy;
    if (mag2 > kBailout2) {
        Sample r = escaped(ref_len, mag2, pc.inv_ref, pc.inv_exact, v.ln_bailout_log);
        r.glitch = glitch;
        r.plain = plain;
        r.bla = nbla;
        return r;
    }
    if (ref_len < v.max_iter) glitch = 1;
    return stage2<T>(v, pc, zx, zy, cx, cy, k, glitch, plain, nbla);
}

template <int T>
struct EngFexp64 {
    using Data = PertData;
    static constexpr bool kIterInSubsamples = true;
    __device__ static Sample sample(const ap_view& v, const Data& d, double si, double sj) {
        return fexp64_sample<T>(v, d, si, sj);
    }
};

// =================================================================== FEXP32
// State: float mantissa pair (dzx, dzy) and int exponent z_exp, actual delta
// = dz_m * 2^(z_exp - E).  Z_k is read as float2 from an FP32 copy of the
// orbit.  A BLA leap leaves max|dz_m| in [1/2, 1); a plain step renormalises
// by the binary exponent only when the mantissa leaves [2^-32, 2^32] (FP32
// has the range to wait).  BLA validity compares the ordered int32 key of
// |dz|^2 with the key of r^2 stored in the entry.
struct Fexp32Data {
    const double* zr;
    const double* zi;
    const float2* zf;
    BlaView bla;
};

// floor(log2 m) of a finite, normal, positive float from its exponent field.
__device__ inline int exponent_of(float m) { return __builtin_amdgcn_frexp_expf(m) - 1; }

__device__ inline void renorm32(float& dzx, float& dzy, int& z_exp) {
    float m = fmaxf(fabsf(dzx), fabsf(dzy));
    if ((m > 4294967296.0f || m < 2.3283064365386963e-10f) && m > 0.0f && isfinite(m)) {
        int e = exponent_of(m);
        dzx = ldexpf(dzx, -e);
        dzy = ldexpf(dzy, -e);
        z_exp += e;
    }
}

// BLA leaps of the FP32 engine.  The highest level a leap from k may take is
// hi = top_level(k); the entry at level hi is fetched together with the key
// test, and in deep views it is the one taken, so a leap costs one round trip
// to the cache.  When it is not valid, leap32_search finds the highest valid
// level below it (-1 when level 0 is not valid either).
//
// dz' = A dz + B dc with dz = m 2^z_exp, A = am 2^ae, B = bm 2^be and dc in
// units of 2^-E (dc_m = dc 2^E, of order one): the A term has exponent
// e1 = z_exp + ae + em, 2^em a bound on |m| read off the key's |m|^2, the B
// term e2 = be.  Both inputs are scaled down to ne = max(e1, e2) before the
// products, which run as FMAs; the result, at exponent ne, is bounded by a
// few units (|am| < 2 sqrt 2, scaled |m| < sqrt 2, plus |bm| |dc_m|), so no
// normalisation step follows.  kNoDc: the B dc term is zero (Julia sets), so
// its exponent takes no part.
__device__ inline int leap32_search(const Fexp32Data& D, int k, int hi, int key) {
    if (!(key < entry(D.bla.c32, (unsigned)k).key)) return -1;
    return highest_valid_level(hi - 1, [&](int m) { return key < entry(D.bla.c32, D.bla.at(k, m)).key; });
}

template <bool kNoDc>
__device__ inline void leap32_terms(const Coef32& c, float q, float dzx, float dzy, int z_exp,
                                    float dpx, float dpy, float& rx, float& ry, int& ne) {
    // |m| < 2^em: q = |m|^2 has biased exponent field b, so |m|^2 < 2^(b - 126)
    // and em = (b >> 1) - 62 bounds it (q = 0 gives a small em).
    const int em = (int)(__float_as_uint(q) >> 24) - 62;
    const int za = z_exp + c.ae;
    if (kNoDc) {
        ne = za + em;
        const float mx = ldexpf(dzx, -em), my = ldexpf(dzy, -em);
        rx = fmaf(c.ar, mx, -c.ai * my);
        ry = fmaf(c.ar, my, c.ai * mx);
    } else {
        ne = max(za + em, c.be);
        const float mx = ldexpf(dzx, za - ne), my = ldexpf(dzy, za - ne);
        const float px = ldexpf(dpx, c.be - ne), py = ldexpf(dpy, c.be - ne);
        rx = fmaf(c.ar, mx, fmaf(-c.ai, my, fmaf(c.br, px, -c.bi * py)));
        ry = fmaf(c.ar, my, fmaf(c.ai, mx, fmaf(c.br, py, c.bi * px)));
    }
}

// One leap from k if level 0 is valid: returns the level taken or -1.
template <bool kNoDc>
__device__ inline int bla_leap32(const Fexp32Data& D, float& dzx, float& dzy, int& z_exp,
                                 float dpx, float dpy, int k, int hi, int E) {
    Coef32 c = entry(D.bla.c32, D.bla.at(k, hi));
    const float q = fmaf(dzx, dzx, dzy * dzy);
    const int key = dz_key(q, z_exp - E);
    int d = hi;
    if (!(key < c.key)) {
        d = leap32_search(D, k, hi, key);
        if (d < 0) return -1;
        c = entry(D.bla.c32, D.bla.at(k, d));
    }
    float rx, ry;
    int ne;
    leap32_terms<kNoDc>(c, q, dzx, dzy, z_exp, dpx, dpy, rx, ry, ne);
    dzx = rx;
    dzy = ry;
    z_exp = ne;
    return d;
}

// Consecutive leaps while z_exp < z_deep, in a loop with a single exit and
// a single (rare) branch, so the exec-mask bookkeeping per leap stays small:
// a lane whose level 0 is not valid at its k applies the identity entry
// (A = 1, B = 0, which rescales its mantissa exactly) with next_hi = -2, and
// leaves the loop; the -2 then sets no_leap.
template <bool kNoDc>
__device__ inline void deep_leaps32(const Fexp32Data& D, float& dzx, float& dzy, int& z_exp,
                                    float dpx, float dpy, int& k, int& nbla, bool& no_leap,
                                    int E, int z_deep, int ref_len) {
    int hi = top_level(k, ref_len, D.bla.levels);
    while (z_exp < z_deep && hi >= 0) {
        Coef32 c = entry(D.bla.c32, D.bla.at(k, hi));
        const float q = fmaf(dzx, dzx, dzy * dzy);
        const int key = dz_key(q, z_exp - E);
        int d = hi;
        if (!(key < c.key)) {
            d = leap32_search(D, k, hi, key);
            if (d >= 0) {
                c = entry(D.bla.c32, D.bla.at(k, d));
            } else {
                c.ar = 1.0f;
                c.ai = c.br = c.bi = 0.0f;
                c.ae = 0;
                c.be = kZeroExp;
                c.next_hi = -2;
            }
        }
        float rx, ry;
        int ne;
        leap32_terms<kNoDc>(c, q, dzx, dzy, z_exp, dpx, dpy, rx, ry, ne);
        dzx = rx;
        dzy = ry;
        z_exp = ne;
        k += d >= 0 ? 1 << d : 0;
        nbla += d >= 0 ? 1 : 0;
        hi = c.next_hi;
    }
    no_leap = hi == -2;
}

// Hand-off point of the FP32 loop: once |dz| >= 2^kTail32 the delta continues
// in FP64 with plain steps.  Near escape a chaotic excursion amplifies the
// error already in dz by the product of |2z| over the excursion; a 24-bit
// mantissa, and the 1e-6 error of BLA leaps taken that late, do not survive
// it.  Below 2^-32 the delta is a linear correction whose relative error does
// not grow, so the FP32 loop and its BLA leaps cover all the deep iterations.
// (Measured on seeded deep views: p99 smooth error against exact FP64
// perturbation 1.4e-5..1.8e-5, versus 3e-4..4e-4 for the FP64-mantissa loop
// with BLA to the end.)
constexpr int kTail32 = -32;

// FP64 perturbation continuation of the FP32 loop from iteration k, with the
// delta now in actual units (dc * 2^-E underflows to 0 beyond FP64 range,
// where it is negligible next to |dz| >= 2^kTail32).
template <int T>
__device__ Sample fexp32_tail(const ap_view& v, const Fexp32Data& D, const PertConsts& pc,
                              double dzx, double dzy, double dcx_d, double dcy_d, int k,
                              int glitch, int plain, int nbla) {
    const bool julia = (T == 2 || T == 3);
    const int ref_len = v.ref_len;
    const bool strict = v.strict_glitch > 0;
    const double sc_e = v.e_scale <= 1074 ? ldexp(1.0, -v.e_scale) : 0.0;
    const double dpx = julia ? 0.0 : dcx_d * sc_e, dpy = julia ? 0.0 : dcy_d * sc_e;
    const int n = v.gen_n;
    while (k < ref_len) {
        const double rx = D.zr[k], ry = D.zi[k];
        const double zx = rx + dzx, zy = ry + dzy;
        const double mag2 = zx * zx + zy * zy;
        if (mag2 > kBailout2) {
            Sample r = escaped(k, mag2, pc.inv_ref, pc.inv_exact, v.ln_bailout_log);
            r.glitch = glitch;
            r.plain = plain;
            r.bla = nbla;
            return r;
        }
        if (strict && k > 15) {
            double r_mag2 = rx * rx + ry * ry;
            if (mag2 < 1e-6 * r_mag2 && r_mag2 > 1e-8) glitch = 1;
        }
        if (T == 0 || T == 2) {
            double ndx = (2.0 * rx + dzx) * dzx - (2.0 * ry + dzy) * dzy + dpx;
            dzy = 2.0 * ((rx + dzx) * dzy + ry * dzx) + dpy;
            dzx = ndx;
        } else if (T == 1 || T == 3) {
            double ndx = (2.0 * rx + dzx) * dzx - (2.0 * ry + dzy) * dzy + dpx;
            double sxs = rx >= 0.0 ? 1.0 : -1.0;
            double sys = ry >= 0.0 ? 1.0 : -1.0;
            double ddx = (rx + dzx) * sxs >= 0.0 ? sxs * dzx : -sxs * (2.0 * rx + dzx);
            double ddy = (ry + dzy) * sys >= 0.0 ? sys * dzy : -sys * (2.0 * ry + dzy);
            dzy = 2.0 * (fabs(rx) * ddy + fabs(ry) * ddx + ddx * ddy) - dpy;
            dzx = ndx;
        } else {
            double pr = dzx, pi = dzy, qr = rx, qi = ry;
            for (int e = 0; e < n - 1; ++e) {
                double t1r = pr * rx - pi * ry, t1i = pr * ry + pi * rx;
                double sr = pr + qr, sim = pi + qi;
                double t2r = sr * dzx - sim * dzy, t2i = sr * dzy + sim * dzx;
                pr = t1r + t2r;
                pi = t1i + t2i;
                double nq = qr * rx - qi * ry;
                qi = qr * ry + qi * rx;
                qr = nq;
            }
            double kdr = v.gen_kr * dzx - v.gen_ki * dzy, kdi = v.gen_kr * dzy + v.gen_ki * dzx;
            dzx = pr + kdr + dpx;
            dzy = pi + kdi + dpy;
        }
        ++k;
        ++plain;
    }
    const double zx = D.zr[ref_len] + dzx, zy = D.zi[ref_len] + dzy;
    const double scfb = v.e_scale <= 1020 ? pow2_f64(-v.e_scale) : 0.0;
    const double cx = julia ? v.julia_cx : v.ref_cx + dcx_d * scfb;
    const double cy = julia ? v.julia_cy : v.ref_cy + dcy_d * scfb;
    const double mag2 = zx * zx + zy * zy;
    if (mag2 > kBailout2) {
        Sample r = escaped(ref_len, mag2, pc.inv_ref, pc.inv_exact, v.ln_bailout_log);
        r.glitch = glitch;
        r.plain = plain;
        r.bla = nbla;
        return r;
    }
    if (ref_len < v.max_iter) glitch = 1;
    return stage2<T>(v, pc, zx, zy, cx, cy, k, glitch, plain, nbla);
}

template <int T>
__device__ Sample fexp32_sample(const ap_view& v, const Fexp32Data& D, double si, double sj) {
    const bool julia = (T == 2 || T == 3);
    const double dcy_d = (v.ref_py - sj) * v.dy;
    const double dcx_d = (si - v.ref_px) * v.dx;
    const float dcx = (float)dcx_d, dcy = (float)dcy_d;
    float dzx = julia ? dcx : 0.0f, dzy = julia ? dcy : 0.0f;
    const float dpx = julia ? 0.0f : dcx, dpy = julia ? 0.0f : dcy;
    int z_exp = 0;
    renorm32(dzx, dzy, z_exp);
    const int E = v.e_scale;
    const int ref_len = v.ref_len;
    const bool strict = v.strict_glitch > 0;
    const bool has_bla = D.bla.levels > 1 && (T == 0 || T == 2 || T == 4);
    const PertConsts pc = pert_consts(v);
    int k = 0, glitch = 0, plain = 0, nbla = 0;
    const int n = v.gen_n;
    const float kr = (float)v.gen_kr, ki = (float)v.gen_ki;
    const int z_deep = E + kTail32 - 40;
    if (mandel_start<T>(v, D.zr, D.zi)) {
        dzx = dpx;
        dzy = dpy;
        renorm32(dzx, dzy, z_exp);
        k = plain = 1;
    }

    while (k < ref_len) {
        // Deep regime (z_exp < z_deep, i.e. eff < kTail32 - 40): max|dz_m| < 2^32,
        // so |dz| < 2^(eff + 32) < 2^(kTail32 - 8): no hand-off, no glitch test
        // (it needs eff >= -40) and no escape can happen; only BLA leaps run,
        // in a loop of their own.  A leap of finite mantissas (inputs scaled
        // to at most 1, |am|, |bm| < 2) is finite, so the loop has no NaN test.
        bool no_leap = false;  // level 0 already failed at this k and state
        if ((T == 0 || T == 2 || T == 4) && has_bla) {
            deep_leaps32<julia>(D, dzx, dzy, z_exp, dpx, dpy, k, nbla, no_leap, E, z_deep, ref_len);
            if (k >= ref_len) break;
        }
        const int eff = z_exp - E;
        // |dz| = max|dz_m| 2^eff >= 2^kTail32: continue in FP64.
        if (eff >= kTail32 - 40) {
            float m = fmaxf(fabsf(dzx), fabsf(dzy));
            if (m > 0.0f && eff + ilogbf(m) >= kTail32)
                return fexp32_tail<T>(v, D, pc, ldexp((double)dzx, eff), ldexp((double)dzy, eff),
                                      dcx_d, dcy_d, k, glitch, plain, nbla);
        }
        const float2 Z = D.zf[k];
        if (strict && k > 15 && eff >= -40) {
            // |Z + dz|^2 < 1e-6 |Z|^2 in actual units, for |Z| >= 1e-4.
            float rmax = fmaxf(fabsf(Z.x), fabsf(Z.y));
            if (rmax >= 1e-4f && (dzx != 0.0f || dzy != 0.0f)) {
                float zx = Z.x + ldexpf(dzx, eff), zy = Z.y + ldexpf(dzy, eff);
                float m2 = zx * zx + zy * zy;
                float rm2 = Z.x * Z.x + Z.y * Z.y;
                if (m2 < 1e-6f * rm2) glitch = 1;
            }
        }
        if (T == 0 || T == 2 || T == 4) {
            if (has_bla && !no_leap && k < ref_len - 1) {
                const int d = bla_leap32<julia>(D, dzx, dzy, z_exp, dpx, dpy, k,
                                                top_level(k, ref_len, D.bla.levels), E);
                if (d >= 0) {
                    ++nbla;
                    k += 1 << d;
                    continue;
                }
            }
        }
        const float s_nl = ldexpf(1.0f, max(z_exp - E, -200));
        const float cxs = ldexpf(dpx, max(-z_exp, -200)), cys = ldexpf(dpy, max(-z_exp, -200));
        if (T == 0 || T == 2) {
            // 2 Z dz + dz^2 2^(z_exp-E) + dc 2^-z_exp
            float tx = dzx * s_nl, ty = dzy * s_nl;
            float lx = 2.0f * (Z.x * dzx - Z.y * dzy) + (tx * dzx - ty * dzy) + cxs;
            float ly = 2.0f * (Z.x * dzy + Z.y * dzx) + 2.0f * tx * dzy + cys;
            dzx = lx;
            dzy = ly;
        } else if (T == 1 || T == 3) {
            float sxs = Z.x >= 0.0f ? 1.0f : -1.0f;
            float sys = Z.y >= 0.0f ? 1.0f : -1.0f;
            float ax = ldexpf(dzx, max(z_exp - E, -200)), ay = ldexpf(dzy, max(z_exp - E, -200));
            float back = ldexpf(1.0f, min(E - z_exp, 200));
            float ddx = (Z.x + ax) * sxs >= 0.0f ? sxs * dzx : -sxs * (2.0f * Z.x * back + dzx);
            float ddy = (Z.y + ay) * sys >= 0.0f ? sys * dzy : -sys * (2.0f * Z.y * back + dzy);
            float lx = (2.0f * Z.x + dzx * s_nl) * dzx - (2.0f * Z.y + dzy * s_nl) * dzy + cxs;
            float ly = 2.0f * (fabsf(Z.x) * ddy + fabsf(Z.y) * ddx + ddx * ddy * s_nl) - cys;
            dzx = lx;
            dzy = ly;
        } else {
            float pr = dzx, pi = dzy, qr = Z.x, qi = Z.y;
            for (int e = 0; e < n - 1; ++e) {
                float t1r = pr * Z.x - pi * Z.y, t1i = pr * Z.y + pi * Z.x;
                float t2r = qr * dzx - qi * dzy, t2i = qr * dzy + qi * dzx;
                float t3r = pr * dzx - pi * dzy, t3i = pr * dzy + pi * dzx;
                pr = t1r + t2r + t3r * s_nl;
                pi = t1i + t2i + t3i * s_nl;
                float nq = qr * Z.x - qi * Z.y;
                qi = qr * Z.y + qi * Z.x;
                qr = nq;
            }
            float ndx = pr + (kr * dzx - ki * dzy) + cxs;
            dzy = pi + (kr * dzy + ki * dzx) + cys;
            dzx = ndx;
        }
        renorm32(dzx, dzy, z_exp);
        ++k;
        ++plain;
    }
    // The reference escaped with the delta still tiny: Stage 2 from
    // Z_ref_len + dz in FP64.
    return fexp32_tail<T>(v, D, pc, ldexp((double)dzx, z_exp - E), ldexp((double)dzy, z_exp - E),
                          dcx_d, dcy_d, k, glitch, plain, nbla);
}

template <int T>
struct EngFexp32 {
    using Data = Fexp32Data;
    static constexpr bool kIterInSubsamples = true;
    __device__ static Sample sample(const ap_view& v, const Data& d, double si, double sj) {
        return fexp32_sample<T>(v, d, si, sj);
    }
};

// ------------------------------------------------------------- launchers --
__global__ __launch_bounds__(256) void orbit_to_f32_kernel(const double* __restrict__ zr,
                                                           const double* __restrict__ zi, int n,
                                                           float2* __restrict__ out) {
    int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) out[k] = make_float2((float)zr[k], (float)zi[k]);
}

template <template <int> class E, class D>
static hipError_t dispatch_pert(const ap_view& v, const ap_outputs& o, const D& d, hipStream_t s) {
    switch (v.fractal_type) {
    case 0: return launch_pass<E<0>>(v, o, d, s);
    case 1: return launch_pass<E<1>>(v, o, d, s);
    case 2: return launch_pass<E<2>>(v, o, d, s);
    case 3: return launch_pass<E<3>>(v, o, d, s);
    case 4: return launch_pass<E<4>>(v, o, d, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_pert64(const ap_view& v, const ap_orbit& orb, const void* bla,
                         const ap_outputs& o, hipStream_t s) {
    PertData d;
    d.zr = orb.re;
    d.zi = orb.im;
    d.bla = bla_view(v.bla_levels > 1 ? bla : nullptr, orb.len, v.bla_levels);
    return dispatch_pert<EngPert64>(v, o, d, s);
}

hipError_t launch_fexp(int mant, const ap_view& v, const ap_orbit& orb, const void* bla,
                       const ap_outputs& o, void* scratch_f32_orbit, hipStream_t s) {
    BlaView b = bla_view(v.bla_levels > 1 ? bla : nullptr, orb.len, v.bla_levels);
    if (mant == 64) {
        PertData d;
        d.zr = orb.re;
        d.zi = orb.im;
        d.bla = b;
        return dispatch_pert<EngFexp64>(v, o, d, s);
    }
    Fexp32Data d;
    d.zr = orb.re;
    d.zi = orb.im;
    d.bla = b;
    if (b.levels > 0) {
        d.zf = b.zf;  // written with the table's FP32 form
    } else {
        d.zf = (const float2*)scratch_f32_orbit;
        int n = orb.len + 1;
        hipLaunchKernelGGL(orbit_to_f32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, orb.re,
                           orb.im, n, (float2*)scratch_f32_orbit);
    }
    return dispatch_pert<EngFexp32>(v, o, d, s);
}

}  // namespace ap
