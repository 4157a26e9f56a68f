// Device Voigt profile H(a, u) = Re w(u + i a), w the Faddeeva function.
//
// The reference evaluates Re w with the FP64 Faddeeva package
// (fake_spectra/singleabs.h:56-61 -> Faddeeva.cpp:679-971).  The fused
// deposit kernel evaluates it in FP32 instead, for damping parameters
// 0 <= a <= kA32Max, from two expansions:
//
//    H = exp(a^2 - u^2) cos(2au) - (2a/sqrt(pi)) G(u) - (a^3/sqrt(pi)) Z(u)
//  with G = 1 - 2u D(u) = D'(u) (D: Dawson's integral).  This is the Taylor
//  series of w about the real axis, whose even terms sum to Re exp(-z^2) and
//  whose odd terms are those of the entire function (2i/sqrt(pi)) D(z); the
//  a^5 term is below 1e-7 relative for a <= 0.04.
//  |u| < 7:  G from a piecewise polynomial table, Z = -D'''/3 with
//            D''' = (4u^2 - 6) G + 2;
//  |u| >= 7: G = x R(x), x = 1/u^2, R a polynomial that follows the
//            asymptotic series w ~ (i/sqrt(pi)) sum_n (2n-1)!!/2^n z^-(2n+1)
//            (the last table row holds x R(x) as a polynomial in x); Z = x^2 Q(x), Q = sum C(2n+3,3) (2n-1)!!/2^n x^n (3 terms),
//            the a^3 term of the same series, free of the cancellation in D'''.
//  Each lane picks its table row (a core piece or the wing row) and its a^3
//  form by selects, and every lane runs the same instructions, so the
//  wavefront never diverges on |u|.
//
// u arrives in FP32.  The deposit forms it as (c_hi - s_hi) + (c_lo - s_lo)
// from two-float splits of the kernel-sample offset c and of the sub-sample
// velocity s (s from the pixel's FP64 position), so u carries ~1 ulp relative
// error even where c and s nearly cancel; u^2 and the exponent of
// exp(a^2 - u^2) are FP32, and exp amplifies their relative error by ~2u^2
// ln2: the Gaussian term carries ~1.2e-7 u^2 relative error on top of the
// expansions' 3.6e-7 (measured against scipy.special.wofz, a <= 0.04,
// 0 <= u <= 3000).  Summed over a pixel's samples this stays two orders
// below the 1e-5 tau tolerance and the 2e-5 tail-cut guard band.
//
// An FP64 Faddeeva w(z) (voigt64) backs the FP32 path: pixels whose tail-cut
// decision sits within a guard band of tautail, and pairs with a > kA32Max.
// It is accurate to ~1e-13 relative on the path's domain:
//   real axis: w(x) = exp(-x^2) + (2i/sqrt(pi)) D(x), D by Rybicki's sum;
//   |z| >= 8 or y >= 2: Laplace continued fraction;
//   otherwise: Taylor series in y about the real axis.
#pragma once
#include <hip/hip_runtime.h>
#include "dawson_table.h"

namespace fsamd {

constexpr float kVoigtCoreMax = 7.0f;
constexpr float kA32Max = 0.04f;
constexpr float kTwoOverSqrtPiF = 1.12837916709551257390f;
constexpr float kInvSqrtPiF = 0.56418958354775628695f;
constexpr double kSqrtPi = 1.77245385090551602729;
constexpr double kTwoOverSqrtPi = 1.12837916709551257390;
constexpr double kLog2E = 1.44269504088896340736;
constexpr float kLog2EF = 1.44269504088896340736f;
constexpr int kGTableSize = FSAMD_G_NROW * FSAMD_G_STRIDE;

// (2n+1)!!/2^n, n = 0..8
#define FSAMD_WING_P {1.0f, 1.5f, 3.75f, 13.125f, 59.0625f, 324.84375f, 2111.484375f, 15836.1328125f, 134607.12890625f}
// C(2n+3,3) (2n-1)!!/2^n, n = 0..2
#define FSAMD_WING_Q {1.0f, 5.0f, 26.25f}

// Per-pair constants of the FP32 profile.
struct VoigtA {
    float a;            // damping parameter
    float a2;           // a^2
    float a2_6;         // a^2 / 6
    float two_a_sqpi;   // 2a/sqrt(pi)
    float a_sqpi;       // a/sqrt(pi)
    float a3_sqpi3;     // a^3/(3 sqrt(pi))
    float a2_log2e;     // a^2 log2(e)
    // log2 of Re exp(-(u + ia)^2) = a2_log2e + lc1 u^2 + lc2 u^4 + lc3 u^6 (voigt32)
    float lc1, lc2, lc3;
};

__device__ inline VoigtA make_voigt_a(double a)
{
    VoigtA v;
    v.a = (float)a;
    v.a2 = (float)(a * a);
    v.a2_6 = (float)(a * a * (1.0 / 6.0));
    v.two_a_sqpi = (float)(a * kTwoOverSqrtPi);
    v.a_sqpi = (float)(a * (1.0 / kSqrtPi));
    v.a3_sqpi3 = (float)(a * a * a * (1.0 / (3.0 * kSqrtPi)));
    v.a2_log2e = (float)(a * a * kLog2E);
    const double a2 = a * a;
    v.lc1 = (float)(-(1.0 + 2.0 * a2) * kLog2E);
    v.lc2 = (float)(-(4.0 / 3.0) * a2 * a2 * kLog2E);
    v.lc3 = (float)(-(64.0 / 45.0) * a2 * a2 * a2 * kLog2E);
    return v;
}

// Damping parameters up to which the deposit uses the light form of the
// profile: without the a^3 term and with the Gaussian exponent cut to
// a^2 - (1 + 2a^2) u^2.  For a <= kA3Drop the two together change H by at
// most 5.5e-7 relative over all u (the a^3 term at u ~ 3.2; measured against
// scipy.special.wofz).
constexpr float kA3Drop = 2e-3f;

// H(a, u) with u the sample's offset from the line centre in units of b,
// any u.  The table `tab` is 16-byte aligned.  kA3 = false: the light form
// (for a <= kA3Drop), which equals the full one when a3_sqpi3, lc2 and lc3
// are zero.
template <bool kA3 = true>
__device__ inline float voigt32(const VoigtA &v, float u, const float *tab)
{
    static_assert(FSAMD_G_DEG == 6 && FSAMD_G_STRIDE % 4 == 0 && FSAMD_G_STRIDE >= 8, "voigt32 evaluates 7 coefficients per row");
    const float uh = __builtin_fabsf(u);
    const float p = u * u;
    // Re exp(-(u + ia)^2) = exp(a^2 - u^2) cos(2au) = 2^y, with ln cos x =
    // -x^2/2 - x^4/12 - x^6/45 - ...:  y = log2(e) (a^2 - (1 + 2a^2) u^2 -
    // (4/3) a^4 u^4 - (64/45) a^6 u^6).  The dropped term, (17/2520)(2au)^8,
    // changes H by < 5e-9 relative for a <= 0.04 (against scipy's wofz); the
    // u^4 and u^6 terms are small next to the u^2 one, so y keeps the
    // rounding of its first fma.  2^y is 0 once y passes the FP32 range.
    const float y = __builtin_fmaf(p, v.lc1, v.a2_log2e);
    const float re_exp = __builtin_amdgcn_exp2f(kA3 ? __builtin_fmaf(p * p, __builtin_fmaf(p, v.lc3, v.lc2), y) : y);
    // G: table piece (|u| < 7) or the wing row, selected per lane; both rows
    // are polynomials in their variable, so G is the row's Horner sum
    const float x = __builtin_amdgcn_rcpf(p);
    const bool core = uh < kVoigtCoreMax;
    const float q = uh * (1.0f / FSAMD_G_WIDTH);
    // piece q (variable: the fractional part of q) or the wing row (variable x)
    constexpr int kRowBytes = FSAMD_G_STRIDE * (int)sizeof(float);
    const int off = core ? (int)q * kRowBytes : FSAMD_G_NPIECE * kRowBytes;
    const float var = core ? __builtin_amdgcn_fractf(q) : x;
    const float *row = (const float *)((const char *)tab + off);
    const float4 lo = *(const float4 *)row;
    const float4 hi = *(const float4 *)(row + 4);
    float r = __builtin_fmaf(hi.z, var, hi.y);
    r = __builtin_fmaf(r, var, hi.x);
    r = __builtin_fmaf(r, var, lo.w);
    r = __builtin_fmaf(r, var, lo.z);
    r = __builtin_fmaf(r, var, lo.y);
    const float G = __builtin_fmaf(r, var, lo.x);
    if constexpr (!kA3) return __builtin_fmaf(-v.two_a_sqpi, G, re_exp);
    // a^3 term, -(a^3/sqrt(pi)) Z = (a^3/(3 sqrt(pi))) D with D = D''' =
    // (4u^2 - 6) G + 2, or D = -3 Z = -3 x^2 (1 + 5x + 26.25x^2) in the wing
    const float Dc = __builtin_fmaf(__builtin_fmaf(4.0f, p, -6.0f), G, 2.0f);
    const float Dw = x * x * __builtin_fmaf(__builtin_fmaf(-78.75f, x, -15.0f), x, -3.0f);
    const float D = core ? Dc : Dw;
    return __builtin_fmaf(v.a3_sqpi3, D, __builtin_fmaf(-v.two_a_sqpi, G, re_exp));
}

// ---------------------------------------------------------------------------
// Profile derivatives for the adjoint, FP32.  With z = u + ia, w' = -2zw + 2i/sqrt(pi):
//   dH/du = Re w' = -2(uH - aL),   dH/da = -Im w' = 2(uL + aH) - 2/sqrt(pi),
// L = Im w.  The adjoint needs H, Hu = dH/du and F = u dH/du + a dH/da
// (the response to the thermal width b, all sample offsets scaling as 1/b):
//   F = -2(u^2 - a^2) H + 4 a u L - 2a/sqrt(pi).
// Core |u| < 7: L = -exp(a^2-u^2) sin(2au) + (2/sqrt(pi)) [D(u) - (a^2/2) D''(u)],
// D = (1 - G)/(2u), D'' = -2D - 2uG.  Wing |u| >= 7: the asymptotic series
// H = (a/sqrt(pi)) sum p_k u^(-2k-2) gives Hu and F termwise, without the
// cancellation of the closed forms (a^2-relative corrections dropped there).
struct VoigtGrad {
    float H, Hu, F;
};

__device__ inline VoigtGrad voigt32_grad(const VoigtA &v, float u, const float *gtab)
{
    const float sgn = u < 0 ? -1.0f : 1.0f;
    const float uh = __builtin_fabsf(u);
    VoigtGrad r;
    if (uh >= kVoigtCoreMax) {
        const float x = __builtin_amdgcn_rcpf(uh * uh);
        const float P[9] = FSAMD_WING_P;
        const float Q[3] = FSAMD_WING_Q;
        float s = P[8], d = 18.0f * P[8], f = -17.0f * P[8];
#pragma unroll
        for (int k = 7; k >= 0; --k) {
            s = __builtin_fmaf(s, x, P[k]);
            d = __builtin_fmaf(d, x, (2.0f * k + 2.0f) * P[k]);
            f = __builtin_fmaf(f, x, -(2.0f * k + 1.0f) * P[k]);
        }
        float q = __builtin_fmaf(Q[2], x, Q[1]);
        q = __builtin_fmaf(q, x, Q[0]);
        r.H = v.a_sqpi * x * __builtin_fmaf(-v.a2 * x, q, s);
        r.Hu = -sgn * v.a_sqpi * x * d * __builtin_amdgcn_rcpf(uh);
        r.F = v.a_sqpi * x * f;
        return r;
    }
    // exp(a^2 - u^2) as in voigt32
    const float p = uh * uh;
    const float E = __builtin_amdgcn_exp2f(__builtin_fmaf(p, -kLog2EF, v.a2_log2e));
    const float c = 2.0f * v.a * uh;
    const float c2 = c * c;
    float cs = __builtin_fmaf(c2, -1.0f / 40320.0f, 1.0f / 720.0f);
    cs = __builtin_fmaf(c2, -cs, 1.0f / 24.0f);
    cs = __builtin_fmaf(c2, -cs, 0.5f);
    cs = __builtin_fmaf(c2, -cs, 1.0f);
    float sn = __builtin_fmaf(c2, -1.0f / 5040.0f, 1.0f / 120.0f);
    sn = __builtin_fmaf(c2, sn, -1.0f / 6.0f);
    sn = c * __builtin_fmaf(c2, sn, 1.0f);
    // G from the piece's row (two 16-byte reads), in the fractional part of u/width
    const float q = uh * (1.0f / FSAMD_G_WIDTH);
    const int idx = min((int)q, FSAMD_G_NPIECE - 1);
    const float var = q - (float)idx;
    const float *row = gtab + idx * FSAMD_G_STRIDE;
    const float4 lo = *(const float4 *)row;
    const float4 hi = *(const float4 *)(row + 4);
    float G = __builtin_fmaf(hi.z, var, hi.y);
    G = __builtin_fmaf(G, var, hi.x);
    G = __builtin_fmaf(G, var, lo.w);
    G = __builtin_fmaf(G, var, lo.z);
    G = __builtin_fmaf(G, var, lo.y);
    G = __builtin_fmaf(G, var, lo.x);
    float D;
    if (uh < 0.25f) {
        // D(u) = u - 2u^3/3 + 4u^5/15 - 8u^7/105 + 16u^9/945
        float t = __builtin_fmaf(p, 16.0f / 945.0f, -8.0f / 105.0f);
        t = __builtin_fmaf(p, t, 4.0f / 15.0f);
        t = __builtin_fmaf(p, t, -2.0f / 3.0f);
        D = uh * __builtin_fmaf(p, t, 1.0f);
    } else {
        D = (1.0f - G) * 0.5f * __builtin_amdgcn_rcpf(uh);
    }
    const float D2 = -2.0f * D - 2.0f * uh * G;
    const float D3 = __builtin_fmaf(__builtin_fmaf(4.0f, p, -6.0f), G, 2.0f);
    const float H = __builtin_fmaf(-v.two_a_sqpi, __builtin_fmaf(-v.a2_6, D3, G), E * cs);
    const float L = __builtin_fmaf(kTwoOverSqrtPiF, __builtin_fmaf(-0.5f * v.a2, D2, D), -E * sn);
    r.H = H;
    r.Hu = sgn * (-2.0f * __builtin_fmaf(uh, H, -v.a * L));
    r.F = __builtin_fmaf(-2.0f * (p - v.a2), H, __builtin_fmaf(4.0f * v.a * uh, L, -v.two_a_sqpi));
    return r;
}

// ---------------------------------------------------------------------------
// FP64 Faddeeva w(z), y >= 0.

__device__ __noinline__ double dawson64(double x)
{
    const double ax = fabs(x);
    if (ax < 0.2) {
        double t = 1.0, s = 1.0;
        const double m = -2.0 * ax * ax;
#pragma unroll 1
        for (int k = 1; k < 14; k++) {
            t *= m / (2 * k + 1);
            s += t;
        }
        return x * s;
    }
    const double h = 0.2;
    const double n0 = 2.0 * floor(0.5 * ax / h + 0.5);
    const double xp = ax - n0 * h;
    const double e1 = exp(2.0 * xp * h), e2 = e1 * e1;
    double up = e1, dn = 1.0 / e1, s = 0.0;
#pragma unroll 1
    for (int i = 0; i < 22; i++) {
        const double n = 2 * i + 1;
        const double g = exp(-(n * h) * (n * h));
        s += g * (up / (n0 + n) + dn / (n0 - n));
        up *= e2;
        dn /= e2;
    }
    const double d = exp(-xp * xp) * s / kSqrtPi;
    return x < 0 ? -d : d;
}

__device__ __noinline__ void w64_contfrac(double x, double y, int nterms, double &wr, double &wi)
{
    double tr = x, ti = y;
    for (int k = nterms; k >= 1; k--) {
        const double den = tr * tr + ti * ti;
        const double c = 0.5 * k / den;
        tr = x - c * tr;
        ti = y + c * ti;
    }
    const double den = tr * tr + ti * ti;
    wr = ti / den / kSqrtPi;
    wi = tr / den / kSqrtPi;
}

__device__ __noinline__ void w64_taylor_y(double x, double y, double &wr, double &wi)
{
    const double w0r = exp(-x * x), w0i = kTwoOverSqrtPi * dawson64(x);
    const double w1r = -2.0 * x * w0r, w1i = -2.0 * x * w0i + kTwoOverSqrtPi;
    double sr = w0r - y * w1i, si = w0i + y * w1r;
    double cr = 0.0, ci = y;
    double pr = w0r, pim = w0i, qr = w1r, qi = w1i;
    int small = 0;
    for (int k = 1; k < 200; k++) {
        const double nr = -2.0 * x * qr - 2.0 * k * pr;
        const double ni = -2.0 * x * qi - 2.0 * k * pim;
        pr = qr;
        pim = qi;
        qr = nr;
        qi = ni;
        const double t = cr;
        cr = -ci * y / (k + 1);
        ci = t * y / (k + 1);
        const double tr = cr * qr - ci * qi, ti = cr * qi + ci * qr;
        sr += tr;
        si += ti;
        if (fabs(tr) + fabs(ti) <= 1e-19 * (fabs(sr) + fabs(si))) {
            if (++small >= 3) break;
        } else {
            small = 0;
        }
    }
    wr = sr;
    wi = si;
}

__device__ inline void w64(double x, double y, double &wr, double &wi)
{
    const double ax = fabs(x);
    double r, i;
    if (y == 0.0) {
        r = exp(-ax * ax);
        i = kTwoOverSqrtPi * dawson64(ax);
    } else if (ax * ax + y * y >= 64.0) {
        w64_contfrac(ax, y, 60, r, i);
    } else if (y >= 2.0) {
        w64_contfrac(ax, y, 400, r, i);
    } else {
        w64_taylor_y(ax, y, r, i);
    }
    wr = r;
    wi = x < 0 ? -i : i;
}

__device__ inline double voigt64(double a, double u)
{
    double r, i;
    w64(u, a, r, i);
    return r;
}

}  // namespace fsamd
