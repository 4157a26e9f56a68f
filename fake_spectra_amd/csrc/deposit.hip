// Fused SPH-kernel x Voigt deposit into LDS-staged sightline rows.
//
// Reference: ParticleInterp::compute_tau (fake_spectra/part_int.cpp:20-51)
// loops over sightlines and, per nearby particle, calls
// LineAbsorption::add_tau_particle (absorption.cpp:212-279), which walks
// pixels outward from the particle in both directions, adds
//     taulast = amp * dens * tau_kern_outer(vlow, vlow + dv) / velfac
// and stops after the first pixel with taulast < tautail.
// tau_kern_outer/tau_kern_inner (singleabs.h:104-167) average 7 interior
// kernel samples of K(q) * Re w over the pixel, with odd-point trapezoid
// sub-sampling when the pixel is wider than b/2.
//
// GPU mapping:
//  * one work item = (sightline, line, range of its pairs); a persistent grid
//    pulls items heaviest-first from a queue; the row of the sightline lives
//    in LDS as float64 and is flushed with coalesced stores (or FP64 atomics
//    when a sightline's pairs are split across several items);
//  * a wavefront handles one pair at a time with lanes on consecutive pixels;
//    the sequential tail cut becomes a ballot: the first lane whose taulast
//    falls below tautail ends the walk, lanes up to it deposit;
//  * the mapping position -> velocity -> pixel, the pixel and sub-sample
//    offsets and the accumulation are FP64, as in the reference; the kernel
//    weights and the Voigt profile are FP32 (voigt.h);
//  * a pixel whose FP32 taulast lies within 2e-5 relative of tautail is
//    re-evaluated in FP64 so the tail-cut decision matches the reference;
//    pairs with a damping parameter above kA32Max are evaluated in FP64.
// Column density (compute_colden, part_int.cpp:53-84 with
// add_colden_particle, absorption.cpp:167-210) is all FP64, one lane per pair.
#include "fsamd_internal.h"
#include "plan.h"
#include "voigt.h"
#include "deposit.h"
#include "wing_table.h"

#include <mutex>
#include <unordered_map>

namespace fsamd {

__constant__ __attribute__((aligned(16))) float c_gtab[kGTableSize] = FSAMD_G_COEF_INIT;

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr double kGuard = 2e-5;
constexpr int kWingTerms = 10;

// Far-wing coefficient tables.  The profile's asymptotic series in x = 1/u^2,
//   H(a,u) = (a/sqrt(pi)) sum_k (p_k - a^2 q_{k-1}) u^(-2k-2),
// p_k = (2k+1)!!/2^k, q_k = C(2k+3,3) (2k-1)!!/2^k (voigt.h), is expanded
// about the pixel centre U over the sample offsets c (Taylor in c/U); the
// even moments m_j = sum w c^(2j) of the pixel's weighted samples then give
//   sum_samples w H = sum_n X^(n+1) sum_{k<=n} g_k C(2n+1, 2(n-k)) m_(n-k).
__constant__ double c_wing_p[kWingTerms] = {1.0, 1.5, 3.75, 13.125, 59.0625, 324.84375, 2111.484375,
                                             15836.1328125, 134607.12890625, 1278767.724609375};
__constant__ double c_wing_q[kWingTerms] = {1.0, 5.0, 26.25, 157.5, 1082.8125, 8445.9375, 73901.953125,
                                             717904.6875, 7672606.34765625, 89513740.72265625};
static_assert(FSAMD_WING_NTERMS == kWingTerms, "wing_table.h");
__constant__ double c_wing_T[(FSAMD_WING_TMAX + 1) * kWingTerms] = FSAMD_WING_T_INIT;
__constant__ double c_binom_even[kWingTerms * kWingTerms] = FSAMD_BINOM_EVEN_INIT;
__constant__ double c_binom_odd[kWingTerms * kWingTerms] = FSAMD_BINOM_ODD_INIT;

// singleabs.h:17-26
__device__ inline double sph_cubic_kernel(double q)
{
    const double norm = 32. / 4 / kPi;
    if (q >= 1) return 0;
    if (q < 0.5) return norm * (1 - 6 * q * q + 6 * q * q * q);
    const double r = 1. - q;
    return norm * (2 * (r * r * r));
}

// singleabs.h:31-42: quintic spline of support 1 as a sum of truncated fifth
// powers, 9/(40 pi) sum_k c_k max(0, r_k - 3q)^5, (r, c) = (3, 1), (2, -6), (1, 15)
__device__ inline double sph_quintic_kernel(double q)
{
    const double norm = 9. / 40 / kPi;
    if (q >= 1) return 0;
    const double t3 = 3. - 3. * q, t2 = 2. - 3. * q, t1 = 1. - 3. * q;
    double f = t3 * t3 * t3 * t3 * t3;
    if (t2 > 0) f -= 6. * (t2 * t2 * t2 * t2 * t2);
    if (t1 > 0) f += 15. * (t1 * t1 * t1 * t1 * t1);
    return norm * f;
}

// Kernel weight of tau_kern_inner (singleabs.h:158-163): the cubic or quintic
// spline, 3/(4 pi) for the tophat, and no weight (1) for any other id, which
// takes none of the reference's branches there.
__device__ inline double kernel_weight(int kernel, double q)
{
    if (kernel == kCubicSpline) return sph_cubic_kernel(q);
    if (kernel == kSphQuintic) return sph_quintic_kernel(q);
    if (kernel == kTopHat) return 3. / 4. / kPi;
    return 1.0;
}

__device__ inline LineDesc line_of(const DepositArgs &A, int k) { return A.lines[k]; }

// Sightline data needed to recompute dr2 exactly as the binning did.
struct LosInfo {
    int ax;
    double key, lproj2;
};

__device__ inline LosInfo los_info(const DepositArgs &A, int s)
{
    LosInfo L;
    L.ax = A.axis[s];
    const double *c = &A.cofm[3 * (size_t)s];
    if (L.ax == 1) {
        L.key = c[1];
        L.lproj2 = c[2];
    } else if (L.ax == 2) {
        L.key = c[0];
        L.lproj2 = c[2];
    } else {
        L.key = c[0];
        L.lproj2 = c[1];
    }
    return L;
}

// index_table.cpp:70-87 with the particle coordinates of index_table.cpp:33-40
__device__ inline double pair_dr2(const LosInfo &L, const float *pp, double box)
{
    float first, second;
    if (L.ax == 1) {
        first = pp[1];
        second = pp[2];
    } else if (L.ax == 2) {
        first = pp[0];
        second = pp[2];
    } else {
        first = pp[0];
        second = pp[1];
    }
    double dr = fabs((double)first - L.key);
    if (dr > 0.5 * box) dr = box - dr;
    double d2 = dr * dr;
    dr = fabs((double)second - L.lproj2);
    if (dr > 0.5 * box) dr = box - dr;
    d2 += dr * dr;
    return d2;
}

// z mod nbins in [0, nbins) (absorption.cpp:257-259); walks stay within half a
// box of a pixel inside the box, so one add or subtract covers all but
// particles with extreme peculiar velocities.
__device__ inline int wrap_pixel(int z, int nbins)
{
    if (z >= 0 && z < nbins) return z;
    if (z < 0 && z >= -nbins) return z + nbins;
    if (z >= nbins && z < 2 * nbins) return z - nbins;
    int j = z % nbins;
    return j < 0 ? j + nbins : j;
}

// Per-pair constants of LineAbsorption::add_tau_particle and SingleAbsorber.
// Each wavefront keeps the constants of its current pair in an LDS slot: the
// values are uniform across the wave, and reading them with broadcast LDS
// loads keeps them out of the vector registers of the per-pixel loop.
// One interior kernel sample of a pair (singleabs.h:146-157): its offset
// vv_i / b as a two-float split (cv_hi + cv_lo, ~48 bits) and its weight
// K(q_i), read together (one 16-byte LDS read) in the sample loop.
struct KernelSample {
    float cv_hi, cv_lo;
    float w;
    float pad;
};

__device__ inline void set_sample_offset(KernelSample &k, double cv)
{
    k.cv_hi = (float)cv;
    k.cv_lo = (float)(cv - (double)k.cv_hi);
}

__device__ inline double sample_offset(const KernelSample &k) { return (double)k.cv_hi + (double)k.cv_lo; }

// The sub-sample velocity in units of b as a two-float split.
struct Split {
    float hi, lo;
};

__device__ inline Split split_s(double s)
{
    Split r;
    r.hi = (float)s;
    r.lo = (float)(s - (double)r.hi);
    return r;
}

// u = (vv_i - vouter)/b in FP32 from the two splits: the high parts cancel
// exactly where u is small (Sterbenz), so u keeps ~1 ulp relative error.
__device__ inline float sample_u(const KernelSample &k, Split s)
{
    return (k.cv_hi - s.hi) + (k.cv_lo - s.lo);
}

// The constants a walk of the main kernel reads, in one layout in the pair
// set-up record and in the group's PairTau slot, so that taking a record
// copies them with one 8-byte move per lane.
struct WalkHead {
    double vel;      // velfac*ppos + pvel (absorption.cpp:234)
    double inv_b;
    double scale;    // amp * dens * dvk / velfac: taulast = scale * (sample sum)
    double inv_nm1;  // 1 / (np - 1)
    // Far wing: for a pixel centred at U (units of b) with |U| >= ufar the
    // 7 x np samples sum to Y (C0 + C1 Y + ... + C9 Y^9), Y = (ufar/U)^2.
    double ufar;
    int zmax;
    int nm1, rounds, pstep;   // near-step shape of the main kernel (step_shape)
};
static_assert(sizeof(WalkHead) == 56, "WalkHead is seven 8-byte words");

struct PairTau : WalkHead {
    KernelSample smp[kNGrid - 1];   // kernel sample velocities / b and weights
    double b;        // btherm (absorption.cpp:220)
    double dvk;      // 2 vhigh / NGRID (singleabs.h:146)
    double vhigh, vdr2, vsmooth, aa;
    float ufar2;     // ufar^2 (fix-up and adjoint kernels)
    double scale_nd; // amp * dvk / velfac (scale without the density)
    double temp;     // particle temperature
    int part;        // particle index
    float Cw[kWingTerms];
    float CwF[kWingTerms];   // adjoint: far-wing coefficients of sum w (u dH/du + a dH/da)
    VoigtA va;
    int fp64;
    int np;          // sub-samples per pixel (1 = midpoint)
    int wing_ready;  // Cw computed
    double dv;       // sub-sample spacing bintov / (np - 1)
};

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Shuffle of a double from lane `src` of the same kG-lane group.
template <int kG>
__device__ inline double shfl_d(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __shfl((int)(b & 0xffffffffll), src, kG);
    const int hi = __shfl((int)(b >> 32), src, kG);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum_i K(q_i) Re w((vv_i - vouter)/b + i a), FP32 profile; the sample
// offsets from the FP64 anchor vouter/b (one split per sum).  kA3 = false:
// without the profile's a^3 term (pairs with a <= kA3Drop).
template <bool kA3 = true>
__device__ inline double inner32(const PairTau &P, double vouter, const float *gtab)
{
    const Split s = split_s(vouter * P.inv_b);
    const VoigtA va = P.va;
    float acc = 0.0f;
#pragma unroll  // full: 5% faster than unrolling by 2 at 256^3 and 512^3 (same operation order)
    for (int i = 0; i < kNGrid - 1; ++i) {
        const KernelSample k = P.smp[i];
        acc = __builtin_fmaf(k.w, voigt32<kA3>(va, sample_u(k, s), gtab), acc);
    }
    return (double)acc;
}

// Same sum with the FP64 profile and FP64 kernel weights; out of line and
// with scalar arguments because it only runs for rare pixels.
__device__ __noinline__ double inner64(double aa, double b, double dvk, double vhigh, double vdr2,
                                       double vsmooth, double vouter, int kernel)
{
    double total = 0;
    for (int i = 1; i < kNGrid; ++i) {
        const double vv = i * dvk - vhigh;
        const double q = sqrt(vdr2 + vv * vv) / vsmooth;
        const double t0 = (vv - vouter) / b;
        total += voigt64(aa, t0) * kernel_weight(kernel, q);
    }
    return total;
}

__device__ inline double inner64(const PairTau &P, double vouter, int kernel)
{
    return inner64(P.aa, P.b, P.dvk, P.vhigh, P.vdr2, P.vsmooth, vouter, kernel);
}

// amp*dens*tau_kern_outer(vlow, vlow+bintov)/velfac  (absorption.cpp:255; singleabs.h:104-126)
template <bool k64>
__device__ inline double taulast_pixel(const PairTau &P, double vlow, double bintov, double velfac, int kernel,
                                       const float *gtab)
{
    const double vhighp = vlow + bintov;
    const double width = vhighp - vlow;
    if (!k64) {
        const double U = (vlow + 0.5 * bintov) * P.inv_b;
        if (fabs(U) >= P.ufar) {
            const float Y = P.ufar2 * __builtin_amdgcn_rcpf((float)(U * U));
            float S = P.Cw[kWingTerms - 1];
#pragma unroll
            for (int t = kWingTerms - 2; t >= 0; --t) S = __builtin_fmaf(S, Y, P.Cw[t]);
            return P.scale * (double)(S * Y);
        }
    }
    if (width < P.b / 2.) {
        const double mid = (vhighp + vlow) / 2.;
        return P.scale * (k64 ? inner64(P, mid, kernel) : inner32(P, mid, gtab));
    }
    const int np = 2 * ceil(width / (P.b / 2.) / 2) + 1.;
    const double dv = width / (np - 1);
    double total = 0.5 * (k64 ? inner64(P, vlow, kernel) : inner32(P, vlow, gtab));
    for (int i = 1; i < np - 1; ++i) {
        const double vv = i * dv + vlow;
        total += k64 ? inner64(P, vv, kernel) : inner32(P, vv, gtab);
    }
    total += 0.5 * (k64 ? inner64(P, vhighp, kernel) : inner32(P, vhighp, gtab));
    return P.scale * (total / (np - 1));
}

// x^n for 0 <= n < 16 by squaring, with a uniform instruction stream.
__device__ inline double ipow(double x, int n)
{
    double r = 1.0;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        if (n & (1 << bit)) r *= x;
        x *= x;
    }
    return r;
}

// Far-wing coefficients of one pair.  The sample offsets from the pixel
// centre are c = x_i - y_j with x_i = vv_i/b the kernel samples (weights W_i)
// and y_j the sub-sample offsets of tau_kern_outer (singleabs.h:110-125,
// trapezoid weights t_j); both sets are symmetric, so the even moments of c
// factor:  m_n = sum_k C(2n,2k) X_k Y_(n-k),  X_k = sum W_i x_i^2k,
// Y_k = sum t_j y_j^2k.  Offsets are measured in units of ufar so that all
// coefficients stay within FP32 range:  C'_n = sum_(k+j=n) g_k ufar^(-2k-2)
// C(2n+1,2j) m'_j with m'_j the moments of c/ufar.  Lane n < 10 builds the
// n-th moments and coefficient and stores C'_n.  The kernel samples are
// x_i = (vhigh/b)(i-3)/4 (singleabs.h:146-149), so X_k takes the three
// positive offsets with the summed weights of each +- pair; Y_k = h^2k T[m][k]
// for a pixel of half-width h in np = m + 1 sub-samples, with T tabulated
// exactly (wing_table.h) up to m = FSAMD_WING_TMAX.
template <int kG, bool kAdj = false>
__device__ inline void wing_setup(PairTau &P, double bintov, int lane)
{
    const double inv_u = 1.0 / P.ufar;
    const int np = P.np;
    const int m1 = np - 1;
    const int n = lane < kWingTerms ? lane : 0;
    // Y_n: pixel sub-samples
    double Yn;
    if (m1 == 0) {
        Yn = n == 0 ? 1.0 : 0.0;
    } else if (m1 <= FSAMD_WING_TMAX) {
        const double h = 0.5 * bintov * P.inv_b * inv_u;
        Yn = ipow(h * h, n) * c_wing_T[m1 * kWingTerms + n];
    } else {
        const double dv = bintov / m1;
        Yn = 0.0;
        for (int j = 0; j < np; ++j) {
            const double y = (j * dv - 0.5 * bintov) * P.inv_b * inv_u;
            const double t = ((j == 0 || j == m1) ? 0.5 : 1.0) / m1;
            Yn += t * ipow(y * y, n);
        }
    }
    // m'_n = sum_k C(2n,2k) X_k Y_(n-k), X_k by running powers
    const double c4 = sample_offset(P.smp[4]) * inv_u, c5 = sample_offset(P.smp[5]) * inv_u,
                 c6 = sample_offset(P.smp[6]) * inv_u;
    const double x4 = c4 * c4, x5 = c5 * c5, x6 = c6 * c6;
    const double w4 = (double)P.smp[4].w + (double)P.smp[2].w;
    const double w5 = (double)P.smp[5].w + (double)P.smp[1].w;
    const double w6 = (double)P.smp[6].w + (double)P.smp[0].w;
    double p4 = 1.0, p5 = 1.0, p6 = 1.0, mom = 0.0;
#pragma unroll 1
    for (int k = 0; k < kWingTerms; ++k) {
        double xk = w4 * p4 + w5 * p5 + w6 * p6;
        if (k == 0) xk += (double)P.smp[3].w;
        p4 *= x4;
        p5 *= x5;
        p6 *= x6;
        const double yk = shfl_d<kG>(Yn, n - k < 0 ? 0 : n - k);
        if (k <= n) mom += c_binom_even[n * kWingTerms + k] * xk * yk;
    }
    // C'_n = sum_(k+j=n) g_k ufar^(-2k-2) C(2n+1, 2j) m'_j
    const double a = P.aa, a2 = a * a, iu2 = inv_u * inv_u, u2 = P.ufar * P.ufar;
    // j = 0..n walks k = n..0: start from ufar^(-2n-2)
    double C = 0.0, CF = 0.0, gs = a * (1.0 / kSqrtPi) * iu2 * ipow(iu2, n);
#pragma unroll 1
    for (int j = 0; j < kWingTerms; ++j) {
        const double mj = shfl_d<kG>(mom, j);
        if (j > 0) gs *= u2;
        if (j <= n) {
            const int k = n - j;
            const double binom = c_binom_odd[n * kWingTerms + j];
            const double qk = k > 0 ? c_wing_q[k - 1] : 0.0;
            C += gs * (c_wing_p[k] - a2 * qk) * binom * mj;
            // u dH/du + a dH/da of the term a p_k u^(-2k-2) - a^3 q_(k-1) u^(-2k-2)
            if (kAdj) CF += gs * (-(2.0 * k + 1.0) * c_wing_p[k] + (2.0 * k - 1.0) * a2 * qk) * binom * mj;
        }
    }
    if (lane < kWingTerms) {
        P.Cw[lane] = (float)C;
        if (kAdj) P.CwF[lane] = (float)CF;
    }
}

// Sub-sample grid of one walk step in the main kernel.  A step evaluates one
// 7-sample kernel sum per lane and round, on the sub-sample grid of
// tau_kern_outer (singleabs.h:104-126) laid out across the lanes: adjacent
// pixels share their boundary sub-sample, so a step of p pixels with np
// sub-samples each needs p (np - 1) + 1 sums; midpoint pixels (np = 1) need
// one each.  Pairs whose grid exceeds kRoundsMax rounds take the FP64 path.
constexpr int kRoundsMax = 4;

struct StepShape {
    int nm1;      // np - 1 (0: midpoint)
    int rounds;   // lane rounds per step
    int pstep;    // pixels per step
};

__device__ inline StepShape step_shape(int np, int kg)
{
    StepShape sh;
    sh.nm1 = np - 1;
    if (sh.nm1 == 0) {
        sh.rounds = 1;
        sh.pstep = kg;
    } else if (sh.nm1 < kg) {
        sh.rounds = 1;
        sh.pstep = (kg - 1) / sh.nm1;
    } else {
        sh.rounds = min(kRoundsMax, (2 * sh.nm1 + kg) / kg);
        sh.pstep = (kg * sh.rounds - 1) / sh.nm1;
    }
    return sh;
}

// Append a walk to the fix-up list.  The counter is 64-bit, so it cannot wrap
// within a launch; entries past the capacity only raise the overflow flag.
__device__ inline void push_fixup(const DepositArgs &A, long long k, int item, int z0, int count, int dir)
{
    const unsigned long long slot = atomicAdd(&A.fix_count[0], 1ull);
    if (slot < (unsigned long long)A.fix_cap)
        A.fix[slot] = Fixup{k, item, z0, count, dir};
    else
        atomicOr(&A.fix_count[1], 1ull);
}

// Lanes of a wave are split into groups of kG lanes; each group walks its own
// (pair, direction) stream, kG consecutive pixels per step.
template <int kG>
__device__ inline unsigned long long group_bits(unsigned long long m, int lane)
{
    if (kG == 64) return m;
    return (m >> ((lane / kG) * kG)) & ((1ull << kG) - 1ull);
}

struct LaneStats {
    long long pairs = 0, pixels = 0, px64 = 0, pairs64 = 0, samples = 0, evaluated = 0, sums = 0, far = 0;
    // wave cycles (s_memtime, counted by lane 0 of each wavefront) in pair
    // set-up, far-wing set-up and walk steps, wave iterations, and the
    // iterations in which at least one group set up a pair
    long long cyc_setup = 0, cyc_wing = 0, cyc_walk = 0, iters = 0, setup_iters = 0;
};

// Per-pair scalars of LineAbsorption::add_tau_particle (absorption.cpp:218-247)
// and SingleAbsorber's constructor (singleabs.h:81-90).
struct PairConsts {
    double vel, b, inv_b, dvk, vhigh, vdr2, vsmooth, aa, scale, scale_nd, ufar;
    float temp;
    int zmax, np, fp64;
};

// Pair kk (the particle at pair_part[kk]).  Returns false if the particle does
// not reach the sightline (absorption.cpp:232; on the Voronoi mesh, if it owns
// no segment of it, absorption.cpp:221-229).  fp64 routes the pair to the
// FP64 path.
__device__ inline bool pair_consts(const DepositArgs &A, int line, const LineDesc &L, const LosInfo &los,
                                   long long kk, int max_sums, PairConsts &C)
{
    const int p = A.pair_part[kk];
    const float4 r0 = *(const float4 *)&A.rec[p].x;
    const float4 r1 = *(const float4 *)&A.rec[p].vx;
    const float pp[3] = {r0.x, r0.y, r0.z};
    const float temp = r1.w;
    const float dens = A.rec[p].dens[line];
    const float ppos = pp[los.ax - 1];
    const float pvel = los.ax == 1 ? r1.x : (los.ax == 2 ? r1.y : r1.z);
    const double b = L.bfac * sqrt((double)temp);
    double pos1 = (double)ppos, vdr2, vsmooth, vhigh;
    if (A.seg) {
        // Voronoi mesh: the segment ends take the places of dr2 and h
        // (part_int.cpp:42-43), the segment centre that of the position and
        // the kernel spans half the segment (absorption.cpp:221-237,
        // singleabs.h:84-88)
        const float2 sg = A.seg[kk];
        const double lo = (double)sg.x;
        if (lo > 2 * A.vbox / A.velfac || (double)sg.y > 2 * A.vbox / A.velfac) return false;
        pos1 = (lo + (double)sg.y) / 2.;
        vdr2 = A.velfac * lo;
        vsmooth = A.velfac * (double)sg.y;
        vhigh = (vdr2 > 0 && vsmooth > 0) ? (vsmooth - vdr2) / 2. : 0;
    } else {
        const double dr2 = pair_dr2(los, pp, A.box);
        const float smooth = r0.w;
        const float sm2 = smooth * smooth;
        if ((double)sm2 - dr2 <= 0) return false;            // absorption.cpp:232
        double val1 = A.velfac * dr2;
        val1 *= A.velfac;
        vdr2 = val1;
        vsmooth = A.velfac * (double)smooth;
        vhigh = vsmooth * vsmooth > vdr2 ? sqrt(vsmooth * vsmooth - vdr2) : 0;   // singleabs.h:83
    }
    const double vel = A.velfac * pos1 + (double)pvel;
    const double inv_b = 1.0 / b;
    // a = voigt_fac/b, amp = sigma_a/sqrt(pi) (c/1e5)/b (absorption.cpp:238,244),
    // by one reciprocal: within 2 ulp of the reference's quotients
    const double aa = L.voigt_fac * inv_b;
    const double bintov = A.vbox / A.nbins;
    const double amp = L.sigma_a * (kLight / 1e5 / kSqrtPi) * inv_b;
    const double dvk = 2. * vhigh / kNGrid;
    const double xs = bintov * inv_b;     // bintov / (b/2) / 2 (singleabs.h:114)
    int np = 1;
    if (!(bintov < b / 2.)) np = 2 * ceil(xs) + 1.;
    // The reference sizes the sub-sample grid per pixel from vhigh - vlow,
    // which equals bintov up to the rounding of vlow + bintov (~1e-13
    // relative); the walk uses one grid per pair, so a pair whose grid could
    // differ between pixels (bintov/b at a switch point to 1e-12) goes to the
    // FP64 path, which follows the reference pixel by pixel.
    const bool grid_edge = fabs(bintov - b / 2.) <= 1e-12 * bintov || (np > 1 && fabs(xs - rint(xs)) <= 1e-12 * xs);
    // pixels wider than max_sums sub-samples do not fit one walk step
    C.fp64 = A.force_fp64 || !(aa <= (double)kA32Max) || grid_edge || np - 1 > max_sums - 1;
    const double dmax = (vhigh - dvk) * inv_b + (np > 1 ? 0.5 * bintov * inv_b : 0.0);
    C.ufar = fmax(7.0 + dmax, 3.0 * dmax);
    C.vel = vel;
    C.b = b;
    C.inv_b = inv_b;
    C.dvk = dvk;
    C.vhigh = vhigh;
    C.vdr2 = vdr2;
    C.vsmooth = vsmooth;
    C.aa = aa;
    C.scale = amp * (double)dens * dvk * A.inv_velfac;
    C.scale_nd = amp * dvk * A.inv_velfac;
    C.temp = temp;
    C.zmax = (int)floor(vel / bintov);
    C.np = np;
    return true;
}

// Kernel weight of interior sample i (singleabs.h:145-152).
__device__ inline float sample_weight(const DepositArgs &A, const PairConsts &C, int i)
{
    const double vv = (i + 1) * C.dvk - C.vhigh;
    const double q = sqrt(C.vdr2 + vv * vv) / C.vsmooth;
    return (float)kernel_weight(A.kernel, q);
}

// Set up one pair in the wave's LDS slot; returns false if the particle does
// not reach the sightline.  Every lane of the wave calls it.
template <int kG>
__device__ inline bool setup_pair(const DepositArgs &A, int line, const LineDesc &L, const LosInfo &los, long long kk,
                                  PairTau &P, int lane, int max_sums = 1 << 30)
{
    PairConsts C;
    if (!pair_consts(A, line, L, los, kk, max_sums, C)) return false;
    const int p = A.pair_part[kk];
    const double bintov = A.vbox / A.nbins;
    wave_sync();   // the previous pair's constants are no longer read
    if (lane < kNGrid - 1) {
        set_sample_offset(P.smp[lane], ((lane + 1) * C.dvk - C.vhigh) * C.inv_b);
        P.smp[lane].w = sample_weight(A, C, lane);
    }
    if (lane == 0) {
        P.vel = C.vel;
        P.b = C.b;
        P.inv_b = C.inv_b;
        P.dvk = C.dvk;
        P.vhigh = C.vhigh;
        P.vdr2 = C.vdr2;
        P.vsmooth = C.vsmooth;
        P.aa = C.aa;
        P.zmax = C.zmax;
        P.fp64 = C.fp64;
        P.va = make_voigt_a(C.fp64 ? 0.0 : C.aa);
        P.scale = C.scale;
        P.scale_nd = C.scale_nd;
        P.temp = (double)C.temp;
        P.part = p;
        P.np = C.np;
        P.ufar = C.ufar;
        P.ufar2 = (float)(C.ufar * C.ufar);
        P.wing_ready = 0;
        P.inv_nm1 = C.np > 1 ? 1.0 / (C.np - 1) : 1.0;
        P.dv = C.np > 1 ? bintov * P.inv_nm1 : 0.0;
        const StepShape sh = step_shape(C.np, kG);
        P.nm1 = sh.nm1;
        P.rounds = sh.rounds;
        P.pstep = sh.pstep;
    }
    wave_sync();
    return true;
}

// The main kernel sets pairs up in batches, one pair per lane of a
// wavefront, into a per-wavefront ring of compact records: the FP64 set-up
// (pair constants, kernel weights and the ten far-wing coefficients) runs on
// up to kWaveQueue lanes at once instead of on the few lanes of one lane
// group, and a group taking the next record only expands it into its PairTau
// slot.  The ring is shared by the wavefront's groups.
constexpr int kWaveQueue = kWaveClaim;

struct PairLite {
    WalkHead h;
    double vhigh;
    long long k;      // pair index
    float W[4];       // kernel weights of the centre sample and the 3 above it
    float Cw[kWingTerms];   // far-wing coefficients (wing_coeffs)
    VoigtA va;        // constants of the FP32 profile
};
static_assert(sizeof(PairLite) == 168, "PairLite layout");
static_assert(sizeof(VoigtA) == kWingTerms * sizeof(float), "the activation copies VoigtA with the Cw lanes");

// The far-wing coefficients of wing_setup for one pair, all ten on one lane
// (the batched set-up of the main kernel), in FP32: every term is positive,
// so the coefficients keep ~4e-7 of the FP64 values and the far-wing sum
// ~4e-7 (checked against the FP64 computation over sub-sample counts 1..255,
// kernel extents 0..30 b and a <= 0.04).  c0 = vhigh/b and h = half the
// pixel width over b, both in units of ufar; W the kernel weights at offsets
// 0, 1/4, 1/2, 3/4 of vhigh; m1 = sub-samples - 1 (at most FSAMD_WING_TMAX).
__device__ __noinline__ void wing_coeffs(float c0, float h, int m1, float W0, float W1, float W2, float W3, float a,
                                         float ufar, float *Cw)
{
    float Y[kWingTerms];
    if (m1 == 0) {
        Y[0] = 1.0f;
#pragma unroll
        for (int n = 1; n < kWingTerms; ++n) Y[n] = 0.0f;
    } else {
        const float h2 = h * h;
        float pw = 1.0f;
#pragma unroll
        for (int n = 0; n < kWingTerms; ++n) {
            Y[n] = pw * (float)c_wing_T[m1 * kWingTerms + n];
            pw *= h2;
        }
    }
    // kernel samples at +-(c0/4, c0/2, 3c0/4) and 0
    const float x1 = 0.0625f * c0 * c0, x2 = 0.25f * c0 * c0, x3 = 0.5625f * c0 * c0;
    const float w1 = 2.0f * W1, w2 = 2.0f * W2, w3 = 2.0f * W3;
    float X[kWingTerms];
    float p1 = 1.0f, p2 = 1.0f, p3 = 1.0f;
#pragma unroll
    for (int k = 0; k < kWingTerms; ++k) {
        X[k] = w1 * p1 + w2 * p2 + w3 * p3 + (k == 0 ? W0 : 0.0f);
        p1 *= x1;
        p2 *= x2;
        p3 *= x3;
    }
    float m[kWingTerms];
#pragma unroll
    for (int n = 0; n < kWingTerms; ++n) {
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k <= n; ++k) t += (float)c_binom_even[n * kWingTerms + k] * X[k] * Y[n - k];
        m[n] = t;
    }
    const float iu = 1.0f / ufar, iu2 = iu * iu, u2 = ufar * ufar, a2 = a * a;
    float lead = a * kInvSqrtPiF * iu2;   // a/sqrt(pi) ufar^(-2n-2) at n = 0
#pragma unroll
    for (int n = 0; n < kWingTerms; ++n) {
        float cn = 0.0f, gs = lead;
#pragma unroll
        for (int j = 0; j <= n; ++j) {
            if (j > 0) gs *= u2;
            const int k = n - j;
            const float qk = k > 0 ? (float)c_wing_q[k - 1] : 0.0f;
            cn += gs * ((float)c_wing_p[k] - a2 * qk) * (float)c_binom_odd[n * kWingTerms + j] * m[j];
        }
        Cw[n] = cn;
        lead *= iu2;
    }
}

// Claim up to `want` pairs of item w for the wavefront, set up one per lane,
// and append those walked in FP32 to the wavefront's ring (tail advances);
// FP64 pairs go straight to the fix-up list.  Every lane of the wavefront
// calls it; `more` says whether the item has pairs left.
template <int kG, bool kStats>
__device__ inline void refill_wave(const DepositArgs &A, const WorkItem &w, int gitem, const LineDesc &L,
                                   const LosInfo &los, unsigned long long *next, PairLite *q, int lane, int want,
                                   int &tail, bool &more, LaneStats &st)
{
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(next, (unsigned long long)want);
    base = (unsigned long long)__shfl((long long)base, 0, 64);
    const long long kk = (long long)base + lane;
    const double bintov = A.vbox / A.nbins;
    bool enq = false;
    PairConsts C;
    if (lane < want && kk < w.p1 && pair_consts(A, w.line, L, los, kk, kG * kRoundsMax, C)) {
        if (kStats) st.pairs++;
        if (C.fp64) {
            if (kStats) st.pairs64++;
            push_fixup(A, kk, gitem, C.zmax, A.nbins / 2, 0);
        } else {
            enq = true;
        }
    }
    const unsigned long long m = __ballot(enq);
    wave_sync();   // the records taken earlier have been read
    if (enq) {
        PairLite &r = q[(tail + __popcll(m & ((1ull << lane) - 1ull))) % kWaveQueue];
        r.h.vel = C.vel;
        r.h.inv_b = C.inv_b;
        r.h.scale = C.scale;
        r.vhigh = C.vhigh;
        r.va = make_voigt_a(C.aa);
        // the walk uses the light form of the profile for a <= kA3Drop, as a
        // property of the pair: its coefficients are zero, so the full form
        // gives the same values and the result does not depend on which
        // pairs share a wavefront
        if (C.aa <= (double)kA3Drop) {
            r.va.a3_sqpi3 = 0.0f;
            r.va.lc2 = 0.0f;
            r.va.lc3 = 0.0f;
        }
        r.h.inv_nm1 = C.np > 1 ? 1.0 / (C.np - 1) : 1.0;
        r.k = kk;
        const double inv_vs = 1.0 / C.vsmooth;
        float W[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double vv = (i + 4) * C.dvk - C.vhigh;
            W[i] = (float)kernel_weight(A.kernel, sqrt(C.vdr2 + vv * vv) * inv_vs);
            r.W[i] = W[i];
        }
        // the walk compares with ufar rounded to FP32, the unit of the coefficients
        const float ufar = (float)C.ufar;
        r.h.ufar = (double)ufar;
        r.h.zmax = C.zmax;
        const StepShape sh = step_shape(C.np, kG);
        r.h.nm1 = sh.nm1;
        r.h.rounds = sh.rounds;
        r.h.pstep = sh.pstep;
        // far-wing coefficients, in units of that ufar
        const double iu = 1.0 / (double)ufar;
        wing_coeffs((float)(C.vhigh * C.inv_b * iu), (float)(0.5 * bintov * C.inv_b * iu), sh.nm1, W[0], W[1], W[2],
                    W[3], (float)C.aa, ufar, r.Cw);
    }
    wave_sync();
    tail += __popcll(m);
    more = (long long)base + want < w.p1;
}

// Expand queued record r into the group's PairTau slot: the fields the main
// kernel's walk reads, one 8-byte word, weight, coefficient or sample offset
// per lane (setup_pair and wing_setup fill the same values for the other
// kernels).  Only the lanes of the groups taking a record call it.
template <bool kStats>
__device__ inline void activate_pair(const PairLite &r, PairTau &P, int glane)
{
    if (glane < kNGrid - 1) {
        const double dvk = 2. * r.vhigh / kNGrid;
        set_sample_offset(P.smp[glane], ((glane + 1) * dvk - r.vhigh) * r.h.inv_b);
        P.smp[glane].w = r.W[glane < 3 ? 3 - glane : glane - 3];   // K(q) is even in the sample offset
    }
    if (glane < kWingTerms) {
        P.Cw[glane] = r.Cw[glane];
        reinterpret_cast<float *>(&P.va)[glane] = reinterpret_cast<const float *>(&r.va)[glane];
    }
    constexpr int kHeadWords = sizeof(WalkHead) / sizeof(long long);
    if (glane < kHeadWords)
        reinterpret_cast<long long *>(static_cast<WalkHead *>(&P))[glane] =
            reinterpret_cast<const long long *>(&r.h)[glane];
    if (kStats && glane == kHeadWords) {
        // read by the counters (pixel_samples)
        P.b = 1.0 / r.h.inv_b;
    }
}

// Voigt samples the reference evaluates for one pixel (singleabs.h:110-125).
__device__ inline int pixel_samples(const PairTau &P, double vlow, double bintov)
{
    const double vhighp = vlow + bintov;
    const double width = vhighp - vlow;
    if (width < P.b / 2.) return kNGrid - 1;
    const int np = 2 * ceil(width / (P.b / 2.) / 2) + 1.;
    return (kNGrid - 1) * np;
}

template <bool kLds>
__device__ inline void row_add(double *p, double v)
{
    if (kLds)
        atomicAdd(p, v);          // ds_add_f64
    else
        unsafeAtomicAdd(p, v);    // global_atomic_add_f64
}

// Walk modes: the FP32 deposit kernel defers any decision it cannot make
// exactly; the fix-up kernel re-evaluates guard-band pixels in FP64, or
// evaluates whole pairs in FP64.
enum { kWalkMain = 0, kWalkGuarded = 1, kWalkFull64 = 2 };



// One step of kG consecutive pixels of a walk (lane glane of the group), each
// pixel evaluated on its own as tau_kern_outer does (singleabs.h:104-126,
// sub-sample grid from the pixel's own vhigh - vlow): the fix-up kernel's
// walk (walk_fix), in the guarded or the all-FP64 mode.  Returns true when the
// walk is over (tail cut reached, end of the window, or in kWalkMain mode the
// decision deferred).
template <int kG, int kMode, bool kLds, bool kStats>
__device__ inline bool walk_chunk(const DepositArgs &A, const PairTau &P, double *row, const float *gtab, int z0,
                                  int dir, int count, int c, int lane, LaneStats &st, long long k, int item,
                                  bool active, int &walked)
{
    const int glane = lane % kG;
    const double bintov = A.vbox / A.nbins;
    const double tt = A.tautail;
    const int idx = c + glane;
    const bool valid = active && idx < count;
    const int z = z0 + dir * idx;
    double t = 0, vlow = 0;
    bool guard = false;
    if (valid) {
        vlow = z * bintov - P.vel;                                  // absorption.cpp:252
        t = taulast_pixel<kMode == kWalkFull64>(P, vlow, bintov, A.velfac, A.kernel, gtab);
        if (kStats) {
            st.evaluated++;
            if (kMode != kWalkFull64 && fabs((vlow + 0.5 * bintov) * P.inv_b) >= P.ufar)
                st.far++;
            else
                st.sums += pixel_samples(P, vlow, bintov) / (kNGrid - 1);
        }
        if (kMode != kWalkFull64 && fabs(t - tt) <= kGuard * tt) {
            if (kMode == kWalkGuarded) {
                t = taulast_pixel<true>(P, vlow, bintov, A.velfac, A.kernel, gtab);
                if (kStats) st.px64++;
            } else {
                guard = true;
            }
        }
    }
    const unsigned long long mb = group_bits<kG>(__ballot(valid && !guard && t < tt), lane);
    const unsigned long long mg = group_bits<kG>(__ballot(guard), lane);
    const unsigned long long mall = mb | mg;
    const int first = mall ? __builtin_ctzll(mall) : kG;
    if (valid && (glane < first || (glane == first && ((mb >> glane) & 1ull)))) {
        row_add<kLds>(&row[wrap_pixel(z, A.nbins)], t);
        if (kStats) {
            st.pixels++;
            st.samples += pixel_samples(P, vlow, bintov);
        }
    }
    if (kMode == kWalkMain && first < kG && ((mg >> first) & 1ull)) {
        // the FP32 value of the deciding pixel is too close to tautail
        if (active && glane == 0) push_fixup(A, k, item, z0 + dir * (c + first), count - (c + first), dir);
        walked = -1;
        return true;
    }
    if (mb != 0) {
        walked = c + first + 1;
        return true;
    }
    if (c + kG >= count) {
        walked = count;
        return true;
    }
    return false;
}

// Far steps cover this many groups of kG pixels (the moment expansion is a
// few FMAs per pixel, so a far walk advances several sub-steps while the
// other groups of the wavefront evaluate their near pixels).
constexpr int kFarRounds = 4;

// End of one sub-step of a walk (lanes glane < pcount hold the pixels
// cstep + glane of the walk, value t): deposit up to and including the first
// pixel below tautail (absorption.cpp:260-263), or hand the walk to the
// fix-up kernel when that decision falls in the guard band.  Returns true
// when the walk is over, with `walked` set.
template <int kG, bool kLds, bool kStats>
__device__ inline bool settle(const DepositArgs &A, const PairTau &P, double *row, int lane, bool valid, double t,
                              int z, double vlow, int z0, int dir, int cstep, int pcount, int count, LaneStats &st,
                              long long k, int item, int &walked)
{
    const int glane = lane % kG;
    const double tt = A.tautail;
    const bool guard = valid && fabs(t - tt) <= kGuard * tt;
    const unsigned long long mb = group_bits<kG>(__ballot(valid && !guard && t < tt), lane);
    const unsigned long long mg = group_bits<kG>(__ballot(guard), lane);
    const unsigned long long mall = mb | mg;
    const int first = mall ? __builtin_ctzll(mall) : kG;
    if (valid && (glane < first || (glane == first && ((mb >> glane) & 1ull)))) {
        row_add<kLds>(&row[wrap_pixel(z, A.nbins)], t);
        if (kStats) {
            st.pixels++;
            st.samples += pixel_samples(P, vlow, A.vbox / A.nbins);
        }
    }
    if (first < kG && ((mg >> first) & 1ull)) {
        // the FP32 value of the deciding pixel is too close to tautail
        if (glane == 0) push_fixup(A, k, item, z0 + dir * (cstep + first), count - (cstep + first), dir);
        walked = -1;
        return true;
    }
    if (mb != 0) {
        walked = cstep + first + 1;
        return true;
    }
    if (cstep + pcount >= count) {
        walked = count;
        return true;
    }
    return false;
}

// One step of a walk in the main kernel (lane glane of a kG-lane group):
// the next pixels z0 + dir*(c + p) of the walk.  Pixels whose centre lies in
// the far wing take the moment expansion: a step is a far step when its
// first pixel is (|U| only grows along a walk) and then covers kFarRounds
// sub-steps of kG pixels.  A near step covers P.pstep pixels on the shared
// sub-sample grid; `sbuf` is the group's LDS scratch of kG * kRoundsMax
// floats.  Returns true when the walk is over (tail cut, end of the window,
// or the decision deferred to the fix-up kernel); `adv` is the number of
// pixels the step covered.
template <int kG, bool kLds, bool kStats>
__device__ inline bool walk_step(const DepositArgs &A, const PairTau &P, float *sbuf, double *row,
                                 const float *gtab, int z0, int dir, int count, int c, bool far, int lane,
                                 LaneStats &st, long long k, int item, bool active, int &walked, int &adv)
{
    adv = 0;
    if (!active) return false;
    const int glane = lane % kG;
    const double bintov = A.vbox / A.nbins;
    if (far) {
        for (int r = 0; r < kFarRounds; ++r) {
            const int cc = c + r * kG;
            const int pc = min(kG, count - cc);
            const bool valid = glane < pc;
            const int z = z0 + dir * (cc + glane);
            const double vlow = z * bintov - P.vel;
            double t = 0;
            if (valid) {
                const double U = (vlow + 0.5 * bintov) * P.inv_b;
                const float Y = (float)(P.ufar * P.ufar) * __builtin_amdgcn_rcpf((float)(U * U));
                float S = P.Cw[kWingTerms - 1];
#pragma unroll
                for (int q = kWingTerms - 2; q >= 0; --q) S = __builtin_fmaf(S, Y, P.Cw[q]);
                t = P.scale * (double)(S * Y);
                if (kStats) {
                    st.evaluated++;
                    st.far++;
                }
            }
            if (settle<kG, kLds, kStats>(A, P, row, lane, valid, t, z, vlow, z0, dir, cc, pc, count, st, k, item,
                                         walked))
                return true;
        }
        adv = kFarRounds * kG;
        return false;
    }
    const int nm1 = P.nm1;
    const int pcount = min(P.pstep, count - c);
    // pairs with a <= kA3Drop carry the light form of the profile (zero a^3
    // and exponent coefficients); when every pair of this near step of the
    // wavefront does, the light code runs (a wavefront-uniform branch with
    // the same results)
    const bool a3 = __ballot(P.va.a3_sqpi3 != 0.0f) != 0ull;
    wave_sync();   // the previous step's reads of sbuf are done
    {
        const int nsub = nm1 == 0 ? pcount : pcount * nm1 + 1;
        const int zlo = dir > 0 ? z0 + c : z0 - c - (pcount - 1);
        const double vlo = zlo * bintov - P.vel;   // vlow of the step's lowest pixel (absorption.cpp:252)
        for (int r = 0; r < P.rounds; ++r) {
            const int m = glane + r * kG;
            if (m < nsub) {
                double v;
                if (nm1 == 0) {
                    const double vlow = (z0 + dir * (c + m)) * bintov - P.vel;   // absorption.cpp:252
                    v = ((vlow + bintov) + vlow) / 2.;                           // singleabs.h:113
                } else {
                    // sub-sample m of the step, in increasing velocity: sub-sample
                    // m mod (np-1) of its pixel (singleabs.h:118-121), up to the
                    // rounding of m dv against the pixel's own vlow
                    v = __builtin_fma((double)m, bintov * P.inv_nm1, vlo);
                }
                sbuf[m] = (float)(a3 ? inner32<true>(P, v, gtab) : inner32<false>(P, v, gtab));
                if (kStats) st.sums++;
            }
        }
    }
    wave_sync();
    const bool valid = glane < pcount;
    const int z = z0 + dir * (c + glane);
    const double vlow = z * bintov - P.vel;
    double t = 0;
    if (valid) {
        if (nm1 == 0) {
            t = P.scale * (double)sbuf[glane];
        } else {
            // trapezoid over the pixel's sub-samples (singleabs.h:116-125)
            const int j = dir > 0 ? glane : pcount - 1 - glane;
            const float *sp = sbuf + j * nm1;
            double tot = 0.5 * (double)sp[0];
            for (int kk = 1; kk < nm1; ++kk) tot += (double)sp[kk];
            tot += 0.5 * (double)sp[nm1];
            t = P.scale * (tot * P.inv_nm1);
        }
        if (kStats) st.evaluated++;
    }
    adv = pcount;
    return settle<kG, kLds, kStats>(A, P, row, lane, valid, t, z, vlow, z0, dir, c, pcount, count, st, k, item,
                                    walked);
}

// Record the pixels walked in one direction of pair k (for the adjoint).
__device__ inline void record_window(const DepositArgs &A, int line, long long k, int dir, int len)
{
    if (A.win) A.win[((long long)line * A.npairs + k) * 2 + (dir < 0 ? 1 : 0)] = len;
}

// Whether the next kg pixels of a fix-up walk reach pixels that use the
// far-wing expansion (|U| only grows along a walk after its first pixel).
__device__ inline bool step_needs_wing(const PairTau &P, double bintov, int z0, int dir, int c, int kg, int count)
{
    const int last = min(c + kg - 1, count - 1);
    const double U0 = ((z0 + dir * c) * bintov - P.vel + 0.5 * bintov) * P.inv_b;
    const double U1 = ((z0 + dir * last) * bintov - P.vel + 0.5 * bintov) * P.inv_b;
    return fmax(fabs(U0), fabs(U1)) >= P.ufar;
}

// Walk one direction with a kG-lane group (fix-up kernel): pixels z0,
// z0+dir, ... for at most `count` pixels, until and including the tail cut.
template <int kG, int kMode, bool kStats>
__device__ inline int walk_fix(const DepositArgs &A, PairTau &P, double *row, const float *gtab, int z0, int dir,
                               int count, int lane, LaneStats &st, long long k, int item)
{
    const double bintov = A.vbox / A.nbins;
    const int glane = lane % kG;
    int walked = 0;
    for (int c = 0; c < count; c += kG) {
        if (kMode != kWalkFull64 && !P.wing_ready && step_needs_wing(P, bintov, z0, dir, c, kG, count)) {
            wing_setup<kG>(P, bintov, glane);
            if (glane == 0) P.wing_ready = 1;
            wave_sync();
        }
        if (walk_chunk<kG, kMode, false, kStats>(A, P, row, gtab, z0, dir, count, c, lane, st, k, item, true, walked))
            return walked;
    }
    return walked;
}

// Lanes per fix-up entry: most deferred walks end a few pixels past the
// deciding pixel, so a wavefront works on four entries at once.
constexpr int kFixG = 16;

// 4 waves per SIMD: measured 8% faster at c1 than the 3 the register count
// allows by default (a few dwords of spill in the pair setup).  kStats adds
// the per-lane event counts (A.stats set).
template <int kG, bool kLds, bool kStats, int kBlock>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void tau_kernel(DepositArgs A, const WorkItem *items, int item_base,
                                                       int use_queue)
{
    constexpr int kGroups = kBlock / kG;
    constexpr int kWavesPerBlock = kBlock / 64;
    constexpr int kGroupsPerWave = 64 / kG;
    extern __shared__ double smem[];
    __shared__ __attribute__((aligned(16))) float gtab[kGTableSize];
    __shared__ PairTau pslot[kGroups];
    __shared__ float sub_sums[kBlock * kRoundsMax];
    __shared__ PairLite wqueue[kWavesPerBlock * kWaveQueue];
    __shared__ int s_item;
    __shared__ unsigned long long s_next;
    for (int i = threadIdx.x; i < kGTableSize; i += kBlock) gtab[i] = c_gtab[i];
    const int lane = threadIdx.x & 63;
    const int glane = lane % kG;
    const int gid = threadIdx.x / kG;
    PairTau &P = pslot[gid];
    float *sbuf = sub_sums + gid * kG * kRoundsMax;
    PairLite *q = wqueue + (threadIdx.x / 64) * kWaveQueue;
    // the wavefront's leader lanes of its groups, below this lane's group
    const unsigned long long leaders = kG == 64 ? 1ull : (kG == 32 ? 0x0000000100000001ull :
                                                         0x0001000100010001ull);
    const unsigned long long leaders_below = leaders & ((1ull << ((lane / kG) * kG)) - 1ull);
    const double bintov = A.vbox / A.nbins;
    const int half = A.nbins / 2;
    LaneStats st;
    int it = use_queue ? 0 : blockIdx.x;
    for (;;) {
        int item;
        __syncthreads();
        if (use_queue) {
            if (threadIdx.x == 0) s_item = atomicAdd(A.queue, 1);
            __syncthreads();
            item = s_item;
        } else {
            item = it;
            it += gridDim.x;
        }
        if (item >= A.nitems) break;
        const int gitem = item_base + item;
        const WorkItem w = items[gitem];
        const LineDesc L = line_of(A, w.line);
        double *out_row = L.out + (size_t)w.los * A.nbins;
        double *row = kLds ? smem : out_row;
        if (kLds)
            for (int j = threadIdx.x; j < A.nbins; j += kBlock) row[j] = 0.0;
        if (threadIdx.x == 0) s_next = (unsigned long long)w.p0;
        __syncthreads();
        const LosInfo los = los_info(A, w.los);
        // group state: dir 0 = needs a pair, +1/-1 = walking, 2 = no pairs left
        long long k = 0;
        int dir = 0, c = 0, z0 = 0;
        // the wavefront's ring of set-up pairs: records head .. tail - 1
        int head = 0, tail = 0;
        bool more = true;
        for (;;) {
            long long tc0 = 0;
            const unsigned long long needing = __ballot(dir == 0) & leaders;
            if (kStats) {
                tc0 = clock64();
                if (lane == 0) {
                    st.iters++;
                    st.setup_iters += needing != 0ull;
                }
            }
            if (needing) {
                const int nneed = __popcll(needing);
                if (tail - head < nneed && more) {
                    // claim up to a ring's worth, fewer near the end of the item so
                    // that the block's wavefronts finish together
                    const long long rem = (long long)w.p1 - (long long)__atomic_load_n(&s_next, __ATOMIC_RELAXED);
                    const long long share = rem / (2 * kWavesPerBlock);
                    int want = kWaveQueue - (tail - head);
                    if (share < want) want = (int)(share > 2 * kGroupsPerWave ? share : 2 * kGroupsPerWave);
                    if (A.claim < want) want = A.claim;
                    if (want < 1) want = 1;
                    want = __builtin_amdgcn_readfirstlane(want);
                    refill_wave<kG, kStats>(A, w, gitem, L, los, &s_next, q, lane, want, tail, more, st);
                }
                const int avail = tail - head;
                const int rank = __popcll(needing & leaders_below);
                bool fresh = false;
                wave_sync();   // the previous pairs' constants are no longer read
                if (dir == 0) {
                    if (rank < avail) {
                        const PairLite &r = q[(head + rank) % kWaveQueue];
                        activate_pair<kStats>(r, P, glane);
                        k = r.k;
                        dir = 1;
                        c = 0;
                        fresh = true;
                    } else if (!more) {
                        dir = 2;
                    }
                }
                wave_sync();
                if (fresh) z0 = P.zmax;
                head += nneed < avail ? nneed : avail;
            }
            const bool walking = (dir == 1 || dir == -1);
            if (!__ballot(dir != 2)) break;
            long long tc1 = 0;
            if (kStats) tc1 = clock64();
            // the step's first pixel in the far wing: |U| only grows along a walk
            // (the record brought the far-wing coefficients)
            const bool far = walking && fabs(((z0 + dir * c) * bintov - P.vel + 0.5 * bintov) * P.inv_b) >= P.ufar;
            long long tc2 = 0;
            if (kStats) tc2 = clock64();
            int walked = 0, adv = 0;
            const bool done = walk_step<kG, kLds, kStats>(A, P, sbuf, row, gtab, z0, dir, half, c, far, lane, st, k,
                                                          gitem, walking, walked, adv);
            if (kStats && lane == 0) {
                const long long tc3 = clock64();
                st.cyc_setup += tc1 - tc0;
                st.cyc_wing += tc2 - tc1;
                st.cyc_walk += tc3 - tc2;
            }
            if (walking) {
                if (done && walked >= 0 && glane == 0) record_window(A, w.line, k, dir, walked);
                if (!done) {
                    c += adv;
                } else if (dir == 1) {
                    dir = -1;
                    z0 = P.zmax - 1;
                    c = 0;
                } else {
                    dir = 0;
                }
            }
        }
        if (kLds) {
            __syncthreads();
            if (w.direct) {
                for (int j = threadIdx.x; j < A.nbins; j += kBlock) out_row[j] = row[j];
            } else {
                for (int j = threadIdx.x; j < A.nbins; j += kBlock) {
                    const double v = row[j];
                    if (v != 0.0) unsafeAtomicAdd(&out_row[j], v);
                }
            }
        }
    }
    if (kStats && A.stats) {
        atomicAdd((unsigned long long *)&A.stats[0], (unsigned long long)st.pairs);
        atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)st.pixels);
        atomicAdd((unsigned long long *)&A.stats[3], (unsigned long long)st.pairs64);
        atomicAdd((unsigned long long *)&A.stats[4], (unsigned long long)st.samples);
        atomicAdd((unsigned long long *)&A.stats[5], (unsigned long long)st.evaluated);
        atomicAdd((unsigned long long *)&A.stats[6], (unsigned long long)st.sums);
        atomicAdd((unsigned long long *)&A.stats[7], (unsigned long long)st.far);
        if (lane == 0) {
            atomicAdd((unsigned long long *)&A.stats[8], (unsigned long long)st.cyc_setup);
            atomicAdd((unsigned long long *)&A.stats[9], (unsigned long long)st.cyc_wing);
            atomicAdd((unsigned long long *)&A.stats[10], (unsigned long long)st.cyc_walk);
            atomicAdd((unsigned long long *)&A.stats[11], (unsigned long long)st.iters);
            atomicAdd((unsigned long long *)&A.stats[12], (unsigned long long)st.setup_iters);
        }
    }
}

// Deferred walks, after the deposit kernel has written its rows: one wave per
// fix-up entry, deposits straight into the output with FP64 atomics.
template <bool kStats>
__global__ __launch_bounds__(kThreads) void tau_fixup_kernel(DepositArgs A, const WorkItem *items)
{
    constexpr int kGroups = kThreads / kFixG;
    __shared__ __attribute__((aligned(16))) float gtab[kGTableSize];
    __shared__ PairTau pslot[kGroups];
    for (int i = threadIdx.x; i < kGTableSize; i += kThreads) gtab[i] = c_gtab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, gid = threadIdx.x / kFixG, glane = lane % kFixG;
    PairTau &P = pslot[gid];
    LaneStats st;
    const int n = (int)min(A.fix_count[0], (unsigned long long)A.fix_cap);
    const int half = A.nbins / 2;
    for (int e = blockIdx.x * kGroups + gid; e < n; e += gridDim.x * kGroups) {
        const Fixup f = A.fix[e];
        const WorkItem w = items[f.item];
        const LineDesc L = line_of(A, w.line);
        double *out_row = L.out + (size_t)w.los * A.nbins;
        const LosInfo los = los_info(A, w.los);
        if (!setup_pair<kFixG>(A, w.line, L, los, f.k, P, glane)) continue;
        if (f.dir == 0) {
            const int lf = walk_fix<kFixG, kWalkFull64, kStats>(A, P, out_row, gtab, P.zmax, 1, half, lane, st, f.k,
                                                                  f.item);
            const int lb = walk_fix<kFixG, kWalkFull64, kStats>(A, P, out_row, gtab, P.zmax - 1, -1, half, lane, st,
                                                                  f.k, f.item);
            if (glane == 0) {
                record_window(A, w.line, f.k, 1, lf);
                record_window(A, w.line, f.k, -1, lb);
            }
        } else {
            const int l = walk_fix<kFixG, kWalkGuarded, kStats>(A, P, out_row, gtab, f.z0, f.dir, f.count, lane, st,
                                                                  f.k, f.item);
            if (glane == 0) record_window(A, w.line, f.k, f.dir, (half - f.count) + l);
        }
    }
    if (kStats && A.stats) {
        atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)st.pixels);
        atomicAdd((unsigned long long *)&A.stats[2], (unsigned long long)st.px64);
        atomicAdd((unsigned long long *)&A.stats[4], (unsigned long long)st.samples);
        atomicAdd((unsigned long long *)&A.stats[5], (unsigned long long)st.evaluated);
        atomicAdd((unsigned long long *)&A.stats[6], (unsigned long long)st.sums);
        atomicAdd((unsigned long long *)&A.stats[7], (unsigned long long)st.far);
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the optical depth.
//
// For a pair, taulast(z) = scale * S(z), S = sum_samples t_j W_i H(a, u_ij),
// u_ij = (vv_i - vouter_j)/b, vouter_j = z*bintov - vel + e_j, scale =
// amp dens dvk/velfac with amp ~ 1/b, b = bfac sqrt(T), a = voigt_fac/b.
// With g(z) = dL/dtau at the pixel:
//   dL/d dens = sum_z g scale/dens S
//   dL/d vel  = sum_z g scale/b SU,          SU = sum t W dH/du
//   dL/d T    = -sum_z g scale/(2T) (S + SF), SF = sum t W (u dH/du + a dH/da)
// over exactly the pixels the forward deposited (recorded windows); the
// discrete choices (window ends, pixel index, sub-sample count) are held
// fixed.  Far-wing pixels use the same moment expansion as the forward, with
// a second coefficient set for SF and the derivative of the first for SU.

__device__ inline void adj_inner32(const PairTau &P, double vouter, const float *gtab, float &S, float &SU,
                                   float &SF)
{
    const Split s = split_s(vouter * P.inv_b);
    const VoigtA va = P.va;
#pragma unroll  // full: the adjoint kernel 6% faster than unrolling by 2 at 128^3
    for (int i = 0; i < kNGrid - 1; ++i) {
        const KernelSample k = P.smp[i];
        const VoigtGrad g = voigt32_grad(va, sample_u(k, s), gtab);
        const float w = k.w;
        S = __builtin_fmaf(w, g.H, S);
        SU = __builtin_fmaf(w, g.Hu, SU);
        SF = __builtin_fmaf(w, g.F, SF);
    }
}

// The same three sums with the FP64 Faddeeva function and FP64 kernel weights,
// for the pairs the forward evaluates in FP64 (damping parameter above
// kA32Max, or a sub-sample count on a switch point).  With L = Im w:
//   dH/du = -2(uH - aL),  u dH/du + a dH/da = -2(u^2 - a^2)H + 4auL - 2a/sqrt(pi).
__device__ __noinline__ void adj_inner64(double aa, double b, double dvk, double vhigh, double vdr2, double vsmooth,
                                         double vouter, int kernel, double &S, double &SU, double &SF)
{
    for (int i = 1; i < kNGrid; ++i) {
        const double vv = i * dvk - vhigh;
        const double W = kernel_weight(kernel, sqrt(vdr2 + vv * vv) / vsmooth);
        const double u = (vv - vouter) / b;
        double H, L;
        w64(u, aa, H, L);
        S += W * H;
        SU += W * (-2.0 * (u * H - aa * L));
        SF += W * (-2.0 * (u * u - aa * aa) * H + 4.0 * aa * u * L - aa * kTwoOverSqrtPi);
    }
}

// One pixel of an FP64 pair: the sums averaged over the pixel's own
// sub-sample grid, as tau_kern_outer builds it (singleabs.h:104-126) and as
// the forward's FP64 walk evaluates it.
__device__ inline void adj_pixel64(const PairTau &P, double vlow, double bintov, int kernel, double &S, double &SU,
                                   double &SF)
{
    const double vhighp = vlow + bintov;
    const double width = vhighp - vlow;
    S = SU = SF = 0.0;
    if (width < P.b / 2.) {
        adj_inner64(P.aa, P.b, P.dvk, P.vhigh, P.vdr2, P.vsmooth, (vhighp + vlow) / 2., kernel, S, SU, SF);
        return;
    }
    const int np = 2 * ceil(width / (P.b / 2.) / 2) + 1.;
    const double dv = width / (np - 1);
    double s0 = 0, u0 = 0, f0 = 0;
    adj_inner64(P.aa, P.b, P.dvk, P.vhigh, P.vdr2, P.vsmooth, vlow, kernel, s0, u0, f0);
    adj_inner64(P.aa, P.b, P.dvk, P.vhigh, P.vdr2, P.vsmooth, vhighp, kernel, s0, u0, f0);
    S = 0.5 * s0;
    SU = 0.5 * u0;
    SF = 0.5 * f0;
    for (int i = 1; i < np - 1; ++i)
        adj_inner64(P.aa, P.b, P.dvk, P.vhigh, P.vdr2, P.vsmooth, i * dv + vlow, kernel, S, SU, SF);
    const double inv = 1.0 / (np - 1);
    S *= inv;
    SU *= inv;
    SF *= inv;
}

// Far-wing pixel (|U| >= ufar) of the adjoint: S, SU and SF by the moment
// expansion of the forward and its derivative.
__device__ inline void adj_far(const PairTau &P, double vlow, double bintov, double &S, double &SU, double &SF)
{
    const double U = (vlow + 0.5 * bintov) * P.inv_b;
    const float Y = P.ufar2 * __builtin_amdgcn_rcpf((float)(U * U));
    float s = P.Cw[kWingTerms - 1], d = kWingTerms * P.Cw[kWingTerms - 1], f = P.CwF[kWingTerms - 1];
#pragma unroll
    for (int t = kWingTerms - 2; t >= 0; --t) {
        s = __builtin_fmaf(s, Y, P.Cw[t]);
        d = __builtin_fmaf(d, Y, (t + 1.0f) * P.Cw[t]);
        f = __builtin_fmaf(f, Y, P.CwF[t]);
    }
    S = (double)(s * Y);
    SU = 2.0 / U * (double)(d * Y);
    SF = (double)(f * Y);
}

template <int kG>
__device__ inline double group_sum(double v)
{
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kG);
    return v;
}

template <int kG>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void tau_adjoint_kernel(
    DepositArgs A, const WorkItem *items, int use_queue)
{
    constexpr int kGroups = kThreads / kG;
    extern __shared__ double grow[];   // dL/dtau of the item's sightline and line
    __shared__ __attribute__((aligned(16))) float gtab[kGTableSize];
    __shared__ PairTau pslot[kGroups];
    __shared__ int s_item;
    __shared__ unsigned long long s_next;
    for (int i = threadIdx.x; i < kGTableSize; i += kThreads) gtab[i] = c_gtab[i];
    const int lane = threadIdx.x & 63;
    const int glane = lane % kG;
    const int gid = threadIdx.x / kG;
    const int gleader = (lane / kG) * kG;
    PairTau &P = pslot[gid];
    const double bintov = A.vbox / A.nbins;
    int it = use_queue ? 0 : blockIdx.x;
    for (;;) {
        int item;
        __syncthreads();
        if (use_queue) {
            if (threadIdx.x == 0) s_item = atomicAdd(A.queue, 1);
            __syncthreads();
            item = s_item;
        } else {
            item = it;
            it += gridDim.x;
        }
        if (item >= A.nitems) break;
        const WorkItem w = items[item];
        const LineDesc L = line_of(A, w.line);
        const double *grad_row = L.out + (size_t)w.los * A.nbins;
        for (int j = threadIdx.x; j < A.nbins; j += kThreads) grow[j] = grad_row[j];
        if (threadIdx.x == 0) s_next = (unsigned long long)w.p0;
        __syncthreads();
        const LosInfo los = los_info(A, w.los);
        for (;;) {
            unsigned long long kk = 0;
            if (glane == 0) kk = atomicAdd(&s_next, 1ull);
            kk = (unsigned long long)__shfl((long long)kk, gleader, 64);
            const bool have = (long long)kk < w.p1;
            if (!__ballot(have)) break;
            if (!have) continue;
            const long long k = (long long)kk;
            if (!setup_pair<kG>(A, w.line, L, los, k, P, glane)) continue;
            const long long wi = ((long long)w.line * A.npairs + k) * 2;
            const int lf = A.win[wi], lb = A.win[wi + 1];
            const int n = lf + lb;
            const int zlo = P.zmax - lb;   // the window: pixels zlo .. zlo + n - 1
            double aS = 0, aU = 0, aF = 0;
            if (P.fp64) {
                // the forward walked this pair with the FP64 profile, pixel by
                // pixel: tau_adjoint64_kernel does the same (the FP64 Faddeeva
                // code stays out of this kernel's registers)
                if (glane == 0 && n > 0) push_fixup(A, k, item, 0, 0, 0);
                continue;
            }
            // far-wing pixels (|U| >= ufar) sit at the two ends of the window
            int nfl = 0, nfr = 0;
            for (int c = 0; c < n; c += kG) {
                const int j = c + glane;
                const double U = ((zlo + j) * bintov - P.vel + 0.5 * bintov) * P.inv_b;
                const bool far = j < n && fabs(U) >= P.ufar;
                nfl += __popcll(group_bits<kG>(__ballot(far && U < 0), lane));
                nfr += __popcll(group_bits<kG>(__ballot(far && U > 0), lane));
            }
            if (nfl + nfr > 0) {
                if (!P.wing_ready) {
                    wing_setup<kG, true>(P, bintov, glane);
                    if (glane == 0) P.wing_ready = 1;
                    wave_sync();
                }
                for (int c = 0; c < nfl + nfr; c += kG) {
                    const int f = c + glane;
                    if (f < nfl + nfr) {
                        const int z = zlo + (f < nfl ? f : n - nfr + (f - nfl));
                        double S, SU, SF;
                        adj_far(P, z * bintov - P.vel, bintov, S, SU, SF);
                        const double g = grow[wrap_pixel(z, A.nbins)];
                        aS += g * S;
                        aU += g * SU;
                        aF += g * SF;
                    }
                }
            }
            // the other pixels on the shared sub-sample grid of tau_kern_outer
            // (singleabs.h:110-125): each grid point once, weighted by the
            // trapezoid weights and dL/dtau of the pixels it belongs to
            const int j0 = nfl, nn = n - nfl - nfr;
            if (nn > 0) {
                const int nm1 = P.np - 1;
                const int npts = nm1 == 0 ? nn : nn * nm1 + 1;
                for (int m = glane; m < npts; m += kG) {
                    double v, wgt;
                    if (nm1 == 0) {
                        const int z = zlo + j0 + m;
                        const double vlow = z * bintov - P.vel;   // absorption.cpp:252
                        v = ((vlow + bintov) + vlow) / 2.;        // singleabs.h:113
                        wgt = grow[wrap_pixel(z, A.nbins)];
                    } else {
                        const int jj = m / nm1, kk = m - jj * nm1;
                        const int z = zlo + j0 + jj;
                        const double vlow = z * bintov - P.vel;
                        v = kk == 0 ? vlow : kk * P.dv + vlow;    // singleabs.h:118-121
                        if (kk == 0) {
                            const double gl = jj > 0 ? grow[wrap_pixel(z - 1, A.nbins)] : 0.0;
                            const double gr = jj < nn ? grow[wrap_pixel(z, A.nbins)] : 0.0;
                            wgt = 0.5 * (gl + gr) * P.inv_nm1;
                        } else {
                            wgt = grow[wrap_pixel(z, A.nbins)] * P.inv_nm1;
                        }
                    }
                    float s = 0, su = 0, sf = 0;
                    adj_inner32(P, v, gtab, s, su, sf);
                    aS += wgt * s;
                    aU += wgt * su;
                    aF += wgt * sf;
                }
            }
            aS = group_sum<kG>(aS);
            aU = group_sum<kG>(aU);
            aF = group_sum<kG>(aF);
            if (glane == 0 && n > 0) {
                const int p = P.part;
                unsafeAtomicAdd(&L.grad_dens[p], P.scale_nd * aS);
                unsafeAtomicAdd(&A.grad_vel[3 * (size_t)p + los.ax - 1], P.scale * P.inv_b * aU);
                unsafeAtomicAdd(&A.grad_temp[p], -P.scale / (2.0 * P.temp) * (aS + aF));
            }
        }
    }
}

// The adjoint of the pairs the forward evaluated in FP64 (listed by
// tau_adjoint_kernel in the fix-up list): one kFixG-lane group per pair, the
// window's pixels with the FP64 Faddeeva function and FP64 kernel weights,
// dL/dtau read from HBM.
__global__ __launch_bounds__(kThreads) void tau_adjoint64_kernel(DepositArgs A, const WorkItem *items)
{
    constexpr int kGroups = kThreads / kFixG;
    __shared__ PairTau pslot[kGroups];
    const int lane = threadIdx.x & 63, gid = threadIdx.x / kFixG, glane = lane % kFixG;
    PairTau &P = pslot[gid];
    const double bintov = A.vbox / A.nbins;
    const int n = (int)min(A.fix_count[0], (unsigned long long)A.fix_cap);
    for (int e = blockIdx.x * kGroups + gid; e < n; e += gridDim.x * kGroups) {
        const Fixup f = A.fix[e];
        const WorkItem w = items[f.item];
        const LineDesc L = line_of(A, w.line);
        const double *grad_row = L.out + (size_t)w.los * A.nbins;
        const LosInfo los = los_info(A, w.los);
        if (!setup_pair<kFixG>(A, w.line, L, los, f.k, P, glane)) continue;
        const long long wi = ((long long)w.line * A.npairs + f.k) * 2;
        const int lf = A.win[wi], lb = A.win[wi + 1];
        const int nw = lf + lb;
        const int zlo = P.zmax - lb;   // the window: pixels zlo .. zlo + nw - 1
        double aS = 0, aU = 0, aF = 0;
        for (int j = glane; j < nw; j += kFixG) {
            const int z = zlo + j;
            double S, SU, SF;
            adj_pixel64(P, z * bintov - P.vel, bintov, A.kernel, S, SU, SF);   // absorption.cpp:252
            const double g = grad_row[wrap_pixel(z, A.nbins)];
            aS += g * S;
            aU += g * SU;
            aF += g * SF;
        }
        aS = group_sum<kFixG>(aS);
        aU = group_sum<kFixG>(aU);
        aF = group_sum<kFixG>(aF);
        if (glane == 0) {
            const int p = P.part;
            unsafeAtomicAdd(&L.grad_dens[p], P.scale_nd * aS);
            unsafeAtomicAdd(&A.grad_vel[3 * (size_t)p + los.ax - 1], P.scale * P.inv_b * aU);
            unsafeAtomicAdd(&A.grad_temp[p], -P.scale / (2.0 * P.temp) * (aS + aF));
        }
    }
}

int launch_tau_adjoint64(const DepositArgs &A, const WorkItem *items, int nblocks, hipStream_t s)
{
    tau_adjoint64_kernel<<<nblocks, kThreads, 0, s>>>(A, items);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int launch_tau_adjoint(int gsize, const DepositArgs &A, const WorkItem *items, int nblocks, hipStream_t s)
{
    const size_t shm = deposit_lds_bytes(A.nbins);
    if (gsize == 16) {
        if (int rc = raise_lds_limit((const void *)tau_adjoint_kernel<16>, shm)) return rc;
        tau_adjoint_kernel<16><<<nblocks, kThreads, shm, s>>>(A, items, 1);
    } else if (gsize == 32) {
        if (int rc = raise_lds_limit((const void *)tau_adjoint_kernel<32>, shm)) return rc;
        tau_adjoint_kernel<32><<<nblocks, kThreads, shm, s>>>(A, items, 1);
    } else {
        if (int rc = raise_lds_limit((const void *)tau_adjoint_kernel<64>, shm)) return rc;
        tau_adjoint_kernel<64><<<nblocks, kThreads, shm, s>>>(A, items, 1);
    }
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int adjoint_blocks_per_cu(int gsize, int nbins)
{
    int nb = 0;
    const size_t shm = deposit_lds_bytes(nbins);
    hipError_t e = gsize == 16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_adjoint_kernel<16>, kThreads, shm)
                 : gsize == 32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_adjoint_kernel<32>, kThreads, shm)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_adjoint_kernel<64>, kThreads, shm);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb < 1 ? 1 : nb;
}

// ---------------------------------------------------------------------------
// Column density

// absorption.cpp:53-74 (cubic), 76-97 (quintic) / 109-115 (tophat, also the
// reference's default for any other id, absorption.cpp:136-148)
__device__ inline double kern_frac(int kernel, double zlow, double zhigh, double smooth, double dr2,
                                   double zrange)
{
    zlow = fmax(zlow, -zrange);
    zhigh = fmin(zhigh, zrange);
    if (kernel == kVoronoiMesh) return fmax(0., zhigh - zlow);     // absorption.cpp:129-134
    if (kernel != kCubicSpline && kernel != kSphQuintic) return 3. / 4. / kPi * fmax(0., zhigh - zlow);
    if (zlow > zhigh) return 0;
    double total = kernel_weight(kernel, sqrt(dr2 + zlow * zlow) / smooth) / 2.;
    const double dz = (zhigh - zlow) / kNGrid;
    for (int i = 1; i < kNGrid; ++i) {
        const double zz = i * dz + zlow;
        total += kernel_weight(kernel, sqrt(dr2 + zz * zz) / smooth);
    }
    total += kernel_weight(kernel, sqrt(dr2 + zhigh * zhigh) / smooth) / 2.;
    return dz * total;
}

template <bool kLds>
__global__ __launch_bounds__(kThreads) void colden_kernel(DepositArgs A, const WorkItem *items, int item_base,
                                                          int use_queue)
{
    extern __shared__ double smem[];
    __shared__ int s_item;
    int it = use_queue ? 0 : blockIdx.x;
    for (;;) {
        int item;
        if (use_queue) {
            __syncthreads();
            if (threadIdx.x == 0) s_item = atomicAdd(A.queue, 1);
            __syncthreads();
            item = s_item;
        } else {
            item = it;
            it += gridDim.x;
        }
        if (item >= A.nitems) break;
        const WorkItem w = items[item_base + item];
        const LineDesc L = line_of(A, w.line);
        double *out_row = L.out + (size_t)w.los * A.nbins;
        double *row = kLds ? smem : out_row;
        if (kLds) {
            for (int j = threadIdx.x; j < A.nbins; j += kThreads) row[j] = 0.0;
            __syncthreads();
        }
        const LosInfo los = los_info(A, w.los);
        const double boxtokpc = A.vbox / A.nbins / A.velfac;   // absorption.cpp:188
        for (long long k = w.p0 + threadIdx.x; k < w.p1; k += kThreads) {
            const int p = A.pair_part[k];
            const float *pp = &A.pos[3 * (size_t)p];
            const float dens = L.dens[p];
            double pos1 = pp[los.ax - 1], zrange, dr2;
            float smooth;
            if (A.seg) {
                // Voronoi mesh: the segment ends replace dr2 and h
                // (absorption.cpp:170-187): centre and half length
                const float2 sg = A.seg[k];
                dr2 = (double)sg.x;
                smooth = sg.y;
                if (dr2 > 2 * A.vbox / A.velfac || (double)smooth > 2 * A.vbox / A.velfac) continue;
                pos1 = (dr2 + (double)smooth) / 2.;
                zrange = ((double)smooth - dr2) / 2.;
            } else {
                dr2 = pair_dr2(los, pp, A.box);
                smooth = A.h[p];
                const float sm2 = smooth * smooth;
                if ((double)sm2 - dr2 <= 0) continue;              // absorption.cpp:182
                zrange = sqrt((double)sm2 - dr2);
            }
            const int zlow = floor((pos1 - zrange) / boxtokpc);
            const int zhigh = ceil((pos1 + zrange) / boxtokpc);
            for (int z = zlow; z <= zhigh; ++z) {
                const double plow = (boxtokpc * z - pos1);
                const double v = (double)dens * kern_frac(A.kernel, plow, plow + boxtokpc, smooth, dr2, zrange);
                row_add<kLds>(&row[wrap_pixel(z, A.nbins)], v);
            }
        }
        if (kLds) {
            __syncthreads();
            if (w.direct) {
                for (int j = threadIdx.x; j < A.nbins; j += kThreads) out_row[j] = row[j];
            } else {
                for (int j = threadIdx.x; j < A.nbins; j += kThreads) {
                    const double v = row[j];
                    if (v != 0.0) unsafeAtomicAdd(&out_row[j], v);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Isolated throughput of the FP32 sample evaluator (kernel weight x profile),
// no memory traffic: each thread walks u = (tid + i * stride) * du over
// [0, umax), the mix of core and wing samples of a pixel window.

__global__ __launch_bounds__(kThreads) void voigt_peak_kernel(long long per_thread, double du, double umax,
                                                              float a, float *sink)
{
    __shared__ __attribute__((aligned(16))) float gtab[kGTableSize];
    for (int i = threadIdx.x; i < kGTableSize; i += kThreads) gtab[i] = c_gtab[i];
    __syncthreads();
    const VoigtA va = make_voigt_a(a);
    const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long nthreads = (long long)gridDim.x * blockDim.x;
    float acc = 0.0f;
    const float w = 0.25f;
    double u = fmod((double)tid * du, umax);
    const double step = fmod((double)nthreads * du, umax);
    for (long long i = 0; i < per_thread; ++i) {
        acc = __builtin_fmaf(w, voigt32(va, (float)u, gtab), acc);
        u += step;
        if (u >= umax) u -= umax;
    }
    if (acc == -1.0f) sink[0] = acc;   // keeps the loop live; never true
}

// H(a, u_i) by the deposit kernel's FP32 profile at the points u_i rounded
// to FP32, one lane per point (the profile's accuracy check); the damping
// parameter is set up as for a pair.
__global__ __launch_bounds__(kThreads) void voigt_eval_kernel(const double *u, long long n, double a, float *out)
{
    __shared__ __attribute__((aligned(16))) float gtab[kGTableSize];
    for (int i = threadIdx.x; i < kGTableSize; i += kThreads) gtab[i] = c_gtab[i];
    __syncthreads();
    const VoigtA va = make_voigt_a(a);
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        out[i] = voigt32(va, (float)u[i], gtab);
}

int launch_voigt_eval(const double *u, long long n, double a, float *out, hipStream_t s)
{
    if (n <= 0) return FSAMD_OK;
    const long long nb = std::min<long long>((n + kThreads - 1) / kThreads, 4096);
    voigt_eval_kernel<<<(int)nb, kThreads, 0, s>>>(u, n, a, out);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

// ---------------------------------------------------------------------------
// Launchers

size_t deposit_lds_bytes(int nbins) { return sizeof(double) * (size_t)nbins; }

// A kernel's static LDS, or a conservative 64 KiB if the runtime cannot say.
static size_t static_lds(const void *fn)
{
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, fn) != hipSuccess) {
        (void)hipGetLastError();
        return 65536;
    }
    return a.sharedSizeBytes;
}

template <int kG>
static size_t tau_static_lds(int waves)
{
    switch (waves) {
    case 2: return static_lds((const void *)tau_kernel<kG, true, false, 128>);
    case 8: return static_lds((const void *)tau_kernel<kG, true, false, 512>);
    default: return static_lds((const void *)tau_kernel<kG, true, false, 256>);   // the counting variant's too
    }
}

size_t deposit_static_lds(bool tau, int gsize, int waves)
{
    if (!tau) return static_lds((const void *)colden_kernel<true>);
    return gsize == 16 ? tau_static_lds<16>(waves) : gsize == 32 ? tau_static_lds<32>(waves) : tau_static_lds<64>(waves);
}

size_t adjoint_static_lds(int gsize)
{
    return static_lds(gsize == 16 ? (const void *)tau_adjoint_kernel<16>
                      : gsize == 32 ? (const void *)tau_adjoint_kernel<32> : (const void *)tau_adjoint_kernel<64>);
}

int raise_lds_limit(const void *fn, size_t shm)
{
    static std::mutex m;
    static std::unordered_map<const void *, size_t> limit;
    std::lock_guard<std::mutex> g(m);
    size_t &cur = limit[fn];
    if (shm > cur) {
        FSAMD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        cur = shm;
    }
    return FSAMD_OK;
}

template <int kG, bool kLds, bool kStats, int kBlock>
static int launch_tau_k(const DepositArgs &A, const WorkItem *items, int item_base, int nblocks, int use_queue,
                        size_t shm, hipStream_t s)
{
    int rc;
    if (kLds && (rc = raise_lds_limit((const void *)tau_kernel<kG, kLds, kStats, kBlock>, shm))) return rc;
    tau_kernel<kG, kLds, kStats, kBlock><<<nblocks, kBlock, shm, s>>>(A, items, item_base, use_queue);
    return FSAMD_OK;
}

// Wavefronts per block (per sightline work item) of the tau kernel: 2, 4 or
// 8 with the row in LDS; the counting variant and the global-row variant
// are built for 4 only.
int effective_waves(bool lds, bool stats, int waves)
{
    if (stats || !lds) return 4;
    return waves == 2 || waves == 8 ? waves : 4;
}

template <int kG, bool kLds>
static int launch_tau(const DepositArgs &A, const WorkItem *items, int item_base, int nblocks, int use_queue,
                      size_t shm, int waves, hipStream_t s)
{
    if (A.stats) return launch_tau_k<kG, kLds, true, 256>(A, items, item_base, nblocks, use_queue, shm, s);
    if constexpr (!kLds) {
        return launch_tau_k<kG, kLds, false, 256>(A, items, item_base, nblocks, use_queue, shm, s);
    } else {
        switch (effective_waves(kLds, false, waves)) {
        case 2: return launch_tau_k<kG, kLds, false, 128>(A, items, item_base, nblocks, use_queue, shm, s);
        case 8: return launch_tau_k<kG, kLds, false, 512>(A, items, item_base, nblocks, use_queue, shm, s);
        default: return launch_tau_k<kG, kLds, false, 256>(A, items, item_base, nblocks, use_queue, shm, s);
        }
    }
}

int launch_deposit(bool tau, bool lds, int gsize, int waves, const DepositArgs &A, const WorkItem *items,
                   int item_base, int nblocks, int use_queue, hipStream_t s)
{
    const size_t shm = lds ? deposit_lds_bytes(A.nbins) : 0;
    int rc = FSAMD_OK;
    if (tau) {
        if (gsize == 16)
            rc = lds ? launch_tau<16, true>(A, items, item_base, nblocks, use_queue, shm, waves, s)
                     : launch_tau<16, false>(A, items, item_base, nblocks, use_queue, shm, waves, s);
        else if (gsize == 32)
            rc = lds ? launch_tau<32, true>(A, items, item_base, nblocks, use_queue, shm, waves, s)
                     : launch_tau<32, false>(A, items, item_base, nblocks, use_queue, shm, waves, s);
        else
            rc = lds ? launch_tau<64, true>(A, items, item_base, nblocks, use_queue, shm, waves, s)
                     : launch_tau<64, false>(A, items, item_base, nblocks, use_queue, shm, waves, s);
        if (rc) return rc;
    } else {
        if (lds) {
            if (int rc2 = raise_lds_limit((const void *)colden_kernel<true>, shm)) return rc2;
            colden_kernel<true><<<nblocks, kThreads, shm, s>>>(A, items, item_base, use_queue);
        } else {
            colden_kernel<false><<<nblocks, kThreads, 0, s>>>(A, items, item_base, use_queue);
        }
    }
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int launch_tau_fixup(const DepositArgs &A, const WorkItem *items, int nblocks, hipStream_t s)
{
    // the counting variant only when counters are collected
    if (A.stats)
        tau_fixup_kernel<true><<<nblocks, kThreads, 0, s>>>(A, items);
    else
        tau_fixup_kernel<false><<<nblocks, kThreads, 0, s>>>(A, items);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int launch_voigt_peak(long long per_thread, int nblocks, double du, double umax, float a, float *sink,
                      hipStream_t s)
{
    voigt_peak_kernel<<<nblocks, kThreads, 0, s>>>(per_thread, du, umax, a, sink);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

template <int kG, int kBlock>
static int occ_tau_k(bool lds, size_t shm)
{
    int nb = 0;
    hipError_t e = lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_kernel<kG, true, false, kBlock>, kBlock, shm)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_kernel<kG, false, false, 256>, 256, 0);
    // (the global-row variant is built with 256 threads only: effective_waves)
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return nb;
}

template <int kG>
static int occ_tau(bool lds, size_t shm, int waves)
{
    switch (effective_waves(lds, false, waves)) {
    case 2: return occ_tau_k<kG, 128>(lds, shm);
    case 8: return occ_tau_k<kG, 512>(lds, shm);
    default: return occ_tau_k<kG, 256>(lds, shm);
    }
}

int deposit_blocks_per_cu(bool tau, bool lds, int gsize, int waves, int nbins)
{
    int nb = 0;
    const size_t shm = lds ? deposit_lds_bytes(nbins) : 0;
    if (tau) {
        nb = gsize == 16 ? occ_tau<16>(lds, shm, waves)
             : gsize == 32 ? occ_tau<32>(lds, shm, waves) : occ_tau<64>(lds, shm, waves);
    } else {
        hipError_t e = lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, colden_kernel<true>, kThreads, shm)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, colden_kernel<false>, kThreads, 0);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            nb = 1;
        }
    }
    return nb < 1 ? 1 : nb;
}

}  // namespace fsamd
