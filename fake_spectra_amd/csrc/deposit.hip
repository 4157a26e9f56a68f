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

namespace fsamd {

__constant__ float c_gtab[kGTableSize] = FSAMD_G_COEF_INIT;

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr double kGuard = 2e-5;

// singleabs.h:17-26
__device__ inline double sph_cubic_kernel(double q)
{
    const double norm = 32. / 4 / kPi;
    if (q >= 1) return 0;
    if (q < 0.5) return norm * (1 - 6 * q * q + 6 * q * q * q);
    const double r = 1. - q;
    return norm * (2 * (r * r * r));
}

__device__ inline double kernel_weight(int kernel, double q)
{
    return kernel == kCubicSpline ? sph_cubic_kernel(q) : 3. / 4. / kPi;
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

__device__ inline int wrap_pixel(int z, int nbins)
{
    int j = z % nbins;
    return j < 0 ? j + nbins : j;
}

// Per-pair constants of LineAbsorption::add_tau_particle and SingleAbsorber.
struct PairTau {
    double vel;      // velfac*ppos + pvel (absorption.cpp:234)
    double b;        // btherm (absorption.cpp:220)
    double inv_b;
    double ampdens;  // amp * dens (absorption.cpp:244,255)
    double dvk;      // 2 vhigh / NGRID (singleabs.h:146)
    double vhigh, vdr2, vsmooth, aa;
    double cv[kNGrid - 1];   // kernel sample velocities / b
    float W[kNGrid - 1];     // kernel weights K(q_i)
    VoigtA va;
    int zmax;
    bool fp64;
};

// sum_i K(q_i) Re w((vv_i - vouter)/b + i a), FP32 profile
__device__ inline double inner32(const PairTau &P, double vouter, const float *gtab)
{
    const double s = vouter * P.inv_b;
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < kNGrid - 1; ++i) acc = __builtin_fmaf(P.W[i], voigt32(P.va, P.cv[i] - s, gtab), acc);
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
    double outer;
    if (width < P.b / 2.) {
        const double mid = (vhighp + vlow) / 2.;
        outer = P.dvk * (k64 ? inner64(P, mid, kernel) : inner32(P, mid, gtab));
    } else {
        const int np = 2 * ceil(width / (P.b / 2.) / 2) + 1.;
        const double dv = width / (np - 1);
        double total = 0.5 * (k64 ? inner64(P, vlow, kernel) : inner32(P, vlow, gtab));
        for (int i = 1; i < np - 1; ++i) {
            const double vv = i * dv + vlow;
            total += k64 ? inner64(P, vv, kernel) : inner32(P, vv, gtab);
        }
        total += 0.5 * (k64 ? inner64(P, vhighp, kernel) : inner32(P, vhighp, gtab));
        outer = P.dvk * total / (np - 1);
    }
    return P.ampdens * outer / velfac;
}

// Set up one pair; returns false if the particle does not reach the sightline.
__device__ inline bool setup_pair(const DepositArgs &A, const LineDesc &L, const LosInfo &los, int p,
                                  PairTau &P)
{
    const float *pp = &A.pos[3 * (size_t)p];
    const double dr2 = pair_dr2(los, pp, A.box);
    const float smooth = A.h[p];
    const float temp = A.temp[p];
    const float dens = L.dens[p];
    const float ppos = pp[los.ax - 1];
    const float pvel = A.vel[3 * (size_t)p + los.ax - 1];
    P.b = L.bfac * sqrt((double)temp);
    const float sm2 = smooth * smooth;
    if ((double)sm2 - dr2 <= 0) return false;            // absorption.cpp:232
    P.vel = A.velfac * (double)ppos + (double)pvel;
    double val1 = A.velfac * dr2;
    val1 *= A.velfac;
    P.vdr2 = val1;
    P.vsmooth = A.velfac * (double)smooth;
    P.aa = L.voigt_fac / P.b;
    P.vhigh = P.vsmooth * P.vsmooth > P.vdr2 ? sqrt(P.vsmooth * P.vsmooth - P.vdr2) : 0;   // singleabs.h:83
    P.inv_b = 1.0 / P.b;
    const double bintov = A.vbox / A.nbins;
    const double amp = L.sigma_a / sqrt(kPi) * (kLight / 1e5 / P.b);
    P.ampdens = amp * (double)dens;
    P.zmax = (int)floor(P.vel / bintov);
    P.dvk = 2. * P.vhigh / kNGrid;
#pragma unroll
    for (int i = 1; i < kNGrid; ++i) {
        const double vv = i * P.dvk - P.vhigh;
        const double q = sqrt(P.vdr2 + vv * vv) / P.vsmooth;
        P.cv[i - 1] = vv * P.inv_b;
        P.W[i - 1] = (float)kernel_weight(A.kernel, q);
    }
    P.fp64 = !(P.aa <= (double)kA32Max);
    P.va = make_voigt_a(P.fp64 ? 0.0 : P.aa);
    return true;
}

struct LaneStats {
    long long pairs = 0, pixels = 0, px64 = 0, pairs64 = 0;
};

template <bool kLds>
__device__ inline void row_add(double *p, double v)
{
    if (kLds)
        atomicAdd(p, v);          // ds_add_f64
    else
        unsafeAtomicAdd(p, v);    // global_atomic_add_f64
}

// Walk one direction: pixels z = z0, z0 + dir, ... for at most `count` pixels.
template <bool k64, bool kLds>
__device__ inline void walk(const DepositArgs &A, const PairTau &P, double *row, const float *gtab, int z0,
                            int dir, int count, int lane, LaneStats &st)
{
    const double bintov = A.vbox / A.nbins;
    for (int c = 0; c < count; c += 64) {
        const int idx = c + lane;
        const bool valid = idx < count;
        const int z = z0 + dir * idx;
        double t = 0;
        if (valid) {
            const double vlow = z * bintov - P.vel;                 // absorption.cpp:252
            t = taulast_pixel<k64>(P, vlow, bintov, A.velfac, A.kernel, gtab);
            if (!k64 && fabs(t - A.tautail) <= kGuard * A.tautail) {
                t = taulast_pixel<true>(P, vlow, bintov, A.velfac, A.kernel, gtab);
                st.px64++;
            }
        }
        const unsigned long long below = __ballot(valid && t < A.tautail);
        const int last = below ? __builtin_ctzll(below) : 63;
        if (valid && lane <= last) {
            row_add<kLds>(&row[wrap_pixel(z, A.nbins)], t);
            st.pixels++;
        }
        if (below) break;
    }
}

template <bool kLds>
__global__ __launch_bounds__(kThreads) void tau_kernel(DepositArgs A, const WorkItem *items, int item_base,
                                                       int use_queue)
{
    extern __shared__ double smem[];
    __shared__ float gtab[kGTableSize];
    __shared__ int s_item;
    for (int i = threadIdx.x; i < kGTableSize; i += kThreads) gtab[i] = c_gtab[i];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    LaneStats st;
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
        for (long long k = w.p0 + wave; k < w.p1; k += kWaves) {
            PairTau P;
            if (!setup_pair(A, L, los, A.pair_part[k], P)) continue;
            st.pairs++;
            const int half = A.nbins / 2;
            if (P.fp64) {
                st.pairs64++;
                walk<true, kLds>(A, P, row, gtab, P.zmax, 1, half, lane, st);
                walk<true, kLds>(A, P, row, gtab, P.zmax - 1, -1, half, lane, st);
            } else {
                walk<false, kLds>(A, P, row, gtab, P.zmax, 1, half, lane, st);
                walk<false, kLds>(A, P, row, gtab, P.zmax - 1, -1, half, lane, st);
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
    if (A.stats) {
        // lanes count pixels individually; pair counts are per wave (lane 0)
        atomicAdd((unsigned long long *)&A.stats[1], (unsigned long long)st.pixels);
        atomicAdd((unsigned long long *)&A.stats[2], (unsigned long long)st.px64);
        if (lane == 0) {
            atomicAdd((unsigned long long *)&A.stats[0], (unsigned long long)st.pairs);
            atomicAdd((unsigned long long *)&A.stats[3], (unsigned long long)st.pairs64);
        }
    }
}

// ---------------------------------------------------------------------------
// Column density

// absorption.cpp:53-74 (cubic) / 109-115 (tophat)
__device__ inline double kern_frac(int kernel, double zlow, double zhigh, double smooth, double dr2,
                                   double zrange)
{
    zlow = fmax(zlow, -zrange);
    zhigh = fmin(zhigh, zrange);
    if (kernel != kCubicSpline) return 3. / 4. / kPi * fmax(0., zhigh - zlow);
    if (zlow > zhigh) return 0;
    double total = sph_cubic_kernel(sqrt(dr2 + zlow * zlow) / smooth) / 2.;
    const double dz = (zhigh - zlow) / kNGrid;
    for (int i = 1; i < kNGrid; ++i) {
        const double zz = i * dz + zlow;
        total += sph_cubic_kernel(sqrt(dr2 + zz * zz) / smooth);
    }
    total += sph_cubic_kernel(sqrt(dr2 + zhigh * zhigh) / smooth) / 2.;
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
            const double dr2 = pair_dr2(los, pp, A.box);
            const float smooth = A.h[p];
            const float dens = L.dens[p];
            const double pos1 = pp[los.ax - 1];
            const float sm2 = smooth * smooth;
            if ((double)sm2 - dr2 <= 0) continue;              // absorption.cpp:182
            const double zrange = sqrt((double)sm2 - dr2);
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
// Launchers

size_t deposit_lds_bytes(int nbins) { return sizeof(double) * (size_t)nbins; }

int launch_deposit(bool tau, bool lds, const DepositArgs &A, const WorkItem *items, int item_base, int nblocks,
                   int use_queue, hipStream_t s)
{
    const size_t shm = lds ? deposit_lds_bytes(A.nbins) : 0;
    if (tau) {
        if (lds) {
            FSAMD_HIP(hipFuncSetAttribute((const void *)tau_kernel<true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            tau_kernel<true><<<nblocks, kThreads, shm, s>>>(A, items, item_base, use_queue);
        } else {
            tau_kernel<false><<<nblocks, kThreads, 0, s>>>(A, items, item_base, use_queue);
        }
    } else {
        if (lds) {
            FSAMD_HIP(hipFuncSetAttribute((const void *)colden_kernel<true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            colden_kernel<true><<<nblocks, kThreads, shm, s>>>(A, items, item_base, use_queue);
        } else {
            colden_kernel<false><<<nblocks, kThreads, 0, s>>>(A, items, item_base, use_queue);
        }
    }
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int deposit_blocks_per_cu(bool tau, bool lds, int nbins)
{
    int nb = 0;
    const size_t shm = lds ? deposit_lds_bytes(nbins) : 0;
    hipError_t e;
    if (tau)
        e = lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_kernel<true>, kThreads, shm)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tau_kernel<false>, kThreads, 0);
    else
        e = lds ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, colden_kernel<true>, kThreads, shm)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, colden_kernel<false>, kThreads, 0);
    if (e != hipSuccess || nb < 1) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb;
}

}  // namespace fsamd
