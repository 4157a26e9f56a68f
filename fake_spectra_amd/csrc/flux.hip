// Flux statistics on device-resident optical depths: the mean-flux rescale
// (py_module.cpp:235-262, fluxstatistics.py:26-39), spectrograph smoothing
// (spec_utils.py:5-25), the 1D flux power (fluxstatistics.py:71-108) and the
// flux PDF (fluxstatistics.py:41-49).  All FP64, HBM-bound streaming kernels.
#include "fsamd_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace fsamd {

constexpr int kFluxThreads = 256;

template <typename T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block-wide sum of kN values per thread; thread 0 returns the totals.
template <int kN>
__device__ inline void block_sum(double (&v)[kN])
{
    __shared__ double part[kN][kFluxThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kN; ++i) {
        const double s = wave_sum(v[i]);
        if (lane == 0) part[i][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < kN; ++i) {
            double s = 0;
            for (int w = 0; w < kFluxThreads / 64; ++w) s += part[i][w];
            v[i] = s;
        }
    }
    __syncthreads();
}

// Per-block partial sums of exp(-s tau), tau exp(-s tau) and the count of
// pixels with tau <= thresh (py_module.cpp:242-251).
__global__ __launch_bounds__(kFluxThreads) void flux_moments_kernel(const double *tau, long long n, double scale,
                                                                   double thresh, double *partial)
{
    double v[3] = {0, 0, 0};
    const long long stride = (long long)gridDim.x * kFluxThreads;
    for (long long i = blockIdx.x * (long long)kFluxThreads + threadIdx.x; i < n; i += stride) {
        const double t = tau[i];
        if (t > thresh) continue;
        const double e = exp(-scale * t);
        v[0] += e;
        v[1] += e * t;
        v[2] += 1.0;
    }
    block_sum<3>(v);
    if (threadIdx.x == 0)
        for (int i = 0; i < 3; ++i) partial[3 * blockIdx.x + i] = v[i];
}

// Gaussian smoothing along rows with periodic wrap, weights w[0..2r]
// (the correlation of scipy.ndimage.gaussian_filter1d, mode 'wrap').
__global__ __launch_bounds__(kFluxThreads) void res_corr_kernel(const double *in, double *out, int nrow, int ncol,
                                                               const double *w, int r)
{
    for (int row = blockIdx.x; row < nrow; row += gridDim.x) {
        const double *src = in + (size_t)row * ncol;
        double *dst = out + (size_t)row * ncol;
        for (int i = threadIdx.x; i < ncol; i += kFluxThreads) {
            double acc = 0;
            for (int k = -r; k <= r; ++k) {
                int j = (i + k) % ncol;
                if (j < 0) j += ncol;
                acc += w[k + r] * src[j];
            }
            dst[i] = acc;
        }
    }
}

// Direct-sum power spectrum, kept for npix too large for the FFT's LDS
// buffers: power[row][k] = |sum_j d_j exp(-2 pi i j k / npix)|^2 with
// d = exp(-s tau)/mean - 1, over an LDS twiddle table.  One block per
// sightline, O(npix^2).
__global__ __launch_bounds__(kFluxThreads) void flux_dft_kernel(const double *tau, int nlos, int npix, double scale,
                                                               double inv_mean, double *power)
{
    extern __shared__ double sh[];
    double *d = sh;                 // [npix]
    double *cs = sh + npix;         // [npix] cos(2 pi m / npix)
    double *sn = sh + 2 * npix;     // [npix] sin(2 pi m / npix)
    for (int m = threadIdx.x; m < npix; m += kFluxThreads) {
        double s, c;
        sincospi(2.0 * m / npix, &s, &c);
        cs[m] = c;
        sn[m] = s;
    }
    const int nk = npix / 2 + 1;
    for (int row = blockIdx.x; row < nlos; row += gridDim.x) {
        __syncthreads();
        for (int j = threadIdx.x; j < npix; j += kFluxThreads)
            d[j] = exp(-scale * tau[(size_t)row * npix + j]) * inv_mean - 1.0;
        __syncthreads();
        for (int k = threadIdx.x; k < nk; k += kFluxThreads) {
            double re = 0, im = 0;
            unsigned int m = 0;        // j k mod npix
            for (int j = 0; j < npix; ++j) {
                re += d[j] * cs[m];
                im -= d[j] * sn[m];
                m += (unsigned int)k;
                if (m >= (unsigned int)npix) m -= npix;
            }
            power[(size_t)row * nk + k] = re * re + im * im;
        }
    }
}

// ---- mixed-radix FFT power spectrum ---------------------------------------
// Self-sorting (Stockham) decimation-in-time FFT of length n = R_1 R_2 ...
// R_S in LDS, two buffers ping-ponged.  Stage s with radix R, after the
// transforms of length p = R_1 ... R_{s-1}: for i in [0, n/R), k = i mod p,
//   u_r = a[i + r n/R] W_n^{r k n/(p R)},   b[(i - k) R + k + q p] = sum_r u_r W_R^{r q}.
// Two real sightlines a, b share one complex transform Z of a + i b; the sum
// of their powers at mode k is (|Z_k|^2 + |Z_{n-k}|^2) / 2, which is all the
// flux power needs (it averages the power over sightlines).

constexpr int kFftMaxStages = 40;
struct FftPlan {
    int nst;
    int radix[kFftMaxStages];
};

// Radices 4, 2, 3, 5, 7 first, then any remaining prime factors.
static FftPlan fft_plan(int n)
{
    FftPlan f{};
    for (int r : {4, 2, 3, 5, 7})
        while (n % r == 0) { f.radix[f.nst++] = r; n /= r; }
    for (int q = 11; n > 1; q += 2)
        while (n % q == 0) { f.radix[f.nst++] = q; n /= q; }
    return f;
}

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// One stage of radix R (compile-time), tw[m] = W_n^m = exp(-2 pi i m / n).
template <int R>
__device__ inline void fft_stage(const double2 *a, double2 *b, const double2 *tw, int n, int p)
{
    const int m = n / R;
    const int tstep = n / (p * R);
    for (int i = threadIdx.x; i < m; i += kFluxThreads) {
        const int k = i % p;
        double2 u[R];
        u[0] = a[i];
#pragma unroll
        for (int r = 1; r < R; ++r) u[r] = cmul(a[i + r * m], tw[r * k * tstep]);  // r k tstep < n
        const int j = (i - k) * R + k;
        if constexpr (R == 2) {
            b[j] = cadd(u[0], u[1]);
            b[j + p] = csub(u[0], u[1]);
        } else if constexpr (R == 4) {
            // W_4 = -i
            const double2 s02 = cadd(u[0], u[2]), d02 = csub(u[0], u[2]);
            const double2 s13 = cadd(u[1], u[3]), d13 = csub(u[1], u[3]);
            b[j] = cadd(s02, s13);
            b[j + p] = make_double2(d02.x + d13.y, d02.y - d13.x);       // d02 - i d13
            b[j + 2 * p] = csub(s02, s13);
            b[j + 3 * p] = make_double2(d02.x - d13.y, d02.y + d13.x);   // d02 + i d13
        } else {
            const int wstep = n / R;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double2 acc = u[0];
#pragma unroll
                for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(u[r], tw[((r * q) % R) * wstep]));
                b[j + q * p] = acc;
            }
        }
    }
}

// A stage of any radix R: one output per thread iteration, O(R) terms each.
__device__ inline void fft_stage_any(const double2 *a, double2 *b, const double2 *tw, int n, int p, int R)
{
    const int m = n / R;
    const int tstep = n / (p * R);
    const int wstep = n / R;
    for (int o = threadIdx.x; o < n; o += kFluxThreads) {
        const int q = o / m, i = o - q * m;
        const int k = i % p;
        double2 acc = make_double2(0.0, 0.0);
        for (int r = 0; r < R; ++r) {
            int t = r * k * tstep + ((r * q) % R) * wstep;   // both terms < n
            if (t >= n) t -= n;
            acc = cadd(acc, cmul(a[i + r * m], tw[t]));
        }
        b[(i - k) * R + k + q * p] = acc;
    }
}

__global__ void fft_twiddle_kernel(double2 *tw, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(2.0 * m / n, &s, &c);
    tw[m] = make_double2(c, -s);
}

// Each block transforms sightline pairs blockIdx.x, blockIdx.x + gridDim.x, ...
// and writes the sum of their powers to part[blockIdx.x][k]; the blocks'
// rows are then summed in block order, so the result does not depend on
// scheduling.  LDS: two complex buffers of n and the accumulator of n/2 + 1.
__global__ __launch_bounds__(kFluxThreads) void flux_fft_kernel(const double *tau, int nlos, int npix, double scale,
                                                               double inv_mean, const double2 *tw, FftPlan fp,
                                                               double *part)
{
    extern __shared__ double2 shc[];
    double2 *A = shc, *B = shc + npix;
    double *acc = reinterpret_cast<double *>(shc + 2 * npix);
    const int nk = npix / 2 + 1;
    for (int k = threadIdx.x; k < nk; k += kFluxThreads) acc[k] = 0.0;
    const int npair = (nlos + 1) / 2;
    for (int pr = blockIdx.x; pr < npair; pr += gridDim.x) {
        const size_t r0 = 2 * (size_t)pr;
        const bool two = r0 + 1 < (size_t)nlos;
        __syncthreads();   // the last pair's power reads are done
        for (int j = threadIdx.x; j < npix; j += kFluxThreads) {
            const double x = exp(-scale * tau[r0 * npix + j]) * inv_mean - 1.0;
            const double y = two ? exp(-scale * tau[(r0 + 1) * npix + j]) * inv_mean - 1.0 : 0.0;
            A[j] = make_double2(x, y);
        }
        double2 *src = A, *dst = B;
        int p = 1;
        for (int st = 0; st < fp.nst; ++st) {
            __syncthreads();
            const int R = fp.radix[st];
            switch (R) {
            case 2: fft_stage<2>(src, dst, tw, npix, p); break;
            case 3: fft_stage<3>(src, dst, tw, npix, p); break;
            case 4: fft_stage<4>(src, dst, tw, npix, p); break;
            case 5: fft_stage<5>(src, dst, tw, npix, p); break;
            case 7: fft_stage<7>(src, dst, tw, npix, p); break;
            default: fft_stage_any(src, dst, tw, npix, p, R); break;
            }
            double2 *t = src; src = dst; dst = t;
            p *= R;
        }
        __syncthreads();
        // each thread owns modes k = threadIdx.x + c kFluxThreads of acc
        for (int k = threadIdx.x; k < nk; k += kFluxThreads) {
            const double2 z = src[k], w = src[k == 0 ? 0 : npix - k];
            acc[k] += 0.5 * ((z.x * z.x + z.y * z.y) + (w.x * w.x + w.y * w.y));
        }
    }
    for (int k = threadIdx.x; k < nk; k += kFluxThreads) part[(size_t)blockIdx.x * nk + k] = acc[k];
}

// sum over rows of power[row][k], in row order (deterministic)
__global__ void column_sum_kernel(const double *power, int nrow, int nk, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    double s = 0;
    for (int r = 0; r < nrow; ++r) s += power[(size_t)r * nk + k];
    out[k] = s;
}

// Histogram of exp(-s tau) over nb equal bins of [0, 1], bin edges e_k =
// k / nb as numpy.histogram places them (searchsorted on the edges, the last
// bin closed); values outside [0, 1] are not counted.
__device__ inline int flux_bin(double f, int nb)
{
    int b = (int)(f * nb);
    b = b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
    if (f < (double)b / (double)nb && b > 0) --b;
    if (b < nb - 1 && f >= (double)(b + 1) / (double)nb) ++b;
    return b;
}

// Counts are kept per block in LDS and added to the global bins once per
// block: nb <= kFluxPrivBins bins get one private row per thread (no atomics;
// a flux field piles up in few bins, which would serialize atomics), larger
// nb one shared row with LDS atomics.
constexpr int kFluxPrivBins = 32;

__global__ __launch_bounds__(kFluxThreads) void flux_hist_kernel(const double *tau, long long n, double scale, int nb,
                                                                unsigned long long *counts)
{
    extern __shared__ unsigned int hist[];   // [nb][kFluxThreads] private rows, or [nb] shared
    const bool priv = nb <= kFluxPrivBins;
    const int rows = priv ? nb * kFluxThreads : nb;
    for (int i = threadIdx.x; i < rows; i += kFluxThreads) hist[i] = 0;
    __syncthreads();
    const long long stride = (long long)gridDim.x * kFluxThreads;
    for (long long i = blockIdx.x * (long long)kFluxThreads + threadIdx.x; i < n; i += stride) {
        const double f = exp(-scale * tau[i]);
        if (!(f >= 0.0 && f <= 1.0)) continue;
        const int b = flux_bin(f, nb);
        if (priv)
            ++hist[b * kFluxThreads + threadIdx.x];
        else
            atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kFluxThreads) {
        unsigned long long c = 0;
        if (priv)
            for (int t = 0; t < kFluxThreads; ++t) c += hist[b * kFluxThreads + t];
        else
            c = hist[b];
        if (c) atomicAdd(&counts[b], c);
    }
}

}  // namespace fsamd

using namespace fsamd;

extern "C" {

int fsamd_mean_flux_scale(const double *tau, long long n, double mean_flux_desired, double tol, double thresh,
                          double *scale_out, void *stream)
{
    if (!scale_out || n < 0 || (n > 0 && !tau)) return fail(FSAMD_ERR_ARG, "NULL argument or negative size");
    if (n == 0) {
        *scale_out = 0.0;    // fluxstatistics.py:38-39
        return FSAMD_OK;
    }
    hipStream_t s = (hipStream_t)stream;
    const int nblocks = (int)std::min<long long>(1024, (n + kFluxThreads - 1) / kFluxThreads);
    double *partial = nullptr;
    FSAMD_HIP(hipMalloc(&partial, sizeof(double) * 3 * nblocks));
    std::vector<double> host(3 * nblocks);
    double scale, newscale = 1;
    int rc = FSAMD_OK;
    for (int it = 0; it < 1000; ++it) {
        scale = newscale;
        flux_moments_kernel<<<nblocks, kFluxThreads, 0, s>>>(tau, n, scale, thresh, partial);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(host.data(), partial, sizeof(double) * 3 * nblocks,
                                                hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            rc = fail(FSAMD_ERR_HIP, std::string("flux moments: ") + hipGetErrorString(e));
            break;
        }
        double mf = 0, tmf = 0, used = 0;
        for (int b = 0; b < nblocks; ++b) {
            mf += host[3 * b];
            tmf += host[3 * b + 1];
            used += host[3 * b + 2];
        }
        // Newton-Raphson on <exp(-s tau)> = mean_flux_desired (py_module.cpp:252-257)
        newscale = scale + (mf - mean_flux_desired * used) / tmf;
        if (newscale <= 0) newscale = 1e-10;
        if (!(fabs(newscale - scale) > tol * newscale)) break;
        if (it == 999) rc = fail(FSAMD_ERR_ARG, "mean flux rescale did not converge");
    }
    (void)hipFree(partial);
    if (rc) return rc;
    *scale_out = newscale;
    return FSAMD_OK;
}

int fsamd_res_corr(const double *in, double *out, int nrow, int ncol, double sigma, void *stream)
{
    if (nrow < 0 || ncol < 1 || !(sigma > 0) || (nrow > 0 && (!in || !out)))
        return fail(FSAMD_ERR_ARG, "bad arguments to fsamd_res_corr");
    if (nrow == 0) return FSAMD_OK;
    // scipy.ndimage.gaussian_filter1d: truncate = 4, radius = int(4 sigma + 0.5)
    const int r = (int)(4.0 * sigma + 0.5);
    if (r > (1 << 20)) return fail(FSAMD_ERR_UNSUPPORTED, "smoothing radius too large");
    std::vector<double> w(2 * r + 1);
    double tot = 0;
    for (int k = -r; k <= r; ++k) tot += (w[k + r] = std::exp(-0.5 * (double)k * k / (sigma * sigma)));
    for (double &x : w) x /= tot;
    hipStream_t s = (hipStream_t)stream;
    double *dw = nullptr;
    FSAMD_HIP(hipMalloc(&dw, sizeof(double) * w.size()));
    hipError_t e = hipMemcpyAsync(dw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        res_corr_kernel<<<std::min(nrow, 65536), kFluxThreads, 0, s>>>(in, out, nrow, ncol, dw, r);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dw);
    if (e != hipSuccess) return fail(FSAMD_ERR_HIP, std::string("res_corr: ") + hipGetErrorString(e));
    return FSAMD_OK;
}

int fsamd_flux_power(const double *tau, int nlos, int npix, double scale, double mean_flux, double vmax,
                     double *power, void *stream)
{
    if (nlos < 1 || npix < 1 || !tau || !power || !(mean_flux != 0))
        return fail(FSAMD_ERR_ARG, "bad arguments to fsamd_flux_power");
    constexpr size_t kLdsBytes = 160 * 1024;
    hipStream_t s = (hipStream_t)stream;
    const int nk = npix / 2 + 1;
    const size_t fft_shm = sizeof(double2) * 2 * (size_t)npix + sizeof(double) * (size_t)nk;
    const size_t dft_shm = sizeof(double) * 3 * (size_t)npix;
    const bool use_fft = fft_shm <= kLdsBytes;
    if (!use_fft && dft_shm > kLdsBytes)
        return fail(FSAMD_ERR_UNSUPPORTED, "flux power: npix too large for the LDS buffers");
    // FFT: one partial row per block (at most 1024 blocks), DFT: one per sightline
    const int npair = (nlos + 1) / 2;
    const int nblk = use_fft ? std::min(npair, 1024) : nlos;
    const size_t nrow = (size_t)nblk;
    double *per = nullptr;
    FSAMD_HIP(hipMalloc(&per, sizeof(double) * nrow * nk + (use_fft ? sizeof(double2) * (size_t)npix : 0)));
    hipError_t e = hipSuccess;
    if (use_fft) {
        double2 *tw = reinterpret_cast<double2 *>(per + nrow * nk);
        if (int rc = raise_lds_limit((const void *)flux_fft_kernel, fft_shm)) {
            (void)hipFree(per);
            return rc;
        }
        fft_twiddle_kernel<<<(npix + 255) / 256, 256, 0, s>>>(tw, npix);
        e = hipGetLastError();
        if (e == hipSuccess) {
            flux_fft_kernel<<<nblk, kFluxThreads, fft_shm, s>>>(tau, nlos, npix, scale, 1.0 / mean_flux, tw,
                                                               fft_plan(npix), per);
            e = hipGetLastError();
        }
    } else {
        if (int rc = raise_lds_limit((const void *)flux_dft_kernel, dft_shm)) {
            (void)hipFree(per);
            return rc;
        }
        flux_dft_kernel<<<std::min(nlos, 65536), kFluxThreads, dft_shm, s>>>(tau, nlos, npix, scale,
                                                                            1.0 / mean_flux, per);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        column_sum_kernel<<<(nk + 255) / 256, 256, 0, s>>>(per, (int)nrow, nk, power);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(per);
    if (e != hipSuccess) return fail(FSAMD_ERR_HIP, std::string("flux power: ") + hipGetErrorString(e));
    // mean over sightlines of vmax |F_k|^2 / npix^2 (fluxstatistics.py:49-56, 95-102)
    std::vector<double> h(nk);
    FSAMD_HIP(hipMemcpy(h.data(), power, sizeof(double) * nk, hipMemcpyDeviceToHost));
    const double norm = vmax / ((double)npix * (double)npix) / nlos;
    for (double &x : h) x *= norm;
    FSAMD_HIP(hipMemcpy(power, h.data(), sizeof(double) * nk, hipMemcpyHostToDevice));
    return FSAMD_OK;
}

int fsamd_flux_hist(const double *tau, long long n, double scale, int nbins, unsigned long long *counts,
                    void *stream)
{
    if (n < 0 || nbins < 1 || !counts || (n > 0 && !tau)) return fail(FSAMD_ERR_ARG, "bad arguments to fsamd_flux_hist");
    hipStream_t s = (hipStream_t)stream;
    FSAMD_HIP(hipMemsetAsync(counts, 0, sizeof(unsigned long long) * nbins, s));
    if (n > 0) {
        const int nblocks = (int)std::min<long long>(4096, (n + kFluxThreads - 1) / kFluxThreads);
        const size_t shm = sizeof(unsigned int) * (size_t)(nbins <= kFluxPrivBins ? nbins * kFluxThreads : nbins);
        if (shm > 160 * 1024) return fail(FSAMD_ERR_UNSUPPORTED, "flux histogram: too many bins for the LDS");
        if (int rc = raise_lds_limit((const void *)flux_hist_kernel, shm)) return rc;
        flux_hist_kernel<<<nblocks, kFluxThreads, shm, s>>>(tau, n, scale, nbins, counts);
        FSAMD_HIP(hipGetLastError());
    }
    FSAMD_HIP(hipStreamSynchronize(s));
    return FSAMD_OK;
}

}  // extern "C"
