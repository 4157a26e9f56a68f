// Sightline <-> particle binning on the GPU.
//
// The reference (IndexTable, fake_spectra/index_table.cpp:7-150) sorts
// sightlines into two std::multimaps keyed by one transverse coordinate,
// range-queries them per particle, and inserts every accepted
// (sightline, particle, dr2) into per-sightline std::maps under an OpenMP
// critical section.  Here:
//   1. sightlines are bucketed into one square grid per axis over their two
//      transverse coordinates (a superset search structure only);
//   2. one thread per particle visits the grid cells its support touches and
//      applies the reference's acceptance test exactly (float key window with
//      periodic wrap, float second-coordinate window, FP64 minimum-image dr2
//      compared with the float product h*h), counting then emitting pairs
//      into a particle-major list at offsets from an exclusive scan;
//   3. a stable LSD radix sort on the sightline id (LDS digit histograms,
//      LDS prefix scans, wave-ballot ranking) turns it into per-sightline
//      rows in ascending particle order, the std::map order of the reference.
// All arithmetic that decides membership is compiled without FP contraction.
#include "fsamd_internal.h"
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

namespace fsamd {

// ---------------------------------------------------------------------------
// Small utilities

__global__ void fill_i32_kernel(int *p, int v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_f64_kernel(double *p, double v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

static int grid_for(long long n, int threads, int cap = 65536)
{
    long long b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

int launch_fill_i32(int *p, int value, long long n, hipStream_t s)
{
    if (n <= 0) return FSAMD_OK;
    fill_i32_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, value, n);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int launch_fill_f64(double *p, double value, long long n, hipStream_t s)
{
    if (n <= 0) return FSAMD_OK;
    fill_f64_kernel<<<grid_for(n, 256), 256, 0, s>>>(p, value, n);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

__global__ void pack_records_kernel(const float *pos, const float *vel, const float *temp, const float *h,
                                    const LineDesc *lines, int nlines, long long n, PartRec *out)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
        float4 a, b, d0, d1;
        a.x = pos[3 * i];
        a.y = pos[3 * i + 1];
        a.z = pos[3 * i + 2];
        a.w = h[i];
        b.x = vel[3 * i];
        b.y = vel[3 * i + 1];
        b.z = vel[3 * i + 2];
        b.w = temp[i];
        float d[kMaxLines];
#pragma unroll
        for (int k = 0; k < kMaxLines; ++k) d[k] = k < nlines ? lines[k].dens[i] : 0.0f;
        d0 = make_float4(d[0], d[1], d[2], d[3]);
        d1 = make_float4(d[4], d[5], d[6], d[7]);
        float4 *o = (float4 *)&out[i];
        o[0] = a;
        o[1] = b;
        o[2] = d0;
        o[3] = d1;
    }
}

int launch_pack_records(const float *pos, const float *vel, const float *temp, const float *h,
                        const LineDesc *lines, int nlines, long long n, PartRec *out, hipStream_t s)
{
    if (n <= 0) return FSAMD_OK;
    pack_records_kernel<<<grid_for(n, 256, 1 << 16), 256, 0, s>>>(pos, vel, temp, h, lines, nlines, n, out);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan (tile reduce -> scan of tile sums -> apply).

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline long long wave_incl_scan(long long v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        long long t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan over the 256 threads of a block; *total gets the block sum.
__device__ inline long long block_excl_scan(long long v, long long *lds4, long long *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long inc = wave_incl_scan(v);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        if (i < w) off += lds4[i];
        tot += lds4[i];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

template <typename T>
__global__ void scan_reduce_kernel(const T *in, long long n, long long *sums)
{
    __shared__ long long lds4[kScanThreads / 64];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) s += (long long)in[base + i];
    long long tot;
    block_excl_scan(s, lds4, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// out[i] = tile_off[tile] + exclusive prefix within the tile; out may alias in
template <typename T>
__global__ void scan_apply_kernel(const T *in, long long *out, long long n, const long long *tile_off)
{
    __shared__ long long lds4[kScanThreads / 64];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    long long v[kScanItems];
    long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? (long long)in[base + i] : 0;
        s += v[i];
    }
    long long tot;
    long long off = block_excl_scan(s, lds4, &tot) + (tile_off ? tile_off[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = off;
        off += v[i];
    }
}

size_t exclusive_scan_ws_bytes(long long n)
{
    // every level holds ceil(n/tile)+1 tile sums and recurses on that many
    size_t total = 0;
    n += 1;
    while (n > kScanTile) {
        long long nt = (n + kScanTile - 1) / kScanTile + 1;
        total += sizeof(long long) * (size_t)nt;
        n = nt;
    }
    return total + 256;
}

// In-place exclusive scan of int64 data[0..n); ws from exclusive_scan_ws_bytes.
static int scan_i64_rec(long long *data, long long n, char *ws, size_t ws_bytes, hipStream_t s)
{
    if (n <= 0) return FSAMD_OK;
    long long nt = (n + kScanTile - 1) / kScanTile;
    if (nt == 1) {
        scan_apply_kernel<long long><<<1, kScanThreads, 0, s>>>(data, data, n, nullptr);
        FSAMD_HIP(hipGetLastError());
        return FSAMD_OK;
    }
    size_t need = sizeof(long long) * (size_t)(nt + 1);
    if (need > ws_bytes) return fail(FSAMD_ERR_ARG, "scan workspace too small");
    long long *sums = (long long *)ws;
    scan_reduce_kernel<long long><<<(int)nt, kScanThreads, 0, s>>>(data, n, sums);
    FSAMD_HIP(hipGetLastError());
    int rc = scan_i64_rec(sums, nt, ws + need, ws_bytes - need, s);
    if (rc) return rc;
    scan_apply_kernel<long long><<<(int)nt, kScanThreads, 0, s>>>(data, data, n, sums);
    FSAMD_HIP(hipGetLastError());
    return FSAMD_OK;
}

int exclusive_scan_i64_inplace(long long *data, long long n, void *ws, size_t ws_bytes, hipStream_t s)
{
    return scan_i64_rec(data, n, (char *)ws, ws_bytes, s);
}

// out[0..n] = exclusive scan of int32 in[0..n), out[n] = total.
int exclusive_scan_i32_to_i64(const int *in, long long *out, long long n, void *ws, size_t ws_bytes,
                              hipStream_t s)
{
    long long nt = (n + kScanTile - 1) / kScanTile;
    if (n <= 0) {
        FSAMD_HIP(hipMemsetAsync(out, 0, sizeof(long long), s));
        return FSAMD_OK;
    }
    size_t need = sizeof(long long) * (size_t)(nt + 1);
    if (need > ws_bytes) return fail(FSAMD_ERR_ARG, "scan workspace too small");
    long long *sums = (long long *)ws;
    scan_reduce_kernel<int><<<(int)nt, kScanThreads, 0, s>>>(in, n, sums);
    FSAMD_HIP(hipGetLastError());
    // sums[nt] := 0 so that after the scan sums[nt] holds the grand total
    FSAMD_HIP(hipMemsetAsync(sums + nt, 0, sizeof(long long), s));
    int rc = scan_i64_rec(sums, nt + 1, (char *)ws + need, ws_bytes - need, s);
    if (rc) return rc;
    scan_apply_kernel<int><<<(int)nt, kScanThreads, 0, s>>>(in, out, n, sums);
    FSAMD_HIP(hipGetLastError());
    FSAMD_HIP(hipMemcpyAsync(out + n, sums + nt, sizeof(long long), hipMemcpyDeviceToDevice, s));
    return FSAMD_OK;
}

// ---------------------------------------------------------------------------
// Sightline grid

// Plane coordinates of a sightline: (key, second) as compared by the reference.
__device__ inline void los_plane(int ax, const double *c, double &k1, double &k2)
{
    if (ax == 1) {
        k1 = c[1];
        k2 = c[2];
    } else if (ax == 2) {
        k1 = c[0];
        k2 = c[2];
    } else {
        k1 = c[0];
        k2 = c[1];
    }
}

__device__ inline int wrap_cell(double k, double box, double inv_cw, int nc)
{
    double kw = k - floor(k / box) * box;
    int c = (int)floor(kw * inv_cw);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void los_axis_count_kernel(const int *axis, int nlos, int *counts3)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nlos) {
        int a = axis[i];
        if (a >= 1 && a <= 3) atomicAdd(&counts3[a - 1], 1);
    }
}

__global__ void los_cell_kernel(const int *axis, const double *cofm, int nlos, double box, LosGrid g,
                                int *cell_of, int *cell_count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlos) return;
    int a = axis[i];
    double k1, k2;
    los_plane(a, &cofm[3 * (size_t)i], k1, k2);
    // sightlines with non-finite coordinates never pass the reference's tests
    if (a < 1 || a > 3 || !isfinite(k1) || !isfinite(k2)) {
        cell_of[i] = -1;
        return;
    }
    const int ai = a - 1;
    const int c1 = wrap_cell(k1, box, g.inv_cw[ai], g.nc[ai]);
    const int c2 = wrap_cell(k2, box, g.inv_cw[ai], g.nc[ai]);
    const int cell = g.base[ai] + c2 * g.nc[ai] + c1;
    cell_of[i] = cell;
    atomicAdd(&cell_count[cell], 1);
}

__global__ void los_scatter_kernel(const int *axis, const double *cofm, int nlos, const int *cell_of,
                                   int *cursor, int *cell_ids, double2 *cell_kl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlos) return;
    int cell = cell_of[i];
    if (cell < 0) return;
    int at = atomicAdd(&cursor[cell], 1);
    double k1, k2;
    los_plane(axis[i], &cofm[3 * (size_t)i], k1, k2);
    cell_ids[at] = i;
    cell_kl[at] = make_double2(k1, k2);
}

__global__ void i64_to_i32_kernel(const long long *in, int *out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (int)in[i];
}

// ---------------------------------------------------------------------------
// The reference acceptance test (index_table.cpp:22-113), exact.
//   first/second: the particle's float coordinates along the sightline's key
//   and second axes; key/lproj2: the sightline's coordinates; hh: float h.
__device__ inline bool los_accept(float first, float second, float hh, double key, double lproj2,
                                  double box, double &dr2)
{
    // get_nearby: half-open key window [ffm, ffp) with periodic wrap (93-112)
    float ffp = first + hh;
    if ((double)ffp > box) ffp = (float)((double)ffp - box);
    float ffm = first - hh;
    if (ffm < 0) ffm = (float)((double)ffm + box);
    const bool inwin = (ffm <= ffp) ? (key >= (double)ffm && key < (double)ffp)
                                    : (key < (double)ffp || key >= (double)ffm);
    if (!inwin) return false;
    // second_close (52-68)
    bool close = false;
    float sp = second + hh;
    if ((double)sp > box && lproj2 < (double)sp - box) close = true;
    float sm = second - hh;
    if (!close && sm < 0 && lproj2 > (double)sm + box) close = true;
    if (!close && lproj2 > (double)sm && lproj2 < (double)sp) close = true;
    if (!close) return false;
    // calc_dr2 (70-87) and the acceptance dr2 <= hh*hh with a float product (45)
    double dr = fabs((double)first - key);
    if (dr > 0.5 * box) dr = box - dr;
    double d2 = dr * dr;
    dr = fabs((double)second - lproj2);
    if (dr > 0.5 * box) dr = box - dr;
    d2 += dr * dr;
    dr2 = d2;
    const float hh2 = hh * hh;
    return d2 <= (double)hh2;
}

// Particle coordinates (first, second) for a sightline axis (index_table.cpp:33-40, 117-127).
__device__ inline void particle_plane(int ax, const float *p, float &first, float &second)
{
    if (ax == 1) {
        first = p[1];
        second = p[2];
    } else if (ax == 2) {
        first = p[0];
        second = p[2];
    } else {
        first = p[0];
        second = p[1];
    }
}

// Visit every sightline accepted for particle p; f(sightline, dr2) returns
// false to stop the visit early.
template <class F>
__device__ inline void visit_accepted(const BinArgs &A, const float *p, float hh, F &&f)
{
    const double ah = fabs((double)hh);
    if (!(ah < 1e300)) return;  // NaN / inf support: the reference's windows are empty or undefined
    for (int ai = 0; ai < 3; ++ai) {
        if (A.grid.nlos_axis[ai] == 0) continue;
        const int ax = ai + 1;
        float first, second;
        particle_plane(ax, p, first, second);
        const int nc = A.grid.nc[ai];
        const double inv = A.grid.inv_cw[ai];
        const double hr = ah * (1.0 + 1e-5) + 1e-3 * A.grid.cw[ai];
        long long lo1 = (long long)floor(((double)first - hr) * inv);
        long long hi1 = (long long)floor(((double)first + hr) * inv);
        long long lo2 = (long long)floor(((double)second - hr) * inv);
        long long hi2 = (long long)floor(((double)second + hr) * inv);
        if (hi1 - lo1 + 1 >= nc) {
            lo1 = 0;
            hi1 = nc - 1;
        }
        if (hi2 - lo2 + 1 >= nc) {
            lo2 = 0;
            hi2 = nc - 1;
        }
        // a window narrower than the grid lies within (-nc, 2nc): one add or
        // subtract wraps a cell index (the modulo stays for anything else)
        auto wrap = [nc](long long c) {
            if (c >= 0 && c < nc) return (int)c;
            if (c < 0 && c >= -nc) return (int)(c + nc);
            if (c >= nc && c < 2 * (long long)nc) return (int)(c - nc);
            return (int)(((c % nc) + nc) % nc);
        };
        for (long long c2 = lo2; c2 <= hi2; ++c2) {
            const int w2 = wrap(c2);
            for (long long c1 = lo1; c1 <= hi1; ++c1) {
                const int w1 = wrap(c1);
                const int cell = A.grid.base[ai] + w2 * nc + w1;
                const int e = A.grid.cell_start[cell + 1];
                for (int k = A.grid.cell_start[cell]; k < e; ++k) {
                    const double2 kl = A.grid.cell_kl[k];
                    double dr2;
                    if (los_accept(first, second, hh, kl.x, kl.y, A.box, dr2) && !f(A.grid.cell_ids[k], dr2))
                        return;
                }
            }
        }
    }
}

__global__ void bin_count_kernel(BinArgs A, int *cnt)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.npart;
         i += (long long)gridDim.x * blockDim.x) {
        const float p[3] = {A.pos[3 * i], A.pos[3 * i + 1], A.pos[3 * i + 2]};
        int c = 0;
        visit_accepted(A, p, A.h[i], [&](int, double) {
            ++c;
            return true;
        });
        cnt[i] = c;
    }
}

__global__ void bin_emit_kernel(BinArgs A, const long long *poff, int *keys, int *vals)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.npart;
         i += (long long)gridDim.x * blockDim.x) {
        const float p[3] = {A.pos[3 * i], A.pos[3 * i + 1], A.pos[3 * i + 2]};
        long long at = poff[i];
        visit_accepted(A, p, A.h[i], [&](int s, double) {
            keys[at] = s;
            vals[at] = (int)i;
            ++at;
            return true;
        });
    }
}

// flag[i] = 1 if particle i reaches at least one sightline: the visit stops at
// the first accepted sightline (the _near_lines test, py_module.cpp:58-76).
__global__ void bin_any_kernel(BinArgs A, int *flag)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < A.npart;
         i += (long long)gridDim.x * blockDim.x) {
        const float p[3] = {A.pos[3 * i], A.pos[3 * i + 1], A.pos[3 * i + 2]};
        int any = 0;
        visit_accepted(A, p, A.h[i], [&](int, double) {
            any = 1;
            return false;
        });
        flag[i] = any;
    }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, value) int32 pairs by key digits.

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr long long kRsTile = (long long)kRsThreads * kRsItems;
constexpr int kRsMaxBits = 8;
constexpr int kRsWaves = kRsThreads / 64;

__global__ void rs_upsweep_kernel(const int *keys, long long n, int shift, int bits, long long ntiles,
                                  long long *table)
{
    __shared__ int hist[1 << kRsMaxBits];
    const int nb = 1 << bits;
    for (int d = threadIdx.x; d < nb; d += kRsThreads) hist[d] = 0;
    __syncthreads();
    const long long base = blockIdx.x * kRsTile;
    const unsigned mask = (unsigned)nb - 1;
    for (int r = 0; r < kRsItems; ++r) {
        const long long e = base + r * kRsThreads + threadIdx.x;
        if (e < n) atomicAdd(&hist[((unsigned)keys[e] >> shift) & mask], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < nb; d += kRsThreads) table[(long long)d * ntiles + blockIdx.x] = hist[d];
}

__global__ void rs_downsweep_kernel(const int *kin, const int *vin, int *kout, int *vout, long long n,
                                    int shift, int bits, long long ntiles, const long long *table)
{
    __shared__ long long base_off[1 << kRsMaxBits];
    __shared__ int run[1 << kRsMaxBits];
    __shared__ int wcnt[kRsWaves][1 << kRsMaxBits];
    const int nb = 1 << bits;
    const unsigned mask = (unsigned)nb - 1;
    for (int d = threadIdx.x; d < nb; d += kRsThreads) {
        base_off[d] = table[(long long)d * ntiles + blockIdx.x];
        run[d] = 0;
        for (int w = 0; w < kRsWaves; ++w) wcnt[w][d] = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const long long tbase = blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; ++r) {
        const long long e = tbase + r * kRsThreads + threadIdx.x;
        const bool valid = e < n;
        int k = 0, v = 0;
        unsigned d = 0;
        if (valid) {
            k = kin[e];
            v = vin[e];
            d = ((unsigned)k >> shift) & mask;
        }
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            long long off = base_off[d] + run[d] + rank;
            for (int w = 0; w < wave; ++w) off += wcnt[w][d];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        for (int dd = threadIdx.x; dd < nb; dd += kRsThreads) {
            int s = 0;
            for (int w = 0; w < kRsWaves; ++w) {
                s += wcnt[w][dd];
                wcnt[w][dd] = 0;
            }
            run[dd] += s;
        }
        __syncthreads();
    }
}

// offsets[s] = first position of key >= s in the sorted keys (s = 0..nlos)
__global__ void row_offsets_kernel(const int *keys, long long n, int nlos, long long *offsets)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nlos) return;
    long long lo = 0, hi = n;
    while (lo < hi) {
        long long mid = (lo + hi) >> 1;
        if (keys[mid] < s) lo = mid + 1;
        else hi = mid;
    }
    offsets[s] = lo;
}

__global__ void near_flag_kernel(const long long *poff, long long n, int *flag)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        flag[i] = poff[i + 1] > poff[i] ? 1 : 0;
}

__global__ void near_scatter_kernel(const long long *poff, const long long *fscan, long long n, int *out)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        if (poff[i + 1] > poff[i]) out[fscan[i]] = (int)i;
}

__global__ void flag_scatter_kernel(const int *flag, const long long *fscan, long long n, int *out)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        if (flag[i]) out[fscan[i]] = (int)i;
}

// ---------------------------------------------------------------------------
// Voronoi mesh (kernel 2): IndexTable::assign_cells (index_table.cpp:152-223).
// Each sightline is cut into npts = int(box/RESO) points at (i + 1/2) reso;
// point i goes to the nearest particle of the sightline's neighbour list
// (periodic distance, the first of the ascending row on ties), and a
// particle's segment runs from its first to its last point, +- reso/2.  The
// reference walks the points in order; here the same result comes from
// three parallel passes:
//   A: over the points past the middle of the box (xp > box/2 + reso/2), the
//      first one owned by the owner of point 0 (the periodic wrap, where the
//      reference restarts that particle's segment and stops the walk);
//   B: over the points before it, each particle's first and last point and
//      count (atomics);
//   C: per particle, the segment in FP32 with the reference's roundings; a
//      particle whose points are not contiguous marks the sightline in
//      error, where the reference exits (index_table.cpp:199-206).

struct VorArgs {
    double box, reso;
    int npts;                  // int(box / RESO)
    int imin;                  // first point with xp > box/2 + reso/2
    const float *pos;
    const int *axis;
    const double *cofm;
    const long long *line_off;
    const int *pair_part;
    int nlos;
    int *nn0;                  // [L] owner of point 0 (row index)
    int *istar;                // [L] first point >= imin owned by nn0; npts if none
    int *first, *last, *count; // [M] per pair
    int *err;                  // [L]
    float2 *seg;               // [M] segment ends
};

constexpr double kReso = 0.1;     // index_table.h:8
constexpr int kVorThreads = 256;
constexpr int kVorTile = 1024;   // row particles staged in LDS at a time

// Owner of the points handled by this block's threads (xp per thread): the
// row's particles are staged in LDS (coordinates along the sightline and the
// two transverse ones), every thread scans them in row order.
__device__ inline void vor_owners(const VorArgs &V, int s, double xp, bool active, int &owner)
{
    __shared__ float3 tile[kVorTile];
    const int ax = V.axis[s];
    const double *c = &V.cofm[3 * (size_t)s];
    const double yp = c[ax % 3], zp = c[(ax + 1) % 3];
    const long long p0 = V.line_off[s], n = V.line_off[s + 1] - p0;
    const double half = V.box / 2.;
    double best = V.box, best2 = V.box * V.box;
    int bi = 0;
    for (long long t0 = 0; t0 < n; t0 += kVorTile) {
        const int tn = (int)min((long long)kVorTile, n - t0);
        __syncthreads();
        for (int j = threadIdx.x; j < tn; j += kVorThreads) {
            const float *pp = &V.pos[3 * (size_t)V.pair_part[p0 + t0 + j]];
            tile[j] = make_float3(pp[ax - 1], pp[ax % 3], pp[(ax + 1) % 3]);
        }
        __syncthreads();
        if (active) {
            for (int j = 0; j < tn; ++j) {
                const float3 q = tile[j];
                double dx = fabs((double)q.x - xp);
                if (dx > half) dx = V.box - dx;
                double dy = fabs((double)q.y - yp);
                if (dy > half) dy = V.box - dy;
                double dz = fabs((double)q.z - zp);
                if (dz > half) dz = V.box - dz;
                const double d2 = dx * dx + dy * dy + dz * dz;
                // sqrt is monotonic: only a smaller d2 can give a smaller distance
                if (d2 < best2) {
                    const double d = sqrt(d2);
                    if (d < best) {
                        best = d;
                        best2 = d2;
                        bi = (int)(t0 + j);
                    }
                }
            }
        }
    }
    owner = bi;
}

// Owner of point 0 of each sightline (one block per sightline).
__global__ __launch_bounds__(kVorThreads) void vor_first_kernel(VorArgs V)
{
    const int s = blockIdx.x;
    if (V.line_off[s + 1] == V.line_off[s]) {
        if (threadIdx.x == 0) V.nn0[s] = -1;
        return;
    }
    int owner = 0;
    vor_owners(V, s, 0.5 * V.reso, threadIdx.x == 0, owner);
    if (threadIdx.x == 0) {
        V.nn0[s] = owner;
        V.istar[s] = V.npts;
    }
}

// Pass A: blocks over (chunk of points from imin, sightline).
__global__ __launch_bounds__(kVorThreads) void vor_wrap_kernel(VorArgs V)
{
    __shared__ int s_skip;
    const int s = blockIdx.y;
    const int i0 = V.imin + blockIdx.x * kVorThreads;
    if (V.nn0[s] < 0 || i0 >= V.npts) return;
    // an earlier point already wraps: read once for the whole block
    if (threadIdx.x == 0) s_skip = __atomic_load_n(&V.istar[s], __ATOMIC_RELAXED) < i0;
    __syncthreads();
    if (s_skip) return;
    const int i = i0 + threadIdx.x;
    int owner = 0;
    vor_owners(V, s, (i + 0.5) * V.reso, i < V.npts, owner);
    if (i < V.npts && owner == V.nn0[s]) atomicMin(&V.istar[s], i);
}

// Pass B: blocks over (chunk of points from 0, sightline), points before istar.
__global__ __launch_bounds__(kVorThreads) void vor_cells_kernel(VorArgs V)
{
    const int s = blockIdx.y;
    const int i0 = blockIdx.x * kVorThreads;
    if (V.nn0[s] < 0 || i0 >= V.istar[s]) return;
    const int i = i0 + threadIdx.x;
    const bool active = i < V.istar[s];
    int owner = 0;
    vor_owners(V, s, (i + 0.5) * V.reso, active, owner);
    if (active) {
        const long long k = V.line_off[s] + owner;
        atomicMin(&V.first[k], i);
        atomicMax(&V.last[k], i);
        atomicAdd(&V.count[k], 1);
    }
}

// Pass C: one block per sightline, threads over its particles.
__global__ __launch_bounds__(kVorThreads) void vor_segments_kernel(VorArgs V)
{
    const int s = blockIdx.x;
    const long long p0 = V.line_off[s], n = V.line_off[s + 1] - p0;
    const double reso = V.reso, box = V.box;
    for (long long j = threadIdx.x; j < n; j += kVorThreads) {
        const long long k = p0 + j;
        const int c = V.count[k];
        float lo = (float)(3 * box), hi = (float)(3 * box);
        if (c > 0) {
            const int f = V.first[k], l = V.last[k];
            if (c != l - f + 1) atomicOr(&V.err[s], 1);   // two separate runs: the reference exits
            lo = (float)((f + 0.5) * reso);
            hi = (float)((l + 0.5) * reso);
            if (j == V.nn0[s] && V.istar[s] < V.npts) {
                // the wrap (index_table.cpp:187-196): the segment restarts at
                // istar and its end moves on by one box length
                lo = (float)((V.istar[s] + 0.5) * reso);
                hi = (float)((double)hi + box);
            }
        }
        // index_table.cpp:218-221
        V.seg[k] = make_float2((float)((double)lo - 0.5 * reso), (float)((double)hi + 0.5 * reso));
    }
}

int plan_assign_cells(fsamd_plan *P, const float *pos, hipStream_t s)
{
    int rc;
    const long long M = P->npairs;
    const int L = P->nlos;
    if ((rc = P->seg.alloc(sizeof(float2) * (size_t)std::max<long long>(M, 1), P->arena))) return rc;
    P->seg_ready = false;
    if (L == 0 || M == 0) {
        P->seg_ready = true;
        return FSAMD_OK;
    }
    const int npts = (int)(P->box / kReso);
    if (npts < 1) return fail(FSAMD_ERR_ARG, "the box is shorter than the Voronoi grid spacing (0.1)");
    VorArgs V;
    V.box = P->box;
    V.reso = P->box / npts;
    V.npts = npts;
    // first point past the middle of the box (the wrap test of index_table.cpp:187)
    int imin = 0;
    while (imin < npts && !((imin + 0.5) * V.reso > P->box / 2. + 0.5 * V.reso)) ++imin;
    V.imin = imin;
    V.pos = pos;
    V.axis = P->axis.as<int>();
    V.cofm = P->cofm.as<double>();
    V.line_off = P->line_off.as<long long>();
    V.pair_part = P->pair_part.as<int>();
    V.nlos = L;
    DevBuf nn0, istar, first, last, count, err;
    if ((rc = nn0.alloc(sizeof(int) * (size_t)L, P->arena)) || (rc = istar.alloc(sizeof(int) * (size_t)L, P->arena)) ||
        (rc = err.alloc(sizeof(int) * (size_t)L, P->arena)) ||
        (rc = first.alloc(sizeof(int) * (size_t)M, P->arena)) || (rc = last.alloc(sizeof(int) * (size_t)M, P->arena)) ||
        (rc = count.alloc(sizeof(int) * (size_t)M, P->arena)))
        return rc;
    V.nn0 = nn0.as<int>();
    V.istar = istar.as<int>();
    V.err = err.as<int>();
    V.first = first.as<int>();
    V.last = last.as<int>();
    V.count = count.as<int>();
    V.seg = P->seg.as<float2>();
    if ((rc = launch_fill_i32(V.err, 0, L, s)) || (rc = launch_fill_i32(V.first, INT32_MAX, M, s)) ||
        (rc = launch_fill_i32(V.last, -1, M, s)) || (rc = launch_fill_i32(V.count, 0, M, s)))
        return rc;
    vor_first_kernel<<<L, kVorThreads, 0, s>>>(V);
    FSAMD_HIP(hipGetLastError());
    const int chunks_a = (npts - imin + kVorThreads - 1) / kVorThreads;
    if (chunks_a > 0) {
        vor_wrap_kernel<<<dim3(chunks_a, L), kVorThreads, 0, s>>>(V);
        FSAMD_HIP(hipGetLastError());
    }
    vor_cells_kernel<<<dim3((npts + kVorThreads - 1) / kVorThreads, L), kVorThreads, 0, s>>>(V);
    FSAMD_HIP(hipGetLastError());
    vor_segments_kernel<<<L, kVorThreads, 0, s>>>(V);
    FSAMD_HIP(hipGetLastError());
    std::vector<int> herr(L);
    FSAMD_HIP(hipMemcpyAsync(herr.data(), V.err, sizeof(int) * L, hipMemcpyDeviceToHost, s));
    FSAMD_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < L; ++i)
        if (herr[i])
            return fail(FSAMD_ERR_ARG, "Voronoi cell assignment: a particle owns two separate runs of sightline " +
                                           std::to_string(i) + " (the reference exits here, index_table.cpp:199-206)");
    P->seg_ready = true;
    return FSAMD_OK;
}

// ---------------------------------------------------------------------------
// Host side

// Device memory kept across plans.  Buffers recycled where no queued work can
// still use them (after a synchronization: a plan's own buffers when it is
// destroyed, the binning's temporaries at its end) go to a per-device free
// list instead of hipFree, so that plans of one shape made one after another
// (the drop-in's plan per call, a loop over snapshots) reuse the same blocks
// instead of allocating and freeing some 20 GB of pair buffers each time.
// Sizes are rounded (256 B, or 2 MiB above 2 MiB); a request takes the
// smallest pooled block that fits with at most 1/4 + 2 MiB to spare.  The
// pool keeps at most kPoolCap bytes per device; a hipMalloc that fails
// empties the calling device's pool and tries again.
namespace {
constexpr size_t kPoolCap = 64ull << 30;
constexpr size_t kBig = 2ull << 20;
std::mutex g_pool_mu;
std::map<int, std::multimap<size_t, void *>> g_pool;
std::map<int, size_t> g_pool_bytes;

size_t pool_round(size_t n) { return n >= kBig ? (n + kBig - 1) & ~(kBig - 1) : (n + 255) & ~(size_t)255; }

bool pool_take(int dev, size_t n, void *&p, size_t &bytes)
{
    std::lock_guard<std::mutex> g(g_pool_mu);
    auto &m = g_pool[dev];
    auto it = m.lower_bound(n);
    if (it == m.end() || it->first > n + n / 4 + kBig) return false;
    p = it->second;
    bytes = it->first;
    g_pool_bytes[dev] -= bytes;
    m.erase(it);
    return true;
}

bool pool_put(int dev, void *p, size_t bytes)
{
    std::lock_guard<std::mutex> g(g_pool_mu);
    if (g_pool_bytes[dev] + bytes > kPoolCap) return false;
    g_pool[dev].emplace(bytes, p);
    g_pool_bytes[dev] += bytes;
    return true;
}

void pool_flush(int dev)
{
    std::multimap<size_t, void *> m;
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        m.swap(g_pool[dev]);
        g_pool_bytes[dev] = 0;
    }
    for (auto &e : m) (void)hipFree(e.second);
}
}  // namespace

}  // namespace fsamd

extern "C" int fsamd_pool_release(void)
{
    int d = 0;
    FSAMD_HIP(hipGetDevice(&d));
    FSAMD_HIP(hipDeviceSynchronize());
    fsamd::pool_flush(d);
    return FSAMD_OK;
}

namespace fsamd {

DevBuf::~DevBuf() { release(); }

void DevBuf::release()
{
    if (ptr && !borrowed) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
    dev = -1;
}

void DevBuf::recycle()
{
    if (ptr && !borrowed && dev >= 0 && pool_put(dev, ptr, bytes)) {
        ptr = nullptr;
        bytes = 0;
        dev = -1;
        return;
    }
    release();
}

int DevBuf::alloc(size_t n, Arena *arena)
{
    release();
    if (n == 0) n = 16;
    if (arena) {
        const size_t at = (arena->used + 255) & ~(size_t)255;
        if (at > arena->cap || n > arena->cap - at)
            return fail(FSAMD_ERR_UNSUPPORTED, "this call needs more device memory than the " +
                                                   std::to_string(arena->cap >> 20) +
                                                   " MiB scratch of the allocation-free entry; "
                                                   "use the allocating entry point");
        ptr = arena->base + at;
        bytes = n;
        borrowed = true;
        arena->used = at + n;
        return FSAMD_OK;
    }
    int d = 0;
    FSAMD_HIP(hipGetDevice(&d));
    n = pool_round(n);
    if (pool_take(d, n, ptr, bytes)) {
        dev = d;
        return FSAMD_OK;
    }
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        pool_flush(d);
        e = hipMalloc(&ptr, n);
    }
    if (e != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        return fail(FSAMD_ERR_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
    dev = d;
    return FSAMD_OK;
}

static int choose_cells(int n)
{
    if (n <= 0) return 1;
    int c = (int)std::ceil(std::sqrt((double)n));
    return std::max(1, std::min(c, 4096));
}

// The sightline grid of a plan (IndexTable ctor, index_table.cpp:7-18):
// copies of axis/cofm, per-axis cell counts, cell starts, sightline ids and
// (key, second) coordinates in cell order.
static int grid_build(fsamd_plan *P, double box, const int *axis, const double *cofm, int nlos, hipStream_t s)
{
    int rc;
    P->box = box;
    P->nlos = nlos;
    if ((rc = P->axis.alloc(sizeof(int) * (size_t)nlos, P->arena))) return rc;
    if ((rc = P->cofm.alloc(sizeof(double) * 3 * (size_t)nlos, P->arena))) return rc;
    if (nlos > 0) {
        FSAMD_HIP(hipMemcpyAsync(P->axis.ptr, axis, sizeof(int) * nlos, hipMemcpyDeviceToDevice, s));
        FSAMD_HIP(hipMemcpyAsync(P->cofm.ptr, cofm, sizeof(double) * 3 * nlos, hipMemcpyDeviceToDevice, s));
    }
    // --- sightline grid
    DevBuf counts3;
    if ((rc = counts3.alloc(sizeof(int) * 4, P->arena))) return rc;
    FSAMD_HIP(hipMemsetAsync(counts3.ptr, 0, sizeof(int) * 4, s));
    if (nlos > 0) {
        los_axis_count_kernel<<<(nlos + 255) / 256, 256, 0, s>>>(P->axis.as<int>(), nlos, counts3.as<int>());
        FSAMD_HIP(hipGetLastError());
    }
    int hc[4] = {0, 0, 0, 0};
    FSAMD_HIP(hipMemcpyAsync(hc, counts3.ptr, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
    FSAMD_HIP(hipStreamSynchronize(s));
    LosGrid &g = P->grid;
    int ncells = 0;
    for (int a = 0; a < 3; ++a) {
        g.nlos_axis[a] = hc[a];
        g.nc[a] = choose_cells(hc[a]);
        g.cw[a] = box / g.nc[a];
        g.inv_cw[a] = g.nc[a] / box;
        g.base[a] = ncells;
        ncells += g.nc[a] * g.nc[a];
    }
    P->ncells = ncells;
    DevBuf cell_of, cell_count, cell_cursor, scan_ws, cell_start64;
    if ((rc = cell_of.alloc(sizeof(int) * (size_t)std::max(nlos, 1), P->arena))) return rc;
    if ((rc = cell_count.alloc(sizeof(int) * (size_t)(ncells + 1), P->arena))) return rc;
    if ((rc = cell_start64.alloc(sizeof(long long) * (size_t)(ncells + 1), P->arena))) return rc;
    if ((rc = P->cell_start.alloc(sizeof(int) * (size_t)(ncells + 1), P->arena))) return rc;
    if ((rc = P->cell_ids.alloc(sizeof(int) * (size_t)std::max(nlos, 1), P->arena))) return rc;
    if ((rc = P->cell_kl.alloc(sizeof(double2) * (size_t)std::max(nlos, 1), P->arena))) return rc;
    if ((rc = scan_ws.alloc(exclusive_scan_ws_bytes((long long)ncells + 1), P->arena))) return rc;
    if ((rc = launch_fill_i32(cell_count.as<int>(), 0, ncells + 1, s))) return rc;
    if (nlos > 0) {
        los_cell_kernel<<<(nlos + 255) / 256, 256, 0, s>>>(P->axis.as<int>(), P->cofm.as<double>(), nlos, box, g,
                                                           cell_of.as<int>(), cell_count.as<int>());
        FSAMD_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan_i32_to_i64(cell_count.as<int>(), cell_start64.as<long long>(), ncells,
                                        scan_ws.ptr, scan_ws.bytes, s)))
        return rc;
    i64_to_i32_kernel<<<grid_for(ncells + 1, 256), 256, 0, s>>>(cell_start64.as<long long>(),
                                                                P->cell_start.as<int>(), ncells + 1);
    FSAMD_HIP(hipGetLastError());
    if ((rc = cell_cursor.alloc(sizeof(int) * (size_t)(ncells + 1), P->arena))) return rc;
    FSAMD_HIP(hipMemcpyAsync(cell_cursor.ptr, P->cell_start.ptr, sizeof(int) * (ncells + 1),
                             hipMemcpyDeviceToDevice, s));
    if (nlos > 0) {
        los_scatter_kernel<<<(nlos + 255) / 256, 256, 0, s>>>(P->axis.as<int>(), P->cofm.as<double>(), nlos,
                                                              cell_of.as<int>(), cell_cursor.as<int>(),
                                                              P->cell_ids.as<int>(), P->cell_kl.as<double2>());
        FSAMD_HIP(hipGetLastError());
    }
    g.cell_start = P->cell_start.as<int>();
    g.cell_ids = P->cell_ids.as<int>();
    g.cell_kl = P->cell_kl.as<double2>();
    return FSAMD_OK;
}

int plan_build(fsamd_plan *P, double box, const float *pos, const float *h, long long npart,
               const int *axis, const double *cofm, int nlos, hipStream_t s)
{
    int rc;
    P->npart = npart;
    if ((rc = grid_build(P, box, axis, cofm, nlos, s))) return rc;
    const LosGrid &g = P->grid;
    DevBuf scan_ws;
    if ((rc = scan_ws.alloc(exclusive_scan_ws_bytes(npart + 1), P->arena))) return rc;

    // --- count, scan, emit
    BinArgs A;
    A.box = box;
    A.pos = pos;
    A.h = h;
    A.npart = npart;
    A.grid = g;
    DevBuf cnt;
    if ((rc = cnt.alloc(sizeof(int) * (size_t)std::max<long long>(npart, 1), P->arena))) return rc;
    if ((rc = P->poff.alloc(sizeof(long long) * (size_t)(npart + 1), P->arena))) return rc;
    if (npart > 0) {
        bin_count_kernel<<<grid_for(npart, 256, 1 << 20), 256, 0, s>>>(A, cnt.as<int>());
        FSAMD_HIP(hipGetLastError());
    }
    if ((rc = exclusive_scan_i32_to_i64(cnt.as<int>(), P->poff.as<long long>(), npart, scan_ws.ptr,
                                        scan_ws.bytes, s)))
        return rc;
    long long M = 0;
    FSAMD_HIP(hipMemcpyAsync(&M, P->poff.as<long long>() + npart, sizeof(long long), hipMemcpyDeviceToHost, s));
    FSAMD_HIP(hipStreamSynchronize(s));
    cnt.recycle();
    P->npairs = M;
    DevBuf k0, v0, k1, v1;
    const size_t mbytes = sizeof(int) * (size_t)std::max<long long>(M, 1);
    if ((rc = k0.alloc(mbytes, P->arena)) || (rc = v0.alloc(mbytes, P->arena)) || (rc = k1.alloc(mbytes, P->arena)) || (rc = v1.alloc(mbytes, P->arena)))
        return rc;
    if (npart > 0 && M > 0) {
        bin_emit_kernel<<<grid_for(npart, 256, 1 << 20), 256, 0, s>>>(A, P->poff.as<long long>(), k0.as<int>(),
                                                                      v0.as<int>());
        FSAMD_HIP(hipGetLastError());
    }
    // --- stable radix sort by sightline id, ping-pong between (k0,v0) and (k1,v1)
    int total_bits = 1;
    while ((1ll << total_bits) < (long long)nlos) ++total_bits;
    const int passes = (total_bits + kRsMaxBits - 1) / kRsMaxBits;
    const int bits_per = (total_bits + passes - 1) / passes;
    DevBuf *kin = &k0, *vin = &v0, *kout = &k1, *vout = &v1;
    DevBuf table, tws;
    if (M > 0) {
        const long long ntiles = (M + kRsTile - 1) / kRsTile;
        if ((rc = table.alloc(sizeof(long long) * (size_t)((1 << bits_per) * ntiles + 1), P->arena))) return rc;
        if ((rc = tws.alloc(exclusive_scan_ws_bytes((1 << bits_per) * ntiles + 1), P->arena))) return rc;
        for (int ps = 0; ps < passes; ++ps) {
            const int shift = ps * bits_per;
            const int bits = std::min(bits_per, total_bits - shift);
            rs_upsweep_kernel<<<(int)ntiles, kRsThreads, 0, s>>>(kin->as<int>(), M, shift, bits, ntiles,
                                                                 table.as<long long>());
            FSAMD_HIP(hipGetLastError());
            if ((rc = exclusive_scan_i64_inplace(table.as<long long>(), (1ll << bits) * ntiles, tws.ptr,
                                                 tws.bytes, s)))
                return rc;
            rs_downsweep_kernel<<<(int)ntiles, kRsThreads, 0, s>>>(kin->as<int>(), vin->as<int>(), kout->as<int>(),
                                                                   vout->as<int>(), M, shift, bits, ntiles,
                                                                   table.as<long long>());
            FSAMD_HIP(hipGetLastError());
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
    }
    // (kin, vin) hold the sorted sightline ids and particle ids
    P->pair_part.take(*vin);
    if ((rc = P->line_off.alloc(sizeof(long long) * (size_t)(nlos + 1), P->arena))) return rc;
    row_offsets_kernel<<<(nlos + 1 + 255) / 256, 256, 0, s>>>(kin->as<int>(), M, nlos,
                                                             P->line_off.as<long long>());
    FSAMD_HIP(hipGetLastError());

    // --- work order: sightlines by descending pair count (stable)
    std::vector<long long> off(nlos + 1);
    FSAMD_HIP(hipMemcpyAsync(off.data(), P->line_off.ptr, sizeof(long long) * (nlos + 1),
                             hipMemcpyDeviceToHost, s));
    FSAMD_HIP(hipStreamSynchronize(s));
    // the binning's temporaries are idle now
    for (DevBuf *b : {&scan_ws, kin, kout, vout, &table, &tws}) b->recycle();
    P->max_row = 0;
    for (int i = 0; i < nlos; ++i) P->max_row = std::max(P->max_row, off[i + 1] - off[i]);
    std::vector<int> order(nlos);
    if (P->max_row <= 4 * (long long)nlos + (1 << 20)) {
        // counting sort on the pair count, descending, stable
        std::vector<int> first((size_t)P->max_row + 2, 0);
        for (int i = 0; i < nlos; ++i) ++first[(size_t)(P->max_row - (off[i + 1] - off[i])) + 1];
        for (size_t c = 1; c < first.size(); ++c) first[c] += first[c - 1];
        for (int i = 0; i < nlos; ++i) order[first[(size_t)(P->max_row - (off[i + 1] - off[i]))]++] = i;
    } else {
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return off[a + 1] - off[a] > off[b + 1] - off[b];
        });
    }
    if ((rc = P->order.alloc(sizeof(int) * (size_t)std::max(nlos, 1), P->arena))) return rc;
    if (nlos > 0)
        FSAMD_HIP(hipMemcpy(P->order.ptr, order.data(), sizeof(int) * nlos, hipMemcpyHostToDevice));
    P->order_host = std::move(order);
    P->off_host = std::move(off);
    return FSAMD_OK;
}

int near_lines_direct(double box, const float *pos, const float *h, long long npart, const int *axis,
                      const double *cofm, int nlos, int *out, long long *n_out, hipStream_t s)
{
    int rc;
    *n_out = 0;
    if (npart == 0 || nlos == 0) return FSAMD_OK;
    fsamd_plan G;   // holds the grid's buffers for this call
    G.npart = npart;
    if ((rc = grid_build(&G, box, axis, cofm, nlos, s))) return rc;
    BinArgs A;
    A.box = box;
    A.pos = pos;
    A.h = h;
    A.npart = npart;
    A.grid = G.grid;
    DevBuf flag, fscan, ws;
    if ((rc = flag.alloc(sizeof(int) * (size_t)npart))) return rc;
    if ((rc = fscan.alloc(sizeof(long long) * (size_t)(npart + 1)))) return rc;
    if ((rc = ws.alloc(exclusive_scan_ws_bytes(npart + 1)))) return rc;
    bin_any_kernel<<<grid_for(npart, 256, 1 << 20), 256, 0, s>>>(A, flag.as<int>());
    FSAMD_HIP(hipGetLastError());
    if ((rc = exclusive_scan_i32_to_i64(flag.as<int>(), fscan.as<long long>(), npart, ws.ptr, ws.bytes, s)))
        return rc;
    flag_scatter_kernel<<<grid_for(npart, 256), 256, 0, s>>>(flag.as<int>(), fscan.as<long long>(), npart, out);
    FSAMD_HIP(hipGetLastError());
    long long cnt = 0;
    FSAMD_HIP(hipMemcpyAsync(&cnt, fscan.as<long long>() + npart, sizeof(long long), hipMemcpyDeviceToHost, s));
    FSAMD_HIP(hipStreamSynchronize(s));
    for (DevBuf *b : {&flag, &fscan, &ws}) b->recycle();
    *n_out = cnt;
    return FSAMD_OK;
}

int plan_near_lines(const fsamd_plan *P, int *out, long long *n_out, hipStream_t s)
{
    int rc;
    const long long n = P->npart;
    if (n == 0) {
        *n_out = 0;
        return FSAMD_OK;
    }
    DevBuf flag, fscan, ws;
    if ((rc = flag.alloc(sizeof(int) * (size_t)n))) return rc;
    if ((rc = fscan.alloc(sizeof(long long) * (size_t)(n + 1)))) return rc;
    if ((rc = ws.alloc(exclusive_scan_ws_bytes(n + 1)))) return rc;
    near_flag_kernel<<<grid_for(n, 256), 256, 0, s>>>(P->poff.as<long long>(), n, flag.as<int>());
    FSAMD_HIP(hipGetLastError());
    if ((rc = exclusive_scan_i32_to_i64(flag.as<int>(), fscan.as<long long>(), n, ws.ptr, ws.bytes, s)))
        return rc;
    near_scatter_kernel<<<grid_for(n, 256), 256, 0, s>>>(P->poff.as<long long>(), fscan.as<long long>(), n, out);
    FSAMD_HIP(hipGetLastError());
    long long cnt = 0;
    FSAMD_HIP(hipMemcpyAsync(&cnt, fscan.as<long long>() + n, sizeof(long long), hipMemcpyDeviceToHost, s));
    FSAMD_HIP(hipStreamSynchronize(s));
    for (DevBuf *b : {&flag, &fscan, &ws}) b->recycle();
    *n_out = cnt;
    return FSAMD_OK;
}

}  // namespace fsamd
