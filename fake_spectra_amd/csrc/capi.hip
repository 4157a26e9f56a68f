// C ABI of libfsamd (declared in include/fake_spectra_amd.h).
#include "fsamd_internal.h"
#include "plan.h"
#include "deposit.h"
#include "voigt.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace fsamd {

static thread_local std::string g_error;
static thread_local long long g_stats[kNStats] = {};
static thread_local int g_stats_on = 0;    // device counters (and timing)
static thread_local int g_timing_on = 0;   // hipEvents around the launches
static thread_local double g_last_ms = 0;          // whole call
static thread_local double g_last_kernel_ms = 0;   // deposit kernel launches only
static std::atomic<int> g_group_size{16};   // pixels per lane group of the tau kernel (16, 32 or 64)
static std::atomic<int> g_block_waves{8};   // wavefronts per block (per sightline work item): 2, 4 or 8
static thread_local int g_fp64 = 0;         // every pair on the FP64 path (fsamd_set_fp64)

void set_error(const std::string &msg) { g_error = msg; }

int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

// Pairs per work item in the one-sightline-per-launch form: a launch spreads
// one sightline's pairs over blocks of this many pairs.
constexpr long long kPlainChunk = 64;

struct DeviceInfo {
    int ncu = 0;
    int max_lds = 0;
};

static int device_info(DeviceInfo &d)
{
    int dev = 0;
    FSAMD_HIP(hipGetDevice(&dev));
    FSAMD_HIP(hipDeviceGetAttribute(&d.ncu, hipDeviceAttributeMultiprocessorCount, dev));
    FSAMD_HIP(hipDeviceGetAttribute(&d.max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return FSAMD_OK;
}

static void line_consts(const fsamd_line &l, LineDesc &d)
{
    // absorption.cpp:152-161
    d.sigma_a = std::sqrt(3.0 * kPi * kSigmaT / 8.0) * l.lambda_cm * l.fosc;
    d.bfac = std::sqrt(2.0 * kBoltzmann / (l.amumass * kProtonMass)) / 1e5;
    d.voigt_fac = l.gamma * l.lambda_cm / (4. * kPi) / 1e5;
    d.dens = l.dens;
    d.out = l.out;
}

static bool find_item_set(const fsamd_plan *P, int K, long long chunk)
{
    for (auto &it : P->item_sets)
        if (it->K == K && it->chunk == chunk) return true;
    return false;
}

static long long tune_key(int nbins, int nlines, int kernel)
{
    return ((long long)nbins << 20) ^ ((long long)nlines << 4) ^ kernel;
}

// Work items for K lines in the plan's heaviest-first order; a sightline
// with more than `chunk` pairs is split into non-direct items.  Cached.
static int item_set(fsamd_plan *P, int K, long long chunk, fsamd_plan::ItemSet **out)
{
    for (auto &it : P->item_sets)
        if (it->K == K && it->chunk == chunk) {
            *out = it.get();
            return FSAMD_OK;
        }
    auto set = std::make_unique<fsamd_plan::ItemSet>();
    set->K = K;
    set->chunk = chunk;
    set->split = false;
    std::vector<WorkItem> &items = set->host;
    const long long stride = chunk < 0 ? -chunk : 1;
    if (chunk < 0) chunk = 1ll << 62;
    long long nsel = 0;
    for (int s : P->order_host) {
        const long long a = P->off_host[s], b = P->off_host[s + 1];
        const long long n = b - a;
        if (nsel++ % stride) continue;
        for (int k = 0; k < K; ++k) {
            if (n <= chunk) {
                items.push_back(WorkItem{s, k, a, b, 1, 0});
            } else {
                set->split = true;
                for (long long c = a; c < b; c += chunk)
                    items.push_back(WorkItem{s, k, c, std::min(b, c + chunk), 0, 0});
            }
        }
    }
    if (items.size() >= (size_t)INT32_MAX) return fail(FSAMD_ERR_ARG, "too many work items");
    int rc = set->buf.alloc(sizeof(WorkItem) * std::max<size_t>(items.size(), 1), P->arena);
    if (rc) return rc;
    if (!items.empty())
        FSAMD_HIP(hipMemcpy(set->buf.ptr, items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice));
    P->item_sets.push_back(std::move(set));
    *out = P->item_sets.back().get();
    return FSAMD_OK;
}

// (Re)allocate a launch-state buffer.  A buffer that a captured graph may
// still address is retired rather than freed.
static int regrow(fsamd_plan *P, DevBuf &b, size_t bytes)
{
    if (P->captured && b.ptr && !b.borrowed) P->retired.push_back(b.disown());
    return b.alloc(bytes, P->arena);
}

static int grow_fixups(fsamd_plan *P, long long cap, hipStream_t s)
{
    cap = std::min<long long>(cap, INT32_MAX / 2);
    FSAMD_HIP(hipStreamSynchronize(s));
    int rc = regrow(P, P->fix, sizeof(Fixup) * (size_t)cap);
    if (rc) return rc;
    P->fix_cap = (int)cap;
    return FSAMD_OK;
}

// The device LineDesc table for this call.  Eager calls rewrite the shared
// table when the lines (or their output pointers) change; a captured call
// needs a table with exactly its contents made by an earlier eager call, and
// pins it so that later calls never rewrite what the graph reads.
static int line_table(fsamd_plan *P, const std::vector<LineDesc> &ld, bool capturing, hipStream_t s,
                      const LineDesc **out)
{
    const size_t bytes = sizeof(LineDesc) * ld.size();
    auto same = [&](const fsamd_plan::LineTable &t) {
        return t.host.size() == ld.size() && std::memcmp(t.host.data(), ld.data(), bytes) == 0;
    };
    for (auto &t : P->pinned_lines)
        if (same(*t)) {
            *out = t->buf.as<LineDesc>();
            return FSAMD_OK;
        }
    fsamd_plan::LineTable &e = P->eager_lines;
    if (capturing) {
        if (!same(e)) return fail(FSAMD_ERR_ARG, "run one uncaptured call with the same lines and outputs before capturing");
        auto t = std::make_unique<fsamd_plan::LineTable>();
        t->buf.take(e.buf);
        t->host.swap(e.host);
        *out = t->buf.as<LineDesc>();
        P->pinned_lines.push_back(std::move(t));
        return FSAMD_OK;
    }
    if (!same(e)) {
        FSAMD_HIP(hipStreamSynchronize(s));   // the previous table may still be in use
        if (e.buf.bytes < bytes) {
            int rc = e.buf.alloc(bytes, P->arena);
            if (rc) return rc;
        }
        FSAMD_HIP(hipMemcpy(e.buf.ptr, ld.data(), bytes, hipMemcpyHostToDevice));
        e.host = ld;
    }
    *out = e.buf.as<LineDesc>();
    return FSAMD_OK;
}

static int run_deposit(bool tau, fsamd_plan *P, int nbins, int kernel, double velfac, double tautail,
                       const float *pos, const float *vel, const float *temp, const float *h,
                       const fsamd_line *lines, int nlines, int mode, hipStream_t s, long long sample_stride = 0,
                       int force_g = 0, int force_waves = 0)
{
    if (!P) return fail(FSAMD_ERR_ARG, "plan is NULL");
    if (nbins < 1) return fail(FSAMD_ERR_ARG, "nbins must be positive");
    if (kernel == kVoronoiMesh && !P->seg_ready)
        return fail(FSAMD_ERR_ARG, "kernel 2 (Voronoi mesh) needs fsamd_plan_assign_cells on the plan first");
    if (nlines < 1 || !lines) return fail(FSAMD_ERR_ARG, "at least one line is required");
    if (mode != FSAMD_MODE_BATCHED && mode != FSAMD_MODE_PLAIN) return fail(FSAMD_ERR_ARG, "unknown mode");
    if ((long long)P->nlos * nbins >= (1ll << 62)) return fail(FSAMD_ERR_ARG, "output too large");
    for (int i = 0; i < nlines; ++i)
        if (!lines[i].out || !lines[i].dens) return fail(FSAMD_ERR_ARG, "line without dens or out");
    static thread_local DeviceInfo dev;
    int rc;
    if (dev.ncu == 0 && (rc = device_info(dev))) return rc;
    const bool tuned = P->tuned_g > 0 && P->tuned_key == tune_key(nbins, nlines, kernel);
    const int gsize = force_g ? force_g : tuned ? P->tuned_g : g_group_size.load(std::memory_order_relaxed);
    int waves = force_waves ? force_waves : tuned ? P->tuned_waves : g_block_waves.load(std::memory_order_relaxed);
    // rows in LDS when the row and the kernel's own LDS fit a CU; a block
    // too big for that falls back to 4 wavefronts before rows go to HBM
    auto fits = [&](int wv) {
        return deposit_lds_bytes(nbins) + deposit_static_lds(tau, gsize, effective_waves(true, g_stats_on, wv)) <=
               (size_t)dev.max_lds;
    };
    bool lds = fits(waves);
    if (!lds && fits(4)) {
        waves = 4;
        lds = true;
    }
    const int ngroups = (nlines + kMaxLines - 1) / kMaxLines;

    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    FSAMD_HIP(hipStreamIsCapturing(s, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    const bool plain = (mode == FSAMD_MODE_PLAIN);
    if (capturing && (g_timing_on || plain))
        return fail(FSAMD_ERR_ARG, "graph capture supports the batched mode without statistics only");
    std::vector<LineDesc> ld(nlines);
    for (int i = 0; i < nlines; ++i) line_consts(lines[i], ld[i]);
    const LineDesc *table = nullptr;
    if ((rc = line_table(P, ld, capturing, s, &table))) return rc;
    if (P->queue.bytes < sizeof(int) * (size_t)ngroups) {
        if (capturing) return fail(FSAMD_ERR_ARG, "run one uncaptured call before capturing");
        if ((rc = regrow(P, P->queue, sizeof(int) * ngroups))) return rc;
    }
    if (g_stats_on) {
        if (!P->stats.ptr && (rc = P->stats.alloc(sizeof(long long) * kNStats, P->arena))) return rc;
        if ((rc = launch_fill_f64((double *)P->stats.ptr, 0.0, kNStats, s))) return rc;   // all-zero bits
    }
    if (g_timing_on) {
        for (auto &e : P->ev)
            if (!e) FSAMD_HIP(hipEventCreate(&e));
        FSAMD_HIP(hipEventRecord(P->ev[0], s));
        g_last_kernel_ms = 0;
    }
    const long long chunk_key = plain ? kPlainChunk : (sample_stride > 1 ? -sample_stride : (1ll << 62));
    const bool rec = tau && P->keep_windows;
    if (rec) {
        const size_t wbytes = sizeof(int) * 2 * (size_t)std::max<long long>(P->npairs, 1) * nlines;
        if (P->win.bytes < wbytes) {
            if (capturing) return fail(FSAMD_ERR_ARG, "run one uncaptured fsamd_tau call before capturing");
            FSAMD_HIP(hipStreamSynchronize(s));
            if ((rc = regrow(P, P->win, wbytes))) return rc;
        }
        if ((rc = launch_fill_i32(P->win.as<int>(), 0, (long long)(wbytes / sizeof(int)), s))) return rc;
        P->win_lines = nlines;
        P->win_nbins = nbins;
        P->win_kernel = kernel;
        ++P->win_gen;
    }
    if (tau) {
        // room for deferred FP64 walks: a few percent of pair-directions
        const long long want = g_fp64 ? P->npairs * std::min(nlines, kMaxLines) + 65536
                                      : std::max<long long>(65536, 2 * P->npairs * std::min(nlines, kMaxLines) / 32);
        if ((P->fix_cap < want || !P->fix_count.ptr) && capturing)
            return fail(FSAMD_ERR_ARG, "run one uncaptured fsamd_tau call before capturing");
        if (P->fix_cap < want && (rc = grow_fixups(P, want, s))) return rc;
        if (!P->fix_count.ptr && (rc = P->fix_count.alloc(sizeof(unsigned long long) * 2, P->arena))) return rc;
    }
    if (tau) {
        const size_t rbytes = sizeof(PartRec) * (size_t)std::max<long long>(P->npart, 1);
        if (P->rec.bytes < rbytes) {
            if (capturing) return fail(FSAMD_ERR_ARG, "run one uncaptured fsamd_tau call before capturing");
            FSAMD_HIP(hipStreamSynchronize(s));
            if ((rc = regrow(P, P->rec, rbytes))) return rc;
        }
    }
    if (capturing) P->captured = true;
    for (int g = 0; g < ngroups; ++g) {
        const int k0 = g * kMaxLines;
        const int K = std::min(kMaxLines, nlines - k0);
        DepositArgs A;
        std::memset(&A, 0, sizeof(A));
        A.nbins = nbins;
        A.kernel = kernel;
        A.velfac = velfac;
        A.inv_velfac = 1.0 / velfac;
        A.vbox = P->box * velfac;     // absorption.cpp:157
        A.box = P->box;
        A.tautail = tautail;
        A.force_fp64 = g_fp64;
        A.pos = pos;
        A.vel = vel;
        A.temp = temp;
        A.h = h;
        A.axis = P->axis.as<int>();
        A.cofm = P->cofm.as<double>();
        A.line_off = P->line_off.as<long long>();
        A.pair_part = P->pair_part.as<int>();
        A.seg = kernel == kVoronoiMesh ? P->seg.as<float2>() : nullptr;
        A.nlines = K;
        A.lines = table + k0;
        A.rec = tau ? P->rec.as<PartRec>() : nullptr;
        if (tau && (rc = launch_pack_records(pos, vel, temp, h, A.lines, K, P->npart, P->rec.as<PartRec>(), s)))
            return rc;
        A.queue = P->queue.as<int>() + g;
        // whole-sightline items: a wavefront sets up to kWaveClaim pairs at a
        // time; the one-launch-per-sightline mode, whose items are short, one
        // per lane group
        A.claim = plain ? 64 / gsize : kWaveClaim;
        A.stats = g_stats_on ? P->stats.as<long long>() : nullptr;
        A.npairs = P->npairs;
        A.win = rec ? P->win.as<int>() + (size_t)k0 * 2 * P->npairs : nullptr;
        fsamd_plan::ItemSet *set = nullptr;
        if (capturing && !find_item_set(P, K, chunk_key))
            return fail(FSAMD_ERR_ARG, "run one uncaptured fsamd_tau call before capturing");
        if ((rc = item_set(P, K, chunk_key, &set))) return rc;
        const std::vector<WorkItem> &items = set->host;
        for (int attempt = 0;; ++attempt) {
            A.fix = P->fix.as<Fixup>();
            A.fix_cap = P->fix_cap;
            A.fix_count = P->fix_count.as<unsigned long long>();
            if (set->split || !lds) {
                for (int k = 0; k < K; ++k)
                    if ((rc = launch_fill_f64(ld[k0 + k].out, 0.0, (long long)P->nlos * nbins, s))) return rc;
            }
            if (items.empty()) break;
            // captured graphs keep the overflow flag sticky across replays
            if (tau && (rc = launch_fill_i32((int *)A.fix_count, 0, capturing ? 2 : 4, s))) return rc;
            if (g_timing_on) FSAMD_HIP(hipEventRecord(P->ev[2], s));
            if (!plain) {
                A.nitems = (int)items.size();
                if ((rc = launch_fill_i32(A.queue, 0, 1, s))) return rc;
                const int per_cu = deposit_blocks_per_cu(tau, lds, gsize, effective_waves(lds, g_stats_on, waves), nbins);
                const int nblocks = (int)std::min<long long>((long long)per_cu * dev.ncu, A.nitems);
                if (g_timing_on) FSAMD_HIP(hipEventRecord(P->ev[2], s));
                if ((rc = launch_deposit(tau, lds, gsize, waves, A, set->buf.as<WorkItem>(), 0, nblocks, 1, s)))
                    return rc;
            } else {
                // one launch per (sightline, line), in the plan's item order
                size_t i = 0;
                while (i < items.size()) {
                    size_t j = i;
                    while (j < items.size() && items[j].los == items[i].los && items[j].line == items[i].line) ++j;
                    A.nitems = (int)(j - i);
                    if ((rc = launch_deposit(tau, lds, gsize, waves, A, set->buf.as<WorkItem>(), (int)i, A.nitems, 0, s)))
                        return rc;
                    i = j;
                }
            }
            if (g_timing_on) {
                FSAMD_HIP(hipEventRecord(P->ev[3], s));
                FSAMD_HIP(hipEventSynchronize(P->ev[3]));
                float kms = 0;
                FSAMD_HIP(hipEventElapsedTime(&kms, P->ev[2], P->ev[3]));
                g_last_kernel_ms += kms;
            }
            if (!tau) break;
            if ((rc = launch_tau_fixup(A, set->buf.as<WorkItem>(), dev.ncu * 2, s))) return rc;
            if (capturing) break;   // overflow is checked after replay (fsamd_plan_overflow)
            unsigned long long fc[2] = {0, 0};
            FSAMD_HIP(hipMemcpyAsync(fc, A.fix_count, sizeof(fc), hipMemcpyDeviceToHost, s));
            FSAMD_HIP(hipStreamSynchronize(s));
            if (!fc[1]) break;
            // the list overflowed: grow it and run this line group again
            if (attempt > 2) return fail(FSAMD_ERR_HIP, "fix-up list overflow persists");
            if (fc[0] > (unsigned long long)(INT32_MAX / 2))
                return fail(FSAMD_ERR_UNSUPPORTED, "more than 2^30 walks of one line group need the FP64 path; "
                                                   "split the sightlines over several plans");
            if ((rc = grow_fixups(P, 2ll * (long long)fc[0] + 65536, s))) return rc;
            if (g_stats_on && (rc = launch_fill_f64((double *)P->stats.ptr, 0.0, kNStats, s))) return rc;
            if (g_timing_on) {
                float kms = 0;
                FSAMD_HIP(hipEventElapsedTime(&kms, P->ev[2], P->ev[3]));
                g_last_kernel_ms -= kms;    // the rerun's launch replaces this one
            }
        }
    }
    if (g_timing_on) {
        FSAMD_HIP(hipEventRecord(P->ev[1], s));
        if (g_stats_on)
            FSAMD_HIP(hipMemcpyAsync(g_stats, P->stats.ptr, sizeof(long long) * kNStats, hipMemcpyDeviceToHost, s));
        FSAMD_HIP(hipStreamSynchronize(s));
        float ms = 0;
        FSAMD_HIP(hipEventElapsedTime(&ms, P->ev[0], P->ev[1]));
        g_last_ms = ms;
    }
    return FSAMD_OK;
}

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free entry (fsamd_particle_interpolate_noalloc):
// a file-scope array in the code object, one copy per GPU, held until the
// process ends.  Host code never names the array in device arguments: it
// uses the address hipGetSymbolAddress returns for the current device, looked
// up once per device on that device's first call.

constexpr size_t kScratchBytes = FSAMD_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;                        // guards g_scratch_ptr
static void *g_scratch_ptr[kMaxDevices];               // device address of `scratch`, per device
static std::mutex g_scratch_call_mu[kMaxDevices];      // one call at a time per device

static int scratch_for_device(int dev, void **ptr)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(FSAMD_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(FSAMD_ERR_HIP, std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return FSAMD_OK;
}

// Arguments of a plan, checked before anything runs (fsamd_plan_create).
static int check_plan_args(double box, const float *pos, const float *h, long long npart, const int *axis,
                           const double *cofm, int nlos)
{
    if (npart < 0 || nlos < 0) return fail(FSAMD_ERR_ARG, "negative size");
    if (npart >= (1ll << 31)) return fail(FSAMD_ERR_ARG, "particle indices must fit in int (N < 2^31)");
    if (!(box > 0) || !std::isfinite(box)) return fail(FSAMD_ERR_ARG, "box must be positive and finite");
    if ((npart > 0 && (!pos || !h)) || (nlos > 0 && (!axis || !cofm))) return fail(FSAMD_ERR_ARG, "NULL array");
    return FSAMD_OK;
}

}  // namespace fsamd

using namespace fsamd;

extern "C" {

const char *fsamd_last_error(void) { return g_error.c_str(); }

int fsamd_version(void) { return 100; }

int fsamd_plan_create(double box, const float *pos, const float *h, long long npart, const int *axis,
                      const double *cofm, int nlos, void *stream, fsamd_plan **out)
{
    if (!out) return fail(FSAMD_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (int rc = check_plan_args(box, pos, h, npart, axis, cofm, nlos)) return rc;
    fsamd_plan *P = new fsamd_plan();
    int rc = plan_build(P, box, pos, h, npart, axis, cofm, nlos, (hipStream_t)stream);
    if (rc) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        delete P;
        return rc;
    }
    *out = P;
    return FSAMD_OK;
}

int fsamd_plan_destroy(fsamd_plan *plan)
{
    delete plan;
    return FSAMD_OK;
}

long long fsamd_plan_npairs(const fsamd_plan *plan) { return plan ? plan->npairs : -1; }

int fsamd_plan_pairs(const fsamd_plan *plan, const long long **offsets, const int **parts)
{
    if (!plan) return fail(FSAMD_ERR_ARG, "plan is NULL");
    if (offsets) *offsets = plan->line_off.as<long long>();
    if (parts) *parts = plan->pair_part.as<int>();
    return FSAMD_OK;
}

int fsamd_plan_assign_cells(fsamd_plan *plan, const float *pos, void *stream)
{
    if (!plan) return fail(FSAMD_ERR_ARG, "plan is NULL");
    if (plan->npart > 0 && !pos) return fail(FSAMD_ERR_ARG, "pos is NULL");
    return plan_assign_cells(plan, pos, (hipStream_t)stream);
}

int fsamd_plan_near_lines(const fsamd_plan *plan, int *out, long long *n_out, void *stream)
{
    if (!plan || !n_out) return fail(FSAMD_ERR_ARG, "NULL argument");
    return plan_near_lines(plan, out, n_out, (hipStream_t)stream);
}

int fsamd_near_lines(double box, const float *pos, const float *h, long long npart, const int *axis,
                     const double *cofm, int nlos, int *out, long long *n_out, void *stream)
{
    if (!n_out) return fail(FSAMD_ERR_ARG, "n_out is NULL");
    *n_out = 0;
    if (int rc = check_plan_args(box, pos, h, npart, axis, cofm, nlos)) return rc;
    if (npart > 0 && nlos > 0 && !out) return fail(FSAMD_ERR_ARG, "out is NULL");
    return near_lines_direct(box, pos, h, npart, axis, cofm, nlos, out, n_out, (hipStream_t)stream);
}

int fsamd_tau(fsamd_plan *plan, int nbins, int kernel, double velfac, double tautail, const float *pos,
              const float *vel, const float *temp, const float *h, const fsamd_line *lines, int nlines,
              int mode, void *stream)
{
    return run_deposit(true, plan, nbins, kernel, velfac, tautail, pos, vel, temp, h, lines, nlines, mode,
                       (hipStream_t)stream);
}

int fsamd_colden(fsamd_plan *plan, int nbins, int kernel, double velfac, const float *pos, const float *h,
                 const fsamd_line *lines, int nlines, void *stream)
{
    return run_deposit(false, plan, nbins, kernel, velfac, 0.0, pos, nullptr, nullptr, h, lines, nlines,
                       FSAMD_MODE_BATCHED, (hipStream_t)stream);
}

int fsamd_particle_interpolate(int compute_tau, int nbins, int kernel, double box100, double velfac, double atime,
                               double lambda, double gamma, double fosc, double amumass, double tautail,
                               const float *pos, const float *vel, const float *dens, const float *temp,
                               const float *h, long long npart, const int *axis, const double *cofm, int nlos,
                               double *out, void *stream)
{
    (void)atime;
    fsamd_plan *plan = nullptr;
    int rc = fsamd_plan_create(box100, pos, h, npart, axis, cofm, nlos, stream, &plan);
    if (rc) return rc;
    if (kernel == kVoronoiMesh && (rc = plan_assign_cells(plan, pos, (hipStream_t)stream))) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        fsamd_plan_destroy(plan);
        return rc;
    }
    fsamd_line l;
    l.lambda_cm = lambda;
    l.gamma = gamma;
    l.fosc = fosc;
    l.amumass = amumass;
    l.dens = dens;
    l.out = out;
    if (compute_tau)
        rc = fsamd_tau(plan, nbins, kernel, velfac, tautail, pos, vel, temp, h, &l, 1, FSAMD_MODE_BATCHED, stream);
    else
        rc = fsamd_colden(plan, nbins, kernel, velfac, pos, h, &l, 1, stream);
    (void)hipStreamSynchronize((hipStream_t)stream);
    fsamd_plan_destroy(plan);
    return rc;
}

int fsamd_particle_interpolate_noalloc(int compute_tau, int nbins, int kernel, double box100, double velfac,
                                       double atime, double lambda, double gamma, double fosc, double amumass,
                                       double tautail, const float *pos, const float *vel, const float *dens,
                                       const float *temp, const float *h, long long npart, const int *axis,
                                       const double *cofm, int nlos, double *out, void *stream)
{
    (void)atime;
    hipStream_t s = (hipStream_t)stream;
    if (int rc = check_plan_args(box100, pos, h, npart, axis, cofm, nlos)) return rc;
    // per-particle and per-sightline buffers alone (packed records, pair
    // offsets and counts; sightline tables, grid, row offsets and work items)
    const double need = (double)npart * (sizeof(PartRec) + 2 * sizeof(long long) + sizeof(int)) +
                        (double)nlos * (sizeof(int) * 4 + sizeof(double) * 3 + sizeof(double2) + sizeof(long long) +
                                        sizeof(WorkItem));
    if (need > (double)kScratchBytes)
        return fail(FSAMD_ERR_UNSUPPORTED, "the allocation-free entry's scratch holds " +
                                               std::to_string(kScratchBytes >> 20) + " MiB and " +
                                               std::to_string(npart) + " particles and " + std::to_string(nlos) +
                                               " sightlines need more; use fsamd_particle_interpolate");
    int dev = 0;
    FSAMD_HIP(hipGetDevice(&dev));
    void *base = nullptr;
    int rc = scratch_for_device(dev, &base);
    if (rc) return rc;
    std::lock_guard<std::mutex> one_call(g_scratch_call_mu[dev]);
    // no counters or events in this entry
    const int stats_on = g_stats_on, timing_on = g_timing_on;
    g_stats_on = 0;
    g_timing_on = 0;
    Arena arena;
    arena.base = (char *)base;
    arena.cap = kScratchBytes;
    {
        fsamd_plan P;
        P.arena = &arena;
        rc = plan_build(&P, box100, pos, h, npart, axis, cofm, nlos, s);
        if (!rc && kernel == kVoronoiMesh) rc = plan_assign_cells(&P, pos, s);
        if (!rc) {
            fsamd_line l;
            l.lambda_cm = lambda;
            l.gamma = gamma;
            l.fosc = fosc;
            l.amumass = amumass;
            l.dens = dens;
            l.out = out;
            rc = compute_tau ? run_deposit(true, &P, nbins, kernel, velfac, tautail, pos, vel, temp, h, &l, 1,
                                           FSAMD_MODE_BATCHED, s)
                             : run_deposit(false, &P, nbins, kernel, velfac, 0.0, pos, nullptr, nullptr, h, &l, 1,
                                           FSAMD_MODE_BATCHED, s);
        }
        // the next call on this device reuses the scratch: drain this one first
        const hipError_t e = hipStreamSynchronize(s);
        if (!rc && e != hipSuccess) rc = fail(FSAMD_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    }
    g_stats_on = stats_on;
    g_timing_on = timing_on;
    return rc;
}

int fsamd_plan_overflow(fsamd_plan *plan, int *flag, void *stream)
{
    if (!plan || !flag) return fail(FSAMD_ERR_ARG, "NULL argument");
    *flag = 0;
    if (!plan->fix_count.ptr) return FSAMD_OK;
    unsigned long long fc[2] = {0, 0};
    FSAMD_HIP(hipMemcpyAsync(fc, plan->fix_count.ptr, sizeof(fc), hipMemcpyDeviceToHost, (hipStream_t)stream));
    FSAMD_HIP(hipStreamSynchronize((hipStream_t)stream));
    *flag = fc[1] != 0;
    return FSAMD_OK;
}

int fsamd_tau_autotune(fsamd_plan *P, int nbins, int kernel, double velfac, double tautail, const float *pos,
                       const float *vel, const float *temp, const float *h, const fsamd_line *lines, int nlines,
                       void *stream, int *best_g, int *best_waves, float *ms7)
{
    if (!P || !best_g || !best_waves) return fail(FSAMD_ERR_ARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    // Large plans are timed on every stride-th sightline of the heaviest-first
    // order (the same mix of heavy and light work items), at least 16384
    // sightlines so that each block still drains several items: with fewer,
    // the last items dominate and the widths time within noise of each other.
    const long long stride = std::max<long long>(1, P->nlos / 16384);
    const int widths[3] = {16, 32, 64};
    hipEvent_t e0, e1;
    FSAMD_HIP(hipEventCreate(&e0));
    FSAMD_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bg = 16, bw = 4;
    int rc = FSAMD_OK;
    // one untimed run makes the allocations; then the three widths at 4 and
    // at 8 wavefronts per block, then 2 wavefronts at the best width
    for (int i = -1; i < 7 && rc == FSAMD_OK; ++i) {
        const int g = i < 0 ? 16 : i < 6 ? widths[i % 3] : bg;
        const int w = i < 3 ? 4 : i < 6 ? 8 : 2;
        FSAMD_HIP(hipEventRecord(e0, s));
        rc = run_deposit(true, P, nbins, kernel, velfac, tautail, pos, vel, temp, h, lines, nlines,
                         FSAMD_MODE_BATCHED, s, stride, g, w);
        if (rc) break;
        FSAMD_HIP(hipEventRecord(e1, s));
        FSAMD_HIP(hipEventSynchronize(e1));
        float ms = 0;
        FSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (i < 0) continue;
        if (ms7) ms7[i] = ms;
        if (ms < best) {
            best = ms;
            bg = g;
            bw = w;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    P->tuned_g = bg;
    P->tuned_waves = bw;
    P->tuned_key = tune_key(nbins, nlines, kernel);
    *best_g = bg;
    *best_waves = bw;
    return FSAMD_OK;
}

int fsamd_plan_keep_windows(fsamd_plan *plan, int enable)
{
    if (!plan) return fail(FSAMD_ERR_ARG, "plan is NULL");
    plan->keep_windows = enable != 0;
    if (!enable) {
        // a graph captured while recording was on still writes this buffer
        if (plan->captured && plan->win.ptr && !plan->win.borrowed)
            plan->retired.push_back(plan->win.disown());
        else
            plan->win.release();
        plan->win_lines = 0;
        ++plan->win_gen;
    }
    return FSAMD_OK;
}

int fsamd_plan_window_generation(const fsamd_plan *plan, long long *gen)
{
    if (!plan || !gen) return fail(FSAMD_ERR_ARG, "NULL argument");
    *gen = plan->win_gen;
    return FSAMD_OK;
}

int fsamd_tau_adjoint(fsamd_plan *P, int nbins, int kernel, double velfac, const float *pos, const float *vel,
                      const float *temp, const float *h, const fsamd_line_grad *lines, int nlines, double *grad_temp,
                      double *grad_vel, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (!P) return fail(FSAMD_ERR_ARG, "plan is NULL");
    if (!lines || nlines < 1 || !grad_temp || !grad_vel) return fail(FSAMD_ERR_ARG, "NULL argument");
    if (kernel == kVoronoiMesh && !P->seg_ready)
        return fail(FSAMD_ERR_ARG, "kernel 2 (Voronoi mesh) needs fsamd_plan_assign_cells on the plan first");
    if (!P->keep_windows || P->win_lines != nlines || P->win_nbins != nbins || P->win_kernel != kernel)
        return fail(FSAMD_ERR_ARG, "fsamd_tau_adjoint needs the windows of a matching fsamd_tau call on this plan "
                                   "(fsamd_plan_keep_windows(plan, 1) before the forward)");
    static thread_local DeviceInfo dev;
    int rc;
    if (dev.ncu == 0 && (rc = device_info(dev))) return rc;
    if (deposit_lds_bytes(nbins) + adjoint_static_lds(g_group_size.load(std::memory_order_relaxed)) >
        (size_t)dev.max_lds)
        return fail(FSAMD_ERR_UNSUPPORTED, "adjoint rows must fit in LDS (nbins too large)");
    std::vector<LineDesc> ld(nlines);
    for (int i = 0; i < nlines; ++i) {
        if (!lines[i].dens || !lines[i].grad_tau || !lines[i].grad_dens)
            return fail(FSAMD_ERR_ARG, "line without dens, grad_tau or grad_dens");
        fsamd_line l{lines[i].lambda_cm, lines[i].gamma, lines[i].fosc, lines[i].amumass, lines[i].dens,
                     const_cast<double *>(lines[i].grad_tau)};
        line_consts(l, ld[i]);
        ld[i].grad_dens = lines[i].grad_dens;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    FSAMD_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone) return fail(FSAMD_ERR_ARG, "fsamd_tau_adjoint cannot be graph-captured");
    const LineDesc *table = nullptr;
    if ((rc = line_table(P, ld, false, s, &table))) return rc;
    const int ngroups = (nlines + kMaxLines - 1) / kMaxLines;
    if (P->queue.bytes < sizeof(int) * (size_t)ngroups && (rc = regrow(P, P->queue, sizeof(int) * ngroups))) return rc;
    const size_t rbytes = sizeof(PartRec) * (size_t)std::max<long long>(P->npart, 1);
    if (P->rec.bytes < rbytes) {
        FSAMD_HIP(hipStreamSynchronize(s));
        if ((rc = regrow(P, P->rec, rbytes))) return rc;
    }
    // the list of pairs for the FP64 adjoint kernel (the forward's fix-up list)
    if (!P->fix_count.ptr && (rc = P->fix_count.alloc(sizeof(unsigned long long) * 2, P->arena))) return rc;
    if (P->fix_cap < 65536 && (rc = grow_fixups(P, 65536, s))) return rc;
    const int gsize = g_group_size.load(std::memory_order_relaxed);
    for (int g = 0; g < ngroups; ++g) {
        const int k0 = g * kMaxLines;
        const int K = std::min(kMaxLines, nlines - k0);
        DepositArgs A;
        std::memset(&A, 0, sizeof(A));
        A.nbins = nbins;
        A.kernel = kernel;
        A.velfac = velfac;
        A.inv_velfac = 1.0 / velfac;
        A.vbox = P->box * velfac;
        A.box = P->box;
        A.force_fp64 = g_fp64;
        A.pos = pos;
        A.vel = vel;
        A.temp = temp;
        A.h = h;
        A.axis = P->axis.as<int>();
        A.cofm = P->cofm.as<double>();
        A.line_off = P->line_off.as<long long>();
        A.pair_part = P->pair_part.as<int>();
        A.seg = kernel == kVoronoiMesh ? P->seg.as<float2>() : nullptr;
        A.nlines = K;
        A.lines = table + k0;
        A.rec = P->rec.as<PartRec>();
        if ((rc = launch_pack_records(pos, vel, temp, h, A.lines, K, P->npart, P->rec.as<PartRec>(), s))) return rc;
        A.queue = P->queue.as<int>() + g;
        A.npairs = P->npairs;
        A.win = P->win.as<int>() + (size_t)k0 * 2 * P->npairs;
        A.grad_temp = grad_temp;
        A.grad_vel = grad_vel;
        fsamd_plan::ItemSet *set = nullptr;
        if ((rc = item_set(P, K, 1ll << 62, &set))) return rc;
        if (set->host.empty()) continue;
        A.nitems = (int)set->host.size();
        A.fix = P->fix.as<Fixup>();
        A.fix_count = P->fix_count.as<unsigned long long>();
        A.fix_cap = P->fix_cap;
        if ((rc = launch_fill_i32(A.queue, 0, 1, s))) return rc;
        if ((rc = launch_fill_i32((int *)A.fix_count, 0, 4, s))) return rc;
        const int nblocks = (int)std::min<long long>((long long)adjoint_blocks_per_cu(gsize, nbins) * dev.ncu, A.nitems);
        if ((rc = launch_tau_adjoint(gsize, A, set->buf.as<WorkItem>(), nblocks, s))) return rc;
        if ((rc = launch_tau_adjoint64(A, set->buf.as<WorkItem>(), dev.ncu * 2, s))) return rc;
        unsigned long long fc[2] = {0, 0};
        FSAMD_HIP(hipMemcpyAsync(fc, A.fix_count, sizeof(fc), hipMemcpyDeviceToHost, s));
        FSAMD_HIP(hipStreamSynchronize(s));
        // the forward listed every FP64 pair of the group in this list, so it
        // holds them unless the FP64 setting changed in between
        if (fc[1])
            return fail(FSAMD_ERR_UNSUPPORTED, "adjoint: more FP64 pairs than the forward's fix-up list holds; "
                                               "run the forward with the same fsamd_set_fp64 setting");
    }
    return FSAMD_OK;
}

int fsamd_set_stats(int flags)
{
    if (flags < 0 || flags > 3) return fail(FSAMD_ERR_ARG, "flags must be 0..3");
    g_stats_on = flags & 1;
    g_timing_on = flags != 0;
    return FSAMD_OK;
}

int fsamd_last_stats(long long *out, int n)
{
    if (!out || n < 0) return fail(FSAMD_ERR_ARG, "NULL argument or negative count");
    for (int i = 0; i < std::min(n, kNStats); ++i) out[i] = g_stats[i];
    return FSAMD_OK;
}

double fsamd_last_ms(void) { return g_last_ms; }
double fsamd_last_kernel_ms(void) { return g_last_kernel_ms; }

int fsamd_set_fp64(int on)
{
    g_fp64 = on != 0;
    return FSAMD_OK;
}

int fsamd_set_group_size(int g)
{
    if (g != 16 && g != 32 && g != 64) return fail(FSAMD_ERR_ARG, "group size must be 16, 32 or 64");
    g_group_size = g;
    return FSAMD_OK;
}

int fsamd_set_block_waves(int w)
{
    if (w != 2 && w != 4 && w != 8) return fail(FSAMD_ERR_ARG, "wavefronts per block must be 2, 4 or 8");
    g_block_waves = w;
    return FSAMD_OK;
}

int fsamd_voigt_eval(const double *u, long long n, double a, float *out, void *stream)
{
    if (n < 0 || (n > 0 && (!u || !out))) return fail(FSAMD_ERR_ARG, "bad argument");
    if (!(a >= 0.0 && a <= (double)kA32Max))
        return fail(FSAMD_ERR_ARG, "the FP32 profile takes 0 <= a <= 0.04 (larger a take the FP64 path)");
    return launch_voigt_eval(u, n, a, out, (hipStream_t)stream);
}

int fsamd_voigt_peak(double umax, double a, double *samples_per_s, void *stream)
{
    if (!samples_per_s || !(umax > 0)) return fail(FSAMD_ERR_ARG, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    DeviceInfo dev;
    int rc = device_info(dev);
    if (rc) return rc;
    DevBuf sink;
    if ((rc = sink.alloc(16))) return rc;
    const int nblocks = dev.ncu * 8;
    const long long per_thread = 4096;
    const double du = 0.001;
    hipEvent_t e0, e1;
    FSAMD_HIP(hipEventCreate(&e0));
    FSAMD_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        FSAMD_HIP(hipEventRecord(e0, s));
        if ((rc = launch_voigt_peak(per_thread, nblocks, du, umax, (float)a, sink.as<float>(), s))) return rc;
        FSAMD_HIP(hipEventRecord(e1, s));
        FSAMD_HIP(hipEventSynchronize(e1));
        float ms = 0;
        FSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *samples_per_s = (double)nblocks * 256 * per_thread / (best * 1e-3);
    return FSAMD_OK;
}

}  // extern "C"
