// C ABI of libfsamd (declared in include/fake_spectra_amd.h).
#include "fsamd_internal.h"
#include "plan.h"
#include "deposit.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace fsamd {

static thread_local std::string g_error;
static thread_local long long g_stats[4] = {0, 0, 0, 0};
static thread_local int g_stats_on = 0;

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

// Work items for lines [k0, k0+K) in the plan's heaviest-first order; a
// sightline with more than `chunk` pairs is split into non-direct items.
static void build_items(const fsamd_plan *P, int K, long long chunk, std::vector<WorkItem> &items,
                        bool &any_split)
{
    items.clear();
    any_split = false;
    for (int s : P->order_host) {
        const long long a = P->off_host[s], b = P->off_host[s + 1];
        const long long n = b - a;
        for (int k = 0; k < K; ++k) {
            if (n <= chunk) {
                items.push_back(WorkItem{s, k, a, b, 1, 0});
            } else {
                any_split = true;
                for (long long c = a; c < b; c += chunk)
                    items.push_back(WorkItem{s, k, c, std::min(b, c + chunk), 0, 0});
            }
        }
    }
}

static int run_deposit(bool tau, const fsamd_plan *P, int nbins, int kernel, double velfac, double tautail,
                       const float *pos, const float *vel, const float *temp, const float *h,
                       const fsamd_line *lines, int nlines, int mode, hipStream_t s)
{
    if (!P) return fail(FSAMD_ERR_ARG, "plan is NULL");
    if (nbins < 1) return fail(FSAMD_ERR_ARG, "nbins must be positive");
    if (kernel != kTopHat && kernel != kCubicSpline)
        return fail(FSAMD_ERR_UNSUPPORTED, "kernel must be 0 (tophat) or 1 (cubic spline)");
    if (nlines < 1 || !lines) return fail(FSAMD_ERR_ARG, "at least one line is required");
    if (mode != FSAMD_MODE_BATCHED && mode != FSAMD_MODE_PLAIN) return fail(FSAMD_ERR_ARG, "unknown mode");
    if ((long long)P->nlos * nbins >= (1ll << 62)) return fail(FSAMD_ERR_ARG, "output too large");
    DeviceInfo dev;
    int rc = device_info(dev);
    if (rc) return rc;
    const bool lds = deposit_lds_bytes(nbins) + 2048 <= (size_t)dev.max_lds;
    for (int i = 0; i < nlines; ++i)
        if (!lines[i].out || !lines[i].dens) return fail(FSAMD_ERR_ARG, "line without dens or out");

    DevBuf queue, stats;
    if ((rc = queue.alloc(sizeof(int) * 4))) return rc;
    if (g_stats_on) {
        if ((rc = stats.alloc(sizeof(long long) * 4))) return rc;
        FSAMD_HIP(hipMemsetAsync(stats.ptr, 0, sizeof(long long) * 4, s));
    }
    for (int k0 = 0; k0 < nlines; k0 += kMaxLines) {
        const int K = std::min(kMaxLines, nlines - k0);
        DepositArgs A;
        std::memset(&A, 0, sizeof(A));
        A.nbins = nbins;
        A.kernel = kernel;
        A.velfac = velfac;
        A.vbox = P->box * velfac;     // absorption.cpp:157
        A.box = P->box;
        A.tautail = tautail;
        A.pos = pos;
        A.vel = vel;
        A.temp = temp;
        A.h = h;
        A.axis = P->axis.as<int>();
        A.cofm = P->cofm.as<double>();
        A.line_off = P->line_off.as<long long>();
        A.pair_part = P->pair_part.as<int>();
        A.nlines = K;
        std::vector<LineDesc> ldesc(K);
        for (int k = 0; k < K; ++k) line_consts(lines[k0 + k], ldesc[k]);
        DevBuf dlines;
        if ((rc = dlines.alloc(sizeof(LineDesc) * K))) return rc;
        FSAMD_HIP(hipMemcpyAsync(dlines.ptr, ldesc.data(), sizeof(LineDesc) * K, hipMemcpyHostToDevice, s));
        A.lines = dlines.as<LineDesc>();
        A.queue = queue.as<int>();
        A.stats = g_stats_on ? stats.as<long long>() : nullptr;

        const bool plain = (mode == FSAMD_MODE_PLAIN);
        std::vector<WorkItem> items;
        bool split = false;
        build_items(P, K, plain ? kPlainChunk : (1ll << 62), items, split);
        if (split || !lds) {
            for (int k = 0; k < K; ++k)
                if ((rc = launch_fill_f64(ldesc[k].out, 0.0, (long long)P->nlos * nbins, s))) return rc;
        }
        if (items.empty()) continue;
        DevBuf ditems;
        if ((rc = ditems.alloc(sizeof(WorkItem) * items.size()))) return rc;
        FSAMD_HIP(hipMemcpyAsync(ditems.ptr, items.data(), sizeof(WorkItem) * items.size(),
                                 hipMemcpyHostToDevice, s));
        if (!plain) {
            A.nitems = (int)items.size();
            if ((rc = launch_fill_i32(A.queue, 0, 1, s))) return rc;
            const int per_cu = deposit_blocks_per_cu(tau, lds, nbins);
            const int nblocks = (int)std::min<long long>((long long)per_cu * dev.ncu, A.nitems);
            if ((rc = launch_deposit(tau, lds, A, ditems.as<WorkItem>(), 0, nblocks, 1, s))) return rc;
        } else {
            // one launch per (sightline, line), in sightline index order
            std::vector<int> first(P->nlos * (size_t)K + 1, -1);
            size_t i = 0;
            while (i < items.size()) {
                size_t j = i;
                while (j < items.size() && items[j].los == items[i].los && items[j].line == items[i].line) ++j;
                A.nitems = (int)(j - i);
                if ((rc = launch_deposit(tau, lds, A, ditems.as<WorkItem>(), (int)i, A.nitems, 0, s))) return rc;
                i = j;
            }
        }
        FSAMD_HIP(hipStreamSynchronize(s));   // items / queue buffers are freed on return
    }
    if (g_stats_on) {
        FSAMD_HIP(hipMemcpyAsync(g_stats, stats.ptr, sizeof(long long) * 4, hipMemcpyDeviceToHost, s));
        FSAMD_HIP(hipStreamSynchronize(s));
    }
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
    if (npart < 0 || nlos < 0) return fail(FSAMD_ERR_ARG, "negative size");
    if (npart >= (1ll << 31)) return fail(FSAMD_ERR_ARG, "particle indices must fit in int (N < 2^31)");
    if (!(box > 0) || !std::isfinite(box)) return fail(FSAMD_ERR_ARG, "box must be positive and finite");
    if ((npart > 0 && (!pos || !h)) || (nlos > 0 && (!axis || !cofm)))
        return fail(FSAMD_ERR_ARG, "NULL array");
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

int fsamd_plan_near_lines(const fsamd_plan *plan, int *out, long long *n_out, void *stream)
{
    if (!plan || !n_out) return fail(FSAMD_ERR_ARG, "NULL argument");
    return plan_near_lines(plan, out, n_out, (hipStream_t)stream);
}

int fsamd_tau(const fsamd_plan *plan, int nbins, int kernel, double velfac, double tautail, const float *pos,
              const float *vel, const float *temp, const float *h, const fsamd_line *lines, int nlines,
              int mode, void *stream)
{
    return run_deposit(true, plan, nbins, kernel, velfac, tautail, pos, vel, temp, h, lines, nlines, mode,
                       (hipStream_t)stream);
}

int fsamd_colden(const fsamd_plan *plan, int nbins, int kernel, double velfac, const float *pos, const float *h,
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

int fsamd_set_stats(int enable)
{
    g_stats_on = enable ? 1 : 0;
    return FSAMD_OK;
}

int fsamd_last_stats(long long *out4)
{
    if (!out4) return fail(FSAMD_ERR_ARG, "NULL argument");
    for (int i = 0; i < 4; ++i) out4[i] = g_stats[i];
    return FSAMD_OK;
}

}  // extern "C"
