// The binning plan: device neighbour lists and the buffers that own them.
#pragma once
#include "fsamd_internal.h"
#include "deposit.h"

#include <memory>
#include <vector>

namespace fsamd {

// Owning device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    int alloc(size_t n);
    void release();
    void take(DevBuf &o)
    {
        release();
        ptr = o.ptr;
        bytes = o.bytes;
        o.ptr = nullptr;
        o.bytes = 0;
    }
    template <typename T>
    T *as() const { return (T *)ptr; }
};

int plan_build(fsamd_plan *P, double box, const float *pos, const float *h, long long npart,
               const int *axis, const double *cofm, int nlos, hipStream_t s);
int plan_near_lines(const fsamd_plan *P, int *out, long long *n_out, hipStream_t s);

}  // namespace fsamd

struct fsamd_plan {
    double box = 0;
    long long npart = 0;
    int nlos = 0;
    long long npairs = 0;
    long long max_row = 0;
    int ncells = 0;
    fsamd::LosGrid grid{};
    fsamd::DevBuf axis, cofm;             // copies of the sightline table
    fsamd::DevBuf cell_start, cell_ids, cell_kl;
    fsamd::DevBuf poff;                   // [npart+1] particle-major pair offsets
    fsamd::DevBuf line_off;               // [nlos+1]
    fsamd::DevBuf pair_part;              // [npairs]
    fsamd::DevBuf order;                  // sightlines by descending pair count
    std::vector<int> order_host;
    std::vector<long long> off_host;
    // Deposit launch state, reused across calls so that a call allocates,
    // copies and synchronizes nothing once warm (and can be graph-captured).
    struct ItemSet {
        int K;
        long long chunk;
        bool split;
        fsamd::DevBuf buf;
        std::vector<fsamd::WorkItem> host;
    };
    std::vector<std::unique_ptr<ItemSet>> item_sets;
    fsamd::DevBuf queue;        // work-queue counters, one per launch group
    fsamd::DevBuf lines_dev;    // LineDesc table
    std::vector<fsamd::LineDesc> lines_host;
    fsamd::DevBuf stats;        // kNStats counters
    fsamd::DevBuf fix;          // fix-up list (FP64 walks)
    fsamd::DevBuf fix_count;    // [count, overflow] per line group
    int fix_cap = 0;
    // pixels walked per (line, pair, direction) by the last fsamd_tau call,
    // kept for the adjoint when keep_windows is set
    bool keep_windows = false;
    fsamd::DevBuf win;
    int win_lines = 0, win_nbins = 0, win_kernel = -1;
    // lane-group width chosen by fsamd_tau_autotune for (nbins, nlines, kernel)
    int tuned_g = 0;
    long long tuned_key = -1;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~fsamd_plan()
    {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};
