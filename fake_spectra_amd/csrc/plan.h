// The binning plan: device neighbour lists and the buffers that own them.
#pragma once
#include "fsamd_internal.h"
#include "deposit.h"

#include <deque>
#include <memory>
#include <vector>

namespace fsamd {

// A bump allocator over device memory the caller owns (the allocation-free
// entry's scratch array): buffers carved from it are never freed one by one.
struct Arena {
    char *base = nullptr;
    size_t cap = 0;
    size_t used = 0;
};

// Device allocation: owned (from the device pool or hipMalloc, see
// DevBuf::alloc), or borrowed from an Arena.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    int dev = -1;          // device of an owned allocation
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    // at least n bytes: from the arena when one is given, else a block of
    // the calling device's pool, else hipMalloc
    int alloc(size_t n, Arena *arena = nullptr);
    // hipFree (which waits for the device)
    void release();
    // back to the device's pool, for a later alloc; only where no queued work
    // can still use the buffer (after a synchronization)
    void recycle();
    void take(DevBuf &o)
    {
        release();
        ptr = o.ptr;
        bytes = o.bytes;
        borrowed = o.borrowed;
        dev = o.dev;
        o.ptr = nullptr;
        o.bytes = 0;
        o.borrowed = false;
        o.dev = -1;
    }
    template <typename T>
    T *as() const { return (T *)ptr; }
    // Hand an owned allocation over to the caller (who frees it), leaving
    // this empty; borrowed memory is not handed over (nullptr).
    void *disown()
    {
        void *p = borrowed ? nullptr : ptr;
        ptr = nullptr;
        bytes = 0;
        borrowed = false;
        dev = -1;
        return p;
    }
};

int plan_build(fsamd_plan *P, double box, const float *pos, const float *h, long long npart,
               const int *axis, const double *cofm, int nlos, hipStream_t s);
int plan_near_lines(const fsamd_plan *P, int *out, long long *n_out, hipStream_t s);
// Voronoi mesh: the segment ends of every pair (IndexTable::assign_cells).
int plan_assign_cells(fsamd_plan *P, const float *pos, hipStream_t s);
int near_lines_direct(double box, const float *pos, const float *h, long long npart, const int *axis,
                      const double *cofm, int nlos, int *out, long long *n_out, hipStream_t s);

}  // namespace fsamd

struct fsamd_plan {
    // Where the plan's device buffers come from: hipMalloc (nullptr), or an
    // arena (the allocation-free entry, whose plans never call hipMalloc).
    fsamd::Arena *arena = nullptr;
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
    fsamd::DevBuf seg;                    // [npairs] float2 Voronoi segment ends (kernel 2)
    bool seg_ready = false;
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
    // LineDesc tables.  Eager calls share one table, rewritten when the lines
    // change; a captured call pins the table it reads, which is then never
    // rewritten or freed before the plan.
    struct LineTable {
        fsamd::DevBuf buf;
        std::vector<fsamd::LineDesc> host;
    };
    LineTable eager_lines;
    std::vector<std::unique_ptr<LineTable>> pinned_lines;
    fsamd::DevBuf stats;        // kNStats counters
    fsamd::DevBuf rec;          // [npart] packed particle records of the current launch
    fsamd::DevBuf fix;          // fix-up list (FP64 walks)
    fsamd::DevBuf fix_count;    // [count, overflow] (uint64 each)
    int fix_cap = 0;
    // pixels walked per (line, pair, direction) by the last fsamd_tau call,
    // kept for the adjoint when keep_windows is set
    bool keep_windows = false;
    fsamd::DevBuf win;
    int win_lines = 0, win_nbins = 0, win_kernel = -1;
    long long win_gen = 0;      // bumped whenever the recorded windows change
    // lane-group width chosen by fsamd_tau_autotune for (nbins, nlines, kernel)
    int tuned_g = 0;
    int tuned_waves = 4;
    long long tuned_key = -1;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // call, deposit-kernel brackets
    // Once a call on this plan has been captured into a graph, buffers that
    // graph's kernels address are retired (kept until the plan is destroyed)
    // instead of freed when a later call needs them larger.
    bool captured = false;
    std::vector<void *> retired;
    ~fsamd_plan()
    {
        // the buffers go back to their device's pool once no queued work can
        // use them: wait for that device (which need not be the current one)
        int cur = -1;
        (void)hipGetDevice(&cur);
        const int d = axis.dev >= 0 ? axis.dev : cur;
        if (d != cur) (void)hipSetDevice(d);
        (void)hipDeviceSynchronize();
        for (fsamd::DevBuf *b : {&axis, &cofm, &cell_start, &cell_ids, &cell_kl, &poff, &line_off, &pair_part,
                                 &order, &seg, &queue, &eager_lines.buf, &stats, &rec, &fix, &fix_count, &win})
            b->recycle();
        for (auto &it : item_sets) it->buf.recycle();
        for (auto &t : pinned_lines) t->buf.recycle();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (void *p : retired)
            if (p) (void)hipFree(p);
        if (d != cur) (void)hipSetDevice(cur);
    }
};
