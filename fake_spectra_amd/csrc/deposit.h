// Work items and launchers of the deposit kernels.
#pragma once
#include "fsamd_internal.h"

namespace fsamd {

// Pairs [p0, p1) of sightline `los` for line `line`.  `direct` items own the
// whole sightline and store the row; others add theirs with FP64 atomics
// into a zero-filled output.
struct WorkItem {
    int los;
    int line;
    long long p0;
    long long p1;
    int direct;
    int pad;
};

size_t deposit_lds_bytes(int nbins);
// LDS a deposit kernel declares itself (queues, per-group state, tables):
// the tau kernel at group width gsize and `waves` wavefronts per block, or
// the column-density kernel; the adjoint kernel at group width gsize.
size_t deposit_static_lds(bool tau, int gsize, int waves);
size_t adjoint_static_lds(int gsize);
int launch_deposit(bool tau, bool lds, int gsize, int waves, const DepositArgs &A, const WorkItem *items, int item_base,
                   int nblocks, int use_queue, hipStream_t s);
int deposit_blocks_per_cu(bool tau, bool lds, int gsize, int waves, int nbins);
int effective_waves(bool lds, bool stats, int waves);
int launch_tau_adjoint(int gsize, const DepositArgs &A, const WorkItem *items, int nblocks, hipStream_t s);
// the FP64 pairs tau_adjoint_kernel listed in A.fix (after it, on the same stream)
int launch_tau_adjoint64(const DepositArgs &A, const WorkItem *items, int nblocks, hipStream_t s);
int adjoint_blocks_per_cu(int gsize, int nbins);
int launch_tau_fixup(const DepositArgs &A, const WorkItem *items, int nblocks, hipStream_t s);
int launch_voigt_eval(const double *u, long long n, double a, float *out, hipStream_t s);
int launch_voigt_peak(long long per_thread, int nblocks, double du, double umax, float a, float *sink,
                      hipStream_t s);

}  // namespace fsamd
