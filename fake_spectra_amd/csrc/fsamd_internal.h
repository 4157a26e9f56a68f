// Internal declarations shared by the HIP translation units of libfsamd.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "fake_spectra_amd.h"

namespace fsamd {

// Physical constants (fake_spectra/absorption.cpp:21-33, absorption.h:4).
constexpr double kSigmaT = 6.652458558e-25;
constexpr double kBoltzmann = 1.3806504e-16;
constexpr double kLight = 2.99792458e10;
constexpr double kProtonMass = 1.67262178e-24;
constexpr double kPi = 3.14159265358979323846;

constexpr int kTopHat = 0;       // singleabs.h:9
constexpr int kCubicSpline = 1;  // singleabs.h:10
constexpr int kVoronoiMesh = 2;  // singleabs.h:11
constexpr int kSphQuintic = 3;   // singleabs.h:12
constexpr int kNGrid = 8;        // singleabs.h:8

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define FSAMD_HIP(call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::fsamd::fail(FSAMD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ----------------------------------------------------------------------------
// Sightline grid: sightlines bucketed by their transverse coordinates, one
// square grid per axis.  The plane coordinates of a sightline are the two
// quantities the reference's IndexTable compares against (index_table.cpp:7-50):
//   axis 1: key = cofm.y, second = cofm.z ; axis 2: key = cofm.x, second = cofm.z
//   axis 3: key = cofm.x, second = cofm.y

struct LosGrid {
    int nc[3];             // cells per side
    int base[3];           // first cell id of each axis grid
    int nlos_axis[3];
    double cw[3];          // cell width
    double inv_cw[3];
    const int *cell_start;     // [ncells + 1]
    const int *cell_ids;       // sightline ids in cell order
    const double2 *cell_kl;    // (key, second) of each entry
};

struct BinArgs {
    double box;
    const float *pos;      // [N,3]
    const float *h;        // [N]
    long long npart;
    LosGrid grid;
};

// Deposit arguments shared by the tau and colden kernels.
struct LineDesc {
    double sigma_a;    // absorption.cpp:154
    double bfac;       // absorption.cpp:155
    double voigt_fac;  // absorption.cpp:156
    const float *dens; // [N] species density
    double *out;       // [nlos, nbins]; in the adjoint: dL/dtau (read only)
    double *grad_dens; // adjoint: [N] +=
};

constexpr int kMaxLines = 8;   // lines per deposit launch
constexpr int kWaveClaim = 32; // most pairs a wavefront of the tau kernel sets up at once

// One particle's inputs to the optical-depth kernels in one 64-byte record
// (a single cache line per pair instead of one per input array): position,
// support radius, velocity, temperature and the density of each line of the
// launch.  Packed once per launch from the caller's arrays.
struct PartRec {
    float x, y, z, h;
    float vx, vy, vz, temp;
    float dens[kMaxLines];
};
static_assert(sizeof(PartRec) == 64, "PartRec is one 64-byte record");

// A walk handed from the FP32 deposit kernel to the FP64-capable fix-up
// kernel: pixels z0, z0 + dir, ... (count of them) of pair k in work item
// `item`; dir == 0 means the whole pair (both directions) in FP64.
struct Fixup {
    long long k;
    int item;
    int z0;
    int count;
    int dir;
};
constexpr int kNStats = FSAMD_NSTATS;

struct DepositArgs {
    int nbins;
    int kernel;
    double velfac;
    double inv_velfac;  // 1 / velfac
    double vbox;       // box * velfac
    double box;
    double tautail;
    const PartRec *rec; // [N] packed inputs (tau and adjoint kernels)
    const float *pos;   // [N,3]
    const float *vel;   // [N,3]
    const float *temp;  // [N]
    const float *h;     // [N]
    const int *axis;    // [L]
    const double *cofm; // [L,3]
    const long long *line_off;  // [L+1]
    const int *pair_part;       // [M]
    const float2 *seg;          // [M] Voronoi segment ends (kernel 2), else nullptr
    const LineDesc *lines;      // device [nlines], lines of this launch
    int nlines;
    int nitems;
    int claim;                  // most pairs a wavefront of the tau kernel claims at once (1..kWaveClaim)
    int *queue;                 // work-queue counter (zeroed before launch)
    int *win;                   // optional [nlines][npairs][2] pixels walked per direction
    long long npairs;
    // adjoint only
    double *grad_temp;          // [N] +=
    double *grad_vel;           // [N,3] +=
    Fixup *fix;                 // fix-up list
    unsigned long long *fix_count;   // [0] entries appended (64-bit: cannot wrap), [1] overflow flag
    int fix_cap;
    int force_fp64;             // every pair on the FP64 path (fsamd_set_fp64)
    long long *stats;           // optional counters, kNStats entries (see fsamd_last_stats)
};

// Host-side launchers (return FSAMD_OK or an error code).
// Raise kernel fn's dynamic-LDS limit to shm bytes if it is lower.  The limit
// only grows, under a lock, so a launch made after its own call here stays
// valid whatever sizes other host threads ask for meanwhile.
int raise_lds_limit(const void *fn, size_t shm);
int launch_fill_i32(int *p, int value, long long n, hipStream_t s);
int launch_fill_f64(double *p, double value, long long n, hipStream_t s);
int launch_pack_records(const float *pos, const float *vel, const float *temp, const float *h,
                        const LineDesc *lines, int nlines, long long n, PartRec *out, hipStream_t s);
int exclusive_scan_i32_to_i64(const int *in, long long *out, long long n, void *ws, size_t ws_bytes,
                              hipStream_t s);
size_t exclusive_scan_ws_bytes(long long n);
int exclusive_scan_i64_inplace(long long *data, long long n, void *ws, size_t ws_bytes, hipStream_t s);

}  // namespace fsamd
