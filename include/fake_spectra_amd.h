/* fake_spectra_amd: C ABI of the MI355X (gfx950) sightline spectral-synthesis
 * library.  Every pointer argument documented "device" is a HIP device
 * pointer; "stream" is a hipStream_t passed as void* (0 = null stream).
 * All functions return FSAMD_OK (0) or an error code; fsamd_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Threads: calls on different plans may run at the same time from different
 * host threads.  Calls on one plan must not overlap: a plan holds the scratch
 * of its calls (work items, queue counters, the fix-up list, packed particle
 * records), so give each thread its own plan or serialise.  The tuning knobs
 * (fsamd_set_group_size, fsamd_set_block_waves) are process-wide.
 *
 * The entry points replace the reference's native path behind
 * Spectra.get_tau / Spectra.get_col_density:
 *   fsamd_plan_create        IndexTable ctor + get_near_particles
 *                            (fake_spectra/index_table.cpp:7-18, 130-150)
 *   fsamd_near_lines,        _near_lines (fake_spectra/py_module.cpp:25-99)
 *   fsamd_plan_near_lines
 *   fsamd_tau                ParticleInterp::compute_tau (part_int.cpp:20-51)
 *                            for several lines at once
 *   fsamd_colden             ParticleInterp::compute_colden (part_int.cpp:53-84)
 *   fsamd_particle_interpolate  _Particle_Interpolate (py_module.cpp:103-233)
 */
#ifndef FAKE_SPECTRA_AMD_H
#define FAKE_SPECTRA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FSAMD_OK 0
#define FSAMD_ERR_ARG 1          /* invalid argument (ValueError in Python) */
#define FSAMD_ERR_OOM 2          /* device allocation failed (MemoryError) */
#define FSAMD_ERR_HIP 3          /* HIP runtime error */
#define FSAMD_ERR_UNSUPPORTED 4  /* unsupported kernel id or size */

#define FSAMD_MODE_BATCHED 0     /* all sightlines x lines in one persistent launch */
#define FSAMD_MODE_PLAIN 1       /* one launch per sightline and line */

const char *fsamd_last_error(void);
int fsamd_version(void);

/* Sightline <-> particle binning: the on-device counterpart of IndexTable.
 * Holds, for every sightline, the indices of particles whose kernel support
 * reaches it (index_table.cpp:22-50 acceptance test, applied exactly),
 * ascending within each sightline. */
typedef struct fsamd_plan fsamd_plan;

/* Builds the neighbour lists.  pos [npart,3] f32, h [npart] f32, axis [nlos]
 * i32 with values 1..3, cofm [nlos,3] f64: all device.  Synchronizes the
 * stream once to size the pair list.  Replaces IndexTable::IndexTable and
 * IndexTable::get_near_particles (index_table.cpp:7-18, 130-150). */
int fsamd_plan_create(double box, const float *pos, const float *h, long long npart,
                      const int *axis, const double *cofm, int nlos, void *stream,
                      fsamd_plan **out);
int fsamd_plan_destroy(fsamd_plan *plan);
long long fsamd_plan_npairs(const fsamd_plan *plan);
/* Device pointers to the lists: offsets [nlos+1] (int64), particle ids [npairs]. */
int fsamd_plan_pairs(const fsamd_plan *plan, const long long **offsets, const int **parts);
/* Sorted unique ids of particles with at least one nearby sightline, written
 * to out (device, capacity npart); *n_out receives the count (host).
 * Replaces Py_near_lines (py_module.cpp:25-99). */
int fsamd_plan_near_lines(const fsamd_plan *plan, int *out, long long *n_out, void *stream);
/* The same list without building neighbour lists (what Py_near_lines,
 * py_module.cpp:25-99, computes): the sightline grid, then one acceptance
 * test per particle that stops at the first sightline it reaches, and a
 * compaction.  pos, h, axis, cofm: device, as in fsamd_plan_create; out:
 * device, capacity npart; *n_out: host.  Synchronizes the stream. */
int fsamd_near_lines(double box, const float *pos, const float *h, long long npart, const int *axis,
                     const double *cofm, int nlos, int *out, long long *n_out, void *stream);

/* Voronoi mesh (kernel 2): cut every sightline of the plan into segments,
 * one per particle of its neighbour list, at the points 0.1 apart along the
 * line that each particle is the nearest to (IndexTable::assign_cells,
 * index_table.cpp:152-223, with its periodic wrap).  Needed before
 * fsamd_tau / fsamd_colden / fsamd_tau_adjoint with kernel 2, whose pairs then
 * take the segment ends in place of dr2 and h (part_int.cpp:28-46).  pos:
 * device [npart,3], the plan's particles.  Returns FSAMD_ERR_ARG where the
 * reference exits (a particle owning two separate runs of one sightline).
 * Synchronizes the stream. */
int fsamd_plan_assign_cells(fsamd_plan *plan, const float *pos, void *stream);

/* One absorption line of a multi-line call: the (lambda, gamma, fosc,
 * amumass) quadruple of _Particle_Interpolate with its species density. */
typedef struct fsamd_line {
    double lambda_cm;   /* rest wavelength [cm] */
    double gamma;       /* damping constant [1/s] */
    double fosc;        /* oscillator strength */
    double amumass;     /* species mass [amu] */
    const float *dens;  /* device [npart] species density */
    double *out;        /* device [nlos, nbins] float64; fully overwritten */
} fsamd_line;

/* Optical depth of nlines lines (ParticleInterp::compute_tau,
 * part_int.cpp:20-51, with LineAbsorption::add_tau_particle,
 * absorption.cpp:212-279).  kernel: 0 tophat, 1 cubic spline, 2 Voronoi mesh
 * (after fsamd_plan_assign_cells), 3 quintic spline (singleabs.h:9-12); any
 * other id leaves the kernel samples unweighted, as the reference does.
 * mode: FSAMD_MODE_BATCHED or FSAMD_MODE_PLAIN. */
int fsamd_tau(fsamd_plan *plan, int nbins, int kernel, double velfac, double tautail,
              const float *pos, const float *vel, const float *temp, const float *h,
              const fsamd_line *lines, int nlines, int mode, void *stream);

/* Gradients of a scalar loss L through fsamd_tau (no counterpart in the
 * reference, which has no gradients).  Before the forward call, enable
 * fsamd_plan_keep_windows(plan, 1): the forward then records how many pixels
 * each (line, pair, direction) deposited, and the adjoint visits exactly
 * those pixels, holding the discrete choices (window ends, pixel index,
 * sub-sample count) fixed.  For each line, grad_tau is dL/dtau (device
 * [nlos, nbins] float64) and grad_dens receives dL/d dens (device [npart]
 * float64, accumulated with +=); grad_temp [npart] and grad_vel [npart, 3]
 * (float64, +=) receive dL/dT and dL/d vel summed over the lines.  The
 * lines, nbins and kernel must match the recorded forward call. */
typedef struct fsamd_line_grad {
    double lambda_cm, gamma, fosc, amumass;
    const float *dens;        /* device [npart] */
    const double *grad_tau;   /* device [nlos, nbins] */
    double *grad_dens;        /* device [npart], += */
} fsamd_line_grad;

int fsamd_plan_keep_windows(fsamd_plan *plan, int enable);
/* Counter bumped each time the plan's recorded windows change (a recording
 * fsamd_tau call, or recording switched off).  A caller that runs the adjoint
 * later (torch.autograd) saves it after the forward and checks it before the
 * adjoint, so that windows overwritten by another call in between are caught
 * instead of silently used.  No counterpart in the reference. */
int fsamd_plan_window_generation(const fsamd_plan *plan, long long *gen);
int fsamd_tau_adjoint(fsamd_plan *plan, int nbins, int kernel, double velfac, const float *pos,
                      const float *vel, const float *temp, const float *h,
                      const fsamd_line_grad *lines, int nlines, double *grad_temp,
                      double *grad_vel, void *stream);

/* Column density (ParticleInterp::compute_colden, part_int.cpp:53-84, with
 * LineAbsorption::add_colden_particle, absorption.cpp:167-210).  Only dens
 * and out of each line are used. */
int fsamd_colden(fsamd_plan *plan, int nbins, int kernel, double velfac, const float *pos,
                 const float *h, const fsamd_line *lines, int nlines, void *stream);

/* Drop-in of _Particle_Interpolate (py_module.cpp:103-233) on device
 * pointers: out is device [nlos, nbins] float64.  vel and temp are unused
 * when compute_tau == 0.  atime is accepted and unused, as in the
 * reference (absorption.cpp:157). */
int fsamd_particle_interpolate(int compute_tau, int nbins, int kernel, double box100,
                               double velfac, double atime, double lambda, double gamma,
                               double fosc, double amumass, double tautail, const float *pos,
                               const float *vel, const float *dens, const float *temp,
                               const float *h, long long npart, const int *axis,
                               const double *cofm, int nlos, double *out, void *stream);

/* fsamd_particle_interpolate without any device allocation.  No counterpart
 * in the reference: it serves callers that cannot call hipMalloc (serving
 * loops, harnesses that forbid allocation).  Same arguments, results and
 * status codes as fsamd_particle_interpolate.  Every device buffer of the
 * call (sightline grid, neighbour lists and their sort and scan space,
 * packed particle records, work items, the fix-up list) is carved from one
 * FSAMD_NOALLOC_SCRATCH_BYTES array in the library's device code, one copy
 * per GPU, held until the process ends; a call that needs more returns
 * FSAMD_ERR_UNSUPPORTED before touching it (about 100 bytes per particle,
 * 16 per sightline-particle pair and 100 per sightline: e.g. 32^3 particles
 * x 64 sightlines use ~4 MiB).  Calls on one GPU share the array: the
 * library serialises them with a lock per device, and each call
 * synchronises its stream before returning, so a caller must not launch
 * work that still uses the array's contents; calls that overlapped on one
 * GPU would race.  The counters and timing of fsamd_set_stats are not
 * collected.  The output is undefined when a call fails. */
#define FSAMD_NOALLOC_SCRATCH_BYTES (64ull << 20)
int fsamd_particle_interpolate_noalloc(int compute_tau, int nbins, int kernel, double box100,
                                       double velfac, double atime, double lambda, double gamma,
                                       double fosc, double amumass, double tautail, const float *pos,
                                       const float *vel, const float *dens, const float *temp,
                                       const float *h, long long npart, const int *axis,
                                       const double *cofm, int nlos, double *out, void *stream);

/* Counters of the last fsamd_tau call on this thread (host), FSAMD_NSTATS
 * entries: pairs deposited, pixels deposited, pixels re-evaluated in FP64 at
 * the tail cut, pairs evaluated entirely in FP64, Voigt samples of the
 * deposited pixels (7 x sub-samples, as the reference counts them), pixels
 * evaluated (deposited or not), 7-sample kernel sums evaluated (one per
 * pixel sub-sample), pixels evaluated by the far-wing expansion, then the
 * deposit kernel's wavefront cycles (s_memtime) in pair set-up, in far-wing
 * set-up and in walk steps, its wavefront loop iterations and the iterations
 * that set up a pair; fsamd_last_stats copies the first min(n, FSAMD_NSTATS).  And the GPU time of its launches in ms (hipEvents on
 * the call's stream).  fsamd_set_stats(flags): bit 0 collects the counters
 * and the time, bit 1 the time only (the kernels then run without their
 * counting code); either makes fsamd_tau synchronize the stream. */
int fsamd_set_stats(int flags);
#define FSAMD_NSTATS 13
int fsamd_last_stats(long long *out, int n);
double fsamd_last_ms(void);
/* GPU time of the deposit kernel launches alone (the tau or colden kernel,
 * without the fills and the FP64 fix-up pass) of the last timed call, ms. */
double fsamd_last_kernel_ms(void);

/* Autotuner (no counterpart in the reference): times the batched tau launch
 * on every k-th sightline of the plan's heaviest-first order (k = nlos /
 * 16384, all sightlines below 32768) for lane-group widths 16, 32 and 64 at 4
 * and at 8 wavefronts per sightline work item, then 2 wavefronts at the best
 * width, and keeps the fastest pair for later fsamd_tau calls with the same
 * nbins, line count and kernel.  Writes partial results into the lines'
 * outputs.  ms7 (optional, host) receives the seven times in that order. */
int fsamd_tau_autotune(fsamd_plan *plan, int nbins, int kernel, double velfac, double tautail,
                       const float *pos, const float *vel, const float *temp, const float *h,
                       const fsamd_line *lines, int nlines, void *stream, int *best_g, int *best_waves,
                       float *ms7);

/* fsamd_tau inside a hipGraph capture (the stream reports capturing):
 * allowed in the batched mode without statistics after one uncaptured call
 * with the same plan, lines and sizes; the captured sequence allocates,
 * copies and synchronizes nothing.  A fix-up list overflow cannot be handled
 * inside a replay: it sets a sticky flag that fsamd_plan_overflow reads
 * (synchronizing the stream); if set, rerun the call uncaptured. */
int fsamd_plan_overflow(fsamd_plan *plan, int *flag, void *stream);

/* Test and diagnostic switch with no counterpart in the reference: with
 * on != 0, fsamd_tau and fsamd_tau_adjoint calls made afterwards on this host
 * thread evaluate every pair with the FP64 Faddeeva function and FP64 kernel
 * weights (the path the FP32 kernel otherwise keeps for pairs with a damping
 * parameter above 0.04), e.g. for finite-difference gradients free of FP32
 * rounding.  Much slower; default off. */
int fsamd_set_fp64(int on);

/* Memory management with no counterpart in the reference: the device
 * buffers of destroyed plans (and the binning's temporaries) stay in a pool
 * of the device they were allocated on, up to 64 GiB per device, and later
 * plans reuse them.  This frees the calling device's pooled buffers (after
 * waiting for the device); the pool also empties itself when an allocation
 * fails. */
int fsamd_pool_release(void);

/* Tuning knob with no counterpart in the reference: pixels walked per lane
 * group of the tau kernel (16, 32 or 64; a 64-lane wavefront walks 64/g
 * (pair, direction) streams at once).  Process-wide; default 16. */
int fsamd_set_group_size(int g);

/* Tuning knob with no counterpart in the reference: wavefronts per block of
 * the tau kernel, i.e. per (sightline, line) work item (2, 4 or 8; the
 * counting variant runs with 4).  Process-wide; default 8, the fastest
 * untuned choice at 256^3-512^3 particles. */
int fsamd_set_block_waves(int w);

/* H(a, u) = Re w(u + ia) at n points u (device pointer, FP64) by the deposit
 * kernel's FP32 profile, into out (device, FP32); 0 <= a <= 0.04.  No
 * counterpart in the reference's interface: it exposes the evaluator that
 * replaces profile() (singleabs.h:56-61) for accuracy checks. */
int fsamd_voigt_eval(const double *u, long long n, double a, float *out, void *stream);
/* Diagnostic with no counterpart in the reference: isolated throughput of the
 * FP32 sample evaluator (kernel weight x Voigt profile) in samples/s, over
 * arguments u swept uniformly in [0, umax) with damping parameter a, no
 * memory traffic.  The denominator of the deposit's VALU fraction. */
int fsamd_voigt_peak(double umax, double a, double *samples_per_s, void *stream);

/* Flux statistics on device optical depths (FP64).
 * fsamd_mean_flux_scale: the scale s with <exp(-s tau)> = mean_flux_desired
 *   over the pixels with tau <= thresh, Newton-Raphson from s = 1 until
 *   |s' - s| <= tol s' (_rescale_mean_flux, py_module.cpp:235-282); 0 for
 *   n = 0 (fluxstatistics.py:38-39).  tau: device [n].
 * fsamd_res_corr: Gaussian smoothing along each row with periodic wrap,
 *   standard deviation sigma pixels, weights exp(-k^2/2sigma^2) normalised over
 *   |k| <= int(4 sigma + 0.5) (spec_utils.py:5-25, which calls
 *   scipy.ndimage.gaussian_filter1d(mode='wrap')).  in, out: device [nrow, ncol].
 * fsamd_flux_power: mean over sightlines of vmax |rfft(exp(-scale tau)/mean_flux - 1)|^2 / npix^2
 *   at k = 0..npix/2 (fluxstatistics.py:49-56, 71-108 before the window
 *   correction).  tau: device [nlos, npix]; power: device [npix/2 + 1].
 * fsamd_flux_hist: counts of exp(-scale tau) in nbins equal bins of [0, 1]
 *   placed as numpy.histogram places them (fluxstatistics.py:41-49).
 *   counts: device [nbins]. */
int fsamd_mean_flux_scale(const double *tau, long long n, double mean_flux_desired, double tol, double thresh,
                          double *scale, void *stream);
int fsamd_res_corr(const double *in, double *out, int nrow, int ncol, double sigma, void *stream);
int fsamd_flux_power(const double *tau, int nlos, int npix, double scale, double mean_flux, double vmax,
                     double *power, void *stream);
int fsamd_flux_hist(const double *tau, long long n, double scale, int nbins, unsigned long long *counts,
                    void *stream);

#ifdef __cplusplus
}
#endif
#endif
