/* Test infrastructure only (the oracle).  FP64 CPU restatement of the
 * reference's sightline spectral-synthesis path.  Never linked into the
 * product; used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker. */
#ifndef FSORACLE_H
#define FSORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

void fso_set_threads(int n);
int fso_max_threads(void);

/* Faddeeva w(z) = exp(-z^2) erfc(-iz) for z = x + iy, y >= 0. */
void fso_w(double x, double y, double *wr, double *wi);
double fso_profile(double u, double a);
double fso_sph_cubic_kernel(double q);
double fso_kern_frac(int kernel, double zlow, double zhigh, double smooth, double dr2, double zrange);
double fso_sph_quintic_kernel(double q);
double fso_tau_kern_outer(double btherm, double vdr2, double vsmooth, double aa, int kernel,
                          double vlow, double vhigh);
void fso_add_tau_particle(double lambda, double gamma, double fosc, double amumass, double velfac,
                          double box, double atime, int kernel, double tautail, double *tau,
                          int nbins, double dr2, float dens, float ppos, float pvel, float temp,
                          float smooth);
void fso_add_colden_particle(double lambda, double gamma, double fosc, double amumass,
                             double velfac, double box, double atime, int kernel, double tautail,
                             double *colden, int nbins, double dr2, float dens, float ppos,
                             float smooth);
int fso_get_near_lines(double box, const int *axis, const double *cofm, int nlos,
                       const float *pos3, float hh, int *lines, double *dr2, int cap);
long long fso_near_particles(double box, const float *pos, const float *h, long long npart,
                             const int *axis, const double *cofm, int nlos, long long *offsets,
                             int *parts, double *dr2, long long cap);
long long fso_near_lines(double box, const float *pos, const float *h, long long npart,
                         const int *axis, const double *cofm, int nlos, int *out, long long cap);
int fso_interpolate(int compute_tau, int nbins, int kernel, double box100, double velfac,
                    double atime, double lambda, double gamma, double fosc, double amumass,
                    double tautail, const float *pos, const float *vel, const float *dens,
                    const float *temp, const float *h, long long npart, const int *axis,
                    const double *cofm, int nlos, double *out);
/* Voronoi mesh: segment ends (FP32, 2 per particle) of one sightline's
 * neighbour list parts[0..n); returns 1 where the reference exits. */
int fso_assign_cells(double box, const float *pos, const int *parts, long long n, int axis, const double *cofm3,
                     float *seg);
/* pairs, tau window pixels, Voigt samples, colden pixels of the last fso_interpolate */
void fso_counters(long long *out4);

#ifdef __cplusplus
}
#endif
#endif
