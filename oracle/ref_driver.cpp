/* Test infrastructure only (the oracle).  Never linked into the product.
 *
 * C-ABI driver around the reference's own C++ hot path.  It is compiled by
 * oracle/Makefile together with the reference sources where they lie
 * (fake_spectra/{index_table,absorption,part_int,Faddeeva}.cpp), with the
 * flags of the reference extension (setup.py:57,75 plus Python's
 * -O2 -fwrapv -DNDEBUG), into oracle/_ref/libfsref.so.  Nothing of the
 * reference is copied here: this file only calls its public classes.
 *
 * Entry points mirror what fake_spectra/py_module.cpp does around them:
 *   fsref_interpolate      <- Py_Particle_Interpolation (py_module.cpp:103-233)
 *   fsref_near_lines       <- Py_near_lines (py_module.cpp:25-99)
 *   fsref_near_particles   <- IndexTable::get_near_particles (index_table.cpp:130-150)
 * plus unit-level probes of the pieces the reference's own Boost tests
 * exercise (test.cpp:37-404).
 */
#include <complex>
#include <cstring>
#include <map>
#include <set>
#include <valarray>
#include <omp.h>

#include "Faddeeva.h"
#include "absorption.h"
#include "index_table.h"
#include "part_int.h"
#include "singleabs.h"

/* Defined with external linkage in absorption.cpp:109, not declared in a header. */
double tophat_kern_frac(double zlow, double zhigh, const double smooth, const double dr2, const double zrange);
double sph_quintic_kern_frac(double zlow, double zhigh, const double smooth, const double dr2, const double zrange);

extern "C" {

void fsref_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int fsref_max_threads(void) { return omp_get_max_threads(); }

/* Same work as _Particle_Interpolate: zero-filled float64 [nlos, nbins] output. */
int fsref_interpolate(int compute_tau, int nbins, int kernel, double box100, double velfac,
                      double atime, double lambda, double gamma, double fosc, double amumass,
                      double tautail, const float *pos, const float *vel, const float *dens,
                      const float *temp, const float *h, long long npart, const int *axis,
                      const double *cofm, int nlos, double *out)
{
    std::memset(out, 0, sizeof(double) * (size_t)nlos * (size_t)nbins);
    ParticleInterp pint(nbins, lambda, gamma, fosc, amumass, box100, velfac, atime, cofm, axis,
                        nlos, kernel, tautail);
    if (compute_tau)
        pint.compute_tau(out, pos, vel, dens, temp, h, npart);
    else
        pint.compute_colden(out, pos, dens, h, npart);
    return 0;
}

/* Neighbour lists per sightline, ascending particle index (std::map order).
 * offsets has nlos+1 entries.  Returns the pair count; parts/dr2 are only
 * written when cap >= pair count. */
long long fsref_near_particles(double box, const float *pos, const float *h, long long npart,
                               const int *axis, const double *cofm, int nlos,
                               long long *offsets, int *parts, double *dr2, long long cap)
{
    IndexTable tab(cofm, axis, nlos, box);
    std::valarray<std::map<int, double> > near = tab.get_near_particles(pos, h, npart);
    long long total = 0;
    for (int i = 0; i < nlos; i++) {
        offsets[i] = total;
        total += (long long)near[i].size();
    }
    offsets[nlos] = total;
    if (cap < total) return total;
    long long k = 0;
    for (int i = 0; i < nlos; i++)
        for (std::map<int, double>::const_iterator it = near[i].begin(); it != near[i].end(); ++it) {
            parts[k] = it->first;
            dr2[k] = it->second;
            k++;
        }
    return total;
}

/* Sorted unique indices of particles with at least one nearby sightline. */
long long fsref_near_lines(double box, const float *pos, const float *h, long long npart,
                           const int *axis, const double *cofm, int nlos, int *out, long long cap)
{
    IndexTable tab(cofm, axis, nlos, box);
    std::set<int> near;
#pragma omp parallel for
    for (long long i = 0; i < npart; i++) {
        std::map<int, double> nb = tab.get_near_lines(&pos[3 * i], h[i]);
        if (!nb.empty()) {
#pragma omp critical
            near.insert((int)i);
        }
    }
    long long n = (long long)near.size();
    if (cap < n) return n;
    long long k = 0;
    for (std::set<int>::const_iterator it = near.begin(); it != near.end(); ++it) out[k++] = *it;
    return n;
}

/* Sightlines near one particle: IndexTable::get_near_lines (index_table.cpp:117-127). */
int fsref_get_near_lines(double box, const int *axis, const double *cofm, int nlos,
                         const float *pos3, float hh, int *lines, double *dr2, int cap)
{
    IndexTable tab(cofm, axis, nlos, box);
    std::map<int, double> nb = tab.get_near_lines(pos3, hh);
    int k = 0;
    for (std::map<int, double>::const_iterator it = nb.begin(); it != nb.end() && k < cap; ++it) {
        lines[k] = it->first;
        dr2[k] = it->second;
        k++;
    }
    return (int)nb.size();
}

void fsref_w(double re, double im, double *wr, double *wi)
{
    std::complex<double> r = Faddeeva::w(std::complex<double>(re, im));
    *wr = r.real();
    *wi = r.imag();
}

double fsref_profile(double u, double a) { return profile(u, a); }
double fsref_sph_cubic_kernel(double q) { return sph_cubic_kernel(q); }
double fsref_sph_quintic_kernel(double q) { return sph_quintic_kernel(q); }

double fsref_kern_frac(int kernel, double zlow, double zhigh, double smooth, double dr2, double zrange)
{
    if (kernel == SPH_CUBIC_SPLINE) return sph_cubic_kern_frac(zlow, zhigh, smooth, dr2, zrange);
    if (kernel == SPH_QUINTIC_SPLINE) return sph_quintic_kern_frac(zlow, zhigh, smooth, dr2, zrange);
    return tophat_kern_frac(zlow, zhigh, smooth, dr2, zrange);
}

double fsref_tau_kern_outer(double btherm, double vdr2, double vsmooth, double aa, int kernel,
                            double vlow, double vhigh)
{
    SingleAbsorber s(btherm, vdr2, vsmooth, aa, kernel);
    return s.tau_kern_outer(vlow, vhigh);
}

void fsref_add_tau_particle(double lambda, double gamma, double fosc, double amumass, double velfac,
                            double box, double atime, int kernel, double tautail, double *tau,
                            int nbins, double dr2, float dens, float ppos, float pvel, float temp,
                            float smooth)
{
    LineAbsorption la(lambda, gamma, fosc, amumass, velfac, box, atime, kernel, tautail);
    la.add_tau_particle(tau, nbins, dr2, dens, ppos, pvel, temp, smooth);
}

void fsref_add_colden_particle(double lambda, double gamma, double fosc, double amumass,
                               double velfac, double box, double atime, int kernel, double tautail,
                               double *colden, int nbins, double dr2, float dens, float ppos,
                               float smooth)
{
    LineAbsorption la(lambda, gamma, fosc, amumass, velfac, box, atime, kernel, tautail);
    la.add_colden_particle(colden, nbins, dr2, dens, ppos, smooth);
}

} /* extern "C" */
