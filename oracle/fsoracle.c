/* Test infrastructure only (the oracle).
 *
 * FP64 C restatement of fake_spectra's _Particle_Interpolate path, written
 * from the reference's algorithm (file:line citations below), for checking
 * the GPU product.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, and only as the checker; the product never links it.
 *
 * Pinning (what checks each function independently of this file):
 *   fso_w / fso_profile   scipy.special.wofz (the same Faddeeva package the
 *                         reference vendors), the reference's 57-point w(z)
 *                         table (Faddeeva.cpp:1919-2106), test.cpp:284-302
 *                         known answers, and the compiled reference.
 *   fso_sph_cubic_kernel  closed form + test.cpp:37-46
 *   fso_sph_quintic_kernel closed forms (normalisation 4 pi int K q^2 dq = 1,
 *                         value and continuity at q = 0, 1/3, 2/3, 1) + the
 *                         compiled reference
 *   fso_kern_frac         test.cpp:48-79 + exact integral of the kernel
 *   fso_tau_kern_outer    test.cpp:304-332 + compiled reference
 *   fso_add_*_particle    test.cpp:83-151, 339-404 + compiled reference
 *   binning               test.cpp:164-282 + compiled reference (exact sets, dr2)
 *   fso_interpolate       compiled reference (whole path), additivity over
 *                         particle subsets, tau/colden ratio
 *
 * w(z) is not computed with the reference's algorithm (Zaghloul-Ali Alg. 916
 * plus a continued fraction, Faddeeva.cpp:39-62).  This file uses:
 *   - real axis: w(x) = exp(-x^2) + i 2/sqrt(pi) D(x), Dawson's integral D by
 *     Rybicki's exponentially convergent sum (Rybicki 1989, Computers in
 *     Physics 3, 85), Taylor series near 0;
 *   - |z| >= 8 or y >= 2: the Laplace continued fraction
 *     w(z) = (i/sqrt(pi)) / (z - (1/2)/(z - 1/(z - (3/2)/(z - ...))));
 *   - otherwise: Taylor series in y about the real axis, with the derivatives
 *     from w' = -2zw + 2i/sqrt(pi), w^(k+1) = -2z w^(k) - 2k w^(k-1).
 */
#include "fsoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define NGRID 8                 /* singleabs.h:8 */
#define TOP_HAT_KERNEL 0        /* singleabs.h:9 */
#define SPH_CUBIC_SPLINE 1      /* singleabs.h:10 */
#define VORONOI_MESH 2          /* singleabs.h:11 */
#define SPH_QUINTIC_SPLINE 3    /* singleabs.h:12 */
#define RESO 0.1                /* index_table.h:8, grid spacing of the mesh assignment */

/* absorption.cpp:21-33, absorption.h:4 */
#define SIGMA_T 6.652458558e-25
#define BOLTZMANN 1.3806504e-16
#define LIGHT 2.99792458e10
#define PROTONMASS 1.67262178e-24

static const double SQRT_PI = 1.77245385090551602729;
static const double TWO_OVER_SQRT_PI = 1.12837916709551257390;

void fso_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int fso_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* Faddeeva function                                                         */

#define RYB_H 0.2
#define RYB_NT 22

static double ryb_gauss[RYB_NT]; /* exp(-(n h)^2), n = 1,3,5,... */
static int ryb_init_done = 0;

static void ryb_init(void)
{
    if (ryb_init_done) return;
    for (int i = 0; i < RYB_NT; i++) {
        double n = 2 * i + 1;
        ryb_gauss[i] = exp(-(n * RYB_H) * (n * RYB_H));
    }
    ryb_init_done = 1;
}

/* Dawson's integral D(x) = exp(-x^2) int_0^x exp(t^2) dt. */
static double dawson(double x)
{
    double ax = fabs(x);
    if (ax < 0.2) {
        /* D(x) = x sum_k (-2x^2)^k / (2k+1)!! */
        double t = 1.0, s = 1.0, m = -2.0 * ax * ax;
        for (int k = 1; k < 14; k++) {
            t *= m / (2 * k + 1);
            s += t;
        }
        return x * s;
    }
    /* Rybicki: D(x) = lim_{h->0} pi^{-1/2} sum_{n odd} exp(-(x - n h)^2) / n,
     * re-centred on the even multiple n0 of h nearest to x. */
    double n0 = 2.0 * floor(0.5 * ax / RYB_H + 0.5);
    double xp = ax - n0 * RYB_H;
    double e1 = exp(2.0 * xp * RYB_H), e2 = e1 * e1;
    double up = e1, dn = 1.0 / e1, s = 0.0;
    for (int i = 0; i < RYB_NT; i++) {
        double n = 2 * i + 1;
        s += ryb_gauss[i] * (up / (n0 + n) + dn / (n0 - n));
        up *= e2;
        dn /= e2;
    }
    double d = exp(-xp * xp) * s / SQRT_PI;
    return x < 0 ? -d : d;
}

static void w_contfrac(double x, double y, int nterms, double *wr, double *wi)
{
    /* T = z - (k/2)/T, bottom-up; w = (i/sqrt(pi)) / T */
    double tr = x, ti = y;
    for (int k = nterms; k >= 1; k--) {
        double den = tr * tr + ti * ti;
        double c = 0.5 * k / den;
        tr = x - c * tr;
        ti = y + c * ti;
    }
    double den = tr * tr + ti * ti;
    /* i/T = i conj(T)/|T|^2 = (ti + i tr)/|T|^2 */
    *wr = ti / den / SQRT_PI;
    *wi = tr / den / SQRT_PI;
}

static void w_taylor_y(double x, double y, double *wr, double *wi)
{
    /* w(x+iy) = sum_k (iy)^k/k! w^(k)(x) */
    double w0r = exp(-x * x), w0i = TWO_OVER_SQRT_PI * dawson(x);
    double w1r = -2.0 * x * w0r, w1i = -2.0 * x * w0i + TWO_OVER_SQRT_PI;
    double sr = w0r, si = w0i;
    double cr = 0.0, ci = y; /* (iy)^1 / 1! */
    sr += cr * w1r - ci * w1i;
    si += cr * w1i + ci * w1r;
    double pr = w0r, pi = w0i, qr = w1r, qi = w1i;
    int small = 0;
    for (int k = 1; k < 200; k++) {
        double nr = -2.0 * x * qr - 2.0 * k * pr;
        double ni = -2.0 * x * qi - 2.0 * k * pi;
        pr = qr; pi = qi; qr = nr; qi = ni;
        /* c *= iy/(k+1) */
        double t = cr;
        cr = -ci * y / (k + 1);
        ci = t * y / (k + 1);
        double tr = cr * qr - ci * qi, ti = cr * qi + ci * qr;
        sr += tr;
        si += ti;
        double mag = fabs(tr) + fabs(ti), ref = fabs(sr) + fabs(si);
        if (mag <= 1e-19 * ref) {
            if (++small >= 3) break;
        } else {
            small = 0;
        }
    }
    *wr = sr;
    *wi = si;
}

void fso_w(double x, double y, double *wr, double *wi)
{
    ryb_init();
    double ax = fabs(x), r, i;
    if (y == 0.0) {
        r = exp(-ax * ax);
        i = TWO_OVER_SQRT_PI * dawson(ax);
    } else if (ax * ax + y * y >= 64.0) {
        w_contfrac(ax, y, 60, &r, &i);
    } else if (y >= 2.0) {
        w_contfrac(ax, y, 400, &r, &i);
    } else {
        w_taylor_y(ax, y, &r, &i);
    }
    *wr = r;
    *wi = x < 0 ? -i : i;
}

/* singleabs.h:56-61 */
double fso_profile(double u, double a)
{
    double r, i;
    fso_w(u, a, &r, &i);
    return r;
}

/* singleabs.h:17-26 */
double fso_sph_cubic_kernel(double q)
{
    const double norm = 32. / 4 / M_PI;
    if (q >= 1) return 0;
    if (q < 0.5) return norm * (1 - 6 * q * q + 6 * q * q * q);
    return norm * (2 * pow(1. - q, 3));
}

/* singleabs.h:31-42: the quintic spline of support 1, written as a sum of
   truncated fifth powers, K = 9/(40 pi) sum_k c_k max(0, r_k - 3q)^5 with
   (r, c) = (3, 1), (2, -6), (1, 15) */
double fso_sph_quintic_kernel(double q)
{
    const double norm = 9. / 40 / M_PI;
    if (q >= 1) return 0;
    const double t3 = 3. - 3. * q, t2 = 2. - 3. * q, t1 = 1. - 3. * q;
    double f = t3 * t3 * t3 * t3 * t3;
    if (t2 > 0) f -= 6. * (t2 * t2 * t2 * t2 * t2);
    if (t1 > 0) f += 15. * (t1 * t1 * t1 * t1 * t1);
    return norm * f;
}

/* singleabs.h:158-163: the weight of one kernel sample; an id that matches
   none of the branches leaves the sample unweighted */
double fso_kernel_weight(int kernel, double q)
{
    if (kernel == SPH_CUBIC_SPLINE) return fso_sph_cubic_kernel(q);
    if (kernel == SPH_QUINTIC_SPLINE) return fso_sph_quintic_kernel(q);
    if (kernel == TOP_HAT_KERNEL) return 3. / 4. / M_PI;
    return 1.0;
}

/* absorption.cpp:53-74 (cubic), 76-97 (quintic) and 109-115 (tophat):
   8-interval trapezoid of the kernel over [zlow, zhigh] clamped to the support */
double fso_kern_frac(int kernel, double zlow, double zhigh, double smooth, double dr2, double zrange)
{
    zlow = fmax(zlow, -zrange);
    zhigh = fmin(zhigh, zrange);
    if (kernel == VORONOI_MESH) return fmax(0., zhigh - zlow);           /* absorption.cpp:129-134 */
    if (kernel != SPH_CUBIC_SPLINE && kernel != SPH_QUINTIC_SPLINE) return 3. / 4. / M_PI * fmax(0., zhigh - zlow);
    if (zlow > zhigh) return 0;
    double total = fso_kernel_weight(kernel, sqrt(dr2 + zlow * zlow) / smooth) / 2.;
    const double dz = (zhigh - zlow) / NGRID;
    for (int i = 1; i < NGRID; i++) {
        const double zz = i * dz + zlow;
        total += fso_kernel_weight(kernel, sqrt(dr2 + zz * zz) / smooth);
    }
    total += fso_kernel_weight(kernel, sqrt(dr2 + zhigh * zhigh) / smooth) / 2.;
    return dz * total;
}

/* ------------------------------------------------------------------------ */
/* SingleAbsorber (singleabs.h:63-175)                                       */

typedef struct {
    double btherm, vdr2, vsmooth, aa, vhigh;
    int kernel;
    long long samples;
} absorber_t;

static void absorber_init(absorber_t *s, double b, double vdr2, double vsm, double aa, int kernel)
{
    s->btherm = b;
    s->vdr2 = vdr2;
    s->vsmooth = vsm;
    s->aa = aa;
    s->kernel = kernel;
    s->vhigh = vsm * vsm > vdr2 ? sqrt(vsm * vsm - vdr2) : 0;   /* singleabs.h:83 */
    /* singleabs.h:84-88: on the mesh, vdr2 and vsm are the segment's lower and
       upper ends in velocity, and the kernel spans half the segment */
    if (kernel == VORONOI_MESH) s->vhigh = (vdr2 > 0 && vsm > 0) ? (vsm - vdr2) / 2. : 0;
    s->samples = 0;
}

/* singleabs.h:143-167: 7 interior points of the 8-interval grid over the kernel */
static double tau_kern_inner(absorber_t *s, double vouter)
{
    const double dv = 2. * s->vhigh / NGRID;
    double total = 0;
    for (int i = 1; i < NGRID; i++) {
        const double vv = i * dv - s->vhigh;
        const double q = sqrt(s->vdr2 + vv * vv) / s->vsmooth;
        const double t0 = (vv - vouter) / s->btherm;
        const double tbin = fso_profile(t0, s->aa) * fso_kernel_weight(s->kernel, q);
        total += tbin;
    }
    s->samples += NGRID - 1;
    return dv * total;
}

/* singleabs.h:104-126 */
static double tau_kern_outer(absorber_t *s, double vlow, double vhigh)
{
    if ((vhigh - vlow) < s->btherm / 2.) return tau_kern_inner(s, (vhigh + vlow) / 2.);
    const int npoints = 2 * ceil((vhigh - vlow) / (s->btherm / 2.) / 2) + 1.;
    double total = tau_kern_inner(s, vlow) / 2.;
    const double dv = (vhigh - vlow) / (npoints - 1);
    for (int i = 1; i < npoints - 1; i++) total += tau_kern_inner(s, i * dv + vlow);
    total += tau_kern_inner(s, vhigh) / 2.;
    return total / (npoints - 1);
}

double fso_tau_kern_outer(double btherm, double vdr2, double vsmooth, double aa, int kernel,
                          double vlow, double vhigh)
{
    absorber_t s;
    absorber_init(&s, btherm, vdr2, vsmooth, aa, kernel);
    return tau_kern_outer(&s, vlow, vhigh);
}

/* ------------------------------------------------------------------------ */
/* LineAbsorption (absorption.cpp:152-279)                                   */

typedef struct {
    double tautail, sigma_a, bfac, voigt_fac, velfac, vbox;
    int kernel;
} line_t;

static void line_init(line_t *l, double lambda, double gamma, double fosc, double amumass,
                      double velfac, double box, int kernel, double tautail)
{
    l->tautail = tautail;
    l->sigma_a = sqrt(3.0 * M_PI * SIGMA_T / 8.0) * lambda * fosc;
    l->bfac = sqrt(2.0 * BOLTZMANN / (amumass * PROTONMASS)) / 1e5;
    l->voigt_fac = gamma * lambda / (4. * M_PI) / 1e5;
    l->velfac = velfac;
    l->vbox = box * velfac;
    l->kernel = kernel;
}

typedef struct { long long pairs, tau_pixels, samples, colden_pixels; } counters_t;

static void add_colden(const line_t *l, double *colden, int nbins, double dr2, float dens,
                       float pos, float smooth, counters_t *c)
{
    double pos1 = pos;
    double zrange;
    if (l->kernel == VORONOI_MESH) {
        /* absorption.cpp:170-186: dr2 and smooth are the segment's ends */
        if (dr2 > 2 * l->vbox / l->velfac || smooth > 2 * l->vbox / l->velfac) return;
        pos1 = (dr2 + smooth) / 2.;
        zrange = (smooth - dr2) / 2.;
    } else {
        if (smooth * smooth - dr2 <= 0) return;                  /* absorption.cpp:182 */
        zrange = sqrt(smooth * smooth - dr2);
    }
    const double boxtokpc = l->vbox / nbins / l->velfac;
    const int zlow = floor((pos1 - zrange) / boxtokpc);
    const int zhigh = ceil((pos1 + zrange) / boxtokpc);
    if (c) c->pairs++;
    for (int z = zlow; z <= zhigh; z++) {
        const double plow = (boxtokpc * z - pos1);
        int j = z % nbins;
        if (j < 0) j += nbins;
        colden[j] += dens * fso_kern_frac(l->kernel, plow, plow + boxtokpc, smooth, dr2, zrange);
        if (c) c->colden_pixels++;
    }
}

static void add_tau(const line_t *l, double *tau, int nbins, double dr2, float dens, float ppos,
                    float pvel, float temp, float smooth, counters_t *c)
{
    double pos1 = ppos;
    const double btherm = l->bfac * sqrt(temp);
    if (l->kernel == VORONOI_MESH) {
        /* absorption.cpp:221-230: dr2 and smooth are the segment's ends */
        if (dr2 > 2 * l->vbox / l->velfac || smooth > 2 * l->vbox / l->velfac) return;
        pos1 = (dr2 + smooth) / 2.;
    } else if (smooth * smooth - dr2 <= 0) {
        return;                                                  /* absorption.cpp:232 */
    }
    const double vel = l->velfac * pos1 + pvel;
    double val1 = l->velfac * dr2;
    if (l->kernel != VORONOI_MESH) val1 *= l->velfac;           /* absorption.cpp:236-237 */
    absorber_t s;
    absorber_init(&s, btherm, val1, l->velfac * smooth, l->voigt_fac / btherm, l->kernel);
    const double bintov = l->vbox / nbins;
    const double amp = l->sigma_a / sqrt(M_PI) * (LIGHT / 1e5 / btherm);
    const int zmax = floor(vel / bintov);
    long long npix = 0;
    for (int z = zmax; z < zmax + nbins / 2; ++z) {
        double vlow = z * bintov - vel;
        const double taulast = amp * dens * tau_kern_outer(&s, vlow, vlow + bintov) / l->velfac;
        int j = z % nbins;
        if (j < 0) j += nbins;
        tau[j] += taulast;
        npix++;
        if (taulast < l->tautail) break;
    }
    for (int z = zmax - 1; z >= zmax - nbins / 2; --z) {
        double vlow = z * bintov - vel;
        const double taulast = amp * dens * tau_kern_outer(&s, vlow, vlow + bintov) / l->velfac;
        int j = z % nbins;
        if (j < 0) j += nbins;
        tau[j] += taulast;
        npix++;
        if (taulast < l->tautail) break;
    }
    if (c) {
        c->pairs++;
        c->tau_pixels += npix;
        c->samples += s.samples;
    }
}

void fso_add_tau_particle(double lambda, double gamma, double fosc, double amumass, double velfac,
                          double box, double atime, int kernel, double tautail, double *tau,
                          int nbins, double dr2, float dens, float ppos, float pvel, float temp,
                          float smooth)
{
    (void)atime;
    ryb_init();
    line_t l;
    line_init(&l, lambda, gamma, fosc, amumass, velfac, box, kernel, tautail);
    add_tau(&l, tau, nbins, dr2, dens, ppos, pvel, temp, smooth, NULL);
}

void fso_add_colden_particle(double lambda, double gamma, double fosc, double amumass,
                             double velfac, double box, double atime, int kernel, double tautail,
                             double *colden, int nbins, double dr2, float dens, float ppos,
                             float smooth)
{
    (void)atime;
    line_t l;
    line_init(&l, lambda, gamma, fosc, amumass, velfac, box, kernel, tautail);
    add_colden(&l, colden, nbins, dr2, dens, ppos, smooth, NULL);
}

/* ------------------------------------------------------------------------ */
/* IndexTable (index_table.cpp:7-150): two key-sorted tables of sightlines.  */

typedef struct { double key; int idx; } kv_t;

typedef struct {
    kv_t *xx;  /* axis 1, keyed by cofm y */
    kv_t *t;   /* axis 2,3, keyed by cofm x */
    int nxx, nt;
    const double *cofm;
    const int *axis;
    double box;
} table_t;

static int kv_cmp(const void *a, const void *b)
{
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return x->idx - y->idx;   /* multimap keeps equal keys in insertion order */
}

static void table_init(table_t *tb, const double *cofm, const int *axis, int nlos, double box)
{
    tb->xx = (kv_t *)malloc(sizeof(kv_t) * (nlos + 1));
    tb->t = (kv_t *)malloc(sizeof(kv_t) * (nlos + 1));
    tb->nxx = tb->nt = 0;
    for (int i = 0; i < nlos; i++) {
        if (axis[i] == 1) {
            tb->xx[tb->nxx].key = cofm[3 * i + 1];
            tb->xx[tb->nxx++].idx = i;
        } else {
            tb->t[tb->nt].key = cofm[3 * i];
            tb->t[tb->nt++].idx = i;
        }
    }
    qsort(tb->xx, tb->nxx, sizeof(kv_t), kv_cmp);
    qsort(tb->t, tb->nt, sizeof(kv_t), kv_cmp);
    tb->cofm = cofm;
    tb->axis = axis;
    tb->box = box;
}

static void table_free(table_t *tb)
{
    free(tb->xx);
    free(tb->t);
}

/* first position with key >= v (std::multimap::lower_bound) */
static int lower_bound(const kv_t *a, int n, double v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid].key < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* index_table.cpp:52-68 */
static int second_close(double box, const float second, const double lproj2, const float hh)
{
    float ffp = second + hh;
    if (ffp > box)
        if (lproj2 < ffp - box) return 1;
    float ffm = second - hh;
    if (ffm < 0)
        if (lproj2 > ffm + box) return 1;
    return lproj2 > ffm && lproj2 < ffp;
}

/* index_table.cpp:70-87 */
static double calc_dr2(double box, const double d1, const double d2)
{
    double dr = fabs(d1), dr2;
    if (dr > 0.5 * box) dr = box - dr;
    dr2 = dr * dr;
    dr = fabs(d2);
    if (dr > 0.5 * box) dr = box - dr;
    dr2 += (dr * dr);
    return dr2;
}

typedef struct { int line; double dr2; } hit_t;

/* index_table.cpp:22-50 */
static int from_range(const table_t *tb, const kv_t *a, int lo, int hi, const float *pos,
                      const float hh, const float first, hit_t *out, int n)
{
    for (int k = lo; k < hi; k++) {
        const int iproc = a[k].idx;
        const int iaxis = tb->axis[iproc];
        float second;
        double lproj2;
        if (iaxis == 3) {
            second = pos[1];
            lproj2 = tb->cofm[3 * iproc + 1];
        } else {
            second = pos[2];
            lproj2 = tb->cofm[3 * iproc + 2];
        }
        const double lproj = a[k].key;
        if (second_close(tb->box, second, lproj2, hh)) {
            double dr2 = calc_dr2(tb->box, first - lproj, second - lproj2);
            if (dr2 <= hh * hh) {
                out[n].line = iproc;
                out[n].dr2 = dr2;
                n++;
            }
        }
    }
    return n;
}

/* index_table.cpp:89-113 */
static int get_nearby(const table_t *tb, float first, const kv_t *a, int na, const float *pos,
                      const float hh, hit_t *out, int n)
{
    float ffp = first + hh;
    if (ffp > tb->box) ffp -= tb->box;
    float ffm = first - hh;
    if (ffm < 0) ffm += tb->box;
    int low = lower_bound(a, na, ffm);
    int high = lower_bound(a, na, ffp);
    if (ffm <= ffp) {
        n = from_range(tb, a, low, high, pos, hh, first, out, n);
    } else {
        n = from_range(tb, a, 0, high, pos, hh, first, out, n);
        n = from_range(tb, a, low, na, pos, hh, first, out, n);
    }
    return n;
}

static int hit_cmp(const void *a, const void *b)
{
    return ((const hit_t *)a)->line - ((const hit_t *)b)->line;
}

/* index_table.cpp:117-127; out must hold nlos entries; result sorted by line */
static int near_lines_of(const table_t *tb, const float *pos, float hh, hit_t *out)
{
    int n = 0;
    if (tb->nt > 0) n = get_nearby(tb, pos[0], tb->t, tb->nt, pos, hh, out, n);
    if (tb->nxx > 0) n = get_nearby(tb, pos[1], tb->xx, tb->nxx, pos, hh, out, n);
    qsort(out, n, sizeof(hit_t), hit_cmp);
    return n;
}

int fso_get_near_lines(double box, const int *axis, const double *cofm, int nlos,
                       const float *pos3, float hh, int *lines, double *dr2, int cap)
{
    table_t tb;
    table_init(&tb, cofm, axis, nlos, box);
    hit_t *h = (hit_t *)malloc(sizeof(hit_t) * (nlos + 1));
    int n = near_lines_of(&tb, pos3, hh, h);
    for (int k = 0; k < n && k < cap; k++) {
        lines[k] = h[k].line;
        dr2[k] = h[k].dr2;
    }
    free(h);
    table_free(&tb);
    return n;
}

/* CSR of index_table.cpp:130-150, rows in ascending particle order.
 * Returns malloc'd offsets[nlos+1], parts[M], dr2[M]. */
static long long build_pairs(const table_t *tb, int nlos, const float *pos, const float *h,
                             long long npart, long long **offs_out, int **parts_out,
                             double **dr2_out)
{
    int nth = omp_get_max_threads();
    long long *cnt = (long long *)calloc((size_t)nth * (nlos + 1), sizeof(long long));
    long long chunk = (npart + nth - 1) / nth;
    /* pass 1: per-thread, per-line counts over contiguous particle ranges */
#pragma omp parallel num_threads(nth)
    {
        int t = omp_get_thread_num();
        hit_t *buf = (hit_t *)malloc(sizeof(hit_t) * (nlos + 1));
        long long lo = t * chunk, hi = lo + chunk < npart ? lo + chunk : npart;
        for (long long p = lo; p < hi; p++) {
            int n = near_lines_of(tb, &pos[3 * p], h[p], buf);
            for (int k = 0; k < n; k++) cnt[(size_t)t * (nlos + 1) + buf[k].line]++;
        }
        free(buf);
    }
    long long *offs = (long long *)malloc(sizeof(long long) * (nlos + 1));
    long long *cur = (long long *)malloc(sizeof(long long) * (size_t)nth * (nlos + 1));
    long long total = 0;
    for (int l = 0; l < nlos; l++) {
        offs[l] = total;
        for (int t = 0; t < nth; t++) {
            cur[(size_t)t * (nlos + 1) + l] = total;
            total += cnt[(size_t)t * (nlos + 1) + l];
        }
    }
    offs[nlos] = total;
    int *parts = (int *)malloc(sizeof(int) * (total + 1));
    double *dr2 = (double *)malloc(sizeof(double) * (total + 1));
    /* pass 2: fill, so each row is ordered by particle index */
#pragma omp parallel num_threads(nth)
    {
        int t = omp_get_thread_num();
        hit_t *buf = (hit_t *)malloc(sizeof(hit_t) * (nlos + 1));
        long long lo = t * chunk, hi = lo + chunk < npart ? lo + chunk : npart;
        for (long long p = lo; p < hi; p++) {
            int n = near_lines_of(tb, &pos[3 * p], h[p], buf);
            for (int k = 0; k < n; k++) {
                long long at = cur[(size_t)t * (nlos + 1) + buf[k].line]++;
                parts[at] = (int)p;
                dr2[at] = buf[k].dr2;
            }
        }
        free(buf);
    }
    free(cnt);
    free(cur);
    *offs_out = offs;
    *parts_out = parts;
    *dr2_out = dr2;
    return total;
}

long long fso_near_particles(double box, const float *pos, const float *h, long long npart,
                             const int *axis, const double *cofm, int nlos, long long *offsets,
                             int *parts, double *dr2, long long cap)
{
    table_t tb;
    table_init(&tb, cofm, axis, nlos, box);
    long long *o;
    int *p;
    double *d;
    long long m = build_pairs(&tb, nlos, pos, h, npart, &o, &p, &d);
    memcpy(offsets, o, sizeof(long long) * (nlos + 1));
    if (cap >= m) {
        memcpy(parts, p, sizeof(int) * m);
        memcpy(dr2, d, sizeof(double) * m);
    }
    free(o);
    free(p);
    free(d);
    table_free(&tb);
    return m;
}

/* py_module.cpp:25-99: sorted indices of particles near at least one line */
long long fso_near_lines(double box, const float *pos, const float *h, long long npart,
                         const int *axis, const double *cofm, int nlos, int *out, long long cap)
{
    table_t tb;
    table_init(&tb, cofm, axis, nlos, box);
    char *flag = (char *)calloc((size_t)npart + 1, 1);
#pragma omp parallel
    {
        hit_t *buf = (hit_t *)malloc(sizeof(hit_t) * (nlos + 1));
#pragma omp for schedule(static)
        for (long long p = 0; p < npart; p++) flag[p] = near_lines_of(&tb, &pos[3 * p], h[p], buf) > 0;
        free(buf);
    }
    long long n = 0;
    for (long long p = 0; p < npart; p++) {
        if (flag[p]) {
            if (n < cap) out[n] = (int)p;
            n++;
        }
    }
    free(flag);
    table_free(&tb);
    return n;
}

/* ------------------------------------------------------------------------ */
/* IndexTable::assign_cells (index_table.cpp:152-223): the Voronoi mesh's cut
   of sightline `cofm3` (axis `axis`) into segments, one per particle of the
   line's neighbour list parts[0..n).  Points at (i + 1/2) reso along the line
   go to the nearest particle (periodic distance; the first in list order on
   ties); a particle's segment runs from its first to its last point, +-reso/2.
   The particle owning the first point may own the last points as well
   (periodic wrap): at its first point past the middle of the box its segment
   is restarted there and extended by a box length, and the walk stops.
   seg[2k], seg[2k+1]: segment ends of particle k in FP32, 3 box - reso/2 and
   3 box + reso/2 when it owns no point.  Returns 1 where the reference
   exits: a particle owning two separate runs of points. */
int fso_assign_cells(double box, const float *pos, const int *parts, long long n, int axis, const double *cofm3,
                     float *seg)
{
    const int npts = (int)(box / RESO);
    const double reso = box / npts;
    for (long long k = 0; k < 2 * n; k++) seg[k] = 3 * box;
    if (n == 0) return 0;
    const double yp = cofm3[axis % 3], zp = cofm3[(axis + 1) % 3];
    for (int i = 0; i < npts; i++) {
        const double xp = (i + 0.5) * reso;
        double min_dist = box;
        long long min_ind = 0;
        for (long long k = 0; k < n; k++) {
            const float *pp = &pos[3 * (size_t)parts[k]];
            double dx = fabs(pp[axis - 1] - xp);
            if (dx > box / 2.) dx = box - dx;
            double dy = fabs(pp[axis % 3] - yp);
            if (dy > box / 2.) dy = box - dy;
            double dz = fabs(pp[(axis + 1) % 3] - zp);
            if (dz > box / 2.) dz = box - dz;
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < min_dist) {
                min_dist = dist;
                min_ind = k;
            }
        }
        if (seg[2 * min_ind] < reso && xp > box / 2. + 0.5 * reso) {
            seg[2 * min_ind] = xp;
            seg[2 * min_ind + 1] += box;
            break;
        }
        if (xp > 1.5 * reso + seg[2 * min_ind + 1]) return 1;
        if (seg[2 * min_ind] > 2 * box) seg[2 * min_ind] = xp;
        seg[2 * min_ind + 1] = xp;
    }
    for (long long k = 0; k < n; k++) {
        seg[2 * k] -= 0.5 * reso;
        seg[2 * k + 1] += 0.5 * reso;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ParticleInterp::compute_tau / compute_colden (part_int.cpp:20-84)         */

static counters_t last_counters;

void fso_counters(long long *out4)
{
    out4[0] = last_counters.pairs;
    out4[1] = last_counters.tau_pixels;
    out4[2] = last_counters.samples;
    out4[3] = last_counters.colden_pixels;
}

int fso_interpolate(int compute_tau, int nbins, int kernel, double box100, double velfac,
                    double atime, double lambda, double gamma, double fosc, double amumass,
                    double tautail, const float *pos, const float *vel, const float *dens,
                    const float *temp, const float *h, long long npart, const int *axis,
                    const double *cofm, int nlos, double *out)
{
    (void)atime;
    ryb_init();
    memset(out, 0, sizeof(double) * (size_t)nlos * (size_t)nbins);
    table_t tb;
    table_init(&tb, cofm, axis, nlos, box100);
    long long *offs;
    int *parts;
    double *dr2;
    build_pairs(&tb, nlos, pos, h, npart, &offs, &parts, &dr2);
    line_t l;
    line_init(&l, lambda, gamma, fosc, amumass, velfac, box100, kernel, tautail);
    long long pairs = 0, tpix = 0, samp = 0, cpix = 0;
    int failed = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : pairs, tpix, samp, cpix, failed)
    for (int i = 0; i < nlos; i++) {
        const int ax = axis[i];
        double *row = &out[(size_t)i * nbins];
        counters_t c = {0, 0, 0, 0};
        const long long n = offs[i + 1] - offs[i];
        float *seg = NULL;
        if (kernel == VORONOI_MESH) {
            seg = (float *)malloc(sizeof(float) * (size_t)(2 * n + 2));
            if (fso_assign_cells(box100, pos, &parts[offs[i]], n, ax, &cofm[3 * (size_t)i], seg)) {
                failed++;
                free(seg);
                continue;
            }
        }
        for (long long k = offs[i]; k < offs[i + 1]; k++) {
            const int p = parts[k];
            const float ppos = pos[3 * (size_t)p + ax - 1];
            /* part_int.cpp:42-45: on the mesh, the segment ends replace dr2 and h */
            const double d2 = seg ? (double)seg[2 * (k - offs[i])] : dr2[k];
            const float sm = seg ? seg[2 * (k - offs[i]) + 1] : h[p];
            if (compute_tau)
                add_tau(&l, row, nbins, d2, dens[p], ppos, vel[3 * (size_t)p + ax - 1], temp[p], sm, &c);
            else
                add_colden(&l, row, nbins, d2, dens[p], ppos, sm, &c);
        }
        free(seg);
        pairs += c.pairs;
        tpix += c.tau_pixels;
        samp += c.samples;
        cpix += c.colden_pixels;
    }
    last_counters.pairs = pairs;
    last_counters.tau_pixels = tpix;
    last_counters.samples = samp;
    last_counters.colden_pixels = cpix;
    free(offs);
    free(parts);
    free(dr2);
    table_free(&tb);
    return failed ? 1 : 0;
}
