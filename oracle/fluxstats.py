"""Test infrastructure only (the oracle): numpy restatement of the flux
statistics that follow the optical depth, for checking the GPU versions.

  mean_flux_scale  py_module.cpp:235-262 (Newton-Raphson on the mean flux)
  flux_hist        fluxstatistics.py:41-49 (numpy.histogram, density)
  flux_power       fluxstatistics.py:49-56, 71-108, 197-215
  res_corr         spec_utils.py:5-25 (scipy.ndimage.gaussian_filter1d, wrap)
  observer_lines   spectra.py:914-938 (get_observer_tau's choice of line)

Pinned by the reference's own tests, restated in tests/test_fluxstats.py
(test_statistics.py:8-72, test_spectra.py: testrescorr), by Parseval's
theorem, and by the library routines numpy.fft and scipy.ndimage.
observer_lines is pinned by hand-built cases of each branch of the rule
(tests/test_fluxstats.py).
"""
import math

import numpy as np
from scipy.ndimage import gaussian_filter1d


def mean_flux_scale(tau, mean_flux_desired, tol=1e-5, thresh=1e30):
    tau = np.asarray(tau, dtype=np.float64).ravel()
    if tau.size == 0:
        return 0.0
    t = tau[tau <= thresh]
    newscale = 1.0
    while True:
        scale = newscale
        e = np.exp(-scale * t)
        mf, tmf = e.sum(), (e * t).sum()
        newscale = scale + (mf - mean_flux_desired * t.size) / tmf
        if newscale <= 0:
            newscale = 1e-10
        if not abs(newscale - scale) > tol * newscale:
            return newscale


def flux_hist(tau, nbins=20, scale=1.0):
    flux = np.exp(-scale * np.asarray(tau, dtype=np.float64))
    bins = np.arange(nbins + 1) / (1.0 * nbins)
    fpdf, _ = np.histogram(flux, bins=bins, density=True)
    return (bins[1:] + bins[:-1]) / 2.0, fpdf


def flux_power(tau, vmax, scale=1.0, mean_flux=None):
    tau = np.asarray(tau, dtype=np.float64)
    if mean_flux is None:
        mean_flux = np.mean(np.exp(-tau))
    nspec, npix = tau.shape
    d = np.exp(-scale * tau) / mean_flux - 1.0
    p = np.abs(np.fft.rfft(d, axis=1)) ** 2 / npix ** 2
    power = vmax * p.sum(axis=0) / nspec
    kf = np.fft.rfftfreq(npix) * 2.0 * math.pi * npix / vmax
    return kf, power


def res_corr(flux, dvbin, fwhm=8):
    sigma = (fwhm / dvbin) / (2 * np.sqrt(2 * np.log(2)))
    return gaussian_filter1d(np.asarray(flux, dtype=np.float64), sigma, axis=-1, mode="wrap")


def observer_lines(maxtaus):
    """Index of the line get_observer_tau keeps for each sightline, from
    maxtaus [nlines, nlos] (the largest smoothed optical depth of each line on
    each sightline), one sightline at a time as spectra.py:918-936 does: the
    strongest line with 0.1 < maxtau < 3; failing that the weakest line with
    maxtau > 0.1; failing that the strongest line; the first of equal values."""
    m = np.asarray(maxtaus, dtype=np.float64)
    nl, nlos = m.shape
    out = np.zeros(nlos, dtype=np.int64)
    for ii in range(nlos):
        col = m[:, ii]
        band = [l for l in range(nl) if 0.1 < col[l] < 3]
        if band:
            target = max(col[l] for l in band)
        else:
            above = [l for l in range(nl) if col[l] > 0.1]
            target = min(col[l] for l in above) if above else max(col)
        out[ii] = next(l for l in range(nl) if col[l] == target)
    return out
