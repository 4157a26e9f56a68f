"""Flux statistics on the GPU, with the names, arguments and results of the
reference's fluxstatistics.py / spec_utils.py for the functions that follow
the optical depth:

    mean_flux(tau, mean_flux_desired, tol=1e-5, thresh=1e30)   fluxstatistics.py:26-39
    flux_pdf(tau, nbins=20, mean_flux_desired=None)            fluxstatistics.py:41-49
    flux_power(tau, vmax, spec_res=8, mean_flux_desired=None, window=False)
                                                               fluxstatistics.py:71-108
    res_corr(flux, dvbin, fwhm=8)                              spec_utils.py:5-25

tau may be a numpy array (uploaded; results come back as numpy, as the
reference returns them) or a float64 CUDA tensor (used in place; results stay
on the device).  The arithmetic is FP64 throughout.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _native as N
from .spectra_priv import _ptr, _stream


def _on_device(a):
    """(float64 contiguous CUDA tensor, whether the input was numpy)."""
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda":
            raise ValueError("tau must be a CUDA tensor or a numpy array")
        return a.to(torch.float64).contiguous(), False
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return torch.from_numpy(arr).cuda(), True


def mean_flux(tau, mean_flux_desired, tol=1e-5, thresh=1e30):
    """Scale factor s for tau such that <exp(-s tau)> = mean_flux_desired over
    the pixels with tau <= thresh, by Newton-Raphson (0 for an empty tau)."""
    t, _ = _on_device(tau)
    out = C.c_double()
    N.check(N.lib().fsamd_mean_flux_scale(_ptr(t), t.numel(), float(mean_flux_desired), float(tol), float(thresh),
                                          C.byref(out), _stream()))
    return out.value


def flux_pdf(tau, nbins=20, mean_flux_desired=None):
    """Normalised histogram of the flux exp(-s tau) in nbins bins of [0, 1]:
    (bin centres, density)."""
    scale = 1.0
    if mean_flux_desired is not None:
        scale = mean_flux(tau, mean_flux_desired)
    t, was_np = _on_device(tau)
    counts = torch.zeros(nbins, dtype=torch.int64, device=t.device)
    N.check(N.lib().fsamd_flux_hist(_ptr(t), t.numel(), float(scale), int(nbins), _ptr(counts), _stream()))
    bins = np.arange(nbins + 1) / (1.0 * nbins)
    width = np.diff(bins)
    c = counts.cpu().numpy().astype(np.float64)
    total = c.sum()
    fpdf = c / (total * width) if total > 0 else np.full(nbins, np.nan)
    cbins = (bins[1:] + bins[:-1]) / 2.0
    return cbins, fpdf


def _window_function(k, *, R, dv):
    """Spectrograph response: Gaussian of FWHM R times the pixel sinc."""
    sigma = R / (2 * np.sqrt(2 * np.log(2)))
    return np.exp(-0.5 * (k * sigma) ** 2) * np.sinc(k * dv / 2 / math.pi)


def _flux_power_bins(vmax, npix):
    """k of the rfft modes in s/km: 2 pi k / vmax, k = 0..npix/2."""
    return np.fft.rfftfreq(npix) * 2.0 * math.pi * npix / vmax


def flux_power(tau, vmax, spec_res=8, mean_flux_desired=None, window=False):
    """Mean over sightlines of the 1D power of dF = exp(-s tau)/<F> - 1,
    vmax |rfft(dF)|^2 / npix^2; returns (k in s/km, power in km/s)."""
    scale = 1.0
    t, _ = _on_device(tau)
    if t.dim() != 2:
        raise ValueError("tau must have shape (NumLos, npix)")
    if mean_flux_desired is not None:
        scale = mean_flux(t, mean_flux_desired)
    else:
        mean_flux_desired = float(torch.exp(-t).mean().item())
    nspec, npix = t.shape
    power = torch.empty(npix // 2 + 1, dtype=torch.float64, device=t.device)
    N.check(N.lib().fsamd_flux_power(_ptr(t), int(nspec), int(npix), float(scale), float(mean_flux_desired),
                                     float(vmax), _ptr(power), _stream()))
    p = power.cpu().numpy()
    kf = _flux_power_bins(vmax, npix)
    if window and spec_res > 0:
        p = p / _window_function(kf, R=spec_res, dv=vmax / npix) ** 2
    return kf, p


def res_corr(flux, dvbin, fwhm=8):
    """Smooth each spectrum (last axis) with a Gaussian of FWHM fwhm km/s,
    pixels of dvbin km/s, periodic at the ends."""
    t, was_np = _on_device(flux)
    shape = t.shape
    ncol = shape[-1]
    nrow = t.numel() // ncol if ncol else 0
    sigma = (fwhm / dvbin) / (2 * np.sqrt(2 * np.log(2)))
    out = torch.empty_like(t)
    N.check(N.lib().fsamd_res_corr(_ptr(t), _ptr(out), int(nrow), int(ncol), float(sigma), _stream()))
    return out.cpu().numpy().reshape(shape) if was_np else out
