"""GPU flux statistics (fake_spectra_amd.fluxstatistics, csrc/flux.hip)
against the numpy oracle and the reference's own test assertions."""
import math

import numpy as np
import pytest
import torch

import fluxstats as FS
from fake_spectra_amd import fluxstatistics as GF, spectra_priv as SP, synth
from gpu_util import dev, plan_for

pytestmark = pytest.mark.gpu


def _tau_c0():
    w = synth.config(0)
    plan = plan_for(w)
    t = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines,
               [dev(w.dens_for(ln)) for ln in w.lines], w.tautail)[0]
    return w, t


def test_mean_flux_known_answers_and_oracle():
    tol = 1e-4
    nn = np.arange(1, 101)
    tau = np.log(nn)
    for n in (1, 2, 0.5):
        assert abs(GF.mean_flux(tau, np.mean(nn ** (-float(n))), tol) - n) < tol
    assert GF.mean_flux(np.zeros(0), 0.5) == 0.0
    _, t = _tau_c0()
    tn = t.cpu().numpy()
    for mfd in (0.3, 0.7, 0.95):
        got = GF.mean_flux(t, mfd, 1e-10)
        ref = FS.mean_flux_scale(tn, mfd, 1e-10)
        assert abs(got - ref) <= 1e-8 * ref
        assert abs(np.mean(np.exp(-got * tn)) - mfd) < 1e-8
    # the drop-in of _rescale_mean_flux, with a threshold
    got = SP._rescale_mean_flux(tn.ravel(), 0.7, tn.size, 1e-10, 5.0)
    assert abs(got - FS.mean_flux_scale(tn, 0.7, 1e-10, thresh=5.0)) <= 1e-8 * got
    with pytest.raises(TypeError):
        SP._rescale_mean_flux(tn.astype(np.float32), 0.7, tn.size, 1e-5, 1e30)
    with pytest.raises(AttributeError):
        SP._rescale_mean_flux([1.0], 0.7, 1, 1e-5, 1e30)


def test_flux_pdf_matches_numpy_histogram():
    tau = np.log(np.arange(1, 101, dtype=np.double))
    bins, hist = GF.flux_pdf(tau, 20)
    rb, rh = FS.flux_hist(tau, 20)
    np.testing.assert_array_equal(bins, rb)
    np.testing.assert_allclose(hist, rh, rtol=1e-12)
    _, t = _tau_c0()
    tn = t.cpu().numpy()
    for nb in (7, 20, 64):
        for mfd in (None, 0.8):
            b, h = GF.flux_pdf(tn, nb, mfd)
            s = 1.0 if mfd is None else GF.mean_flux(tn, mfd)
            rb, rh = FS.flux_hist(tn, nb, s)
            np.testing.assert_allclose(h, rh, rtol=1e-12, err_msg="nbins=%d" % nb)


def test_flux_power_matches_oracle():
    w, t = _tau_c0()
    tn = t.cpu().numpy()
    vmax = w.box * w.velfac
    for mfd in (None, 0.75):
        kf, p = GF.flux_power(t, vmax, mean_flux_desired=mfd)
        if mfd is None:
            rk, rp = FS.flux_power(tn, vmax)
        else:
            rk, rp = FS.flux_power(tn, vmax, scale=FS.mean_flux_scale(tn, mfd), mean_flux=mfd)
        np.testing.assert_allclose(kf, rk, rtol=1e-14)
        np.testing.assert_allclose(p, rp, rtol=1e-8, atol=1e-12 * rp.max())
    # the reference's window test (test_statistics.py:54-72), odd and even npix
    for bb in (200, 201):
        xx = np.linspace(0, 51, bb)
        inn = np.sin(2 * math.pi * xx) + 1.5
        ff = np.abs(np.fft.rfft(np.exp(-inn) - 1)) ** 2 / bb ** 2
        bins, power = GF.flux_power(np.vstack([inn] * 10), vmax=1., spec_res=0.01, window=True)
        power /= 12.5569
        wind = GF._window_function(bins[1:], R=0.01, dv=1 / bb)
        assert np.all(np.abs(power[1:] * wind ** 2 - ff[1:]) < 0.01 * ff[1:])
        assert power[0] < 1e-20


def test_res_corr_matches_scipy():
    tau = np.zeros((2, 50))
    tau[0, 25] = 2
    tau[1, 23] = 3
    tau2 = GF.res_corr(tau, 2, 8)
    for i in (0, 1):
        assert abs(np.sum(tau2[i]) / np.sum(tau[i]) - 1) < 1e-6
        assert np.size(np.where(tau2[i] > 0)) == 15
    np.testing.assert_allclose(tau2, FS.res_corr(tau, 2, 8), rtol=1e-12, atol=1e-300)
    _, t = _tau_c0()
    for dv, fwhm in ((2.0, 8.0), (0.5, 30.0), (10.0, 8.0)):
        got = GF.res_corr(t, dv, fwhm)
        assert isinstance(got, torch.Tensor)
        np.testing.assert_allclose(got.cpu().numpy(), FS.res_corr(t.cpu().numpy(), dv, fwhm), rtol=1e-12,
                                   atol=1e-14 * float(t.max()))
