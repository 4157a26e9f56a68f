"""GPU flux statistics (fake_spectra_amd.fluxstatistics, csrc/flux.hip)
against the numpy oracle and the reference's own test assertions."""
import math

import numpy as np
import pytest
import torch

import fluxstats as FS
from fake_spectra_amd import fluxstatistics as GF, spectra_priv as SP, synth
from gpu_util import TAU_REL, assert_tau_close, dev, plan_for

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


def test_observer_tau_matches_oracle_choice(oracle):
    """get_observer_tau (spectra.py:894-943) over five H I lines of different
    strengths (Lyman alpha with its oscillator strength scaled by 0.05, 0.004
    and 1e-4 besides the real 1215 and 1025 lines, so that different
    sightlines choose different lines) on the 32^3-particle field: the line chosen per
    sightline and its optical depth against the oracle's taus, smoothing and
    choice.  Sightlines whose smoothed maxima sit within the tau tolerance of a band edge
    or of each other could go either way and are left out of the index check."""
    from fake_spectra_amd import products
    from fake_spectra_amd.lines import HI_1215, HI_1025, Line
    w = synth.config(0)
    lines = [HI_1215, HI_1025] + [Line("scaled %g" % f, "H", 1, 1215.6701, 0.4164 * f, 6.265e8)
                                  for f in (0.05, 0.004, 1e-4)]
    dvbin = w.velfac * w.box / w.nbins
    plan = plan_for(w)
    got, choice = products.observer_tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), lines,
                                        dev(w.dens_for(HI_1215)), dvbin, 8.0)
    ref_taus = np.stack([oracle.interpolate(*w.interp_args(ln, True)) for ln in lines])
    maxt = FS.res_corr(ref_taus, dvbin, 8.0).max(axis=-1)
    ref_choice = FS.observer_lines(maxt)
    choice = choice.cpu().numpy()
    srt = np.sort(np.concatenate([maxt, np.full((1, maxt.shape[1]), 0.1), np.full((1, maxt.shape[1]), 3.0)]), axis=0)
    close = (np.diff(srt, axis=0) <= 2 * TAU_REL * np.maximum(srt[1:], 1e-30)).any(axis=0)
    assert close.mean() < 0.01
    np.testing.assert_array_equal(choice[~close], ref_choice[~close])
    assert len(np.unique(ref_choice)) >= 2, "the workload should exercise more than one line"
    keep = np.nonzero(~close)[0]
    assert_tau_close(got.cpu().numpy()[keep], ref_taus[ref_choice[keep], keep], "observer tau")
    with pytest.raises(ValueError):
        products.observer_tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), lines,
                              dev(w.dens_for(HI_1215)), dvbin, 0.0)


@pytest.mark.parametrize("npix,nlos", [(1, 3), (2, 2), (97, 5), (210, 7), (572, 4), (1024, 33), (4096, 3), (6000, 2)])
def test_flux_power_fft_lengths(npix, nlos):
    """The mixed-radix FFT over lengths with radices 2, 3, 4, 5, 7, larger
    primes (97, 11 and 13) and a lone sightline left over from the pairing;
    6000 is past the FFT's LDS buffers and takes the direct sum."""
    rng = np.random.default_rng(npix)
    tau = rng.exponential(0.5, size=(nlos, npix))
    kf, p = GF.flux_power(tau, 1000.0, mean_flux_desired=None)
    rk, rp = FS.flux_power(tau, 1000.0)
    np.testing.assert_allclose(kf, rk, rtol=1e-14)
    np.testing.assert_allclose(p, rp, rtol=1e-10, atol=1e-13 * rp.max())
