"""The flux-statistics oracle (oracle/fluxstats.py) against the reference's
own tests of the same functions (fake_spectra/tests/test_statistics.py:8-72,
test_spectra.py testrescorr), restated."""
import math

import numpy as np

import fluxstats as FS


def test_mean_flux_known_answers():          # test_statistics.py:8-19
    tol = 1e-4
    nn = np.arange(1, 101)
    tau = np.log(nn)
    for n in (1, 2, 0.5):
        mf = np.mean(nn ** (-float(n)))
        assert abs(FS.mean_flux_scale(tau, mf, tol) - n) < tol
    assert FS.mean_flux_scale(np.zeros(0), 0.5) == 0.0


def test_mean_flux_threshold_excludes_pixels():
    tau = np.concatenate([np.log(np.arange(1, 101)), [1e40, 1e40]])
    mf = np.mean(np.arange(1, 101) ** -2.0)
    assert abs(FS.mean_flux_scale(tau, mf, 1e-6, thresh=1e30) - 2) < 1e-5


def test_flux_pdf_known_answers():           # test_statistics.py:21-36
    tau = np.log(np.arange(1, 101, dtype=np.double))
    bins, hist = FS.flux_hist(tau, 20)
    assert bins[0] == 0. + 1 / 40. and bins[-1] == 1. - 1. / 40.
    expected = np.array([16., 2.2, 0.6, 0.2, 0.2, 0.2, 0.2, 0., 0., 0., 0.2, 0., 0., 0., 0., 0., 0., 0., 0., 0.2])
    assert abs(np.sum(expected) - np.sum(hist)) < 1e-3
    assert np.all(np.abs(hist[3:] - expected[3:]) < 1e-3)
    assert abs(hist[0] - expected[0]) < 1e-2


def test_flux_power_parseval():              # test_statistics.py:38-52
    xx = np.tile(np.arange(0, 1, 0.01) ** 2, (10, 1))
    n = xx.shape[1]
    # tau = xx, mean flux 1 -> dF = exp(-xx) - 1; flux_power sums over rows / nspec
    _, p = FS.flux_power(xx, vmax=1.0, mean_flux=1.0)
    d = np.exp(-xx[0]) - 1
    full = p[0] + 2 * p[1:(n + 1) // 2].sum() + (p[n // 2] if n % 2 == 0 else 0)
    assert abs(full - np.sum(d ** 2) / n) < 1e-12


def test_flux_power_window():                # test_statistics.py:54-72
    """The reference divides by the window and multiplies it back, so its
    assertion is: power / 12.5569 matches |rfft(exp(-inn) - 1)|^2 / n^2 to 1%
    (12.5569 ~ 1 / <F>^2 for this field), and the k = 0 mode vanishes."""
    for bb in (200, 201):
        xx = np.linspace(0, 51, bb)
        inn = np.sin(2 * math.pi * xx) + 1.5
        ff = np.abs(np.fft.rfft(np.exp(-inn) - 1)) ** 2 / bb ** 2
        taus = np.vstack([inn] * 10)
        kf, power = FS.flux_power(taus, vmax=1.0)
        power = power / 12.5569
        assert np.all(np.abs(power[1:] - ff[1:]) < 0.01 * ff[1:])
        assert power[0] < 1e-20
        assert abs(kf[1] - 2 * math.pi) < 1e-12


def test_res_corr_known_answers():           # test_spectra.py testrescorr
    tau = np.zeros((2, 50))
    tau[0, 25] = 2
    tau[1, 23] = 3
    tau2 = FS.res_corr(tau, 2, 8)
    for i in (0, 1):
        assert abs(np.sum(tau2[i]) / np.sum(tau[i]) - 1) < 1e-6
        assert np.size(np.where(tau2[i] > 0)) == 15


def test_observer_line_choice_each_branch():    # spectra.py:918-936
    m = np.array([
        # in band: the strongest of 0.5, 2.0 (5.0 is saturated)
        [0.5, 5.0, 0.05, 0.05, 1.0],
        [2.0, 7.0, 0.02, 0.05, 1.0],
        [5.0, 4.0, 0.08, 0.01, 0.3],
    ])
    # col 1: nothing in band, the weakest above 0.1 -> 4.0 (line 2)
    # col 2: nothing above 0.1, the strongest -> 0.08 (line 2)
    # col 3: ties at the maximum -> the first (line 0)
    # col 4: ties in band -> the first (line 0)
    assert list(FS.observer_lines(m)) == [1, 2, 2, 0, 0]
    # the band is open at both ends
    assert list(FS.observer_lines(np.array([[0.1], [3.0], [0.09]]))) == [1]
    assert list(FS.observer_lines(np.array([[0.1], [3.0], [2.99]]))) == [2]


def test_device_line_choice_matches_oracle():
    """products.observer_lines (torch, runs wherever the tensor lives) against
    the one-sightline-at-a-time restatement, on values full of ties and band
    edges."""
    import torch
    from fake_spectra_amd import products
    rng = np.random.default_rng(3)
    vals = np.array([0.02, 0.05, 0.1, 0.5, 2.0, 2.0, 3.0, 4.0, 9.0])
    for _ in range(300):
        m = rng.choice(vals, size=(int(rng.integers(1, 6)), 11))
        np.testing.assert_array_equal(products.observer_lines(torch.from_numpy(m)).numpy(), FS.observer_lines(m))
