"""The caller's unit pipeline (spectra.py:210-241, gas_properties.py:105-110,
spectra.py:576-615) against textbook physics rather than the same formulas:
the Hubble flow across a comoving kpc/h, and the critical density 3 H0^2 / (8 pi G)."""
import math

import numpy as np

from fake_spectra_amd import synth
from fake_spectra_amd.lines import PROTONMASS

G_CGS = 6.674e-8                 # cm^3 g^-1 s^-2
MPC_CM = 3.085678e24


def test_velfac_is_the_hubble_flow_across_a_comoving_kpc():
    for z, h, om in ((3.0, 0.7, 0.3), (2.0, 0.6774, 0.3089), (0.0, 0.7, 0.3)):
        c = synth.Cosmology(redshift=z, hubble=h, omega_m=om, omega_l=1 - om)
        a = 1 / (1 + z)
        hz = 100 * h * math.sqrt(om * (1 + z) ** 3 + (1 - om))     # km/s/Mpc
        # one comoving kpc/h is a/h physical kpc = 1e-3 a/h Mpc
        assert abs(c.velfac() / (hz * 1e-3 * a / h) - 1) < 1e-7
    # 60 Mpc/h at z = 3, h = 0.7, Omega_m = 0.3: 0.111524 km/s per kpc/h
    assert abs(synth.Cosmology().velfac() - 0.111524) < 1e-6


def _rho_crit(hubble):
    """3 H0^2 / (8 pi G) in g/cm^3."""
    h0 = 100 * hubble * 1e5 / MPC_CM                              # 1/s
    return 3 * h0 ** 2 / (8 * math.pi * G_CGS)


def test_critical_density_known_answers():
    """The reference's own known answers (tests/test_spectra.py:93-97)."""
    assert abs(_rho_crit(0.7) / 9.204285430050004e-30 - 1) < 1e-6
    assert abs(_rho_crit(1.0) / 1.8784255979693885e-29 - 1) < 1e-6


def test_gas_density_is_omega_b_times_the_critical_density():
    c = synth.Cosmology()
    rho_crit = _rho_crit(c.hubble)                                # g/cm^3 today
    n_mean = c.omega_b * rho_crit * (1 + c.redshift) ** 3 / PROTONMASS
    w = synth.config(0, scale=1 / 64)
    # the field's own mean: den = rho * density_conv with rho = Omega_b rho_crit (1 + delta)
    rho_bar = c.omega_b * synth.RHO_CRIT_INTERNAL
    got = rho_bar * c.density_conv()
    assert abs(got / n_mean - 1) < 5e-4          # 27.755 vs 27.754 (G to 4 digits)
    # and H I density per particle carries rscale, X_H, x_HI and 1/m_H (spectra.py:593-615)
    d = w.dens_for(w.lines[0]).astype(np.float64)
    expect = (n_mean * w.overdensity * c.rscale() * synth.HYDROGEN_MASS_FRAC * w.neutral_fraction
              / 1.00794)
    assert np.max(np.abs(d / expect - 1)) < 6e-4

