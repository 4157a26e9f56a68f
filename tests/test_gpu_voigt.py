"""The deposit kernel's FP32 Voigt profile against scipy's Faddeeva function.

H(a, u) = Re w(u + ia) replaces the reference's profile() (singleabs.h:56-61,
Faddeeva::w with relerr DBL_EPSILON).  The FP32 evaluator (voigt.h) is
checked point by point over the damping range it serves, 0 <= a < 0.04
(larger a take the FP64 path), and over u from the core to the far wing,
to 5e-7 of H (the deposit's error budget leaves 1e-5 relative on tau).
"""
import numpy as np
import pytest
from scipy.special import wofz

from fake_spectra_amd import spectra_priv as SP

pytestmark = pytest.mark.gpu

U = np.concatenate([np.linspace(0.0, 12.0, 48001), np.geomspace(12.0, 5e4, 4000), -np.linspace(0.0, 9.0, 901)])


@pytest.mark.parametrize("a", [0.0, 1e-5, 1e-4, 5e-4, 1e-3, 3e-3, 1e-2, 0.02, 0.0399])
def test_profile_matches_wofz(a):
    got = SP.voigt_profile(U, a).cpu().numpy().astype(np.float64)
    ref = wofz(U + 1j * a).real if a > 0 else np.exp(-U * U)
    m = ref > 1e-300
    rel = np.abs(got[m] - ref[m]) / ref[m]
    assert np.all(np.isfinite(got))
    # a = 0: H is exp(-u^2) alone, zero in FP32 beyond u ~ 10
    if a == 0.0:
        keep = ref > 1e-37
        assert np.max(rel[keep[m]]) <= 5e-7, np.max(rel[keep[m]])
        assert np.all(got[~keep] <= 1e-37)
    else:
        assert np.max(rel) <= 5e-7, (np.max(rel), U[m][np.argmax(rel)])


def test_profile_rejects_large_a():
    with pytest.raises(ValueError):
        SP.voigt_profile(np.zeros(4), 0.05)
