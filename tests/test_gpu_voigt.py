"""The deposit kernel's FP32 Voigt profile against scipy's Faddeeva function.

H(a, u) = Re w(u + ia) replaces the reference's profile() (singleabs.h:56-61,
Faddeeva::w with relerr DBL_EPSILON).  The FP32 evaluator (voigt.h) is
checked point by point over the damping range it serves, 0 <= a < 0.04
(larger a take the FP64 path), and over u from the core to the far wing.
It takes u in FP32, so the points are rounded to FP32 and scipy is
evaluated at the same rounded points.  The bound is 5e-7 of H for the
expansions plus 1.5e-7 u^2 for the Gaussian term, whose FP32 exponent
-(u^2 - a^2) log2(e) carries ~2 ulp: exp turns an absolute exponent error e
into a relative error e ln2 (the deposit's error budget leaves 1e-5 relative
on tau).
"""
import numpy as np
import pytest
from scipy.special import wofz

from fake_spectra_amd import spectra_priv as SP

pytestmark = pytest.mark.gpu

U = np.concatenate([np.linspace(0.0, 12.0, 48001), np.geomspace(12.0, 5e4, 4000), -np.linspace(0.0, 9.0, 901)])


@pytest.mark.parametrize("a", [0.0, 1e-5, 1e-4, 5e-4, 1e-3, 3e-3, 1e-2, 0.02, 0.0399])
def test_profile_matches_wofz(a):
    u = U.astype(np.float32).astype(np.float64)
    got = SP.voigt_profile(u, a).cpu().numpy().astype(np.float64)
    ref = wofz(u + 1j * a).real if a > 0 else np.exp(-u * u)
    m = ref > 1e-300
    rel = np.abs(got[m] - ref[m]) / ref[m]
    bound = 5e-7 + 1.5e-7 * u[m] ** 2
    assert np.all(np.isfinite(got))
    # a = 0: H is exp(-u^2) alone, zero in FP32 beyond u ~ 10
    if a == 0.0:
        keep = ref > 1e-37
        assert np.all(rel[keep[m]] <= bound[keep[m]]), np.max(rel[keep[m]] / bound[keep[m]])
        assert np.all(got[~keep] <= 1e-37)
    else:
        worst = np.argmax(rel / bound)
        assert np.all(rel <= bound), (rel[worst], u[m][worst])


def test_profile_rejects_large_a():
    with pytest.raises(ValueError):
        SP.voigt_profile(np.zeros(4), 0.05)
