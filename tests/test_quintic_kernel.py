"""The quintic SPH kernel (singleabs.h:31-42, absorption.cpp:76-97) in the
oracle and, where built, the compiled reference: pinned by closed forms that
a dropped term, a wrong coefficient or a wrong break point would fail."""
import math

import numpy as np
from scipy import integrate


def test_quintic_normalisation(impl):
    """4 pi int_0^1 K(q) q^2 dq = 1 (the normalisation of absorption.cpp:35-46)."""
    val, _ = integrate.quad(lambda q: 4 * math.pi * impl.sph_quintic_kernel(q) * q * q, 0, 1,
                            points=[1 / 3, 2 / 3], epsabs=1e-14, epsrel=1e-13)
    assert abs(val - 1) < 1e-12


def test_quintic_values_and_continuity(impl):
    norm = 9 / (40 * math.pi)
    k = impl.sph_quintic_kernel
    assert abs(k(0) - 66 * norm) < 1e-13          # 3^5 - 6 2^5 + 15
    assert k(1) == 0 and k(1.5) == 0
    assert abs(k(2 / 3) - norm) < 1e-12            # only (3 - 3q)^5 = 1 is left
    for qb in (1 / 3, 2 / 3):
        lo, hi = k(qb - 1e-9), k(qb + 1e-9)
        assert abs(lo - hi) < 1e-6 * abs(lo)
    # first derivative vanishes at q = 0 and the kernel decreases monotonically
    q = np.linspace(0, 1, 2001)
    v = np.array([k(x) for x in q])
    assert np.all(np.diff(v) <= 1e-15)
    assert abs(k(1e-4) - k(0)) < 1e-6 * k(0)


def test_quintic_kern_frac_is_the_line_integral(impl):
    """kern_frac over the whole support is the 8-interval trapezoid of K along
    the chord; with a fine check against the exact chord integral."""
    for dr2 in (0.0, 0.1, 0.5):
        zr = math.sqrt(1 - dr2)
        got = impl.kern_frac(3, -zr, zr, 1.0, dr2, zr)
        dz = 2 * zr / 8
        zs = -zr + dz * np.arange(9)
        vals = np.array([impl.sph_quintic_kernel(math.sqrt(dr2 + z * z)) for z in zs])
        trap = dz * (vals[0] / 2 + vals[1:-1].sum() + vals[-1] / 2)
        assert abs(got - trap) < 1e-13 * max(1.0, abs(trap))
        exact, _ = integrate.quad(lambda z: impl.sph_quintic_kernel(math.sqrt(dr2 + z * z)), -zr, zr, limit=200)
        assert abs(got - exact) < 0.05 * exact     # the 8-point rule is coarse, not wrong


def test_oracle_quintic_matches_reference(oracle, ref):
    q = np.linspace(0, 1.1, 1001)
    a = np.array([oracle.sph_quintic_kernel(x) for x in q])
    b = np.array([ref.sph_quintic_kernel(x) for x in q])
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
