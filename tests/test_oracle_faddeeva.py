"""Pins the oracle's FP64 Faddeeva function w(z) (and hence the profile
Re w(u + ia) of singleabs.h:56-61) against the reference's own 57-point
self-test table, scipy.special.wofz, and closed forms."""
import json
import math
import os

import numpy as np
import pytest
from scipy.special import erfcx, wofz, dawsn

from conftest import GOLDEN


def _rows():
    with open(os.path.join(GOLDEN, "faddeeva_w_table.json")) as f:
        return json.load(f)["rows"]


def test_reference_w_table_upper_half_plane(oracle):
    """Faddeeva.cpp:1919-2106: every table point with Im z >= 0 (the path's
    domain), to the table's 1e-13 relative tolerance (Faddeeva.cpp:2100)."""
    n = 0
    for zr, zi, wr, wi in _rows():
        if zi < 0 or not all(map(math.isfinite, (zr, zi, wr, wi))):
            continue
        w = oracle.w(zr, zi)
        ref = complex(wr, wi)
        err = abs(w - ref) / abs(ref)
        assert err < 1e-13, (zr, zi, w, ref)
        n += 1
    assert n >= 25


def test_profile_vs_scipy_grid(oracle):
    rng = np.random.default_rng(3)
    us = np.concatenate([np.linspace(0, 12, 97), rng.uniform(0, 1e4, 60), 10 ** rng.uniform(-3, 4, 60)])
    for a in [0, 1e-7, 1e-5, 4.7e-5, 1e-3, 4.7e-3, 0.03, 0.1, 0.3, 1.0, 3.0]:
        for u in us:
            ref = wofz(u + 1j * a).real
            got = oracle.profile(u, a)
            if ref == 0:
                assert got == 0
            else:
                assert abs(got - ref) <= 2e-13 * abs(ref), (u, a, got, ref)


def test_closed_forms(oracle):
    # a = 0: Re w(u) = exp(-u^2) exactly (Faddeeva.cpp:684-686 branch)
    for u in [0, 0.3, 1, 2.5, 5, 9, 26]:
        assert oracle.profile(u, 0) == pytest.approx(math.exp(-u * u), rel=1e-15, abs=0)
    # pure imaginary argument: w(iy) = erfcx(y)
    for y in [1e-6, 0.01, 0.5, 1, 1.9, 2.1, 7, 30]:
        assert oracle.profile(0, y) == pytest.approx(erfcx(y), rel=1e-13)
    # real axis imaginary part: Im w(x) = 2/sqrt(pi) D(x)
    for x in [0.05, 0.19, 0.21, 1, 3.3, 7.9, 8.1, 40]:
        assert oracle.w(x, 0).imag == pytest.approx(2 / math.sqrt(math.pi) * dawsn(x), rel=1e-13)
    # far wing: Re w ~ a / (sqrt(pi) u^2) (1 + 3/(2u^2))
    for u, a in [(300, 1e-4), (1e4, 4e-3)]:
        lead = a / (math.sqrt(math.pi) * u * u) * (1 + 1.5 / u ** 2)
        assert oracle.profile(u, a) == pytest.approx(lead, rel=1e-8)
    # symmetry w(-x + iy) = conj w(x + iy)
    for x, y in [(1.3, 0.01), (9, 0.2), (0.4, 3)]:
        assert oracle.w(-x, y) == pytest.approx(complex(oracle.w(x, y).real, -oracle.w(x, y).imag), rel=1e-15)
