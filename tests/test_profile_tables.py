"""The generated constant tables of the device profile (CPU checks).

fake_spectra_amd/csrc/dawson_table.h holds the pieces of the FP32 evaluator's
G(u) = 1 - 2u D(u) (D: Dawson's integral) and its wing row, written by
gen_dawson_table.py; wing_table.h holds the exact sub-sample moments and
binomials of the far-wing expansion, written by gen_wing_table.py.  Each
header is compared with a fresh run of its generator (no hand edits), and
its numbers with values computed independently here: G from scipy's
dawsn, the moments summed directly over the trapezoid points.  The
tolerances sit just above the FP32 rounding of the stored coefficients, so
a wrong row order, row variable or dropped coefficient fails by orders of
magnitude.
"""
import os
import re
import subprocess
import sys
from math import comb

import numpy as np
import pytest
from scipy.special import dawsn

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fake_spectra_amd", "csrc")


def _read(name):
    with open(os.path.join(CSRC, name)) as f:
        return f.read()


def _define(text, name):
    m = re.search(r"#define %s\s+([-0-9.e+]+)f?\b" % name, text)
    assert m, name
    return float(m.group(1))


def _init(text, name):
    m = re.search(r"#define %s \{(.*?)\}" % name, text, re.S)
    assert m, name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S).replace("\\", " ")
    return np.array([float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()])


def _g64(u):
    return 1.0 - 2.0 * u * dawsn(u)


@pytest.mark.parametrize("gen,header", [("gen_dawson_table.py", "dawson_table.h"),
                                        ("gen_wing_table.py", "wing_table.h")])
def test_header_is_generator_output(gen, header):
    out = subprocess.run([sys.executable, os.path.join(CSRC, gen)], capture_output=True, text=True, check=True)
    assert out.stdout == _read(header)


def _g_table():
    t = _read("dawson_table.h")
    npiece = int(_define(t, "FSAMD_G_NPIECE"))
    deg = int(_define(t, "FSAMD_G_DEG"))
    stride = int(_define(t, "FSAMD_G_STRIDE"))
    nrow = int(_define(t, "FSAMD_G_NROW"))
    width = _define(t, "FSAMD_G_WIDTH")
    umax = _define(t, "FSAMD_G_UMAX")
    coef = _init(t, "FSAMD_G_COEF_INIT").astype(np.float32).astype(np.float64)
    assert nrow == npiece + 1 and coef.size == nrow * stride
    assert stride % 4 == 0 and stride >= deg + 1          # two 16-byte reads per row
    assert np.all(coef.reshape(nrow, stride)[:, deg + 1:] == 0.0)
    return coef.reshape(nrow, stride)[:, :deg + 1], npiece, width, umax, t


def _horner(c, v):
    r = np.full_like(v, c[-1])
    for k in range(len(c) - 2, -1, -1):
        r = r * v + c[k]
    return r


def test_core_pieces_follow_dawson():
    rows, npiece, width, umax, _ = _g_table()
    assert npiece * width == umax
    s = np.linspace(0.0, 1.0, 257)[:-1]
    worst = 0.0
    for p in range(npiece):
        u = (p + s) * width
        g = _g64(u)
        err = np.abs(_horner(rows[p], s) - g) / np.maximum(np.abs(g), 1e-2)
        worst = max(worst, float(err.max()))
    # float64 fit error 2e-9 plus the FP32 rounding of the coefficients
    # (measured 2.4e-7 of max(|G|, 0.01), where G crosses zero near u = 0.92)
    assert worst < 3e-7, worst


def test_wing_row_is_g_as_polynomial_in_x():
    rows, npiece, width, umax, t = _g_table()
    wing = rows[npiece]
    assert wing[0] == 0.0 and wing[1] == -0.5      # G ~ -x/2 - 3x^2/4 - ... (asymptotic series)
    u = np.geomspace(umax, 1e3, 2000)
    x = 1.0 / (u * u)
    g = _g64(u)
    rel = np.abs(_horner(wing, x) - g) / np.abs(g)
    assert rel.max() < 1.5e-7, rel.max()
    # continuity at the switch from the last core piece to the wing row
    last = _horner(rows[npiece - 1], np.array([1.0]))[0]
    assert abs(last - _horner(wing, np.array([1.0 / umax ** 2]))[0]) < 2e-7 * abs(last)


def test_wing_moments_and_binomials():
    t = _read("wing_table.h")
    nterms = int(_define(t, "FSAMD_WING_NTERMS"))
    tmax = int(_define(t, "FSAMD_WING_TMAX"))
    T = _init(t, "FSAMD_WING_T_INIT").reshape(tmax + 1, nterms)
    assert np.all(T[0] == 0.0)
    for m in list(range(1, 20)) + [31, 64, 100, 127, 200, tmax]:
        y = 2.0 * np.arange(m + 1) / m - 1.0
        w = np.ones(m + 1)
        w[0] = w[-1] = 0.5
        direct = np.array([np.sum(w * y ** (2 * n)) / m for n in range(nterms)])
        np.testing.assert_allclose(T[m], direct, rtol=1e-12, atol=0)
    even = _init(t, "FSAMD_BINOM_EVEN_INIT").reshape(nterms, nterms)
    odd = _init(t, "FSAMD_BINOM_ODD_INIT").reshape(nterms, nterms)
    for n in range(nterms):
        for k in range(nterms):
            assert even[n, k] == comb(2 * n, 2 * k)
            assert odd[n, k] == comb(2 * n + 1, 2 * k)
