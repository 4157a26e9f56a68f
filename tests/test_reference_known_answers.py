"""The reference's own C++ unit tests (fake_spectra/test.cpp, Boost.Test,
known answers from Mathematica), restated and run against the FP64 oracle
and, where it is built, the compiled reference itself."""
import math

import numpy as np
import pytest


def near(x, y, rel=1e-5):
    """FLOATS_NEAR_TO (test.cpp:29-31)"""
    assert np.isfinite(x) and np.isfinite(y)
    assert abs(x - y) <= max(abs(x), abs(y)) * rel, (x, y)


def approx(x, y):
    """FLOATS_APPROX_NEAR_TO (test.cpp:33-35)"""
    near(x, y, 1e-2)


NORM = 32.0 / 4 / math.pi
CORR = 3 / (4 * math.pi)


def test_sph_kern(impl):                 # test.cpp:37-46
    assert impl.sph_cubic_kernel(1) == 0
    near(impl.sph_cubic_kernel(0), NORM)
    near(impl.sph_cubic_kernel(0.5), 0.25 * NORM)
    near(impl.sph_cubic_kernel(0.25), 0.71875 * NORM)
    near(impl.sph_cubic_kernel(0.75), 0.03125 * NORM)


def test_sph_kern_frac(impl):            # test.cpp:48-79
    kf = lambda *a: impl.kern_frac(1, *a)
    near(kf(-1, 1, 1, 0, 1), 3 * 2 / math.pi)
    near(kf(-3, 10, 1, 0, 1), 3 * 2 / math.pi)
    approx(kf(-0.15, -0.1, 1, 0, 1), CORR * 0.489167)
    approx(kf(0.05, 0.1, 1, 0.4, 0.774597), CORR * 0.051004)
    approx(kf(0.15, 0.16, 1, 0.8, 0.447214), CORR * 0.000167423)
    approx(kf(-0.05, 0.1, 1, 0.9, 0.316228), CORR * 0.00040101)
    approx(kf(0.3, 1, 1, 0.3, 0.83666), CORR * 0.177801)
    assert kf(1.5, 2, 1, 0, 1) == 0


VELFAC = 414.50523718485636 / 1e3 * 0.2 / 0.71
LYA = (1215.6701e-10, 6.265e8, 0.416400, 1.00794)


def test_compute_colden(impl):           # test.cpp:83-151
    n = 2000
    col = np.zeros(n)
    args = LYA + (VELFAC, 10000, 1, 1, 1e-5)

    def add(dr2, dens, pos, smooth):
        impl.add_colden_particle(*args, col, dr2, dens, pos, smooth)

    nonzero = set()
    add(0, 1, 5002.5, 1)
    total = 8 * 3. / (4 * math.pi)
    assert col[999] == 0 and col[1001] == 0
    near(col[1000], total)
    nonzero.add(1000)
    total /= 4
    add(0, 1 / 8., 1002.5, 2)
    assert col[199] == 0 and col[201] == 0
    near(col[200], total)
    nonzero.add(200)
    add(0, 10 / 8., 1012.5, 2)
    assert col[201] == 0 and col[203] == 0
    near(col[202], 10 * total)
    nonzero.add(202)
    add(0, 1 / 8., 1030, 2)
    assert col[204] == 0 and col[207] == 0
    near(col[205], total / 2.)
    near(col[206], total / 2.)
    nonzero |= {205, 206}
    add(0, 1 / 8., 10000, 2)
    assert col[1998] == 0 and col[1] == 0
    near(col[1999], total / 2.)
    near(col[0], total / 2.)
    nonzero |= {1999, 0}
    add(0.7, 1, 4852.5, 1)
    assert abs(col[969]) < 1e-300
    approx(col[970], 0.0451531 * 3 / (4 * math.pi))
    assert col[971] == 0
    nonzero.add(970)
    add(1.0, 1, 4802.5, 1)
    assert abs(col[960]) < 1e-300
    for i in range(n):
        if i not in nonzero:
            assert col[i] + i == i


def _sightlines():                      # test.cpp:164-194
    cofm, axis = [], []
    for a, x, y, z in [(1, 4000, 4000, 4000), (1, 4000, 4000, 4000), (1, 4000, 4020, 4010), (1, 4000, 4000, 4010),
                       (1, 0, 0.4, 0.1), (1, 0, 10000 - 0.4, 0.3), (1, 0, 2000, 1000), (3, 1000, 2000, 500),
                       (2, 1000, 2000, 500), (1, 1000, 2000, 500), (3, 3000, 5000, 550), (3, 8000, 5500, 9000),
                       (1, 6000, 5500, 9000)]:
        cofm.append((x, y, z))
        axis.append(a)
    return np.array(axis, np.int32), np.array(cofm, np.float64)


def test_index_table(impl):             # test.cpp:164-282
    axis, cofm = _sightlines()
    box = 10000
    g = lambda p, h: impl.get_near_lines(box, axis, cofm, np.array(p, np.float32), h)
    nb = g([500, 2000, 1000], 1)
    assert nb == {6: 0.0}
    nb = g([5000, 2000, 1000], 1)
    assert nb == {6: 0.0}
    nb = g([5010, 2010, 990], 20)
    assert nb == {6: 10 * 10 + 10 * 10.}
    assert g([5010, 2010, 990], 10) == {}
    assert g([4000, 4000, 4000], 1) == {0: 0.0, 1: 0.0}
    nb = g([4000, 4000, 4000], 25)
    assert nb == {0: 0.0, 1: 0.0, 2: 20 * 20 + 10 * 10., 3: 10 * 10.}
    nb = g([1000, 9999.9, 9999.9], 0.6)
    assert sorted(nb) == [4, 5]
    approx(nb[4], 0.5 * 0.5 + 0.2 * 0.2)
    approx(nb[5], 0.3 * 0.3 + 0.4 * 0.4)
    nb = g([1000.5, 2000, 501], 1.5)
    assert sorted(nb) == [7, 8, 9]
    approx(nb[7], 0.25)
    approx(nb[8], 0.25 + 1)
    approx(nb[9], 1)
    poses = np.array([500, 2000, 1000, 5000, 2000.0, 1000.0, 5010.0, 2010, 990, 4000, 4000, 4000, 1000, 9999.9,
                      9999.9, 1000.5, 2000, 501, 7500, 7500, 7500, 4008, 4008.0, 4008.0, 2000.0, 9999, 9999.8],
                     np.float32).reshape(9, 3)
    hh = np.array([1, 1, 20, 25, 0.6, 1.5, 7, 10, 0.8], np.float32)
    off, parts, dr2 = impl.near_particles(box, poses, hh, axis, cofm)
    sizes = np.diff(off)
    assert list(sizes) == [1, 1, 1, 2, 1, 2, 3, 1, 1, 1, 0, 0, 0]
    row = lambda i: list(parts[off[i]:off[i + 1]])
    assert row(0)[0] == 3
    assert 3 in row(3)
    assert 4 in row(5) and 8 in row(5)
    assert row(6) == [0, 1, 2]
    assert row(8)[0] == 5


def test_profile(impl):                 # test.cpp:284-302
    p = impl.profile
    near(p(0, 0), 1)
    near(p(0, 0.1), 0.896457)
    near(p(0.1, 1e-6), 0.990048)
    near(p(0.1, 1e-4), 0.989939)
    near(p(15, 1e-6), 2.52441e-9)
    near(p(15, 1e-4), 2.52441e-7)
    near(p(20, 1e-7), 1.4158e-10)
    near(p(1, 1e-4), 0.367888)
    near(p(1.5, 1e-6), 0.1054)
    near(p(2, 1e-7), 0.0183157)
    near(p(1, 1e-3), 0.367965)


def test_single_absorber(impl):         # test.cpp:304-332
    bb = 0.128557 * math.sqrt(2e4 / 1)
    t = lambda vdr2, vsm, aa, lo, hi, b=bb: impl.tau_kern_outer(b, vdr2, vsm, aa, 1, lo, hi)
    near(t(0, 10, 1e-4, 0, 0), CORR * 78.0409)
    near(t(0, 10, 1e-4, 5, 5), CORR * 72.6216)
    near(t(0, 10, 1e-4, 10, 10), CORR * 58.5185)
    near(t(0, 10, 1e-4, 20, 20), CORR * 24.6696)
    near(t(0, 2, 1e-4, 5, 5), CORR * 14.8203)
    approx(t(25, 10, 1e-4, 0, 0), CORR * 18.218)
    approx(t(25, 10, 1e-4, 5, 5), CORR * 16.9436)
    bc = 0.128557 * math.sqrt(2e4 / 16)
    approx(t(0, 5, 1e-6, 0, 10, bc), CORR * 16.0403)
    approx(t(0, 5, 1e-6, -5, 5, bc), CORR * 27.1978)


BOLTZMANN = 1.3806504e-16
PROTONMASS = 1.67262178e-24


def test_add_tau(impl):                 # test.cpp:337-381
    n = 2000
    args = LYA + (VELFAC, 10000, 0.25, 1, 1e-5)
    temp = 2e4
    bb = math.sqrt(2.0 * BOLTZMANN / PROTONMASS) / 1e5 * math.sqrt(temp / 1.00794)
    smooth = 3
    amp = 7.57973e-15
    voigt_fac = 1215.6701e-10 * 6.265e8 / (4. * math.pi) / 1e5
    rscale = 3.085678e21 * 0.25 / 0.7
    vbin = lambda b, pos: 10000 * VELFAC / n * b - VELFAC * pos
    SA = lambda b, pos, dens: amp * dens / bb * impl.tau_kern_outer(bb, 0, VELFAC * smooth, voigt_fac / bb, 1,
                                                                    vbin(b, pos), vbin(b + 1, pos)) / VELFAC
    tau = np.zeros(n)
    d = float(np.float32(1e-3 * rscale))
    impl.add_tau_particle(*args, tau, 0, d, 5002.5, 0, temp, smooth)
    near(tau[999], SA(999, 5002.5, d))
    near(tau[1000], SA(1000, 5002.5, d))
    near(tau[1001], tau[999])
    near(tau[400], SA(400, 5002.5, d))
    near(tau[401], SA(401, 5002.5, d))
    near(tau[402], SA(402, 5002.5, d))
    tau[:] = 0
    pecvel = float(np.float32(10000 / n * 3 * VELFAC))
    impl.add_tau_particle(*args, tau, 0, d, 5002.5, pecvel, temp, smooth)
    near(tau[1002], SA(999, 5002.5, d))
    near(tau[1003], SA(1000, 5002.5, d))
    near(tau[1004], tau[1002])
    tau[:] = 0
    impl.add_tau_particle(*args, tau, 1, 1, 5002.5, 0, 10000, 1)
    assert tau[1000] == 0 and tau[1501] == 0


def test_tau_colden_consistency(impl):  # test.cpp:383-404
    n = 2000
    args = LYA + (VELFAC, 10000, 0.25, 1, 1e-5)
    rscale = 3.085678e21 * 0.25 / 0.7
    d = float(np.float32(0.1 * rscale))
    tau = np.zeros(n)
    col = np.zeros(n)
    impl.add_tau_particle(*args, tau, 0, d, 5002.5, 0, 2e4, 3)
    impl.add_colden_particle(*args, col, 0, d, 5002.5, 3)
    near(tau.sum() / col.sum(), 7.57973e-15 / VELFAC / 2.81809)
