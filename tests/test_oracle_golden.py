"""The FP64 oracle against the reference's outputs committed under
tests/golden (produced by the compiled reference C++)."""
import numpy as np
import pytest

from golden_cases import c0_fixture, crafted, tau_keys
from fake_spectra_amd.lines import HI_1215


def test_c0_subset(oracle):
    w, z = c0_fixture()
    sel = z["sel"]
    args = list(w.interp_args(HI_1215, True))
    args[16], args[17] = w.axis[sel], w.cofm[sel]
    tau = oracle.interpolate(*args)
    np.testing.assert_allclose(tau, z["tau"], rtol=1e-12, atol=0)
    cargs = list(w.interp_args(None, False))
    cargs[16], cargs[17] = w.axis[sel], w.cofm[sel]
    np.testing.assert_allclose(oracle.interpolate(*cargs), z["colden"], rtol=1e-12, atol=0)
    off, parts, dr2 = oracle.near_particles(w.box, w.pos, w.hh, w.axis[sel], w.cofm[sel])
    assert np.array_equal(off, z["offsets"]) and np.array_equal(parts, z["parts"])
    assert np.array_equal(dr2, z["dr2"])
    assert np.array_equal(oracle.near_lines(w.box, w.pos, w.hh, w.axis, w.cofm), z["near_lines"])


def test_crafted_cases(oracle):
    cases, z = crafted()
    velfac = float(z["velfac"])
    for key, case, kern, line, gs, nb in tau_keys(z):
        c = cases[case]
        tau = oracle.interpolate(1, nb, kern, float(c["box"]), velfac, 0.25, line.lambda_cm, line.gamma * gs,
                                 line.fosc, line.amumass, 1e-7, c["pos"], c["vel"], c["dens"], c["temp"], c["hh"],
                                 c["axis"], c["cofm"])
        np.testing.assert_allclose(tau, z[key], rtol=1e-11, atol=1e-300, err_msg=key)
    for case, c in cases.items():
        for kern in (0, 1, 3):
            col = oracle.interpolate(0, 257, kern, float(c["box"]), velfac, 0.25, HI_1215.lambda_cm, HI_1215.gamma,
                                     HI_1215.fosc, HI_1215.amumass, 1e-7, c["pos"], np.zeros(1, np.float32),
                                     c["dens"], np.zeros(1, np.float32), c["hh"], c["axis"], c["cofm"])
            np.testing.assert_allclose(col, z["%s_k%d_colden" % (case, kern)], rtol=1e-12, atol=0)
        off, parts, dr2 = oracle.near_particles(float(c["box"]), c["pos"], c["hh"], c["axis"], c["cofm"])
        assert np.array_equal(off, z[case + "_offsets"])
        assert np.array_equal(parts, z[case + "_parts"])
        assert np.array_equal(dr2, z[case + "_dr2"])
        assert np.array_equal(oracle.near_lines(float(c["box"]), c["pos"], c["hh"], c["axis"], c["cofm"]),
                              z[case + "_near"])


def test_additivity_over_particle_subsets(oracle):
    """tau is a sum of independent per-pair deposits (part_int.cpp:35-47), so
    splitting the particles in two sums to the whole (spectra.py:819-824)."""
    cases, _ = crafted()
    c = cases["clump"]
    n = c["pos"].shape[0]
    idx = np.arange(n)
    a, b = idx % 3 == 0, idx % 3 != 0

    def run(m):
        return oracle.interpolate(1, 400, 1, float(c["box"]), 0.11, 0.25, HI_1215.lambda_cm, HI_1215.gamma,
                                  HI_1215.fosc, HI_1215.amumass, 1e-7, c["pos"][m], c["vel"][m], c["dens"][m],
                                  c["temp"][m], c["hh"][m], c["axis"], c["cofm"])
    np.testing.assert_allclose(run(a) + run(b), run(idx >= 0), rtol=1e-13, atol=0)


def test_whole_path_matches_reference(oracle, ref):
    """Whole c0 workload: the restatement against the compiled reference."""
    w, _ = c0_fixture()
    args = w.interp_args(HI_1215, True)
    np.testing.assert_allclose(oracle.interpolate(*args), ref.interpolate(*args), rtol=1e-12, atol=0)
