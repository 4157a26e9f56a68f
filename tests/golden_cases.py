"""Loaders for the golden fixtures written by tests/golden/make_golden.py."""
import os

import numpy as np

from conftest import GOLDEN
from fake_spectra_amd import synth
from fake_spectra_amd.lines import HI_1215, CIV_1548

LINES = {"HI1215": HI_1215, "CIV1548": CIV_1548}


def c0_fixture():
    z = np.load(os.path.join(GOLDEN, "ref_c0.npz"))
    w = synth.config(0)
    cs = np.array([w.pos.astype(np.float64).sum(), w.hh.astype(np.float64).sum(),
                   w.temp.astype(np.float64).sum(), w.vel.astype(np.float64).sum(),
                   w.dens["H1"].astype(np.float64).sum(), w.cofm.sum()])
    assert np.array_equal(cs, z["checksums"]), "synthetic generator no longer reproduces the fixture inputs"
    return w, z


def crafted():
    """[(case name, inputs dict)] and the fixture archive."""
    z = np.load(os.path.join(GOLDEN, "ref_cases.npz"))
    cases = {}
    for k in z.files:
        if "_in_" in k:
            name, field = k.split("_in_")
            cases.setdefault(name, {})[field] = z[k]
    return cases, z


def tau_keys(z):
    """(key, case, kernel, line, gamma scale, nbins) for each tau fixture."""
    out = []
    for k in z.files:
        if k.endswith("_tau"):
            case, kk, ln, g, n, _ = k.split("_")
            out.append((k, case, int(kk[1:]), LINES[ln], float(g[1:]), int(n[1:])))
    return out
