"""Helpers for the GPU parity tests: device copies and the stated tolerances."""
import numpy as np
import torch

from fake_spectra_amd import spectra_priv as SP

# Tolerances against the FP64 reference, per output.
TAU_REL = 1e-5        # relative tau where tau > 1e-3
TAU_FLOOR = 1e-3
FLUX_ABS = 1e-6       # absolute error of exp(-tau)
COLDEN_REL = 1e-6     # relative column density


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def assert_tau_close(got, ref, what=""):
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape
    assert np.all(np.isfinite(got)), what
    m = ref > TAU_FLOOR
    rel = np.max(np.abs(got[m] - ref[m]) / ref[m]) if m.any() else 0.0
    flux = np.max(np.abs(np.exp(-got) - np.exp(-ref))) if ref.size else 0.0
    assert rel <= TAU_REL, "%s: max relative tau error %.3g" % (what, rel)
    assert flux <= FLUX_ABS, "%s: max flux error %.3g" % (what, flux)
    # pixels the reference leaves at exactly zero stay zero
    assert np.all(got[ref == 0] == 0), what
    return rel, flux


def assert_colden_close(got, ref, what=""):
    got = np.asarray(got)
    ref = np.asarray(ref)
    scale = np.maximum(np.abs(ref), 1e-300)
    rel = np.max(np.abs(got - ref) / scale) if ref.size else 0.0
    assert rel <= COLDEN_REL, "%s: max relative colden error %.3g" % (what, rel)
    assert np.all(got[ref == 0] == 0), what
    return rel


def plan_for(w, axis=None, cofm=None):
    axis = w.axis if axis is None else axis
    cofm = w.cofm if cofm is None else cofm
    return SP.Plan(w.box, dev(w.pos), dev(w.hh), dev(axis), dev(cofm))
