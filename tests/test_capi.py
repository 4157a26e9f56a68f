"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared symbol, and the drop-in shim rejects bad arguments with the
reference's exception classes (py_module.cpp:115-151, 33-54)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from fake_spectra_amd import _native, spectra_priv, synth
from fake_spectra_amd.lines import HI_1215


def header_symbols():
    src = open(os.path.join(ROOT, "include", "fake_spectra_amd.h")).read()
    return sorted(set(re.findall(r"\b(fsamd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_header():
    lib = _native.lib()
    syms = header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_native.EXPORTS)
    assert lib.fsamd_version() >= 100


def test_error_string_roundtrip():
    lib = _native.lib()
    plan = ctypes.c_void_p()
    rc = lib.fsamd_plan_create(-1.0, None, None, 0, None, None, 0, None, ctypes.byref(plan))
    assert rc == _native.FSAMD_ERR_ARG
    assert b"box" in lib.fsamd_last_error()
    with pytest.raises(ValueError):
        _native.check(rc)


def _args():
    w = synth.config(0, nlos=4, n_side=6)
    return list(w.interp_args(HI_1215, True))


def test_parse_errors_are_attribute_errors():
    a = _args()
    with pytest.raises(AttributeError):
        spectra_priv._Particle_Interpolate(*a[:-1])
    b = list(a)
    b[1] = 1024.0               # 'i' rejects floats
    with pytest.raises(AttributeError):
        spectra_priv._Particle_Interpolate(*b)
    b = list(a)
    b[11] = b[11].tolist()      # 'O!' needs ndarrays
    with pytest.raises(AttributeError):
        spectra_priv._Particle_Interpolate(*b)


def test_dtype_errors_are_type_errors():
    a = _args()
    for i, bad in [(11, np.float64), (12, np.float64), (13, np.float64), (14, np.float16), (15, np.float64)]:
        b = list(a)
        b[i] = b[i].astype(bad)
        with pytest.raises(TypeError):
            spectra_priv._Particle_Interpolate(*b)
    b = list(a)
    b[17] = b[17].astype(np.float32)
    with pytest.raises(TypeError):
        spectra_priv._Particle_Interpolate(*b)
    b = list(a)
    b[16] = b[16].astype(np.int64)
    with pytest.raises(TypeError):
        spectra_priv._Particle_Interpolate(*b)


def test_shape_errors_are_value_errors():
    a = _args()
    b = list(a)
    b[13] = b[13][:-1]
    with pytest.raises(ValueError):
        spectra_priv._Particle_Interpolate(*b)
    b = list(a)
    b[15] = b[15][:-1]
    with pytest.raises(ValueError):
        spectra_priv._Particle_Interpolate(*b)
    b = list(a)
    b[16] = b[16][:-1]
    with pytest.raises(ValueError):
        spectra_priv._Particle_Interpolate(*b)
    b = list(a)
    b[17] = np.ascontiguousarray(b[17][:, :2])
    with pytest.raises(ValueError):
        spectra_priv._Particle_Interpolate(*b)



def test_empty_inputs_give_zero_arrays():
    a = _args()
    b = list(a)
    for i in (11, 12):
        b[i] = np.zeros((0, 3), np.float32)
    for i in (13, 14, 15):
        b[i] = np.zeros(0, np.float32)
    out = spectra_priv._Particle_Interpolate(*b)
    assert out.shape == (4, a[1]) and out.dtype == np.float64 and not out.any()
    b = list(a)
    b[16] = np.zeros(0, np.int32)
    b[17] = np.zeros((0, 3), np.float64)
    out = spectra_priv._Particle_Interpolate(*b)
    assert out.shape == (0, a[1])


def test_near_lines_validation():
    a = _args()
    box, pos, h, axis, cofm = a[3], a[11], a[15], a[16], a[17]
    with pytest.raises(ValueError):
        spectra_priv._near_lines(box, pos, h, axis, cofm[:, :2].copy())
    with pytest.raises(ValueError):
        spectra_priv._near_lines(box, pos, h, axis.astype(np.int64), cofm)
    with pytest.raises(TypeError):
        spectra_priv._near_lines(box, pos.astype(np.float64), h, axis, cofm)
