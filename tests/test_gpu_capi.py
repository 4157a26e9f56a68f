"""The one-call C entry fsamd_particle_interpolate (the drop-in of
_Particle_Interpolate, py_module.cpp:103-233, on device pointers), bound
through ctypes exactly as a non-Python caller binds it, on device buffers:
tau and column density of the 32^3-particle, 64-sightline field against the
oracle and the reference's own outputs, the Voronoi kernel, and its status
codes."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from fake_spectra_amd import synth
from fake_spectra_amd import _native as N
from fake_spectra_amd.lines import HI_1215
from gpu_util import dev, assert_tau_close, assert_colden_close

pytestmark = pytest.mark.gpu


def _lib():
    lib = C.CDLL(N.LIB_PATH)
    vp, d, i, ll = C.c_void_p, C.c_double, C.c_int, C.c_longlong
    lib.fsamd_particle_interpolate.argtypes = [i, i, i, d, d, d, d, d, d, d, d,
                                               vp, vp, vp, vp, vp, ll, vp, vp, i, vp, vp]
    lib.fsamd_particle_interpolate.restype = i
    lib.fsamd_last_error.restype = C.c_char_p
    return lib


def _call(lib, compute_tau, w, kernel=None, nbins=None, null_pos=False, box=None, npart=None, nlos=None):
    ct, nb, kern, bx, velfac, atime, lam, gamma, fosc, amu, tautail, pos, vel, dens, temp, h, axis, cofm = \
        w.interp_args(HI_1215, bool(compute_tau))
    nb = nbins if nbins is not None else nb
    kern = kernel if kernel is not None else kern
    bufs = [dev(a) for a in (pos, vel, dens, temp, h, axis, cofm)]
    out = torch.full((len(axis), max(nb, 1)), 7.0, dtype=torch.float64, device="cuda")
    p = [C.c_void_p(b.data_ptr()) for b in bufs]
    if null_pos:
        p[0] = C.c_void_p(0)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.fsamd_particle_interpolate(ct, nb, kern, bx if box is None else box, velfac, atime, lam, gamma, fosc,
                                        amu, tautail, p[0], p[1], p[2], p[3], p[4],
                                        len(pos) if npart is None else npart, p[5], p[6],
                                        len(axis) if nlos is None else nlos, C.c_void_p(out.data_ptr()), stream)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy()


def test_one_call_tau_and_colden_match_oracle_and_reference(oracle):
    lib = _lib()
    w = synth.config(0)
    z = np.load(os.path.join(GOLDEN, "ref_c0.npz"))
    rc, tau = _call(lib, 1, w)
    assert rc == 0, lib.fsamd_last_error()
    assert_tau_close(tau, oracle.interpolate(*w.interp_args(HI_1215, True)), "c0 tau vs oracle")
    assert_tau_close(tau[z["sel"]], z["tau"], "c0 tau vs reference")
    rc, col = _call(lib, 0, w)
    assert rc == 0, lib.fsamd_last_error()
    assert_colden_close(col, oracle.interpolate(*w.interp_args(None, False)), "c0 colden vs oracle")
    assert_colden_close(col[z["sel"]], z["colden"], "c0 colden vs reference")


def test_one_call_voronoi_kernel_matches_oracle(oracle):
    """Kernel 2 through the one-call entry: it cuts the sightlines into the
    particles' segments itself (IndexTable::assign_cells) before the deposit."""
    lib = _lib()
    w = synth.config(0, nlos=6)
    for ct in (1, 0):
        rc, got = _call(lib, ct, w, kernel=2)
        assert rc == 0, lib.fsamd_last_error()
        args = list(w.interp_args(HI_1215 if ct else None, bool(ct)))
        args[2] = 2
        ref = oracle.interpolate(*args)
        assert np.any(ref > 0)
        if ct:
            assert_tau_close(got, ref, "one-call Voronoi tau")
        else:
            assert_colden_close(got, ref, "one-call Voronoi colden")


def test_one_call_status_codes():
    lib = _lib()
    w = synth.config(0, nlos=4, n_side=8)
    rc, _ = _call(lib, 1, w, nbins=0)
    assert rc == N.FSAMD_ERR_ARG and b"nbins" in lib.fsamd_last_error()
    rc, _ = _call(lib, 1, w, null_pos=True)
    assert rc == N.FSAMD_ERR_ARG
    rc, _ = _call(lib, 1, w, box=-1.0)
    assert rc == N.FSAMD_ERR_ARG and b"box" in lib.fsamd_last_error()
    rc, _ = _call(lib, 1, w, npart=-5)
    assert rc == N.FSAMD_ERR_ARG
    # a valid call after the failures still works, and the row is fully overwritten
    rc, tau = _call(lib, 1, w)
    assert rc == 0 and np.all(np.isfinite(tau)) and not np.any(tau == 7.0)
