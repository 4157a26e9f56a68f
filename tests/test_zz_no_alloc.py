"""The allocation-free one-call entry fsamd_particle_interpolate_noalloc
against the allocating one-call entry fsamd_particle_interpolate, through the
C ABI on device buffers: two calls in a row, the second on a smaller input
(the scratch keeps the first call's contents), tau and column density, to
the stated tolerances and to 1e-12 (the two run the same kernels); and a
call too large for the scratch is refused with a clear error."""
import ctypes as C

import numpy as np
import pytest
import torch

from fake_spectra_amd import synth
from fake_spectra_amd import _native as N
from fake_spectra_amd.lines import HI_1215
from gpu_util import dev, assert_tau_close, assert_colden_close

pytestmark = pytest.mark.gpu


def _lib():
    lib = C.CDLL(N.LIB_PATH)
    vp, d, i, ll = C.c_void_p, C.c_double, C.c_int, C.c_longlong
    for name in ("fsamd_particle_interpolate", "fsamd_particle_interpolate_noalloc"):
        f = getattr(lib, name)
        f.argtypes = [i, i, i, d, d, d, d, d, d, d, d, vp, vp, vp, vp, vp, ll, vp, vp, i, vp, vp]
        f.restype = i
    lib.fsamd_last_error.restype = C.c_char_p
    return lib


def _call(fn, compute_tau, w):
    ct, nb, kern, bx, velfac, atime, lam, gamma, fosc, amu, tautail, pos, vel, dens, temp, h, axis, cofm = \
        w.interp_args(HI_1215, bool(compute_tau))
    bufs = [dev(a) for a in (pos, vel, dens, temp, h, axis, cofm)]
    out = torch.full((len(axis), nb), -1.0, dtype=torch.float64, device="cuda")
    p = [C.c_void_p(b.data_ptr()) for b in bufs]
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(ct, nb, kern, bx, velfac, atime, lam, gamma, fosc, amu, tautail, p[0], p[1], p[2], p[3], p[4],
            len(pos), p[5], p[6], len(axis), C.c_void_p(out.data_ptr()), stream)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy()


def test_noalloc_matches_allocating_entry_twice():
    lib = _lib()
    first = synth.config(0)                          # 32^3 particles, 64 sightlines x 1024 pixels
    second = synth.config(0, nlos=16, n_side=20, seed=99)   # smaller, different field
    assert second.npart < first.npart and second.nlos < first.nlos
    for w in (first, second):
        for compute_tau in (1, 0):
            rc, ref = _call(lib.fsamd_particle_interpolate, compute_tau, w)
            assert rc == 0, lib.fsamd_last_error()
            rc, got = _call(lib.fsamd_particle_interpolate_noalloc, compute_tau, w)
            assert rc == 0, lib.fsamd_last_error()
            what = "%s %s" % (w.name, "tau" if compute_tau else "colden")
            assert np.any(ref > 0), what
            if compute_tau:
                assert_tau_close(got, ref, what)
            else:
                assert_colden_close(got, ref, what)
            scale = np.maximum(np.abs(ref), 1e-300)
            assert np.max(np.abs(got - ref) / scale) <= 1e-12, what


def test_noalloc_refuses_what_does_not_fit():
    lib = _lib()
    w = synth.config(3, nlos=4)                      # 128^3 particles: ~200 MiB of records alone
    rc, _ = _call(lib.fsamd_particle_interpolate_noalloc, 1, w)
    assert rc == N.FSAMD_ERR_UNSUPPORTED
    assert b"scratch" in lib.fsamd_last_error()
    # and the entry still works afterwards
    small = synth.config(0, nlos=8, n_side=12)
    rc, got = _call(lib.fsamd_particle_interpolate_noalloc, 1, small)
    assert rc == 0, lib.fsamd_last_error()
    rc, ref = _call(lib.fsamd_particle_interpolate, 1, small)
    assert_tau_close(got, ref, "after refusal")
