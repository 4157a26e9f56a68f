"""Drop-in for the reference's native module ``fake_spectra._spectra_priv``.

``_Particle_Interpolate`` and ``_near_lines`` keep the reference's positional
arguments, array layouts, dtype rules, exception classes and return values
(fake_spectra/py_module.cpp:25-233); the work runs on the GPU through
libfsamd.  ``Plan``, ``tau`` and ``colden`` are the device-resident API behind
them: torch tensors on the GPU in, torch tensors out, one binning shared by
any number of lines.
"""
from __future__ import annotations

import ctypes as C
import operator
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from .lines import Line

_USAGE = ("Incorrect arguments: use compute_tau, nbins, boxsize, velfac, atime, lambda, gamma, fosc, "
          "species mass (amu), min. tau,  pos, vel, dens, temp, h, axis, cofm\n")

KERNELS = {"tophat": 0, "cubic": 1, "sph": 1}


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("fake_spectra_amd runs on a ROCm GPU and none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())


class _DeviceArray:
    """Zero-copy view of a libfsamd device buffer for torch.as_tensor."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr,
                                         "data": (int(ptr), False), "version": 3, "strides": None}


def _require(t: torch.Tensor, dtype: torch.dtype, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s must have dtype %s" % (name, dtype))
    if t.device.type != "cuda":
        raise ValueError("%s must be on the GPU" % name)
    return t.contiguous()


class Plan:
    """Sightline <-> particle neighbour lists on the device (IndexTable,
    fake_spectra/index_table.cpp:7-150).  Build once, deposit any number of
    lines and column densities from it."""

    def __init__(self, box: float, pos: torch.Tensor, h: torch.Tensor, axis: torch.Tensor,
                 cofm: torch.Tensor):
        pos = _require(pos, torch.float32, "pos")
        h = _require(h, torch.float32, "h")
        axis = _require(axis, torch.int32, "axis")
        cofm = _require(cofm, torch.float64, "cofm")
        if pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError("pos must have shape (N, 3)")
        if h.numel() != pos.shape[0]:
            raise ValueError("h must have one entry per particle")
        if cofm.dim() != 2 or cofm.shape[1] != 3 or cofm.shape[0] != axis.numel():
            raise ValueError("cofm must have dimensions (np.size(axis),3)")
        self.box = float(box)
        self.npart = int(pos.shape[0])
        self.nlos = int(axis.numel())
        self._handle = C.c_void_p()
        N.check(N.lib().fsamd_plan_create(self.box, _ptr(pos), _ptr(h), self.npart, _ptr(axis), _ptr(cofm),
                                          self.nlos, _stream(), C.byref(self._handle)))
        self.pos, self.h = pos, h

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            N.lib().fsamd_plan_destroy(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        if not self._handle.value:
            raise RuntimeError("plan is closed")
        return self._handle

    @property
    def npairs(self) -> int:
        return int(N.lib().fsamd_plan_npairs(self.handle))

    def pairs(self) -> Tuple[torch.Tensor, torch.Tensor]:
        """(offsets int64 [nlos+1], particle ids int32 [npairs]) as device
        tensors (copies), each row ascending in particle index."""
        off, parts = C.c_void_p(), C.c_void_p()
        N.check(N.lib().fsamd_plan_pairs(self.handle, C.byref(off), C.byref(parts)))
        torch.cuda.current_stream().synchronize()
        o = torch.as_tensor(_DeviceArray(off.value, (self.nlos + 1,), "<i8"), device=self.pos.device).clone()
        m = self.npairs
        if m == 0:
            return o, torch.zeros(0, dtype=torch.int32, device=self.pos.device)
        p = torch.as_tensor(_DeviceArray(parts.value, (m,), "<i4"), device=self.pos.device).clone()
        return o, p

    def offsets(self) -> torch.Tensor:
        """Row offsets int64 [nlos+1] of the neighbour lists (a device copy)."""
        off = C.c_void_p()
        N.check(N.lib().fsamd_plan_pairs(self.handle, C.byref(off), None))
        torch.cuda.current_stream().synchronize()
        return torch.as_tensor(_DeviceArray(off.value, (self.nlos + 1,), "<i8"), device=self.pos.device).clone()

    def rows(self, sightlines) -> list:
        """Particle ids (numpy int32) of the given sightlines' neighbour lists,
        copying only those rows to the host."""
        off, parts = C.c_void_p(), C.c_void_p()
        N.check(N.lib().fsamd_plan_pairs(self.handle, C.byref(off), C.byref(parts)))
        torch.cuda.current_stream().synchronize()
        o = self.offsets().cpu().numpy()
        m = self.npairs
        if m == 0:
            return [np.zeros(0, dtype=np.int32) for _ in sightlines]
        p = torch.as_tensor(_DeviceArray(parts.value, (m,), "<i4"), device=self.pos.device)
        return [p[int(o[s]):int(o[s + 1])].cpu().numpy() for s in sightlines]

    def assign_cells(self):
        """Voronoi mesh (kernel 2): cut each sightline into the segments its
        particles own (IndexTable::assign_cells, index_table.cpp:152-223);
        needed before tau / colden with kernel 2."""
        N.check(N.lib().fsamd_plan_assign_cells(self.handle, _ptr(self.pos), _stream()))

    def keep_windows(self, enable: bool = True):
        """Record the deposited window of every (line, pair, direction) in the
        next tau() call, as the adjoint needs."""
        N.check(N.lib().fsamd_plan_keep_windows(self.handle, 1 if enable else 0))

    def window_generation(self) -> int:
        """Counter bumped whenever the recorded windows change."""
        g = C.c_longlong()
        N.check(N.lib().fsamd_plan_window_generation(self.handle, C.byref(g)))
        return int(g.value)

    def near_lines(self) -> torch.Tensor:
        """Sorted ids of particles near any sightline (_near_lines)."""
        out = torch.empty(max(self.npart, 1), dtype=torch.int32, device=self.pos.device)
        n = C.c_longlong()
        N.check(N.lib().fsamd_plan_near_lines(self.handle, _ptr(out), C.byref(n), _stream()))
        return out[: n.value]


def near_lines(box: float, pos: torch.Tensor, h: torch.Tensor, axis: torch.Tensor, cofm: torch.Tensor) -> torch.Tensor:
    """Sorted ids (int32 device tensor) of particles whose support reaches at
    least one sightline, without building the neighbour lists (the work of
    Py_near_lines, py_module.cpp:25-99, on device tensors)."""
    pos = _require(pos, torch.float32, "pos")
    h = _require(h, torch.float32, "h")
    axis = _require(axis, torch.int32, "axis")
    cofm = _require(cofm, torch.float64, "cofm")
    n = int(pos.shape[0]) if pos.dim() >= 1 else 0
    out = torch.empty(max(n, 1), dtype=torch.int32, device=pos.device)
    cnt = C.c_longlong()
    N.check(N.lib().fsamd_near_lines(float(box), _ptr(pos), _ptr(h), n, _ptr(axis), _ptr(cofm), int(axis.numel()),
                                     _ptr(out), C.byref(cnt), _stream()))
    return out[: cnt.value]


def _line_structs(lines, dens_list, outs):
    arr = (N.FsamdLine * len(lines))()
    for i, (ln, d, o) in enumerate(zip(lines, dens_list, outs)):
        lam, gam, fosc, amu = ln
        arr[i] = N.FsamdLine(float(lam), float(gam), float(fosc), float(amu), d.data_ptr(), o.data_ptr())
    return arr


def _as_quad(line) -> Tuple[float, float, float, float]:
    if isinstance(line, Line):
        return (line.lambda_cm, line.gamma, line.fosc, line.amumass)
    lam, gam, fosc, amu = line
    return (float(lam), float(gam), float(fosc), float(amu))


def tau(plan: Plan, nbins: int, kernel: int, velfac: float, vel: torch.Tensor, temp: torch.Tensor,
        lines: Sequence, dens: Sequence[torch.Tensor], tautail: float = 1e-7, mode: int = N.MODE_BATCHED,
        out: Optional[Sequence[torch.Tensor]] = None):
    """Optical depth float64 [nlos, nbins] for each line, from one binning.
    lines: Line objects or (lambda_cm, gamma, fosc, amumass) tuples;
    dens: one float32 [N] device tensor per line."""
    vel = _require(vel, torch.float32, "vel")
    temp = _require(temp, torch.float32, "temp")
    if vel.numel() != 3 * plan.npart or temp.numel() != plan.npart:
        raise ValueError("vel and temp must have one entry per particle")
    quads = [_as_quad(x) for x in lines]
    dens = [_require(d, torch.float32, "dens") for d in dens]
    if len(dens) != len(quads):
        raise ValueError("one density array per line is required")
    for d in dens:
        if d.numel() != plan.npart:
            raise ValueError("Dens, pos and h must have the same length")
    if out is None:
        out = [torch.empty((plan.nlos, nbins), dtype=torch.float64, device=plan.pos.device) for _ in quads]
    arr = _line_structs(quads, dens, out)
    N.check(N.lib().fsamd_tau(plan.handle, int(nbins), int(kernel), float(velfac), float(tautail), _ptr(plan.pos),
                              _ptr(vel), _ptr(temp), _ptr(plan.h), arr, len(quads), int(mode), _stream()))
    return list(out)


def tau_adjoint(plan: Plan, nbins: int, kernel: int, velfac: float, vel: torch.Tensor, temp: torch.Tensor,
                lines: Sequence, dens: Sequence[torch.Tensor], grad_tau: Sequence[torch.Tensor]):
    """dL/d dens (one float64 [N] per line), dL/dT (float64 [N]) and dL/d vel
    (float64 [N,3]) from dL/dtau of each line, through the last tau() call on
    this plan (made with plan.keep_windows(True))."""
    vel = _require(vel, torch.float32, "vel")
    temp = _require(temp, torch.float32, "temp")
    quads = [_as_quad(x) for x in lines]
    dens = [_require(d, torch.float32, "dens") for d in dens]
    gts = [_require(g, torch.float64, "grad_tau") for g in grad_tau]
    for g in gts:
        if g.shape != (plan.nlos, nbins):
            raise ValueError("grad_tau must have shape (nlos, nbins)")
    dv = plan.pos.device
    gd = [torch.zeros(plan.npart, dtype=torch.float64, device=dv) for _ in quads]
    gT = torch.zeros(plan.npart, dtype=torch.float64, device=dv)
    gv = torch.zeros((plan.npart, 3), dtype=torch.float64, device=dv)
    arr = (N.FsamdLineGrad * len(quads))()
    for i, ((lam, gam, fosc, amu), d, g, o) in enumerate(zip(quads, dens, gts, gd)):
        arr[i] = N.FsamdLineGrad(lam, gam, fosc, amu, d.data_ptr(), g.data_ptr(), o.data_ptr())
    N.check(N.lib().fsamd_tau_adjoint(plan.handle, int(nbins), int(kernel), float(velfac), _ptr(plan.pos), _ptr(vel),
                                      _ptr(temp), _ptr(plan.h), arr, len(quads), _ptr(gT), _ptr(gv), _stream()))
    return gd, gT, gv


def autotune(plan: Plan, nbins: int, kernel: int, velfac: float, vel: torch.Tensor, temp: torch.Tensor,
             lines: Sequence, dens: Sequence[torch.Tensor], tautail: float = 1e-7):
    """Pick the tau kernel's lane-group width and wavefronts per sightline
    for this plan and call shape by timing them on a sample of the work;
    returns (width, waves, {(width, waves): ms})."""
    vel = _require(vel, torch.float32, "vel")
    temp = _require(temp, torch.float32, "temp")
    quads = [_as_quad(x) for x in lines]
    dens = [_require(d, torch.float32, "dens") for d in dens]
    scratch = [torch.empty((plan.nlos, nbins), dtype=torch.float64, device=plan.pos.device) for _ in quads]
    arr = _line_structs(quads, dens, scratch)
    best, waves = C.c_int(), C.c_int()
    ms = (C.c_float * 7)()
    N.check(N.lib().fsamd_tau_autotune(plan.handle, int(nbins), int(kernel), float(velfac), float(tautail),
                                       _ptr(plan.pos), _ptr(vel), _ptr(temp), _ptr(plan.h), arr, len(quads), _stream(),
                                       C.byref(best), C.byref(waves), ms))
    times = {(16, 4): ms[0], (32, 4): ms[1], (64, 4): ms[2], (16, 8): ms[3], (32, 8): ms[4], (64, 8): ms[5]}
    times[(best.value, 2)] = ms[6]    # 2 wavefronts are timed at the best width
    return best.value, waves.value, times


class TauGraph:
    """One batched tau call (all sightlines x lines in one persistent launch
    plus its fix-up pass) captured in a HIP graph and replayed.  The outputs
    are fixed tensors (self.out); inputs are read from the tensors given here
    at every replay."""

    def __init__(self, plan: Plan, nbins: int, kernel: int, velfac: float, vel: torch.Tensor,
                 temp: torch.Tensor, lines: Sequence, dens: Sequence[torch.Tensor], tautail: float = 1e-7):
        self.plan = plan
        self.args = (nbins, kernel, velfac, _require(vel, torch.float32, "vel"),
                     _require(temp, torch.float32, "temp"), list(lines),
                     [_require(d, torch.float32, "dens") for d in dens], tautail)
        self.out = [torch.empty((plan.nlos, nbins), dtype=torch.float64, device=plan.pos.device) for _ in lines]
        nb, kern, vf, v, t, ln, d, tt = self.args
        tau(plan, nb, kern, vf, v, t, ln, d, tt, out=self.out)      # allocations and tables, uncaptured
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            tau(plan, nb, kern, vf, v, t, ln, d, tt, out=self.out)

    def replay(self):
        self.graph.replay()
        return self.out

    def overflowed(self) -> bool:
        """True if a replay overflowed the fix-up list (results then invalid)."""
        flag = C.c_int()
        N.check(N.lib().fsamd_plan_overflow(self.plan.handle, C.byref(flag), _stream()))
        return bool(flag.value)


def colden(plan: Plan, nbins: int, kernel: int, velfac: float, dens: Sequence[torch.Tensor],
           out: Optional[Sequence[torch.Tensor]] = None):
    """Column density float64 [nlos, nbins] for each density array."""
    dens = [_require(d, torch.float32, "dens") for d in dens]
    for d in dens:
        if d.numel() != plan.npart:
            raise ValueError("Dens, pos and h must have the same length")
    if out is None:
        out = [torch.empty((plan.nlos, nbins), dtype=torch.float64, device=plan.pos.device) for _ in dens]
    quads = [(0.0, 0.0, 0.0, 1.0)] * len(dens)
    arr = _line_structs(quads, dens, out)
    N.check(N.lib().fsamd_colden(plan.handle, int(nbins), int(kernel), float(velfac), _ptr(plan.pos), _ptr(plan.h),
                                 arr, len(dens), _stream()))
    return list(out)


def last_stats():
    buf = (C.c_longlong * 13)()
    N.check(N.lib().fsamd_last_stats(buf, 13))
    return {"pairs": buf[0], "pixels": buf[1], "fp64_pixels": buf[2], "fp64_pairs": buf[3],
            "voigt_samples": buf[4], "evaluated_pixels": buf[5], "kernel_sums": buf[6], "far_pixels": buf[7],
            "cycles_setup": buf[8], "cycles_wing": buf[9], "cycles_walk": buf[10], "wave_iterations": buf[11],
            "setup_iterations": buf[12],
            "gpu_ms": float(N.lib().fsamd_last_ms()), "kernel_ms": float(N.lib().fsamd_last_kernel_ms())}


def voigt_profile(u, a: float) -> torch.Tensor:
    """H(a, u) = Re w(u + ia) by the deposit kernel's FP32 profile at the
    points u (a CUDA tensor or array, evaluated in FP64 offsets), 0 <= a <= 0.04."""
    dev = device()
    uu = torch.as_tensor(np.asarray(u, dtype=np.float64) if not isinstance(u, torch.Tensor) else u,
                         dtype=torch.float64).to(dev).contiguous().reshape(-1)
    out = torch.empty(uu.numel(), dtype=torch.float32, device=dev)
    N.check(N.lib().fsamd_voigt_eval(_ptr(uu), uu.numel(), float(a), _ptr(out), _stream()))
    return out


def set_fp64(on: bool):
    """Evaluate every pair of later tau / tau_adjoint calls on this thread with
    the FP64 profile and kernel weights (a test and diagnostic switch)."""
    N.check(N.lib().fsamd_set_fp64(1 if on else 0))


def release_pool():
    """Free the device buffers the library keeps from destroyed plans on the
    current device (fsamd_pool_release)."""
    N.check(N.lib().fsamd_pool_release())


def set_stats(enable: bool, timing_only: bool = False):
    """Collect the counters and GPU time of each tau call (last_stats), or
    with timing_only just the time (the kernels then run without counting)."""
    N.check(N.lib().fsamd_set_stats(0 if not enable else (2 if timing_only else 1)))


# ---------------------------------------------------------------------------
# The reference's module functions (py_module.cpp)

def _is_f32(a: np.ndarray) -> bool:
    return a.dtype == np.dtype(np.float32)


def _parse(args):
    """PyArg_ParseTuple(args, "iiiddddddddO!O!O!O!O!O!O!") (py_module.cpp:115);
    any failure becomes AttributeError with the usage string (116-118)."""
    try:
        if len(args) != 18:
            raise TypeError
        ints = []
        for a in args[:3]:
            if isinstance(a, (float, np.floating)):
                raise TypeError
            ints.append(operator.index(a))
        dbl = [float(a) for a in args[3:11]]
        arrays = list(args[11:18])
        for a in arrays:
            if not isinstance(a, np.ndarray):
                raise TypeError
        return ints, dbl, arrays
    except (TypeError, ValueError, OverflowError):
        raise AttributeError(_USAGE) from None


def _to_dev(a: np.ndarray, dev) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _Particle_Interpolate(*args):
    """Optical depth (compute_tau=1) or column density (0) along sightlines
    (Py_Particle_Interpolation, py_module.cpp:103-233).  Returns a fresh
    float64 array [NumLos, nbins]."""
    (compute_tau, nbins, kernel), (box100, velfac, atime, lam, gamma, fosc, amumass, tautail), arrays = _parse(args)
    pos, vel, dens, temp, h, axis, cofm = arrays
    if not (_is_f32(pos) and _is_f32(vel) and _is_f32(dens) and _is_f32(temp) and _is_f32(h)):
        raise TypeError("One of the data arrays does not have 32-bit float type\n")
    if cofm.dtype != np.dtype(np.float64):
        raise TypeError("Sightline positions must have 64-bit float type\n")
    if axis.dtype != np.dtype(np.int32):
        raise TypeError("Axis must be a 32-bit integer\n")
    numlos = cofm.shape[0] if cofm.ndim >= 1 else 0
    npart = pos.shape[0] if pos.ndim >= 1 else 0
    if dens.ndim < 1 or h.ndim < 1 or npart != dens.shape[0] or npart != h.shape[0]:
        raise ValueError(" Dens, pos and h must have the same length\n")
    if axis.ndim < 1 or cofm.ndim != 2 or numlos != axis.shape[0] or cofm.shape[1] != 3:
        raise ValueError("cofm must have dimensions (np.size(axis),3) \n")
    # Beyond the reference, which would read out of bounds:
    if npart > 0 and (pos.ndim != 2 or pos.shape[1] != 3):
        raise ValueError("pos must have shape (N, 3)")
    if compute_tau and (vel.size < 3 * npart or temp.size < npart):
        raise ValueError("vel and temp must have one entry per particle when compute_tau is set")
    # 0 tophat, 1 cubic SPH, 2 Voronoi mesh, 3 quintic SPH; any other id takes
    # the reference's fallbacks (unweighted kernel samples in tau, the tophat
    # fraction in column density; singleabs.h:158-163, absorption.cpp:136-148)
    if numlos and (axis.min() < 1 or axis.max() > 3):
        raise ValueError("axis values must be 1, 2 or 3")
    if nbins < 1:
        raise ValueError("nbins must be positive")
    if numlos == 0 or npart == 0:
        return np.zeros((numlos, nbins), dtype=np.float64)
    dev = device()
    tpos = _to_dev(pos.reshape(npart, 3), dev)
    th = _to_dev(h, dev)
    plan = Plan(box100, tpos, th, _to_dev(axis, dev), _to_dev(cofm, dev))
    try:
        if kernel == 2:
            plan.assign_cells()     # part_int.cpp:28 (the mesh's segments)
        td = _to_dev(dens, dev)
        if compute_tau:
            out = tau(plan, nbins, kernel, velfac, _to_dev(vel.reshape(-1)[: 3 * npart], dev),
                      _to_dev(temp.reshape(-1)[:npart], dev), [(lam, gamma, fosc, amumass)], [td], tautail)[0]
        else:
            out = colden(plan, nbins, kernel, velfac, [td])[0]
        res = out.cpu().numpy()
    finally:
        plan.close()
    return res


def _near_lines(*args):
    """Sorted unique int32 indices of particles within their smoothing length
    of at least one sightline (Py_near_lines, py_module.cpp:25-99)."""
    if len(args) != 5:
        raise TypeError("_near_lines takes exactly 5 arguments (%d given)" % len(args))
    box100, pos, hh, axis, cofm = args
    box100 = float(box100)
    for a in (pos, hh, axis, cofm):
        if not isinstance(a, np.ndarray):
            raise TypeError("argument must be numpy.ndarray, not %s" % type(a).__name__)
    if cofm.ndim < 2 or axis.ndim < 1:
        raise ValueError("cofm must have dimensions (np.size(axis),3) \n")
    numlos = cofm.shape[0]
    npart = pos.shape[0] if pos.ndim >= 1 else 0
    if numlos != axis.shape[0] or cofm.shape[1] != 3:
        raise ValueError("cofm must have dimensions (np.size(axis),3) \n")
    if cofm.dtype != np.dtype(np.float64) or axis.dtype != np.dtype(np.int32):
        raise ValueError("cofm must have 64-bit float type and axis must be a 32-bit integer\n")
    if not (_is_f32(pos) and _is_f32(hh)):
        raise TypeError("pos and h must have 32-bit float type\n")
    if npart > 0 and (pos.ndim != 2 or pos.shape[1] != 3 or hh.shape[0] != npart):
        raise ValueError("pos must have shape (N, 3) and h length N")
    if numlos and (axis.min() < 1 or axis.max() > 3):
        raise ValueError("axis values must be 1, 2 or 3")
    if numlos == 0 or npart == 0:
        return np.zeros(0, dtype=np.int32)
    dev = device()
    return near_lines(box100, _to_dev(pos, dev), _to_dev(hh, dev), _to_dev(axis, dev), _to_dev(cofm, dev)).cpu().numpy()


def _rescale_mean_flux(*args):
    """Scale factor for tau reaching a desired mean flux (Py_mean_flux,
    py_module.cpp:263-282): (tau float64 array, mean_flux_desired, nbins,
    tol, thresh) -> float; the first nbins entries of tau are used."""
    if len(args) != 5 or not isinstance(args[0], np.ndarray):
        raise AttributeError("Incorrect arguments: use tau (array), mean_flux_desired (double), nbins (int), "
                             "tol (double), thresh (double)\n")
    tau, mfd, nbins, tol, thresh = args
    try:
        mfd, nbins, tol, thresh = float(mfd), int(nbins), float(tol), float(thresh)
    except (TypeError, ValueError):
        raise AttributeError("Incorrect arguments: use tau (array), mean_flux_desired (double), nbins (int), "
                             "tol (double), thresh (double)\n")
    if tau.dtype != np.dtype(np.float64):
        raise TypeError("Optical depth must have 64-bit float type\n")
    flat = np.ascontiguousarray(tau).ravel()
    if nbins < 0 or nbins > flat.size:
        raise ValueError("nbins must lie between 0 and the size of tau")
    t = torch.from_numpy(flat[:nbins].copy()).to(device())
    out = C.c_double()
    N.check(N.lib().fsamd_mean_flux_scale(_ptr(t), nbins, mfd, tol, thresh, C.byref(out), _stream()))
    return out.value
