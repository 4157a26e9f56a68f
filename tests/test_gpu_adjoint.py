"""The hand-written adjoint (fsamd_tau_adjoint behind torch.autograd).

Loss, as in the differentiable workload: L = sum_pixels (exp(-tau) - F)^2,
F the flux of the same field with temperatures scaled by 1.1.  Checked, for
the tophat (0), cubic (1) and quintic (3) kernels:

* against an FP64 pure-PyTorch reference gradient (tests/torch_reference.py)
  on a small field, and on a sightline through the full 128^3-particle field
  with 1024 pixels;
* against the same adjoint with every pair on the FP64 path
  (fsamd_set_fp64: FP64 Faddeeva function, pixel by pixel) on 64 sightlines
  of the full field: the FP32 profile derivatives and the far-wing moment
  derivatives against FP64 arithmetic;
* against central finite differences at the full workload (128^3 particles,
  1024 sightlines x 1024 pixels) on 64 random particles near the sightlines,
  for density, temperature, line-of-sight velocity and neutral fraction; the
  same for the Voronoi kernel (2) on a 64^3-particle field.
  The difference quotients are taken with the FP64 forward (fsamd_set_fp64)
  on the sightlines the particle reaches, so they carry no FP32 rounding
  noise.  The forward is piecewise smooth: a perturbation can move a pixel
  across the tail cut (absorption.cpp:260-263), shift the pixel a particle
  falls in (absorption.cpp:246) or change the sub-sample count
  (singleabs.h:110-114), and the adjoint differentiates the current piece
  only.  Each quotient is therefore taken at two steps, 1e-4 and 5e-5
  relative; a jump J inside the wider step changes the two quotients by
  J/(2h), so quotients that disagree by more than 1e-4 relative straddle a
  jump and are set aside (at most 8 of 64 per field), and every other one
  must match the adjoint to 1e-3.
* Cold gas (damping parameter above 0.04): the FP64 adjoint path against the
  torch reference.
* Windows overwritten between forward and backward are caught.
"""
import numpy as np
import pytest
import torch

from fake_spectra_amd import synth, spectra_priv as SP
from fake_spectra_amd.autograd import tau_autograd
from fake_spectra_amd.lines import HI_1215
from gpu_util import dev, plan_for
import torch_reference as TR

pytestmark = pytest.mark.gpu

LINE = (HI_1215.lambda_cm, HI_1215.gamma, HI_1215.fosc, HI_1215.amumass)
KERNELS = [0, 1, 3]


def _workload(nlos, n_side, nbins, seed=1234, kernel=1):
    w = synth.config(3, nlos=nlos, n_side=n_side, seed=seed)
    w.nbins = nbins
    w.kernel = kernel
    return w


def _target(w, plan):
    t = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp * np.float32(1.1)), [HI_1215],
               [dev(w.dens["H1"])])[0]
    return torch.exp(-t)


def _gpu_grads(w, plan, F, temp=None):
    d = dev(w.dens["H1"]).requires_grad_(True)
    T = dev(w.temp if temp is None else temp).requires_grad_(True)
    v = dev(w.vel).requires_grad_(True)
    tau = tau_autograd(plan, w.nbins, w.kernel, w.velfac, v, T, [HI_1215], [d])[0]
    loss = ((torch.exp(-tau) - F) ** 2).sum()
    loss.backward()
    return loss.item(), d.grad.double().cpu().numpy(), T.grad.double().cpu().numpy(), v.grad.double().cpu().numpy()


def _torch_grads(oracle, w, axis, cofm, F, temp=None, chunk=None):
    pairs = TR.pairs_from_oracle(oracle, w, axis, cofm)
    t = w.temp if temp is None else temp
    d64 = torch.tensor(w.dens["H1"].astype(np.float64), device="cuda", requires_grad=True)
    T64 = torch.tensor(t.astype(np.float64), device="cuda", requires_grad=True)
    v64 = torch.tensor(w.vel.astype(np.float64).ravel(), device="cuda", requires_grad=True)
    tau64 = TR.tau_torch(pairs, LINE, w.nbins, w.box, w.velfac, w.kernel, d64, T64, v64, chunk=chunk)
    loss64 = ((torch.exp(-tau64) - F) ** 2).sum()
    loss64.backward()
    return loss64.item(), d64.grad.cpu().numpy(), T64.grad.cpu().numpy(), v64.grad.cpu().numpy()


def _compare(got, ref, what, rel=2e-3):
    scale = np.max(np.abs(ref))
    assert scale > 0, what
    err = np.max(np.abs(got - ref)) / scale
    assert err <= rel, "%s: max |grad - ref| / max |ref| = %.3g" % (what, err)
    m = np.abs(ref) > 1e-2 * scale
    comp = np.max(np.abs(got[m] - ref[m]) / np.abs(ref[m]))
    assert comp <= 10 * rel, "%s: largest componentwise relative error %.3g" % (what, comp)


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_matches_fp64_torch_reference(oracle, kernel):
    w = _workload(nlos=2, n_side=16, nbins=256, kernel=kernel)
    plan = plan_for(w)
    F = _target(w, plan)
    loss, gd, gT, gv = _gpu_grads(w, plan, F)
    loss64, rd, rT, rv = _torch_grads(oracle, w, w.axis, w.cofm, F)
    assert abs(loss - loss64) <= 1e-5 * abs(loss64)
    _compare(gd, rd, "k%d dL/d dens" % kernel)
    _compare(gT, rT, "k%d dL/dT" % kernel)
    _compare(gv.ravel(), rv, "k%d dL/d vel" % kernel)


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_full_field_sightline_vs_torch_reference(oracle, kernel):
    """One sightline through the whole 128^3-particle field, 1024 pixels."""
    w = synth.config(3)
    w.kernel = kernel
    sel = np.array([5])
    axis, cofm = w.axis[sel], w.cofm[sel]
    plan = plan_for(w, axis, cofm)
    F = _target(w, plan)
    d = dev(w.dens["H1"]).requires_grad_(True)
    T = dev(w.temp).requires_grad_(True)
    v = dev(w.vel).requires_grad_(True)
    tau = tau_autograd(plan, w.nbins, w.kernel, w.velfac, v, T, [HI_1215], [d])[0]
    ((torch.exp(-tau) - F) ** 2).sum().backward()
    assert plan.npairs > 2000
    loss64, rd, rT, rv = _torch_grads(oracle, w, axis, cofm, F, chunk=256)
    _compare(d.grad.double().cpu().numpy(), rd, "k%d dL/d dens" % kernel)
    _compare(T.grad.double().cpu().numpy(), rT, "k%d dL/dT" % kernel)
    _compare(v.grad.double().cpu().numpy().ravel(), rv, "k%d dL/d vel" % kernel)


def _adjoint(plan, w, F, fp64):
    SP.set_fp64(fp64)
    try:
        return _gpu_grads(w, plan, F)
    finally:
        SP.set_fp64(False)


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_fp32_matches_fp64_path(kernel):
    w = synth.config(3)
    w.kernel = kernel
    sel = np.arange(0, 1024, 16)
    plan = plan_for(w, w.axis[sel], w.cofm[sel])
    F = _target(w, plan)
    _, d32, T32, v32 = _adjoint(plan, w, F, False)
    _, d64, T64, v64 = _adjoint(plan, w, F, True)
    for name, a, b in (("dens", d32, d64), ("temp", T32, T64), ("vel", v32.ravel(), v64.ravel())):
        scale = np.max(np.abs(b))
        assert scale > 0, name
        err = np.abs(a - b)
        assert np.all(err <= 1e-4 * np.abs(b) + 1e-6 * scale), \
            "k%d %s: worst %.3g relative" % (kernel, name, np.max(err / (np.abs(b) + 1e-6 * scale)))


def _rows_of(plan, nlos):
    off, parts = plan.pairs()
    off, parts = off.cpu().numpy(), parts.cpu().numpy()
    los = np.repeat(np.arange(nlos), np.diff(off))
    return los, parts


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_central_differences_full_workload(kernel):
    w = synth.config(3)
    assert w.npart == 128 ** 3 and w.nlos == 1024 and w.nbins == 1024
    w.kernel = kernel
    _central_differences(w, seed=7 + kernel)


def test_voronoi_adjoint_central_differences():
    """Kernel 2 (each particle deposited over the segment of the sightline it
    owns, IndexTable::assign_cells): the adjoint against central differences
    on a 64^3-particle field with 64 sightlines, segments assigned once per
    plan (they depend on positions only)."""
    w = synth.config(3, nlos=64, n_side=64)
    w.kernel = 2
    _central_differences(w, seed=11, assign=True)


def _central_differences(w, seed, assign=False):
    kernel = w.kernel
    pos, hh, axis, cofm = dev(w.pos), dev(w.hh), dev(w.axis), dev(w.cofm)
    plan = SP.Plan(w.box, pos, hh, axis, cofm)
    if assign:
        plan.assign_cells()
    F = _target(w, plan)
    xh = dev(w.neutral_fraction)
    nfac_h = (w.dens["H1"] / w.neutral_fraction).astype(np.float32)
    nfac = dev(nfac_h)
    dvel, dtemp = dev(w.vel), dev(w.temp)
    x = xh.clone().requires_grad_(True)
    T = dtemp.clone().requires_grad_(True)
    v = dvel.clone().requires_grad_(True)
    dens = x * nfac
    dens.retain_grad()
    tau = tau_autograd(plan, w.nbins, w.kernel, w.velfac, v, T, [HI_1215], [dens])[0]
    ((torch.exp(-tau) - F) ** 2).sum().backward()
    grads = {"dens": dens.grad.double().cpu().numpy(), "temp": T.grad.double().cpu().numpy(),
             "vel": v.grad.double().cpu().numpy()[:, 0], "xhi": x.grad.double().cpu().numpy()}
    los, parts = _rows_of(plan, w.nlos)
    del tau
    rng = np.random.default_rng(seed)
    sel = rng.choice(np.unique(parts), size=64, replace=False)
    base = {"dens": (x * nfac).detach(), "temp": dtemp, "vel": dvel, "xhi": xh}
    fd = {k: [] for k in grads}
    SP.set_fp64(True)
    try:
        for p in sel:
            rows = np.unique(los[parts == p])
            sub = SP.Plan(w.box, pos, hh, axis[torch.as_tensor(rows, device="cuda")],
                          cofm[torch.as_tensor(rows, device="cuda")])
            if assign:
                sub.assign_cells()
            Fs = F[torch.as_tensor(rows, device="cuda")]

            def loss_of(name, val):
                dd = base["dens"]
                tt, vv = dtemp, dvel
                if name == "dens":
                    dd = dd.clone()
                    dd[p] = val
                elif name == "xhi":
                    xx = xh.clone()
                    xx[p] = val
                    dd = (xx * nfac).contiguous()
                elif name == "temp":
                    tt = tt.clone()
                    tt[p] = val
                else:
                    vv = vv.clone()
                    vv[p, 0] = val
                t = SP.tau(sub, w.nbins, w.kernel, w.velfac, vv, tt, [HI_1215], [dd])[0]
                return ((torch.exp(-t) - Fs) ** 2).sum().item()

            for name in grads:
                b = base[name]
                x0 = float(b[p, 0] if name == "vel" else b[p])
                quot = []
                for rel in (1e-4, 5e-5):
                    h = rel * (max(abs(x0), 1.0) if name == "vel" else abs(x0))
                    xp, xm = np.float32(x0 + h), np.float32(x0 - h)
                    dl = loss_of(name, float(xp)) - loss_of(name, float(xm))
                    if name == "xhi":
                        # the density x * nfac is rounded to FP32: divide by the
                        # density step taken and apply d dens / dx = nfac
                        nf = np.float32(nfac_h[p])
                        quot.append(dl / (float(xp * nf) - float(xm * nf)) * float(nf))
                    else:
                        quot.append(dl / (float(xp) - float(xm)))
                fd[name].append(quot)
            sub.close()
    finally:
        SP.set_fp64(False)
    for name in grads:
        q = np.asarray(fd[name])
        a = grads[name][sel]
        scale = np.max(np.abs(q[:, 1]))
        assert scale > 0, name
        smooth = np.abs(q[:, 0] - q[:, 1]) <= 1e-4 * np.abs(q[:, 1]) + 1e-9 * scale
        assert smooth.sum() >= 56, "k%d %s: %d of 64 quotients straddle a jump" % (kernel, name, 64 - smooth.sum())
        f = q[smooth, 1]
        bad = np.abs(a[smooth] - f) > 1e-3 * np.abs(f) + 1e-5 * scale
        assert not bad.any(), "k%d %s: %d of %d particles disagree: %s vs %s" % (
            kernel, name, bad.sum(), smooth.sum(), a[smooth][bad][:4], f[bad][:4])


def test_cold_gas_adjoint_uses_fp64_path(oracle):
    """Pairs with a damping parameter above 0.04 (H I below ~1.4 K) take the
    FP64 path in the forward and in the adjoint."""
    w = _workload(nlos=2, n_side=16, nbins=256, seed=31)
    cold = w.temp.copy()
    cold[::3] = np.float32(1.0)
    plan = plan_for(w)
    F = _target(w, plan)
    loss, gd, gT, gv = _gpu_grads(w, plan, F, temp=cold)
    loss64, rd, rT, rv = _torch_grads(oracle, w, w.axis, w.cofm, F, temp=cold, chunk=16)
    assert abs(loss - loss64) <= 1e-5 * abs(loss64)
    _compare(gd, rd, "cold dL/d dens")
    _compare(gT, rT, "cold dL/dT")
    _compare(gv.ravel(), rv, "cold dL/d vel")


def test_adjoint_requires_recorded_windows():
    w = _workload(nlos=2, n_side=8, nbins=128)
    plan = plan_for(w)
    SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), [HI_1215], [dev(w.dens["H1"])])
    g = torch.zeros((plan.nlos, w.nbins), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        SP.tau_adjoint(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), [HI_1215],
                       [dev(w.dens["H1"])], [g])


def test_backward_after_another_forward_raises():
    """Two differentiable forwards on one plan: the second overwrites the
    first's windows, and the first's backward refuses to use them."""
    w = _workload(nlos=4, n_side=12, nbins=256)
    plan = plan_for(w)
    d1 = dev(w.dens["H1"]).requires_grad_(True)
    t1 = tau_autograd(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), [HI_1215], [d1])[0]
    d2 = dev(w.dens["H1"]).requires_grad_(True)
    t2 = tau_autograd(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp * np.float32(1.2)), [HI_1215],
                      [d2])[0]
    with pytest.raises(RuntimeError, match="windows"):
        t1.sum().backward()
    t2.sum().backward()
    assert torch.isfinite(d2.grad).all() and d2.grad.abs().max() > 0


def test_graph_replay_after_windows_released():
    """A graph captured while windows were recorded keeps its window buffer
    after recording is switched off: replays still match."""
    w = _workload(nlos=8, n_side=12, nbins=256)
    plan = plan_for(w)
    plan.keep_windows(True)
    g = SP.TauGraph(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), [HI_1215], [dev(w.dens["H1"])])
    first = g.replay()[0].clone()
    torch.cuda.synchronize()
    plan.keep_windows(False)
    again = g.replay()[0].clone()
    torch.cuda.synchronize()
    assert torch.allclose(first, again, rtol=1e-12, atol=0)
    assert not g.overflowed()
