"""The hand-written adjoint (fsamd_tau_adjoint behind torch.autograd) against
an FP64 pure-PyTorch reference gradient and central finite differences.

Loss, as in the differentiable workload: L = sum_pixels (exp(-tau) - F)^2,
F the flux of the same field with temperatures scaled by 1.1."""
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


def _workload(nlos, n_side, nbins, seed=1234):
    w = synth.config(3, nlos=nlos, n_side=n_side, seed=seed)
    w.nbins = nbins
    return w


def _target(w, plan):
    t = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp * np.float32(1.1)), [HI_1215],
               [dev(w.dens["H1"])])[0]
    return torch.exp(-t)


def _gpu_grads(w, plan, F):
    d = dev(w.dens["H1"]).requires_grad_(True)
    T = dev(w.temp).requires_grad_(True)
    v = dev(w.vel).requires_grad_(True)
    tau = tau_autograd(plan, w.nbins, w.kernel, w.velfac, v, T, [HI_1215], [d])[0]
    loss = ((torch.exp(-tau) - F) ** 2).sum()
    loss.backward()
    return loss.item(), d.grad.double().cpu().numpy(), T.grad.double().cpu().numpy(), v.grad.double().cpu().numpy()


def _compare(got, ref, what, rel=2e-3):
    scale = np.max(np.abs(ref))
    assert scale > 0, what
    err = np.max(np.abs(got - ref)) / scale
    assert err <= rel, "%s: max |grad - ref| / max |ref| = %.3g" % (what, err)
    m = np.abs(ref) > 1e-2 * scale
    comp = np.max(np.abs(got[m] - ref[m]) / np.abs(ref[m]))
    assert comp <= 10 * rel, "%s: largest componentwise relative error %.3g" % (what, comp)


def test_adjoint_matches_fp64_torch_reference(oracle):
    w = _workload(nlos=2, n_side=16, nbins=256)
    plan = plan_for(w)
    F = _target(w, plan)
    loss, gd, gT, gv = _gpu_grads(w, plan, F)
    pairs = TR.pairs_from_oracle(oracle, w, w.axis, w.cofm)
    d64 = torch.tensor(w.dens["H1"].astype(np.float64), device="cuda", requires_grad=True)
    T64 = torch.tensor(w.temp.astype(np.float64), device="cuda", requires_grad=True)
    v64 = torch.tensor(w.vel.astype(np.float64).ravel(), device="cuda", requires_grad=True)
    tau64 = TR.tau_torch(pairs, LINE, w.nbins, w.box, w.velfac, w.kernel, d64, T64, v64)
    loss64 = ((torch.exp(-tau64) - F) ** 2).sum()
    loss64.backward()
    assert abs(loss - loss64.item()) <= 1e-5 * abs(loss64.item())
    _compare(gd, d64.grad.cpu().numpy(), "dL/d dens")
    _compare(gT, T64.grad.cpu().numpy(), "dL/dT")
    _compare(gv.ravel(), v64.grad.cpu().numpy(), "dL/d vel")


def test_adjoint_central_differences():
    w = _workload(nlos=16, n_side=32, nbins=1024, seed=99)
    plan = plan_for(w)
    F = _target(w, plan)
    _, gd, gT, gv = _gpu_grads(w, plan, F)
    near = plan.near_lines().cpu().numpy()
    rng = np.random.default_rng(7)
    sel = rng.choice(near, size=min(64, near.size), replace=False)
    dvel, dtemp = dev(w.vel), dev(w.temp)

    def loss_of(dens=None, temp=None, vel=None):
        t = SP.tau(plan, w.nbins, w.kernel, w.velfac, vel if vel is not None else dvel,
                   temp if temp is not None else dtemp, [HI_1215],
                   [dens if dens is not None else dev(w.dens["H1"])])[0]
        return ((torch.exp(-t) - F) ** 2).sum().item()

    fd = {"dens": [], "temp": [], "vel": []}
    ad = {"dens": gd[sel], "temp": gT[sel], "vel": gv[sel, 0]}
    for p in sel:
        for name, base in (("dens", w.dens["H1"]), ("temp", w.temp), ("vel", w.vel)):
            x = base.copy()
            if name == "vel":
                x0 = float(x[p, 0])
                h = 1e-4 * max(abs(x0), 1.0)
                x[p, 0] = x0 + h
                lp = loss_of(vel=dev(x))
                x[p, 0] = x0 - h
                lm = loss_of(vel=dev(x))
                # the float32 input rounds the step: use the steps actually taken
                hp = float(np.float32(x0 + h)) - x0
                hm = x0 - float(np.float32(x0 - h))
            else:
                x0 = float(x[p])
                h = 1e-4 * abs(x0)
                x[p] = x0 + h
                lp = loss_of(**{name: dev(x)})
                x[p] = x0 - h
                lm = loss_of(**{name: dev(x)})
                hp = float(np.float32(x0 + h)) - x0
                hm = x0 - float(np.float32(x0 - h))
            fd[name].append((lp - lm) / (hp + hm))
    for name in fd:
        f = np.asarray(fd[name])
        a = ad[name]
        scale = np.max(np.abs(f))
        assert scale > 0, name
        bad = np.abs(a - f) > 2e-2 * np.abs(f) + 2e-3 * scale
        assert bad.sum() <= 2, "%s: %d of %d particles disagree with finite differences: %s vs %s" % (
            name, bad.sum(), f.size, a[bad][:4], f[bad][:4])


def test_adjoint_requires_recorded_windows():
    w = _workload(nlos=2, n_side=8, nbins=128)
    plan = plan_for(w)
    SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), [HI_1215], [dev(w.dens["H1"])])
    g = torch.zeros((plan.nlos, w.nbins), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        SP.tau_adjoint(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), [HI_1215],
                       [dev(w.dens["H1"])], [g])


def test_adjoint_neutral_fraction_through_torch():
    """d loss / d x_HI through the density pipeline built in torch
    (dens = x_HI * n_H rscale 0.76 / amu, spectra.py:593-615): autograd chains
    the hand-written adjoint into torch ops; checked against central
    differences in x_HI on random particles near the sightlines."""
    w = _workload(nlos=8, n_side=32, nbins=512, seed=5)
    plan = plan_for(w)
    F = _target(w, plan)
    xh = w.neutral_fraction
    nfac = (w.dens["H1"] / xh).astype(np.float32)
    nfac_t = dev(nfac)
    x = dev(xh).requires_grad_(True)
    dvel, dtemp = dev(w.vel), dev(w.temp)
    tau = tau_autograd(plan, w.nbins, w.kernel, w.velfac, dvel, dtemp, [HI_1215], [x * nfac_t])[0]
    ((torch.exp(-tau) - F) ** 2).sum().backward()
    gx = x.grad.double().cpu().numpy()

    def loss_of(xv):
        d = dev(xv) * nfac_t
        t = SP.tau(plan, w.nbins, w.kernel, w.velfac, dvel, dtemp, [HI_1215], [d])[0]
        return ((torch.exp(-t) - F) ** 2).sum().item()

    near = plan.near_lines().cpu().numpy()
    sel = np.random.default_rng(11).choice(near, size=min(32, near.size), replace=False)
    fd = []
    for p in sel:
        xv = xh.copy()
        x0 = float(xv[p])
        h = 1e-4 * x0
        xv[p] = x0 + h
        lp = loss_of(xv)
        xv[p] = x0 - h
        lm = loss_of(xv)
        hp = float(np.float32(x0 + h)) - x0
        hm = x0 - float(np.float32(x0 - h))
        fd.append((lp - lm) / (hp + hm))
    fd = np.asarray(fd)
    a = gx[sel]
    scale = np.max(np.abs(fd))
    assert scale > 0
    bad = np.abs(a - fd) > 2e-2 * np.abs(fd) + 2e-3 * scale
    assert bad.sum() <= 1, "x_HI: %d of %d disagree: %s vs %s" % (bad.sum(), fd.size, a[bad][:4], fd[bad][:4])
