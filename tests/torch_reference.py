"""FP64 pure-PyTorch restatement of the optical-depth path, differentiable by
autograd: the reference gradient for the hand-written adjoint.

It follows LineAbsorption::add_tau_particle (absorption.cpp:212-279) and
SingleAbsorber (singleabs.h:63-175) for a small set of pairs, vectorised over
pairs x candidate pixels x samples, with its own Faddeeva function:
  |u| < 8: Taylor series in a about the real axis, w^(k) by the recursion
           w' = -2zw + 2i/sqrt(pi), w^(k+1) = -2u w^(k) - 2k w^(k-1), with
           Dawson's integral by Rybicki's sum;
  |u| >= 8: Laplace continued fraction, 40 terms.
The tail cut is decided in FP64 on detached values (a discrete choice, held
fixed under differentiation).  Meant for small cases only.
"""
import math

import numpy as np
import torch
import torch.utils.checkpoint

SQPI = math.sqrt(math.pi)
TWO_SQPI = 2.0 / SQPI
SIGMA_T = 6.652458558e-25
BOLTZMANN = 1.3806504e-16
LIGHT = 2.99792458e10
PROTONMASS = 1.67262178e-24


def dawson(x):
    """D(x) for x >= 0, float64 tensor."""
    out = torch.empty_like(x)
    small = x < 0.2
    if small.any():
        xs = x[small]
        m = -2.0 * xs * xs
        t = torch.ones_like(xs)
        s = torch.ones_like(xs)
        for k in range(1, 14):
            t = t * m / (2 * k + 1)
            s = s + t
        out[small] = xs * s
    big = ~small
    if big.any():
        xb = x[big]
        h = 0.2
        n0 = (2.0 * torch.floor(0.5 * xb / h + 0.5)).detach()
        xp = xb - n0 * h
        n = torch.arange(1, 44, 2, dtype=x.dtype, device=x.device)
        e_up = torch.exp(-(xp[:, None] - n * h) ** 2) / (n0[:, None] + n)
        e_dn = torch.exp(-(xp[:, None] + n * h) ** 2) / (n0[:, None] - n)
        out[big] = (e_up + e_dn).sum(1) / SQPI
    return out


def faddeeva_re(u, a, nterms=16):
    """Re w(u + ia), a >= 0, float64 tensors of one shape."""
    au = u.abs()
    out = torch.empty_like(u)
    near = au < 8.0
    if near.any():
        x, y = au[near], a[near]
        w0 = torch.complex(torch.exp(-x * x), TWO_SQPI * dawson(x))
        w1 = -2.0 * x * w0 + 1j * TWO_SQPI
        iy = 1j * y
        coef = iy
        acc = w0 + coef * w1
        wkm1, wk = w0, w1
        for k in range(1, nterms):
            wkp1 = -2.0 * x * wk - 2.0 * k * wkm1
            coef = coef * iy / (k + 1)
            acc = acc + coef * wkp1
            wkm1, wk = wk, wkp1
        out[near] = acc.real
    far = ~near
    if far.any():
        z = torch.complex(au[far], a[far])
        T = z
        for k in range(40, 0, -1):
            T = z - (0.5 * k) / T
        out[far] = ((1j / SQPI) / T).real
    return out


def sph_cubic(q):
    norm = 32.0 / 4 / math.pi
    q = np.asarray(q, dtype=np.float64)
    r = np.where(q < 0.5, norm * (1 - 6 * q * q + 6 * q ** 3), norm * 2 * (1 - q) ** 3)
    return np.where(q >= 1, 0.0, r)


def sph_quintic(q):
    """singleabs.h:31-42: 9/(40 pi) sum_k c_k max(0, r_k - 3q)^5, (r, c) = (3, 1), (2, -6), (1, 15)."""
    q = np.asarray(q, dtype=np.float64)
    t = lambda r: np.maximum(r - 3.0 * q, 0.0) ** 5
    return np.where(q >= 1, 0.0, 9.0 / (40.0 * math.pi) * (t(3.0) - 6.0 * t(2.0) + 15.0 * t(1.0)))


def kernel_weight(kernel, q):
    """Weight of a kernel sample in tau_kern_inner (singleabs.h:158-163)."""
    if kernel == 1:
        return sph_cubic(q)
    if kernel == 3:
        return sph_quintic(q)
    return np.full_like(np.asarray(q, dtype=np.float64), 3.0 / 4.0 / math.pi)


def tau_torch(pairs, line, nbins, box, velfac, kernel, dens, temp, vlos, tautail=1e-7, chunk=None):
    """Optical depth [nlos, nbins] for the listed pairs.

    pairs: dict of numpy arrays over pairs: los (sightline), part (particle),
      dr2 (float64), ppos (float32 LOS position), h (float32); nlos.
    line: (lambda_cm, gamma, fosc, amumass).
    dens, temp, vlos: float64 torch tensors over particles (leaves for grads);
    vlos is the velocity component along each pair's sightline, indexed by
    pairs['vidx'] into a flat [N*3] tensor.
    chunk: if set, pairs are processed in chunks of this many under
    torch.utils.checkpoint (recomputed in backward), which bounds the memory
    of the autograd graph for a few thousand pairs."""
    lam, gamma, fosc, amu = line
    sigma_a = math.sqrt(3.0 * math.pi * SIGMA_T / 8.0) * lam * fosc
    bfac = math.sqrt(2.0 * BOLTZMANN / (amu * PROTONMASS)) / 1e5
    voigt_fac = gamma * lam / (4.0 * math.pi) / 1e5
    vbox = box * velfac
    bintov = vbox / nbins
    dv = dens.device
    part = torch.as_tensor(pairs["part"], dtype=torch.long, device=dv)
    dr2 = pairs["dr2"]
    h = pairs["h"].astype(np.float32)
    sm2 = (h * h).astype(np.float32).astype(np.float64)
    keep_pair = sm2 - dr2 > 0
    T = temp[part]
    b = bfac * torch.sqrt(T)
    vel = velfac * torch.as_tensor(pairs["ppos"].astype(np.float64), device=dv) + \
        vlos[torch.as_tensor(pairs["vidx"], device=dv)]
    vdr2 = velfac * dr2 * velfac
    vsm = velfac * h.astype(np.float64)
    vhigh = np.where(vsm * vsm > vdr2, np.sqrt(np.maximum(vsm * vsm - vdr2, 0)), 0.0)
    dvk = 2.0 * vhigh / 8
    ii = np.arange(1, 8)
    vv = ii[None, :] * dvk[:, None] - vhigh[:, None]                       # [P,7]
    q = np.sqrt(vdr2[:, None] + vv * vv) / vsm[:, None]
    W = kernel_weight(kernel, q)
    aa = voigt_fac / b
    amp = sigma_a / SQPI * (LIGHT / 1e5 / b)
    zmax = torch.floor(vel.detach() / bintov).long()
    half = nbins // 2
    idx = torch.arange(half, device=dv)
    zf = zmax[:, None] + idx[None, :]
    zb = zmax[:, None] - 1 - idx[None, :]
    z = torch.cat([zf, zb], 1)                                              # [P, 2*half]
    bd = b.detach().cpu().numpy()
    npts = np.where(bintov < bd / 2, 1, 2 * np.ceil(bintov / (bd / 2) / 2) + 1).astype(int)
    tau = torch.zeros(pairs["nlos"] * nbins, dtype=torch.float64, device=dv)
    vv_t = torch.as_tensor(vv, device=dv)
    W_t = torch.as_tensor(W, device=dv)

    def contribution(npv, sel, dens):
        st = torch.as_tensor(sel, device=dv)
        vlow = z[st].double() * bintov - vel[st][:, None]                   # [p, 2h]
        if npv == 1:
            offs = torch.tensor([0.5 * bintov], dtype=torch.float64, device=dv)
            wts = torch.tensor([1.0], dtype=torch.float64, device=dv)
        else:
            offs = torch.arange(npv, dtype=torch.float64, device=dens.device) * (bintov / (npv - 1))
            wts = torch.ones(npv, dtype=torch.float64, device=dens.device)
            wts[0] = wts[-1] = 0.5
            wts = wts / (npv - 1)
        vout = vlow[:, :, None] + offs[None, None, :]                       # [p, 2h, j]
        u = (vv_t[st][:, None, None, :] - vout[..., None]) / b[st][:, None, None, None]
        a = aa[st][:, None, None, None].expand_as(u)
        H = faddeeva_re(u.reshape(-1), a.reshape(-1)).reshape(u.shape)
        S = (H * W_t[st][:, None, None, :]).sum(-1)
        S = (S * wts[None, None, :]).sum(-1)                                # [p, 2h]
        t = amp[st][:, None] * dens[part[st]][:, None] * torch.as_tensor(dvk[sel], device=dv)[:, None] * S / velfac
        # tail cut: each direction keeps pixels up to and including the first below tautail
        td = t.detach()
        above = td >= tautail
        keep = torch.ones_like(above)
        for d in range(2):
            blk = above[:, d * half:(d + 1) * half]
            keep[:, d * half + 1:(d + 1) * half] = torch.cumprod(blk[:, :-1].long(), 1).bool()
        j = torch.remainder(z[st], nbins)
        rows = torch.as_tensor(pairs["los"][sel], device=dv)[:, None].expand_as(j)
        flat = (rows * nbins + j)[keep]
        return torch.zeros(pairs["nlos"] * nbins, dtype=torch.float64, device=dv).index_add(0, flat, t[keep])

    for npv in np.unique(npts):
        sel_all = np.nonzero((npts == npv) & keep_pair)[0]
        if sel_all.size == 0:
            continue
        step = sel_all.size if chunk is None else chunk
        for c0 in range(0, sel_all.size, step):
            sel = sel_all[c0:c0 + step]
            if chunk is None:
                tau = tau + contribution(npv, sel, dens)
            else:
                tau = tau + torch.utils.checkpoint.checkpoint(contribution, npv, sel, dens, use_reentrant=False)
    return tau.reshape(pairs["nlos"], nbins)


def pairs_from_oracle(oracle, w, axis, cofm):
    """Neighbour lists and per-pair data of a workload subset (via the oracle)."""
    off, parts, dr2 = oracle.near_particles(w.box, w.pos, w.hh, axis, cofm)
    L = axis.shape[0]
    los = np.repeat(np.arange(L), np.diff(off))
    ax = axis[los]
    return {"los": los, "part": parts.astype(np.int64), "dr2": dr2,
            "ppos": w.pos[parts, ax - 1], "h": w.hh[parts], "vidx": 3 * parts.astype(np.int64) + ax - 1,
            "nlos": L}
