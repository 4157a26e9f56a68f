"""The GPU path at full workload sizes.

* Every sightline of the full launches through a size-independent property:
  each pair's contribution depends on that pair alone (part_int.cpp:33-47
  adds them one by one), so the spectra of the even- and odd-indexed
  particles sum to the spectra of all particles up to the order of the FP64
  additions (compared on the device, no host copy of the full arrays):
    256^3 uniform, 5000 sightlines x 2048 pixels, H I 1215 + 1025;
    256^3 clustered, 3 x 224^2 gridded sightlines x 1024 pixels,
      H I 1215 + 1025 + C IV 1548 (the 150,528-sightline batched launch);
    512^3 clustered, 100,000 sightlines on mixed axes x 2048 pixels,
      H I 1215 tophat + column density.
* Additivity holds for any deterministic per-pair deposit, right or wrong, so
  the same full launches are also held to the oracle on sightlines taken
  from both ends of the pair-count order (the heaviest and the lightest
  non-empty rows): neighbour lists bit-exact, tau and column density to the
  stated tolerances.
* _near_lines without the neighbour lists (fsamd_near_lines) equals the
  particles flagged by the full lists, at each full size.
* The 128^3-particle differentiable workload (1024 sightlines x 1024 pixels)
  on sampled sightlines against the oracle.
"""
import numpy as np
import pytest
import torch

from fake_spectra_amd import spectra_priv as SP, synth
from gpu_util import dev, assert_tau_close, assert_colden_close

pytestmark = pytest.mark.gpu

_CACHE = {}


def workload(idx):
    """One full-size workload at a time (the 512^3 field alone takes ~6 GB)."""
    if idx not in _CACHE:
        _CACHE.clear()
        _CACHE[idx] = synth.config(idx)
    return _CACHE[idx]


class Field:
    """Device copies of a workload's particle arrays."""

    def __init__(self, w, sel=slice(None)):
        self.pos, self.hh = dev(w.pos[sel]), dev(w.hh[sel])
        self.vel, self.temp = dev(w.vel[sel]), dev(w.temp[sel])
        self.dens = [dev(w.dens_for(ln)[sel]) for ln in w.lines]
        self.axis, self.cofm = dev(w.axis), dev(w.cofm)


def _run(w, sel, colden):
    f = Field(w, sel)
    plan = SP.Plan(w.box, f.pos, f.hh, f.axis, f.cofm)
    outs = SP.tau(plan, w.nbins, w.kernel, w.velfac, f.vel, f.temp, w.lines, f.dens, w.tautail)
    if colden:
        outs.append(SP.colden(plan, w.nbins, w.kernel, w.velfac, [f.dens[0]])[0])
    return plan, outs


def _additive(full, even, odd, names):
    for name, a, e, o in zip(names, full, even, odd):
        assert bool(torch.isfinite(a).all()), name
        s = e + o
        amax = a.abs().max().item()
        assert amax > 0, name
        big = a.abs() > 1e-3 * amax
        rel = ((s - a).abs()[big] / a.abs()[big]).max().item()
        assert rel <= 1e-10, "%s: %.3g" % (name, rel)
        assert (s - a).abs().max().item() <= 1e-10 * amax, name
        assert bool((s[a == 0] == 0).all()), name


def _extremes(plan, nheavy, nlight):
    cnt = np.diff(plan.offsets().cpu().numpy())
    order = np.argsort(-cnt, kind="stable")
    nonempty = order[cnt[order] > 0]
    return np.unique(np.concatenate([nonempty[:nheavy], nonempty[-nlight:]])), cnt


def _check_rows(oracle, w, plan, outs, rows, colden):
    got = plan.rows(rows)
    roff, rparts, _ = oracle.near_particles(w.box, w.pos, w.hh, w.axis[rows], w.cofm[rows])
    for i, s in enumerate(rows):
        assert np.array_equal(got[i], rparts[roff[i]:roff[i + 1]]), "pairs of sightline %d" % s
    idx = torch.as_tensor(rows, device="cuda")
    for k, ln in enumerate(w.lines):
        args = list(w.interp_args(ln, True))
        args[16], args[17] = w.axis[rows], w.cofm[rows]
        assert_tau_close(outs[k][idx].cpu().numpy(), oracle.interpolate(*args), "%s %s" % (w.name, ln.name))
    if colden:
        args = list(w.interp_args(None, False))
        args[16], args[17] = w.axis[rows], w.cofm[rows]
        assert_colden_close(outs[-1][idx].cpu().numpy(), oracle.interpolate(*args), "%s colden" % w.name)


def _full_launch(oracle, idx, colden, nheavy, nlight):
    w = workload(idx)
    n = w.npart
    plan, full = _run(w, slice(0, n), colden)
    m_all = plan.npairs
    # _near_lines without neighbour lists equals the flags of the full lists
    f = Field(w)
    direct = SP.near_lines(w.box, f.pos, f.hh, f.axis, f.cofm)
    assert torch.equal(direct, plan.near_lines())
    del f, direct
    rows, cnt = _extremes(plan, nheavy, nlight)
    assert cnt.max() > cnt[cnt > 0].min()
    _check_rows(oracle, w, plan, full, rows, colden)
    plan.close()
    pe, even = _run(w, slice(0, n, 2), colden)
    m_even = pe.npairs
    pe.close()
    po, odd = _run(w, slice(1, n, 2), colden)
    m_odd = po.npairs
    po.close()
    assert m_even + m_odd == m_all          # every pair lands in exactly one half
    _additive(full, even, odd, [ln.name for ln in w.lines] + (["colden"] if colden else []))


@pytest.mark.timeout(600)
def test_c1_uniform_full_launch(oracle):
    w = workload(1)
    assert w.npart == 256 ** 3 and w.nlos == 5000 and w.nbins == 2048
    _full_launch(oracle, 1, True, 8, 8)


@pytest.mark.timeout(600)
def test_c2_clustered_150528_sightlines_full_launch(oracle):
    w = workload(2)
    assert w.npart == 256 ** 3 and w.nlos == 3 * 224 ** 2 and w.nbins == 1024 and len(w.lines) == 3
    _full_launch(oracle, 2, False, 12, 20)


@pytest.mark.timeout(600)
def test_c3_differentiable_workload_sampled_sightlines(oracle):
    w = workload(3)
    assert w.npart == 128 ** 3 and w.nlos == 1024 and w.nbins == 1024
    f = Field(w)
    plan = SP.Plan(w.box, f.pos, f.hh, f.axis, f.cofm)
    outs = SP.tau(plan, w.nbins, w.kernel, w.velfac, f.vel, f.temp, w.lines, f.dens, w.tautail)
    outs.append(SP.colden(plan, w.nbins, w.kernel, w.velfac, [f.dens[0]])[0])
    rows, _ = _extremes(plan, 8, 8)
    rows = np.unique(np.concatenate([rows, np.arange(0, w.nlos, 64)]))
    _check_rows(oracle, w, plan, outs, rows, True)


@pytest.mark.timeout(900)
def test_c4_clustered_512cubed_full_launch(oracle):
    w = workload(4)
    assert w.npart == 512 ** 3 and w.nlos == 100000 and w.nbins == 2048 and w.kernel == 0
    assert set(np.unique(w.axis)) == {1, 2, 3}
    _full_launch(oracle, 4, True, 8, 16)
