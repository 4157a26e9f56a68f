"""GPU path against the oracle and the reference's golden outputs.
Binning is index work: bit-exact.  tau/colden: the stated tolerances."""
import numpy as np
import pytest
import torch

from fake_spectra_amd import spectra_priv as SP, synth
from fake_spectra_amd.lines import HI_1215, HI_1025, CIV_1548
from golden_cases import c0_fixture, crafted, tau_keys
from gpu_util import dev, assert_tau_close, assert_colden_close, plan_for

pytestmark = pytest.mark.gpu


def _pairs(plan):
    off, parts = plan.pairs()
    return off.cpu().numpy(), parts.cpu().numpy()


def test_binning_c0_matches_reference():
    w, z = c0_fixture()
    sel = z["sel"]
    plan = plan_for(w, w.axis[sel], w.cofm[sel])
    off, parts = _pairs(plan)
    assert np.array_equal(off, z["offsets"])
    assert np.array_equal(parts, z["parts"])
    near = SP._near_lines(w.box, w.pos, w.hh, w.axis, w.cofm)
    assert np.array_equal(near, z["near_lines"])


def test_binning_crafted_matches_reference():
    cases, z = crafted()
    for name, c in cases.items():
        plan = SP.Plan(float(c["box"]), dev(c["pos"]), dev(c["hh"]), dev(c["axis"]), dev(c["cofm"]))
        off, parts = _pairs(plan)
        assert np.array_equal(off, z[name + "_offsets"]), name
        assert np.array_equal(parts, z[name + "_parts"]), name
        near = SP._near_lines(float(c["box"]), c["pos"], c["hh"], c["axis"], c["cofm"])
        assert np.array_equal(near, z[name + "_near"]), name


@pytest.mark.parametrize("idx,scale,nlos", [(1, 1 / 64, 24), (2, 1 / 64, None), (4, 1 / 512, 300)])
def test_binning_configs_match_oracle(oracle, idx, scale, nlos):
    w = synth.config(idx, scale=scale, nlos=nlos)
    plan = plan_for(w)
    off, parts = _pairs(plan)
    roff, rparts, _ = oracle.near_particles(w.box, w.pos, w.hh, w.axis, w.cofm)
    assert np.array_equal(off, roff) and np.array_equal(parts, rparts)
    assert plan.npairs == len(rparts)


def test_tau_c0_full(oracle):
    w, _ = c0_fixture()
    args = w.interp_args(HI_1215, True)
    got = SP._Particle_Interpolate(*args)
    assert_tau_close(got, oracle.interpolate(*args), "c0 tau")


def test_colden_c0_full(oracle):
    w, z = c0_fixture()
    args = w.interp_args(None, False)
    got = SP._Particle_Interpolate(*args)
    assert_colden_close(got, oracle.interpolate(*args), "c0 colden")
    sel = z["sel"]
    assert_colden_close(got[sel], z["colden"], "c0 colden vs reference")


def test_tau_c0_vs_reference_fixture():
    w, z = c0_fixture()
    args = list(w.interp_args(HI_1215, True))
    args[16], args[17] = w.axis[z["sel"]], w.cofm[z["sel"]]
    assert_tau_close(SP._Particle_Interpolate(*args), z["tau"], "c0 subset vs reference")


def test_crafted_cases_vs_reference_fixture():
    cases, z = crafted()
    velfac = float(z["velfac"])
    for key, case, kern, line, gs, nb in tau_keys(z):
        c = cases[case]
        got = SP._Particle_Interpolate(1, nb, kern, float(c["box"]), velfac, 0.25, line.lambda_cm, line.gamma * gs,
                                       line.fosc, line.amumass, 1e-7, c["pos"], c["vel"], c["dens"], c["temp"],
                                       c["hh"], c["axis"], c["cofm"])
        assert_tau_close(got, z[key], key)
    for case, c in cases.items():
        for kern in (0, 1, 3):
            got = SP._Particle_Interpolate(0, 257, kern, float(c["box"]), velfac, 0.25, HI_1215.lambda_cm,
                                           HI_1215.gamma, HI_1215.fosc, HI_1215.amumass, 1e-7, c["pos"],
                                           np.zeros(1, np.float32), c["dens"], np.zeros(1, np.float32), c["hh"],
                                           c["axis"], c["cofm"])
            assert_colden_close(got, z["%s_k%d_colden" % (case, kern)], case)


@pytest.mark.parametrize("idx,scale,nlos", [(1, 1 / 64, 16), (2, 1 / 64, 6), (3, 1 / 8, 16), (4, 1 / 512, 64)])
def test_tau_configs_vs_oracle(oracle, idx, scale, nlos):
    """Every line of each workload, from one binning, in one batched launch."""
    w = synth.config(idx, scale=scale, nlos=nlos)
    plan = plan_for(w)
    outs = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines,
                  [dev(w.dens_for(ln)) for ln in w.lines], w.tautail)
    for ln, o in zip(w.lines, outs):
        ref = oracle.interpolate(*w.interp_args(ln, True))
        assert_tau_close(o.cpu().numpy(), ref, "c%d %s" % (idx, ln.name))
    col = SP.colden(plan, w.nbins, w.kernel, w.velfac, [dev(w.dens_for(w.lines[0]))])[0]
    assert_colden_close(col.cpu().numpy(), oracle.interpolate(*w.interp_args(None, False)), "c%d colden" % idx)


@pytest.mark.parametrize("kernel", [0, 3])
def test_kernels_on_a_workload_match_oracle(oracle, kernel):
    """The tophat and the quintic SPH kernel (singleabs.h:31-42) on the
    c1-like field, both lines, in one batched launch."""
    w = synth.config(1, scale=1 / 64, nlos=12)
    w.kernel = kernel
    plan = plan_for(w)
    outs = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines,
                  [dev(w.dens_for(ln)) for ln in w.lines], w.tautail)
    for ln, o in zip(w.lines, outs):
        assert_tau_close(o.cpu().numpy(), oracle.interpolate(*w.interp_args(ln, True)), "kernel %d %s" % (kernel, ln.name))


@pytest.mark.parametrize("kernel", [5, -1])
def test_unknown_kernel_ids_take_the_reference_fallbacks(oracle, kernel):
    """An id other than 0-3 leaves the kernel samples unweighted in tau
    (singleabs.h:158-163 take no branch) and takes the tophat fraction in
    column density (absorption.cpp:136-148), as the reference does; the
    oracle's fallback is pinned to the compiled reference on the CPU."""
    w = synth.config(0, nlos=8)
    for ct in (True, False):
        args = list(w.interp_args(HI_1215 if ct else None, ct))
        args[2] = kernel
        got = SP._Particle_Interpolate(*args)
        ref = oracle.interpolate(*args)
        if ct:
            assert_tau_close(got, ref, "kernel %d tau" % kernel)
        else:
            assert_colden_close(got, ref, "kernel %d colden" % kernel)


def test_voronoi_mesh_matches_oracle(oracle):
    """Kernel 2: the sightlines cut into the segments their particles own
    (IndexTable::assign_cells) on the GPU, then the mesh branches of the tau
    and column-density deposits, against the oracle (pinned to the compiled
    reference); through the drop-in and through one plan serving two lines."""
    w = synth.config(0, nlos=6)
    for ct in (True, False):
        args = list(w.interp_args(HI_1215 if ct else None, ct))
        args[2] = 2
        got = SP._Particle_Interpolate(*args)
        ref = oracle.interpolate(*args)
        assert np.any(ref > 0)
        if ct:
            assert_tau_close(got, ref, "Voronoi tau")
        else:
            assert_colden_close(got, ref, "Voronoi colden")
    plan = plan_for(w)
    plan.assign_cells()
    d = dev(w.dens["H1"])
    outs = SP.tau(plan, w.nbins, 2, w.velfac, dev(w.vel), dev(w.temp), [HI_1215, HI_1025], [d, d], w.tautail)
    for ln, o in zip((HI_1215, HI_1025), outs):
        args = list(w.interp_args(ln, True))
        args[2] = 2
        assert_tau_close(o.cpu().numpy(), oracle.interpolate(*args), "Voronoi plan %s" % ln.name)


def test_voronoi_needs_assigned_cells():
    w = synth.config(0, nlos=2, n_side=8)
    plan = plan_for(w)
    with pytest.raises(ValueError, match="assign_cells"):
        SP.tau(plan, w.nbins, 2, w.velfac, dev(w.vel), dev(w.temp), [HI_1215], [dev(w.dens["H1"])])


@pytest.mark.parametrize("g,waves", [(16, 4), (32, 4), (64, 4), (16, 2), (32, 8)])
def test_group_sizes_match_oracle(oracle, g, waves):
    """Every lane-group width and wavefront count of the tau kernel gives
    the same spectra."""
    from fake_spectra_amd import _native as N
    w = synth.config(1, scale=1 / 64, nlos=8)
    plan = plan_for(w)
    N.set_group_size(g)
    N.set_block_waves(waves)
    try:
        outs = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines,
                      [dev(w.dens_for(ln)) for ln in w.lines], w.tautail)
    finally:
        N.set_group_size(16)
        N.set_block_waves(8)
    for ln, o in zip(w.lines, outs):
        assert_tau_close(o.cpu().numpy(), oracle.interpolate(*w.interp_args(ln, True)),
                         "G=%d waves=%d %s" % (g, waves, ln.name))


def test_graph_replays_match_eager():
    """A captured batched launch, replayed several times with eager work and
    synchronizations in between, reproduces the eager result every time."""
    w = synth.config(1, scale=1 / 64, nlos=12)
    plan = plan_for(w)
    vel, temp = dev(w.vel), dev(w.temp)
    dens = [dev(w.dens_for(ln)) for ln in w.lines]
    eager = [o.clone() for o in SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens)]
    g = SP.TauGraph(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens)
    for rep in range(4):
        for o in g.out:
            o.fill_(-1.0)
        torch.cuda.synchronize()
        outs = g.replay()
        torch.cuda.synchronize()
        for a, b in zip(outs, eager):
            np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-12, atol=1e-300)
        # eager work between replays
        SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens)
    assert not g.overflowed()


def test_autotune_picks_a_width(oracle):
    w = synth.config(1, scale=1 / 64, nlos=8)
    plan = plan_for(w)
    vel, temp = dev(w.vel), dev(w.temp)
    dens = [dev(w.dens_for(ln)) for ln in w.lines]
    best, waves, ms = SP.autotune(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens)
    assert best in (16, 32, 64) and waves in (2, 4, 8) and all(v > 0 for v in ms.values())
    assert len(ms) == 7
    outs = SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens)
    for ln, o in zip(w.lines, outs):
        assert_tau_close(o.cpu().numpy(), oracle.interpolate(*w.interp_args(ln, True)), "tuned " + ln.name)


def test_plain_mode_equals_batched():
    w = synth.config(2, scale=1 / 64, nlos=4)
    plan = plan_for(w)
    dens = [dev(w.dens_for(ln)) for ln in w.lines]
    a = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines, dens, mode=0)
    b = SP.tau(plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines, dens, mode=1)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x.cpu().numpy(), y.cpu().numpy(), rtol=1e-12, atol=1e-300)


def test_rows_wider_than_lds(oracle):
    """nbins too large for an LDS row: deposits go straight to HBM."""
    w = synth.config(0, nlos=3, n_side=12)
    w.nbins = 40000
    args = w.interp_args(HI_1215, True)
    assert_tau_close(SP._Particle_Interpolate(*args), oracle.interpolate(*args), "nbins=40000")
    cargs = w.interp_args(None, False)
    assert_colden_close(SP._Particle_Interpolate(*cargs), oracle.interpolate(*cargs), "nbins=40000 colden")


@pytest.mark.parametrize("nbins,waves", [(19000, 4), (15000, 8)])
def test_rows_near_the_lds_limit(oracle, nbins, waves):
    """Rows that fit LDS only without the kernel's own LDS: 19000 pixels go
    to HBM at 4 wavefronts, 15000 pixels at 8 wavefronts fall back to 4 and
    stay in LDS."""
    from fake_spectra_amd import _native as N
    w = synth.config(0, nlos=3, n_side=12)
    w.nbins = nbins
    N.set_block_waves(waves)
    try:
        args = w.interp_args(HI_1215, True)
        got = SP._Particle_Interpolate(*args)
    finally:
        N.set_block_waves(8)
    assert_tau_close(got, oracle.interpolate(*args), "nbins=%d waves=%d" % (nbins, waves))


def test_concurrent_calls_match_sequential():
    """Drop-in calls from several host threads at once, each with its own
    pixel count (and so its own LDS size), give the results of the same calls
    made one after another."""
    from concurrent.futures import ThreadPoolExecutor
    w = synth.config(0, nlos=8, n_side=16)
    sizes = (512, 1024, 2048, 4096, 3000, 700)

    def run(nbins):
        args = list(w.interp_args(HI_1215, True))
        args[1] = nbins
        return SP._Particle_Interpolate(*args)

    want = [run(n) for n in sizes]
    for _ in range(3):
        with ThreadPoolExecutor(max_workers=len(sizes)) as ex:
            got = list(ex.map(run, sizes))
        for n, a, b in zip(sizes, got, want):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=0, err_msg="nbins=%d" % n)


def test_cold_gas_uses_fp64_profile(oracle):
    """Damping parameters above the FP32 profile's range go through FP64."""
    w = synth.config(0, nlos=6, n_side=10)
    w.temp[:] = np.float32(1.5)        # H I at 1.5 K: a ~ 0.04
    w.temp[::2] = np.float32(1.0)
    args = w.interp_args(HI_1215, True)
    SP.set_stats(True)
    try:
        got = SP._Particle_Interpolate(*args)
        st = SP.last_stats()
    finally:
        SP.set_stats(False)
    assert st["fp64_pairs"] > 0
    assert_tau_close(got, oracle.interpolate(*args), "cold gas")


@pytest.mark.parametrize("temps", [(3000.0, 20000.0), (1500.0, 3000.0, 20000.0), (1500.0, 1800.0)])
def test_profile_regimes_match_oracle(oracle, temps):
    """H I pairs at damping parameters from 4e-4 to 1.2e-3 (20000 K down
    to 1500 K), one value per sightline set and mixed within wavefronts."""
    w = synth.config(0, nlos=8, n_side=12)
    for i, t in enumerate(temps):
        w.temp[i::len(temps)] = np.float32(t)
    args = w.interp_args(HI_1215, True)
    assert_tau_close(SP._Particle_Interpolate(*args), oracle.interpolate(*args), "temps %s" % (temps,))


def test_stats_counters_consistent():
    """The counting kernels deposit the same spectra as the plain ones (up to
    the order of the FP64 additions), and
    their counters add up: every evaluated pixel is either a far-wing pixel or
    one or more 7-sample kernel sums, and at least one per deposited pixel."""
    w = synth.config(1, scale=1 / 64, nlos=16)
    plan = plan_for(w)
    args = (plan, w.nbins, w.kernel, w.velfac, dev(w.vel), dev(w.temp), w.lines,
            [dev(w.dens_for(ln)) for ln in w.lines], w.tautail)
    base = [o.cpu().numpy() for o in SP.tau(*args)]
    for timing_only in (False, True):
        SP.set_stats(True, timing_only=timing_only)
        try:
            outs = [o.cpu().numpy() for o in SP.tau(*args)]
            st = SP.last_stats()
        finally:
            SP.set_stats(False)
        for a, b in zip(outs, base):
            # FP64 LDS atomics add in launch-dependent order
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
        assert st["gpu_ms"] > 0
        assert 0 < st["kernel_ms"] <= st["gpu_ms"]
        if timing_only:
            continue
        assert st["pairs"] > 0 and st["pixels"] > 0
        assert st["evaluated_pixels"] >= st["pixels"]
        assert st["far_pixels"] < st["evaluated_pixels"]
        assert st["kernel_sums"] >= st["evaluated_pixels"] - st["far_pixels"]
        assert st["voigt_samples"] >= 7 * (st["pixels"] - st["far_pixels"])


def test_weighted_fields_match_oracle(oracle):
    """Column-density-weighted velocity and temperature from one binning and
    one batched colden launch equal the quotient of the oracle's column
    densities of dens*w and dens (spectra.py:945-1010)."""
    from fake_spectra_amd import products
    w = synth.config(0, nlos=8)
    plan = plan_for(w)
    dens = w.dens["H1"]
    r = products.velocity_temperature(plan, w.nbins, w.kernel, w.velfac, dev(dens), dev(w.vel), dev(w.temp),
                                      w.atime)

    def col(d):
        a = list(w.interp_args(None, False))
        a[13] = np.ascontiguousarray(d, dtype=np.float32)
        return oracle.interpolate(*a)

    base = col(dens)
    assert_colden_close(r["colden"].cpu().numpy(), base, "colden")
    den = np.where(base == 0, 1.0, base)
    v = (w.vel * np.float32(np.sqrt(w.atime))).astype(np.float32)
    checks = [("temp", r["temp"], dens * w.temp)] + \
             [("vel%d" % a, r["velocity"][..., a], dens * v[:, a]) for a in range(3)]
    for name, got, weighted in checks:
        ref = col(weighted) / den
        g = got.cpu().numpy()
        scale = np.max(np.abs(ref))
        np.testing.assert_allclose(g, ref, rtol=3e-6, atol=3e-6 * scale, err_msg=name)
