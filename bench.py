#!/usr/bin/env python3
"""Throughput of the sightline spectral-synthesis path on one MI355X.

A step is one pass of the hot path over the workload: build the
sightline<->particle neighbour lists (IndexTable) from device-resident
particles, then deposit the optical depth of every line of every sightline
in one batched launch.  The default workload is 256^3 clustered particles
in a 60 Mpc/h box (70% uniform, 30% in 2000 Gaussian blobs), 3 x 224^2
gridded sightlines along the three axes, 1024 pixels each, H I 1215, H I 1025
and C IV 1548 (synthetic gas field, seeded).  Prints one JSON line, which
also carries the parity of the timed rows against the reference's rows of
the CPU baseline's sightlines.

    python bench.py [--steps K] [--warmup W] [--config 0..4] [--no-cpu] [--no-plain]

Defaults: 3 untimed warm-up steps, 10 timed steps (value from their total
time; the median step is reported beside it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def progress(msg):
    """One line per finished phase on stderr (the JSON line alone goes to stdout)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


T_START = time.perf_counter()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2,
                    help="workload (2: 256^3 clustered, 3 x 224^2 sightlines x 1024 pixels, 3 lines)")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the _near_lines host sample and of the whole-job table sample")
    ap.add_argument("--cpu-sightlines", type=int, default=1000,
                    help="sightlines of the CPU baseline's deposit sample (and of the parity check)")
    ap.add_argument("--graph-replays", type=int, default=2)
    ap.add_argument("--tune", default="", metavar="GxW",
                    help="lane-group width x wavefronts per sightline instead of autotuning (e.g. 32x8)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-near-lines", action="store_true")
    ap.add_argument("--no-drop-in", action="store_true")
    ap.add_argument("--no-flux", action="store_true", help="skip the flux-statistics leg")
    ap.add_argument("--adjoint-config", type=int, default=3,
                    help="workload of the forward vs forward+adjoint leg (3: 128^3, 1024 x 1024, H I 1215)")
    ap.add_argument("--colden", action="store_true",
                    help="also time the column density of the first line's density (binning included)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes) and exit")
    return ap.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "copies": 3}))


def cpu_sample(w, nsel):
    """The sightlines the CPU baseline runs: every k-th, at least nsel of them."""
    import numpy as np
    k = max(1, w.nlos // max(1, nsel))
    return np.arange(0, w.nlos, k)[:max(nsel, 1)], k


def cpu_baseline(w, sel, k, table_budget_s):
    """The reference C++ (oracle/_ref, OpenMP, all host threads; the oracle's
    C restatement where the reference library is not built) on the full
    particle set, one _Particle_Interpolate call per line as the reference
    makes them (spectra.py:673), scaled to the whole job.

    A reference call is IndexTable + get_near_particles over all particles
    and the call's sightlines (index_table.cpp:130-150; its pair inserts are
    serialised under omp critical), then the per-sightline deposit
    (part_int.cpp:20-51).  Both are estimated separately:
      * the deposit from the sampled sightlines: each sampled call's time
        minus the time of its table alone (measured on the same sightlines);
      * the table of a whole-job call (all sightlines, so all M pairs) from
        every kp-th particle with all sightlines, times kp, kp grown until
        that sample takes about table_budget_s.
    Job time = per line: table(all) + deposit(sample) x L / L_sample.
    Returns the baseline and the sampled rows (for the parity check)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    lib = O.load_ref() if kind == "reference" else O.load_oracle()
    cores = lib.max_threads()
    rows = []
    t_calls = []
    for line in w.lines:
        args = list(w.interp_args(line, True))
        args[16], args[17] = w.axis[sel], w.cofm[sel]
        t0 = time.perf_counter()
        rows.append(lib.interpolate(*args))
        t_calls.append(time.perf_counter() - t0)
        progress("cpu baseline: %s on %d sightlines, %.1f s" % (line.name, len(sel), t_calls[-1]))
    t_tab_s, m_s = lib.table_seconds(w.box, w.pos, w.hh, w.axis[sel], w.cofm[sel])
    kp = 1024
    while True:
        kp = max(1, kp)
        t_tab_p, m_p = lib.table_seconds(w.box, w.pos[::kp], w.hh[::kp], w.axis, w.cofm)
        if t_tab_p >= table_budget_s / 4 or kp == 1:
            break
        kp = max(1, int(kp * t_tab_p / table_budget_s) if t_tab_p > 0.05 else kp // 8)
    t_tab_full = t_tab_p * kp
    dep = [max(tc - t_tab_s, 0.0) for tc in t_calls]
    job_s = sum(t_tab_full + d * w.nlos / len(sel) for d in dep)
    units = len(w.lines) * w.nlos * w.nbins
    value = units / job_s
    sample = ("%s, lines %s (one call each, all %d particles): deposit from %d of %d sightlines (every %d-th), "
              "%.1f s of calls less %.2f s of table per call; whole-job table (all sightlines, %d pairs) from "
              "every %d-th particle (%d pairs, %.2f s) x %d = %.1f s per call; job estimate %.0f s" % (
                  "reference C++ -O2 -ffast-math OpenMP" if kind == "reference" else "C restatement",
                  "+".join(ln.name for ln in w.lines), w.npart, len(sel), w.nlos, k, sum(t_calls), t_tab_s,
                  m_p * kp, kp, m_p, t_tab_p, kp, t_tab_full, job_s))
    return {"value": value, "unit": "sightline-pixels/s", "cores": cores, "kind": kind, "sample": sample,
            "sightlines_sampled": int(len(sel)), "job_seconds_estimate": job_s,
            "table_seconds_per_call": t_tab_full,
            "deposit_seconds_per_call": [d * w.nlos / len(sel) for d in dep]}, rows


def parity_rows(w, gpu_rows, ref_rows):
    """Timed GPU rows against the CPU baseline's rows for the same sightlines,
    at the stated tolerances (relative tau where tau > 1e-3, flux absolute)."""
    import numpy as np
    tau_rel = 0.0
    flux = 0.0
    zero_ok = True
    for g, r in zip(gpu_rows, ref_rows):
        m = r > 1e-3
        if m.any():
            tau_rel = max(tau_rel, float(np.max(np.abs(g[m] - r[m]) / r[m])))
        flux = max(flux, float(np.max(np.abs(np.exp(-g) - np.exp(-r)))))
        zero_ok = zero_ok and bool(np.all(g[r == 0] == 0))
    return {"sightlines": int(gpu_rows[0].shape[0]), "lines": [ln.name for ln in w.lines],
            "tau_rel_max": tau_rel, "flux_abs_max": flux, "zero_pixels_stay_zero": zero_ok,
            "tolerance": {"tau_rel": 1e-5, "tau_floor": 1e-3, "flux_abs": 1e-6},
            "pass": tau_rel <= 1e-5 and flux <= 1e-6 and zero_ok}


def near_lines_leg(w, pos, hh, ax, cofm, with_cpu, budget_s):
    """_near_lines (py_module.cpp:25-99), the prefilter timed on its own: the
    GPU on the full, device-resident particle set (sightline grid, one
    early-exit acceptance test per particle and a compaction; median of 3
    calls) and, on the host, the reference on every k-th particle with all
    sightlines (the sample grown to about budget_s), both per particle."""
    import numpy as np
    import torch
    from fake_spectra_amd import spectra_priv as SP

    def run():
        return SP.near_lines(w.box, pos, hh, ax, cofm)

    ids = run()
    torch.cuda.synchronize()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        ids = run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    gpu_s = sorted(ts)[1]
    out = {"unit": "particles/s", "particles": w.npart, "near_particles": int(ids.numel()),
           "gpu_ms": 1e3 * gpu_s, "gpu_value": w.npart / gpu_s}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        kind = "reference" if O.ref_available() else "port"
        lib = O.load_ref() if kind == "reference" else O.load_oracle()
        n = min(w.npart, 200000)
        while True:
            k = max(1, w.npart // n)
            t0 = time.perf_counter()
            lib.near_lines(w.box, w.pos[::k], w.hh[::k], w.axis, w.cofm)
            t = time.perf_counter() - t0
            m = len(range(0, w.npart, k))
            if t >= budget_s / 3 or k == 1:
                break
            n = min(w.npart, max(2 * n, int(n * budget_s / 3 / max(t, 1e-3))))
        out["cpu_value"] = m / t
        out["cpu_kind"] = kind
        out["cpu_sample"] = "every %d-th particle (%d), all %d sightlines, %.1f s, %d threads" % (
            k, m, w.nlos, t, lib.max_threads())
        out["gpu_over_cpu"] = out["gpu_value"] / out["cpu_value"]
    return out


def drop_in_leg(w):
    """The numpy-in, numpy-out _Particle_Interpolate call (the reference's
    boundary, py_module.cpp:103-233), one per line as spectra.py:673 makes
    them: host-to-device copies of the particle arrays, binning, deposit and
    the copy of the result back to the host, all timed (one pass)."""
    from fake_spectra_amd import spectra_priv as SP
    t0 = time.perf_counter()
    for ln in w.lines:
        SP._Particle_Interpolate(*w.interp_args(ln, True))
    t = time.perf_counter() - t0
    return {"value": len(w.lines) * w.nlos * w.nbins / t, "unit": "sightline-pixels/s", "ms_per_step": 1e3 * t,
            "note": "host numpy in and out (PCIe included), one call per line"}


def host_cpu():
    """The host's CPU model and logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def adjoint_leg(steps, config=3, w=None):
    """Forward and forward+adjoint of the first line: loss
    sum (exp(-tau) - F)^2, F from the same field with T scaled by 1.1
    (by default the differentiable workload, 128^3 particles, 1024 sightlines
    x 1024 pixels, H I 1215)."""
    import numpy as np
    import torch
    from fake_spectra_amd import synth, spectra_priv as SP
    if w is None:
        w = synth.config(config)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    pos, hh, ax, cofm = dev(w.pos), dev(w.hh), dev(w.axis), dev(w.cofm)
    vel, temp, dens = dev(w.vel), dev(w.temp), dev(w.dens["H1"])
    line = w.lines[0]
    plan = SP.Plan(w.box, pos, hh, ax, cofm)
    F = torch.exp(-SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, dev(w.temp * np.float32(1.1)), [line], [dens])[0])
    plan.close()
    out = [torch.empty((w.nlos, w.nbins), dtype=torch.float64, device="cuda")]

    def fwd():
        pl = SP.Plan(w.box, pos, hh, ax, cofm)
        SP.tau(pl, w.nbins, w.kernel, w.velfac, vel, temp, [line], [dens], w.tautail, out=out)
        return pl

    def fwd_adj():
        pl = SP.Plan(w.box, pos, hh, ax, cofm)
        pl.keep_windows(True)
        t = SP.tau(pl, w.nbins, w.kernel, w.velfac, vel, temp, [line], [dens], w.tautail, out=out)[0]
        e = torch.exp(-t)
        g = (-2.0 * (e - F) * e).contiguous()      # dL/dtau
        SP.tau_adjoint(pl, w.nbins, w.kernel, w.velfac, vel, temp, [line], [dens], [g])
        return pl

    res = {}
    for name, fn in (("forward", fwd), ("forward_adjoint", fwd_adj)):
        fn().close()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn().close()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / steps
    units = w.nlos * w.nbins
    return {"workload": "c%d: %d^3 particles, %d sightlines x %d pixels, %s, kernel %s, loss sum (exp(-tau)-F)^2" % (
                config, round(w.npart ** (1 / 3)), w.nlos, w.nbins, line.name, "cubic" if w.kernel == 1 else "tophat"),
            "forward_ms": 1e3 * res["forward"], "forward_adjoint_ms": 1e3 * res["forward_adjoint"],
            "forward_value": units / res["forward"], "forward_adjoint_value": units / res["forward_adjoint"],
            "adjoint_over_forward": res["forward_adjoint"] / res["forward"], "unit": "sightline-pixels/s",
            "note": "binning included in both; fwd+adjoint records windows and runs the adjoint kernel"}


def colden_leg(w, pos, hh, ax, cofm, dens, steps):
    """Column density of one density field on the workload's sightlines,
    binning included (the _Particle_Interpolate colden call)."""
    import torch
    from fake_spectra_amd import spectra_priv as SP
    out = [torch.empty((w.nlos, w.nbins), dtype=torch.float64, device="cuda")]

    def run():
        plan = SP.Plan(w.box, pos, hh, ax, cofm)
        SP.colden(plan, w.nbins, w.kernel, w.velfac, [dens], out=out)
        plan.close()

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"value": w.nlos * w.nbins / dt, "unit": "sightline-pixels/s", "ms_per_call": 1e3 * dt,
            "note": "binning + batched colden launch"}


def flux_leg(w, tau, reps=3):
    """The flux statistics on the device (fsamd_mean_flux_scale,
    fsamd_flux_power, fsamd_res_corr, fsamd_flux_hist) over one line's
    optical depths as the deposit left them in HBM ([nlos, nbins] float64):
    milliseconds per call (host round trip included, as the reference's
    fluxstatistics.py functions return to the host) and the tau bytes read
    per millisecond."""
    import torch
    from fake_spectra_amd import fluxstatistics as GF
    vmax = w.box * w.velfac
    dvbin = vmax / w.nbins
    calls = {
        "mean_flux": lambda: GF.mean_flux(tau, 0.7),
        "flux_power": lambda: GF.flux_power(tau, vmax, mean_flux_desired=0.7),
        "res_corr": lambda: GF.res_corr(tau, dvbin, 8.0),
        "flux_pdf": lambda: GF.flux_pdf(tau, 20, 0.7),
    }
    nbytes = tau.numel() * 8
    res = {"tau_bytes": nbytes, "rows": int(tau.shape[0]), "pixels": int(tau.shape[1])}
    for name, f in calls.items():
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        res[name] = {"ms": ms, "tau_GB_per_s": nbytes / (ms * 1e-3) / 1e9}
    res["note"] = ("mean_flux: Newton-Raphson, one reduction pass per iteration; flux_power and flux_pdf "
                   "include their mean-flux rescale; flux_power is the mixed-radix FFT, two sightlines per "
                   "complex transform")
    return res


def cyc_split(st):
    """Fractions of the counting kernel's wavefront cycles in pair set-up,
    far-wing set-up and walk steps, and the share of wave iterations that set
    up at least one pair."""
    tot = st["cycles_setup"] + st["cycles_wing"] + st["cycles_walk"]
    if tot <= 0 or st["wave_iterations"] <= 0:
        return None
    return {"setup": st["cycles_setup"] / tot, "wing": st["cycles_wing"] / tot, "walk": st["cycles_walk"] / tot,
            "iterations_with_setup": st["setup_iterations"] / st["wave_iterations"],
            "wave_iterations": st["wave_iterations"]}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (sightlines are independent, so the "
              "sightline work queue would be split across GPUs). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import numpy as np
    import torch
    from fake_spectra_amd import synth, spectra_priv as SP
    from fake_spectra_amd import _native as N
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    w = synth.config(args.config, scale=args.scale)
    progress("workload c%d generated: %d particles, %d sightlines" % (args.config, w.npart, w.nlos))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    pos, hh, ax, cofm = dev(w.pos), dev(w.hh), dev(w.axis), dev(w.cofm)
    vel, temp = dev(w.vel), dev(w.temp)
    dens = [dev(w.dens_for(ln)) for ln in w.lines]
    outs = [torch.empty((w.nlos, w.nbins), dtype=torch.float64, device="cuda") for _ in w.lines]
    K = len(w.lines)
    sel, sel_k = cpu_sample(w, args.cpu_sightlines)
    torch.cuda.synchronize()

    # pick the lane-group width on this workload (kept by the plan for later calls),
    # or take it from --tune GxW (counter runs: no sample launches of the same kernel)
    if args.tune:
        best_g, best_w = (int(x) for x in args.tune.lower().split("x"))
        tune_ms = {}
    else:
        tplan = SP.Plan(w.box, pos, hh, ax, cofm)
        best_g, best_w, tune_ms = SP.autotune(tplan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens,
                                              w.tautail)
        tplan.close()
    N.set_group_size(best_g)
    N.set_block_waves(best_w)
    progress("autotune: group width %d, %d wavefronts per sightline %s" % (best_g, best_w, tune_ms))

    state = {}

    def step():
        ts = time.perf_counter()
        plan = SP.Plan(w.box, pos, hh, ax, cofm)
        SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens, w.tautail, out=outs)
        st = SP.last_stats()
        state["npairs"] = plan.npairs
        state.setdefault("ms", []).append(st["gpu_ms"])
        state.setdefault("kernel_ms", []).append(st["kernel_ms"])
        state["stats"] = st
        plan.close()
        state.setdefault("wall", []).append(time.perf_counter() - ts)   # the tau call synchronized

    # timed steps carry only hipEvents around the deposit; the counters come
    # from one more, untimed step
    SP.set_stats(True, timing_only=True)
    for _ in range(args.warmup):
        step()
    state["ms"] = []
    state["kernel_ms"] = []
    state["wall"] = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_median_ms = 1e3 * float(np.median(state["wall"]))
    dep_ms = float(np.mean(state["ms"]))              # deposit call: fills + tau kernel + FP64 fix-up
    ker_ms = float(np.mean(state["kernel_ms"]))       # the tau kernel launches alone
    progress("timed steps: %.1f ms per step" % (1e3 * elapsed / args.steps))
    # the timed output rows of the sightlines the CPU baseline runs (parity check)
    sel_t = torch.as_tensor(sel, device="cuda")
    gpu_rows = [o[sel_t].cpu().numpy() for o in outs]
    SP.set_stats(True)
    step()
    SP.set_stats(False)

    ms_per_step = 1e3 * elapsed / args.steps
    units = K * w.nlos * w.nbins
    value = units / (elapsed / args.steps)
    M = state["npairs"]
    st = state["stats"]
    # algorithmic HBM bytes of one deposit launch: per (pair, line) the pair
    # index (4 B) and the particle gather pos/vel/T/h/dens (28 B); per
    # sightline-pixel the float64 output (8 B)
    dep_bytes = 32.0 * M * K + 8.0 * units
    achieved = dep_bytes / (ker_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "tau_deposit_c%d_traffic.json" % args.config)
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    peak = C_double = None
    import ctypes
    C_double = ctypes.c_double()
    N.check(N.lib().fsamd_voigt_peak(64.0, 5e-4, ctypes.byref(C_double), ctypes.c_void_p(0)))
    peak = C_double.value
    samples_per_s = st["voigt_samples"] / (ker_ms * 1e-3)

    out = {
        "metric": "sightline-pixels/sec (tau fwd, fwd+adjoint) at stated tol, batched vs plain, 1 GPU",
        "value": value,
        "unit": "sightline-pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 profile / f64 mapping+accumulation",
        "data": "synthetic (seeded lognormal gas field)",
        "config": {"workload": "c%d: %d^3 particles, %d sightlines x %d pixels, lines %s, kernel %s, batched tau "
                               "(binning included)" % (args.config, round(w.npart ** (1 / 3)), w.nlos, w.nbins,
                                                       "+".join(ln.name for ln in w.lines),
                                                       "cubic" if w.kernel == 1 else "tophat"),
                   "global_batch": w.nlos, "seq_len": w.nbins, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "kernel": "tau_kernel (fused deposit)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes_per_launch": dep_bytes, "launch_ms": ker_ms},
        "valu": {"kernel": "tau_kernel", "achieved_samples_per_s": samples_per_s,
                 "isolated_peak_samples_per_s": peak, "frac": samples_per_s / peak,
                 "voigt_samples_per_launch": st["voigt_samples"],
                 "pixels_deposited": st["pixels"], "pixels_evaluated": st["evaluated_pixels"],
                 "fp64_tailcut_pixels": st["fp64_pixels"], "fp64_pairs": st["fp64_pairs"],
                 "kernel_sums": st["kernel_sums"], "far_wing_pixels": st["far_pixels"],
                 "wave_cycles": cyc_split(st)},
        "pairs": M,
        "step_median_ms": step_median_ms,
        "deposit_ms": dep_ms,
        "binning_ms": ms_per_step - dep_ms,
    }
    # the same deposit captured once in a HIP graph and replayed (--graph-replays 0 skips it)
    if args.graph_replays > 0:
        gplan = SP.Plan(w.box, pos, hh, ax, cofm)
        graph = SP.TauGraph(gplan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens, w.tautail)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        graph.replay()
        torch.cuda.synchronize()
        ev0.record()
        for _ in range(args.graph_replays):
            graph.replay()
        ev1.record()
        torch.cuda.synchronize()
        out["graph"] = {"deposit_ms": ev0.elapsed_time(ev1) / args.graph_replays,
                        "replays": args.graph_replays, "overflow": graph.overflowed(),
                        "note": "batched deposit + fix-up pass replayed from one captured hipGraph (binning excluded)"}
        del graph
        gplan.close()
        progress("graph replays done")
    out["autotune"] = {"group_width": best_g, "waves_per_sightline": best_w,
                       "sample_ms": {"%dx%d" % k: v for k, v in tune_ms.items()}}
    if not args.no_plain:
        # the plain form at the library's default lane-group width and block
        # size (16 lanes, 8 wavefronts), whatever the autotuner picked for the
        # batched launch, so that its time does not move with the tuning
        N.set_group_size(16)
        N.set_block_waves(8)
        plan = SP.Plan(w.box, pos, hh, ax, cofm)
        # one untimed call builds the per-sightline work items
        SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens, w.tautail, out=outs, mode=N.MODE_PLAIN)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        SP.tau(plan, w.nbins, w.kernel, w.velfac, vel, temp, w.lines, dens, w.tautail, out=outs, mode=N.MODE_PLAIN)
        torch.cuda.synchronize()
        tplain = time.perf_counter() - t0
        tbatch = dep_ms * 1e-3
        out["plain"] = {"value": units / tplain, "deposit_s": tplain,
                        "batched_over_plain": tplain / tbatch, "group_width": 16, "waves_per_sightline": 8,
                        "note": "deposit only, same binning; one launch per sightline and line, fixed "
                                "16-lane groups and 8-wavefront blocks"}
        plan.close()
        N.set_group_size(best_g)
        N.set_block_waves(best_w)
        progress("plain: %.2f s" % tplain)
    if not args.no_drop_in:
        out["drop_in"] = drop_in_leg(w)
        progress("drop-in calls done")
    if not args.no_near_lines:
        out["near_lines"] = near_lines_leg(w, pos, hh, ax, cofm, not args.no_cpu, args.cpu_seconds / 3)
        progress("near_lines done")
    if args.colden:
        out["colden"] = colden_leg(w, pos, hh, ax, cofm, dens[0], args.steps)
        progress("colden done")
    if not args.no_flux:
        out["flux"] = flux_leg(w, outs[0])
        progress("flux statistics done")
    if not args.no_adjoint:
        same = args.adjoint_config == args.config
        out["adjoint"] = adjoint_leg(max(2, args.steps), args.adjoint_config, w if same else None)
        progress("adjoint leg done")
    if not args.no_cpu:
        cb, ref_rows = cpu_baseline(w, sel, sel_k, args.cpu_seconds)
        cb["host"] = host_cpu()
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
        out["parity"] = parity_rows(w, gpu_rows, ref_rows)
        out["parity"]["against"] = "cpu_baseline rows (%s) of the same sightlines" % cb["kind"]
        progress("cpu baseline and parity done: %s" % json.dumps(out["parity"]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
