#!/usr/bin/env python
"""Benchmark: ns/day of the 95,000-atom water + ions NVE step on one MI355X.

Workload: cubic box L = 99 A, 31,000 rigid 3-site waters + 1000 Na+ +
1000 Cl-, cutoff 10 A, skin 2 A, PME 96^3 order 4, dt = 2 fs, FP32.  A step
is one full MD step through the HIP library: neighbor-list check (and
rebuild when an atom moved more than skin/2), PME excluded-pair correction,
LJ + PME direct pair forces, PME reciprocal (spread, FFT convolution,
gather), leap-frog and SETTLE.  Inputs are synthetic (seeded lattice water,
random orientations, Maxwell-Boltzmann velocities) and resident on the GPU
before the timed region.

The timed region replays the step as a HIP graph with no timers: the
reciprocal half of each step (spread, FFT convolution) runs on a second
stream beside the list and pair kernels.  A second pass of the same number
of steps replays the step with HIP events recorded on the MD stream around
the pair kernel in every replay (timers split the graph there and keep the
step on one stream), which gives the pair kernel's average launch duration
for the roofline.  A short eager pass afterwards times the other kernels for
the breakdown.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day @96k-atom water+ions; learned-vs-plain speedup at ≤1% recip RMSE"
FP32_PEAK_TFLOPS = 157.3          # MI355X FP32 vector, dual-issue packed
FP32_PEAK_TFLOPS_UNPACKED = 78.6  # without packed FP32 instructions
HBM_PEAK_GBS = 8000.0
# HBM bytes per force-only pair-kernel launch on this workload: rocprofv3
# FETCH_SIZE and WRITE_SIZE runs scaled by a 256 MiB copy (2.0 x FETCH_SIZE,
# 1.0 x WRITE_SIZE); commands and table in profiles/traffic_water_ions_95k.md
PAIR_KERNEL_TRAFFIC_BYTES = 159.5e6
# equilibration of the seeded lattice before any timed or warmup step: 100
# steps rescaled every step, then 900 rescaled every 10 steps
EQUIL_STEPS = 1000
EQUIL_STRONG = 100
TIMERS = ("lj_pairs", "nl_search", "pme_spread", "pme_gather", "pme_fft_z_r2c", "pme_fft_y",
          "pme_fft_x_conv", "pme_ifft_y", "pme_fft_z_c2r")


def pair_kernel_flops(n_in, n_list):
    """Algorithmic FP32 FLOP of one force-only pair-kernel launch: 45 per
    in-cutoff pair (distance, minimum image, LJ and erfc-Coulomb force) and
    17 per list entry beyond the cutoff (distance and minimum image)."""
    return 45.0 * n_in + 17.0 * (n_list - n_in)


def pme_hbm_rates(kernel, n_atoms, fft):
    """Effective HBM rates of the reciprocal pieces from the eager-pass timers
    (each timer spans its kernels' launches): algorithmic bytes per
    evaluation -- spread 16 N + 4 G (coordinates and charges read, grid
    written), gather 40 N + 4 G (coordinates and charges, grid, force
    read-modify-write), convolution 8 G + 28 Nfft over the five FFT passes --
    divided by the time."""
    G = int(np.prod(fft))
    nfft = fft[0] * fft[1] * (fft[2] // 2 + 1)
    out = {}
    pieces = {"spread": (16 * n_atoms + 4 * G, ["pme_spread"]),
              "gather": (40 * n_atoms + 4 * G, ["pme_gather"]),
              "fft_convolution": (8 * G + 28 * nfft, ["pme_fft_z_r2c", "pme_fft_y", "pme_fft_x_conv",
                                                      "pme_ifft_y", "pme_fft_z_c2r"])}
    for name, (nbytes, timers) in pieces.items():
        if not all(kernel.get(t, (0, 0))[1] for t in timers):
            continue
        us = sum(kernel[t][0] * 1e3 / kernel[t][1] for t in timers)
        gbs = nbytes / (us * 1e-6) / 1e9
        out[name] = dict(bytes=nbytes, us=us, GB_s=gbs, frac=gbs / HBM_PEAK_GBS)
    return out


def pair_counts(crd, counts, serial, box, cutoff, chunk=1 << 22):
    """(pairs within the cutoff, entries) of a reference-format half list:
    FP32 minimum-image distances of every entry."""
    N = serial.shape[0]
    i_idx = np.repeat(np.arange(N), counts)
    j_idx = serial[np.arange(serial.shape[1])[None, :] < counts[:, None]]
    crd = np.asarray(crd, np.float32)
    box = np.asarray(box, np.float32)
    rbox = np.float32(1.0) / box
    rc2 = np.float32(cutoff) * np.float32(cutoff)
    n_in = 0
    for c in range(0, len(i_idx), chunk):
        d = crd[j_idx[c:c + chunk]] - crd[i_idx[c:c + chunk]]
        d -= np.floor(d * rbox + np.float32(0.5)) * box
        n_in += int(((d * d).sum(1) < rc2).sum())
    return n_in, int(len(i_idx))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(s, gpu_rebuild_interval, budget_s=20.0):
    """The oracle's baseline build (the reference CPU path restated with its
    OpenMP + atomic structure, -O3 -ffast-math) on the same coordinates:
    force evaluations of the full system, list builds amortised over the
    rebuild interval the GPU run observed; every host core, then one thread,
    with the per-module split (list build, LJ + PME direct, excluded-pair
    correction, reciprocal spread / FFT convolution / gather)."""
    from oracle import oracle as O  # the CPU baseline leg only
    lib = O.lib(fast=True)
    all_threads = lib.orc_num_threads()
    beta = O.get_beta(s.cutoff, s.tolerance)

    def run(threads, budget, max_evals):
        lib.orc_set_num_threads(threads)
        t0 = time.perf_counter()
        c, ser, _ = O.nl_build(s, fast=True)
        t_nl = time.perf_counter() - t0
        pm = O.PME(s, s.fft, beta, fast=True)
        split = dict(direct=0.0, excluded=0.0, spread=0.0, fft_convolution=0.0, gather=0.0)
        n = 0
        start = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            O.lj_direct(s, c, ser, beta, need_energy=False, need_virial=False, fast=True)
            t1 = time.perf_counter()
            O.pme_excluded(s, beta, need_energy=False, need_virial=False)
            t2 = time.perf_counter()
            pm.reciprocal(need_energy=False, need_virial=False)
            rs = pm.last_seconds()
            split["direct"] += t1 - t0
            split["excluded"] += t2 - t1
            for k in ("spread", "fft_convolution", "gather"):
                split[k] += rs[k]
            n += 1
            if time.perf_counter() - start > budget or n >= max_evals:
                break
        per_eval = sum(split.values()) / n
        per_step = per_eval + t_nl / max(gpu_rebuild_interval, 1.0)
        ms = {k: v * 1e3 / n for k, v in split.items()}
        ms["list_build_per_build"] = t_nl * 1e3
        ms["list_build_amortised"] = t_nl * 1e3 / max(gpu_rebuild_interval, 1.0)
        return dict(ns_day=86.4 * s.dt_ps / per_step, ms_per_step=per_step * 1e3,
                    force_evals_per_s=1.0 / per_eval, evals=n, threads=threads, ms_split=ms)

    full = run(all_threads, budget_s, 50)
    one = run(1, 0.0, 1)
    lib.orc_set_num_threads(all_threads)
    return dict(value=full["ns_day"], unit="ns/day", cores=all_threads, kind="port",
                sample=f"{full['evals']} force evaluations (LJ + PME direct, excluded-pair "
                       f"correction, PME reciprocal with the oracle's own FFT) of the "
                       f"{s.n_atoms}-atom system on {all_threads} threads plus one list build "
                       f"amortised over {gpu_rebuild_interval:.1f} steps; one evaluation and one "
                       f"build on 1 thread",
                ms_per_step=full["ms_per_step"], force_evals_per_s=full["force_evals_per_s"],
                ms_split=full["ms_split"], cpu_model=cpu_model(), one_thread=one)


def learned_form(md, s, torch, abi, n_train=200, n_heldout=16, stride=25, epochs=30, lr=2e-3,
                 period=4, check_interval=8, bound=0.01, timed_steps=400, drift_steps=10000):
    """The learned multiple-time-step form on the same system: training
    frames from the plain engine (stride steps apart) and consecutive
    held-out frames, the surrogate trained on the GPU, its reciprocal-force
    error 1-3 steps after an exact PME step (the schedule's lags), the graph-
    replayed ms/step of the schedules, and the NVE energy drift over
    drift_steps of: the plain step, the reference's own impulse schedule (PME
    update_interval = period), the surrogate in the impulse schedule (trained
    and untrained) and the anchored schedule, all from the same state."""
    beta = abi.load().sponge_amd_get_beta(float(s.cutoff), float(s.tolerance))
    rs = abi.ReciprocalSurrogate(s.n_atoms, beta, feature_cutoff=s.cutoff,
                                 max_frames=n_train + n_heldout)
    model0 = rs.get_model()
    t0 = time.perf_counter()
    md.collect_frames(rs, n_train, stride)
    md.collect_frames(rs, n_heldout, 1)
    t_collect = time.perf_counter() - t0
    rs.fit_normalization(n_train)
    t0 = time.perf_counter()
    losses = rs.train(n_train, epochs, lr)
    t_train = time.perf_counter() - t0
    trained = rs.get_model()
    ev = {lag: rs.evaluate(n_train + 3, n_train + n_heldout, lag) for lag in (1, 2, 3)}
    crd0, vel0, _ = md.state()

    def timed(steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        md.run(steps, use_graph=True)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps

    md.set_surrogate(None)
    md.run(8, use_graph=True)
    plain_ms = timed(timed_steps)
    md.set_surrogate(rs, period, check_interval, 1e9)
    md.run(2 * period * check_interval, use_graph=True)
    learned_ms = timed(timed_steps)
    # kernel breakdown of the learned schedule (eager, timers on)
    abi.set_kernel_timing_filter(["rsur_features", "rsur_mlp"])
    abi.reset_kernel_timing()
    abi.set_kernel_timing(True)
    md.run(4 * period, use_graph=False)
    torch.cuda.synchronize()
    abi.set_kernel_timing(False)
    kt = {k: abi.kernel_timing(k) for k in ("rsur_features", "rsur_mlp")}
    abi.set_kernel_timing_filter(None)

    every = drift_steps // 10

    def drift(system, setup):
        system.set_state(crd0, vel0)
        setup(system)
        e = []
        for k in range(11):
            system.run(1, energy_interval=1)
            en = system.energies()
            e.append(en["potential"] + en["kinetic"])
            if k < 10:
                system.run(every - 1, use_graph=True)
        x = np.arange(11) * every
        slope = float(np.polyfit(x, np.asarray(e), 1)[0])
        return slope * 1000.0

    def learned(mode, model):
        def f(system):
            rs.set_model(model)
            system.set_surrogate_mode(mode)
            system.set_surrogate(rs, period, check_interval, 1e9)
        return f

    d = {"plain": drift(md, lambda m: m.set_surrogate(None)),
         "learned_impulse": drift(md, learned(1, trained)),
         "untrained_impulse": drift(md, learned(1, model0)),
         "learned_anchored": drift(md, learned(0, trained))}
    stats = md.surrogate_stats()
    md.set_surrogate(None)
    md.set_surrogate_mode(1)
    rs.set_model(trained)
    # the reference's own impulse schedule: PME every period-th step, kicked
    ref = abi.MDSystem(s, pme_opts={"update_interval": period})
    d["pme_update_interval"] = drift(ref, lambda m: None)
    ref.clear()
    # the fallback test at the stated bound: a sampled step whose anchored
    # error exceeds it takes PME's force and switches the surrogate off
    md.set_state(crd0, vel0)
    md.set_surrogate(rs, period, check_interval, bound)
    md.run(4 * period * check_interval, use_graph=True)
    fb = md.surrogate_stats()
    md.set_surrogate(None)
    rmse = {f"{lag}_step": ev[lag]["model_anchored"] for lag in (1, 2, 3)}
    return dict(
        speedup_vs_plain=plain_ms / learned_ms, ms_per_step_learned=learned_ms,
        ms_per_step_plain=plain_ms,
        recip_rel_rmse_anchored=rmse,
        recip_rel_rmse_schedule_mean=float(np.mean(list(rmse.values()))),
        meets_1pct_rmse=bool(max(rmse.values()) <= 0.01),
        ablation={f"{lag}_step": dict(stale_pme=ev[lag]["stale"],
                                      baseline_anchored=ev[lag]["baseline_anchored"],
                                      network_anchored=ev[lag]["model_anchored"])
                  for lag in (1, 2, 3)},
        network_alone_rel_rmse=ev[1]["model"], baseline_alone_rel_rmse=ev[1]["baseline"],
        energy_drift_kcal_mol_per_1000_steps=d, drift_steps=drift_steps,
        checks=stats["checks"], fallback_bound=bound,
        fallback_run=dict(checks=fb["checks"], fallbacks=fb["fallbacks"],
                          switched_off=fb["switched_off"], last_error=fb["last_error"]),
        kernel_avg_us={k: (v[0] * 1e3 / v[1] if v[1] else None) for k, v in kt.items()},
        training=dict(frames=n_train, heldout=n_heldout, stride_steps=stride, epochs=epochs,
                      lr=lr, loss_first_epoch=losses[0], loss_last_epoch=losses[-1],
                      collect_s=t_collect, train_s=t_train),
        schedule=f"impulse: PME every {period}th step (and on energy steps) with the kick "
                 f"{period} x (F_PME - F_surrogate), the surrogate's force on the other steps; "
                 f"every {check_interval}th period one surrogate step is checked against PME")


def batched_form(torch, abi, S, steps=500, steps_per_launch=10, n_boxes=48):
    """The batched form: n_boxes independent water boxes (500-6000 waters,
    log-uniform, 1.0 g/cm^3, cutoff 9 A, each with its own ~1 A PME grid and
    autotuned FFT tiling) advanced by one graph launch per steps_per_launch
    steps, against the same systems run one after another (each replaying
    its own step graph)."""
    t0 = time.perf_counter()
    boxes = S.batched_boxes(n_boxes)
    t_build = time.perf_counter() - t0
    ms = [abi.MDSystem(s) for s in boxes]
    for m in ms:
        m.equilibrate(100, temperature=300.0, every=10)
    b = abi.Batch(ms)
    b.run(1)
    tiles = b.autotune(3)
    b.run(2 * steps_per_launch, steps_per_launch)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(steps, steps_per_launch)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    # the same systems one after another, each replaying its own step graph
    for m in ms:
        m.run(4, use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in ms:
        m.run(steps, use_graph=True)
    for m in ms:
        m.energies()  # synchronises each system's stream
    t_seq = time.perf_counter() - t0
    atoms = int(sum(s.n_atoms for s in boxes))
    dt_ps = boxes[0].dt_ps
    ms_step = t_batch * 1e3 / steps
    out = dict(workload=f"{n_boxes} water boxes, {min(s.n_waters for s in boxes)}-"
                        f"{max(s.n_waters for s in boxes)} waters (log-uniform, seeded), "
                        f"1.0 g/cm^3, cutoff 9 A, skin 2 A, PME ~1 A order 4, dt 2 fs, SETTLE",
               systems=n_boxes, atoms=atoms, steps=steps, steps_per_launch=steps_per_launch,
               ms_per_step_all_systems=ms_step,
               ns_day_per_system=86.4 * dt_ps / (ms_step * 1e-3),
               aggregate_ns_day=n_boxes * 86.4 * dt_ps / (ms_step * 1e-3),
               atom_steps_per_s=atoms * steps / t_batch,
               ms_per_step_sequential=t_seq * 1e3 / steps,
               speedup_vs_sequential=t_seq / t_batch,
               fft_tilings=sorted({str(t) for t in tiles}),
               build_s=t_build)
    b.clear()
    for m in ms:
        m.clear()
    return out


def lambda_windows_form(torch, abi, S, steps=500, steps_per_launch=10):
    """The lambda-window alternative of the batched form: 16 soft-core windows
    (alpha 0.5, p 1, sigma 3 A) decoupling one 24-atom LJ / charge solute from
    2000 waters (charges first, then LJ), each an NVE system of the step
    driver in soft-core mode, all
    advanced by one batched graph launch per steps_per_launch steps, against
    the same windows run one after another (each replaying its own graph)."""
    systems, lambdas, st = S.lambda_windows()
    ms, scs = [], []
    for s, lam in zip(systems, lambdas):
        m = abi.MDSystem(s)
        sc = abi.LJSoftCore(s.n_atoms, st.AA, st.AB, st.BA, st.BB, st.typeA, st.typeB, s.cutoff,
                            alpha=st.alpha, p=st.p, sigma=st.sigma, sigma_min=0.0,
                            **{"lambda": float(lam)})
        m.set_soft_core(sc)
        m.equilibrate(100, temperature=300.0, every=10)
        ms.append(m)
        scs.append(sc)
    b = abi.Batch(ms)
    b.run(1)
    b.autotune(3)
    b.run(2 * steps_per_launch, steps_per_launch)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(steps, steps_per_launch)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    for m in ms:
        m.run(4, use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for m in ms:
        m.run(steps, use_graph=True)
    for m in ms:
        m.energies()
    t_seq = time.perf_counter() - t0
    # NVE energies of every window after the runs (finite, per window)
    pots = []
    for m in ms:
        m.run(1, energy_interval=1)
        pots.append(m.energies()["potential"])
    ms_step = t_batch * 1e3 / steps
    dt_ps = systems[0].dt_ps
    out = dict(workload=f"{len(systems)} soft-core lambda windows (alpha 0.5, p 1, sigma 3 A), "
                        f"one 24-atom LJ / charge solute in {systems[0].n_waters} waters "
                        f"({systems[0].n_atoms} atoms) decoupled in two stages (charges over "
                        f"the first 8 windows, LJ over the last 8), cutoff "
                        f"{systems[0].cutoff:g} A, PME ~1 A, dt 2 fs, SETTLE",
               systems=len(systems), atoms_per_window=systems[0].n_atoms, steps=steps,
               steps_per_launch=steps_per_launch, ms_per_step_all_windows=ms_step,
               ns_day_per_window=86.4 * dt_ps / (ms_step * 1e-3),
               ms_per_step_sequential=t_seq * 1e3 / steps,
               speedup_vs_sequential=t_seq / t_batch,
               potentials_finite=bool(all(np.isfinite(p) for p in pots)))
    b.clear()
    for m in ms:
        m.clear()
    return out


def plain_path(s, torch, abi, lib, steps, warmup, breakdown_steps, eager=False):
    """The plain MD step of one system: fixed equilibration of the seeded
    lattice (outside every timed region), FFT tiling autotune, W warmup
    steps, K timed steps (graph replay, no timers), K more steps with the
    pair-kernel timer, a short eager pass with every timer; pair statistics
    of the list at the end."""
    md = abi.MDSystem(s)
    nl_h, _, _ = md.handles()
    torch.cuda.synchronize()
    # set-up of the synthetic input (outside the timed region, independent of
    # --warmup): the seeded lattice start relaxes to the stated 300 K state
    # under velocity rescaling, every step at first (so the energy the
    # lattice releases cannot heat the box), then every 10 steps
    md.equilibrate(EQUIL_STRONG, temperature=300.0, every=1)
    md.equilibrate(EQUIL_STEPS - EQUIL_STRONG, temperature=300.0, every=10)
    fft_tiling = md.autotune(5)
    torch.cuda.synchronize()

    use_graph = not eager
    md.run(max(1, warmup), use_graph=use_graph)
    torch.cuda.synchronize()
    rb0 = C_rebuilds(lib, nl_h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    md.run(steps, use_graph=use_graph)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rb1 = C_rebuilds(lib, nl_h)

    # pair-kernel timing pass: the same number of steps, events around the
    # pair kernel in every replay
    abi.set_kernel_timing_filter(["lj_pairs"])
    abi.reset_kernel_timing()
    abi.set_kernel_timing(True)
    md.run(1, use_graph=use_graph)  # capture outside the timed pass
    torch.cuda.synchronize()
    abi.reset_kernel_timing()
    t2 = time.perf_counter()
    md.run(steps, use_graph=use_graph)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    lj_ms, lj_n = abi.kernel_timing("lj_pairs")
    abi.set_kernel_timing(False)

    # breakdown of the other kernels: a short eager pass with every timer
    kernel = {}
    if breakdown_steps > 0:
        abi.set_kernel_timing_filter(None)
        abi.reset_kernel_timing()
        abi.set_kernel_timing(True)
        md.run(breakdown_steps, use_graph=False)
        torch.cuda.synchronize()
        abi.set_kernel_timing(False)
        kernel = {k: abi.kernel_timing(k) for k in TIMERS}

    ms_per_step = (t1 - t0) * 1e3 / steps
    rebuilds = rb1 - rb0
    rebuild_interval = steps / max(rebuilds, 1)

    # pair statistics of the current list
    N = s.n_atoms
    M = 1200
    counts = torch.zeros(N, dtype=torch.int32, device="cuda")
    serial = torch.zeros(N * M, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # the MD stream is not torch's: fill first
    rc = lib.sponge_nl_export_half_list(nl_h, abi._ptr(counts), abi._ptr(serial), M)
    assert rc == 0
    crd, _, _ = md.state()
    torch.cuda.synchronize()
    n_in, n_list = pair_counts(crd, counts.cpu().numpy(), serial.view(N, M).cpu().numpy(), s.box,
                               s.cutoff)
    del counts, serial
    assert lj_n == steps, (lj_n, steps)
    lj_avg_s = lj_ms * 1e-3 / max(lj_n, 1)
    flops = pair_kernel_flops(n_in, n_list)
    achieved = flops / lj_avg_s / 1e12
    return dict(md=md, crd=crd, ms_per_step=ms_per_step, ns_day=86.4 * s.dt_ps / (ms_per_step * 1e-3),
                rebuild_interval=rebuild_interval, fft_tiling=list(fft_tiling),
                lj_avg_us=lj_avg_s * 1e6, flops=flops, achieved=achieved, n_in=n_in,
                n_list=n_list, ms_per_step_timing_pass=(t3 - t2) * 1e3 / steps,
                kernel={k: (v[0] * 1e3 / v[1] if v[1] else None) for k, v in kernel.items()},
                pme_hbm=pme_hbm_rates(kernel, N, s.fft), launch=(
                    "hip graph replay, reciprocal half on a second stream" if use_graph
                    else "eager, reciprocal half on a second stream"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--eager", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--breakdown-steps", type=int, default=100)
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned multiple-time-step form")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched form")
    ap.add_argument("--no-cfg2", action="store_true",
                    help="skip the 24,000-atom sub-line (force evaluations/s, pair-kernel VALU "
                         "fraction, spread/gather GB/s)")
    ap.add_argument("--cfg2-steps", type=int, default=1000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the hot path is "
              "measured on one GPU's share, and value is that one GPU's throughput.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    from sponge_amd import abi
    from sponge_amd import system as S

    lib = abi.load()
    s = S.config(args.config)
    r = plain_path(s, torch, abi, lib, args.steps, args.warmup, args.breakdown_steps, args.eager)
    md = r["md"]
    N = s.n_atoms
    traffic = PAIR_KERNEL_TRAFFIC_BYTES if args.config == 3 else None
    roofline = dict(bound="valu", achieved=r["achieved"], peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=r["achieved"] / FP32_PEAK_TFLOPS, traffic=traffic,
                    traffic_note="HBM bytes per launch from rocprofv3 PMC runs (copy-calibrated)",
                    kernel="lj_pairs", flop_per_launch=r["flops"], pairs_in_cutoff=r["n_in"],
                    list_entries=r["n_list"], avg_launch_us=r["lj_avg_us"],
                    frac_of_unpacked_peak=r["achieved"] / FP32_PEAK_TFLOPS_UNPACKED,
                    timing="HIP events on the MD stream around every pair-kernel launch of a "
                           "second pass of the same steps (timers split the graph at the pair "
                           "kernel and keep the step on one stream)",
                    ms_per_step_timing_pass=r["ms_per_step_timing_pass"])

    out = dict(metric=METRIC, value=r["ns_day"], unit="ns/day", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=r["ms_per_step"], higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload="water_ions_95k NVE step (31000 rigid TIP3P-like waters + "
                                    "1000 Na+ + 1000 Cl-, L=99 A, rc=10 A, skin=2 A, PME 96^3 "
                                    "order 4, dt=2 fs, SETTLE)" if args.config == 3 else s.name,
                           atoms=N, box_A=float(s.box[0]), cutoff_A=s.cutoff, skin_A=s.skin,
                           pme_grid=list(s.fft), dt_fs=s.dt_ps * 1e3, parallelism="1 GPU",
                           equilibration_steps=EQUIL_STEPS),
               roofline=roofline,
               force_evals_per_s=1e3 / r["ms_per_step"],
               list_rebuild_interval_steps=r["rebuild_interval"],
               launch=r["launch"],
               pme_fft_tiling=r["fft_tiling"],
               kernel_avg_us_eager_pass=r["kernel"],
               pme_hbm=r["pme_hbm"])
    if not args.no_learned:
        out["learned"] = learned_form(md, s, torch, abi)
    crd = r["crd"]
    md.clear()
    del r
    if not args.no_cfg2 and args.config == 3:
        s2 = S.config(2)
        r2 = plain_path(s2, torch, abi, lib, args.cfg2_steps, args.warmup, args.breakdown_steps,
                        args.eager)
        hbm = r2["pme_hbm"]
        out["cfg2"] = dict(
            workload="water_24k NVE step (8000 rigid waters, L=62.1 A, rc=10 A, skin=2 A, "
                     "PME 64^3 order 4, dt=2 fs, SETTLE)",
            atoms=s2.n_atoms, steps=args.cfg2_steps, ms_per_step=r2["ms_per_step"],
            ns_day=r2["ns_day"], force_evals_per_s=1e3 / r2["ms_per_step"],
            list_rebuild_interval_steps=r2["rebuild_interval"],
            pair_kernel=dict(avg_launch_us=r2["lj_avg_us"], achieved_TFLOP_s=r2["achieved"],
                             valu_frac_of_fp32_peak=r2["achieved"] / FP32_PEAK_TFLOPS,
                             valu_frac_of_unpacked_peak=r2["achieved"] / FP32_PEAK_TFLOPS_UNPACKED,
                             flop_per_launch=r2["flops"], pairs_in_cutoff=r2["n_in"],
                             list_entries=r2["n_list"]),
            spread_GB_s=hbm.get("spread", {}).get("GB_s"),
            gather_GB_s=hbm.get("gather", {}).get("GB_s"),
            fft_convolution_GB_s=hbm.get("fft_convolution", {}).get("GB_s"),
            pme_hbm=hbm, kernel_avg_us_eager_pass=r2["kernel"])
        r2["md"].clear()
        del r2
    if not args.no_batched:
        out["batched"] = batched_form(torch, abi, S)
        out["lambda_windows"] = lambda_windows_form(torch, abi, S)
    if not args.no_cpu_baseline:
        import dataclasses
        s_now = dataclasses.replace(s, crd=crd)
        out["cpu_baseline"] = cpu_baseline(s_now, out["list_rebuild_interval_steps"],
                                           args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


def copy_calibration(torch):
    """A plain device copy of known size, for scaling the HBM byte counters
    (FETCH_SIZE / WRITE_SIZE) to bytes: 256 MiB read + 256 MiB written per
    launch, 5 launches."""
    n = 64 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                      "launches": 5}))


def C_rebuilds(lib, nl_h):
    import ctypes
    n = ctypes.c_longlong()
    rc = lib.sponge_nl_rebuild_count(nl_h, ctypes.byref(n))
    assert rc == 0
    return n.value


if __name__ == "__main__":
    main()
