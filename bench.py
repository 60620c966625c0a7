"""Benchmark: pair-interactions/s of the retarded-interaction step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload n65536]

One step = one full predictor-corrector timestep of every dot (pair sum,
pulse, predictor, 10 corrector iterations, Bloch RHS), as Integrator::solve
does it (src/integrator.cpp:120-185).  The warm-up is at least the steps
that fill the ring's active window with computed history (the largest delay
+ order + 2), so the timed steps run in a long run's steady state; with a
method that forms K steps per pass, the warm-up and the timed steps are
rounded up to multiples of K, so the timed region holds whole passes.
`warmup` and `steps` are the numbers run, `warmup_requested` and
`steps_requested` the numbers asked for.  The pair sums use the time-blocked
method by default (QUEST_PAIR_TIME_BLOCKED: one pass per 8 or 16 steps over
the pairs whose delay is at least that long, list kernels for the shorter
delays); the per-step method's figures follow in `per_step_method`.  The
roofline's flops are the pass's own algorithmic count (flop_per_pair_step);
`per_step_formulation` credits the same launches with the per-step
formulation's 105 flop per pair-step instead.  The workload is N = 65536 dots at
512 um^-3 with random dipoles, 2 rad/ps Gaussian detuning and a pi/2 pulse
(dt = 1e-4 ps, order-4 interpolation, 2000-step run); `value` is the
mixed-precision path (FP32 pair arithmetic, FP64 accumulation and FP64 state)
and the same run in plain FP64 is reported beside it, with the max |d rho|
between the two.  Inputs live in HBM before the timed region starts.  The
timed steps are launched kernel by kernel with HIP events on the context's
stream around every pair-kernel launch, so the roofline's kernel duration is
the mean over the timed region itself; the same number of steps replayed as
one captured graph per step (quest_run's default) follows, reported as
graph_ms_per_step.

After the timed runs, `driven` compares mixed with FP64 on the same system
in a driven state (100 steps under a short, strong pi/2 pulse, |rho01| up to
~0.5), against the metric's 1e-6 bound.

The CPU baseline is the FP64 oracle's reference method (pair table plus 11
scatter sums per step, history_interaction.cpp:47-70 + integrator.cpp:129-140)
on one host core, over a bounded sample: 3 steps of the N = 4096 system at
steady state (past its largest delay, every tap summed; the table of N = 65536
would need 180 GB).  `cpu_c0` is the whole N = 64, 20000-step run of config 0
on one core, `cpu_best` the split gather on up to 16 threads.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F_ALG = 105            # flop per ordered pair per step (algorithmic count of the per-step formulation)
# Time-blocked formulation with K steps per pass, per ordered pair: once per
# pass 158 flop (geometry, dipole products, radial factors, the rotation
# e^{-i theta (n + x)}, the complex q0/q1/q2 coefficients, the weights on the
# M + 1 = 5 samples through the monomial Lagrange basis on centred nodes: 11
# real-by-complex terms and 4 complex mirror sums) plus the sums wr + wi of
# the 5 weights and x + y of the K + 4 window samples; then 30 flop per step
# (5 complex multiply-adds of three real products each): 30 + (167 + K) / K
# per pair-step, 41.4 at K = 16 (with four real products per complex
# multiply-add and the 21 terms of the one-sided basis it would be
# 40 + 190 / K = 51.9).
F_BLK_PAIR = 167
F_BLK_STEP = 30


def flop_per_pair_step(steps_per_launch: int) -> float:
    """Algorithmic flop per ordered pair and step of the pair kernel that
    forms `steps_per_launch` steps per launch (1: the per-step formulation)."""
    k = steps_per_launch
    return float(F_ALG) if k <= 1 else F_BLK_STEP + (F_BLK_PAIR + k) / k
FP32_PEAK = 157.3      # TFLOP/s, MI355X vector FP32 (datasheet)
FP64_PEAK = 78.6       # TFLOP/s, MI355X vector FP64
# TFLOP/s: non-packed FP32 FMAs issue at one 64-lane instruction per cycle per
# CU (256 CUs x 2.4 GHz x 64 lanes x 2 flop), the same rate as FP64; the
# 157.3 datasheet figure needs packed FP32, which this build does not use
FMA_CEILING = 78.6


def fill_warmup(ctx, order=4):
    """Warm-up steps that fill the ring's active window with computed history:
    every pair then reads a slot written by the run, as in a long simulation.
    The pair kernels take the same cycles either way, but with slots still at
    their cleared zeros the GPU holds a higher clock under its power limit
    (measured at N = 65536: 2.38 GHz with an all-zero ring, ~2.07 GHz once
    it is populated; 6.5 against 7.5 ms per mixed launch)."""
    warm = max(10, int(ctx.info()["max_delay_bound"]) + order + 2)
    return (warm + 15) // 16 * 16  # the time-blocked method's passes start at multiples of 8 or 16


def timed_steps(ctx, warmup, steps):
    """Warm-up (at least fill_warmup), then the timed region: `steps` steps
    launched kernel by kernel with HIP events on the context's stream around
    every pair-kernel launch (wall time and the pair kernel's mean duration
    over the same steps); then `steps` steps replayed as one captured graph
    per step (captured outside the timed loop)."""
    # At least the ring-filling warm-up, whatever was asked for: with fewer
    # steps most pairs still read the zero prehistory and the GPU holds a
    # higher clock than in the steady state of a run (same instructions).
    warmup = max(warmup, fill_warmup(ctx))
    # The time-blocked method forms the pair sums of K steps in one pass launched
    # at multiples of K: the timed region starts on a pass boundary and holds
    # whole passes (the requested steps rounded up to a multiple of K), so it
    # pays for every step it times and for no step beyond them.
    spl = max(1, int(ctx.info()["block_steps"]))
    warmup = -(-warmup // spl) * spl
    steps = -(-steps // spl) * spl
    if warmup:
        ctx.run(warmup)
    ctx.sync()
    ctx.set_timing(True)
    ctx.pair_kernel_time(reset=True)
    ctx.short_kernel_time(reset=True)
    t0 = time.perf_counter()
    ctx.run(steps)              # returns after the stream has drained
    ctx.sync()
    secs = time.perf_counter() - t0
    kms, launches = ctx.pair_kernel_time()
    sms, slaunches = ctx.short_kernel_time()
    ctx.set_timing(False)
    ctx.run(16)                 # captures the step graph (one step, or the 8 or 16 steps of a time-blocked pass)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.run(steps)
    ctx.sync()
    graph_secs = time.perf_counter() - t0
    return {"seconds": secs, "pair_ms": kms / max(launches, 1), "launches": launches,
            "short_ms": sms / max(slaunches, 1), "short_launches": slaunches,
            "steps_per_launch": spl, "steps": steps,
            "graph_ms_per_step": 1e3 * graph_secs / max(steps, 1), "warmup": warmup}


def run_gpu(sysd, consts, pulse, precision, warmup, steps, weights, method=0, tune=False):
    """One context on the workload; with tune (not the table method) the work
    split is the autotuner's (quest_autotune, 3 launches per candidate, before
    the warm-up), else the library's default; see timed_steps."""
    from quest_amd import PAIR_TABLE, Context
    ps, cs, fc = weights
    with Context(sysd, consts, precision=precision, pulse=pulse) as ctx:
        if method:
            ctx.set_pair_method(method)
        ctx.set_pc_weights(ps, cs, fc)
        if tune and method != PAIR_TABLE:
            ctx.autotune(3)
        info = ctx.info()
        out = timed_steps(ctx, warmup, steps)
        rho, _ = ctx.state()
    return dict(out, rho=rho, info=info)


def run_ensemble(ens, consts, precision, warmup, steps, weights, method=0):
    """The batched workload: all systems in one context, work split autotuned
    before the warm-up; see timed_steps."""
    from quest_amd import Context
    ps, cs, fc = weights
    with Context([s for s, _ in ens], consts, precision=precision, pulse=[p for _, p in ens]) as ctx:
        ctx.set_pc_weights(ps, cs, fc)
        if method:
            ctx.set_pair_method(method)
        chunk = ctx.autotune(3)
        out = timed_steps(ctx, warmup, steps)
        info = ctx.info()
        rho, _ = ctx.state()
    return dict(out, rho=rho, info=info, chunk=chunk)


def ensemble_main(args):
    from quest_amd import pc_weights, systems
    ens = systems.ensemble(256)
    consts = systems.consts()
    weights = pc_weights(18, 22, 3.15)
    sizes = np.array([len(s["pos"]) for s, _ in ens], dtype=np.float64)
    pairs = float((sizes * sizes).sum())
    ordered = float((sizes * (sizes - 1)).sum())
    method = 2 if args.pair_method == "time-blocked" else 0
    mixed = run_ensemble(ens, consts, 1, args.warmup, args.steps, weights, method)
    fp64 = None if args.no_fp64 else run_ensemble(ens, consts, 0, args.warmup, args.steps, weights, method)
    value = pairs * mixed["steps"] / mixed["seconds"]
    achieved = (flop_per_pair_step(mixed["steps_per_launch"]) * ordered * mixed["steps_per_launch"]
                / (mixed["pair_ms"] * 1e-3) / 1e12)
    out = {
        "metric": "pair-interactions/s (sum of N^2 * steps / s), 256 batched systems",
        "value": value, "unit": "pair-interactions/s", "n_gpus": 1, "steps": mixed["steps"],
        "steps_requested": args.steps, "warmup_requested": args.warmup,
        "warmup": mixed["warmup"], "ms_per_step": 1e3 * mixed["seconds"] / mixed["steps"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair arithmetic + f64 accumulation/state", "data": "synthetic",
        "config": {"workload": "ensemble: 256 systems, N log-uniform in [200, 2000] at 512 um^-3, "
                               "random dipoles, 2 rad/ps detuning, pulse area uniform in [pi/4, pi], "
                               "dt=1e-4 ps, order 4, 5000-step run",
                   "systems": 256, "dots": int(sizes.sum()), "pairs_per_step": pairs,
                   "pair_method": (f"time-blocked ({mixed['steps_per_launch']} steps per pass)" if method
                                   else "on-the-fly"),
                   "source_chunk_autotuned": mixed["chunk"], "work_items": mixed["info"]["work_items"],
                   "persistent_grid": mixed["info"]["grid"],
                   "dynamic_items": mixed["info"]["dynamic_items"]},
        "graph_ms_per_step": mixed["graph_ms_per_step"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK, "traffic": None, "fma_issue_ceiling": FMA_CEILING,
                     "frac_of_fma_issue_ceiling": achieved / FMA_CEILING,
                     "kernel": kernel_name(mixed["info"], 1, method),
                     "kernel_ms": mixed["pair_ms"], "steps_per_launch": mixed["steps_per_launch"],
                     "flop_per_pair": flop_per_pair_step(mixed["steps_per_launch"])},
    }
    if method:
        out["short_list"] = {"kernel": kernel_name(mixed["info"], 1, method, short=True),
                             "us_per_step": 1e3 * mixed["short_ms"], "pairs": mixed["info"]["short_pairs"]}
    if fp64 is not None:
        v64 = pairs * fp64["steps"] / fp64["seconds"]
        out["fp64"] = {"value": v64, "ms_per_step": 1e3 * fp64["seconds"] / fp64["steps"],
                       "steps": fp64["steps"], "graph_ms_per_step": fp64["graph_ms_per_step"],
                       "pair_kernel_ms": fp64["pair_ms"], "steps_per_launch": fp64["steps_per_launch"],
                       "source_chunk_autotuned": fp64["chunk"]}
        out["mixed_over_fp64"] = value / v64
        out["max_abs_drho_vs_fp64"] = float(np.abs(mixed["rho"] - fp64["rho"]).max())
        out["max_abs_rho01"] = float(np.abs(fp64["rho"][:, 1]).max())
        out["rel_rho01_err_vs_fp64"] = rho01_rel_err(mixed["rho"], fp64["rho"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_ensemble(ens)
    print(json.dumps(out))
    return 0


def table_main(args):
    """FP64 coefficient-table method (QUEST_PAIR_TABLE: the reference's
    precomputed per-pair coefficients streamed every step) beside the FP64
    on-the-fly path on the same workload.  The table kernel is HBM-bound: its
    roofline is the table bytes it streams per launch, (M + 1) x 16 B + 2 B per
    ordered pair, over the kernel's mean duration, against 8 TB/s."""
    from quest_amd import PAIR_TABLE, pc_weights, systems
    sysd, full_steps = systems.workload(args.workload)
    consts = systems.consts()
    pulse = systems.pulse(math.pi / 2)
    weights = pc_weights(18, 22, 3.15)
    n = sysd["pos"].shape[0]
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
    tab = run_gpu(sysd, consts, pulse, 0, args.warmup, args.steps, weights, method=PAIR_TABLE)
    fly = run_gpu(sysd, consts, pulse, 0, args.warmup, args.steps, weights)
    pairs = float(n) * n
    value = pairs * args.steps / tab["seconds"]
    table_bytes = pairs * (5 * 16 + 2)
    achieved = table_bytes / (tab["pair_ms"] * 1e-3) / 1e9
    out = {
        "metric": "pair-interactions/s (N²·steps/s), FP64 coefficient-table method",
        "value": value, "unit": "pair-interactions/s", "n_gpus": 1, "steps": args.steps,
        "warmup": tab["warmup"], "ms_per_step": 1e3 * tab["seconds"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload}: N={n} dots, table of {table_bytes / 1e9:.2f} GB, "
                               f"{full_steps}-step run", "n_dots": n},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0,
                     "traffic": (profiled_traffic("pair_table_f64<4, false>", "pair_table_n16384.md")
                                 if args.workload == "n16384" else None),
                     "kernel": "pair_table_f64<4,false>",
                     "kernel_ms": tab["pair_ms"], "bytes_per_pair": 82},
        "on_the_fly_fp64": {"value": pairs * args.steps / fly["seconds"], "pair_kernel_ms": fly["pair_ms"]},
        "rel_rho_table_vs_on_the_fly": float(np.abs(tab["rho"] - fly["rho"]).max() / np.abs(fly["rho"]).max()),
    }
    print(json.dumps(out))
    return 0


def rho01_rel_err(mixed_rho, fp64_rho) -> float:
    """max |rho01(mixed) - rho01(fp64)| / max |rho01(fp64)| over all dots."""
    scale = float(np.abs(fp64_rho[:, 1]).max())
    return float(np.abs(mixed_rho[:, 1] - fp64_rho[:, 1]).max()) / scale if scale > 0 else 0.0


def copy_probe(nbytes: int) -> None:
    """One device-to-device copy of nbytes (the runtime's copy kernel): the
    calibration point for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(nbytes)) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) or hip.hipDeviceSynchronize():  # DeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


PAIR_PMC_PROFILES = ("pair_block_n65536_pmc.md", "pair_kernels_n65536_pmc.md")


def profiled_traffic(kernel: str, profile: str | None = None):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 summary in
    profiles/ (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration):
    the traffic_bytes column of a pair-kernel table (the N = 65536 summaries
    of both pair methods, unless `profile` names one), or HBM read + write
    bytes of the table-kernel one; None if absent."""
    for prof in ((profile,) if profile else PAIR_PMC_PROFILES):
        path = os.path.join(ROOT, "profiles", prof)
        if not os.path.exists(path):
            continue
        for line in open(path):
            if line.startswith("| `" + kernel + "`"):
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                try:
                    if prof in PAIR_PMC_PROFILES:
                        return float(cells[-1])
                    if len(cells) == 7:  # the traffic table (the trace table above it has 5 columns)
                        return float(cells[2]) + float(cells[3])
                except (ValueError, IndexError):
                    return None
    return None


def _oracle_setup():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    import pc_weights
    from quest_amd import systems
    oracle.build()
    c = systems.consts()
    p = oracle.params(c["c0"], c["hbar"], c["mu0"], c["laser_omega"], c["dt"], 4, 22)
    return oracle, p, pc_weights.weights(), c


def _qo_pulse(oracle, pl):
    return oracle.pulse(pl["amplitude"], pl["delay"], pl["width"], pl["freq"], pl["wavevector"],
                        pl["polarization"])


def cpu_baseline(steps=3):
    """The reference method (pair table + 11 scatter sums per step,
    history_interaction.cpp:47-70 + integrator.cpp:129-140) on one host core,
    timed at steady state: the timed steps start past the largest delay plus
    the order, on a non-zero synthetic prehistory, so every tap of every pair
    is summed (the reference skips taps before its prehistory only in the
    first ~D_max steps)."""
    from quest_amd import systems
    oracle, p, (ps, cs, fc), c = _oracle_setup()
    oracle.set_threads(1)
    sysd, _ = systems.workload("n4096")
    n = sysd["pos"].shape[0]
    start = oracle.steady_start(sysd["pos"], c["c0"], c["dt"], 4)
    r = oracle.time_steps(p, sysd, _qo_pulse(oracle, systems.pulse(math.pi / 2)), ps, cs, fc, steps,
                          start_step=start, mode=0)
    full = r["taps"] == 11 * steps * (n * (n - 1) // 2) * 5 * 2
    return {"value": n * n * steps / r["step_seconds"], "unit": "pair-interactions/s", "cores": 1,
            "kind": "port",
            "sample": (f"FP64 oracle, reference method (pair table + 11 scatter sums/step), N={n} system "
                       f"of the same generator, steps {start}-{start + steps - 1} after a synthetic "
                       f"non-zero prehistory ({'every' if full else 'NOT every'} tap summed: "
                       f"{r['taps']} pair-tap terms), {r['step_seconds']:.2f} s on 1 core (table build "
                       f"{r['setup_seconds']:.1f} s excluded); N=65536 needs a 180 GB table")}


def cpu_c0():
    """Config 0 as the reference runs it: N = 64, all 20000 steps from the
    ground state through the pi/2 pulse, reference method, one core."""
    from quest_amd import systems
    oracle, p, (ps, cs, fc), c = _oracle_setup()
    oracle.set_threads(1)
    sysd, steps = systems.workload("n64")
    r = oracle.solve(p, sysd, _qo_pulse(oracle, systems.pulse(math.pi / 2)), ps, cs, fc, steps, mode=0)
    n = sysd["pos"].shape[0]
    return {"value": n * n * steps / r["seconds"], "unit": "pair-interactions/s", "cores": 1, "kind": "port",
            "seconds": r["seconds"],
            "sample": (f"FP64 oracle, reference method, N={n} (0.5 um cube), the whole {steps}-step run "
                       f"from the ground state in {r['seconds']:.1f} s on 1 core (table build included)")}


def cpu_baseline_ensemble(ens, steps=3, every=32):
    """The reference method on one host core at steady state (see
    cpu_baseline), timed per system and summed over a seeded sample of the
    ensemble (every 32nd system, the ensemble's own seeds and pulses): sum of
    N^2 x steps over the summed step seconds."""
    oracle, p, (ps, cs, fc), c = _oracle_setup()
    oracle.set_threads(1)
    pairs = secs = setup = 0.0
    full = True
    sample = ens[::every]
    for sysd, pl in sample:
        n = sysd["pos"].shape[0]
        start = oracle.steady_start(sysd["pos"], c["c0"], c["dt"], 4)
        r = oracle.time_steps(p, sysd, _qo_pulse(oracle, pl), ps, cs, fc, steps, start_step=start, mode=0)
        full = full and r["taps"] == 11 * steps * (n * (n - 1) // 2) * 5 * 2
        pairs += float(n * n * steps)
        secs += r["step_seconds"]
        setup += r["setup_seconds"]
    sizes = [s["pos"].shape[0] for s, _ in sample]
    return {"value": pairs / secs, "unit": "pair-interactions/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 oracle, reference method (pair table + 11 scatter sums/step), "
                       f"{len(sample)} of the 256 systems (every {every}nd, N {min(sizes)}-{max(sizes)}), "
                       f"{steps} steps each past its largest delay after a synthetic prehistory "
                       f"({'every' if full else 'NOT every'} tap summed), {secs:.2f} s of steps on 1 core "
                       f"(table builds {setup:.1f} s excluded), timed per system and summed")}


def cpu_best(steps=10):
    """The oracle's split gather (one pair sum per step plus the zero-delay
    pairs per evaluation, the arithmetic the GPU path uses) on up to 16 host
    threads (the host cores available per GPU), at steady state like
    cpu_baseline: the best CPU figure beside the serial reference method."""
    from quest_amd import systems
    oracle, p, (ps, cs, fc), c = _oracle_setup()
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(threads)
    sysd, _ = systems.workload("n4096")
    start = oracle.steady_start(sysd["pos"], c["c0"], c["dt"], 4)
    r = oracle.time_steps(p, sysd, _qo_pulse(oracle, systems.pulse(math.pi / 2)), ps, cs, fc, steps,
                          start_step=start, mode=1)
    oracle.set_threads(1)
    n = sysd["pos"].shape[0]
    return {"value": n * n * steps / r["step_seconds"], "unit": "pair-interactions/s", "cores": threads,
            "kind": "port",
            "sample": (f"FP64 oracle, split gather (OpenMP), N={n} system of the same generator, steps "
                       f"{start}-{start + steps - 1} after a synthetic prehistory, {r['step_seconds']:.2f} s "
                       f"on {threads} threads (setup excluded)")}


def kernel_name(info, precision, method, short=False):
    """The pair kernel (or the time-blocked method's per-step short-list kernel) a
    context of the bench's workload launches (order 4, rotating frame, short
    rotation series)."""
    real = "double" if precision == 0 else "float"
    dyn = "true" if info["dynamic_items"] else "false"
    if method == 2:
        if short:
            return f"pair_list<{real}, 4, 1, false, false>"
        return f"pair_block<{real}, 4, {int(info.get('block_steps') or 8)}, false, {dyn}, false>"
    if precision == 0:
        return f"pair_past_f64<4, false, {dyn}, false>"
    return "pair_past_f32<4, false, {}, false, {}, false>".format(dyn, "true" if info["pipelined"] else "false")


def driven_error(sysd, consts, weights, steps=100, method=0):
    """Mixed against FP64 in a driven state: the same system under a short,
    strong pulse (pi/2 area, sigma 15 steps, centred at step 50), so that
    |rho01| is large within the 100 steps; rho01 error relative to the FP64
    run's max |rho01|, rho00 error absolute."""
    from quest_amd import Context
    amp = (math.pi / 2) * 0.65821193 / (5.2917721e-4 * 0.0015 * math.sqrt(2 * math.pi))
    pl = {"amplitude": amp, "delay": 0.005, "width": consts["laser_omega"] * 0.0015,
          "freq": consts["laser_omega"], "wavevector": [consts["laser_omega"] / consts["c0"], 0.0, 0.0],
          "polarization": [0.0, 0.0, 1.0]}
    rho = {}
    for precision in (0, 1):
        with Context(sysd, consts, precision=precision, pulse=pl) as ctx:
            ctx.set_pc_weights(*weights)
            if method:
                ctx.set_pair_method(method)
            ctx.run(steps)
            rho[precision], _ = ctx.state()
    err01 = rho01_rel_err(rho[1], rho[0])
    err00 = float(np.abs(rho[1][:, 0] - rho[0][:, 0]).max())
    return {"steps": steps, "pulse": "pi/2 area, sigma 0.0015 ps, delay 0.005 ps",
            "max_abs_rho01": float(np.abs(rho[0][:, 1]).max()), "rel_rho01_err_vs_fp64": err01,
            "max_abs_drho00_vs_fp64": err00, "bound": 1e-6, "within_bound": bool(err01 <= 1e-6 and err00 <= 1e-6)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=192,
                    help="timed steps (default 192: whole passes of the time-blocked method)")
    ap.add_argument("--warmup", type=int, default=-1,
                    help="untimed steps before the timed ones, at least (and by default) enough to fill "
                         "the ring's active window with computed history (max delay + order + 2 steps)")
    ap.add_argument("--workload", default="n65536")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--zero-drive", action="store_true",
                    help="diagnostic: the pulse amplitude set to 0, so rho01 and every interpolant-ring "
                         "slot stay exactly 0 (the same kernels and cycles; shows the GPU clock's "
                         "dependence on the data under its power limit)")
    ap.add_argument("--tune", action="store_true",
                    help="quest_autotune the work split before the warm-up (default: the library's "
                         "default split; measured within 1 %% of the tuned one in sustained runs at N = 65536)")
    ap.add_argument("--no-driven", action="store_true",
                    help="skip the driven-state (short strong pulse) mixed-vs-FP64 error runs")
    ap.add_argument("--pair-method", choices=("time-blocked", "on-the-fly", "table"), default="time-blocked",
                    help="time-blocked (default): one pass per 8 or 16 steps (QUEST_PAIR_TIME_BLOCKED), with the "
                         "per-step on-the-fly method reported beside it; on-the-fly: one pass per step; "
                         "table: the FP64 coefficient-table method beside FP64 on the fly")
    ap.add_argument("--no-per-step", action="store_true",
                    help="time-blocked: skip the per-step method's comparison runs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        return 2

    if args.workload == "ensemble":
        return ensemble_main(args)
    if args.pair_method == "table":
        return table_main(args)
    from quest_amd import PAIR_ON_THE_FLY, PAIR_TIME_BLOCKED, pc_weights, systems
    sysd, full_steps = systems.workload(args.workload)
    consts = systems.consts()
    pulse = systems.pulse(math.pi / 2)
    if args.zero_drive:
        pulse = dict(pulse, amplitude=0.0)
    weights = pc_weights(18, 22, 3.15)
    n = sysd["pos"].shape[0]

    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
    tune = args.tune
    method = PAIR_TIME_BLOCKED if args.pair_method == "time-blocked" else PAIR_ON_THE_FLY
    mixed = run_gpu(sysd, consts, pulse, 1, args.warmup, args.steps, weights, method=method, tune=tune)
    fp64 = (None if args.no_fp64 else
            run_gpu(sysd, consts, pulse, 0, args.warmup, args.steps, weights, method=method, tune=tune))

    pairs = float(n) * n
    value = pairs * mixed["steps"] / mixed["seconds"]
    ms_per_step = 1e3 * mixed["seconds"] / mixed["steps"]
    ordered = float(n) * (n - 1)
    # algorithmic flops of one launch: the formulation's flop per ordered pair
    # and step, times the steps one launch forms
    spl = mixed["steps_per_launch"]
    f_pair = flop_per_pair_step(spl)
    achieved = f_pair * ordered * spl / (mixed["pair_ms"] * 1e-3) / 1e12
    per_step_units = F_ALG * ordered * spl / (mixed["pair_ms"] * 1e-3) / 1e12
    kname = kernel_name(mixed["info"], 1, method)
    out = {
        "metric": f"pair-interactions/s (N²·steps/s), mixed vs FP64 on GPU at N={n}, ρ err≤1e-6",
        "value": value, "unit": "pair-interactions/s", "n_gpus": 1, "steps": mixed["steps"],
        "steps_requested": args.steps, "warmup_requested": args.warmup,
        "warmup": mixed["warmup"], "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair arithmetic + f64 accumulation/state",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: N={n} dots at 512 um^-3, random dipoles, "
                               f"2 rad/ps detuning, pi/2 pulse, dt=1e-4 ps, order 4, "
                               f"{full_steps}-step run", "n_dots": n,
                   "pair_method": ("time-blocked: one pass per {0} steps over the pairs of delay >= {0}, "
                                   "lists for delays 1-3 (every step), 4-7 (4-step windows){1}".format(
                                       mixed["steps_per_launch"],
                                       ", 8-15 (8-step windows)" if mixed["steps_per_launch"] == 16 else "")
                                   if method == PAIR_TIME_BLOCKED else "on-the-fly, one pass per step"),
                   "source_chunk": mixed["info"]["source_chunk"],
                   "source_chunks_per_tile": mixed["info"]["splits"],
                   "persistent_grid": mixed["info"]["grid"],
                   "pipelined_loop": mixed["info"]["pipelined"],
                   "dynamic_items": mixed["info"]["dynamic_items"],
                   "autotuned": tune,
                   "drive": "none: pulse amplitude 0, every ring slot holds zeros" if args.zero_drive
                            else "pi/2 pulse",
                   "ring_slots": mixed["info"]["ring_slots"],
                   "zero_delay_pairs": mixed["info"]["near_pairs"],
                   "short_list_pairs": mixed["info"]["short_pairs"]},
        "graph_ms_per_step": mixed["graph_ms_per_step"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK,
                     "traffic": profiled_traffic(kname) if args.workload == "n65536" else None,
                     "kernel": kname, "kernel_ms": mixed["pair_ms"],
                     "steps_per_launch": spl,
                     "kernel_ms_per_step": mixed["pair_ms"] / spl,
                     "flop_per_pair": f_pair, "fma_issue_ceiling": FMA_CEILING,
                     "frac_of_fma_issue_ceiling": achieved / FMA_CEILING,
                     # the same launches credited with the per-step formulation's 105 flop per
                     # pair-step: a throughput in that formulation's units, not a use of the hardware
                     "per_step_formulation": {"flop_per_pair": F_ALG, "achieved": per_step_units,
                                              "frac_of_peak": per_step_units / FP32_PEAK}},
    }
    if method == PAIR_TIME_BLOCKED:
        out["short_list"] = {"kernel": kernel_name(mixed["info"], 1, method, short=True),
                             "us_per_step": 1e3 * mixed["short_ms"], "pairs": mixed["info"]["short_pairs"]}
    if fp64 is not None:
        v64 = pairs * fp64["steps"] / fp64["seconds"]
        a64 = (flop_per_pair_step(fp64["steps_per_launch"]) * ordered * fp64["steps_per_launch"]
               / (fp64["pair_ms"] * 1e-3) / 1e12)
        out["fp64"] = {"value": v64, "ms_per_step": 1e3 * fp64["seconds"] / fp64["steps"],
                       "steps": fp64["steps"],
                       "graph_ms_per_step": fp64["graph_ms_per_step"],
                       "pair_kernel": kernel_name(fp64["info"], 0, method),
                       "pair_kernel_ms": fp64["pair_ms"],
                       "pair_kernel_ms_per_step": fp64["pair_ms"] / fp64["steps_per_launch"],
                       "achieved_tflops": a64, "frac_of_fp64_peak": a64 / FP64_PEAK,
                       "short_list_us_per_step": 1e3 * fp64["short_ms"],
                       "source_chunk": fp64["info"]["source_chunk"], "persistent_grid": fp64["info"]["grid"],
                       "dynamic_items": fp64["info"]["dynamic_items"]}
        out["mixed_over_fp64"] = value / v64
        out["max_abs_drho_vs_fp64"] = float(np.abs(mixed["rho"] - fp64["rho"]).max())
        out["max_abs_rho01"] = float(np.abs(fp64["rho"][:, 1]).max())
        out["rel_rho01_err_vs_fp64"] = rho01_rel_err(mixed["rho"], fp64["rho"])
        if n > 4096:
            # the same figures over a seeded 4096-dot subsample of the full-system runs
            sub = np.sort(np.random.default_rng(8).choice(n, size=4096, replace=False))
            out["subsample_4096"] = {
                "seed": 8, "max_abs_drho_vs_fp64": float(np.abs(mixed["rho"][sub] - fp64["rho"][sub]).max()),
                "rel_rho01_err_vs_fp64": rho01_rel_err(mixed["rho"][sub], fp64["rho"][sub])}
        if not args.no_driven:
            out["driven"] = driven_error(sysd, consts, weights, method=method)
    if method == PAIR_TIME_BLOCKED and not args.no_per_step:
        # the per-step on-the-fly method on the same workload, for comparison
        out["per_step_method"] = {}
        for prec, key in ((1, "mixed"), (0, "fp64")):
            if prec == 0 and args.no_fp64:
                continue
            r = run_gpu(sysd, consts, pulse, prec, args.warmup, args.steps, weights, tune=tune)
            a = F_ALG * ordered / (r["pair_ms"] * 1e-3) / 1e12
            out["per_step_method"][key] = {
                "value": pairs * r["steps"] / r["seconds"], "ms_per_step": 1e3 * r["seconds"] / r["steps"],
                "pair_kernel": kernel_name(r["info"], prec, PAIR_ON_THE_FLY), "pair_kernel_ms": r["pair_ms"],
                "achieved_tflops": a, "rel_rho01_vs_time_blocked_fp64":
                    rho01_rel_err(r["rho"], fp64["rho"]) if fp64 is not None else None}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_best"] = cpu_best()
        out["cpu_c0"] = cpu_c0()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
