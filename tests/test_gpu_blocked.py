"""The time-blocked pair method (QUEST_PAIR_TIME_BLOCKED) against the oracle
and against the per-step on-the-fly method.

One pass at every step that is a multiple of K = 8 (or 16) forms the past sums
of the pairs of delay >= K for the next K steps, lists the pairs of delay 1..3
step by step, 4..7 in 4-step windows and (K = 16) 8..15 in 8-step windows; the
sums are the reference's (history_interaction.cpp:47-70)
in another order, so FP64 runs meet the oracle at 1e-10 and mixed runs at
1e-6, as the per-step method does.
"""
import math
import os

import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse

pytestmark = pytest.mark.gpu


def _ctx(sysd, consts, weights, precision=0, dyadic=0, order=4, pulse=None, method=2, **kw):
    from quest_amd import Context
    ps, cs, fc = weights
    ctx = Context(sysd, consts, precision=precision, dyadic=dyadic, order=order, pulse=pulse, **kw)
    ctx.set_pc_weights(ps, cs, fc)
    ctx.set_pair_method(method)
    return ctx


PASS_STEPS = pytest.mark.parametrize("block_steps", [8, 16], ids=["K8", "K16"])


@PASS_STEPS
@pytest.mark.parametrize("order", [2, 4, 6])
@pytest.mark.parametrize("precision,dyadic", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_blocked_run_matches_oracle(oracle_lib, ref_weights, order, precision, dyadic, block_steps):
    """A dense system whose delays (0 .. ~20 steps) fill the zero-delay list,
    the short list and the blocked pass, driven through a short strong
    pulse, 300 steps (a length that is not a multiple of 8)."""
    from quest_amd import systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(200, 0.36, seed=70 + order)
    pl = short_pulse()
    steps = 300
    ps, cs, fc = ref_weights
    ref = oracle.solve(oracle_params(oracle, consts, order=order), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                       steps, mode=1, kind=dyadic)
    assert ref["status"] == 0
    with _ctx(sysd, consts, ref_weights, precision, dyadic, order, pl, block_steps=block_steps) as ctx:
        info = ctx.info()
        assert info["pair_method"] == 2 and info["block_steps"] == block_steps
        assert info["short_pairs"] > 0 and info["near_pairs"] > 0
        ctx.run(steps)
        rho, drho = ctx.state()
    err = rel_err(rho, ref["y"]) if precision == 0 else float(np.abs(rho - ref["y"]).max())
    print(f"blocked K={block_steps} order {order} precision {precision} dyadic {dyadic}: err {err:.2e} "
          f"short pairs {info['short_pairs']}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1
    # measured: FP64 <= 9.2e-16, mixed <= 5e-10 (rotating) and 1.9e-11 (lab)
    assert err < (1e-12 if precision == 0 else 5e-9)
    if precision == 0:
        assert rel_err(drho, ref["yp"]) < 1e-9


@pytest.mark.parametrize("precision", [0, 1])
def test_blocked_equals_per_step_method(ref_weights, precision):
    """N = 2048 at 512 um^-3 with random dipoles and detuning, 240 driven
    steps: the blocked pass and the per-step on-the-fly kernels agree to FP64
    rounding (FP64) or within the mixed path's error (mixed, both against the
    per-step FP64 run)."""
    from quest_amd import systems
    consts = systems.consts()
    sysd = systems.system(2048, seed=5, random_dipoles=True, detuning_sigma=2.0)
    side = (2048 / 512.0) ** (1 / 3)
    rng = np.random.default_rng(5)
    sysd["pos"] = rng.uniform(-side / 2, side / 2, size=(2048, 3))
    pl = short_pulse(area=math.pi / 2, delay=0.01, sigma_t=0.002)
    out = {}
    for prec, method in ((0, 0), (precision, 2)):
        with _ctx(sysd, consts, ref_weights, prec, pulse=pl, method=method) as ctx:
            ctx.run(240)
            out[(prec, method)] = ctx.state()[0]
    ref = out[(0, 0)]
    got = out[(precision, 2)]
    err = rel_err(got, ref)
    print(f"N=2048 blocked precision {precision} vs per-step FP64: {err:.2e}, "
          f"max|rho01| {np.abs(ref[:, 1]).max():.3f}")
    assert np.abs(ref[:, 1]).max() > 0.05
    # measured: FP64 1.1e-16, mixed 4.8e-10
    assert err < (1e-13 if precision == 0 else 5e-9)


@PASS_STEPS
@pytest.mark.parametrize("precision", [0, 1])
def test_blocked_runs_split_anywhere_are_bitwise_equal(ref_weights, precision, block_steps):
    """The same 45 steps in one call (graphs of 8 steps plus single steps), in
    calls that start and end inside blocks, with timing on (kernel by kernel)
    and per-call noalloc: identical states bit for bit."""
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(150, 0.3, seed=81)
    pl = short_pulse()
    states = {}
    with _ctx(sysd, consts, ref_weights, precision, pulse=pl, block_steps=block_steps) as ctx:
        ctx.run(45)
        assert ctx.info()["graph"] == 1
        states["one"] = ctx.state()[0]
    with _ctx(sysd, consts, ref_weights, precision, pulse=pl, block_steps=block_steps) as ctx:
        for n in (3, 13, 1, 20, 8):
            ctx.run(n)
        states["split"] = ctx.state()[0]
    with _ctx(sysd, consts, ref_weights, precision, pulse=pl, block_steps=block_steps) as ctx:
        ctx.set_timing(True)
        ctx.run(45)
        ms, launches = ctx.pair_kernel_time()
        sms, slaunches = ctx.short_kernel_time()
        assert launches == (45 + block_steps - 1) // block_steps and slaunches == 45 and ms > 0 and sms > 0
        states["timed"] = ctx.state()[0]
    with _ctx(sysd, consts, ref_weights, precision, pulse=pl, block_steps=block_steps) as ctx:
        ctx.run(5)
        ctx.run_noalloc(6)     # 5 .. 10: the pass under way, formed again in the static scratch
        ctx.run(4)             # 11 .. 14: formed again in the context's buffer
        ctx.run_noalloc(30)
        states["noalloc"] = ctx.state()[0]
    for k in ("split", "timed", "noalloc"):
        assert np.array_equal(states[k], states["one"]), k


@PASS_STEPS
@pytest.mark.parametrize("precision", [0, 1])
def test_blocked_per_call_matches_oracle(oracle_lib, ref_weights, precision, block_steps):
    """quest_hist_evaluate on a time-blocked context, driven as the
    reference's integrator drives evaluate (several calls per step, the last
    one with the step's final value): every evaluation against the oracle's
    evaluate over the same history (an oracle trajectory under a short
    pulse)."""
    from quest_amd import systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(160, 0.33, seed=91)
    steps = 60
    ps, cs, fc = ref_weights
    pl = short_pulse(delay=steps * 0.6e-4, sigma_t=steps * 0.15e-4)
    r = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                     steps, mode=1, record_every=1)
    rho01 = r["rec"][:, :, 1].T.copy()     # [n][steps]
    assert np.abs(rho01).max() > 0.05
    p = oracle_params(oracle, consts)
    delays, coef = oracle.build_table(p, sysd["pos"], sysd["dip"], kind=0)
    n = rho01.shape[0]
    y = np.zeros((n, 22 + steps, 2), dtype=np.complex128)
    y[:, :, 0] = 1.0
    y[:, 22:, 1] = rho01
    worst = 0.0
    with _ctx(sysd, consts, ref_weights, precision, block_steps=block_steps) as ctx:
        assert ctx.info()["short_pairs"] > 0
        for s in range(steps):
            for it in range(2):
                cur = rho01[:, s] * (1.0 + 0.01 * it)
                yy = y.copy()
                yy[:, 22 + s, 1] = cur
                got = ctx.hist_evaluate(s, cur)
                want = oracle.evaluate_history(p, delays, coef, yy, s, kind=0)
                if s > 0:
                    worst = max(worst, rel_err(got, want))
            ctx.hist_evaluate(s, rho01[:, s])
    print(f"per-call blocked K={block_steps} precision {precision}: worst {worst:.2e}")
    # measured: FP64 1.9e-13, mixed 2.4e-7 (the per-step mixed kernel: up to 4.3e-7)
    assert worst < (1e-12 if precision == 0 else 1e-6)


@PASS_STEPS
def test_blocked_autotune_keeps_the_run(ref_weights, block_steps):
    """quest_autotune inside a block (its timed passes overwrite the block's
    partial sums and change the split): the run continues as if untuned."""
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(300, 0.45, seed=93)
    pl = short_pulse()
    with _ctx(sysd, consts, ref_weights, 0, pulse=pl, block_steps=block_steps) as a, \
            _ctx(sysd, consts, ref_weights, 0, pulse=pl, block_steps=block_steps) as b:
        a.run(13)
        b.run(13)
        b.autotune(1)
        a.run(30)
        b.run(30)
        ra, rb = a.state()[0], b.state()[0]
    assert rel_err(rb, ra) < 1e-13


@PASS_STEPS
def test_blocked_batch_matches_single_systems(ref_weights, block_steps):
    """Five systems of uneven size in one batch context against each run
    alone, both time-blocked (FP64)."""
    from quest_amd import Context, systems
    consts = systems.consts()
    sizes = [40, 130, 7, 260, 90]
    syss = [dense_system(n, 0.12 * n ** (1 / 3), seed=100 + i) for i, n in enumerate(sizes)]
    pls = [short_pulse(area=a) for a in (0.5, 1.0, 1.5, 2.0, 2.5)]
    ps, cs, fc = ref_weights
    with Context(syss, consts, pulse=pls, block_steps=block_steps) as ctx:
        ctx.set_pc_weights(ps, cs, fc)
        ctx.set_pair_method(2)
        ctx.run(100)
        batch = ctx.state()[0]
    off = 0
    for s, pl, n in zip(syss, pls, sizes):
        with _ctx(s, consts, ref_weights, 0, pulse=pl, block_steps=block_steps) as one:
            one.run(100)
            r = one.state()[0]
        assert rel_err(batch[off:off + n], r) < 1e-13
        off += n


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_blocked_n4096_full_run_matches_oracle_fixture(precision, tol):
    """The N = 4096 workload (512 um^-3, pi/2 pulse), all 5000 steps, against
    the oracle trajectory in tests/golden/n4096_traj.npz, every 500 steps."""
    from quest_amd import Context, systems
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "n4096_traj.npz"))
    sysd, steps = systems.workload("n4096")
    every = int(g["record_every"])
    worst = 0.0
    with Context(sysd, systems.consts(), precision=precision, pulse=systems.pulse(math.pi / 2)) as ctx:
        ctx.set_pc_weights(g["ps"], g["cs"], float(g["fc"]))
        ctx.set_pair_method(2)
        for r in range(steps // every):
            ctx.run(every)
            rho, _ = ctx.state()
            e = (np.abs(rho - g["rec"][r]).max() / np.abs(g["rec"][r]).max() if precision == 0
                 else np.abs(rho - g["rec"][r]).max())
            worst = max(worst, float(e))
    print(f"blocked N=4096 x {steps} steps, precision {precision}: worst {worst:.3e}")
    assert worst < tol / 20   # measured: FP64 1.2e-13, mixed 5.4e-9


def test_blocked_method_rules():
    """The method is chosen before the first step; noalloc refuses partial
    sums beyond the static scratch; the automatic pass length is 16 steps
    for a large system (N = 65536: 0.3 % of its pairs have delays below 16)
    and 8 for a dense small one; block_steps other than 0, 8, 16 is an
    argument error."""
    from quest_amd import Context, QuestError, systems
    consts = systems.consts()
    with pytest.raises(QuestError):
        Context(dense_system(64, 0.3, seed=3), consts, block_steps=12)
    with Context(dense_system(64, 0.3, seed=3), consts) as ctx:
        ctx.set_pair_method(2)
        assert ctx.info()["block_steps"] == 8
    with Context(dense_system(64, 0.3, seed=3), consts) as ctx:
        ctx.set_pair_method(2)
        ctx.set_pair_method(0)
        ctx.set_pair_method(2)
        ctx.run(1)
        with pytest.raises(QuestError):
            ctx.set_pair_method(0)
        with pytest.raises(QuestError):
            ctx.set_pair_method(7)
    sysd, _ = systems.workload("n65536")
    with Context(sysd, consts, precision=1) as ctx:
        ctx.set_pair_method(2)
        assert ctx.info()["block_steps"] == 16
        with pytest.raises(QuestError, match="static scratch"):
            ctx.run_noalloc(1)
