"""GPU: the coefficient-table pair method (QUEST_PAIR_TABLE) -- the
reference's own precomputed per-pair coefficients streamed every step --
against the oracle's reference method (which is that method, on the CPU),
against the on-the-fly FP64 path, per call and in batches.

Tolerances: 1e-12 relative in rho (trajectories) and in the field of one
evaluation; the table and the oracle use the same FP64 coefficient formulas
and differ only in summation order.
"""
import math

import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse

pytestmark = pytest.mark.gpu


def test_table_run_matches_reference_method(oracle_lib, ref_weights):
    from quest_amd import PAIR_TABLE, Context, systems
    consts = systems.consts()
    sysd = dense_system(48, 0.22, seed=21)
    pl = short_pulse()
    ps, cs, fc = ref_weights
    steps = 200
    with Context(sysd, consts, pulse=pl) as ctx:
        ctx.set_pair_method(PAIR_TABLE)
        ctx.set_pc_weights(ps, cs, fc)
        ctx.run(steps)
        rho, drho = ctx.state()
    ref = oracle_lib.solve(oracle_params(oracle_lib, consts), sysd, oracle_pulse(oracle_lib, pl),
                           ps, cs, fc, steps, mode=0)
    err = rel_err(rho, ref["y"])
    print(f"table vs reference method: {err:.2e}, max|rho01| {np.abs(ref['y'][:, 1]).max():.3f}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1
    assert err < 1e-12 and rel_err(drho, ref["yp"]) < 1e-11


@pytest.mark.parametrize("dyadic", [0, 1])
def test_table_field_equals_on_the_fly(dyadic):
    """The field of every per-call evaluation, table against on the fly, on a
    smooth synthetic history (Gaussian in time, a phase per dot) over N = 1200
    dots: the two FP64 paths compute the same sums in different forms."""
    from quest_amd import PAIR_TABLE, Context, systems
    consts = systems.consts()
    sysd = dense_system(1200, 1.3, seed=41)
    n, steps = len(sysd["pos"]), 150
    rng = np.random.default_rng(3)
    amp = 0.3 * rng.uniform(0.5, 1.0, n) * np.exp(2j * np.pi * rng.uniform(size=n))
    t = np.arange(steps)
    hist = amp[:, None] * np.exp(-((t[None, :] - 90.0) / 40.0) ** 2) * np.exp(-0.02j * t[None, :])
    worst, scale = 0.0, 0.0
    with Context(sysd, consts, dyadic=dyadic) as fly, Context(sysd, consts, dyadic=dyadic) as tab:
        tab.set_pair_method(PAIR_TABLE)
        for s in range(steps):
            a = fly.hist_evaluate(s, hist[:, s])
            b = tab.hist_evaluate(s, hist[:, s])
            scale = max(scale, np.abs(a).max())
            if np.abs(a).max() > 0:
                worst = max(worst, rel_err(b, a))
    print(f"dyadic {dyadic}: table vs on-the-fly field {worst:.2e} (max |field| {scale:.3g})")
    assert scale > 0
    assert worst < 1e-12


def test_table_hist_evaluate_matches_oracle(oracle_lib, ref_weights):
    """Per-call mode over a table: each evaluate against the oracle's
    evaluate (history_interaction.cpp:47-70) on the same history."""
    from quest_amd import PAIR_TABLE, Context, systems
    consts = systems.consts()
    sysd = dense_system(64, 0.25, seed=75)
    steps = 60
    ps, cs, fc = ref_weights
    pl = short_pulse(delay=steps * 0.6e-4, sigma_t=steps * 0.15e-4)
    r = oracle_lib.solve(oracle_params(oracle_lib, consts), sysd, oracle_pulse(oracle_lib, pl), ps, cs, fc,
                         steps, mode=1, record_every=1)
    rho01 = r["rec"][:, :, 1].T.copy()
    p = oracle_params(oracle_lib, consts)
    delays, coef = oracle_lib.build_table(p, sysd["pos"], sysd["dip"])
    y = np.zeros((len(sysd["pos"]), 22 + steps, 2), dtype=np.complex128)
    y[:, :, 0] = 1.0
    y[:, 22:, 1] = rho01
    worst = 0.0
    with Context(sysd, consts) as ctx:
        ctx.set_pair_method(PAIR_TABLE)
        for s in range(steps):
            got = ctx.hist_evaluate(s, rho01[:, s])
            want = oracle_lib.evaluate_history(p, delays, coef, y, s)
            if s > 0:
                worst = max(worst, rel_err(got, want))
    print(f"per-call table vs oracle evaluate: {worst:.2e}")
    assert worst < 1e-12


def test_table_batch_matches_individual(ref_weights):
    from quest_amd import PAIR_TABLE, Context, systems
    consts = systems.consts()
    sizes = [37, 300, 5, 129]
    sys_list = [dense_system(n, 0.12 + 0.002 * n, seed=100 + i, plant_near=n >= 4) for i, n in enumerate(sizes)]
    pulses = [short_pulse(area=a) for a in (math.pi / 2, math.pi, math.pi / 4, 2.0)]
    with Context(sys_list, consts, pulse=pulses) as ctx:
        ctx.set_pair_method(PAIR_TABLE)
        ctx.set_pc_weights(*ref_weights)
        ctx.run(250)
        batch, _ = ctx.state()
    off = 0
    for s, pl in zip(sys_list, pulses):
        n = len(s["pos"])
        with Context(s, consts, pulse=pl) as one:
            one.set_pc_weights(*ref_weights)
            one.run(250)
            rho, _ = one.state()
        assert rel_err(batch[off:off + n], rho) < 1e-12
        off += n


def test_table_mode_errors(ref_weights):
    from quest_amd import PAIR_ON_THE_FLY, PAIR_TABLE, Context, QuestError, systems
    consts = systems.consts()
    sysd = dense_system(32, 0.2, seed=5)
    with Context(sysd, consts, precision=1) as ctx:
        with pytest.raises(QuestError, match="FP64"):
            ctx.set_pair_method(PAIR_TABLE)
        with pytest.raises(QuestError):
            ctx.set_pair_method(7)
    with Context(sysd, consts, pulse=short_pulse()) as ctx:
        ctx.set_pc_weights(*ref_weights)
        ctx.set_pair_method(PAIR_TABLE)
        ctx.set_pair_method(PAIR_ON_THE_FLY)   # back before the first step: allowed
        ctx.set_pair_method(PAIR_TABLE)
        ctx.run(5)
        with pytest.raises(QuestError, match="before the first step"):
            ctx.set_pair_method(PAIR_ON_THE_FLY)
