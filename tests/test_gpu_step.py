"""The one-launch predictor-corrector step against the oracle.

The step kernel runs the predictor and all 11 evaluate/correct iterations of
every dot in one launch, with a block barrier between iterations; the only
coupling inside a step is through zero-delay pairs (history_interaction.cpp:
56-66 with delay 0 reads the current iterate), so the library packs each
connected group of zero-delay partners into one block.  A chain of dots
spaced below c0 dt is one such group; its length picks the block size (64,
128, 256) or, beyond 256 dots, the per-iteration kernels.  Both paths must
give the oracle's trajectory (FP64 tolerance 1e-10 norm-wise relative).
"""
import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse

pytestmark = pytest.mark.gpu

SPACING = 0.02   # um; c0 dt = 0.02998 um, so neighbours have delay 0, next-but-one delay 1


def _chained(n_chain, n_free, seed):
    """n_free random dots plus a straight chain of n_chain dots whose
    neighbours are zero-delay partners (one connected group)."""
    from quest_amd import systems
    rng = np.random.default_rng(seed)
    sysd = dense_system(n_free + n_chain, 1.0, seed=seed)
    start = np.array([0.6, -0.5, 0.1])
    t = np.arange(n_chain)[:, None] * np.array([SPACING, 0.0, 0.0])
    jitter = rng.uniform(-5e-4, 5e-4, size=(n_chain, 3))
    pos = sysd["pos"].copy()
    pos[n_free:] = start + t + jitter
    sysd["pos"] = pos
    sysd["freq"] = np.full(n_free + n_chain, systems.OMEGA0) + rng.normal(0, 2.0, n_free + n_chain)
    return sysd


@pytest.mark.parametrize("n_chain,block", [(0, 64), (40, 64), (100, 128), (200, 256), (300, 0)])
def test_step_block_follows_largest_group(n_chain, block):
    from quest_amd import Context, systems
    sysd = _chained(n_chain, 60, seed=5) if n_chain else dense_system(60, 1.0, seed=5, plant_near=False)
    with Context(sysd, systems.consts(), precision=0) as ctx:
        info = ctx.info()
    assert info["max_group"] == max(n_chain, 1)
    assert info["step_block"] == block


@pytest.mark.parametrize("method,block_steps", [(0, 0), (2, 8), (2, 16)], ids=["per_step", "K8", "K16"])
@pytest.mark.parametrize("n_chain", [100, 200, 300])
def test_grouped_run_matches_oracle(oracle_lib, ref_weights, n_chain, method, block_steps):
    """n_chain 100 / 200: one-launch step with 128- / 256-thread blocks and
    padding; n_chain 300: per-iteration kernels; each with the per-step and
    the time-blocked pair sums (8- and 16-step passes)."""
    from quest_amd import Context, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = _chained(n_chain, 90, seed=17)
    pl = short_pulse()
    steps = 160
    ps, cs, fc = ref_weights
    ref = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                       steps, mode=0)
    assert ref["status"] == 0
    with Context(sysd, consts, precision=0, pulse=pl, block_steps=block_steps) as ctx:
        ctx.set_pc_weights(ps, cs, fc)
        ctx.set_pair_method(method)
        ctx.run(steps // 2 + 3)
        ctx.run(steps - steps // 2 - 3)
        info = ctx.info()
        rho, drho = ctx.state()
    assert info["step_block"] == {100: 128, 200: 256, 300: 0}[n_chain]
    e_rho, e_d = rel_err(rho, ref["y"]), rel_err(drho, ref["yp"])
    print(f"chain {n_chain} method {method} K {block_steps}: block {info['step_block']}, rho rel err "
          f"{e_rho:.3e}, drho {e_d:.3e}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1, "the pulse did not drive the system"
    assert e_rho < 1e-10 and e_d < 1e-9


def test_grouped_batch_matches_individual_contexts(ref_weights):
    """Systems of a batch each hold a group; groups never straddle blocks,
    and the batch gives each system's own trajectory."""
    from quest_amd import Context, systems
    consts = systems.consts()
    ps, cs, fc = ref_weights
    pl = short_pulse()
    sys_list = [_chained(nc, nf, seed=40 + i) for i, (nc, nf) in enumerate([(30, 20), (50, 70), (5, 9)])]
    with Context(sys_list, consts, precision=0, pulse=[pl] * len(sys_list)) as b:
        b.set_pc_weights(ps, cs, fc)
        b.run(120)
        assert b.info()["step_block"] == 64 and b.info()["max_group"] == 50
        got = b.state()[0]
    off = 0
    for sd in sys_list:
        n = len(sd["pos"])
        with Context(sd, consts, precision=0, pulse=pl) as ctx:
            ctx.set_pc_weights(ps, cs, fc)
            ctx.run(120)
            want = ctx.state()[0]
        assert rel_err(got[off:off + n], want) < 1e-12
        off += n
