"""GPU: batches of independent systems in one context (quest_batch_create),
the captured step graph, and the autotuner.

Tolerances: batch vs one-context-per-system 1e-12 relative in FP64 and 1e-8
in mixed precision (the work lists differ, so the FP32 partial sums round
differently); graph replay vs kernel-by-kernel launch
bitwise; a batch member vs the FP64 oracle 1e-10 relative.
"""
import math

import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse

pytestmark = pytest.mark.gpu


def _systems():
    sizes = [37, 300, 5, 129, 1]
    sys_list = [dense_system(n, 0.12 + 0.002 * n, seed=100 + i, plant_near=n >= 4)
                for i, n in enumerate(sizes)]
    pulses = [short_pulse(area=a) for a in (math.pi / 2, math.pi, math.pi / 4, 2.0, 1.0)]
    return sys_list, pulses


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_individual_contexts(ref_weights, precision):
    from quest_amd import Context, systems
    consts = systems.consts()
    sys_list, pulses = _systems()
    steps = 250
    with Context(sys_list, consts, precision=precision, pulse=pulses) as ctx:
        ctx.set_pc_weights(*ref_weights)
        info = ctx.info()
        assert info["systems"] == len(sys_list) and info["n"] == sum(len(s["pos"]) for s in sys_list)
        ctx.run(steps)
        batch_rho, _ = ctx.state()
    off = 0
    for s, pl in zip(sys_list, pulses):
        n = len(s["pos"])
        with Context(s, consts, precision=precision, pulse=pl) as one:
            one.set_pc_weights(*ref_weights)
            one.run(steps)
            rho, _ = one.state()
        err = rel_err(batch_rho[off:off + n], rho)
        print(f"system of {n}: batch vs single {err:.2e}")
        assert err < (1e-12 if precision == 0 else 1e-8)
        off += n


def test_batch_member_matches_oracle(oracle_lib, ref_weights):
    from quest_amd import Context, systems
    consts = systems.consts()
    sys_list, pulses = _systems()
    steps = 300
    with Context(sys_list, consts, pulse=pulses) as ctx:
        ctx.set_pc_weights(*ref_weights)
        ctx.run(steps)
        rho, drho = ctx.state()
    ps, cs, fc = ref_weights
    off = 37
    ref = oracle_lib.solve(oracle_params(oracle_lib, consts), sys_list[1], oracle_pulse(oracle_lib, pulses[1]),
                           ps, cs, fc, steps, mode=1)
    err = rel_err(rho[off:off + 300], ref["y"])
    print(f"batch member (N=300) vs oracle: {err:.2e}, max|rho01| {np.abs(ref['y'][:, 1]).max():.3f}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1
    assert err < 1e-10 and rel_err(drho[off:off + 300], ref["yp"]) < 1e-9


def test_graph_replays_match_kernel_by_kernel(ref_weights):
    """The captured step graph, replayed in chunks with eager state reads and
    synchronisations between, equals launching every kernel of every step."""
    from quest_amd import Context, systems
    consts = systems.consts()
    sys_list, pulses = _systems()
    with Context(sys_list, consts, pulse=pulses) as g, Context(sys_list, consts, pulse=pulses) as e:
        g.set_pc_weights(*ref_weights)
        e.set_pc_weights(*ref_weights)
        e.set_timing(True)                      # timing mode launches kernel by kernel
        for chunk in range(6):
            g.run(25)
            e.run(25)
            assert g.info()["graph"] == 1 and e.info()["graph"] == 0
            rg, dg = g.state()
            re_, de = e.state()
            assert np.array_equal(rg, re_) and np.array_equal(dg, de), f"chunk {chunk}"
            g.sync()


def test_batch_pulse_lists_match_individual_contexts(ref_weights):
    """Per-system pulse lists in a batch (two pulses, none, one) equal the
    same lists on one context per system."""
    from quest_amd import Context, systems
    consts = systems.consts()
    sys_list, pulses = _systems()
    sys_list = sys_list[:3]
    lists = [[pulses[0], short_pulse(area=0.5, delay=0.03, sigma_t=0.004)], [], [pulses[2]]]
    steps = 200
    with Context(sys_list, consts) as ctx:
        ctx.set_pc_weights(*ref_weights)
        for s, pl in enumerate(lists):
            ctx.set_pulse_list(pl, system=s)
        ctx.run(steps)
        batch_rho, _ = ctx.state()
    off = 0
    for s, pl in zip(sys_list, lists):
        n = len(s["pos"])
        with Context(s, consts) as one:
            one.set_pc_weights(*ref_weights)
            one.set_pulse_list(pl)
            one.run(steps)
            rho, _ = one.state()
        assert rel_err(batch_rho[off:off + n], rho) < 1e-12
        if not pl:
            assert np.all(rho[:, 1] == 0)
        off += n


def test_autotune_changes_no_state(ref_weights):
    from quest_amd import Context, systems
    consts = systems.consts()
    sys_list, pulses = _systems()
    with Context(sys_list, consts, pulse=pulses) as a, Context(sys_list, consts, pulse=pulses) as b:
        a.set_pc_weights(*ref_weights)
        b.set_pc_weights(*ref_weights)
        a.run(120)
        b.run(120)
        chunk = a.autotune(2)
        assert chunk == a.info()["source_chunk"] and chunk >= 32
        a.run(120)
        b.run(120)
        ra, _ = a.state()
        rb, _ = b.state()
    assert rel_err(ra, rb) < 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_dynamic_item_claiming_is_bitwise_static(ref_weights, precision):
    """Waves that claim items from the device counter compute the same items
    as the static stride, so the trajectories agree bit for bit; the counters
    reset themselves between launches (step graph replays, the autotuner's
    back-to-back launches; after it the chunk lengths may differ, so the
    batch tolerances apply)."""
    from quest_amd import CLAIM_DYNAMIC, CLAIM_STATIC, Context, systems
    consts = systems.consts()
    sys_list, pulses = _systems()
    first, second = {}, {}
    for claim in (CLAIM_STATIC, CLAIM_DYNAMIC):
        with Context(sys_list, consts, precision=precision, pulse=pulses, source_chunk=48,
                     item_claim=claim) as ctx:
            ctx.set_pc_weights(*ref_weights)
            assert ctx.info()["dynamic_items"] == (claim == CLAIM_DYNAMIC)
            ctx.run(60)
            first[claim] = ctx.state()[0]
            assert ctx.autotune(2) >= 32
            assert ctx.info()["dynamic_items"] == (claim == CLAIM_DYNAMIC)
            ctx.run(60)
            second[claim] = ctx.state()[0]
    assert np.array_equal(first[CLAIM_STATIC], first[CLAIM_DYNAMIC])
    assert rel_err(second[CLAIM_DYNAMIC], second[CLAIM_STATIC]) < (1e-12 if precision == 0 else 1e-8)


@pytest.mark.parametrize("batch,dyadic,dt,acc", [(False, 0, 1e-4, 0), (True, 0, 1e-4, 0), (False, 1, 1e-4, 0),
                                                  (False, 0, 1.3e-4, 0), (False, 0, 1e-4, 1), (True, 1, 1e-4, 1)])
def test_pipelined_source_loop_is_bitwise_plain(ref_weights, batch, dyadic, dt, acc):
    """The mixed-precision pair kernel's two source loops (plain; staged one
    source ahead) do the same arithmetic in the same order: whole runs agree
    bit for bit, for one system and for a batch, in the rotating and the lab
    frame, with theta = omega dt beyond the short rotation series (the
    library-sincos kernel variant, dt = 1.3e-4 ps), and with Kahan-compensated
    block sums (acc = 1)."""
    from quest_amd import LOOP_PIPELINED, LOOP_PLAIN, Context, systems
    consts = systems.consts(dt=dt)
    if batch:
        sys_list, pulses = _systems()
    else:
        sys_list, pulses = dense_system(300, 0.3, seed=71), short_pulse()
    rho = {}
    for loop in (LOOP_PLAIN, LOOP_PIPELINED):
        with Context(sys_list, consts, precision=1, dyadic=dyadic, pulse=pulses, source_loop=loop,
                     accumulation=acc) as ctx:
            assert ctx.info()["pipelined"] == (loop == LOOP_PIPELINED)
            ctx.set_pc_weights(*ref_weights)
            ctx.run(150)
            rho[loop] = ctx.state()[0]
    assert np.abs(rho[LOOP_PLAIN][:, 1]).max() > 0.05
    assert np.array_equal(rho[LOOP_PLAIN], rho[LOOP_PIPELINED])


def test_source_loop_default_and_autotune(ref_weights):
    """QUEST_LOOP_AUTO: pipelined in mixed precision at every size, plain in
    FP64; the autotuner may switch it and changes no state; a pinned loop
    stays pinned."""
    from quest_amd import LOOP_PLAIN, Context, systems
    consts = systems.consts()
    small = dense_system(400, 0.35, seed=72)
    large = dense_system(8193, 1.2, seed=73, plant_near=False)
    with Context(small, consts, precision=1) as ctx:
        assert ctx.info()["pipelined"] == 1
    with Context(small, consts, precision=0) as ctx:
        assert ctx.info()["pipelined"] == 0
    with Context(large, consts, precision=1) as ctx:
        assert ctx.info()["pipelined"] == 1
    with Context(large, consts, precision=1, source_loop=LOOP_PLAIN) as ctx:
        assert ctx.info()["pipelined"] == 0
    with Context(small, consts, precision=1, source_loop=LOOP_PLAIN) as ctx:
        ctx.set_pc_weights(*ref_weights)
        ctx.autotune(2)
        assert ctx.info()["pipelined"] == 0
    with Context(small, consts, precision=1, pulse=short_pulse()) as a, \
            Context(small, consts, precision=1, pulse=short_pulse()) as b:
        for c in (a, b):
            c.set_pc_weights(*ref_weights)
        a.run(60)
        b.run(30)
        b.autotune(2)
        b.run(30)
        assert b.info()["pipelined"] in (0, 1)
        ra, rb = a.state()[0], b.state()[0]
    assert rel_err(rb, ra) < 1e-8


@pytest.mark.parametrize("n", [24576, 49152])
def test_default_split_gives_every_xcd_an_equal_share(n):
    """The default work split cuts the largest system into 48 (mixed) or 24
    (FP64) equal source chunks -- multiples of the 8 XCDs, so the static list
    gives every XCD the same pair work (49152 dots in 4096-source chunks made
    12, half the XCDs doing twice the work) -- and runs one block per CU
    below the occupancy limit; the partial rows are the chunks."""
    from quest_amd import Context, systems
    sysd = dense_system(n, (n / 512) ** (1 / 3), seed=81, plant_near=False)
    grids = {}
    for precision, chunks in ((1, 48), (0, 24)):
        with Context(sysd, systems.consts(), precision=precision) as ctx:
            info = ctx.info()
            assert info["splits"] == chunks, info
            assert info["source_chunk"] == -(-n // chunks)
            assert info["grid"] % 8 == 0
            grids[precision] = info["grid"]
    # mixed: 8 resident blocks per CU less one; FP64: 6 less one (same CU count)
    assert grids[1] * 5 == grids[0] * 7
