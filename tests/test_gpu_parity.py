"""GPU parity: the HIP library (through the C ABI) against the FP64 CPU oracle.

Tolerances (norm-wise relative error max|gpu - oracle| / max|oracle|):
  FP64 field of one evaluation          1e-12
  FP64 trajectories                     1e-10
  mixed field of one evaluation         2e-6   (FP32 pair arithmetic; 1e-6 at N = 131072)
  mixed vs FP64 trajectories, |d rho|   1e-6   absolute
Integer delays are compared bit for bit.
"""
import math

import numpy as np
import pytest

from conftest import OBS_STEP, dense_system, oracle_params, oracle_pulse, rel_err, short_pulse, smooth_history

pytestmark = pytest.mark.gpu


def _physical_history(oracle, sysd, consts, steps, weights):
    """rho01[i][t], t in [0, steps): an oracle trajectory under a short pulse
    centred in the window (smooth in time, as every integrated history is)."""
    ps, cs, fc = weights
    pl = short_pulse(delay=steps * 0.6e-4, sigma_t=steps * 0.15e-4)
    r = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                     steps, mode=1, record_every=1)
    return r["rec"][:, :, 1].T.copy()


def _oracle_y(rho01, window=22):
    n, steps = rho01.shape
    y = np.zeros((n, window + steps, 2), dtype=np.complex128)
    y[:, :, 0] = 1.0
    y[:, window:, 1] = rho01
    return y


@pytest.mark.parametrize("precision,dyadic,order,n,side", [
    (0, 0, 4, 64, 0.25),
    (0, 0, 3, 64, 0.25),
    (0, 1, 4, 64, 0.25),
    (0, 0, 4, 1000, 1.2),
    (1, 0, 4, 64, 0.25),
    (1, 0, 4, 1000, 1.2),
    (1, 1, 4, 64, 0.25),
    (1, 0, 6, 64, 0.25),
])
def test_hist_evaluate_matches_oracle(oracle_lib, ref_weights, precision, dyadic, order, n, side):
    from quest_amd import Context, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(n, side, seed=11 + n)
    steps = 70
    rho01 = _physical_history(oracle, sysd, consts, steps, ref_weights)
    assert np.abs(rho01).max() > 0.05
    p = oracle_params(oracle, consts, order=order)
    delays, coef = oracle.build_table(p, sysd["pos"], sysd["dip"], kind=dyadic)
    assert (delays == 0).sum() >= 1, "the planted zero-delay pair is missing"
    y = _oracle_y(rho01)
    # mixed: measured 6e-8 .. 4.3e-7 (N = 1000); the tolerance is ~5x that
    tol = 1e-12 if precision == 0 else 2e-6
    worst = 0.0
    with Context(sysd, consts, precision=precision, dyadic=dyadic, order=order) as ctx:
        for s in range(steps):
            for it in range(2):   # the integrator calls evaluate repeatedly per step
                cur = rho01[:, s] * (1.0 + 0.01 * it)
                yy = y.copy()
                yy[:, 22 + s, 1] = cur
                got = ctx.hist_evaluate(s, cur)
                want = oracle.evaluate_history(p, delays, coef, yy, s, kind=dyadic)
                if s > 0:
                    worst = max(worst, rel_err(got, want))
            # the final value of step s is the unscaled one
            ctx.hist_evaluate(s, rho01[:, s])
    print(f"precision={precision} dyadic={dyadic} order={order} n={n}: worst rel err {worst:.3e}")
    assert worst < tol


@pytest.mark.parametrize("dyadic", [0, 1])
def test_compensated_block_sums_match_oracle(oracle_lib, ref_weights, dyadic):
    """QUEST_ACC_COMPENSATED (Kahan-compensated FP32 block sums folded into
    FP64): per-call fields of N = 1000 against the oracle within the mixed
    tolerance 2e-6, and within 1e-6 of the plain block sums -- both errors are
    set by the FP32 arithmetic of each pair's term, which compensation of the
    sums does not touch; a whole driven run within 1e-6 of FP64."""
    from quest_amd import ACC_COMPENSATED, Context, systems
    oracle = oracle_lib
    consts = systems.consts()
    n, steps = 1000, 60
    sysd = dense_system(n, 1.2, seed=1011)
    rho01 = _physical_history(oracle, sysd, consts, steps, ref_weights)
    p = oracle_params(oracle, consts)
    delays, coef = oracle.build_table(p, sysd["pos"], sysd["dip"], kind=dyadic)
    y = _oracle_y(rho01)
    got = {}
    for acc in (0, ACC_COMPENSATED):
        with Context(sysd, consts, precision=1, dyadic=dyadic, accumulation=acc) as ctx:
            for s in range(steps):
                got[acc] = ctx.hist_evaluate(s, rho01[:, s])
    want = oracle.evaluate_history(p, delays, coef, y, steps - 1, kind=dyadic)
    e_c, e_b = rel_err(got[ACC_COMPENSATED], want), rel_err(got[0], want)
    print(f"dyadic={dyadic}: compensated {e_c:.3e}, block {e_b:.3e} vs oracle")
    assert e_c < 2e-6 and rel_err(got[ACC_COMPENSATED], got[0]) < 1e-6
    pl = short_pulse()
    r = {}
    for prec, acc in ((0, 0), (1, ACC_COMPENSATED)):
        r[prec] = _run_gpu(sysd, consts, pl, ref_weights, 150, precision=prec, accumulation=acc)[0]
    assert np.abs(r[0][:, 1]).max() > 0.05
    assert np.abs(r[1] - r[0]).max() < 1e-6


def test_pulse_matches_oracle(oracle_lib):
    from quest_amd import Context, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = systems.system(128, seed=9)
    pl = systems.pulse(math.pi / 2)
    with Context(sysd, consts, pulse=pl) as ctx:
        for step in (0, 3000, 5000, 5021, 7000):
            got = ctx.pulse_evaluate(step)
            want = oracle.pulse_evaluate(oracle_pulse(oracle, pl), consts["hbar"], sysd["pos"],
                                         sysd["dip"], step * consts["dt"])
            assert rel_err(got, want) < 1e-13
            assert np.all(got.imag == 0)


def _two_pulses():
    p1 = short_pulse(area=math.pi / 3, delay=0.012, sigma_t=0.003)
    p2 = dict(short_pulse(area=math.pi / 4, delay=0.025, sigma_t=0.004), polarization=[0.0, 0.6, 0.8])
    return [p1, p2]


def test_pulse_list_evaluate_is_the_sum(oracle_lib):
    """quest_set_pulse_list: the pulse field is the sum of the pulses' fields;
    an empty list clears it; bad system or count is an argument error."""
    from quest_amd import Context, QuestError, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = systems.system(96, seed=12)
    pls = _two_pulses() + [systems.pulse(math.pi / 2)]
    with Context(sysd, consts) as ctx:
        assert np.all(ctx.pulse_evaluate(200) == 0)
        ctx.set_pulse_list(pls)
        for step in (0, 120, 250, 400, 5000):
            t = step * consts["dt"]
            want = sum(oracle.pulse_evaluate(oracle_pulse(oracle, pl), consts["hbar"], sysd["pos"],
                                             sysd["dip"], t) for pl in pls)
            got = ctx.pulse_evaluate(step)
            assert rel_err(got, want) < 1e-13, step
        ctx.set_pulse_list([])
        assert np.all(ctx.pulse_evaluate(250) == 0)
        for bad in (lambda: ctx.set_pulse_list(pls, system=1),
                    lambda: ctx.set_pulse_list([pls[0]] * 9)):
            with pytest.raises(QuestError):
                bad()


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_pulse_list_run_matches_oracle(oracle_lib, ref_weights, precision, tol):
    """Two pulses (different areas, delays, polarisations) on a dense system:
    the whole device run against the oracle's sum of pulse interactions."""
    from quest_amd import Context, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(40, 0.2, seed=21)
    pls = _two_pulses()
    ps, cs, fc = ref_weights
    with Context(sysd, consts, precision=precision) as ctx:
        ctx.set_pc_weights(ps, cs, fc)
        ctx.set_pulse_list(pls)
        ctx.run(350)
        rho, drho = ctx.state()
    ref = oracle.solve(oracle_params(oracle, consts), sysd, [oracle_pulse(oracle, pl) for pl in pls],
                       ps, cs, fc, 350, mode=1)
    err = rel_err(rho, ref["y"])
    print(f"two pulses, precision {precision}: {err:.2e}, max|rho01| {np.abs(ref['y'][:, 1]).max():.3f}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1
    assert err < tol


def _run_gpu(sysd, consts, pl, weights, steps, precision=0, chunks=1, accumulation=0, method=0):
    from quest_amd import Context
    ps, cs, fc = weights
    with Context(sysd, consts, precision=precision, pulse=pl, accumulation=accumulation) as ctx:
        ctx.set_pc_weights(ps, cs, fc)
        ctx.set_pair_method(method)
        per = steps // chunks
        for _ in range(chunks):
            ctx.run(per)
        assert ctx.step == steps
        return ctx.state()


@pytest.mark.parametrize("mode", [0, 1])
def test_run_matches_oracle(oracle_lib, ref_weights, mode):
    from quest_amd import systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(48, 0.22, seed=21)
    pl = short_pulse()
    steps = 160 if mode == 0 else 400
    ps, cs, fc = ref_weights
    ref = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                       steps, mode=mode)
    assert ref["status"] == 0
    rho, drho = _run_gpu(sysd, consts, pl, ref_weights, steps, chunks=2)
    e_rho = rel_err(rho, ref["y"])
    e_d = rel_err(drho, ref["yp"])
    print(f"mode {mode}: rho rel err {e_rho:.3e}, drho rel err {e_d:.3e}, "
          f"max|rho01| {np.abs(ref['y'][:, 1]).max():.3f}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1, "the pulse did not drive the system"
    assert e_rho < 1e-10 and e_d < 1e-9


def test_run_mixed_close_to_fp64(ref_weights):
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(256, 0.5, seed=31)
    pl = short_pulse()
    r64, _ = _run_gpu(sysd, consts, pl, ref_weights, 400, precision=0)
    r32, _ = _run_gpu(sysd, consts, pl, ref_weights, 400, precision=1)
    err = float(np.abs(r32 - r64).max())
    print(f"mixed vs fp64 max|d rho| = {err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("precision", [0, 1])
def test_pair_delays_bit_exact(precision):
    """Delay split of every pair at N = 4096 -- the FP64 split, and the mixed
    path's FP32 estimate with its FP64 fallback near integers -- vs an IEEE
    numpy restatement of history_interaction.cpp:29-34
    (sqrt((dx^2 + dy^2) + dz^2) / (c0 dt), trunc)."""
    from quest_amd import Context, systems
    sysd, _ = systems.workload("n4096")
    consts = systems.consts()
    pos = sysd["pos"]
    rows, cols = np.tril_indices(pos.shape[0], -1)
    dr = pos[cols] - pos[rows]
    r2 = (dr[:, 0] * dr[:, 0] + dr[:, 1] * dr[:, 1]) + dr[:, 2] * dr[:, 2]
    want = np.trunc(np.sqrt(r2) / (consts["c0"] * consts["dt"])).astype(np.int32)
    with Context(sysd, consts, precision=precision) as ctx:
        got = ctx.pair_delays()
        part = ctx.pair_delays(first=1234567, count=10000)
    assert got.shape == want.shape
    mism = int((got != want).sum())
    assert mism == 0, f"{mism} delays differ"
    assert np.array_equal(part, want[1234567:1244567])


@pytest.mark.timeout(300)
@pytest.mark.parametrize("workload", ["n65536", "n131072"])
def test_mixed_delays_equal_fp64_full_size(workload):
    """At the two full-size workloads the mixed path classifies every pair
    like the FP64 split (history_interaction.cpp:31-32): the integer delays of
    the mixed pair kernel's split equal the FP64 ones bit for bit -- all
    2.1e9 pairs at N = 65536, 128 seeded ranges of 2^19 pairs (67M) at
    N = 131072 -- and so do the zero-delay lists (same number of pairs) and,
    with the time-blocked method, the short lists and the pass length."""
    from quest_amd import Context, systems
    sysd, _ = systems.workload(workload)
    n = sysd["pos"].shape[0]
    total = n * (n - 1) // 2
    if workload == "n65536":
        step = 1 << 27
        ranges = [(f, min(step, total - f)) for f in range(0, total, step)]
    else:
        rng = np.random.default_rng(17)
        ranges = [(int(f), 1 << 19) for f in np.sort(rng.integers(0, total - (1 << 19), size=128))]
    consts = systems.consts()
    checked = 0
    with Context(sysd, consts, precision=0) as c64, Context(sysd, consts, precision=1) as c32:
        near = c64.info()["near_pairs"]
        assert c32.info()["near_pairs"] == near
        for first, count in ranges:
            a = c64.pair_delays(first, count)
            b = c32.pair_delays(first, count)
            mism = int(np.count_nonzero(a != b))
            assert mism == 0, f"{mism} of pairs [{first}, +{count}) differ"
            checked += count
        c64.set_pair_method(2)
        c32.set_pair_method(2)
        i64, i32 = c64.info(), c32.info()
        assert i32["short_pairs"] == i64["short_pairs"] > 0
        assert i32["block_steps"] == i64["block_steps"] == 16
    print(f"{workload}: {checked} pairs, mixed delays == FP64 delays; {near} zero-delay pairs and "
          f"{i64['short_pairs']} short-list pairs in both")


def test_unbounded_solution_reports_first_bad_step(oracle_lib, ref_weights):
    """A dot with a negative T2 grows without bound; the step the device
    reports equals the oracle's (integrator.cpp:187-207)."""
    from quest_amd import UnboundedSolution, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(16, 0.2, seed=41)
    sysd["t2"][5] = -2e-5
    pl = short_pulse()
    ps, cs, fc = ref_weights
    ref = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                       400, mode=1)
    assert ref["status"] == 1
    with pytest.raises(UnboundedSolution) as ei:
        _run_gpu(sysd, consts, pl, ref_weights, 400)
    assert ei.value.step == ref["bad_step"]
    assert str(ei.value) == f"unbounded history values at or before step {ref['bad_step']}"


def test_per_call_sequence_errors():
    from quest_amd import Context, QuestError, systems
    sysd = systems.system(32, seed=3)
    with Context(sysd, systems.consts()) as ctx:
        z = np.zeros(32, dtype=np.complex128)
        ctx.hist_evaluate(0, z)
        ctx.hist_evaluate(1, z)
        with pytest.raises(QuestError):
            ctx.hist_evaluate(0, z)
        with pytest.raises(QuestError):
            ctx.hist_evaluate(3, z)
        with pytest.raises(QuestError):
            ctx.run(1)
        with pytest.raises(QuestError):   # the rho01 array must hold n values
            ctx.hist_evaluate(1, z[:31])
    with Context(sysd, systems.consts()) as ctx:
        with pytest.raises(QuestError, match="step 0"):
            ctx.hist_evaluate(2, np.zeros(32, dtype=np.complex128))


def test_mixed_ring_plane_limit_is_an_error():
    """The FP32 pair kernel addresses a coefficient plane (16 B per dot and
    ring slot) with 32-bit offsets: a mixed-precision context whose plane
    would pass 4 GiB is refused at create, before any ring is allocated.
    20000 dots spread over 150 um need 16384 slots: 5.2 GB per plane."""
    from quest_amd import Context, QuestError, systems
    rng = np.random.default_rng(8)
    n = 20000
    sysd = {"pos": rng.uniform(-75.0, 75.0, size=(n, 3)), "dip": np.tile([0.0, 0.0, systems.DIPOLE], (n, 1)),
            "freq": np.full(n, systems.OMEGA0), "t1": np.full(n, systems.T1), "t2": np.full(n, systems.T2)}
    with pytest.raises(QuestError, match="exceeds 2"):
        Context(sysd, systems.consts(), precision=1)


@pytest.mark.parametrize("precision", [0, 1])
def test_delay_range_limit_is_an_error(precision):
    """Delays are limited to 1e6 steps (which also keeps the FP32 delay
    split's float rounding constant, exact below 2^22 steps, valid): two dots
    1.5e5 um apart (5e6 steps at c0 dt = 0.03 um) are refused at create."""
    from quest_amd import Context, QuestError, systems
    sysd = {"pos": np.array([[0.0, 0.0, 0.0], [1.5e5, 0.0, 0.0]]), "dip": np.tile([0.0, 0.0, systems.DIPOLE], (2, 1)),
            "freq": np.full(2, systems.OMEGA0), "t1": np.full(2, systems.T1), "t2": np.full(2, systems.T2)}
    with pytest.raises(QuestError, match="too large"):
        Context(sysd, systems.consts(), precision=precision)


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_n4096_full_run_matches_oracle_fixture(precision, tol):
    """The N = 4096 system (512 um^-3, pi/2 pulse), all 5000 steps, against the
    oracle trajectory in tests/golden/n4096_traj.npz (made by
    tests/golden/make_golden.py with the same weights), every 500 steps."""
    import os
    from quest_amd import Context, systems
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "n4096_traj.npz"))
    sysd, steps = systems.workload("n4096")
    assert steps == int(g["steps"])
    every = int(g["record_every"])
    worst = 0.0
    with Context(sysd, systems.consts(), precision=precision, pulse=systems.pulse(math.pi / 2)) as ctx:
        ctx.set_pc_weights(g["ps"], g["cs"], float(g["fc"]))
        for r in range(steps // every):
            ctx.run(every)
            rho, drho = ctx.state()
            e = (np.abs(rho - g["rec"][r]).max() / np.abs(g["rec"][r]).max() if precision == 0
                 else np.abs(rho - g["rec"][r]).max())
            worst = max(worst, float(e))
        assert np.abs(rho - g["y"]).max() == np.abs(rho - g["rec"][-1]).max()
    print(f"N=4096 x {steps} steps, precision {precision}: worst error {worst:.3e} "
          f"(max|rho01| {np.abs(g['y'][:, 1]).max():.3f})")
    assert worst < tol


@pytest.mark.parametrize("method", [0, 2], ids=["per_step", "time_blocked"])
@pytest.mark.parametrize("name", ["n65536", "n131072"])
def test_full_size_mixed_matches_fp64(ref_weights, name, method):
    """The workloads of the measured configurations at their full sizes (the
    CPU oracle cannot run them: the reference method's table alone would be
    180 / 722 GB), mixed against FP64 on the GPU under a short, strong pulse
    so that rho01 is large within 100 steps: rho01 within 1e-6 relative (the
    north-star bound), populations within 1e-6 absolute, no non-finite step."""
    from quest_amd import Context, systems
    sysd, _ = systems.workload(name)
    n = len(sysd["pos"])
    pl = short_pulse(area=math.pi / 2, delay=0.005, sigma_t=0.0015)
    steps = 100
    out = {}
    for precision in (0, 1):
        with Context(sysd, systems.consts(), precision=precision, pulse=pl) as ctx:
            ctx.set_pc_weights(*ref_weights)
            ctx.set_pair_method(method)
            ctx.run(steps)
            rho, _ = ctx.state()
            info = ctx.info()
            assert info["n"] == n and ctx.step == steps
            out[precision] = rho
    big = np.abs(out[0][:, 1]).max()
    err01 = np.abs(out[1][:, 1] - out[0][:, 1]).max() / big
    err00 = np.abs(out[1][:, 0] - out[0][:, 0]).max()
    print(f"{name} method {method}: max|rho01| {big:.3f}, mixed vs FP64 rho01 {err01:.2e} relative, "
          f"rho00 {err00:.2e}")
    assert big > 0.1
    assert np.all(np.isfinite(out[1]))
    assert err01 < 1e-6 and err00 < 1e-6


@pytest.mark.parametrize("method", [0, 2], ids=["per_step", "time_blocked"])
@pytest.mark.parametrize("dt", [1.3e-4, 1e-3])
def test_large_laser_phase_per_step_matches_oracle(oracle_lib, ref_weights, dt, method):
    """theta = omega dt beyond the short rotation series (FP32 series up to
    0.28, FP64 up to 0.33): dt = 1.3e-4 ps (theta = 0.296: FP64 series, FP32
    library sincos) and 1e-3 ps (theta = 2.28: library sincos in both).  The
    FP64 trajectory matches the oracle (complex exp of the reference,
    rotating_green_function.cpp:14) to 1e-10 and mixed stays within 1e-6 of
    it."""
    from quest_amd import systems
    oracle = oracle_lib
    consts = systems.consts(dt=dt)
    side = 0.22 if dt < 5e-4 else 1.6
    sysd = dense_system(48, side, seed=23, plant_near=False)
    steps = 160 if dt < 5e-4 else 120
    pl = short_pulse(delay=steps * dt * 0.4, sigma_t=steps * dt * 0.1)
    ps, cs, fc = ref_weights
    ref = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc, steps, mode=1)
    assert ref["status"] == 0
    r64, d64 = _run_gpu(sysd, consts, pl, ref_weights, steps, precision=0, method=method)
    r32, _ = _run_gpu(sysd, consts, pl, ref_weights, steps, precision=1, method=method)
    e64 = rel_err(r64, ref["y"])
    e32 = float(np.abs(r32 - r64).max())
    print(f"dt={dt} method {method}: theta={consts['laser_omega'] * dt:.3f} fp64 rel err {e64:.2e}, "
          f"mixed |d rho| {e32:.2e}, "
          f"max|rho01| {np.abs(ref['y'][:, 1]).max():.3f}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1, "the pulse did not drive the system"
    assert e64 < 1e-10 and e32 < 1e-6


@pytest.mark.parametrize("order", [1, 2, 3, 5, 6])
@pytest.mark.parametrize("precision,dyadic", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_run_every_order_matches_oracle(oracle_lib, ref_weights, order, precision, dyadic):
    """Whole runs at interpolation orders 1-6 other than 4 (the FP32 ring's
    plane layout differs with the order: pairs of coefficients, plus a last
    plane when the order is even), rotating and lab dyadic: FP64 within 1e-10
    of the oracle, mixed within 1e-6 (absolute, |rho| <= 1)."""
    from quest_amd import Context, systems
    oracle = oracle_lib
    consts = systems.consts()
    sysd = dense_system(40, 0.24, seed=50 + order)
    pl = short_pulse()
    steps = 150
    ps, cs, fc = ref_weights
    ref = oracle.solve(oracle_params(oracle, consts, order=order), sysd, oracle_pulse(oracle, pl), ps, cs, fc,
                       steps, mode=1, kind=dyadic)
    assert ref["status"] == 0
    with Context(sysd, consts, precision=precision, dyadic=dyadic, order=order, pulse=pl) as ctx:
        ctx.set_pc_weights(ps, cs, fc)
        ctx.run(steps)
        rho, _ = ctx.state()
    err = rel_err(rho, ref["y"]) if precision == 0 else float(np.abs(rho - ref["y"]).max())
    print(f"order {order} precision {precision} dyadic {dyadic}: err {err:.2e}")
    assert np.abs(ref["y"][:, 1]).max() > 0.1
    assert err < (1e-10 if precision == 0 else 1e-6)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("method", [0, 2], ids=["per_step", "time_blocked"])
@pytest.mark.parametrize("precision", [0, 1])
def test_n131072_observers_match_oracle(precision, method):
    """Full size, the N = 131072 workload: the per-call field at step 380 of a
    smooth synthetic rho01 history (every delay of the system, ~370 steps at
    most, reaches into it), at 4096 seeded observers, against the FP64
    oracle's table-free evaluation of the same history committed in
    tests/golden/n131072_observers.npz (pinned by tests/test_golden.py).
    Norm-wise relative tolerance: FP64 2e-11 (a sum over 131071 sources;
    measured 6.8e-12, against < 6e-13 for the 1000-dot sums above), mixed 1e-6
    (measured 1.0e-7; the bound of the FP32 chunk sums is ~4e-7 of the
    field's scale)."""
    import os
    from quest_amd import Context, systems
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "n131072_observers.npz"))
    sysd, _ = systems.workload("n131072")
    n = sysd["pos"].shape[0]
    hist = smooth_history(n, OBS_STEP + 1)
    with Context(sysd, systems.consts(), precision=precision) as ctx:
        ctx.set_pair_method(method)
        for s in range(OBS_STEP + 1):
            out = ctx.hist_evaluate(s, hist[s])
    obs = g["observers"]
    err = rel_err(out[obs], g["field"])
    print(f"N={n} method={method} precision={precision}: {len(obs)} observers, rel err {err:.3e}, "
          f"max |field| {np.abs(g['field']).max():.3e}")
    assert err < (2e-11 if precision == 0 else 1e-6)
