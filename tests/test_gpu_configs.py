"""GPU runs of two measured configurations at their full length, against
committed oracle fixtures (tests/golden/make_golden.py, pinned on the CPU by
tests/test_golden.py):

* the N = 64 system (0.5 um cube, dipoles along z, pi/2 pulse at 0.5 ps), all
  20000 steps, so the whole pulse passes through, against the oracle's
  reference method (pair table + 11 scatter sums per step,
  history_interaction.cpp:47-70; integrator.cpp:120-185), every 2000 steps:
  FP64 within 1e-10, mixed within 1e-6 (norm-wise relative);
* the 256-system ensemble (N log-uniform in [200, 2000], a pulse of area
  uniform in [pi/4, pi] per system), one batch context, autotuned, 5000 steps
  in both precisions: three members (smallest, median and largest N) against
  the oracle every 1000 steps in FP64 within 1e-10, and every system's mixed
  rho01 within 1e-6 of its FP64 rho01 (relative to the system's max |rho01|),
  populations within 1e-6 absolute.

Both tests run the per-step pair method (QUEST_PAIR_ON_THE_FLY) and the
time-blocked one (QUEST_PAIR_TIME_BLOCKED) to the same bounds.  The N = 65536
and N = 131072 workloads also run for their full lengths (2000 and 1000
steps) in both precisions, mixed against FP64.
"""
import math
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


METHODS = pytest.mark.parametrize("method", [0, 2], ids=["per_step", "time_blocked"])


@METHODS
@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-6)])
def test_n64_full_run_matches_reference_method_fixture(precision, tol, method):
    from quest_amd import Context, systems
    g = np.load(os.path.join(GOLDEN, "n64_traj.npz"))
    sysd, steps = systems.workload("n64")
    every = int(g["record_every"])
    assert steps == int(g["steps"]) == 20000
    worst = 0.0
    with Context(sysd, systems.consts(), precision=precision, pulse=systems.pulse(math.pi / 2)) as ctx:
        ctx.set_pc_weights(g["ps"], g["cs"], float(g["fc"]))
        ctx.set_pair_method(method)
        for r in range(steps // every):
            ctx.run(every)
            rho, drho = ctx.state()
            worst = max(worst, rel_err(rho, g["rec"][r]))
        final_d = rel_err(drho, g["yp"])
    print(f"N=64 x {steps} steps, method {method}, precision {precision}: worst rel err {worst:.3e} over "
          f"{steps // every} records, final derivative {final_d:.3e}, max|rho01| "
          f"{np.abs(g['rec'][:, :, 1]).max():.3f}")
    assert worst < tol
    assert final_d < 10 * tol


def _member_index():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.ensemble_member_index()


@pytest.mark.timeout(300)
@METHODS
def test_ensemble_256_systems_5000_steps_both_precisions(method):
    from quest_amd import Context, systems
    g = np.load(os.path.join(GOLDEN, "ensemble_members.npz"))
    idx = _member_index()
    assert np.array_equal(idx, g["index"])
    ens = systems.ensemble()
    sys_list = [s for s, _ in ens]
    pulses = [p for _, p in ens]
    sizes = [len(s["pos"]) for s in sys_list]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    steps, every = int(g["steps"]), int(g["record_every"])
    consts = systems.consts()
    final = {}
    for precision in (0, 1):
        with Context(sys_list, consts, precision=precision, pulse=pulses) as ctx:
            ctx.set_pc_weights(g["ps"], g["cs"], float(g["fc"]))
            ctx.set_pair_method(method)
            ctx.autotune(2)
            assert ctx.info()["systems"] == 256
            worst = 0.0
            for r in range(steps // every):
                ctx.run(every)
                rho, _ = ctx.state()
                assert np.all(np.isfinite(rho))
                if precision == 0:
                    for j, i in enumerate(idx):
                        worst = max(worst, rel_err(rho[offs[i]:offs[i + 1]], g[f"rec{j}"][r]))
            final[precision] = rho
            if precision == 0:
                print(f"method {method}: FP64 members {list(idx)} (n = {[sizes[i] for i in idx]}) vs oracle: "
                      f"worst rel err {worst:.3e}")
                assert worst < 1e-10
    e01, e00 = 0.0, 0.0
    for s in range(256):
        a, b = final[1][offs[s]:offs[s + 1]], final[0][offs[s]:offs[s + 1]]
        e01 = max(e01, float(np.abs(a[:, 1] - b[:, 1]).max() / np.abs(b[:, 1]).max()))
        e00 = max(e00, float(np.abs(a[:, 0] - b[:, 0]).max()))
    print(f"ensemble at step {steps}: max|rho01| {np.abs(final[0][:, 1]).max():.3f}; mixed vs FP64 "
          f"rho01 {e01:.2e} (worst system, relative), rho00 {e00:.2e} absolute")
    assert np.abs(final[0][:, 1]).max() > 0.1
    assert e01 < 1e-6 and e00 < 1e-6


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", ["n65536", "n131072"])
def test_large_workload_full_length_mixed_vs_fp64(name):
    """The N = 65536 workload for all 2000 steps and the N = 131072 one for all
    1000, in both precisions with the time-blocked method (the bench's): the
    state stays finite at every record, and at the last step the mixed rho01
    is within 1e-6 of the FP64 rho01 (norm-wise relative; measured 1.0e-8 at
    N = 65536, 7.2e-9 at N = 131072) and the populations within 1e-6
    absolute (measured 2e-15).  The pi/2 pulse peaks at 0.5 ps, after these
    runs end, so |rho01| reaches only its leading tail (1.4e-3 and 3.8e-5 at
    the last step); the accuracy under a strong drive is the 100-step test in
    test_gpu_parity.py."""
    from quest_amd import Context, pc_weights, systems
    sysd, steps = systems.workload(name)
    every = 500
    final = {}
    for precision in (0, 1):
        with Context(sysd, systems.consts(), precision=precision, pulse=systems.pulse(math.pi / 2)) as ctx:
            ctx.set_pc_weights(*pc_weights(18, 22, 3.15))
            ctx.set_pair_method(2)
            for _ in range(steps // every):
                ctx.run(every)
                rho, _ = ctx.state()
                assert np.all(np.isfinite(rho))
            assert ctx.step == steps
            final[precision] = rho
    e01 = rel_err(final[1][:, 1], final[0][:, 1])
    e00 = float(np.abs(final[1][:, 0] - final[0][:, 0]).max())
    print(f"{name} x {steps} steps: max|rho01| {np.abs(final[0][:, 1]).max():.3e}; mixed vs FP64 rho01 {e01:.2e} "
          f"(relative), rho00 {e00:.2e} (absolute)")
    assert np.abs(final[0][:, 1]).max() > 1e-9
    assert e01 < 1e-6 and e00 < 1e-6
