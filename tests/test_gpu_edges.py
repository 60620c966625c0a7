"""GPU: the smallest systems and ragged batches, both pair methods and both
precisions, against the FP64 oracle (reference method, mode 0).

* one dot (no pairs: the pulse and the Bloch equations only; no work items),
  two and three dots, packed (every delay below 4: zero-delay list and short
  list only, the time-blocked pass has no live pair) and spread (delays of
  20 .. 60 steps: the pass carries the pairs);
* a batch of systems of 1, 2, 1, 3 and 50 dots, and a batch of 1-dot systems
  only (no work items at all).

400 steps through a short strong pulse (delay 0.02 ps, sigma 0.005 ps).
Tolerances: FP64 1e-12 relative (norm-wise), mixed 1e-6 absolute (|rho| <= 1).
"""
import math

import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse

pytestmark = pytest.mark.gpu

STEPS = 400
METHODS = pytest.mark.parametrize("method", [0, 2], ids=["per_step", "time_blocked"])
PRECISIONS = pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])


def _oracle_run(oracle, consts, sysd, pl, weights):
    ps, cs, fc = weights
    ref = oracle.solve(oracle_params(oracle, consts), sysd, oracle_pulse(oracle, pl), ps, cs, fc, STEPS, mode=0)
    assert ref["status"] == 0
    return ref


def _err(precision, got, want):
    return rel_err(got, want) if precision == 0 else float(np.abs(got - want).max())


def _tol(precision):
    return 1e-12 if precision == 0 else 1e-6


@METHODS
@PRECISIONS
@pytest.mark.parametrize("n,side", [(1, 0.05), (2, 0.05), (3, 0.05), (2, 1.2), (3, 1.5)],
                         ids=["1", "2_packed", "3_packed", "2_spread", "3_spread"])
def test_smallest_systems_match_oracle(oracle_lib, ref_weights, n, side, precision, method):
    from quest_amd import Context, systems
    consts = systems.consts()
    sysd = dense_system(n, side, seed=40 + n, plant_near=False)
    pl = short_pulse(area=math.pi)
    ref = _oracle_run(oracle_lib, consts, sysd, pl, ref_weights)
    with Context(sysd, consts, precision=precision, pulse=pl) as ctx:
        ctx.set_pc_weights(*ref_weights)
        ctx.set_pair_method(method)
        info = ctx.info()
        ctx.run(STEPS)
        rho, _ = ctx.state()
    err = _err(precision, rho, ref["y"])
    print(f"N={n} side {side} um, method {method}, precision {precision}: err {err:.2e}, "
          f"work items {info['work_items']}, max|rho01| {np.abs(ref['y'][:, 1]).max():.3f}")
    assert (info["work_items"] == 0) == (n == 1)
    assert np.abs(ref["y"][:, 1]).max() > 0.05
    assert err < _tol(precision)


@METHODS
@PRECISIONS
@pytest.mark.parametrize("sizes", [[1, 2, 1, 3, 50], [1, 1, 1, 1]], ids=["ragged", "ones"])
def test_ragged_batches_match_oracle(oracle_lib, ref_weights, sizes, precision, method):
    from quest_amd import Context, systems
    consts = systems.consts()
    syss = [dense_system(n, 0.6 * max(n, 2) ** (1 / 3), seed=60 + i, plant_near=n >= 4)
            for i, n in enumerate(sizes)]
    pls = [short_pulse(area=a) for a in np.linspace(math.pi / 4, math.pi, len(sizes))]
    with Context(syss, consts, precision=precision, pulse=pls) as ctx:
        ctx.set_pc_weights(*ref_weights)
        ctx.set_pair_method(method)
        info = ctx.info()
        ctx.run(STEPS)
        rho, _ = ctx.state()
    assert info["systems"] == len(sizes) and info["n"] == sum(sizes)
    assert (info["work_items"] == 0) == (max(sizes) == 1)
    off = 0
    for s, pl, n in zip(syss, pls, sizes):
        ref = _oracle_run(oracle_lib, consts, s, pl, ref_weights)
        err = _err(precision, rho[off:off + n], ref["y"])
        print(f"batch member of {n}: method {method}, precision {precision}: err {err:.2e}")
        assert err < _tol(precision)
        off += n
