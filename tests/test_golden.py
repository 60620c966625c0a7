"""The committed oracle fixtures (tests/golden/) against the live oracle.

Each fixture is recomputed in part here, so that a change to the oracle or to
the seeded generators that the fixture no longer matches fails on the CPU.
"""
import os

import numpy as np

from conftest import OBS_STEP, observers, oracle_params, rel_err, smooth_history

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_n131072_observer_fixture_matches_oracle(oracle_lib):
    from quest_amd import systems
    g = np.load(os.path.join(GOLDEN, "n131072_observers.npz"))
    sysd, _ = systems.workload("n131072")
    n = sysd["pos"].shape[0]
    assert int(g["step"]) == OBS_STEP
    assert np.array_equal(g["observers"], observers(n))
    pick = np.array([0, 1, 2047, 4095])
    hist = smooth_history(n, OBS_STEP + 1)
    got = oracle_lib.evaluate_observers(oracle_params(oracle_lib, systems.consts()), sysd["pos"],
                                        sysd["dip"], hist, OBS_STEP, g["observers"][pick], kind=0)
    assert rel_err(got, g["field"][pick]) < 1e-14


def test_n4096_fixture_header():
    g = np.load(os.path.join(GOLDEN, "n4096_traj.npz"))
    assert int(g["steps"]) == 5000 and int(g["record_every"]) == 500
    assert g["y"].shape == (4096, 2) and np.isfinite(g["y"]).all()
