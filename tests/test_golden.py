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


def _solve(oracle, sysd, pl, g, steps, mode, threads):
    from quest_amd import systems
    oracle.set_threads(threads)
    c = systems.consts()
    p = oracle.params(c["c0"], c["hbar"], c["mu0"], c["laser_omega"], c["dt"], 4, 22)
    qp = oracle.pulse(pl["amplitude"], pl["delay"], pl["width"], pl["freq"], pl["wavevector"],
                      pl["polarization"])
    r = oracle.solve(p, sysd, qp, g["ps"], g["cs"], float(g["fc"]), steps, mode=mode)
    assert r["status"] == 0
    return r["y"]


def test_n4096_fixture_first_record_matches_oracle(oracle_lib):
    """The first record (step 500) of the N = 4096 trajectory, recomputed."""
    import math
    from quest_amd import systems
    g = np.load(os.path.join(GOLDEN, "n4096_traj.npz"))
    assert int(g["steps"]) == 5000 and int(g["record_every"]) == 500
    assert g["y"].shape == (4096, 2) and np.isfinite(g["y"]).all()
    sysd, _ = systems.workload("n4096")
    y = _solve(oracle_lib, sysd, systems.pulse(math.pi / 2), g, 500, 1, os.cpu_count() or 1)
    assert rel_err(y, g["rec"][0]) < 1e-14
    assert np.abs(y[:, 1]).max() > 0


def test_n64_fixture_first_record_matches_oracle(oracle_lib):
    """Config 0 (N = 64, 20000 steps, reference method): header, the pulse
    passes through (rho01 rises and the populations move), and the first
    record (step 2000) recomputed by the reference method."""
    import math
    from quest_amd import systems
    g = np.load(os.path.join(GOLDEN, "n64_traj.npz"))
    assert int(g["steps"]) == 20000 and int(g["record_every"]) == 2000
    assert g["rec"].shape == (10, 64, 2) and np.isfinite(g["rec"]).all()
    assert np.abs(g["rec"][:, :, 1]).max() > 0.4          # a pi/2 pulse: |rho01| near 1/2
    assert np.abs(g["rec"][-1, :, 0] - 1).max() > 0.3      # excited population
    sysd, _ = systems.workload("n64")
    y = _solve(oracle_lib, sysd, systems.pulse(math.pi / 2), g, 2000, 0, 1)
    assert rel_err(y, g["rec"][0]) < 1e-14


def test_ensemble_fixture_first_record_matches_oracle(oracle_lib):
    """The three ensemble members (smallest, median, largest N): indices as
    the generator picks them; the smallest member's first record (step 1000)
    recomputed."""
    import importlib.util
    from quest_amd import systems
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = np.load(os.path.join(GOLDEN, "ensemble_members.npz"))
    idx = mg.ensemble_member_index()
    assert np.array_equal(g["index"], idx)
    ens = systems.ensemble()
    ns = [ens[i][0]["pos"].shape[0] for i in idx]
    assert ns[0] == min(s["pos"].shape[0] for s, _ in ens) and ns[2] == max(s["pos"].shape[0] for s, _ in ens)
    for j, n in enumerate(ns):
        assert g[f"rec{j}"].shape == (5, n, 2) and np.isfinite(g[f"rec{j}"]).all()
    sysd, pl = ens[idx[0]]
    y = _solve(oracle_lib, sysd, pl, g, 1000, 1, os.cpu_count() or 1)
    assert rel_err(y, g["rec0"][0]) < 1e-14
