"""Writes the oracle fixtures the GPU parity tests compare against.

    python tests/golden/make_golden.py [n4096] [n131072]

n4096_traj.npz: the N = 4096 system (quest_amd.systems.workload("n4096"):
512 um^-3, dipoles along z, pi/2 pulse, dt = 1e-4 ps, order 4) integrated for
5000 steps by the FP64 oracle (split-gather mode, which reproduces the
reference method to 1e-13 -- tests/test_oracle.py::test_solve_modes_agree),
with the oracle's PC weights.  Stores rho00/rho01 of every dot every 500
steps and the final derivative.

n131072_observers.npz: the N = 131072 system (quest_amd.systems.workload(
"n131072"): 512 um^-3, random dipoles, order 4), its smooth synthetic rho01
history (tests/conftest.py smooth_history) and the FP64 oracle's evaluation
(rotating frame, no table) at step 380 for 4096 seeded observers
(conftest.observers).  About two minutes on 8 threads.
"""
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def make_n4096(path):
    import oracle
    import pc_weights
    from quest_amd import systems
    oracle.build()
    oracle.set_threads(os.cpu_count() or 1)
    sysd, steps = systems.workload("n4096")
    c = systems.consts()
    pl = systems.pulse(math.pi / 2)
    ps, cs, fc = pc_weights.weights(18, 22, 3.15)
    p = oracle.params(c["c0"], c["hbar"], c["mu0"], c["laser_omega"], c["dt"], 4, 22)
    t = time.time()
    r = oracle.solve(p, sysd, oracle.pulse(pl["amplitude"], pl["delay"], pl["width"], pl["freq"],
                                           pl["wavevector"], pl["polarization"]),
                     ps, cs, fc, steps, mode=1, record_every=500)
    assert r["status"] == 0
    np.savez_compressed(path, rec=r["rec"], y=r["y"], yp=r["yp"], steps=steps, record_every=500,
                        ps=ps, cs=cs, fc=fc)
    print(f"wrote {path} in {time.time() - t:.0f} s; max|rho01| {np.abs(r['y'][:, 1]).max():.4f}")


def make_n131072_observers(path):
    import oracle
    from quest_amd import systems
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import OBS_STEP, observers, oracle_params, smooth_history
    oracle.build()
    oracle.set_threads(os.cpu_count() or 1)
    sysd, _ = systems.workload("n131072")
    n = sysd["pos"].shape[0]
    hist = smooth_history(n, OBS_STEP + 1)
    obs = observers(n)
    p = oracle_params(oracle, systems.consts())
    t = time.time()
    field = oracle.evaluate_observers(p, sysd["pos"], sysd["dip"], hist, OBS_STEP, obs, kind=0)
    np.savez_compressed(path, observers=obs, field=field, step=OBS_STEP)
    print(f"wrote {path} in {time.time() - t:.0f} s; max|field| {np.abs(field).max():.4e}")


if __name__ == "__main__":
    which = sys.argv[1:] or ["n4096"]
    if "n4096" in which:
        make_n4096(os.path.join(HERE, "n4096_traj.npz"))
    if "n131072" in which:
        make_n131072_observers(os.path.join(HERE, "n131072_observers.npz"))
