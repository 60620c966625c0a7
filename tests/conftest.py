import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ref_weights():
    """Oracle PC weights (numpy gelsd restatement of integrator.cpp:85-102)."""
    import pc_weights
    return pc_weights.weights(18, 22, 3.15)


def short_pulse(area=math.pi / 2, delay=0.02, sigma_t=0.005, omega=2278.9):
    """A pulse that fits in a few hundred steps of dt = 1e-4 ps."""
    from quest_amd import systems
    amp = area * systems.HBAR / (systems.DIPOLE * sigma_t * math.sqrt(2 * math.pi))
    return {"amplitude": amp, "delay": delay, "width": omega * sigma_t, "freq": omega,
            "wavevector": [omega / systems.C0, 0.0, 0.0], "polarization": [0.0, 0.0, 1.0]}


def dense_system(n, side, seed, random_dipoles=True, detuning=2.0, plant_near=True):
    """n dots in a cube of the given side (um); optionally two planted close pairs."""
    from quest_amd import systems
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-side / 2, side / 2, size=(n, 3))
    if plant_near and n >= 4:
        pos[1] = pos[0] + np.array([0.011, 0.004, -0.006])   # delay 0 (r < c dt)
        pos[3] = pos[2] + np.array([0.0, 0.025, 0.02])       # delay 1
    if random_dipoles:
        v = rng.normal(size=(n, 3))
        dip = systems.DIPOLE * v / np.linalg.norm(v, axis=1, keepdims=True)
    else:
        dip = np.zeros((n, 3))
        dip[:, 2] = systems.DIPOLE
    freq = systems.OMEGA0 + (rng.normal(0, detuning, size=n) if detuning > 0 else 0.0)
    return {"pos": pos, "dip": dip, "freq": np.broadcast_to(freq, (n,)).copy(),
            "t1": np.full(n, systems.T1), "t2": np.full(n, systems.T2)}


def oracle_params(oracle, consts, order=4, window=22):
    return oracle.params(consts["c0"], consts["hbar"], consts["mu0"], consts["laser_omega"],
                         consts["dt"], order, window)


def oracle_pulse(oracle, pl):
    return oracle.pulse(pl["amplitude"], pl["delay"], pl["width"], pl["freq"],
                        pl["wavevector"], pl["polarization"])


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# The full-size (N = 131072) single-evaluation fixture: a smooth synthetic
# rho01 history of every dot, evaluated at OBS_STEP for OBS_COUNT observers.
OBS_STEP = 380        # beyond the largest delay (~370 steps) plus the stencil
OBS_COUNT = 4096


def smooth_history(n, n_t, seed=7):
    """rho01[t][b], t = 0 .. n_t - 1: a_b (1 - exp(-t / tau_b)) exp(i (phi_b +
    delta_b t dt)) with a_b in [0.05, 0.5], tau_b in [20, 200] steps, detuning
    delta_b ~ N(0, 2 rad/ps), dt = 1e-4 ps (a rotating-frame envelope)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.05, 0.5, n)
    tau = rng.uniform(20.0, 200.0, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    delta = rng.normal(0.0, 2.0, n) * 1e-4
    t = np.arange(n_t, dtype=np.float64)[:, None]
    return (a * -np.expm1(-t / tau)) * np.exp(1j * (phi + delta * t))


def observers(n, count=OBS_COUNT, seed=8):
    """count distinct dots of n, sorted (seeded)."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n, size=count, replace=False)).astype(np.int32)
