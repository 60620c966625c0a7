"""Seeded generators of the dot systems the benchmarks and tests run.

The reference ships one fixed dot file (examples/dots.cfg) and a generator
seeded from the wall clock (tools/point_gen.cpp:54), so these systems are
defined here: dots uniform in a centred cube of side L = (N / 512)^(1/3) um
(number density 512 um^-3), constants of examples/input.cfg:7-10, T1 = 10 ps,
T2 = 20 ps, |d| = 5.2917721e-4 e um, and a Gaussian pulse along x polarised
along z with sigma_t = 0.1 ps, delay 0.5 ps and amplitude set for a given
pulse area on an isolated dot.
"""
from __future__ import annotations

import math

import numpy as np

C0 = 299.792458          # um / ps
HBAR = 0.65821193        # meV ps
MU0 = 2.0133545e-4       # meV ps^2 / (e^2 um)
OMEGA0 = 2278.9          # rad / ps
DIPOLE = 5.2917721e-4    # e um
T1, T2 = 10.0, 20.0
DT = 1e-4                # ps
ORDER = 4
WINDOW = 22
DENSITY = 512.0          # um^-3
SIGMA_T = 0.1            # ps
PULSE_DELAY = 0.5        # ps


def consts(dt: float = DT, laser_omega: float = OMEGA0) -> dict:
    return {"c0": C0, "hbar": HBAR, "mu0": MU0, "laser_omega": laser_omega, "dt": dt}


def pulse_amplitude(area: float, dipole: float = DIPOLE, sigma_t: float = SIGMA_T) -> float:
    """E0 such that the pulse rotates an isolated resonant dot by `area`.

    With Omega(t) = (E0 / 2) |d| g(t) / hbar and the Bloch rate 2 Omega
    (quantum_dot.cpp:13-17), area = E0 |d| sigma_t sqrt(2 pi) / hbar.
    """
    return area * HBAR / (dipole * sigma_t * math.sqrt(2 * math.pi))


def pulse(area: float = math.pi / 2, omega: float = OMEGA0) -> dict:
    """pulse.cfg fields: E0 delay width freq k pol (README.md:64-75)."""
    return {"amplitude": pulse_amplitude(area), "delay": PULSE_DELAY, "width": omega * SIGMA_T,
            "freq": omega, "wavevector": [omega / C0, 0.0, 0.0], "polarization": [0.0, 0.0, 1.0]}


def cube_side(n: int) -> float:
    return (n / DENSITY) ** (1.0 / 3.0)


def system(n: int, seed: int, random_dipoles: bool = False, detuning_sigma: float = 0.0) -> dict:
    """n dots: positions uniform in [-L/2, L/2)^3, dipoles along z (or random
    unit vectors), omega0 = 2278.9 rad/ps (+ Gaussian detuning)."""
    rng = np.random.default_rng(seed)
    L = cube_side(n)
    pos = rng.uniform(-L / 2, L / 2, size=(n, 3))
    if random_dipoles:
        v = rng.normal(size=(n, 3))
        dip = DIPOLE * v / np.linalg.norm(v, axis=1, keepdims=True)
    else:
        dip = np.zeros((n, 3))
        dip[:, 2] = DIPOLE
    freq = np.full(n, OMEGA0)
    if detuning_sigma > 0:
        freq = freq + rng.normal(0.0, detuning_sigma, size=n)
    return {"pos": pos, "dip": dip, "freq": freq, "t1": np.full(n, T1), "t2": np.full(n, T2)}


# The workloads: (n, steps, random dipoles, detuning sigma).
WORKLOADS = {
    "n64": dict(n=64, steps=20000, random_dipoles=False, detuning_sigma=0.0, seed=1),
    "n4096": dict(n=4096, steps=5000, random_dipoles=False, detuning_sigma=0.0, seed=2),
    "n65536": dict(n=65536, steps=2000, random_dipoles=True, detuning_sigma=2.0, seed=3),
    "n131072": dict(n=131072, steps=1000, random_dipoles=True, detuning_sigma=2.0, seed=5),
    # coefficient-table measurements: the largest size whose FP64 table (22 GB) is
    # far from the memory limit while streaming for milliseconds per step
    "n16384": dict(n=16384, steps=2000, random_dipoles=True, detuning_sigma=2.0, seed=6),
}


def workload(name: str) -> tuple[dict, int]:
    w = WORKLOADS[name]
    return system(w["n"], w["seed"], w["random_dipoles"], w["detuning_sigma"]), w["steps"]


def ensemble(count: int = 256, n_lo: int = 200, n_hi: int = 2000, seed: int = 4):
    """Independent systems with N log-uniform in [n_lo, n_hi], per-system seeds,
    random dipoles, detuning sigma 2 rad/ps, pulse area uniform in [pi/4, pi]."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), size=count)).astype(int)
    areas = rng.uniform(math.pi / 4, math.pi, size=count)
    out = []
    for i, (n, area) in enumerate(zip(ns, areas)):
        s = system(int(n), seed * 100003 + i, random_dipoles=True, detuning_sigma=2.0)
        out.append((s, pulse(float(area))))
    return out


def min_distance(pos: np.ndarray) -> float:
    """Smallest pair distance (O(N^2) in blocks; for the record)."""
    best = np.inf
    n = pos.shape[0]
    for i in range(0, n, 2048):
        d = np.linalg.norm(pos[i:i + 2048, None, :] - pos[None, :, :], axis=-1)
        idx = np.arange(i, min(i + 2048, n))
        d[idx - i, idx] = np.inf
        best = min(best, float(d.min()))
    return best
