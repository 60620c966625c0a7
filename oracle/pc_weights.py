"""Predictor-corrector weights, restated in numpy (TEST INFRASTRUCTURE ONLY).

Restates PredictorCorrector::Weights and WeightsBuilder
(src/integrator.cpp:5-102): the weights of a 22-step linear multistep
predictor and corrector fitted, by complex minimum-norm least squares, to be
exact on exponentials exp(lambda t) for 18 lambdas on a "semidisk" contour of
radius 3.15 (main.cpp:49).

The reference solves with Eigen's JacobiSVD.  The system is ill-conditioned
(cond ~ 3e13), so any two backward-stable solvers agree on the residual
(~1e-13) but not on the weights (up to ~1e-3): the exact Eigen weights are
parity unpinned.  This module uses LAPACK gelsd (numpy.linalg.lstsq).  Its
pins are in tests/test_pc_weights.py: the residual, partition of unity of the
value weights, exactness on low-degree polynomials and on exp(mu t), and the
shapes and step-factor scaling of integrator.cpp:95-101.
"""
from __future__ import annotations

import numpy as np


def semidisk(t: float) -> complex:
    """integrator.cpp:5-14."""
    if 0 <= t < 1:
        return 1j * t
    if 1 <= t < np.pi + 1:
        return np.exp(1j * (t + np.pi / 2 - 1))
    if np.pi + 1 <= t < np.pi + 2:
        return 1j * (t - (np.pi + 2))
    return 0j


def builder(n_lambda: int, n_time: int, radius: float):
    """integrator.cpp:34-43: lambdas, times, future time."""
    xs = np.linspace(0, np.pi + 2, n_lambda + 1)
    lambdas = np.array([radius * semidisk(x) for x in xs[:n_lambda]])
    times = np.linspace(-1, 1, n_time)
    timestep = 2.0 / (n_time - 1)
    future_time = 1 + timestep
    return lambdas, times, future_time


def predictor_matrix(lambdas, times):
    """integrator.cpp:45-56: [exp(l t_j) | l exp(l t_j)]."""
    b = np.exp(np.outer(lambdas, times))
    return np.hstack([b, b * lambdas[:, None]])


def corrector_matrix(lambdas, times, future_time):
    """integrator.cpp:58-67: predictor matrix plus the l exp(l t_f) column."""
    return np.hstack([predictor_matrix(lambdas, times),
                      (lambdas * np.exp(lambdas * future_time))[:, None]])


def rhs_vector(lambdas, future_time):
    """integrator.cpp:69-73."""
    return np.exp(lambdas * future_time)


def min_norm_solve(mat, rhs):
    """integrator.cpp:75-83, with LAPACK gelsd in place of JacobiSVD; real part."""
    sol, *_ = np.linalg.lstsq(mat, rhs, rcond=None)
    return sol.real


def weights(n_lambda: int = 18, n_time: int = 22, radius: float = 3.15):
    """integrator.cpp:85-102.

    Returns (ps, cs, future_coef): ps and cs are [2][n_time] (row 0 multiplies
    the state, row 1 the derivative times dt), with row 1 and future_coef
    scaled by the step factor (n_time - 1) / 2.
    """
    lambdas, times, tf = builder(n_lambda, n_time, radius)
    step_factor = (n_time - 1) / 2.0
    pred = min_norm_solve(predictor_matrix(lambdas, times), rhs_vector(lambdas, tf))
    corr = min_norm_solve(corrector_matrix(lambdas, times, tf), rhs_vector(lambdas, tf))
    # Eigen::Map<ArrayXXd>(data, n_time, 2).transpose(): row j = data[j*n_time:(j+1)*n_time]
    ps = pred[: 2 * n_time].reshape(2, n_time).copy()
    cs = corr[: 2 * n_time].reshape(2, n_time).copy()
    ps[1] *= step_factor
    cs[1] *= step_factor
    future_coef = corr[2 * n_time] * step_factor
    return ps, cs, float(future_coef)


def residuals(ps, cs, future_coef, n_lambda=18, n_time=22, radius=3.15):
    """|M w - b| of the unscaled real weights for the predictor and corrector."""
    lambdas, times, tf = builder(n_lambda, n_time, radius)
    step_factor = (n_time - 1) / 2.0
    wp = np.concatenate([ps[0], ps[1] / step_factor])
    wc = np.concatenate([cs[0], cs[1] / step_factor, [future_coef / step_factor]])
    b = rhs_vector(lambdas, tf)
    rp = np.abs(predictor_matrix(lambdas, times) @ wp - b).max()
    rc = np.abs(corrector_matrix(lambdas, times, tf) @ wc - b).max()
    return float(rp), float(rc)
