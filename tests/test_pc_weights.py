"""PredictorCorrector::Weights (integrator.cpp:5-102): the oracle's numpy
restatement and the product's host Jacobi-SVD solver, pinned by residuals and
exactness properties (the exact Eigen JacobiSVD weights are parity unpinned:
the system has cond ~ 3e13, so valid solvers differ by up to ~1e-2)."""
import numpy as np
import pytest


def _both():
    import pc_weights
    from quest_amd import pc_weights as product
    return {"oracle": pc_weights.weights(18, 22, 3.15), "product": product(18, 22, 3.15)}


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_residuals(which):
    import pc_weights
    ps, cs, fc = _both()[which]
    rp, rc = pc_weights.residuals(ps, cs, fc)
    assert rp < 5e-13 and rc < 5e-13


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_exact_on_low_degree_polynomials(which):
    """A multistep formula fitted to exp(lambda t) near lambda = 0 is exact on
    polynomials of low degree: y_f = sum w y_j + sum w' dt y'_j (+ fc dt y'_f)."""
    ps, cs, fc = _both()[which]
    n = 22
    tj = np.arange(n) - n        # steps -22..-1 relative to the new step 0
    for deg in range(4):
        y = tj.astype(float) ** deg
        yp = deg * tj.astype(float) ** (deg - 1) if deg else np.zeros(n)
        target = 0.0 ** deg
        pred = ps[0] @ y + ps[1] @ yp
        corr = cs[0] @ y + cs[1] @ yp + fc * (deg * 0.0 ** (deg - 1) if deg else 0.0)
        scale = max(1.0, np.abs(y).max())
        assert abs(pred - target) < 1e-10 * scale, (deg, pred)
        assert abs(corr - target) < 1e-10 * scale, (deg, corr)


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_exact_on_slow_exponentials(which):
    ps, cs, fc = _both()[which]
    n = 22
    for mu in (0.01, -0.02, 0.05j, -0.01 + 0.03j):
        tj = (np.arange(n) - n).astype(float)
        y, yp = np.exp(mu * tj), mu * np.exp(mu * tj)
        assert abs(ps[0] @ y + ps[1] @ yp - 1.0) < 1e-9
        assert abs(cs[0] @ y + cs[1] @ yp + fc * mu - 1.0) < 1e-9


def test_shapes_and_scaling():
    """ps, cs are [2][22]; row 0 sums to one; fc carries the step factor 10.5."""
    for ps, cs, fc in _both().values():
        assert ps.shape == (2, 22) and cs.shape == (2, 22)
        assert abs(ps[0].sum() - 1) < 1e-12 and abs(cs[0].sum() - 1) < 1e-12
        assert abs(fc - 0.2815) < 1e-3


def test_solvers_agree_within_conditioning():
    """Both solvers are backward stable; their weights differ only at the
    level the conditioning allows."""
    w = _both()
    dp = np.abs(w["oracle"][0] - w["product"][0]).max()
    dc = np.abs(w["oracle"][1] - w["product"][1]).max()
    assert dp < 5e-2 and dc < 5e-3


def test_trajectory_insensitive_to_the_valid_weight_sets(oracle_lib):
    """The two backward-stable weight sets (numpy gelsd, the product's Jacobi
    SVD) drive the FP64 oracle to the same trajectory far below the parity
    tolerances (1e-10 relative): the weights themselves stay parity unpinned
    against Eigen's JacobiSVD, but any valid set gives the same physics.
    A driven, coupled 48-dot system over 600 steps (the pulse passes)."""
    from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse
    from quest_amd import systems
    w = _both()
    sysd = dense_system(48, 0.24, seed=31)
    pl = oracle_pulse(oracle_lib, short_pulse())
    p = oracle_params(oracle_lib, systems.consts())
    runs = {k: oracle_lib.solve(p, sysd, pl, *w[k], 600, mode=1) for k in w}
    assert all(r["status"] == 0 for r in runs.values())
    y_o, y_p = runs["oracle"]["y"], runs["product"]["y"]
    assert np.abs(y_o[:, 1]).max() > 0.1
    e = rel_err(y_p, y_o)
    print(f"weights differ by {np.abs(w['oracle'][0] - w['product'][0]).max():.1e} (predictor); "
          f"trajectories by {e:.1e}")
    assert e < 1e-11
