"""Pins of the FP64 CPU oracle against facts other than itself: the
reference's own known-answer tests, closed forms, invariants and brute force.
Each test names the reference lines the pinned function restates."""
import math

import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err, short_pulse

C0, HBAR, MU0, OMEGA = 299.792458, 0.65821193, 2.0133545e-4, 2278.9013


def _params(oracle, order=4, dt=1e-4, omega=OMEGA):
    return oracle.params(C0, HBAR, MU0, omega, dt, order, 22)


# ----------------------------------------------------- lagrange_set.cpp:17-41

@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("x", [0.013, 0.25, 0.5, 0.731, 0.999])
def test_lagrange_reproduces_polynomials(oracle_lib, order, x):
    """Weights interpolate, and differentiate, every polynomial of degree <=
    order exactly: sum w0 f(m) = f(x), sum w1 f(m) = -f'(x)/dt (node m is
    m steps in the past), sum w2 f(m) = f''(x)/dt^2."""
    dt = 0.37
    w = oracle_lib.lagrange_weights(x, order, dt)
    nodes = np.arange(order + 1, dtype=float)
    rng = np.random.default_rng(order)
    for deg in range(order + 1):
        c = rng.normal(size=deg + 1)
        f = np.polynomial.Polynomial(c)
        assert abs(w[0] @ f(nodes) - f(x)) < 1e-12 * max(1, abs(f(x)))
        assert abs(w[1] @ f(nodes) - (-f.deriv(1)(x) / dt)) < 1e-10
        d2 = f.deriv(2)(x) / dt ** 2 if deg >= 2 else 0.0
        assert abs(w[2] @ f(nodes) - d2) < 1e-9


def test_lagrange_partition_of_unity_and_shapes(oracle_lib):
    for order in (3, 4, 5):
        w = oracle_lib.lagrange_weights(0.3, order)
        assert w.shape == (3, order + 1)      # lagrange_set_test.cpp:6-23
        assert abs(w[0].sum() - 1) < 1e-14
        assert abs(w[1].sum()) < 1e-13 and abs(w[2].sum()) < 1e-12


def test_lagrange_matches_numpy_fit(oracle_lib):
    """Brute force: the weights are the rows of the inverse Vandermonde
    evaluated at x (numpy polyfit through unit vectors)."""
    order, x = 4, 0.42
    w = oracle_lib.lagrange_weights(x, order)
    nodes = np.arange(order + 1)
    for m in range(order + 1):
        e = np.zeros(order + 1)
        e[m] = 1
        p = np.polynomial.Polynomial.fit(nodes, e, order, domain=[0, order], window=[0, order])
        assert abs(p(x) - w[0][m]) < 1e-12
        assert abs(-p.deriv(1)(x) - w[1][m]) < 1e-11
        assert abs(p.deriv(2)(x) - w[2][m]) < 1e-10


def test_lagrange_singular_at_node(oracle_lib):
    """The reference formula divides by x - m: at x = 0 the derivative
    weights are not finite (lagrange_set.cpp:25,33)."""
    w = oracle_lib.lagrange_weights(0.0, 4)
    assert not np.all(np.isfinite(w[1]))


# ------------------------------------------ history_interaction.cpp:72-97

def test_idx2coord_enumerates_lower_triangle(oracle_lib):
    n = 90
    idx = 0
    for row in range(1, n):
        for col in range(row):
            assert oracle_lib.idx2coord(idx) == (row, col)
            idx += 1


def test_idx2coord_large_index(oracle_lib):
    """Beyond the reference's int range (8*idx overflows for n >= 23171)."""
    row = 131071
    idx = row * (row - 1) // 2 + 17
    assert oracle_lib.idx2coord(idx) == (row, 17)


def test_split_delay_cases(oracle_lib):
    cdt = C0 * 1e-4
    n, x = oracle_lib.split_delay([0, 0, 0], [2.5 * cdt, 0, 0], C0, 1e-4)
    assert n == 2 and abs(x - 0.5) < 1e-12
    n, x = oracle_lib.split_delay([1, 1, 1], [1, 1, 1 + 0.3 * cdt], C0, 1e-4)
    assert n == 0 and abs(x - 0.3) < 1e-12
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)
        d = b - a
        tau = math.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) / (C0 * 1e-4)
        n, x = oracle_lib.split_delay(a, b, C0, 1e-4)
        assert n == math.floor(tau) and x == tau - math.floor(tau)


# ---------------------------- rotating_green_function.cpp / green_function.*

def _w(oracle, x, order=4, dt=1e-4):
    return oracle.lagrange_weights(x, order, dt)


def test_rotating_coefs_analytic_case(oracle_lib):
    """d || z, dr || x: d.(I - 3 rr)d = d.(I - rr)d = |d|^2, so
    coef_m = -K e^{-i w r/c} |d|^2 [c^2/r^3 w0 + c/r^2 (w1 + i w w0)
    + 1/r (w2 + 2 i w w1 - w^2 w0)]."""
    p = _params(oracle_lib)
    r = 0.137
    d = np.array([0, 0, 5.29e-4])
    w = _w(oracle_lib, 0.37)
    got = oracle_lib.rotating_coefs(p, [r, 0, 0], w, d, d)
    K = MU0 / (4 * math.pi * HBAR)
    ph = np.exp(-1j * OMEGA * r / C0)
    want = -K * ph * d[2] ** 2 * (C0 ** 2 / r ** 3 * w[0] + C0 / r ** 2 * (w[1] + 1j * OMEGA * w[0])
                                 + (w[2] + 2j * OMEGA * w[1] - OMEGA ** 2 * w[0]) / r)
    assert rel_err(got, want) < 1e-13


def test_rotating_coefs_axial_case(oracle_lib):
    """d || dr: d.(I - 3 rr)d = -2|d|^2, d.(I - rr)d = 0 (no far field)."""
    p = _params(oracle_lib)
    r = 0.21
    d = np.array([0.0, 0.0, 1.0])
    w = _w(oracle_lib, 0.61)
    got = oracle_lib.rotating_coefs(p, [0, 0, r], w, d, d)
    K = MU0 / (4 * math.pi * HBAR)
    ph = np.exp(-1j * OMEGA * r / C0)
    want = -K * ph * (-2) * (C0 ** 2 / r ** 3 * w[0] + C0 / r ** 2 * (w[1] + 1j * OMEGA * w[0]))
    assert rel_err(got, want) < 1e-13


def test_green_symmetry(oracle_lib):
    """G(dr) = G(-dr) = G^T: swapping source and observer changes nothing."""
    p = _params(oracle_lib)
    rng = np.random.default_rng(3)
    for _ in range(20):
        dr = rng.normal(size=3) * 0.3
        da, db = rng.normal(size=3), rng.normal(size=3)
        w = _w(oracle_lib, rng.uniform(0.01, 0.99))
        a = oracle_lib.rotating_coefs(p, dr, w, da, db)
        b = oracle_lib.rotating_coefs(p, -dr, w, db, da)
        assert rel_err(a, b) < 1e-13
        la = oracle_lib.lab_coefs(p, dr, w, da, db)
        lb = oracle_lib.lab_coefs(p, -dr, w, db, da)
        assert rel_err(la, lb) < 1e-13


def test_rotating_low_frequency_limit_is_minus_lab(oracle_lib):
    """omega -> 0: the rotating coefficients become -(lab coefficients)."""
    p0 = _params(oracle_lib, omega=0.0)
    dr, da, db = [0.1, -0.2, 0.05], [0.3, 0.1, 1.0], [-0.2, 0.5, 0.7]
    w = _w(oracle_lib, 0.3)
    rot = oracle_lib.rotating_coefs(p0, dr, w, da, db)
    lab = oracle_lib.lab_coefs(p0, dr, w, da, db)
    assert rel_err(rot, -lab) < 1e-14


def test_lab_far_field_dominates(oracle_lib):
    """At r >> c dt the 1/r term (D2) dominates the lab coefficients' second
    derivative part; check the exact split into the three dyads."""
    p = _params(oracle_lib)
    dr = np.array([2.0, 0, 0])
    d = np.array([0, 0, 1.0])
    w = _w(oracle_lib, 0.5)
    K = MU0 / (4 * math.pi * HBAR)
    want = K * (C0 ** 2 / 8 * w[0] + C0 / 4 * w[1] + w[2] / 2)
    assert rel_err(oracle_lib.lab_coefs(p, dr, w, d, d), want) < 1e-13


# ---------------------------------------------- history_interaction.cpp:47-70

def _brute_force_field(oracle, p, pos, dip, rho01_hist, step, window=22):
    """O(N^2) numpy restatement, pair by pair, of the retarded field."""
    n = pos.shape[0]
    out = np.zeros(n, dtype=complex)
    for a in range(n):
        for b in range(n):
            if a == b:
                continue
            d = pos[a] - pos[b]
            tau = math.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) / (p.c0 * p.dt)
            nn, x = math.floor(tau), tau - math.floor(tau)
            w = oracle.lagrange_weights(x, p.order, p.dt)
            cf = oracle.rotating_coefs(p, d, w, dip[a], dip[b])
            for i in range(p.order + 1):
                t = step - nn - i
                if t >= -window:
                    out[a] += cf[i] * rho01_hist(b, t)
    return out


def test_evaluate_matches_brute_force(oracle_lib):
    p = _params(oracle_lib)
    sysd = dense_system(12, 0.15, seed=5)
    rng = np.random.default_rng(1)
    steps = 30
    r = rng.normal(size=(12, steps)) + 1j * rng.normal(size=(12, steps))
    y = np.zeros((12, 22 + steps, 2), dtype=complex)
    y[:, 22:, 1] = r
    delays, coef = oracle_lib.build_table(p, sysd["pos"], sysd["dip"])
    step = steps - 1
    got = oracle_lib.evaluate_history(p, delays, coef, y, step)
    want = _brute_force_field(oracle_lib, p, sysd["pos"], sysd["dip"],
                              lambda b, t: r[b, t] if t >= 0 else 0.0, step)
    assert rel_err(got, want) < 1e-13


def test_evaluate_two_dots_closed_form(oracle_lib):
    """N = 2, history rho01 = const: the field at dot 0 is sum_m coef_m rho."""
    p = _params(oracle_lib)
    pos = np.array([[0, 0, 0], [0.05, 0.02, 0.0]])
    dip = np.array([[0, 0, 1.0], [0, 0, 1.0]])
    delays, coef = oracle_lib.build_table(p, pos, dip)
    y = np.zeros((2, 22 + 40, 2), dtype=complex)
    y[1, :, 1] = 0.3 - 0.1j
    y[0, :, 1] = 0.0
    out = oracle_lib.evaluate_history(p, delays, coef, y, 39)
    assert abs(out[0] - coef[0].sum() * (0.3 - 0.1j)) < 1e-12 * abs(out[0])
    assert out[1] == 0


def test_evaluate_linear_and_zero(oracle_lib):
    p = _params(oracle_lib)
    sysd = dense_system(20, 0.3, seed=8)
    delays, coef = oracle_lib.build_table(p, sysd["pos"], sysd["dip"])
    rng = np.random.default_rng(2)
    y1 = np.zeros((20, 22 + 25, 2), dtype=complex)
    y2 = np.zeros_like(y1)
    y1[:, 22:, 1] = rng.normal(size=(20, 25))
    y2[:, 22:, 1] = 1j * rng.normal(size=(20, 25))
    e1 = oracle_lib.evaluate_history(p, delays, coef, y1, 24)
    e2 = oracle_lib.evaluate_history(p, delays, coef, y2, 24)
    e12 = oracle_lib.evaluate_history(p, delays, coef, 2 * y1 + 3 * y2, 24)
    assert rel_err(e12, 2 * e1 + 3 * e2) < 1e-13
    assert np.all(oracle_lib.evaluate_history(p, delays, coef, np.zeros_like(y1), 24) == 0)


@pytest.mark.parametrize("kind", [0, 1])
def test_evaluate_observers_matches_table_evaluate(oracle_lib, kind):
    """The table-free evaluation at a subset of observers equals the table
    evaluation there (same terms, another order of additions), the brute
    force too, and includes the zero-delay pairs with the current sample."""
    p = _params(oracle_lib)
    n, steps = 90, 40
    sysd = dense_system(n, 0.5, seed=21)
    rng = np.random.default_rng(4)
    r = rng.normal(size=(n, steps)) + 1j * rng.normal(size=(n, steps))
    y = np.zeros((n, 22 + steps, 2), dtype=complex)
    y[:, 22:, 1] = r
    delays, coef = oracle_lib.build_table(p, sysd["pos"], sysd["dip"], kind=kind)
    assert (delays == 0).sum() >= 1
    step = steps - 1
    want = oracle_lib.evaluate_history(p, delays, coef, y, step, kind=kind)
    obs = np.array([0, 7, 33, 34, 89, 5], dtype=np.int32)
    got = oracle_lib.evaluate_observers(p, sysd["pos"], sysd["dip"], r.T, step, obs, kind=kind)
    assert rel_err(got, want[obs]) < 1e-13
    all_obs = oracle_lib.evaluate_observers(p, sysd["pos"], sysd["dip"], r.T, step, np.arange(n), kind=kind)
    assert rel_err(all_obs, want) < 1e-13
    if kind == 0:
        brute = _brute_force_field(oracle_lib, p, sysd["pos"], sysd["dip"],
                                   lambda b, t: r[b, t] if t >= 0 else 0.0, step)
        assert rel_err(all_obs, brute) < 1e-13


# ------------------------------------------------ pulse.cpp, quantum_dot.cpp

def test_pulse_closed_form(oracle_lib):
    pl = {"amplitude": 123.0, "delay": 0.4, "width": 50.0, "freq": 2000.0,
          "wavevector": [1.5, -0.5, 0.25], "polarization": [0.0, 0.6, 0.8]}
    pos = np.array([[0.1, 0.2, 0.3], [-1.0, 0.5, 2.0]])
    dip = np.array([[0, 0, 1e-3], [1e-3, 2e-3, -1e-3]])
    t = 0.37
    got = oracle_lib.pulse_evaluate(oracle_pulse(oracle_lib, pl), HBAR, pos, dip, t)
    for i in range(2):
        arg = np.dot(pl["wavevector"], pos[i]) - pl["freq"] * (t - pl["delay"])
        e = pl["amplitude"] / 2 * np.array(pl["polarization"]) * math.exp(-(arg / pl["width"]) ** 2 / 2)
        assert abs(got[i] - e @ dip[i] / HBAR) < 1e-14 * abs(got[i]) + 1e-300
    assert np.all(got.imag == 0)


def test_liouville_rhs_known_answer(oracle_lib):
    """test/quantum_dot_test.cpp:13-43: the RHS equals -i[H, rho] - D for
    H = [[0, Omega], [conj(Omega), 0]], resonant, Omega = 1."""
    omega0, t1, t2 = 2278.9013, 10.0, 20.0
    rabi = 1.0 + 0j
    rho = np.array([[1, 0.5 + 0.5j], [0.5 - 0.5j, 0]])
    H = np.array([[0, rabi], [np.conj(rabi), 0]])
    D = np.array([[(rho[0, 0] - 1) / t1, rho[0, 1] / t2], [rho[1, 0] / t2, rho[1, 1] / t1]])
    m = -1j * (H @ rho - rho @ H) - D
    got = oracle_lib.liouville_rhs([rho[0, 0], rho[0, 1]], rabi, omega0, omega0, t1, t2)
    for g, w in ((got[0], m[0, 0]), (got[1], m[0, 1])):
        # BOOST_CHECK_CLOSE(..., 1e-9) is a percentage
        assert abs(g.real - w.real) <= 1e-11 * abs(w.real) + 1e-300
        assert abs(g.imag - w.imag) <= 1e-11 * abs(w.imag) + 1e-300


@pytest.mark.parametrize("sol,finite", [
    ([0, 0], True), ([np.inf, 0], False), ([0, np.inf], False),
    ([1j * np.inf, 0], False), ([0, 1j * np.inf], False),
    ([np.nan, 0], False), ([0, np.nan], False), ([1j * np.nan, 0], False), ([0, 1j * np.nan], False),
])
def test_isfinite(oracle_lib, sol, finite):
    """test/history_test.cpp:13-53 (with the imaginary-infinity case the
    reference intended: its imag_inf is (inf, 0), history_test.cpp:10-11)."""
    s = np.array([complex(v) if not isinstance(v, complex) else v for v in sol])
    if not np.isfinite(s).all():
        s = np.array([complex(np.real(v), np.imag(v)) for v in sol])
    assert oracle_lib.isfinite(s) == finite


# ------------------------------------------------------ integrator.cpp:104-207

def _single_dot(detuning=0.0, t1=1e9, t2=1e9):
    return {"pos": np.zeros((1, 3)), "dip": np.array([[0, 0, 5.2917721e-4]]),
            "freq": np.array([2278.9 - detuning]), "t1": np.array([t1]), "t2": np.array([t2])}


@pytest.mark.parametrize("area", [math.pi / 2, math.pi, math.pi / 4])
def test_single_dot_pulse_area(oracle_lib, ref_weights, area):
    """N = 1, resonant, no damping: rho00 = cos^2(area / 2) after the pulse."""
    from quest_amd import systems
    consts = systems.consts()
    pl = short_pulse(area=area, delay=0.02, sigma_t=0.004)
    ps, cs, fc = ref_weights
    r = oracle_lib.solve(oracle_params(oracle_lib, consts), _single_dot(),
                         oracle_pulse(oracle_lib, pl), ps, cs, fc, 450, mode=0)
    assert abs(r["y"][0, 0].real - math.cos(area / 2) ** 2) < 1e-6
    assert abs(abs(r["y"][0, 1]) - abs(math.sin(area) / 2)) < 1e-6


def test_single_dot_free_decay(oracle_lib, ref_weights):
    """After the pulse, rho01 decays as exp(-t/T2) and rotates at the detuning."""
    from quest_amd import systems
    consts = systems.consts()
    pl = short_pulse(area=math.pi / 2, delay=0.01, sigma_t=0.002)
    ps, cs, fc = ref_weights
    det = 30.0
    s = _single_dot(detuning=det, t1=1e9, t2=0.05)
    r = oracle_lib.solve(oracle_params(oracle_lib, consts), s, oracle_pulse(oracle_lib, pl),
                         ps, cs, fc, 1000, mode=0, record_every=100)
    rec = r["rec"][:, 0, 1]
    t = (np.arange(1, 11) * 100 - 1) * 1e-4
    # steps 299.. are well past the pulse (t > 0.025 ps)
    ratio = rec[5:] / rec[4:-1]
    want = np.exp(-0.01 / 0.05) * np.exp(-1j * det * 0.01)
    assert np.max(np.abs(ratio - want)) < 1e-7
    assert t[-1] > 0.09


def test_solve_modes_agree(oracle_lib, ref_weights):
    """The split gather (mode 1) reproduces the reference method (mode 0)."""
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(40, 0.2, seed=4)
    pl = oracle_pulse(oracle_lib, short_pulse())
    ps, cs, fc = ref_weights
    p = oracle_params(oracle_lib, consts)
    r0 = oracle_lib.solve(p, sysd, pl, ps, cs, fc, 300, mode=0)
    r1 = oracle_lib.solve(p, sysd, pl, ps, cs, fc, 300, mode=1)
    assert np.abs(r0["y"][:, 1]).max() > 0.1
    assert rel_err(r1["y"], r0["y"]) < 1e-13
    assert rel_err(r1["yp"], r0["yp"]) < 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_pulse_list_of_one_equals_single_pulse(oracle_lib, ref_weights, mode):
    """qo_solve_pulses with a one-pulse list is qo_solve, bit for bit."""
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(24, 0.15, seed=8)
    pl = oracle_pulse(oracle_lib, short_pulse())
    ps, cs, fc = ref_weights
    p = oracle_params(oracle_lib, consts)
    r0 = oracle_lib.solve(p, sysd, pl, ps, cs, fc, 250, mode=mode)
    r1 = oracle_lib.solve(p, sysd, pl, ps, cs, fc, 250, mode=mode, use_list=True)
    assert np.array_equal(r0["y"], r1["y"]) and np.array_equal(r0["yp"], r1["yp"])


def test_two_equal_pulses_equal_one_of_double_amplitude(oracle_lib, ref_weights):
    """The pulse fields add: [p, p] drives exactly as p with twice the amplitude
    (doubling is exact in binary floating point)."""
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(24, 0.15, seed=9)
    base = short_pulse(area=math.pi / 4)
    double = dict(base, amplitude=2 * base["amplitude"])
    ps, cs, fc = ref_weights
    p = oracle_params(oracle_lib, consts)
    r2 = oracle_lib.solve(p, sysd, [oracle_pulse(oracle_lib, base)] * 2, ps, cs, fc, 250, mode=1)
    rd = oracle_lib.solve(p, sysd, oracle_pulse(oracle_lib, double), ps, cs, fc, 250, mode=1)
    assert np.abs(rd["y"][:, 1]).max() > 0.1
    assert np.array_equal(r2["y"], rd["y"])


@pytest.mark.parametrize("areas", [(math.pi / 4, math.pi / 4), (math.pi / 3, math.pi / 2)])
def test_two_pulses_add_areas(oracle_lib, ref_weights, areas):
    """N = 1, resonant, no damping: two separated pulses rotate about the same
    axis, so rho00 = cos^2((theta1 + theta2) / 2) after both."""
    from quest_amd import systems
    consts = systems.consts()
    p1 = short_pulse(area=areas[0], delay=0.015, sigma_t=0.003)
    p2 = short_pulse(area=areas[1], delay=0.035, sigma_t=0.003)
    ps, cs, fc = ref_weights
    r = oracle_lib.solve(oracle_params(oracle_lib, consts), _single_dot(),
                         [oracle_pulse(oracle_lib, p1), oracle_pulse(oracle_lib, p2)],
                         ps, cs, fc, 550, mode=0)
    total = areas[0] + areas[1]
    assert abs(r["y"][0, 0].real - math.cos(total / 2) ** 2) < 1e-6


def test_empty_pulse_list_keeps_ground_state(oracle_lib, ref_weights):
    from quest_amd import systems
    consts = systems.consts()
    sysd = dense_system(12, 0.1, seed=10)
    ps, cs, fc = ref_weights
    r = oracle_lib.solve(oracle_params(oracle_lib, consts), sysd, [], ps, cs, fc, 60, mode=1)
    # rho00 moves only by the rounding of the PC weights (they sum to 1 to ~1e-13)
    assert np.abs(r["y"][:, 0] - 1).max() < 1e-12 and np.all(r["y"][:, 1] == 0)


def test_solve_reports_nonfinite(oracle_lib, ref_weights):
    from quest_amd import systems
    consts = systems.consts()
    s = _single_dot()
    s["freq"][0] = np.nan
    ps, cs, fc = ref_weights
    r = oracle_lib.solve(oracle_params(oracle_lib, consts), s, oracle_pulse(oracle_lib, short_pulse()),
                         ps, cs, fc, 50, mode=0)
    assert r["status"] == 1 and r["bad_step"] == 0


def test_rho00_stays_real(oracle_lib, ref_weights):
    """The reference arithmetic keeps rho00 exactly real from the real ground
    state (its RHS's imaginary part is -Re z - Im rho00 / T1 with Re z the
    difference of two identically rounded products), which is what lets the
    device hold rho00 as a real number: every recorded step of a driven,
    coupled run has Im rho00 == 0 and Im drho00 == 0."""
    from quest_amd import systems
    ps, cs, fc = ref_weights
    sysd = dense_system(40, 0.24, seed=12)
    p = oracle_params(oracle_lib, systems.consts())
    r = oracle_lib.solve(p, sysd, oracle_pulse(oracle_lib, short_pulse()), ps, cs, fc, 400, mode=1,
                         record_every=1)
    assert r["status"] == 0
    assert np.abs(r["y"][:, 1]).max() > 0.1
    assert np.all(r["rec"][:, :, 0].imag == 0)
    assert np.all(r["y"][:, 0].imag == 0) and np.all(r["yp"][:, 0].imag == 0)


def test_timed_steps_at_steady_state_sum_every_tap(oracle_lib, ref_weights):
    """The CPU baseline's timed steps (qo_time_steps) start past the largest
    delay plus the order on a non-zero synthetic prehistory, so the reference
    method's skip of taps before the prehistory (history_interaction.cpp:57)
    never applies: 11 evaluations x N(N-1)/2 pairs x (M+1) taps x 2
    directions per step.  From the ground state at step 0 most taps are
    skipped; the split gather adds each far pair's taps once per step."""
    import math
    from quest_amd import systems
    ps, cs, fc = ref_weights
    # 64 dots in a 2 um cube: delays up to ~115 steps, as at N = 4096
    sysd = dense_system(64, 2.0, seed=3)
    c = systems.consts()
    p = oracle_params(oracle_lib, c)
    pl = oracle_pulse(oracle_lib, systems.pulse(math.pi / 2))
    n, M, K = 64, 4, 3
    start = oracle_lib.steady_start(sysd["pos"], c["c0"], c["dt"], M)
    pairs = n * (n - 1) // 2
    full = 11 * K * pairs * (M + 1) * 2
    r = oracle_lib.time_steps(p, sysd, pl, ps, cs, fc, K, start_step=start, mode=0)
    assert r["taps"] == full and math.isfinite(r["checksum"])
    cold = oracle_lib.time_steps(p, sysd, pl, ps, cs, fc, K, start_step=0, mode=0)
    assert 0 < cold["taps"] < full / 2
    delays, _ = oracle_lib.build_table(p, sysd["pos"], sysd["dip"])
    near = int((delays == 0).sum())
    split = oracle_lib.time_steps(p, sysd, pl, ps, cs, fc, K, start_step=start, mode=1)
    assert split["taps"] == K * (pairs - near) * (M + 1) * 2 + 11 * K * near * (M + 1) * 2
    # the prehistory is what the timed steps integrate from: both modes agree
    assert abs(split["checksum"] - r["checksum"]) < 1e-12 * abs(r["checksum"])
