"""ctypes front end of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (built by oracle/Makefile from quest_oracle.c).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class QoParams(C.Structure):
    _fields_ = [("c0", C.c_double), ("hbar", C.c_double), ("mu0", C.c_double),
                ("omega", C.c_double), ("dt", C.c_double), ("order", C.c_int),
                ("window", C.c_int)]


class QoPulse(C.Structure):
    _fields_ = [("amplitude", C.c_double), ("delay", C.c_double),
                ("width", C.c_double), ("freq", C.c_double),
                ("k", C.c_double * 3), ("pol", C.c_double * 3)]


def build(force: bool = False) -> None:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []),
                       check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.qo_lagrange_weights.argtypes = [C.c_double, C.c_int, C.c_double, dp]
        L.qo_idx2coord.argtypes = [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.qo_split_delay.argtypes = [dp, dp, C.c_double, C.c_double, dp]
        L.qo_split_delay.restype = C.c_int
        L.qo_rotating_coefs.argtypes = [C.POINTER(QoParams), dp, dp, dp, dp, dp]
        L.qo_lab_coefs.argtypes = [C.POINTER(QoParams), dp, dp, dp, dp, dp]
        L.qo_build_table.argtypes = [C.POINTER(QoParams), C.c_int, C.c_int, dp, dp, ip, dp]
        L.qo_evaluate_history.argtypes = [C.POINTER(QoParams), C.c_int, C.c_int, C.c_int,
                                          ip, dp, dp, C.c_int, dp]
        L.qo_evaluate_observers.argtypes = [C.POINTER(QoParams), C.c_int, C.c_int, dp, dp, C.c_int,
                                            dp, C.c_int, C.c_int, ip, dp]
        L.qo_pulse_evaluate.argtypes = [C.POINTER(QoPulse), C.c_double, C.c_int, dp, dp,
                                        C.c_double, dp]
        L.qo_liouville_rhs.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, dp]
        L.qo_isfinite.argtypes = [dp]
        L.qo_isfinite.restype = C.c_int
        L.qo_solve.argtypes = [C.POINTER(QoParams), C.c_int, C.c_int, dp, dp, dp, dp, dp,
                               C.POINTER(QoPulse), dp, dp, C.c_double, C.c_int, C.c_int,
                               C.c_int, dp, dp, dp, C.POINTER(C.c_int), dp]
        L.qo_solve.restype = C.c_int
        L.qo_solve_pulses.argtypes = [C.POINTER(QoParams), C.c_int, C.c_int, dp, dp, dp, dp, dp,
                                      C.POINTER(QoPulse), C.c_int, dp, dp, C.c_double, C.c_int,
                                      C.c_int, C.c_int, dp, dp, dp, C.POINTER(C.c_int), dp]
        L.qo_solve_pulses.restype = C.c_int
        L.qo_time_steps.argtypes = [C.POINTER(QoParams), C.c_int, dp, dp, dp, dp, dp,
                                    C.POINTER(QoPulse), dp, dp, C.c_double, C.c_int, C.c_int,
                                    C.c_int, dp, dp, C.POINTER(C.c_int64)]
        L.qo_time_steps.restype = C.c_double
        L.qo_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def params(c0, hbar, mu0, omega, dt, order, window=22) -> QoParams:
    return QoParams(c0, hbar, mu0, omega, dt, order, window)


def pulse(amplitude, delay, width, freq, k, pol) -> QoPulse:
    return QoPulse(amplitude, delay, width, freq, (C.c_double * 3)(*k),
                   (C.c_double * 3)(*pol))


def set_threads(n: int) -> None:
    lib().qo_set_threads(int(n))


def lagrange_weights(x: float, order: int, dt: float = 1.0) -> np.ndarray:
    w = np.zeros(3 * (order + 1))
    lib().qo_lagrange_weights(x, order, dt, _dp(w))
    return w.reshape(3, order + 1)


def idx2coord(idx: int):
    r, c = C.c_int64(), C.c_int64()
    lib().qo_idx2coord(idx, C.byref(r), C.byref(c))
    return r.value, c.value


def split_delay(pos_src, pos_obs, c0, dt):
    a, b = _f64(pos_src), _f64(pos_obs)
    frac = C.c_double()
    n = lib().qo_split_delay(_dp(a), _dp(b), c0, dt, C.byref(frac))
    return n, frac.value


def rotating_coefs(p: QoParams, dr, w, d_obs, d_src) -> np.ndarray:
    out = np.zeros(2 * (p.order + 1))
    lib().qo_rotating_coefs(C.byref(p), _dp(_f64(dr)), _dp(_f64(w).ravel()),
                            _dp(_f64(d_obs)), _dp(_f64(d_src)), _dp(out))
    return out[0::2] + 1j * out[1::2]


def lab_coefs(p: QoParams, dr, w, d_obs, d_src) -> np.ndarray:
    out = np.zeros(2 * (p.order + 1))
    lib().qo_lab_coefs(C.byref(p), _dp(_f64(dr)), _dp(_f64(w).ravel()),
                       _dp(_f64(d_obs)), _dp(_f64(d_src)), _dp(out))
    return out[0::2] + 1j * out[1::2]


def build_table(p: QoParams, pos, dip, kind: int = 0):
    pos, dip = _f64(pos), _f64(dip)
    n = pos.shape[0]
    npairs = n * (n - 1) // 2
    delays = np.zeros(max(npairs, 1), dtype=np.int32)
    coef = np.zeros(2 * (p.order + 1) * max(npairs, 1))
    lib().qo_build_table(C.byref(p), kind, n, _dp(pos), _dp(dip),
                         delays.ctypes.data_as(C.POINTER(C.c_int32)), _dp(coef))
    coef = coef.reshape(-1, p.order + 1, 2)
    return delays[:npairs], (coef[..., 0] + 1j * coef[..., 1])[:npairs]


def evaluate_history(p: QoParams, delays, coef, y, step: int, kind: int = 0):
    """y: complex [n][window + T][2] (rho00, rho01 at t = slot - window)."""
    y = np.ascontiguousarray(y, dtype=np.complex128)
    n, n_slots = y.shape[0], y.shape[1]
    d = np.ascontiguousarray(delays, dtype=np.int32)
    cf = np.ascontiguousarray(coef, dtype=np.complex128)
    out = np.zeros(n, dtype=np.complex128)
    lib().qo_evaluate_history(C.byref(p), kind, n, n_slots,
                              d.ctypes.data_as(C.POINTER(C.c_int32)),
                              _dp(cf.view(np.float64)), _dp(y.view(np.float64)), step,
                              _dp(out.view(np.float64)))
    return out


def evaluate_observers(p: QoParams, pos, dip, rho01, step: int, observers, kind: int = 0):
    """evaluate_history at the given observers without a table.
    rho01: complex [n_t][n] samples at t = 0 .. n_t - 1 (earlier: 0)."""
    pos, dip = _f64(pos), _f64(dip)
    r = np.ascontiguousarray(rho01, dtype=np.complex128)
    n_t, n = r.shape
    assert n == pos.shape[0] and 0 <= step < n_t
    obs = np.ascontiguousarray(observers, dtype=np.int32)
    out = np.zeros(obs.shape[0], dtype=np.complex128)
    lib().qo_evaluate_observers(C.byref(p), kind, n, _dp(pos), _dp(dip), n_t,
                                _dp(r.view(np.float64)), step, obs.shape[0],
                                obs.ctypes.data_as(C.POINTER(C.c_int32)), _dp(out.view(np.float64)))
    return out


def pulse_evaluate(pl: QoPulse, hbar, pos, dip, t):
    pos, dip = _f64(pos), _f64(dip)
    out = np.zeros(pos.shape[0], dtype=np.complex128)
    lib().qo_pulse_evaluate(C.byref(pl), hbar, pos.shape[0], _dp(pos), _dp(dip), t,
                            _dp(out.view(np.float64)))
    return out


def liouville_rhs(rho, rabi, laser_freq, freq, t1, t2):
    r = np.ascontiguousarray(rho, dtype=np.complex128)
    rb = np.array([rabi], dtype=np.complex128)
    out = np.zeros(2, dtype=np.complex128)
    lib().qo_liouville_rhs(_dp(r.view(np.float64)), _dp(rb.view(np.float64)), laser_freq,
                           freq, t1, t2, _dp(out.view(np.float64)))
    return out


def isfinite(sol) -> bool:
    s = np.ascontiguousarray(sol, dtype=np.complex128)
    return bool(lib().qo_isfinite(_dp(s.view(np.float64))))


def solve(p: QoParams, system, pl, ps, cs, fc, n_steps: int, mode: int = 1,
          record_every: int = 0, kind: int = 0, use_list: bool = False):
    """Run the whole PC integration.  system: dict with pos, dip, freq, t1, t2.
    pl: one QoPulse, or a list of them (their fields add, in list order).
    use_list routes a single pulse through qo_solve_pulses as a list of one.

    Returns dict(rec=[n_rec][n][2] complex or None, y=[n][2], yp=[n][2],
    status, bad_step, seconds).
    """
    pos, dip = _f64(system["pos"]), _f64(system["dip"])
    freq, t1, t2 = _f64(system["freq"]), _f64(system["t1"]), _f64(system["t2"])
    n = pos.shape[0]
    ps, cs = _f64(ps).ravel(), _f64(cs).ravel()
    n_rec = n_steps // record_every if record_every > 0 else 0
    rec = np.zeros((max(n_rec, 1), n, 2), dtype=np.complex128)
    fy = np.zeros((n, 2), dtype=np.complex128)
    fyp = np.zeros((n, 2), dtype=np.complex128)
    bad = C.c_int(-1)
    secs = C.c_double(0)
    tail = (_dp(ps), _dp(cs), float(fc), n_steps, mode, record_every,
            _dp(rec.view(np.float64)) if n_rec else None,
            _dp(fy.view(np.float64)), _dp(fyp.view(np.float64)), C.byref(bad), C.byref(secs))
    if isinstance(pl, (list, tuple)) or use_list:
        pls = list(pl) if isinstance(pl, (list, tuple)) else [pl]
        arr = (QoPulse * max(len(pls), 1))(*pls)
        status = lib().qo_solve_pulses(C.byref(p), kind, n, _dp(pos), _dp(dip), _dp(freq), _dp(t1),
                                       _dp(t2), arr, len(pls), *tail)
    else:
        status = lib().qo_solve(C.byref(p), kind, n, _dp(pos), _dp(dip), _dp(freq), _dp(t1),
                                _dp(t2), C.byref(pl), *tail)
    if status < 0:
        raise MemoryError("oracle allocation failed")
    return {"rec": rec[:n_rec] if n_rec else None, "y": fy, "yp": fyp,
            "status": status, "bad_step": bad.value, "seconds": secs.value}


def time_steps(p: QoParams, system, pl: QoPulse, ps, cs, fc, n_steps: int, start_step: int = 0,
               mode: int = 0):
    """Times steps [start_step, start_step + n_steps) of the reference method
    (mode 0, serial) or the split gather (mode 1, OpenMP) after a synthetic
    non-zero prehistory (qo_time_steps).  Returns checksum, setup_seconds,
    step_seconds and taps (the pair-tap terms the timed evaluations added)."""
    pos, dip = _f64(system["pos"]), _f64(system["dip"])
    freq, t1, t2 = _f64(system["freq"]), _f64(system["t1"]), _f64(system["t2"])
    ps, cs = _f64(ps).ravel(), _f64(cs).ravel()
    setup, steps, taps = C.c_double(0), C.c_double(0), C.c_int64(0)
    chk = lib().qo_time_steps(C.byref(p), pos.shape[0], _dp(pos), _dp(dip), _dp(freq), _dp(t1),
                              _dp(t2), C.byref(pl), _dp(ps), _dp(cs), float(fc), int(mode),
                              int(start_step), int(n_steps), C.byref(setup), C.byref(steps),
                              C.byref(taps))
    return {"checksum": chk, "setup_seconds": setup.value, "step_seconds": steps.value,
            "taps": taps.value}


def steady_start(pos, c0: float, dt: float, order: int) -> int:
    """First step whose evaluations skip no tap: the bounding-box delay bound
    plus the order plus one."""
    pos = np.asarray(pos, dtype=np.float64)
    diag = float(np.linalg.norm(pos.max(axis=0) - pos.min(axis=0)))
    return int(diag / (c0 * dt)) + order + 1
