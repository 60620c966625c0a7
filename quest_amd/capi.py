"""ctypes bindings of libquest_amd.so (the C ABI in include/quest_amd.h).

The Python mirror of the reference's interaction and integrator interface:

* ``Context.hist_evaluate(step, rho01)`` -- HistoryInteraction::evaluate
  (src/interactions/history_interaction.cpp:47-70)
* ``Context.pulse_evaluate(step)``      -- PulseInteraction::evaluate
  (src/interactions/pulse_interaction.cpp:9-20)
* ``Context.run(n_steps)``              -- PredictorCorrector::Integrator::solve
  (src/integrator.cpp:120-127), whole run on the device
* ``pc_weights(18, 22, 3.15)``          -- PredictorCorrector::Weights
  (src/integrator.cpp:85-102)

Errors: a non-finite state raises ``UnboundedSolution`` (a ``ValueError``,
the Python spelling of the reference's std::domain_error) with the
reference's message "unbounded history values at or before step N"
(integrator.cpp:200-207); every other failure raises ``QuestError``
(RuntimeError).  There is no CPU fallback: without the built library every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QUEST_AMD_LIB") or os.path.join(HERE, "lib", "libquest_amd.so")

QUEST_OK, QUEST_ERR_ARG, QUEST_ERR_HIP, QUEST_ERR_DOMAIN, QUEST_ERR_STATE, QUEST_ERR_NOMEM = range(6)
PREC_FP64, PREC_MIXED = 0, 1
DYADIC_ROTATING, DYADIC_LAB = 0, 1

# Every symbol include/quest_amd.h declares.
EXPORTS = [
    "quest_ctx_create", "quest_ctx_destroy", "quest_last_error", "quest_hist_evaluate",
    "quest_set_pulse", "quest_pulse_evaluate", "quest_pc_weights", "quest_set_pc_weights",
    "quest_run", "quest_current_step", "quest_bad_step", "quest_get_state",
    "quest_pair_delays", "quest_set_timing", "quest_pair_kernel_time", "quest_ctx_info",
    "quest_sync", "quest_batch_create", "quest_set_pulses", "quest_autotune",
    "quest_set_pulse_list", "quest_set_pair_method", "quest_pair_delays_range",
    "quest_run_noalloc", "quest_short_kernel_time",
]
PAIR_ON_THE_FLY, PAIR_TABLE, PAIR_TIME_BLOCKED = 0, 1, 2  # QUEST_PAIR_*
CLAIM_AUTO, CLAIM_STATIC, CLAIM_DYNAMIC = 0, 1, 2  # QUEST_CLAIM_*
LOOP_AUTO, LOOP_PLAIN, LOOP_PIPELINED = 0, 1, 2     # QUEST_LOOP_*
ACC_BLOCK, ACC_COMPENSATED = 0, 1                   # QUEST_ACC_*
MAX_PULSES = 8  # QUEST_MAX_PULSES


class QuestError(RuntimeError):
    pass


class UnboundedSolution(ValueError):
    """std::domain_error of Integrator::throw_if_unbounded_solution."""

    def __init__(self, msg, step):
        super().__init__(msg)
        self.step = step


class Params(C.Structure):
    _fields_ = [("c0", C.c_double), ("hbar", C.c_double), ("mu0", C.c_double),
                ("laser_omega", C.c_double), ("dt", C.c_double), ("interp_order", C.c_int),
                ("history_window", C.c_int), ("precision", C.c_int), ("dyadic", C.c_int),
                ("device", C.c_int), ("source_chunk", C.c_int),
                ("item_claim", C.c_int), ("source_loop", C.c_int), ("accumulation", C.c_int),
                ("block_steps", C.c_int)]


class Pulse(C.Structure):
    _fields_ = [("amplitude", C.c_double), ("delay", C.c_double), ("width", C.c_double),
                ("freq", C.c_double), ("wavevector", C.c_double * 3),
                ("polarization", C.c_double * 3)]


_lib = None


def load():
    """The loaded library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QuestError(f"{LIB_PATH} is missing: build it with `python -m quest_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    vp = C.c_void_p
    L.quest_ctx_create.argtypes = [C.POINTER(Params), C.c_int, dp, dp, dp, dp, dp, C.POINTER(vp)]
    L.quest_ctx_destroy.argtypes = [vp]
    L.quest_ctx_destroy.restype = None
    L.quest_last_error.argtypes = [vp]
    L.quest_last_error.restype = C.c_char_p
    L.quest_hist_evaluate.argtypes = [vp, C.c_int, dp, dp]
    L.quest_set_pulse.argtypes = [vp, C.POINTER(Pulse)]
    L.quest_pulse_evaluate.argtypes = [vp, C.c_int, dp]
    L.quest_pc_weights.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp, dp]
    L.quest_set_pc_weights.argtypes = [vp, C.c_int, dp, dp, C.c_double]
    L.quest_run.argtypes = [vp, C.c_int]
    L.quest_run_noalloc.argtypes = [vp, C.c_int]
    L.quest_current_step.argtypes = [vp, ip]
    L.quest_bad_step.argtypes = [vp, ip]
    L.quest_get_state.argtypes = [vp, dp, dp]
    L.quest_pair_delays.argtypes = [vp, C.POINTER(C.c_int32)]
    L.quest_pair_delays_range.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(C.c_int32)]
    L.quest_set_timing.argtypes = [vp, C.c_int]
    L.quest_pair_kernel_time.argtypes = [vp, dp, C.POINTER(C.c_int64), C.c_int]
    L.quest_short_kernel_time.argtypes = [vp, dp, C.POINTER(C.c_int64), C.c_int]
    L.quest_ctx_info.argtypes = [vp, C.POINTER(C.c_int64)]
    L.quest_sync.argtypes = [vp]
    L.quest_batch_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(C.c_int), dp, dp, dp, dp,
                                     dp, C.POINTER(vp)]
    L.quest_set_pulses.argtypes = [vp, C.POINTER(Pulse)]
    L.quest_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.quest_set_pulse_list.argtypes = [vp, C.c_int, C.c_int, C.POINTER(Pulse)]
    L.quest_set_pair_method.argtypes = [vp, C.c_int]
    for name in EXPORTS:
        if name not in ("quest_ctx_destroy", "quest_last_error"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def pc_weights(n_lambda: int = 18, n_time: int = 22, radius: float = 3.15):
    """PredictorCorrector::Weights (integrator.cpp:85-102): (ps [2][n_time], cs, future_coef)."""
    L = load()
    ps = np.zeros((2, n_time))
    cs = np.zeros((2, n_time))
    fc = C.c_double()
    rc = L.quest_pc_weights(n_lambda, n_time, radius, _dp(ps), _dp(cs), C.byref(fc))
    if rc != QUEST_OK:
        raise QuestError(L.quest_last_error(None).decode())
    return ps, cs, fc.value


def _pulse_struct(pl):
    return Pulse(pl["amplitude"], pl["delay"], pl["width"], pl["freq"],
                 (C.c_double * 3)(*pl["wavevector"]), (C.c_double * 3)(*pl["polarization"]))


class Context:
    """One dot system -- or, given a list of systems, a batch of independent
    systems (quest_batch_create) -- on one GPU (a quest_ctx).  Per-dot arrays
    of a batch are the systems' arrays concatenated in order."""

    def __init__(self, system, consts, precision=PREC_FP64, dyadic=DYADIC_ROTATING,
                 order=4, window=22, device=0, source_chunk=0, pulse=None, item_claim=0, source_loop=0,
                 accumulation=0, block_steps=0):
        L = load()
        self._L = L
        self.ptr = None
        batch = isinstance(system, (list, tuple))
        systems = system if batch else [system]
        sizes = [np.asarray(s["pos"]).shape[0] for s in systems]

        def cat(key, default, width):
            return np.ascontiguousarray(np.concatenate(
                [np.asarray(s[key] if key in s else default(n), dtype=np.float64).reshape(n * width)
                 for s, n in zip(systems, sizes)]))

        pos = cat("pos", lambda n: np.zeros((n, 3)), 3)
        dip = cat("dip", lambda n: np.zeros((n, 3)), 3)
        freq = cat("freq", np.zeros, 1)
        t1 = cat("t1", np.ones, 1)
        t2 = cat("t2", np.ones, 1)
        self.n = int(sum(sizes))
        self.sizes = sizes
        p = Params(consts["c0"], consts["hbar"], consts["mu0"], consts["laser_omega"],
                   consts["dt"], order, window, precision, dyadic, device, source_chunk, item_claim,
                   source_loop, accumulation, block_steps)
        out = C.c_void_p()
        if batch:
            arr = (C.c_int * len(sizes))(*sizes)
            rc = L.quest_batch_create(C.byref(p), len(sizes), arr, _dp(pos), _dp(dip), _dp(freq),
                                      _dp(t1), _dp(t2), C.byref(out))
        else:
            rc = L.quest_ctx_create(C.byref(p), self.n, _dp(pos), _dp(dip), _dp(freq), _dp(t1),
                                    _dp(t2), C.byref(out))
        if rc != QUEST_OK:
            raise QuestError(f"quest_ctx_create: {L.quest_last_error(None).decode()}")
        self.ptr = out
        if pulse is not None:
            if isinstance(pulse, (list, tuple)):
                self.set_pulses(pulse)
            else:
                self.set_pulse(pulse)

    def _check(self, rc):
        if rc == QUEST_OK:
            return
        msg = self._L.quest_last_error(self.ptr).decode()
        if rc == QUEST_ERR_DOMAIN:
            bad = C.c_int()
            self._L.quest_bad_step(self.ptr, C.byref(bad))
            raise UnboundedSolution(msg, bad.value)
        raise QuestError(msg)

    def close(self):
        if self.ptr:
            self._L.quest_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_pulse(self, pl):
        if pl is None:
            self._check(self._L.quest_set_pulse(self.ptr, None))
            return
        s = _pulse_struct(pl)
        self._check(self._L.quest_set_pulse(self.ptr, C.byref(s)))

    def set_pulses(self, pulses):
        """One pulse per system of a batch."""
        if len(pulses) != len(self.sizes):
            raise QuestError(f"set_pulses: {len(pulses)} pulses for {len(self.sizes)} systems")
        arr = (Pulse * len(pulses))(*[_pulse_struct(pl) for pl in pulses])
        self._check(self._L.quest_set_pulses(self.ptr, arr))

    def set_pulse_list(self, pulses, system: int = 0):
        """Several pulses on one system (their fields add, in list order); an
        empty list clears the system's pulses."""
        arr = (Pulse * max(len(pulses), 1))(*[_pulse_struct(pl) for pl in pulses])
        self._check(self._L.quest_set_pulse_list(self.ptr, int(system), len(pulses), arr))

    def set_pair_method(self, method: int):
        """PAIR_ON_THE_FLY (default), PAIR_TABLE (the reference's per-pair
        coefficient table, FP64 only) or PAIR_TIME_BLOCKED (one pass over the
        pairs of delay >= K per K = 8 or 16 steps, `block_steps`); before the
        first step."""
        self._check(self._L.quest_set_pair_method(self.ptr, int(method)))

    def autotune(self, launches: int = 3) -> int:
        """quest_autotune: the fastest source-chunk length (kept)."""
        out = C.c_int()
        self._check(self._L.quest_autotune(self.ptr, int(launches), C.byref(out)))
        return out.value

    def set_pc_weights(self, ps, cs, fc):
        ps = np.ascontiguousarray(ps, dtype=np.float64)
        cs = np.ascontiguousarray(cs, dtype=np.float64)
        self._check(self._L.quest_set_pc_weights(self.ptr, ps.shape[1], _dp(ps), _dp(cs), float(fc)))

    def hist_evaluate(self, step: int, rho01) -> np.ndarray:
        r = np.ascontiguousarray(rho01, dtype=np.complex128)
        if r.shape != (self.n,):
            raise QuestError(f"hist_evaluate: rho01 has shape {r.shape}, expected ({self.n},)")
        out = np.zeros(self.n, dtype=np.complex128)
        self._check(self._L.quest_hist_evaluate(self.ptr, step, _dp(r.view(np.float64)),
                                                _dp(out.view(np.float64))))
        return out

    def pulse_evaluate(self, step: int) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.complex128)
        self._check(self._L.quest_pulse_evaluate(self.ptr, step, _dp(out.view(np.float64))))
        return out

    def run(self, n_steps: int) -> None:
        self._check(self._L.quest_run(self.ptr, int(n_steps)))

    def run_noalloc(self, n_steps: int) -> None:
        """quest_run_noalloc: the same steps with no device allocation (the
        pair sums in the library's static per-GPU scratch array)."""
        self._check(self._L.quest_run_noalloc(self.ptr, int(n_steps)))

    def state(self):
        """(rho [n][2], drho [n][2]) complex at the last finished step."""
        rho = np.zeros((self.n, 2), dtype=np.complex128)
        drho = np.zeros((self.n, 2), dtype=np.complex128)
        self._check(self._L.quest_get_state(self.ptr, _dp(rho.view(np.float64)),
                                            _dp(drho.view(np.float64))))
        return rho, drho

    @property
    def step(self) -> int:
        s = C.c_int()
        self._check(self._L.quest_current_step(self.ptr, C.byref(s)))
        return s.value

    def pair_delays(self, first: int | None = None, count: int | None = None) -> np.ndarray:
        """Integer delays of the unordered pairs in the reference's order, from
        the delay split of this context's precision; all pairs, or
        [first, first + count)."""
        total = self.n * (self.n - 1) // 2
        if first is None and count is None:
            out = np.zeros(max(total, 1), dtype=np.int32)
            self._check(self._L.quest_pair_delays(self.ptr, out.ctypes.data_as(C.POINTER(C.c_int32))))
            return out[:total]
        first = 0 if first is None else int(first)
        count = total - first if count is None else int(count)
        out = np.zeros(max(count, 1), dtype=np.int32)
        self._check(self._L.quest_pair_delays_range(self.ptr, first, count,
                                                    out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out[:max(count, 0)]

    def set_timing(self, enable: bool) -> None:
        self._check(self._L.quest_set_timing(self.ptr, int(bool(enable))))

    def pair_kernel_time(self, reset: bool = False):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(self._L.quest_pair_kernel_time(self.ptr, C.byref(ms), C.byref(cnt), int(reset)))
        return ms.value, cnt.value

    def short_kernel_time(self, reset: bool = False):
        """Time-blocked method: accumulated list-kernel time (ms: delays 1-3
        every step, 4-7 every 4 steps, 8-15 every 8 steps with 16-step passes)
        and the number of steps it served."""
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(self._L.quest_short_kernel_time(self.ptr, C.byref(ms), C.byref(cnt), int(reset)))
        return ms.value, cnt.value

    def info(self) -> dict:
        a = (C.c_int64 * 20)()
        self._check(self._L.quest_ctx_info(self.ptr, a))
        keys = ["n", "ring_slots", "max_delay_bound", "near_pairs", "work_items", "grid",
                "state_slots", "precision", "systems", "source_chunk", "splits", "graph",
                "step_block", "max_group", "dynamic_items", "pipelined", "pair_method",
                "block_steps", "short_pairs", "sample_slots"]
        return dict(zip(keys, list(a)))

    def sync(self) -> None:
        self._check(self._L.quest_sync(self.ptr))
