"""Readers of quest_hip's output files.

* ``read_text(path, n_dots)`` -- output.dat in the reference's format
  (src/history.cpp:17-34: one line per recorded step, "(re,im) (re,im) " per
  dot for rho00 and rho01), also the sampled text rows of
  ``--every``/``--subsample`` (n_dots = the dots recorded);
* ``read_binary(path)`` -- the binary records of ``--format binary``:
  little-endian ``char[8] "QUESTOUT"``; ``uint32`` version (1), n_dots,
  n_sel, every, n_records, 0; ``float64`` dt; ``int32 dots[n_sel]``; then
  n_records records of ``int32 step, int32 0, float64 [n_sel][4]`` (re, im
  rho00; re, im rho01).  ``step`` t is output.dat's line t, the state after
  step t.
"""
from __future__ import annotations

import re

import numpy as np

_NUM = re.compile(r"\(([^,]+),([^)]+)\)")
MAGIC = b"QUESTOUT"


def read_text(path: str, n_dots: int) -> np.ndarray:
    """[records][n_dots][2] complex (rho00, rho01)."""
    rows = []
    with open(path) as f:
        for line in f:
            vals = [complex(float(a), float(b)) for a, b in _NUM.findall(line)]
            if len(vals) != 2 * n_dots:
                raise ValueError(f"{path}: a line holds {len(vals) // 2} dots, expected {n_dots}")
            rows.append(np.array(vals).reshape(n_dots, 2))
    return np.array(rows).reshape(len(rows), n_dots, 2)


def read_binary(path: str) -> dict:
    """dict(n_dots, every, dt, dots [n_sel] int, steps [n_records] int,
    rho [n_records][n_sel][2] complex)."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not a quest_hip binary output (magic {raw[:8]!r})")
    version, n_dots, n_sel, every, n_rec, _ = np.frombuffer(raw, dtype="<u4", count=6, offset=8)
    if version != 1:
        raise ValueError(f"{path}: unsupported version {version}")
    dt = float(np.frombuffer(raw, dtype="<f8", count=1, offset=32)[0])
    off = 40
    dots = np.frombuffer(raw, dtype="<i4", count=int(n_sel), offset=off).astype(np.int64)
    off += 4 * int(n_sel)
    rec_bytes = 8 + 32 * int(n_sel)
    if len(raw) != off + int(n_rec) * rec_bytes:
        raise ValueError(f"{path}: {len(raw)} bytes, expected {off + int(n_rec) * rec_bytes} "
                         f"for {n_rec} records of {n_sel} dots")
    rec = np.frombuffer(raw, dtype=np.uint8, count=int(n_rec) * rec_bytes, offset=off).reshape(int(n_rec), rec_bytes)
    steps = rec[:, :4].copy().view("<i4").reshape(-1).astype(np.int64)
    vals = rec[:, 8:].copy().view("<f8").reshape(int(n_rec), int(n_sel), 4)
    rho = (vals[:, :, 0::2] + 1j * vals[:, :, 1::2])
    return {"n_dots": int(n_dots), "every": int(every), "dt": dt, "dots": dots, "steps": steps, "rho": rho}
