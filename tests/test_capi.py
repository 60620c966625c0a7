"""The C ABI library loads on a machine without a GPU and exports every
function include/quest_amd.h declares; host-only entry points work."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "quest_amd.h")).read()
    return sorted(set(re.findall(r"\b(quest_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_bindings():
    from quest_amd import capi
    assert _declared() == sorted(capi.EXPORTS)


def test_library_exports_every_symbol():
    from quest_amd import capi
    lib = C.CDLL(capi.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_create_rejects_bad_arguments():
    from quest_amd import Context, QuestError, systems
    sysd = systems.system(8, seed=1)
    with pytest.raises(QuestError, match="interp_order"):
        Context(sysd, systems.consts(), order=9)
    with pytest.raises(QuestError, match="history_window"):
        Context(sysd, systems.consts(), window=0)
    bad = dict(sysd)
    bad["pos"] = sysd["pos"].copy()
    bad["pos"][0, 0] = np.nan
    with pytest.raises(QuestError, match="non-finite"):
        Context(bad, systems.consts())


def test_pc_weights_argument_errors():
    from quest_amd import QuestError, pc_weights
    with pytest.raises(QuestError):
        pc_weights(18, 1, 3.15)
    with pytest.raises(QuestError):
        pc_weights(18, 65, 3.15)


def test_systems_generator_density():
    from quest_amd import systems
    s = systems.system(4096, seed=2)
    L = systems.cube_side(4096)
    assert abs(L - 2.0) < 1e-12
    assert np.all(np.abs(s["pos"]) <= L / 2)
    assert abs(systems.pulse_amplitude(np.pi / 2) - 7794.6130486) < 1e-3
    assert abs(systems.pulse_amplitude(np.pi) - 15589.226097) < 1e-3


def test_struct_layouts_match_header():
    """The ctypes mirrors of quest_params and quest_pulse list the header's
    fields in order (a field added on one side only would shift every later
    one), and the documented info length is the one the front end reads."""
    from quest_amd import capi
    text = open(os.path.join(ROOT, "include", "quest_amd.h")).read()
    for ctype, cname in ((capi.Params, "quest_params"), (capi.Pulse, "quest_pulse")):
        end = text.index("} " + cname + ";")
        body = text[text.rindex("typedef struct {", 0, end) + len("typedef struct {"):end]
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = []
        for decl in body.split(";"):
            m = re.match(r"\s*(?:double|int)\s+(.*)", decl, re.S)
            if m:
                names += [re.sub(r"\[\d+\]", "", v).strip() for v in m.group(1).split(",")]
        assert names == [f[0] for f in ctype._fields_], cname
    n_info = int(re.search(r"info:\s*\[(\d+)\]\s*int64", text).group(1))
    assert n_info == 20
    import inspect
    assert f"c_int64 * {n_info}" in inspect.getsource(capi.Context.info)
