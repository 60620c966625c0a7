"""The C++ mirror's host-side checks (tests/cpp/mirror_test.cpp): the
reference's known-answer RHS test, finiteness, file formats, weights."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mirror_host_checks(tmp_path):
    exe = os.path.join(ROOT, "quest_amd", "bin", "mirror_test")
    assert os.path.exists(exe), "build first: python -m quest_amd.build"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "all checks passed" in r.stdout
    # the sampled-output files it wrote, read back by quest_amd.outputs
    from quest_amd import outputs
    b = outputs.read_binary(str(tmp_path / "sampled_test.bin"))
    assert b["n_dots"] == 6 and b["every"] == 3 and b["dt"] == 1e-4
    assert list(b["dots"]) == [1, 4] and list(b["steps"]) == [2, 5]
    for r_i, t in enumerate((2, 5)):
        for k, i in enumerate((1, 4)):
            assert b["rho"][r_i, k, 0] == complex(t + i / 10.0, -t)
            assert b["rho"][r_i, k, 1] == complex(i, t / 2.0)
    txt = outputs.read_text(str(tmp_path / "sampled_test.txt"), 2)
    assert txt.shape == (2, 2, 2)
    assert np.allclose(txt, b["rho"], rtol=1e-11, atol=0)
