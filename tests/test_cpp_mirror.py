"""The C++ mirror's host-side checks (tests/cpp/mirror_test.cpp): the
reference's known-answer RHS test, finiteness, file formats, weights."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_mirror_host_checks(tmp_path):
    exe = os.path.join(ROOT, "quest_amd", "bin", "mirror_test")
    assert os.path.exists(exe), "build first: python -m quest_amd.build"
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "all checks passed" in r.stdout
