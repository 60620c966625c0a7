"""The C++ driver quest_hip (main.cpp's replacement over the GPU library) on
the reference's own example inputs and on a denser synthetic system, in both
the device loop and the host loop (reference integrator + GPU interactions),
against the FP64 oracle.  Tolerance: 1e-10 norm-wise relative in rho."""
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import dense_system, oracle_params, oracle_pulse, rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "quest_amd", "bin", "quest_hip")
EXAMPLES = os.path.join(ROOT, "tests", "golden", "examples")
_num = re.compile(r"\(([^,]+),([^)]+)\)")


def read_output(path, n):
    rows = []
    for line in open(path):
        vals = [complex(float(a), float(b)) for a, b in _num.findall(line)]
        assert len(vals) == 2 * n
        rows.append(np.array(vals).reshape(n, 2))
    return np.array(rows)


def read_dots(path, n):
    a = np.loadtxt(path)[:n]
    return {"pos": a[:, 0:3], "freq": a[:, 3], "t1": a[:, 4], "t2": a[:, 5], "dip": a[:, 6:9]}


def run_cli(workdir, extra):
    r = subprocess.run([EXE] + extra, cwd=workdir, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def oracle_traj(oracle, sysd, consts, pl, steps, order, mode):
    """The oracle with the weights quest_hip uses (PredictorCorrector::Weights
    of the library, integrator.cpp:85-102)."""
    from quest_amd import pc_weights
    ps, cs, fc = pc_weights(18, 22, 3.15)
    r = oracle.solve(oracle_params(oracle, consts, order=order), sysd, oracle_pulse(oracle, pl),
                     ps, cs, fc, steps, mode=mode, record_every=1)
    assert r["status"] == 0
    return r["rec"]


@pytest.mark.parametrize("loop", ["device", "host"])
def test_reference_example_inputs(oracle_lib, tmp_path, loop):
    """examples/{input,dots,pulse}.cfg as shipped: N = 2, dt = 5e-3 ps, 4000
    steps, order 3, a pi pulse (every pair is a zero-delay pair here)."""
    for f in ("input.cfg", "dots.cfg", "pulse.cfg"):
        shutil.copy(os.path.join(EXAMPLES, f), tmp_path / f)
    run_cli(tmp_path, ["--host-loop"] if loop == "host" else [])
    got = read_output(tmp_path / "output.dat", 2)
    consts = {"c0": 299.792458, "hbar": 0.65821193, "mu0": 2.0133545e-04,
              "laser_omega": 2278.9013, "dt": 0.5e-2}
    a = np.loadtxt(os.path.join(EXAMPLES, "pulse.cfg"))
    pl = {"amplitude": a[0], "delay": a[1], "width": a[2], "freq": a[3],
          "wavevector": list(a[4:7]), "polarization": list(a[7:10])}
    want = oracle_traj(oracle_lib, read_dots(os.path.join(EXAMPLES, "dots.cfg"), 2),
                       consts, pl, 4000, 3, mode=0)
    assert got.shape == want.shape
    err = rel_err(got, want)
    print(f"{loop} loop: rel err {err:.3e}, final rho00 {want[-1, :, 0].real}")
    assert abs(want[-1, 0, 0].real - 0.5) < 0.5   # the pi pulse moved the state
    assert err < 1e-10


@pytest.mark.parametrize("loop,precision", [("device", "fp64"), ("host", "fp64"), ("device", "mixed"),
                                            ("device", "table"), ("host", "table"),
                                            ("device", "mixed+compensated"), ("device", "mixed+tune"),
                                            ("device", "fp64+tune")])
def test_dense_system(oracle_lib, tmp_path, loop, precision):
    """A 40-dot system with close pairs under a short pi/2 pulse, 300 steps;
    "+compensated" runs --accumulation compensated, "+tune" --tune (the work
    split timed before the run: the results move at rounding level only)."""
    n, steps = 40, 300
    sysd = dense_system(n, 0.2, seed=77)
    with open(tmp_path / "dots.cfg", "w") as f:
        for i in range(n):
            f.write(" ".join(repr(float(v)) for v in (*sysd["pos"][i], sysd["freq"][i], sysd["t1"][i],
                                                       sysd["t2"][i], *sysd["dip"][i])) + "\n")
    sigma = 0.005
    omega = 2278.9
    amp = (math.pi / 2) * 0.65821193 / (5.2917721e-4 * sigma * math.sqrt(2 * math.pi))
    pl = {"amplitude": amp, "delay": 0.015, "width": omega * sigma, "freq": omega,
          "wavevector": [omega / 299.792458, 0.0, 0.0], "polarization": [0.0, 0.0, 1.0]}
    with open(tmp_path / "pulse.cfg", "w") as f:
        f.write(" ".join(repr(float(v)) for v in (pl["amplitude"], pl["delay"], pl["width"], pl["freq"],
                                                   *pl["wavevector"], *pl["polarization"])) + "\n")
    with open(tmp_path / "input.cfg", "w") as f:
        f.write(f"[parameters]\n num_particles = {n}\n total_time = {steps * 1e-4}\n timestep = 1e-4\n"
                f" interpolation_order = 4\n[constants]\n c0 = 299.792458\n hbar = 0.65821193\n"
                f" mu0 = 2.0133545e-04\n laser_frequency = {omega}\n")
    base, *flags = precision.split("+")
    args = (["--pair-method", "table"] if base == "table" else ["--precision", base]) + \
        (["--accumulation", "compensated"] if "compensated" in flags else []) + \
        (["--tune"] if "tune" in flags else []) + (["--host-loop"] if loop == "host" else [])
    run_cli(tmp_path, args)
    got = read_output(tmp_path / "output.dat", n)
    consts = {"c0": 299.792458, "hbar": 0.65821193, "mu0": 2.0133545e-04, "laser_omega": omega, "dt": 1e-4}
    want = oracle_traj(oracle_lib, sysd, consts, pl, steps, 4, mode=1)
    assert got.shape[0] == steps
    err = rel_err(got, want)
    print(f"{loop} loop {precision}: rel err {err:.3e}, max|rho01| {np.abs(want[:, :, 1]).max():.3f}")
    assert err < (1e-6 if base == "mixed" else 1e-10)


@pytest.mark.parametrize("loop", ["device", "host"])
def test_all_pulses(oracle_lib, tmp_path, loop):
    """--all-pulses: every row of pulse.cfg drives the dots (the pulse
    interactions' fields add); without it only the first row does, as in the
    reference.  Both against the oracle's list of pulse interactions."""
    from quest_amd import pc_weights
    n, steps, omega = 30, 400, 2278.9
    sysd = dense_system(n, 0.18, seed=78)
    with open(tmp_path / "dots.cfg", "w") as f:
        for i in range(n):
            f.write(" ".join(repr(float(v)) for v in (*sysd["pos"][i], sysd["freq"][i], sysd["t1"][i],
                                                       sysd["t2"][i], *sysd["dip"][i])) + "\n")
    pls = []
    for area, delay, sigma, pol in ((math.pi / 3, 0.012, 0.003, (0.0, 0.0, 1.0)),
                                    (math.pi / 4, 0.028, 0.004, (0.0, 0.6, 0.8))):
        amp = area * 0.65821193 / (5.2917721e-4 * sigma * math.sqrt(2 * math.pi))
        pls.append({"amplitude": amp, "delay": delay, "width": omega * sigma, "freq": omega,
                    "wavevector": [omega / 299.792458, 0.0, 0.0], "polarization": list(pol)})
    with open(tmp_path / "pulse.cfg", "w") as f:
        for pl in pls:
            f.write(" ".join(repr(float(v)) for v in (pl["amplitude"], pl["delay"], pl["width"], pl["freq"],
                                                       *pl["wavevector"], *pl["polarization"])) + "\n")
    with open(tmp_path / "input.cfg", "w") as f:
        f.write(f"[parameters]\n num_particles = {n}\n total_time = {steps * 1e-4}\n timestep = 1e-4\n"
                f" interpolation_order = 4\n[constants]\n c0 = 299.792458\n hbar = 0.65821193\n"
                f" mu0 = 2.0133545e-04\n laser_frequency = {omega}\n")
    consts = {"c0": 299.792458, "hbar": 0.65821193, "mu0": 2.0133545e-04, "laser_omega": omega, "dt": 1e-4}
    ps, cs, fc = pc_weights(18, 22, 3.15)
    p = oracle_params(oracle_lib, consts)
    loop_args = ["--host-loop"] if loop == "host" else []
    for args, used in ((["--all-pulses"], pls), ([], pls[:1])):
        run_cli(tmp_path, loop_args + args)
        got = read_output(tmp_path / "output.dat", n)
        want = oracle_lib.solve(p, sysd, [oracle_pulse(oracle_lib, pl) for pl in used], ps, cs, fc, steps,
                                mode=1, record_every=1)["rec"]
        err = rel_err(got, want)
        print(f"{loop} loop, {len(used)} pulse(s): rel err {err:.3e}")
        assert np.abs(want[:, :, 1]).max() > 0.1
        assert err < 1e-10


def _write_inputs(tmp_path, sysd, pl, steps, dt, omega, order=4):
    n = len(sysd["pos"])
    with open(tmp_path / "dots.cfg", "w") as f:
        for i in range(n):
            f.write(" ".join(repr(float(v)) for v in (*sysd["pos"][i], sysd["freq"][i], sysd["t1"][i],
                                                       sysd["t2"][i], *sysd["dip"][i])) + "\n")
    with open(tmp_path / "pulse.cfg", "w") as f:
        f.write(" ".join(repr(float(v)) for v in (pl["amplitude"], pl["delay"], pl["width"], pl["freq"],
                                                   *pl["wavevector"], *pl["polarization"])) + "\n")
    with open(tmp_path / "input.cfg", "w") as f:
        f.write(f"[parameters]\n num_particles = {n}\n total_time = {steps * dt!r}\n timestep = {dt!r}\n"
                f" interpolation_order = {order}\n[constants]\n c0 = 299.792458\n hbar = 0.65821193\n"
                f" mu0 = 2.0133545e-04\n laser_frequency = {omega!r}\n")


@pytest.mark.timeout(300)
def test_binary_subsample_n4096_matches_oracle_fixture(tmp_path):
    """--format binary --every 500 --subsample 512:11 on the N = 4096 system
    (dipoles along z, pi/2 pulse, all 5000 steps, device loop): the records
    against the oracle trajectory of tests/golden/n4096_traj.npz at the same
    steps and dots, 1e-10 relative (the fixture's PC weights are numpy's,
    quest_hip's the library's; valid weight sets agree to ~1e-13); the text
    form of the same options holds the same values to 12 digits."""
    from quest_amd import outputs, systems
    g = np.load(os.path.join(ROOT, "tests", "golden", "n4096_traj.npz"))
    sysd, steps = systems.workload("n4096")
    _write_inputs(tmp_path, sysd, systems.pulse(math.pi / 2), steps, 1e-4, systems.OMEGA0)
    opts = ["--every", "500", "--subsample", "512:11"]
    run_cli(tmp_path, opts + ["--format", "binary", "--output", "out.bin"])
    b = outputs.read_binary(str(tmp_path / "out.bin"))
    assert b["n_dots"] == 4096 and b["every"] == 500 and len(b["dots"]) == 512
    assert list(b["steps"]) == [500 * (r + 1) - 1 for r in range(10)]
    assert len(set(b["dots"].tolist())) == 512 and np.all(np.diff(b["dots"]) > 0)
    want = g["rec"][:, b["dots"], :]
    err = rel_err(b["rho"], want)
    print(f"N=4096 binary subsample: 10 records x 512 dots, rel err {err:.3e}, "
          f"max|rho01| {np.abs(want[:, :, 1]).max():.3f}")
    assert np.abs(want[:, :, 1]).max() > 0.1
    assert err < 1e-10
    run_cli(tmp_path, opts + ["--output", "out.txt"])
    txt = outputs.read_text(str(tmp_path / "out.txt"), 512)
    assert txt.shape == b["rho"].shape
    assert np.abs(txt - b["rho"]).max() <= 1e-11 * np.abs(b["rho"]).max()


def test_sampled_output_host_loop(oracle_lib, tmp_path):
    """The host loop (reference integrator, GPU interactions) writes the same
    sampled records from its history: --every 7 --subsample 10:5 --format
    binary on a 40-dot system against the oracle at those steps and dots."""
    from quest_amd import outputs
    n, steps, omega = 40, 300, 2278.9
    sysd = dense_system(n, 0.2, seed=79)
    sigma = 0.005
    amp = (math.pi / 2) * 0.65821193 / (5.2917721e-4 * sigma * math.sqrt(2 * math.pi))
    pl = {"amplitude": amp, "delay": 0.015, "width": omega * sigma, "freq": omega,
          "wavevector": [omega / 299.792458, 0.0, 0.0], "polarization": [0.0, 0.0, 1.0]}
    _write_inputs(tmp_path, sysd, pl, steps, 1e-4, omega)
    run_cli(tmp_path, ["--host-loop", "--every", "7", "--subsample", "10:5", "--format", "binary",
                       "--output", "out.bin"])
    b = outputs.read_binary(str(tmp_path / "out.bin"))
    assert list(b["steps"]) == list(range(6, steps, 7)) and len(b["dots"]) == 10
    consts = {"c0": 299.792458, "hbar": 0.65821193, "mu0": 2.0133545e-04, "laser_omega": omega, "dt": 1e-4}
    want = oracle_traj(oracle_lib, sysd, consts, pl, steps, 4, mode=1)[b["steps"]][:, b["dots"], :]
    err = rel_err(b["rho"], want)
    print(f"host loop sampled: {len(b['steps'])} records, rel err {err:.3e}")
    assert np.abs(want[:, :, 1]).max() > 0.1
    assert err < 1e-10
