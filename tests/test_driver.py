"""The tetfit command-line driver (counterpart of dga_main.cpp:38-67):
argument errors and a missing device end with status 1 (CPU tests); on a
GPU it runs every simulation of an input and writes the reference's text
outputs, which must carry the same samples as the C-ABI path."""
import os
import subprocess

import numpy as np
import pytest

from tetfit_amd import Model, configs
from tetfit_amd.device import Simulation, device_count

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tetfit_amd", "bin", "tetfit")


def run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_usage_error(tmp_path):
    r = run([], str(tmp_path))
    assert r.returncode == 1 and "usage" in r.stderr
    r = run(["a.fdtd", "--bogus"], str(tmp_path))
    assert r.returncode == 1 and "usage" in r.stderr


def test_no_device_or_bad_input(tmp_path):
    r = run([str(tmp_path / "missing.fdtd")], str(tmp_path))
    assert r.returncode == 1
    if device_count() < 1:
        assert "no HIP device" in r.stderr


def read_probe_file(path, n_probes):
    a = np.loadtxt(path)
    return a.reshape(-1, n_probes, 7)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["fit", "gfdtd"])
def test_driver_outputs_match_abi(tmp_path, kind):
    """probe_E / probe_H lines (x y z t + 3 components, setw(15), 6
    significant digits) and poynting.dat equal the C-ABI run of the same
    input to the precision printed."""
    d = str(tmp_path)
    if kind == "fit":
        path = configs.fit_cavity(d, n=10, step=0.1, n_steps=80, n_probes=3, name="drv")
    else:
        path = configs.kuhn_cavity(d, n=4, method="gfdtd", n_steps=30, n_probes=3, block=2, name="drv")
    print(f"case: driver {kind}", flush=True)
    r = run([path], d)
    assert r.returncode == 0, r.stderr
    b = Model(path).build(1)
    sim = Simulation(b, "fp64")
    sim.run()
    t, s = sim.probes()
    poy = sim.poynting()
    sim.close()
    _, pts, npr = b.probes_raw()
    E = read_probe_file(os.path.join(d, f"{b.output_name}{1:05d}_probe_E.dat"), npr)
    H = read_probe_file(os.path.join(d, f"{b.output_name}{1:05d}_probe_H.dat"), npr)
    assert E.shape[0] == len(t) and H.shape[0] == len(t)
    np.testing.assert_allclose(E[:, :, 3], np.repeat(t[:, None], npr, 1), rtol=1e-5)
    np.testing.assert_allclose(H[:, :, 3], np.repeat(t[:, None] - 0.5 * b.dt, npr, 1), rtol=1e-5, atol=1e-5 * b.dt)
    for field, col in ((E, slice(0, 3)), (H, slice(3, 6))):
        ref = s[:, :, col]
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(field[:, :, 4:7] - ref).max() <= 1e-5 * scale
    P = np.loadtxt(os.path.join(d, "poynting.dat"))
    assert P.shape[0] == len(poy)
    assert np.abs(P[:, 1] - poy).max() <= 1e-5 * max(np.abs(poy).max(), 1e-300)
    off = 1.0 if kind == "gfdtd" else 0.5
    np.testing.assert_allclose(P[:, 0], (np.arange(len(poy)) + off) * b.dt, rtol=1e-5)
