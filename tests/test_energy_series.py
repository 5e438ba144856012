"""Energy time series (the l2norm output's discrete energies, ExportFields
D:2773-2775 / ExportFitFields D:3468-3470): the oracle's We, Wh pinned by
closed forms on uniform fields, and the device series (tf_set_energy_output
/ tf_read_energy) against the oracle's series and against tf_energy."""
import os

import numpy as np
import pytest

import pyoracle
from tetfit_amd import Model, configs

EPS0 = 8.854187817e-12
MU0 = 4e-7 * np.pi
E0 = np.array([0.3, -0.7, 0.5])
B0 = np.array([0.2, 0.1, -0.4])


def _tet_build(d, method, jitter):
    os.makedirs(d, exist_ok=True)
    p = configs.kuhn_cavity(d, n=4, method=method, n_steps=3, n_probes=1, block=2, seed=3, dt=1e-12,
                            name=f"u_{method}", eps_range=(1.0, 1.0), jitter=jitter)
    return Model(p).build(1)


# We of the uniform field U_e = E0 . (x_b - x_a) on the unit cube in vacuum,
# in units of 1/2 eps0 |E0|^2 V: the Whitney mass of fem reproduces it
# exactly; the node-star blocks of dga (f = face vector x 2/(3 vol), D:8926-9078)
# give exactly 4 on any mesh; the tet-corner basis of gfdtd (a = edge middle -
# node, omega = vol/4, D:8379-8420) gives exactly 4 on the unjittered Kuhn
# mesh (and about 4.11 once jittered, so that case is not a pin).
@pytest.mark.parametrize("method,jitter,ratio", [("fem", 0.2, 1.0), ("dga", 0.2, 4.0), ("gfdtd", 0.0, 4.0)])
def test_oracle_energy_uniform_fields(tmp_path, method, jitter, ratio):
    b = _tet_build(str(tmp_path / method), method, jitter)
    t = b.tet_tables()
    pts, en, fn = t["pts"], t["edge_nodes"], t["face_nodes"]
    U = (pts[en[:, 1]] - pts[en[:, 0]]) @ E0
    # B_f = B0 . (face area vector of (p0, p1, p2))
    B = 0.5 * np.cross(pts[fn[:, 1]] - pts[fn[:, 0]], pts[fn[:, 2]] - pts[fn[:, 0]]) @ B0
    we, wh = pyoracle.energy(b, U, B)
    assert abs(we / (0.5 * EPS0 * E0 @ E0) - ratio) <= 1e-12 * ratio
    # N = sum of the tets' local mu-block inverses reproduces a uniform B exactly
    assert abs(wh / (0.5 * (B0 @ B0) / MU0) - 1.0) <= 1e-12
    # the 1/2 and both quadratic forms matter: a half-field has a quarter of each
    we2, wh2 = pyoracle.energy(b, 0.5 * U, 0.5 * B)
    assert abs(we2 - 0.25 * we) <= 1e-13 * we and abs(wh2 - 0.25 * wh) <= 1e-13 * wh


def test_oracle_energy_uniform_field_fit(tmp_path):
    """FIT: 1/2 sum ep_vec U^2 with U = E0 . edge vector sums eps L A_dual over
    the edges of each axis: 1/2 eps0 |E0|^2 V exactly on the unit cube."""
    p = configs.fit_cavity(str(tmp_path), n=6, step=1 / 6, n_steps=3, n_probes=1, name="u_fit")
    b = Model(p).build(1)
    t = b.fit_tables()
    U = (t["pts"][t["g_nodes"][:, 1]] - t["pts"][t["g_nodes"][:, 0]]) @ E0
    we, _ = pyoracle.energy(b, U, np.zeros(b.n_faces))
    assert abs(we / (0.5 * EPS0 * E0 @ E0) - 1.0) <= 1e-12


def _cases(workdir):
    return {
        "fdtd": configs.fit_cavity(str(workdir), n=10, step=0.1, n_steps=60, n_probes=2, name="es_fit"),
        "dga": configs.kuhn_cavity(str(workdir), n=5, method="dga", n_steps=60, n_probes=2, block=2,
                                   name="es_dga"),
        "gfdtd": configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=40, n_probes=2, block=2,
                                     name="es_gfdtd"),
    }


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["fdtd", "dga", "gfdtd"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-4)])
def test_energy_series_matches_oracle(workdir, method, prec, tol):
    """The device series at every output sample (same times as the probes)
    against the oracle's fp64 We, Wh of the same samples."""
    from tetfit_amd.device import Simulation
    b = Model(_cases(workdir)[method]).build(1)
    print(f"case: energy series {method} {prec} E={b.n_edges}", flush=True)
    sim = Simulation(b, prec, energy=True)
    sim.run()
    t, we, wh = sim.energy_series()
    pt, _ = sim.probes()
    sim.close()
    ref = pyoracle.run(b, energy=True)
    assert len(t) == len(ref["times"]) > 10
    assert np.array_equal(t, ref["times"]) and np.array_equal(t, pt)
    assert ref["We"].max() > 0 and ref["Wh"].max() > 0
    for got, want in ((we, ref["We"]), (wh, ref["Wh"])):
        assert np.abs(got - want).max() <= tol * np.abs(want).max()


@pytest.mark.gpu
def test_energy_series_equals_tf_energy(workdir):
    """fdtd samples after each step's update: with a sample every step, sample
    k equals tf_energy after k + 1 steps of a context without output, bit for
    bit (the same reductions)."""
    from tetfit_amd.device import Simulation
    b = Model(configs.fit_cavity(str(workdir), n=10, step=0.1, n_steps=8, n_probes=2, name="es_eq")).build(1)
    print("case: energy series vs tf_energy (fdtd)", flush=True)
    a = Simulation(b, "fp64", energy=True)
    from tetfit_amd import _ffi
    from tetfit_amd.device import _check
    _check(_ffi.hip_lib().tf_set_output(a._h, 0.0, 0.0))
    a.run(n=6)
    _, we, wh = a.energy_series()
    a.close()
    assert len(we) == 6
    c = Simulation(b, "fp64", output=False)
    for k in range(6):
        c.run(n=1)
        e = c.energy()
        assert e == (we[k], wh[k]), (k, e, we[k], wh[k])
    c.close()
