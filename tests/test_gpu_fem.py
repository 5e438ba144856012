"""Device path vs the CPU checker for method fem (Whitney edge elements,
second order in time, CG per step; D:1580-1790): the device forms each
step's initial residual from W = U_old - U_older (fem_w_kernel) instead of
the reference's Psi - A x0, the same system in exact arithmetic. Bars as in
test_gpu_parity: fp64 per probe and component within 1e-10 of the peak,
final U and B within 1e-10 relative L2, every solve at the tolerance and
iteration counts within +-1 per step on average; fp32 within 1e-4."""
import numpy as np
import pytest

from tetfit_amd import Model
from tetfit_amd.device import Simulation

import pyoracle
from test_fem import fem_case
from test_gpu_parity import per_probe_err, rel_l2
from test_gpu_sources import add_source, edit

pytestmark = pytest.mark.gpu


def check_fem(b, prec="fp64", tol=1e-10, **kw):
    print(f"case: fem E={b.n_edges} steps={b.n_steps} {prec}", flush=True)
    sim = Simulation(b, prec)
    st = sim.run(**kw)
    t_g, s_g = sim.probes()
    U, _, B = sim.state()
    poy = sim.poynting()
    losses = sim.losses()
    sim.close()
    r = pyoracle.run(b)
    np.testing.assert_allclose(t_g, r["times"], rtol=0, atol=0)
    assert per_probe_err(s_g, r["samples"]) <= tol
    assert rel_l2(U, r["U"]) <= tol and rel_l2(B, r["B"]) <= tol
    if prec == "fp64":
        assert st["cg_not_converged"] == 0 and st["cg_max_relres"] <= b.cg_tol
        n = b.n_steps
        assert abs(st["cg_iters_total"] - r["cg_iters_total"]) <= n
        np.testing.assert_allclose(poy, r["poynting"], rtol=1e-9, atol=1e-12 * max(np.abs(r["poynting"]).max(), 1))
        np.testing.assert_allclose(losses, r["losses"][:len(losses)], rtol=1e-9,
                                   atol=1e-12 * max(np.abs(r["losses"]).max(), 1e-300))
    return st, r


def test_fem_parity(workdir):
    _, b = fem_case(workdir, n=5, seed=3, name="gfem", n_steps=150, dt=1.5e-11)
    check_fem(b)


def test_fem_sigma_row_order(workdir, monkeypatch):
    """The CG's free rows in SELL-C-sigma order instead of the edges'
    locality rank: same bars (the edge-ordered scatter and W kernels)."""
    monkeypatch.setenv("TETFIT_ROW_ORDER", "sigma")
    _, b = fem_case(workdir, n=5, seed=3, name="gfem_sig", n_steps=60, dt=1.5e-11)
    check_fem(b)


def test_fem_lossy_with_j_dipole(workdir):
    """sigma in half of the blocks (the SigMat terms of Psi and the losses)
    and a j dipole beside the h source (-I/dt in Psi)."""
    p, _ = fem_case(workdir, n=5, seed=4, name="gfem_l", sigma=0.5, sigma_frac=0.5, n_steps=120, dt=1.5e-11)
    f = 3e8
    edit(p, lambda s: add_source(s, 2, ["type j", "location {0.43,0.57,0.61}", "direction z", "carrier ricker",
                                        "frequency %r" % f, "delay %r" % (1.5 / f), "amplitude 2", "impedance 3"]))
    check_fem(Model(p).build(1))


def test_fem_fp32(workdir):
    _, b = fem_case(workdir, n=5, seed=3, name="gfem32", n_steps=150, dt=1.5e-11)
    check_fem(b, prec="fp32", tol=1e-4)


def test_fem_learned_preconditioner_keeps_results(workdir):
    """The learned preconditioner on the fem system (A SPD): every solve
    still meets the tolerance and the fields stay within 1e-7 of the
    checker's (plain CG judges convergence)."""
    _, b = fem_case(workdir, n=6, seed=3, name="gfem_lp", n_steps=80, dt=1.5e-11)
    sim = Simulation(b, "fp64")
    st = sim.run(precond="learned", learn_steps=40)
    U, _, B = sim.state()
    sim.close()
    r = pyoracle.run(b)
    assert st["cg_not_converged"] == 0
    assert rel_l2(U, r["U"]) <= 1e-7 and rel_l2(B, r["B"]) <= 1e-7
