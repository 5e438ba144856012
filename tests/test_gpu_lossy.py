"""Device path vs the CPU checker for method dga in lossy (sigma) and
magnetic-loss (chi) media: the fractional unknowns of D:883-920 (edge
classes 1-4, face classes 1-3; operators pinned by tests/test_lossy_dga.py).
Same bars as test_gpu_parity (fp64: per probe and component within 1e-10 of
the peak, final U, F, B within 1e-10 relative L2; fp32: within 1e-4)."""
import numpy as np
import pytest

from tetfit_amd import Model
from tetfit_amd.device import Simulation

import pyoracle
from test_gpu_parity import check_run, per_probe_err, rel_l2
from test_gpu_sources import edit, set_source1
from test_lossy_dga import lossy_case

pytestmark = pytest.mark.gpu


def test_lossy_dga_parity(workdir):
    p, b = lossy_case(workdir, n_steps=300, name="glossy")
    assert b.frac_ops() is not None
    check_run(b)


def test_lossy_dga_e_source(workdir):
    """e source on the x-min wall (class-4 edges written into U_d every step)
    beside every loss class."""
    p, _ = lossy_case(workdir, n_steps=200, name="glossy_e", seed=5)
    edit(p, lambda s: set_source1(s, type="e", carrier="cos"))
    check_run(Model(p).build(1))


def test_lossy_dga_fp32(workdir):
    p, b = lossy_case(workdir, n_steps=300, name="glossy32")
    sim = Simulation(b, "fp32")
    sim.run()
    _, s32 = sim.probes()
    U, F, B = sim.state()
    sim.close()
    r = pyoracle.run(b)
    assert rel_l2(s32, r["samples"]) <= 1e-4
    for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
        assert rel_l2(got, ref) <= 1e-4


def test_lossy_dga_from_written_state(workdir):
    """A random state written into the context (each value split evenly over
    its fractional entries, as the checker does) and 60 steps without
    sources (PEC on every wall)."""
    p, b = lossy_case(workdir, n_steps=60, name="glossy_ws", pec_surfaces="1,2,3,4,5,6")
    rng = np.random.default_rng(9)
    cls = b.dga_ops()["edge_class"]
    U0 = rng.standard_normal(b.n_edges) * (cls != 4)
    F0 = rng.standard_normal(b.n_faces)
    B0 = rng.standard_normal(b.n_faces)
    sim = Simulation(b, "fp64", output=False)
    sim.set_state(U0, F0, B0)
    sim.run(n=60)
    U, F, B = sim.state()
    sim.close()
    r = pyoracle.run(b, n=60, U0=U0, F0=F0, B0=B0, output=False)
    for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
        assert rel_l2(got, ref) <= 1e-10
