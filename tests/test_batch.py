"""Batched dga (many independent simulations in one context, one launch per
phase): every instance against its own run of the CPU checker.

Instances differ in mesh size, jitter seed, eps range, source frequency and
therefore dt; all advance the same number of steps. Bars as for single
simulations: fp64 per probe and component <= 1e-10 of the component's (or
field's) peak and final U, F, B within 1e-10 relative L2; fp32 within 1e-4
relative L2."""
import numpy as np
import pytest

from tetfit_amd import Model, configs

import pyoracle
from test_gpu_parity import per_probe_err, rel_l2

pytestmark = pytest.mark.gpu

STEPS = 80


@pytest.fixture(scope="module")
def instances(tmp_path_factory):
    d = tmp_path_factory.mktemp("batch")
    specs = [(3, 1, (1.0, 4.0), 2e8), (5, 2, (1.0, 10.0), 3e8), (4, 3, (2.0, 6.0), 1.5e8), (6, 4, (1.0, 10.0), 4e8)]
    out = []
    for k, (n, seed, er, f) in enumerate(specs):
        p = configs.kuhn_cavity(str(d), n=n, method="dga", n_steps=STEPS, n_probes=3, block=2, seed=seed,
                                eps_range=er, freq=f, name=f"b{k}")
        out.append(Model(p).build(1))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_matches_each_instance(instances, prec):
    from tetfit_amd.device import BatchSimulation
    dts = sorted({b.dt for b in instances})
    print(f"case: batch {prec} of {len(instances)} dga instances, dts {dts}", flush=True)
    assert len(dts) == len(instances)  # the instances really run on different dt
    sim = BatchSimulation(instances, prec, sample_every=1)
    sim.run(STEPS)
    for k, b in enumerate(instances):
        t, s = sim.probes(k)
        U, F, B = sim.state(k)
        r = pyoracle.run(b, n=STEPS)
        np.testing.assert_allclose(t, r["times"], rtol=1e-15, atol=0)
        if prec == "fp64":
            assert per_probe_err(s, r["samples"]) <= 1e-10, k
            for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
                assert rel_l2(got, ref) <= 1e-10, k
        else:
            assert rel_l2(s, r["samples"]) <= 1e-4, k
            assert rel_l2(U, r["U"]) <= 1e-4, k
    sim.close()


@pytest.fixture(scope="module")
def gf_instances(tmp_path_factory):
    d = tmp_path_factory.mktemp("gfbatch")
    specs = [(3, 1, (1.0, 4.0), 2e8), (5, 2, (1.0, 10.0), 3e8), (4, 3, (2.0, 6.0), 1.5e8), (6, 4, (1.0, 10.0), 4e8),
             (2, 5, (1.0, 3.0), 2.5e8)]
    out = []
    for k, (n, seed, er, f) in enumerate(specs):
        p = configs.kuhn_cavity(str(d), n=n, method="gfdtd", n_steps=STEPS, n_probes=3, block=2, seed=seed,
                                eps_range=er, freq=f, name=f"g{k}")
        if k == 1:
            # an e source (Dirichlet values through dir_new) and a j dipole beside the h source
            from test_gpu_sources import add_source, edit
            edit(p, lambda s: add_source(s, 2, ["type j", "location {0.43,0.57,0.61}", "direction z",
                                                "carrier ricker", "frequency %r" % f, "delay %r" % (1.5 / f),
                                                "amplitude 2", "impedance 3"]))
            edit(p, lambda s: add_source(s, 3, ["type e", "surface 4", "direction y", "carrier gaussian",
                                                "width %r" % (1.0 / f), "delay %r" % (2.0 / f)]))
        out.append(Model(p).build(1))
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gfdtd_batch_matches_each_instance(gf_instances, prec):
    """Batched gfdtd: one Jacobi CG per instance per step inside the stacked
    system (each instance's own alpha, beta and stop test; its rows padded to
    whole 64-row slices with solver-only edges). Every instance against its
    own checker run: fp64 within 1e-10 per probe and final fields, every
    solve at the tolerance, iteration counts within +-1 per step and
    instance on average, the summed Poynting series within 1e-9; fp32
    within 1e-4."""
    from tetfit_amd.device import BatchSimulation
    dts = sorted({b.dt for b in gf_instances})
    print(f"case: gfdtd batch {prec} of {len(gf_instances)} instances, dts {dts}", flush=True)
    assert len(dts) == len(gf_instances)
    sim = BatchSimulation(gf_instances, prec, sample_every=1)
    st = sim.run(STEPS)
    poy = sim.poynting()
    iters, poy_ref = 0, np.zeros(STEPS)
    for k, b in enumerate(gf_instances):
        t, s = sim.probes(k)
        U, F, B = sim.state(k)
        r = pyoracle.run(b, n=STEPS)
        iters += r["cg_iters_total"]
        poy_ref += np.asarray(r["poynting"])[:STEPS]
        np.testing.assert_allclose(t, r["times"], rtol=1e-15, atol=0)
        if prec == "fp64":
            assert per_probe_err(s, r["samples"]) <= 1e-10, k
            for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
                assert rel_l2(got, ref) <= 1e-10, k
        else:
            assert rel_l2(s, r["samples"]) <= 1e-4, k
            assert rel_l2(U, r["U"]) <= 1e-4, k
    if prec == "fp64":
        assert st["cg_not_converged"] == 0 and st["cg_max_relres"] <= gf_instances[0].cg_tol
        assert abs(st["cg_iters_total"] - iters) <= STEPS * len(gf_instances)
        np.testing.assert_allclose(poy, poy_ref, rtol=1e-9, atol=1e-12 * max(np.abs(poy_ref).max(), 1))
    sim.close()
