"""Device path vs the CPU checker for the source kinds beyond the wave-profile
h source of the other parity tests: imposed-field e sources (FIT special
edges, D:1850-1895; dga/gfdtd class-4 / Dirichlet edges, D:821-868,
D:1490-1510), j current dipoles on FIT (D:447-492) and gfdtd (D:429-462), several sources in one
simulation, and the gaussian, ricker and cos carriers (Source.hpp:313-567).
Same bars as test_gpu_parity: fp64 per probe and component within 1e-10 of
the component's (or field's) peak, final fields within 1e-10 relative L2."""
import re

import pytest

from tetfit_amd import Model, configs

import pyoracle
from test_gpu_parity import check_run

pytestmark = pytest.mark.gpu


def edit(path, fn):
    s = open(path).read()
    open(path, "w").write(fn(s))
    return path


def set_source1(s, **kv):
    """Replace / add SET lines of source 1."""
    a = s.index("DEFINE source 1")
    b = s.index("END source 1")
    block = s[a:b]
    for k, v in kv.items():
        if re.search(r"SET %s " % k, block):
            block = re.sub(r"SET %s [^\n]*" % k, "SET %s %s" % (k, v), block)
        else:
            block = block + "    SET %s %s\n" % (k, v)
    return s[:a] + block + s[b:]


def add_source(s, label, lines):
    block = "DEFINE source %d\n%sEND source %d\n" % (label, "".join("    SET %s\n" % l for l in lines), label)
    s = s.replace("DEFINE output 1", block + "DEFINE output 1", 1)
    return s.replace("DEFINE simulation 1\n", "DEFINE simulation 1\n    SET source %d\n" % label, 1)


def test_fit_e_source_gaussian(workdir):
    p = configs.fit_cavity(str(workdir), n=12, step=0.1, n_steps=150, n_probes=4, name="fit_esrc")
    f = 2e8
    edit(p, lambda s: set_source1(s, type="e", carrier="gaussian", width=repr(1.0 / f), delay=repr(2.0 / f)))
    check_run(Model(p).build(1))


def test_fit_h_plus_j_dipole_ricker(workdir):
    p = configs.fit_cavity(str(workdir), n=12, step=0.1, n_steps=150, n_probes=4, name="fit_jsrc")
    f = 3e8
    edit(p, lambda s: add_source(s, 2, ["type j", "location {0.55,0.65,0.45}", "direction z", "carrier ricker",
                                        "frequency %r" % f, "delay %r" % (1.5 / f), "amplitude 2",
                                        "impedance 3"]))
    check_run(Model(p).build(1))


def test_dga_e_source_cos(workdir):
    p = configs.kuhn_cavity(str(workdir), n=5, method="dga", n_steps=150, n_probes=4, block=2, name="dga_esrc")
    edit(p, lambda s: set_source1(s, type="e", carrier="cos"))
    check_run(Model(p).build(1))


def test_gfdtd_e_source(workdir):
    p = configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=50, n_probes=4, block=2, name="gf_esrc")
    edit(p, lambda s: set_source1(s, type="e"))
    st, r = check_run(Model(p).build(1))
    assert st["cg_not_converged"] == 0


def test_gfdtd_h_plus_j_dipole(workdir):
    """gfdtd with the wave-profile h source and a j dipole (-dt I in Psi,
    D:1491-1516): the initial residual carries the current on a free edge."""
    p = configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=50, n_probes=4, block=2, name="gf_jsrc")
    f = 3e8
    edit(p, lambda s: add_source(s, 2, ["type j", "location {0.43,0.57,0.61}", "direction z", "carrier ricker",
                                        "frequency %r" % f, "delay %r" % (1.5 / f), "amplitude 2", "impedance 3"]))
    st, r = check_run(Model(p).build(1))
    assert st["cg_not_converged"] == 0


def test_gfdtd_lossy(workdir):
    """gfdtd with conducting blocks (sigma 0.5 S/m on every other eps block):
    A = E + dt/2 SigMat and Psi's - dt/2 SigMat U_old (D:1516, D:8646)."""
    from test_tet_host import _lossy
    p = configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=50, n_probes=4, block=2, name="gf_lossy")
    st, r = check_run(Model(_lossy(p, 0.5, 2)).build(1))
    assert st["cg_not_converged"] == 0


def test_gfdtd_lossy_losses_column(workdir):
    """tf_read_losses against the checker's Losses vector: the first n
    values, one per step, as poynting.dat pairs them (D:1573-1576)."""
    import numpy as np
    from tetfit_amd.device import Simulation
    from test_tet_host import _lossy
    p = configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=40, n_probes=2, block=2, name="gf_loss")
    b = Model(_lossy(p, 0.5, 2)).build(1)
    sim = Simulation(b, "fp64")
    sim.run(n=15)
    sim.run()  # the column continues across runs
    got = sim.losses()
    sim.close()
    ref = pyoracle.run(b)["losses"][:b.n_steps]
    assert len(got) == b.n_steps and np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


def test_gfdtd_e_source_cos(workdir):
    """An e source whose carrier is not zero at t = 0 (cos, with the amp^2
    quirk): the pass before the loop (D:1434-1453) sets U on its Dirichlet
    edges at t = 0, which step 1 reads through E U_old."""
    p = configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=40, n_probes=4, block=2, name="gf_ecos")
    edit(p, lambda s: set_source1(s, type="e", carrier="cos", amplitude="1.5"))
    st, r = check_run(Model(p).build(1))
    assert st["cg_not_converged"] == 0


@pytest.mark.parametrize("method", ["gfdtd", "dga"])
def test_written_state(workdir, method):
    """tf_write_state then a run from step 1 against the checker started from
    the same fields: gfdtd warm-starts its first solve from the written U on
    the free edges and the pass before the loop resets the Dirichlet edges
    (random values written there too); dga from a state that is zero on its
    class-4 edges."""
    import numpy as np
    from tetfit_amd.device import Simulation
    b = Model(configs.kuhn_cavity(str(workdir), n=5, method=method, n_steps=40, n_probes=4, block=2,
                                  name=f"ws_{method}")).build(1)
    rng = np.random.default_rng(11)
    U0 = rng.standard_normal(b.n_edges)
    B0 = 1e-8 * rng.standard_normal(b.n_faces)
    F0 = None
    if method == "dga":
        U0[b.dga_ops()["edge_class"] == 4] = 0.0
        F0 = 1e-2 * rng.standard_normal(b.n_faces)
    print(f"case: written state {method} E={b.n_edges}", flush=True)
    sim = Simulation(b, "fp64")
    sim.set_state(U=U0, F=F0, B=B0)
    sim.run()
    t_g, s_g = sim.probes()
    U, F, B = sim.state()
    sim.close()
    r = pyoracle.run(b, U0=U0, F0=F0, B0=B0)
    from test_gpu_parity import per_probe_err, rel_l2
    assert per_probe_err(s_g, r["samples"]) <= 1e-10
    for got, ref in ((U, r["U"]), (B, r["B"])):
        assert rel_l2(got, ref) <= 1e-10


@pytest.mark.parametrize("method", ["dga", "gfdtd"])
def test_d_source_wall(workdir, method):
    """A d source on the x-max wall (not PEC) beside the h source on x-min:
    dga resets U on its edges to ComputeEfieldSource's 0 before each H update
    (D:835-848), gfdtd leaves them to the solve (D:1497-1499, D:1540)."""
    from test_source_routing import d_wall_case
    b = Model(d_wall_case(workdir, method, f"gpu_dwall_{method}")).build(1)
    print(f"case: d source wall {method} E={b.n_edges}", flush=True)
    st, r = check_run(b)
    if method == "gfdtd":
        assert st["cg_not_converged"] == 0
