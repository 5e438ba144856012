"""Device path vs the CPU checker on the same assembled operators.

Tolerance (fp64 plain path): per probe and component,
max_t |x_gpu - x_oracle| <= 1e-10 * max_t |x_oracle|; final fields within
1e-10 relative L2. gfdtd additionally checks that every CG solve meets the
relative-residual tolerance and that iteration counts match the checker's
within +-1 per step on average.
"""
import numpy as np
import pytest

from tetfit_amd import Model, configs
from tetfit_amd.device import Simulation

import pyoracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def per_probe_err(got, ref):
    """max over probes and components of max_t|got-ref| / max_t|ref|. A
    component that is zero by symmetry carries only rounding noise, so its
    scale is floored at the largest component of the same field (E or H) at
    that probe."""
    comp = np.abs(ref).max(axis=0)                     # [probe, 6]
    field = np.concatenate([np.repeat(comp[:, :3].max(axis=1, keepdims=True), 3, axis=1),
                            np.repeat(comp[:, 3:].max(axis=1, keepdims=True), 3, axis=1)], axis=1)
    scale = np.maximum(np.maximum(comp, field), 1e-300)
    return (np.abs(got - ref).max(axis=0) / scale).max()


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def check_run(b, n=None, **kw):
    print(f"case: {b.method} E={b.n_edges} F={b.n_faces} steps={b.n_steps}", flush=True)
    sim = Simulation(b, "fp64")
    st = sim.run(n=n, **kw)
    t_g, s_g = sim.probes()
    U, F, B = sim.state()
    poy = sim.poynting()
    sim.close()
    r = pyoracle.run(b, n=n)
    assert len(t_g) == len(r["times"])
    np.testing.assert_allclose(t_g, r["times"], rtol=0, atol=0)
    assert per_probe_err(s_g, r["samples"]) <= TOL
    for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
        assert rel_l2(got, ref) <= TOL
    np.testing.assert_allclose(poy, r["poynting"], rtol=1e-9, atol=1e-12 * max(np.abs(r["poynting"]).max(), 1))
    return st, r


def test_fit_parity(workdir):
    b = Model(configs.fit_cavity(str(workdir), n=12, step=0.1, n_steps=200, n_probes=6, name="fit12")).build(1)
    check_run(b)


def test_dga_parity(workdir):
    b = Model(configs.kuhn_cavity(str(workdir), n=6, method="dga", n_steps=200, n_probes=4, block=2)).build(1)
    check_run(b)


def test_gfdtd_parity(workdir):
    b = Model(configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=60, n_probes=4, block=2)).build(1)
    st, r = check_run(b)
    assert st["cg_max_relres"] <= b.cg_tol
    assert st["cg_not_converged"] == 0
    assert abs(st["cg_iters_total"] - r["cg_iters_total"]) <= b.n_steps


# ---- fp32 fields, fp64 reductions: error stated against the fp64 checker.
# Bar of the fast path: relative L2 over all probe samples <= 1e-4.
FP32_TOL = 1e-4


def fp32_err(b, n=None):
    print(f"case: fp32 {b.method} E={b.n_edges} steps={b.n_steps}", flush=True)
    sim = Simulation(b, "fp32")
    st = sim.run(n=n)
    _, s = sim.probes()
    U, F, B = sim.state()
    sim.close()
    r = pyoracle.run(b, n=n)
    err = rel_l2(s, r["samples"])
    print(f"fp32 {b.method}: probe rel L2 {err:.3e}, U rel L2 {rel_l2(U, r['U']):.3e}", flush=True)
    return err, st


def test_fit_fp32(workdir):
    b = Model(configs.fit_cavity(str(workdir), n=12, step=0.1, n_steps=200, n_probes=6, name="fit12")).build(1)
    err, _ = fp32_err(b)
    assert err <= FP32_TOL


def test_dga_fp32(workdir):
    b = Model(configs.kuhn_cavity(str(workdir), n=6, method="dga", n_steps=200, n_probes=4, block=2)).build(1)
    err, _ = fp32_err(b)
    assert err <= FP32_TOL


def test_gfdtd_fp32(workdir):
    b = Model(configs.kuhn_cavity(str(workdir), n=5, method="gfdtd", n_steps=60, n_probes=4, block=2)).build(1)
    err, st = fp32_err(b)
    assert err <= FP32_TOL


@pytest.mark.parametrize("method", ["dga", "gfdtd"])
def test_autotune_keeps_results_bitwise(workdir, method):
    """The autotuner changes only launch shapes (grid size, gather batch);
    every row is still summed by one lane in column order, so the explicit
    dga fields are bitwise identical with and without tuning. The CG's dot
    products are summed per workgroup, so a tuned grid changes their
    rounding: gfdtd fields then agree to 1e-12 relative with the same
    iteration counts."""
    b = Model(configs.kuhn_cavity(str(workdir), n=5, method=method, n_steps=40, n_probes=2, block=2,
                                  name=f"tune_{method}")).build(1)
    print(f"case: autotune {method} E={b.n_edges}", flush=True)
    a = Simulation(b, "fp64")
    sa = a.run()
    Ua, Fa, Ba = a.state()
    a.close()
    t = Simulation(b, "fp64")
    tuned = t.autotune()
    print(tuned, flush=True)
    st = t.run()
    Ut, Ft, Bt = t.state()
    t.close()
    if method == "dga":
        assert np.array_equal(Ua, Ut) and np.array_equal(Fa, Ft) and np.array_equal(Ba, Bt)
    else:
        assert st["cg_iters_total"] == sa["cg_iters_total"]
        for x, y in ((Ua, Ut), (Fa, Ft), (Ba, Bt)):
            assert rel_l2(y, x) <= 1e-12


def test_fit_random_lossy_parity(workdir):
    """Per-cell random eps_r in [1, 4), a sigma = 0.05 S/m slab (lossy M_h,
    M_q) and the gaussmod carrier: the ingredients of the 512^3 precision
    study at a size the checker runs in seconds. fp64 at the 1e-10 bar,
    fp32 at the 1e-4 bar."""
    b = Model(configs.fit_random_media(str(workdir), n=12, n_steps=160, n_probes=4, name="fitrand12")).build(1)
    check_run(b)
    err, _ = fp32_err(b)
    assert err <= FP32_TOL


def test_fit_energy_diagnostic(workdir):
    """tf_energy (fp64 reductions on the device) equals 1/2 U.(ep U) and
    1/2 B.(nu B) computed on the host from the read-back state (D:3468-3470),
    for fp64 and fp32 fields."""
    b = Model(configs.fit_cavity(str(workdir), n=10, step=0.1, n_steps=120, n_probes=2, name="fit_energy")).build(1)
    f = b.fit_ops()
    for prec in ("fp64", "fp32"):
        print(f"case: energy {prec}", flush=True)
        sim = Simulation(b, prec)
        sim.run()
        we, wh = sim.energy()
        U, _, B = sim.state()
        sim.close()
        ref_we, ref_wh = 0.5 * U @ (f["ep_vec"] * U), 0.5 * B @ (f["nu_vec"] * B)
        assert ref_we > 0 and ref_wh > 0
        assert abs(we - ref_we) <= 1e-12 * ref_we and abs(wh - ref_wh) <= 1e-12 * ref_wh


@pytest.mark.parametrize("method", ["dga", "gfdtd"])
def test_tet_energy_diagnostic(workdir, method):
    """tf_energy on dga / gfdtd contexts: We = 1/2 U.(E U), Wh = 1/2 B.(N B)
    (D:2773-2775) reduced in fp64 on the device, equal to the same forms
    computed on the host from the read-back state and the host's E and N
    (for fp32 contexts with E and N rounded to fp32, as the device stores
    them)."""
    import scipy.sparse as sp
    b = Model(configs.kuhn_cavity(str(workdir), n=4, method=method, n_steps=40, n_probes=2, block=2,
                                  name=f"energy_{method}")).build(1)
    d = b.dga_ops()
    for prec in ("fp64", "fp32"):
        print(f"case: tet energy {method} {prec}", flush=True)
        sim = Simulation(b, prec)
        sim.run()
        we, wh = sim.energy()
        U, _, B = sim.state()
        sim.close()
        cast = (lambda v: v) if prec == "fp64" else (lambda v: v.astype(np.float32).astype(np.float64))
        Em = sp.csr_matrix((cast(d["E"]["val"]), d["E"]["col"], d["E"]["rowptr"]), shape=(b.n_edges, b.n_edges))
        Nm = sp.csr_matrix((cast(d["N"]["val"]), d["N"]["col"], d["N"]["rowptr"]), shape=(b.n_faces, b.n_faces))
        ref_we, ref_wh = 0.5 * U @ (Em @ U), 0.5 * B @ (Nm @ B)
        assert ref_we > 0 and ref_wh > 0
        assert abs(we - ref_we) <= 1e-12 * ref_we and abs(wh - ref_wh) <= 1e-12 * ref_wh


@pytest.mark.parametrize("method", ["fdtd", "dga"])
def test_graph_steps_match_eager(workdir, method):
    """With nothing sampled, fdtd and dga step in HIP graphs of 64 steps
    (captured once, replayed with the carrier rows refilled and the Poynting
    partials drained by stream copies around each replay). Several runs of
    150 steps (graph replays with eager remainder steps and host waits in
    between) give fields and Poynting values bitwise equal to launching
    every step from the host."""
    if method == "fdtd":
        b = Model(configs.fit_cavity(str(workdir), n=10, step=0.1, n_steps=400, n_probes=2, name="graph_fit")).build(1)
    else:
        b = Model(configs.kuhn_cavity(str(workdir), n=5, method="dga", n_steps=400, n_probes=2, block=2,
                                      name="graph_dga")).build(1)
    out = {}
    for graphs in (True, False):
        print(f"case: {method} graphs={graphs}", flush=True)
        sim = Simulation(b, "fp64", output=False)
        sim.set_graphs(graphs)
        for _ in range(2):
            sim.run(n=150)
        out[graphs] = (*sim.state(), sim.poynting())
        sim.close()
    for a, e in zip(out[True], out[False]):
        assert np.array_equal(a, e)
    assert np.abs(out[True][0]).max() > 0


def test_batch_graph_steps_match_eager(workdir):
    from tetfit_amd.device import BatchSimulation
    builds = [Model(configs.kuhn_cavity(str(workdir), n=n, method="dga", n_steps=300, n_probes=1, block=2, seed=n,
                                        freq=f, name=f"gb{n}")).build(1) for n, f in ((3, 2e8), (4, 3e8))]
    out = {}
    for graphs in (True, False):
        print(f"case: batch graphs={graphs}", flush=True)
        sim = BatchSimulation(builds, "fp64")
        sim.set_graphs(graphs)
        for _ in range(2):
            sim.run(140)
        out[graphs] = [sim.state(k) for k in range(len(builds))]
        sim.close()
    for k in range(len(builds)):
        for a, e in zip(out[True][k], out[False][k]):
            assert np.array_equal(a, e)


@pytest.mark.parametrize("layout", ["col32", "win", "wide"])
@pytest.mark.parametrize("method", ["dga", "gfdtd"])
def test_column_layouts_parity(workdir, method, layout, monkeypatch):
    """Every SELL operator stored with int32 column indices, or with 16-bit
    codes over four column windows per slice (the layouts for meshes whose
    slices span more than 65,536 columns), gives the same fp64 parity; so do
    the wide-mesh kernels (non-temporal operator loads) on one-base 16-bit
    columns, the combination of a large mesh numbered compactly."""
    env = {"col32": "TETFIT_SELL_COL32", "win": "TETFIT_SELL_WIN", "wide": "TETFIT_SELL_WIDE"}[layout]
    monkeypatch.setenv(env, "1")
    b = Model(configs.kuhn_cavity(str(workdir), n=5, method=method, n_steps=60, n_probes=4, block=2,
                                  name=f"c{layout}_{method}")).build(1)
    print(f"case: {method} {layout} columns E={b.n_edges}", flush=True)
    sim = Simulation(b, "fp64")
    for op in ("n", "a") if method == "gfdtd" else ("n", "h", "ct"):
        info = sim.sell_info(op)
        want = {"col32": (4, 0), "win": (2, 4), "wide": (2, 1)}[layout]
        assert (info["col_bytes"], info["col_bases"]) == want, (op, info)
    sim.close()
    check_run(b)


@pytest.mark.parametrize("prec,n", [("fp64", 10), ("fp32", 10), ("fp64", 40)])
def test_fit_grid_order_matches_indexed(workdir, prec, n, monkeypatch):
    """The FIT step in grid order (faces of an edge and edges of a face found
    from their grid positions) against the index-table layout on a non-cubic
    grid with an h source, a j dipole and an e source: probes, Poynting
    values, final fields and energy bitwise equal (both sum each edge's four
    faces in the reference's order) and the energy to rounding; the grid
    order holds less memory. The n x 8 x n grids: at n = 10 the x, y and z
    edges (and faces) each fill 4 blocks of 256, so every XCD's slab has
    equal direction lengths; at n = 40 they fill 58, 53 and 58 blocks, so
    the XCDs' longer directions append their extra blocks after the
    interleaved part."""
    import re
    step = 1.0 / n
    p = configs.fit_cavity(str(workdir), n=n, step=step, n_steps=80, n_probes=5, name=f"fgo_{prec}_{n}")
    s = open(p).read()
    s = re.sub(r"ygrid \{0,([^,]*),([^}]*)\}", lambda m: "ygrid {0,%s,%r}" % (m.group(1), 8.25 * step), s)
    s = re.sub(r"SET probe \{([^,]*),([^,]*),([^}]*)\}",
               lambda m: "SET probe {%s,%r,%s}" % (m.group(1), 4.5 * step, m.group(3)), s)
    blocks = ("DEFINE source 2\n    SET type j\n    SET location {0.55,%r,0.35}\n    SET direction z\n"
              "    SET carrier ricker\n    SET frequency 3e8\n    SET delay 5e-9\nEND source 2\n" % (4.5 * step) +
              "DEFINE source 3\n    SET type e\n    SET surface 4\n    SET direction y\n    SET carrier gaussian\n"
              "    SET width 5e-9\n    SET delay 1e-8\nEND source 3\n")
    s = s.replace("DEFINE output 1", blocks + "DEFINE output 1", 1)
    s = s.replace("DEFINE simulation 1\n", "DEFINE simulation 1\n    SET source 2\n    SET source 3\n", 1)
    s = re.sub(r"SET surfaces \{[^}]*\}", "SET surfaces {1,3,5,6}", s, count=1)
    open(p, "w").write(s)
    b = Model(p).build(1)
    out = {}
    for indexed in (False, True):
        monkeypatch.setenv("TETFIT_FIT_INDEXED", "1" if indexed else "0")
        print(f"case: fit {prec} indexed={indexed} E={b.n_edges}", flush=True)
        sim = Simulation(b, prec)
        nbytes = sim.device_bytes()
        sim.run()
        t, smp = sim.probes()
        U, F, B = sim.state()
        out[indexed] = (nbytes, t, smp, U, F, B, sim.poynting(), sim.energy())
        sim.close()
    g, ix = out[False], out[True]
    assert g[0] < ix[0]
    for a, c in zip(g[1:7], ix[1:7]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(c))
    # the energy sums its terms in device order, which differs between layouts
    np.testing.assert_allclose(g[7], ix[7], rtol=1e-13)
    assert np.abs(g[2]).max() > 0


def test_fit_pml_parity(workdir):
    """FIT with perfectly matched layers on +x and +z (upml / fpml below 1 in
    the layers, folded into the grid-order coefficients)."""
    b = Model(configs.fit_cavity(str(workdir), n=12, step=0.1, n_steps=200, n_probes=6, name="fitpml",
                                 pml={"+x": 0.35, "+z": 0.25})).build(1)
    assert np.any(b.fit_ops()["upml"] < 1)
    check_run(b)


@pytest.mark.parametrize("prec,precond", [("fp64", "jacobi"), ("fp32", "learned")])
def test_slice_stream_bitwise(workdir, prec, precond, monkeypatch):
    """The CG kernels stream windowed-column slices through a two-stage
    software pipeline (next batch's entries loaded while the current batch's
    gathers are in flight); one-base columns take the unpipelined slice loop
    of the same wide-mesh kernels. Both sum every row in entry order, so a
    gfdtd run (Jacobi fp64, learned fp32) gives bitwise equal fields and
    iteration counts either way."""
    from tetfit_amd.device import Simulation
    b = Model(configs.kuhn_cavity(str(workdir), n=10, method="gfdtd", n_steps=60, n_probes=4, block=2, seed=5,
                                  eps_range=(1.0, 80.0), log_eps=True, name="stream_bw")).build(1)
    out = {}
    for layout, env in (("win", "TETFIT_SELL_WIN"), ("wide", "TETFIT_SELL_WIDE")):
        monkeypatch.delenv("TETFIT_SELL_WIN", raising=False)
        monkeypatch.delenv("TETFIT_SELL_WIDE", raising=False)
        monkeypatch.setenv(env, "1")
        print(f"case: slice stream {prec} {precond} {layout} E={b.n_edges}", flush=True)
        sim = Simulation(b, prec)
        info = sim.sell_info("a")
        assert info["col_bases"] == (4 if layout == "win" else 1)
        if precond == "learned":
            sim.run(n=30, precond="learned", learn_steps=30, train_iters=40)
            st = sim.run(n=30, precond="learned")
            assert st["pc_state"] == 2
        else:
            st = sim.run()
        out[layout] = (st["cg_iters_total"],) + sim.state()
        sim.close()
    assert out["win"][0] == out["wide"][0]
    for a, c in zip(out["win"][1:], out["wide"][1:]):
        assert np.array_equal(a, c)


@pytest.fixture(scope="module")
def contrast10(tmp_path_factory):
    d = tmp_path_factory.mktemp("extrap")
    return Model(configs.kuhn_cavity(str(d), n=10, method="gfdtd", n_steps=80, n_probes=4, block=2, seed=5,
                                     eps_range=(1.0, 80.0), log_eps=True, name="extrap10")).build(1)


@pytest.mark.parametrize("guess", [1, 2])
def test_extrapolated_start(contrast10, guess):
    """gfdtd solves started from the linear / quadratic extrapolation of the
    previous steps (fp64, Jacobi): every solve still meets the reference's
    test ||r|| <= tol ||b - A U_old||, now with fewer CG iterations in total;
    fields and probes agree with warm-started runs and the oracle to the
    tolerance's accumulated effect; the last solve's fp64 residual, measured
    against a tight reference solve from the same state with the host's A,
    is within tol of the warm-start residual."""
    import scipy.sparse as sp
    b = contrast10
    print(f"case: extrapolated start order {guess} E={b.n_edges}", flush=True)
    sa = Simulation(b, "fp64")
    st_a = sa.run()
    Ua = sa.state()[0]
    sa.close()
    sb = Simulation(b, "fp64")
    st_b1 = sb.run(n=b.n_steps - 1, guess=guess)
    S = sb.state()
    st_b2 = sb.run(n=1, guess=guess)
    Ub = sb.state()[0]
    _, samples = sb.probes()
    sb.close()
    tol = b.cg_tol
    iters_a = st_a["cg_iters_total"]
    iters_b = st_b1["cg_iters_total"] + st_b2["cg_iters_total"]
    print("iterations warm start", iters_a, "extrapolated", iters_b, flush=True)
    assert iters_b < iters_a
    for st in (st_b1, st_b2):
        assert st["cg_not_converged"] == 0 and st["cg_max_relres"] <= tol
    assert rel_l2(Ub, Ua) <= 1e-6
    r = pyoracle.run(b)
    assert rel_l2(samples, r["samples"]) <= 1e-6
    # the last solve against a tight solve from the same state
    ref = Simulation(b, "fp64", output=False)
    ref.set_state(*S)
    ref.run(n=1, first=b.first_step + b.n_steps - 1, cg_tol=1e-13, cg_maxiter=5000)
    Ustar = ref.state()[0]
    ref.close()
    ops = b.dga_ops()
    A, free = ops["A"], ops["free_edges"]
    Am = sp.csr_matrix((A["val"], A["col"], A["rowptr"]), shape=(A["nrows"], A["ncols"]))
    ratio = np.linalg.norm(Am @ (Ustar[free] - Ub[free])) / np.linalg.norm(Am @ (Ustar[free] - S[0][free]))
    assert ratio <= 1.05 * tol, ratio


def test_extrapolated_start_learned(contrast10):
    """The learned fp32 path with quadratic-extrapolated starts: G trained
    and kept (no fallbacks), every solve converged, fields against the
    oracle to the fp32 bar."""
    b = contrast10
    print(f"case: extrapolated start learned fp32 E={b.n_edges}", flush=True)
    sim = Simulation(b, "fp32")
    s1 = sim.run(n=30, precond="learned", learn_steps=30, train_iters=60, guess=2)
    s2 = sim.run(n=b.n_steps - 30, precond="learned", guess=2)
    _, samples = sim.probes()
    U, F, B = sim.state()
    sim.close()
    assert s2["pc_state"] == 2 and s2["pc_fallbacks"] == 0
    assert s1["cg_not_converged"] == 0 and s2["cg_not_converged"] == 0 and s2["cg_max_relres"] <= b.cg_tol
    r = pyoracle.run(b)
    assert rel_l2(samples, r["samples"]) <= 1e-4
    for got, want in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
        assert rel_l2(got, want) <= 1e-4


@pytest.mark.parametrize("order", ["sigma", "locality"])
def test_cg_row_orders_parity(workdir, order, monkeypatch):
    """The CG's free rows in either order (the edges' locality rank, the
    default when the problem carries one, or SELL-C-sigma) meet the fp64
    parity bars, and the edge-ordered scatter writes back every edge."""
    monkeypatch.setenv("TETFIT_ROW_ORDER", order)
    b = Model(configs.kuhn_cavity(str(workdir), n=6, method="gfdtd", n_steps=40, n_probes=4, block=2,
                                  name=f"ord_{order}")).build(1)
    assert b.edge_rank() is not None
    st, _ = check_run(b)
    assert st["cg_not_converged"] == 0 and st["cg_max_relres"] <= b.cg_tol


@pytest.mark.parametrize("method", ["dga", "gfdtd", "fem"])
def test_tet_locality_layout_bitwise(workdir, method, monkeypatch):
    """Fields and operators laid out in the edges' and faces' locality order
    (the default for tet problems that carry the ranks) give the fields of
    the problem's own numbering bit for bit: rows keep their entry order,
    only labels change. Probe samples are bitwise too; the Poynting sums
    (partials over source edges in another order) agree to rounding."""
    if method == "fem":
        from test_fem import fem_case
        _, b = fem_case(workdir, n=5, seed=3, name="lay_fem", n_steps=40, dt=1.5e-11)
    else:
        b = Model(configs.kuhn_cavity(str(workdir), n=6, method=method, n_steps=60, n_probes=4, block=2,
                                      name=f"lay_{method}")).build(1)
    out = {}
    for order in ("numbering", "locality"):
        monkeypatch.setenv("TETFIT_TET_ORDER", order)
        print(f"case: {method} layout {order}", flush=True)
        sim = Simulation(b, "fp64")
        sim.run()
        out[order] = (sim.probes()[1], sim.state(), sim.poynting())
        sim.close()
    (s0, st0, p0), (s1, st1, p1) = out["numbering"], out["locality"]
    assert np.array_equal(s0, s1)
    for a, e in zip(st0, st1):
        assert np.array_equal(a, e)
    np.testing.assert_allclose(p1, p0, rtol=1e-12, atol=1e-14 * max(np.abs(p0).max(), 1e-300))
