"""Learned CG preconditioner (gfdtd): the MLP's fp64 checker pinned by
finite differences, the device MLP forward/backward against it, and the
trained preconditioner's gain over Jacobi-PCG with results unchanged at the
CG tolerance."""
import numpy as np
import pytest

from tetfit_amd import Model, configs
from tetfit_amd._ffi import MLP_PARAMS

import pyoracle

I, H = 16, 64
W1, B1 = 0, H * I
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
BLOCKS = {"W1": slice(W1, B1), "b1": slice(B1, W2), "W2": slice(W2, B2), "b2": slice(B2, W3),
          "w3": slice(W3, B3), "b3": slice(B3, B3 + 1)}


def bf16(a):
    """Round float32 values to the nearest bf16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def random_params(seed, out_scale=0.3):
    rng = np.random.default_rng(seed)
    P = np.zeros(MLP_PARAMS, np.float32)
    P[BLOCKS["W1"]] = rng.uniform(-1, 1, H * I) * np.sqrt(6 / I)
    P[BLOCKS["b1"]] = rng.uniform(-0.2, 0.2, H)
    P[BLOCKS["W2"]] = rng.uniform(-1, 1, H * H) * np.sqrt(6 / H)
    P[BLOCKS["b2"]] = rng.uniform(-0.2, 0.2, H)
    P[BLOCKS["w3"]] = rng.uniform(-out_scale, out_scale, H)
    P[B3] = 0.1
    return P


def preacts(P, x):
    """Hidden pre-activations of one row (used only to pick rows away from
    ReLU kinks for the finite-difference check)."""
    P = P.astype(np.float64)
    z1 = P[BLOCKS["W1"]].reshape(H, I) @ x + P[BLOCKS["b1"]]
    z2 = P[BLOCKS["W2"]].reshape(H, H) @ np.maximum(z1, 0) + P[BLOCKS["b2"]]
    return z1, z2


def test_oracle_mlp_matches_finite_differences():
    """Within one ReLU region y is linear in each single parameter, so
    central differences of L = sum gy*y are exact up to rounding: every
    entry of the hand-written backward pass is pinned (a transposed W2, a
    missing ReLU mask or a dropped bias term all fail)."""
    P = random_params(1)
    rng = np.random.default_rng(2)
    rows = []
    while len(rows) < 12:
        x = rng.standard_normal(I).astype(np.float32)
        z1, z2 = preacts(P, x.astype(np.float64))
        if min(np.abs(z1).min(), np.abs(z2).min()) > 0.05:
            rows.append(x)
    X = np.array(rows, np.float32)
    gy = rng.standard_normal(len(X))
    _, g = pyoracle.mlp(P, X, gy)
    h = np.float32(2.0 ** -12)
    fd = np.zeros(MLP_PARAMS)
    for k in range(MLP_PARAMS):
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        step = float(Pp[k]) - float(Pm[k])
        yp, _ = pyoracle.mlp(Pp, X)
        ym, _ = pyoracle.mlp(Pm, X)
        fd[k] = (gy @ yp - gy @ ym) / step
    for name, sl in BLOCKS.items():
        err = np.abs(g[sl] - fd[sl]).max() / max(np.abs(fd).max(), 1e-30)
        assert err <= 1e-9, (name, err)
    # forward: y of the zero-output network is b3 exactly
    P0 = P.copy()
    P0[BLOCKS["w3"]] = 0
    y0, _ = pyoracle.mlp(P0, X)
    assert np.all(y0 == np.float32(0.1).astype(np.float64))


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def bf16_model(P, X, gy):
    """Where the device kernels round to bf16 (MFMA operands): the layer-1
    output h1, the back-propagated g2 and g1; everything else in fp64 here
    (fp32 on the device). Used to bound the device's deviation from the fp64
    checker to what those roundings explain."""
    b = lambda a: bf16(np.asarray(a, np.float32)).astype(np.float64)
    W1m = P[BLOCKS["W1"]].reshape(H, I).astype(np.float64)
    W2m = P[BLOCKS["W2"]].reshape(H, H).astype(np.float64)
    z1 = X.astype(np.float64) @ W1m.T + P[BLOCKS["b1"]]
    h1 = b(np.maximum(z1, 0))
    z2 = h1 @ W2m.T + P[BLOCKS["b2"]]
    y = np.maximum(z2, 0) @ P[BLOCKS["w3"]].astype(np.float64) + P[B3]
    g = gy.astype(np.float64)[:, None]
    g2 = b(np.where(z2 > 0, g * P[BLOCKS["w3"]], 0))
    g1 = np.where(z1 > 0, g2 @ W2m, 0)
    grad = np.zeros(MLP_PARAMS)
    grad[BLOCKS["W1"]] = (b(g1).T @ X).ravel()
    grad[BLOCKS["b1"]] = g1.sum(0)
    grad[BLOCKS["W2"]] = (g2.T @ h1).ravel()
    grad[BLOCKS["b2"]] = np.where(z2 > 0, g * P[BLOCKS["w3"]], 0).sum(0)
    grad[BLOCKS["w3"]] = (np.maximum(z2, 0) * g).sum(0)
    grad[B3] = g.sum()
    return y, grad


@pytest.mark.gpu
def test_device_mlp_grad_matches_oracle():
    """bf16 MFMA forward/backward (fp32 accumulate) vs the fp64 checker on
    the same bf16-representable weights and features. The device rounds h1,
    g2 and g1 to bf16 for the MFMA operands; over 1000 pairs with
    random-sign gradients that alone gives ~3.5% relative L2 on the weight
    gradients (reproduced by bf16_model), so the bars are: outputs 1e-2 and
    gradient blocks 6e-2 against the fp64 checker, and 1e-4 against the
    bf16 rounding model (which catches a wrong operand layout or tile
    index). n = 1000 is not a multiple of the 16-pair tile (padded tail)."""
    from tetfit_amd.device import mlp_grad
    print("case: tf_mlp_grad n=1000", flush=True)
    P = random_params(3)
    P[BLOCKS["W1"]] = bf16(P[BLOCKS["W1"]])
    P[BLOCKS["W2"]] = bf16(P[BLOCKS["W2"]])
    rng = np.random.default_rng(4)
    X = bf16(rng.standard_normal((1000, I)).astype(np.float32))
    gy = rng.standard_normal(1000).astype(np.float32)
    y, g = mlp_grad(P, X, gy)
    yr, gr = pyoracle.mlp(P, X, gy.astype(np.float64))
    ym, gm = bf16_model(P, X, gy)
    assert rel(y, yr) <= 1e-2 and rel(y, ym) <= 1e-4
    for name, sl in BLOCKS.items():
        assert rel(g[sl], gr[sl]) <= 6e-2, (name, rel(g[sl], gr[sl]))
        assert rel(g[sl], gm[sl]) <= 1e-4, (name, rel(g[sl], gm[sl]))
    # the device result is not the oracle's with a transposed W2
    Pt = P.copy()
    Pt[BLOCKS["W2"]] = P[BLOCKS["W2"]].reshape(H, H).T.ravel()
    yt, _ = pyoracle.mlp(Pt, X)
    assert rel(y, yt) > 0.1


@pytest.fixture(scope="module")
def contrast_cavity(tmp_path_factory):
    """Kuhn 10^3 gfdtd cavity with log-uniform eps contrast [1, 80] per 2^3
    block (the high-contrast case the learned preconditioner targets)."""
    d = tmp_path_factory.mktemp("pc")
    p = configs.kuhn_cavity(str(d), n=10, method="gfdtd", n_steps=70, n_probes=4, block=2, seed=5,
                            eps_range=(1.0, 80.0), log_eps=True, tol=1e-10)
    return Model(p).build(1)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bar", [("fp64", 1e-7), ("fp32", 1e-4)])
def test_learned_preconditioner_gain(contrast_cavity, prec, bar):
    """The first 30 solves run Jacobi-PCG and supply the training pairs; the
    network is trained on the device and the last 40 solves use it. Checked:
    training lowers its loss, no solve falls back, every solve meets the
    reference's stopping test, the learned solves take at most 0.8x the
    Jacobi-PCG iterations of the same steps, and the fields agree with the
    fp64 checker (Jacobi-PCG, the reference algorithm) at the CG tolerance
    1e-10: relative L2 <= 1e-7 for fp64 fields, the fp32 bar for fp32."""
    from tetfit_amd.device import Simulation
    b = contrast_cavity
    print(f"case: learned {prec} E={b.n_edges} steps={b.n_steps}", flush=True)
    sim = Simulation(b, prec)
    s1 = sim.run(n=30, precond="learned", learn_steps=30, train_iters=80)
    s2 = sim.run(n=b.n_steps - 30, precond="learned")
    _, samples = sim.probes()
    U, F, B = sim.state()
    sim.close()
    jac = Simulation(b, prec)
    j1 = jac.run(n=30)
    j2 = jac.run(n=b.n_steps - 30)
    jac.close()
    print(s1, s2, j1, j2, flush=True)
    assert s1["pc_state"] == 2 and s2["pc_state"] == 2
    assert s2["pc_pairs"] >= 8
    assert s2["pc_loss"] < s2["pc_loss0"]
    assert s2["pc_fallbacks"] == 0 and s2["pc_learned_solves"] == b.n_steps - 30
    for s in (s1, s2):
        assert s["cg_not_converged"] == 0 and s["cg_max_relres"] <= b.cg_tol
    assert s2["cg_iters_total"] <= 0.8 * j2["cg_iters_total"]
    r = pyoracle.run(b)
    assert rel(samples, r["samples"]) <= bar
    for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
        assert rel(got, ref) <= bar


@pytest.mark.gpu
@pytest.mark.parametrize("layout", ["col32", "win", "wide"])
def test_learned_column_layouts(contrast_cavity, layout, monkeypatch):
    """The same learned run with A (and G beside it) stored with int32
    column indices (the unpacked G path of cg_pg_kernel and the int32 path of
    the SELL kernels), with 16-bit codes over four column windows per slice
    (packed G decoded through the windows), or with one-base 16-bit columns
    in the wide-mesh kernels (non-temporal operator loads)."""
    from tetfit_amd.device import Simulation
    env = {"col32": "TETFIT_SELL_COL32", "win": "TETFIT_SELL_WIN", "wide": "TETFIT_SELL_WIDE"}[layout]
    monkeypatch.setenv(env, "1")
    b = contrast_cavity
    print(f"case: learned fp32 {layout} columns E={b.n_edges} steps={b.n_steps}", flush=True)
    sim = Simulation(b, "fp32")
    info = sim.sell_info("a")
    assert (info["col_bytes"], info["col_bases"]) == {"col32": (4, 0), "win": (2, 4), "wide": (2, 1)}[layout]
    s1 = sim.run(n=30, precond="learned", learn_steps=30, train_iters=80)
    s2 = sim.run(n=b.n_steps - 30, precond="learned")
    _, samples = sim.probes()
    U, F, B = sim.state()
    sim.close()
    assert s2["pc_state"] == 2 and s2["pc_fallbacks"] == 0
    assert s2["cg_not_converged"] == 0 and s2["cg_max_relres"] <= b.cg_tol
    assert s2["cg_iters_total"] <= 0.8 * s1["cg_iters_total"] * (b.n_steps - 30) / 30
    r = pyoracle.run(b)
    assert rel(samples, r["samples"]) <= 1e-4
    for got, ref in ((U, r["U"]), (F, r["F"]), (B, r["B"])):
        assert rel(got, ref) <= 1e-4


@pytest.mark.gpu
def test_training_is_what_gains(tmp_path):
    """On the 48^3 Kuhn workload the operator the network starts from does
    not beat Jacobi: with training switched off (learning rate 1e-12, so the
    zero-initialised output layer keeps y = 0 and G is the first-order
    Neumann inverse D^-1/2 (2I - Â) D^-1/2, which loses definiteness when the
    largest eigenvalue of Â exceeds 2) its solves take more iterations than
    Jacobi-PCG, or miss the tolerance and are redone, and plain CG switches it
    off. The trained network's G meets the tolerance in every solve with at
    most 0.6x the Jacobi-PCG iterations. (On the small Kuhn 10^3 mesh of the
    gain test the untrained base does about as well as the trained one.)"""
    from tetfit_amd.device import Simulation
    p = configs.kuhn_cavity(str(tmp_path), n=48, method="gfdtd", n_steps=120, n_probes=2, block=8, seed=2,
                            dt=4.777e-11)
    b = Model(p).build(1)
    res = {}
    for name, lr in (("trained", 3e-3), ("untrained", 1e-12)):
        print(f"case: {name} preconditioner E={b.n_edges}", flush=True)
        sim = Simulation(b, "fp32", output=False)
        res[name + "_collect"] = sim.run(n=60, precond="learned", learn_steps=60, train_iters=150, lr=lr)
        res[name] = sim.run(n=20, precond="learned")
        sim.close()
        print(name, res[name], flush=True)
    t, u = res["trained"], res["untrained"]
    jac = res["trained_collect"]["pc_jacobi_mean_iters"]
    assert t["pc_state"] == 2 and t["pc_fallbacks"] == 0 and t["cg_not_converged"] == 0
    assert t["cg_iters_total"] <= 0.6 * jac * 20
    assert t["pc_loss"] < 0.9 * t["pc_loss0"]
    assert u["pc_loss"] >= 0.99 * u["pc_loss0"]  # nothing learned
    assert u["pc_state"] == 3  # switched off
    assert u["pc_fallbacks"] >= 1 or u["pc_learned_iters"] > jac * u["pc_learned_solves"]


@pytest.mark.gpu
def test_learned_quiet_steps(tmp_path):
    """Steps whose right-hand side is exactly zero (a gaussian carrier whose
    delay keeps it at 0.0 for the whole run): every solve starts from a zero
    residual and stops at once, as the reference's 0/0 > tol test does
    (ConjugateGradientSolver.hpp:72). On the learned path such a solve counts
    as converged: no Jacobi redo, no missed tolerance, fields stay zero."""
    from tetfit_amd.device import Simulation
    p = configs.kuhn_cavity(str(tmp_path), n=5, method="gfdtd", n_steps=40, n_probes=2, block=2, dt=2e-11,
                            name="quiet")
    s = open(p).read().replace("SET carrier sin", "SET carrier gaussian\n    SET width 1e-9\n    SET delay 1e-6")
    open(p, "w").write(s)
    b = Model(p).build(1)
    print(f"case: learned quiet steps E={b.n_edges}", flush=True)
    sim = Simulation(b, "fp32")
    s1 = sim.run(n=10, precond="learned", learn_steps=10, train_iters=20)
    s2 = sim.run(n=b.n_steps - 10, precond="learned")
    U, F, B = sim.state()
    sim.close()
    print(s1, s2, flush=True)
    assert s2["pc_fallbacks"] == 0
    assert s1["cg_not_converged"] == 0 and s2["cg_not_converged"] == 0
    assert s1["cg_iters_total"] == 0 and s2["cg_iters_total"] == 0
    assert not np.any(U) and not np.any(F) and not np.any(B)
