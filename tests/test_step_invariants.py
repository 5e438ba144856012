"""The CPU checker's time loops pinned by what the mathematics fixes:

* FIT: the TM110 mode of a PEC box is an exact eigenvector of the discrete
  curl-curl with eigenvalue kappa = 8 sin^2(pi/2N) / (eps0 mu0 h^2); the
  leapfrog then scales it by a_i with a_{i+1} = (2 - dt^2 kappa) a_i - a_{i-1}.
* DGA: W = 1/2 U^T H^-1 U + 1/2 F_-^T N^-1 F_+ is conserved exactly by the
  explicit leapfrog (U += dt H C^T F; F -= dt N C U).
* gfdtd: W = 1/2 U^T A U + 1/2 B_-^T N B_+ is conserved when the CG solves
  are tight.
A sign error, a wrong dt factor, a C that is not the transpose of C^T or a
wrong coefficient breaks each of these.
"""
import math

import numpy as np
import pytest

from tetfit_amd import Model, configs

import pyoracle

MU0 = float(np.float64(4 * np.longdouble(math.pi) * np.longdouble(1e-7)))
EPS0 = 8.854187817e-12


def tm110_mode(b, n, h):
    """U on z-edges: sin(pi x/a) sin(pi y/a) (any scale), 0 elsewhere."""
    t = b.fit_tables()
    pts, g = t["pts"], t["g_nodes"]
    p0, p1 = pts[g[:, 0]], pts[g[:, 1]]
    a = n * h
    zed = np.abs(p1[:, 2] - p0[:, 2]) > 0.5 * h
    u = np.where(zed, np.sin(math.pi * p0[:, 0] / a) * np.sin(math.pi * p0[:, 1] / a), 0.0)
    u[np.abs(u) < 1e-14] = 0.0
    return u


@pytest.fixture(scope="module")
def pec_box(tmp_path_factory):
    d = tmp_path_factory.mktemp("pecbox")
    n, h = 8, 0.125
    p = configs.fit_cavity(str(d), n=n, step=h, n_steps=300, n_probes=2, name="pec8",
                           pec_surfaces="1,2,3,4,5,6", source_surface=99)
    return Model(p).build(1), n, h


def test_fit_eigenmode_closed_form(pec_box):
    b, n, h = pec_box
    u0 = tm110_mode(b, n, h)
    r = pyoracle.run(b, U0=u0)
    kappa = 8 * math.sin(math.pi / (2 * n)) ** 2 / (EPS0 * MU0 * h * h)
    c = 2 - b.dt ** 2 * kappa
    a_prev, a = 1.0, 1.0 - b.dt ** 2 * kappa  # a_0 (F_-1/2 = 0), a_1
    for _ in range(b.n_steps - 2):
        a_prev, a = a, c * a - a_prev
    expect = a * u0
    scale = np.abs(u0).max()
    assert np.abs(r["U"] - expect).max() <= 1e-11 * scale
    # the mode actually oscillates over the run (not a trivially small a)
    omega = math.acos(c / 2) / b.dt
    assert omega * b.dt * b.n_steps > 2 * math.pi


def test_fit_eigenmode_wrong_sign_is_caught(pec_box):
    """A flipped sign in kappa's recurrence (e.g. +dt^2 kappa) would not match."""
    b, n, h = pec_box
    u0 = tm110_mode(b, n, h)
    r = pyoracle.run(b, U0=u0, n=50)
    kappa = 8 * math.sin(math.pi / (2 * n)) ** 2 / (EPS0 * MU0 * h * h)
    c = 2 + b.dt ** 2 * kappa
    a_prev, a = 1.0, 1.0 + b.dt ** 2 * kappa
    for _ in range(48):
        a_prev, a = a, c * a - a_prev
    assert np.abs(r["U"] - a * u0).max() > 1e-3 * np.abs(u0).max()


def dense(csr, nrows=None, ncols=None):
    nr = csr["nrows"] if nrows is None else nrows
    nc = csr["ncols"] if ncols is None else ncols
    M = np.zeros((nr, nc))
    rp, col = csr["rowptr"], csr["col"]
    v = csr["val"] if csr["val"] is not None else csr["sgn"].astype(float)
    for i in range(csr["nrows"]):
        for q in range(rp[i], rp[i + 1]):
            M[i, col[q]] += v[q]
    return M


@pytest.fixture(scope="module")
def kuhn_pec(tmp_path_factory):
    d = tmp_path_factory.mktemp("kuhnpec")
    out = {}
    for m in ("dga", "gfdtd"):
        p = configs.kuhn_cavity(str(d), n=3, method=m, n_steps=40, n_probes=1, block=2, seed=7,
                                pec_surfaces="1,2,3,4,5,6", source_surface=99)
        out[m] = Model(p).build(1)
    return out


def test_dga_energy_conserved(kuhn_pec):
    b = kuhn_pec["dga"]
    d = b.dga_ops()
    E, F = b.n_edges, b.n_faces
    rows = np.where(d["h_row"] >= 0)[0]
    H = dense(d["H"])[:, rows]              # free x free
    Hinv = np.linalg.inv(H)
    N = dense(d["N"])
    Ninv = np.linalg.inv(N)
    rng = np.random.default_rng(1)
    U = np.zeros(E)
    U[rows] = rng.standard_normal(len(rows))
    Fv = rng.standard_normal(F)
    B = np.zeros(F)
    W = []
    for i in range(1, 41):
        r = pyoracle.run(b, first=i, n=1, U0=U, F0=Fv, B0=B, output=False)
        Un, Fn = r["U"], r["F"]
        W.append(0.5 * Un[rows] @ Hinv @ Un[rows] + 0.5 * Fv @ Ninv @ Fn)
        U, Fv, B = Un, Fn, r["B"]
    W = np.array(W)
    assert np.abs(W - W[0]).max() <= 1e-11 * abs(W[0])


def test_dga_symmetric_operators(kuhn_pec):
    d = kuhn_pec["dga"].dga_ops()
    rows = np.where(d["h_row"] >= 0)[0]
    H = dense(d["H"])[:, rows]
    N = dense(d["N"])
    assert np.array_equal(H, H.T)
    assert np.abs(N - N.T).max() <= 1e-15 * np.abs(N).max()
    assert np.linalg.eigvalsh(H).min() > 0 and np.linalg.eigvalsh(N).min() > 0


def test_gfdtd_energy_conserved(kuhn_pec):
    b = kuhn_pec["gfdtd"]
    d = b.dga_ops()
    E, F = b.n_edges, b.n_faces
    free = d["free_edges"]
    A = dense(d["A"])
    N = dense(d["N"])
    rng = np.random.default_rng(3)
    U = np.zeros(E)
    U[free] = rng.standard_normal(len(free))
    B = 1e-9 * rng.standard_normal(F)
    W = []
    for i in range(1, 31):
        Bm = B
        r = pyoracle.run(b, first=i, n=1, U0=U, F0=None, B0=B, output=False, cg_tol=1e-15, cg_maxiter=500)
        Un, Bp = r["U"], r["B"]
        # energy of the state after the B update of step i (B_+) with U^{i-1}
        W.append(0.5 * U[free] @ A @ U[free] + 0.5 * Bm @ N @ Bp)
        U, B = Un, Bp
    W = np.array(W)
    assert np.abs(W - W[0]).max() <= 1e-9 * abs(W[0])
