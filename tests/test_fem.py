"""method fem (D:1580-1790, ConstructFEMaterialMatrices D:7687-8130): the
host's Whitney element mass E and SigMat against the exact integrals, the
system matrix A = E/dt^2 + SigMat/(2 dt) on the free edges, and the
checker's fem step against the energy that the second-order scheme
conserves exactly.

* E: the reference integrates w_j . eps w_k with a degree-3 rule (the four
  nodes at weight 1/40, the four face barycentres at 9/40), which is exact
  for the quadratic integrand, so E must equal the exact Whitney mass
  matrix: with w_ab = l_a grad l_b - l_b grad l_a and
  int l_a l_c dV = V (1 + delta_ac) / 20 every entry has a closed form
  (computed here in numpy from the barycentric gradients).
* Patch test: U = edge integrals of a uniform field E0 reproduces E0 inside
  every tet, so U.E U = sum eps |E0|^2 V.
* Scheme: E (U' - 2U + U_old)/dt^2 + Ct N C U = 0 (lossless, no sources)
  conserves 1/2 V.E V + 1/2 U'.(Ct N C) U with V = (U' - U)/dt.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from tetfit_amd import Model, configs

from test_codecasa_bruteforce import EPS0, bary_grads
from test_lossy_dga import materials


def fem_case(tmp_path, n=3, seed=6, sigma=0.0, sigma_frac=0.0, name="fem", pec="1,3,4,5,6", n_steps=40, dt=1e-11,
             eps_range=(1.0, 6.0)):
    p = configs.kuhn_cavity(str(tmp_path), n=n, method="fem", n_steps=n_steps, n_probes=3, block=1, seed=seed,
                            eps_range=eps_range, sigma=sigma, sigma_frac=sigma_frac, dt=dt, name=name,
                            pec_surfaces=pec)
    return p, Model(p).build(1)


def csr(m, n):
    return sp.csr_matrix((m["val"], m["col"], m["rowptr"]), shape=(m["nrows"], n))


def exact_whitney_mass(b, weight):
    """Dense exact int w_j . (weight w_k) dV, weight[t] a scalar per tet."""
    t = b.tet_tables()
    pts, vols, en = t["pts"], t["vols"], t["edge_nodes"]
    eid = {(int(x), int(y)): i for i, (x, y) in enumerate(en)}
    M = np.zeros((b.n_edges, b.n_edges))
    loc = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for ti, v in enumerate(vols):
        P = pts[v]
        V = abs(np.dot(P[1] - P[0], np.cross(P[2] - P[0], P[3] - P[0]))) / 6
        g = bary_grads(P)
        ll = lambda a, c: V * (1 + (a == c)) / 20
        es = [eid[(int(v[a]), int(v[c]))] for a, c in loc]
        for j, (a, b_) in enumerate(loc):
            for k, (c, d) in enumerate(loc):
                val = (ll(a, c) * g[b_] @ g[d] - ll(a, d) * g[b_] @ g[c] - ll(b_, c) * g[a] @ g[d]
                       + ll(b_, d) * g[a] @ g[c])
                M[es[j], es[k]] += weight[ti] * val
    return M


def test_fem_mass_is_the_exact_whitney_mass(tmp_path):
    p, b = fem_case(tmp_path, sigma=0.7, sigma_frac=0.5)
    eps, _, sig, _ = materials(p)
    lab = b.tet_tables()["vol_material"]
    ops = b.dga_ops()
    E = csr(ops["E"], b.n_edges).toarray()
    Ed = exact_whitney_mass(b, [EPS0 * eps[l] for l in lab])
    np.testing.assert_allclose(E, Ed, rtol=0, atol=1e-13 * np.abs(Ed).max())
    S = csr(ops["Sig"], b.n_edges).toarray()
    Sd = exact_whitney_mass(b, [sig[l] for l in lab])
    assert np.abs(Sd).max() > 0
    np.testing.assert_allclose(S, Sd, rtol=0, atol=1e-13 * np.abs(Sd).max())
    assert np.all(np.linalg.eigvalsh(E) > 0)


def test_fem_uniform_field_patch(tmp_path):
    p, b = fem_case(tmp_path, n=4, seed=2, name="fpatch")
    eps, _, _, _ = materials(p)
    t = b.tet_tables()
    E = csr(b.dga_ops()["E"], b.n_edges)
    E0 = np.array([0.3, -1.2, 0.7])
    en, pts = t["edge_nodes"], t["pts"]
    U = (pts[en[:, 1]] - pts[en[:, 0]]) @ E0
    P = pts[t["vols"]]
    V = np.abs(np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]))) / 6
    ref = sum(EPS0 * eps[l] * v for l, v in zip(t["vol_material"], V)) * (E0 @ E0)
    assert abs(U @ (E @ U) / ref - 1) < 1e-13


def test_fem_system_matrix(tmp_path):
    """A = E/dt^2 + SigMat/(2 dt) on the free edges and RHSmat1 on their
    Dirichlet columns (D:8086-8112), entry for entry."""
    p, b = fem_case(tmp_path, sigma=0.7, sigma_frac=0.5, name="fsys")
    ops = b.dga_ops()
    E = csr(ops["E"], b.n_edges).toarray()
    S = csr(ops["Sig"], b.n_edges).toarray()
    dt = b.dt
    full = E * (1 / (dt * dt)) + (0.5 / dt) * S
    free = ops["free_edges"]
    A = csr(ops["A"], len(free)).toarray()
    np.testing.assert_allclose(A, full[np.ix_(free, free)], rtol=1e-15, atol=0)
    dirich = np.setdiff1d(np.arange(b.n_edges), free)
    R1 = csr(ops["RHS1"], b.n_edges).toarray()
    np.testing.assert_allclose(R1[np.ix_(free, dirich)], full[np.ix_(free, dirich)], rtol=1e-15, atol=0)
    assert np.all(R1[:, free] == 0)


def test_fem_checker_conserves_the_scheme_energy(tmp_path):
    import pyoracle

    p, b = fem_case(tmp_path, n=3, seed=8, name="fen", pec="1,2,3,4,5,6", n_steps=60, dt=2e-11)
    assert len(b.sources()[1]) == 0
    ops = b.dga_ops()
    E = csr(ops["E"], b.n_edges)
    C = sp.csr_matrix((ops["C"]["sgn"].astype(float), ops["C"]["col"], ops["C"]["rowptr"]),
                      shape=(b.n_faces, b.n_edges))
    N = csr(ops["N"], b.n_faces)
    K = C.T @ N @ C
    rng = np.random.default_rng(3)
    U0 = rng.standard_normal(b.n_edges) * (ops["edge_class"] != 4)
    dt = b.dt

    def U_after(k):
        if k == 0:
            return U0
        return pyoracle.run(b, n=k, U0=U0, output=False, cg_tol=1e-15, cg_maxiter=2000)["U"]

    def energy(k):
        u, u1 = U_after(k), U_after(k + 1)
        v = (u1 - u) / dt
        return 0.5 * v @ (E @ v) + 0.5 * u1 @ (K @ u)

    e0 = energy(0)
    for k in (7, 40):
        assert abs(energy(k) / e0 - 1) < 1e-9
