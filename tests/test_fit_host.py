"""Cartesian mesher and FIT operators (host side), pinned by integer
invariants, closed forms and an independent restatement of the reference's
eight cube cases (src/DGA/Discretization.hpp:4643-4934)."""
import math

import numpy as np
import pytest

from tetfit_amd import Model, configs

MU0 = 4 * math.pi * 1e-7
EPS0 = 8.854187817e-12


def cube_cases(Nx, Ny, Nz):
    """Literal restatement of the eight neighbour cases of ReadStructuredMesh:
    returns D (6 faces/cube), E_cluster (12 edges/cube), P_cluster (8 nodes)."""
    D, Ec, Pc = [], [], []
    nf = ne = np_ = 0
    nv = 0
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                bottom = nv + 1 - Nx * Ny if k > 0 else 0
                left = nv + 1 - Nx if j > 0 else 0
                back = nv if i > 0 else 0
                bt = (1 if bottom else 0) + (2 if left else 0) + (4 if back else 0)
                b, l, k_ = bottom - 1, left - 1, back - 1
                if bt == 0:
                    d = [nf + q for q in range(6)]; nf += 6
                    e = [ne + q for q in range(12)]; ne += 12
                    p = [np_ + q for q in range(8)]; np_ += 8
                elif bt == 1:
                    d = [D[b][5]] + [nf + q for q in range(5)]; nf += 5
                    e = [Ec[b][8], Ec[b][9], ne, Ec[b][10], ne + 1, Ec[b][11], ne + 2, ne + 3, ne + 4, ne + 5, ne + 6, ne + 7]; ne += 8
                    p = [Pc[b][4], Pc[b][5], Pc[b][6], Pc[b][7], np_, np_ + 1, np_ + 2, np_ + 3]; np_ += 4
                elif bt == 2:
                    d = [nf, D[l][4], nf + 1, nf + 2, nf + 3, nf + 4]; nf += 5
                    e = [Ec[l][5], ne, Ec[l][6], ne + 1, Ec[l][7], ne + 2, ne + 3, ne + 4, Ec[l][11], ne + 5, ne + 6, ne + 7]; ne += 8
                    p = [Pc[l][2], Pc[l][3], np_, np_ + 1, Pc[l][6], Pc[l][7], np_ + 2, np_ + 3]; np_ += 4
                elif bt == 3:
                    d = [D[b][5], D[l][4], nf, nf + 1, nf + 2, nf + 3]; nf += 4
                    e = [Ec[b][8], Ec[b][9], Ec[l][6], Ec[b][10], Ec[l][7], Ec[b][11], ne, ne + 1, Ec[l][11], ne + 2, ne + 3, ne + 4]; ne += 5
                    p = [Pc[b][4], Pc[b][5], Pc[b][6], Pc[b][7], Pc[l][6], Pc[l][7], np_, np_ + 1]; np_ += 2
                elif bt == 4:
                    d = [nf, nf + 1, D[k_][3], nf + 2, nf + 3, nf + 4]; nf += 5
                    e = [ne, Ec[k_][3], Ec[k_][4], ne + 1, ne + 2, ne + 3, Ec[k_][7], ne + 4, ne + 5, Ec[k_][10], ne + 6, ne + 7]; ne += 8
                    p = [Pc[k_][1], np_, Pc[k_][3], np_ + 1, Pc[k_][5], np_ + 2, Pc[k_][7], np_ + 3]; np_ += 4
                elif bt == 5:
                    d = [D[b][5], nf, D[k_][3], nf + 1, nf + 2, nf + 3]; nf += 4
                    e = [Ec[b][8], Ec[b][9], Ec[k_][4], Ec[b][10], ne, Ec[b][11], Ec[k_][7], ne + 1, ne + 2, Ec[k_][10], ne + 3, ne + 4]; ne += 5
                    p = [Pc[b][4], Pc[b][5], Pc[b][6], Pc[b][7], Pc[k_][5], np_, Pc[k_][7], np_ + 1]; np_ += 2
                elif bt == 6:
                    d = [nf, D[l][4], D[k_][3], nf + 1, nf + 2, nf + 3]; nf += 4
                    e = [Ec[l][5], Ec[k_][3], Ec[l][6], ne, Ec[l][7], ne + 1, Ec[k_][7], ne + 2, Ec[l][11], Ec[k_][10], ne + 3, ne + 4]; ne += 5
                    p = [Pc[l][2], Pc[l][3], Pc[k_][3], np_, Pc[l][6], Pc[l][7], Pc[k_][7], np_ + 1]; np_ += 2
                else:
                    d = [D[b][5], D[l][4], D[k_][3], nf, nf + 1, nf + 2]; nf += 3
                    e = [Ec[b][8], Ec[b][9], Ec[l][6], Ec[b][10], Ec[l][7], Ec[b][11], Ec[k_][7], ne, Ec[l][11], Ec[k_][10], ne + 1, ne + 2]; ne += 3
                    p = [Pc[b][4], Pc[b][5], Pc[b][6], Pc[b][7], Pc[l][6], Pc[l][7], Pc[k_][7], np_]; np_ += 1
                D.append(d); Ec.append(e); Pc.append(p)
                nv += 1
    return np.array(D), np.array(Ec), np.array(Pc)


@pytest.fixture(scope="module")
def fit4(tmp_path_factory):
    d = tmp_path_factory.mktemp("fit4")
    b = Model(configs.fit_cavity(str(d), n=4, step=0.25, n_steps=5, n_probes=2, name="fit4")).build(1)
    return b


def test_counts_closed_form(fit4):
    n = 4
    assert fit4.n_edges == 3 * n * (n + 1) ** 2
    assert fit4.n_faces == 3 * n * n * (n + 1)
    assert fit4.n_vols == n ** 3


def test_numbering_matches_eight_cases(fit4):
    t = fit4.fit_tables()
    D, Ec, _ = cube_cases(4, 4, 4)
    assert np.array_equal(t["d_faces"], D)
    assert np.array_equal(t["e_cluster"], Ec)


def test_numbering_non_cubic():
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = configs.fit_cavity(d, n=3, step=0.2, n_steps=2, n_probes=1, name="f3")
        # rewrite the grid as 3 x 2 x 4 cells
        import re
        s = open(p).read()
        s = re.sub(r"SET ygrid \{[^}]*\}", "SET ygrid {0,0.2,0.45}", s)
        s = re.sub(r"SET zgrid \{[^}]*\}", "SET zgrid {0,0.2,0.85}", s)
        open(p, "w").write(s)
        b = Model(p).build(1)
        t = b.fit_tables()
        D, Ec, _ = cube_cases(3, 2, 4)
        assert np.array_equal(t["d_faces"], D) and np.array_equal(t["e_cluster"], Ec)


def loop_curl_matrices(b):
    """The signed incidences the fdtd loop actually applies:
    H phase  C(f, c_edge[f][k]) = curl[f] * (+,-,+,-)[k]          (D:1910)
    E phase  C^T(e, ct_face[e][k]) = dual_curl[e] * (+,-,+,-)[k]  (D:1874), common edges."""
    f = b.fit_ops()
    E, F = b.n_edges, b.n_faces
    alt = np.array([1, -1, 1, -1])
    C = np.zeros((F, E), dtype=np.int64)
    for fi in f["tbc_faces"]:
        for k in range(4):
            C[fi, f["c_edge"][fi, k]] += f["curl"][fi] * alt[k]
    Ct = np.zeros((E, F), dtype=np.int64)
    for e in f["common_edges"]:
        faces = f["ct_face"][f["ct_ptr"][e]:f["ct_ptr"][e + 1]]
        assert len(faces) == 4
        for k in range(4):
            Ct[e, faces[k]] += f["dual_curl"][e] * alt[k]
    return C, Ct, f


def test_curl_of_gradient_vanishes(fit4):
    """C G = 0 with G the node-to-edge gradient (exact integer identity)."""
    C, _, _ = loop_curl_matrices(fit4)
    g = fit4.fit_tables()["g_nodes"]
    G = np.zeros((fit4.n_edges, int(g.max()) + 1), dtype=np.int64)
    for e, (a, c) in enumerate(g):
        G[e, a] -= 1
        G[e, c] += 1
    assert not (C @ G).any()


def test_hardcoded_alternating_signs_are_the_transpose(fit4):
    """The E phase ignores the stored signs and uses dual_curl*(+,-,+,-) over
    Ct_vec's push order; on every common edge that must equal the transpose
    of the H-phase curl (a wrong push order or sign flips the physics)."""
    C, Ct, f = loop_curl_matrices(fit4)
    common = f["common_edges"]
    assert np.array_equal(Ct[common], C.T[common])


def test_boundary_labels(fit4):
    t = fit4.fit_tables()
    pts, g, bf = t["pts"], t["g_nodes"], t["boundary_face"]
    f = fit4.fit_ops()
    L = 1.0
    for fi in range(fit4.n_faces):
        nodes = g[f["c_edge"][fi]].ravel()
        c = pts[nodes].mean(axis=0)
        lab = 0
        for ax, lo, hi in ((2, 1, 6), (0, 2, 4), (1, 5, 3)):  # z-min 1 z-max 6, x-min 2 x-max 4, y-min 5 y-max 3
            if np.all(np.abs(pts[nodes][:, ax]) < 1e-12):
                lab = lo
            elif np.all(np.abs(pts[nodes][:, ax] - L) < 1e-12):
                lab = hi
        assert bf[fi] == lab


def test_coefficients_closed_form(fit4):
    """Uniform vacuum, cell step h: M_q = eps0*A_dual/h with A_dual = h^2,
    h^2/2, h^2/4 for 4, 2, 1 adjacent cells; M_nu = 1/(mu0*h) interior,
    1/(2*mu0*h) on the boundary; M_h = 1/M_q, M_mu = 1/M_nu."""
    f = fit4.fit_ops()
    h = 0.25
    nfaces_per_edge = np.diff(f["ct_ptr"])
    # cells around an edge: 4 faces -> 4 cells, 3 -> 2, 2 -> 1
    ncell = np.select([nfaces_per_edge == 4, nfaces_per_edge == 3, nfaces_per_edge == 2], [4, 2, 1])
    np.testing.assert_allclose(f["M_q"], EPS0 * ncell * h * h / 4 / h, rtol=1e-14)
    np.testing.assert_allclose(f["M_h"] * f["M_q"], 1.0, rtol=1e-14)
    t = fit4.fit_tables()
    interior = t["boundary_face"] == 0
    np.testing.assert_allclose(f["M_nu"][interior], 1 / (MU0 * h), rtol=1e-14)
    np.testing.assert_allclose(f["M_nu"][~interior], 1 / (2 * MU0 * h), rtol=1e-14)


def test_edge_classes(fit4):
    f = fit4.fit_ops()
    E = fit4.n_edges
    common, unc, bc = set(f["common_edges"]), set(f["uncommon_edges"]), set(f["bc_edges"])
    assert not (common & unc) and not (common & bc) and not (unc & bc)
    assert len(common) + len(unc) + len(bc) == E
    n = 4
    # h-source surface x-min: edges in the wall minus its rim (rim edges touch pec walls)
    assert len(unc) == 2 * n * (n + 1) - 4 * n
    assert len(common) == 3 * n * (n - 1) ** 2


def test_source_weights_closed_form(fit4):
    """h source, direction y, mode {cos,cos,sin}, wavevector {0,0,0.5}:
    weight = sin(pi z_mid) * (bnd dual edge vector)_y; the dual vector of a
    z-edge in the x-min wall is +-h along y, y-edges get weight 0."""
    _, terms = fit4.sources()
    desc, edges = fit4.source_desc()
    h = 0.25
    for e in edges:
        w = terms["weight"][terms["edge"] == e["edge"]][0]
        expect = math.sin(math.pi * e["bary"][2]) * e["vec"][1]
        assert abs(w - expect) <= 1e-15
        assert abs(abs(e["vec"][1]) - h) < 1e-15 or abs(e["vec"][1]) < 1e-15


def cell_uniform(seed, index):
    """The epsilon_random draw: SplitMix64 finalisation of seed*golden+index."""
    m = (1 << 64) - 1
    z = (seed * 0x9E3779B97F4A7C15 + index + 0x632BE59BD9B4E019) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return (z >> 11) * (1.0 / 9007199254740992.0)


def test_random_media_and_lossy_coefficients(tmp_path):
    """Per-cell eps (epsilon_random) and the lossy slab: for every edge,
    eps_bar = sum over its cells of eps0*eps_r(cell)*h^2/4 and sigma_bar the
    same over the slab cells; lossless edges M_q = eps_bar/h, lossy edges
    M_h = h/(eps_bar + dt sigma_bar/2), M_q = (eps_bar - dt sigma_bar/2)/h
    (D:5581-5586)."""
    from tetfit_amd import configs as cf
    n, seed, lo, hi, sig = 6, 7, 1.0, 4.0, 0.05
    b = Model(cf.fit_random_media(str(tmp_path), n=n, n_steps=4, n_probes=1, eps_range=(lo, hi), seed=seed,
                                  sigma=sig, slab=(0.35, 0.65))).build(1)
    f, t = b.fit_ops(), b.fit_tables()
    h = 1.0 / n
    ebar = np.zeros(b.n_edges)
    sbar = np.zeros(b.n_edges)
    for c, edges in enumerate(t["e_cluster"]):
        i = c % n
        er = lo + (hi - lo) * cell_uniform(seed, c)
        xc = (i + 0.5) * h
        for e in edges:
            ebar[e] += EPS0 * er * h * h / 4
            if 0.35 <= xc < 0.65:
                sbar[e] += sig * h * h / 4
    lossy = sbar > 0
    assert lossy.any() and (~lossy).any()
    np.testing.assert_allclose(f["M_q"][~lossy], ebar[~lossy] / h, rtol=1e-13)
    np.testing.assert_allclose(f["M_h"][lossy], h / (ebar[lossy] + 0.5 * b.dt * sbar[lossy]), rtol=1e-13)
    np.testing.assert_allclose(f["M_q"][lossy], (ebar[lossy] - 0.5 * b.dt * sbar[lossy]) / h, rtol=1e-13)
    # the draws really differ per cell
    assert np.unique(np.round(f["M_q"][~lossy] * h / (EPS0 * h * h), 12)).size > 10


def test_grid_positions(tmp_path):
    """edge_sid / face_sid place every edge and face at its grid position
    (the layout documented in tetfit_problem.h), computed here from the node
    coordinates: a permutation of the edges and of the faces, on a
    non-cubic grid."""
    p = configs.fit_cavity(str(tmp_path), n=5, step=0.1, n_steps=4, n_probes=1, name="gp")
    s = open(p).read().replace("ygrid {0,0.1,0.525}", "ygrid {0,0.1,0.725}").replace("zgrid {0,0.1,0.525}",
                                                                                    "zgrid {0,0.1,0.325}")
    open(p, "w").write(s)
    b = Model(p).build(1)
    o, t = b.fit_ops(), b.fit_tables()
    nx, ny, nz = o["cells"]
    assert (nx, ny, nz) == (5, 7, 3)
    pts, g = t["pts"], t["g_nodes"]
    h = t["L"]
    lo = pts.min(axis=0)
    a0, a1 = pts[g[:, 0]], pts[g[:, 1]]
    axis = np.argmax(np.abs(a1 - a0), axis=1)
    ijk = np.rint((np.minimum(a0, a1) - lo) / h).astype(np.int64)
    i, j, k = ijk.T
    nEx, nEy = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    want = np.where(axis == 0, (k * (ny + 1) + j) * nx + i,
                    np.where(axis == 1, nEx + (k * ny + j) * (nx + 1) + i,
                             nEx + nEy + (k * (ny + 1) + j) * (nx + 1) + i))
    np.testing.assert_array_equal(o["edge_sid"], want)
    assert sorted(o["edge_sid"]) == list(range(b.n_edges))
    # faces: the lowest corner of the face's four edges and its normal axis
    ce = o["c_edge"]
    fl = np.minimum.reduce([np.minimum(pts[g[ce[:, q], 0]], pts[g[ce[:, q], 1]]) for q in range(4)])
    fh = np.maximum.reduce([np.maximum(pts[g[ce[:, q], 0]], pts[g[ce[:, q], 1]]) for q in range(4)])
    normal = np.argmin(fh - fl, axis=1)
    i, j, k = np.rint((fl - lo) / h).astype(np.int64).T
    nFx, nFy = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    want = np.where(normal == 0, (k * ny + j) * (nx + 1) + i,
                    np.where(normal == 1, nFx + (k * (ny + 1) + j) * nx + i, nFx + nFy + (k * ny + j) * nx + i))
    np.testing.assert_array_equal(o["face_sid"], want)
    assert sorted(o["face_sid"]) == list(range(b.n_faces))
