"""Pins the checker's probe samplers by what the mathematics fixes: a uniform
field, written into the edge voltages and face fluxes as the line and
surface integrals the discrete unknowns stand for, must come back unchanged
at every probe.

* FIT (GetFitElectricField D:3555-3568, GetFitMagneticField D:3570-3588):
  the cell average of U (dual areas over the 12 edges of the cube) returns
  E0, the average of B over the 6 faces returns mu^-1 B0.
* Tetrahedra (GetWhitneyElectricField D:3768-3875): Whitney edge forms
  reproduce a constant field exactly anywhere inside the tet.
  GetMagneticField (D:3877-3897): the four-face formula returns mu^-1 B0.

Every geometric quantity here (edge vectors from the node coordinates and
the edges' end nodes, face area vectors from the face nodes, the axis of
each FIT entity) is derived in numpy from the mesh tables, independently of
the checker's own per-cube tables; distinct components of E0 and B0, a
non-cubic FIT cell and mu_r != 1 make a wrong axis, a swapped face, a
missing mu^-1 or a wrong volume fail."""
import re

import numpy as np

import pyoracle
from tetfit_amd import Model, configs

MU0 = 4e-7 * np.pi
E0 = np.array([0.7, -1.3, 2.1])
B0 = np.array([-0.4, 0.9, 1.7])


def test_fit_cell_average_reproduces_uniform_field(workdir):
    p = configs.fit_cavity(str(workdir), n=6, step=0.1, n_steps=1, n_probes=5, name="fitprobe")
    s = open(p).read()
    s = re.sub(r"ygrid \{[^}]*\}", "ygrid {0,0.07,0.4300001}", s)
    s = re.sub(r"zgrid \{[^}]*\}", "zgrid {0,0.09,0.5400001}", s)
    s = s.replace("SET mu 1", "SET mu 2.5")
    pts = "".join("    SET probe {%r,%r,%r}\n" % q for q in
                  [(0.03, 0.02, 0.05), (0.31, 0.2, 0.4), (0.55, 0.41, 0.53), (0.12, 0.35, 0.27), (0.49, 0.08, 0.3)])
    s = re.sub(r"(    SET probe \{[^}]*\}\n)+", pts, s)
    open(p, "w").write(s)
    b = Model(p).build(1)
    t = b.fit_tables()
    f = b.fit_ops()
    P = t["pts"]
    g = t["g_nodes"]
    ev = P[g[:, 1]] - P[g[:, 0]]
    U = ev @ E0
    # face area vectors: the cross product of two of the face's edges,
    # oriented along +axis
    ce = f["c_edge"]
    a, c = ev[ce[:, 0]], ev[ce[:, 1]]
    n = np.cross(a, c)
    par = np.linalg.norm(n, axis=1) < 1e-12 * np.linalg.norm(a, axis=1) * np.linalg.norm(c, axis=1)
    n[par] = np.cross(a[par], ev[ce[par, 2]])
    n = np.abs(n)
    assert np.all(np.count_nonzero(n > 1e-14, axis=1) == 1)
    B = n @ B0
    six = pyoracle.export(b, U, B)
    assert six.shape == (5, 6)
    np.testing.assert_allclose(six[:, :3], np.tile(E0, (5, 1)), rtol=0, atol=1e-13 * np.abs(E0).max())
    H0 = B0 / (MU0 * 2.5)
    np.testing.assert_allclose(six[:, 3:], np.tile(H0, (5, 1)), rtol=0, atol=1e-13 * np.abs(H0).max())


def test_tet_whitney_and_face_formula_reproduce_uniform_field(workdir):
    p = configs.kuhn_cavity(str(workdir), n=4, method="dga", n_steps=1, n_probes=6, block=2, name="tetprobe",
                            dt=2e-11)
    s = open(p).read().replace("SET mu 1", "SET mu 1.7")
    rng = np.random.default_rng(3)
    q = rng.uniform(0.05, 0.95, size=(9, 3))
    pts = "".join("    SET probe {%r,%r,%r}\n" % tuple(float(v) for v in x) for x in q)
    s = re.sub(r"(    SET probe \{[^}]*\}\n)+", pts, s)
    open(p, "w").write(s)
    b = Model(p).build(1)
    t = b.tet_tables()
    P = t["pts"]
    en, fn = t["edge_nodes"], t["face_nodes"]
    assert np.all(en[:, 0] < en[:, 1]) and np.all(fn[:, 0] < fn[:, 1]) and np.all(fn[:, 1] < fn[:, 2])
    U = (P[en[:, 1]] - P[en[:, 0]]) @ E0
    area = 0.5 * np.cross(P[fn[:, 1]] - P[fn[:, 0]], P[fn[:, 2]] - P[fn[:, 0]])
    B = area @ B0
    six = pyoracle.export(b, U, B)
    assert six.shape == (9, 6)
    np.testing.assert_allclose(six[:, :3], np.tile(E0, (9, 1)), rtol=0, atol=1e-12 * np.abs(E0).max())
    H0 = B0 / (MU0 * 1.7)
    np.testing.assert_allclose(six[:, 3:], np.tile(H0, (9, 1)), rtol=0, atol=1e-12 * np.abs(H0).max())


def test_uniform_field_pins_fail_on_a_wrong_axis(workdir):
    """The pin is sensitive: the same uniform-field check applied to U built
    with E0's x and y components exchanged does not reproduce E0."""
    p = configs.kuhn_cavity(str(workdir), n=3, method="dga", n_steps=1, n_probes=4, block=2, name="tetprobe2",
                            dt=2e-11)
    b = Model(p).build(1)
    t = b.tet_tables()
    P = t["pts"]
    en = t["edge_nodes"]
    Ew = E0[[1, 0, 2]]
    six = pyoracle.export(b, (P[en[:, 1]] - P[en[:, 0]]) @ Ew, np.zeros(b.n_faces))
    assert np.abs(six[:, :3] - E0).max() > 0.1


def test_find_probe_picks_the_containing_element(workdir):
    """FindProbe (D:3645-3747) and FindFitProbe (D:3590-3643) against a
    brute-force search over every element: for points in general position
    the tet whose barycentric coordinates are all >= 0 (unique), and the FIT
    cube whose box holds the point. A point on a shared face may go to
    either neighbour (the reference resolves that by its BFS order); the
    element picked must still contain it."""
    rng = np.random.default_rng(9)
    q = [tuple(float(v) for v in x) for x in rng.uniform(0.02, 0.98, size=(12, 3))]
    q.append((0.5, 0.5, 0.5))  # a mesh node of the 4^3 Kuhn grid (jittered, so only near one)
    p = configs.kuhn_cavity(str(workdir), n=4, method="dga", n_steps=1, n_probes=2, block=2, name="findtet",
                            dt=2e-11)
    s = re.sub(r"(    SET probe \{[^}]*\}\n)+", "".join("    SET probe {%r,%r,%r}\n" % x for x in q), open(p).read())
    open(p, "w").write(s)
    b = Model(p).build(1)
    t = b.tet_tables()
    P, vols = t["pts"], t["vols"]
    pg = b.probe_geom()
    pts = b.probe_points()
    assert len(pg) == len(q)
    for g, x in zip(pg, pts):
        lam = []
        for v in vols:
            M = np.vstack([np.hstack([P[v].T, np.zeros((3, 0))]), np.ones(4)])
            lam.append(np.linalg.solve(M, np.append(x, 1.0)))
        lam = np.array(lam)
        inside = np.where(lam.min(axis=1) >= -1e-12)[0]
        assert g["elem"] in inside
        if lam[inside].min(axis=1).max() > 1e-9:  # strictly inside one tet
            assert len(inside) == 1
    # FIT: cubes of a non-cubic grid; one probe on an interior cell face
    p = configs.fit_cavity(str(workdir), n=6, step=0.1, n_steps=1, n_probes=2, name="findfit")
    s = open(p).read()
    s = re.sub(r"ygrid \{[^}]*\}", "ygrid {0,0.07,0.4300001}", s)
    fq = [(0.03, 0.02, 0.05), (0.31, 0.2, 0.4), (0.55, 0.41, 0.53), (0.12, 0.35, 0.27), (0.3, 0.105, 0.25)]
    s = re.sub(r"(    SET probe \{[^}]*\}\n)+", "".join("    SET probe {%r,%r,%r}\n" % x for x in fq), s)
    open(p, "w").write(s)
    b = Model(p).build(1)
    ft = b.fit_tables()
    P, G, ec = ft["pts"], ft["g_nodes"], ft["e_cluster"]
    lo = np.array([P[G[c]].reshape(-1, 3).min(axis=0) for c in ec])
    hi = np.array([P[G[c]].reshape(-1, 3).max(axis=0) for c in ec])
    pg = b.probe_geom()
    for g, x in zip(pg, b.probe_points()):
        holds = np.where(np.all((x >= lo - 1e-12) & (x <= hi + 1e-12), axis=1))[0]
        assert g["elem"] in holds
        strict = np.where(np.all((x > lo + 1e-9) & (x < hi - 1e-9), axis=1))[0]
        if len(strict):
            assert list(holds) == [g["elem"]]
