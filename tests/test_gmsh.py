"""gmsh meshes (`SET mesher gmsh`, ConvertFromGMSH D:5692-5866): a gmsh 2.2
ASCII file is turned into the netgen text the reader takes, with the node
coordinates written at the stream's default 6 significant digits. A Kuhn
mesh written in both formats (nodes already at 6 digits, tetrahedra
shuffled and with unsorted nodes in the gmsh file, the two gmsh tags of each
element) must assemble to identical operators; with more digits in the
gmsh file the result equals the netgen mesh with the coordinates rounded to
6 significant digits."""
import os
import re

import numpy as np

from tetfit_amd import Model, configs


def read_netgen(path):
    lines = open(path).read().split("\n")
    k = 0
    n = int(lines[k]); k += 1
    pts = [list(map(float, lines[k + i].split())) for i in range(n)]; k += n
    nt = int(lines[k]); k += 1
    tets = [list(map(int, lines[k + i].split())) for i in range(nt)]; k += nt
    ns = int(lines[k]); k += 1
    tris = [list(map(int, lines[k + i].split())) for i in range(ns)]
    return pts, tets, tris


def write_netgen(path, pts, tets, tris):
    with open(path, "w") as f:
        f.write("%d\n" % len(pts))
        for p in pts:
            f.write("  %r  %r  %r\n" % tuple(p))
        f.write("%d\n" % len(tets))
        for t in tets:
            f.write(" ".join(map(str, t)) + "\n")
        f.write("%d\n" % len(tris))
        for t in tris:
            f.write(" ".join(map(str, t)) + "\n")


def write_gmsh(path, pts, tets, tris, rng, fmt="%r"):
    els = [(2, t[0], t[1:]) for t in tris] + [(4, t[0], list(rng.permutation(t[1:]))) for t in tets]
    order = rng.permutation(len(els))
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(pts))
        for i, p in enumerate(pts):
            f.write(("%d " + " ".join([fmt] * 3) + "\n") % ((i + 1,) + tuple(p)))
        f.write("$EndNodes\n$Elements\n%d\n" % len(els))
        for num, k in enumerate(order):
            typ, lab, nodes = els[k]
            f.write("%d %d 2 %d %d %s\n" % (num + 1, typ, lab, 100 + lab, " ".join(map(str, nodes))))
        f.write("$EndElements\n")


def build(path):
    return Model(path).build(1)


def ops_equal(a, b):
    oa, ob = a.dga_ops(), b.dga_ops()
    for k in ("C", "H", "N", "E"):
        for f in ("rowptr", "col", "val", "sgn"):
            x, y = oa[k][f], ob[k][f]
            if x is None or y is None:
                assert x is None and y is None
                continue
            np.testing.assert_array_equal(x, y, err_msg=f"{k}.{f}")


def setup(tmp_path, method="dga"):
    p = configs.kuhn_cavity(str(tmp_path), n=3, method=method, n_steps=4, n_probes=2, block=2, seed=4, dt=1e-12,
                            name=f"g_{method}")
    mesh = re.search(r"SET file (\S+)", open(p).read()).group(1)
    return p, os.path.join(str(tmp_path), mesh)


def test_gmsh_mesh_assembles_like_netgen(tmp_path):
    rng = np.random.default_rng(0)
    p, mesh = setup(tmp_path)
    pts, tets, tris = read_netgen(mesh)
    pts6 = [[float("%g" % v) for v in q] for q in pts]
    write_netgen(mesh, pts6, tets, tris)
    ref = build(p)
    gm = mesh[:-5] + ".msh"
    write_gmsh(gm, pts6, tets, tris, rng)
    pg = p.replace(".fdtd", "_gmsh.fdtd")
    s = open(p).read().replace("SET mesher netgen", "SET mesher gmsh").replace(os.path.basename(mesh),
                                                                               os.path.basename(gm))
    open(pg, "w").write(s)
    got = build(pg)
    assert got.n_edges == ref.n_edges and got.n_faces == ref.n_faces
    ops_equal(got, ref)
    np.testing.assert_array_equal(got.tet_tables()["vol_material"], ref.tet_tables()["vol_material"])


def test_gmsh_coordinates_round_to_six_digits(tmp_path):
    rng = np.random.default_rng(1)
    p, mesh = setup(tmp_path)
    pts, tets, tris = read_netgen(mesh)
    gm = mesh[:-5] + ".msh"
    write_gmsh(gm, pts, tets, tris, rng)  # full-precision coordinates
    pts6 = [[float("%g" % v) for v in q] for q in pts]
    assert pts6 != pts
    write_netgen(mesh, pts6, tets, tris)
    ref = build(p)
    pg = p.replace(".fdtd", "_gmsh.fdtd")
    s = open(p).read().replace("SET mesher netgen", "SET mesher gmsh").replace(os.path.basename(mesh),
                                                                               os.path.basename(gm))
    open(pg, "w").write(s)
    ops_equal(build(pg), ref)
