"""Python view of the host side (libtetfit_host.so): parse an .fdtd input,
mesh and assemble one simulation, and expose the assembled tables as numpy
arrays (copies, so they outlive the build)."""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import host_lib


class TetfitError(RuntimeError):
    pass


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    ct = np.ctypeslib.as_array(ptr, shape=(n,))
    return np.array(ct, dtype=dtype, copy=True)


def _csr(c):
    """tf_csr -> dict of numpy arrays (rowptr, col, val or sgn)."""
    nr = c.nrows
    if nr == 0 or not c.rowptr:
        return dict(nrows=nr, ncols=c.ncols, rowptr=np.zeros(nr + 1, np.int64),
                    col=np.zeros(0, np.int32), val=None, sgn=None)
    rp = _arr(c.rowptr, nr + 1, np.int64)
    nnz = int(rp[-1])
    return dict(nrows=nr, ncols=c.ncols, rowptr=rp, col=_arr(c.col, nnz, np.int32),
                val=_arr(c.val, nnz, np.float64) if c.val else None,
                sgn=_arr(c.sgn, nnz, np.int8) if c.sgn else None)


class Model:
    """A parsed .fdtd input (Discretization's parser, D:45-211)."""

    def __init__(self, path):
        L = host_lib()
        h = C.c_void_p()
        if L.tf_model_load(str(path).encode(), C.byref(h)) != 0:
            raise TetfitError(L.tfh_last_error().decode())
        self._h = h
        self.path = str(path)

    def simulation_labels(self):
        L = host_lib()
        return [L.tf_model_simulation_label(self._h, i) for i in range(L.tf_model_n_simulations(self._h))]

    def build(self, label):
        return Build(self, label)

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib().tf_model_free(self._h)
            self._h = None


class Build:
    """One simulation meshed and assembled; owns the tf_problem memory."""

    def __init__(self, model, label):
        L = host_lib()
        h = C.c_void_p()
        if L.tf_build_simulation(model._h, int(label), C.byref(h)) != 0:
            raise TetfitError(L.tfh_last_error().decode())
        self._h = h
        self.label = label
        self.problem_ptr = L.tf_build_problem(h)
        p = self.problem_ptr.contents
        self.method = {0: "fdtd", 1: "dga", 2: "gfdtd", 3: "fem"}[p.method]
        self.n_edges, self.n_faces, self.n_vols = p.n_edges, p.n_faces, p.n_vols
        self.dt, self.duration = p.dt, p.duration
        self.n_steps = L.tf_build_n_steps(h)
        self.first_step = L.tf_build_first_step(h)
        self.stable_dt = L.tf_build_stable_dt(h)
        self.possibly_unstable = bool(L.tf_build_possibly_unstable(h))
        self.cg_tol = L.tf_build_cg_tol(h)
        self.output_mode = L.tf_build_output_mode(h)
        self.output_delay = L.tf_build_output_delay(h)
        self.output_period = L.tf_build_output_period(h)
        self.output_name = L.tf_build_output_name(h).decode()

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib().tf_build_free(self._h)
            self._h = None

    # ---- raw C views (valid while the build lives)
    def sources_raw(self):
        L = host_lib()
        s = C.POINTER(_ffi.tf_source)()
        t = C.POINTER(_ffi.tf_source_term)()
        ns, nt = C.c_int32(), C.c_int64()
        L.tf_build_sources(self._h, C.byref(s), C.byref(ns), C.byref(t), C.byref(nt))
        return s, ns.value, t, nt.value

    def probes_raw(self):
        L = host_lib()
        p = C.POINTER(_ffi.tf_probe)()
        pts = _ffi.f64p()
        n = C.c_int32()
        L.tf_build_probes(self._h, C.byref(p), C.byref(pts), C.byref(n))
        return p, pts, n.value

    # ---- numpy copies
    def sources(self):
        s, ns, t, nt = self.sources_raw()
        src = [dict(carrier=s[i].carrier, field=s[i].field, amp=s[i].amp, freq=s[i].freq,
                    delay=s[i].delay, width=s[i].width, impedance=s[i].impedance) for i in range(ns)]
        terms = np.array([(t[i].edge, t[i].source, t[i].weight) for i in range(nt)],
                         dtype=[("edge", np.int32), ("source", np.int32), ("weight", np.float64)])
        return src, terms

    def probe_points(self):
        _, pts, n = self.probes_raw()
        return _arr(pts, 3 * n, np.float64).reshape(n, 3)

    def probe_weights(self):
        p, _, n = self.probes_raw()
        out = []
        for i in range(n):
            q = p[i]
            out.append(dict(e_idx=np.array(q.e_idx[:q.ne], np.int32),
                            e_w=np.array([[q.e_w[k][c] for c in range(3)] for k in range(q.ne)]),
                            h_idx=np.array(q.h_idx[:q.nh], np.int32),
                            h_w=np.array([[q.h_w[k][c] for c in range(3)] for k in range(q.nh)])))
        return out

    def source_desc(self):
        L = host_lib()
        d = C.POINTER(_ffi.tf_source_desc)()
        e = C.POINTER(_ffi.tf_source_edge)()
        nd, ne = C.c_int32(), C.c_int64()
        L.tf_build_source_desc(self._h, C.byref(d), C.byref(nd), C.byref(e), C.byref(ne))
        descs = []
        for i in range(nd.value):
            x = d[i]
            descs.append(dict(type=x.type, carrier=x.carrier, profile=x.profile, dir=x.dir,
                              bfunc=list(x.bfunc), amp=x.amp, freq=x.freq, delay=x.delay,
                              width=x.width, impedance=x.impedance, radius=x.radius,
                              decay=x.decay, kvec=list(x.kvec), center=list(x.center)))
        edges = []
        for i in range(ne.value):
            x = e[i]
            edges.append(dict(edge=x.edge, field=x.field, src=list(x.src[:x.n_src]),
                              bary=np.array(list(x.bary)), vec=np.array(list(x.vec))))
        return descs, edges

    def probe_geom(self):
        L = host_lib()
        g = C.POINTER(_ffi.tf_probe_geom)()
        n = C.c_int32()
        L.tf_build_probe_geom(self._h, C.byref(g), C.byref(n))
        out = []
        for i in range(n.value):
            x = g[i]
            out.append(dict(elem=x.elem, edges=np.array(list(x.edges)), faces=np.array(list(x.faces)),
                            point=np.array(list(x.point)),
                            nodes=np.array([[x.nodes[a][c] for c in range(3)] for a in range(4)]),
                            L=np.array(list(x.L)), mu=np.array(list(x.mu)).reshape(3, 3),
                            volume=x.volume))
        return out

    def fit_ops(self):
        f = self.problem_ptr.contents.fit
        E, F = self.n_edges, self.n_faces
        ct_ptr = _arr(f.ct_ptr, E + 1, np.int64)
        nct = int(ct_ptr[-1])
        return dict(
            ct_ptr=ct_ptr, ct_face=_arr(f.ct_face, nct, np.int32), ct_sgn=_arr(f.ct_sgn, nct, np.int8),
            c_edge=_arr(f.c_edge, 4 * F, np.int32).reshape(F, 4), curl=_arr(f.curl, F, np.int8),
            dual_curl=_arr(f.dual_curl, E, np.int8),
            M_h=_arr(f.M_h, E, np.float64), M_q=_arr(f.M_q, E, np.float64), upml=_arr(f.upml, E, np.float64),
            M_mu=_arr(f.M_mu, F, np.float64), M_nu=_arr(f.M_nu, F, np.float64), fpml=_arr(f.fpml, F, np.float64),
            common_edges=_arr(f.common_edges, f.n_common, np.int32),
            uncommon_edges=_arr(f.uncommon_edges, f.n_uncommon, np.int32),
            bc_edges=_arr(f.bc_edges, f.n_bc, np.int32), bc_value=_arr(f.bc_value, f.n_bc, np.float64),
            src_edges=_arr(f.src_edges, f.n_src, np.int32),
            tbc_faces=_arr(f.tbc_faces, f.n_tbc, np.int32),
            ep_vec=_arr(f.ep_vec, E, np.float64), nu_vec=_arr(f.nu_vec, F, np.float64),
            edge_sid=_arr(f.edge_sid, E, np.int32) if f.edge_sid else None,
            face_sid=_arr(f.face_sid, F, np.int32) if f.face_sid else None, cells=tuple(f.cells))

    def fit_tables(self):
        L = host_lib()
        ec, df, gn, bf = _ffi.i32p(), _ffi.i32p(), _ffi.i32p(), _ffi.i32p()
        pts = _ffi.f64p()
        nc, npt = C.c_int32(), C.c_int32()
        L3 = (C.c_double * 3)()
        L.tf_build_fit_tables(self._h, C.byref(ec), C.byref(df), C.byref(gn), C.byref(pts),
                              C.byref(nc), C.byref(npt), C.byref(bf), L3)
        return dict(e_cluster=_arr(ec, 12 * nc.value, np.int32).reshape(-1, 12),
                    d_faces=_arr(df, 6 * nc.value, np.int32).reshape(-1, 6),
                    g_nodes=_arr(gn, 2 * self.n_edges, np.int32).reshape(-1, 2),
                    pts=_arr(pts, 3 * npt.value, np.float64).reshape(-1, 3),
                    boundary_face=_arr(bf, self.n_faces, np.int32), L=np.array(list(L3)))

    def tet_tables(self):
        L = host_lib()
        vols, vtf, en, fn, bf, vm = (_ffi.i32p() for _ in range(6))
        pts = _ffi.f64p()
        nv, npt = C.c_int32(), C.c_int32()
        L.tf_build_tet_tables(self._h, C.byref(vols), C.byref(pts), C.byref(vtf), C.byref(en),
                              C.byref(fn), C.byref(nv), C.byref(npt), C.byref(bf), C.byref(vm))
        return dict(vols=_arr(vols, 4 * nv.value, np.int32).reshape(-1, 4),
                    pts=_arr(pts, 3 * npt.value, np.float64).reshape(-1, 3),
                    vtf=_arr(vtf, 4 * nv.value, np.int32).reshape(-1, 4),
                    edge_nodes=_arr(en, 2 * self.n_edges, np.int32).reshape(-1, 2),
                    face_nodes=_arr(fn, 3 * self.n_faces, np.int32).reshape(-1, 3),
                    boundary_face=_arr(bf, self.n_faces, np.int32),
                    vol_material=_arr(vm, nv.value, np.int32))

    def dga_ops(self):
        d = self.problem_ptr.contents.dga
        E = self.n_edges
        return dict(C=_csr(d.C), Ct=_csr(d.Ct), H=_csr(d.H), N=_csr(d.N), E=_csr(d.E), A=_csr(d.A),
                    RHS1=_csr(d.RHS1), bnd_edge=_arr(d.bnd_edge, E, np.int32), n_bnd=d.n_bnd,
                    edge_class=_arr(d.edge_class, E, np.int8), fixed_value=_arr(d.fixed_value, E, np.float64),
                    h_row=_arr(d.h_row, E, np.int32) if d.h_row else None,
                    free_edges=_arr(d.free_edges, d.n_free, np.int32), Sig=_csr(d.Sig), frac=self.frac_ops())

    def edge_rank(self):
        """tf_dga_ops.edge_rank (the locality rank of every edge), or None."""
        d = self.problem_ptr.contents.dga
        if self.method == "fdtd" or not d.edge_rank:
            return None
        return _arr(d.edge_rank, self.n_edges, np.int32)

    def face_rank(self):
        """tf_dga_ops.face_rank (the locality rank of every face), or None."""
        d = self.problem_ptr.contents.dga
        if self.method == "fdtd" or not d.face_rank:
            return None
        return _arr(d.face_rank, self.n_faces, np.int32)

    def frac_ops(self):
        """Fractional operators of a lossy dga problem (tf_dga_frac), or None."""
        fr = self.problem_ptr.contents.dga.frac
        if not fr.enabled:
            return None
        sizes = {k: getattr(fr, k) for k in ("h_size", "p_size", "q_size", "b_size", "n_size", "r_size", "s_size")}
        return dict(sizes, Ku=_csr(fr.Ku), Ru=_csr(fr.Ru), Mu=_csr(fr.Mu), Kf=_csr(fr.Kf), Rf=_csr(fr.Rf),
                    Tf=_csr(fr.Tf), slot=_arr(fr.slot, self.n_edges, np.int32))


def write_kuhn_mesh(path, n, jitter=0.2, seed=0, block=8):
    """Writes an n^3 Kuhn mesh (netgen neutral); returns (max material label,
    min |vol| / (h^3/6))."""
    L = host_lib()
    r = C.c_double()
    lab = L.tf_write_kuhn_mesh(str(path).encode(), int(n), float(jitter), int(seed), int(block), C.byref(r))
    if lab < 0:
        raise TetfitError(L.tfh_last_error().decode())
    return lab, r.value
