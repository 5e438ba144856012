"""ctypes front of oracle/liboracle.so (the CPU checker).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


class oracle_out(C.Structure):
    _fields_ = [("times", C.POINTER(C.c_double)), ("vals", C.POINTER(C.c_double)),
                ("cap_samples", C.c_int64), ("n_samples", C.c_int64),
                ("poynting", C.POINTER(C.c_double)), ("cap_poynting", C.c_int64), ("n_poynting", C.c_int64),
                ("U", C.POINTER(C.c_double)), ("F", C.POINTER(C.c_double)), ("B", C.POINTER(C.c_double)),
                ("cg_iters_total", C.c_int64), ("cg_iters_max", C.c_int32), ("cg_not_converged", C.c_int32),
                ("cg_max_relres", C.c_double), ("seconds", C.c_double),
                ("losses", C.POINTER(C.c_double)), ("cap_losses", C.c_int64), ("n_losses", C.c_int64),
                ("energy", C.POINTER(C.c_double))]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is not built (make -C oracle)")
    from tetfit_amd import _ffi
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.oracle_run.argtypes = [C.POINTER(_ffi.tf_problem), C.POINTER(_ffi.tf_source_desc), C.c_int32,
                             C.POINTER(_ffi.tf_source_edge), C.c_int64, C.POINTER(_ffi.tf_probe_geom), C.c_int32,
                             C.c_int32, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int32,
                             C.c_double, C.c_int32, dp, dp, dp, C.POINTER(oracle_out)]
    L.oracle_run.restype = C.c_int32
    L.oracle_last_error.restype = C.c_char_p
    L.oracle_energy.argtypes = [C.POINTER(_ffi.tf_problem), dp, dp, dp, dp]
    L.oracle_energy.restype = None
    L.oracle_source_compute.argtypes = [C.POINTER(_ffi.tf_source_desc), dp, C.c_double, dp]
    L.oracle_source_compute.restype = None
    L.oracle_cg_solve.argtypes = [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32), dp, dp, dp,
                                  C.c_double, C.c_int32, dp, dp]
    L.oracle_cg_solve.restype = C.c_int32
    L.oracle_export.argtypes = [C.c_int32, C.POINTER(_ffi.tf_probe_geom), C.c_int32, C.c_int32, C.c_int32, dp, dp,
                                dp]
    L.oracle_export.restype = None
    fp = C.POINTER(C.c_float)
    L.oracle_mlp.argtypes = [fp, fp, dp, C.c_int64, dp, dp]
    L.oracle_mlp.restype = None
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def run(build, first=None, n=None, faithful=True, cg_tol=None, cg_maxiter=100, U0=None, F0=None, B0=None,
        output=True, max_samples=None, energy=False):
    """Runs the oracle on a tetfit_amd.host.Build; returns a dict with
    times, samples [k, probe, 6], poynting, losses (gfdtd's Losses vector in
    push order), U, F, B and CG statistics; with energy, also We and Wh at
    every sample."""
    from tetfit_amd import _ffi
    L = lib()
    first = build.first_step if first is None else first
    n = build.n_steps if n is None else n
    h = build._h
    H = __import__("tetfit_amd._ffi", fromlist=["host_lib"]).host_lib()
    sd = C.POINTER(_ffi.tf_source_desc)()
    se = C.POINTER(_ffi.tf_source_edge)()
    nsd, nse = C.c_int32(), C.c_int64()
    H.tf_build_source_desc(h, C.byref(sd), C.byref(nsd), C.byref(se), C.byref(nse))
    pg = C.POINTER(_ffi.tf_probe_geom)()
    npg = C.c_int32()
    H.tf_build_probe_geom(h, C.byref(pg), C.byref(npg))
    np_ = npg.value
    cap = max_samples if max_samples is not None else n + 2
    times = np.zeros(cap)
    vals = np.zeros(cap * max(np_, 1) * 6)
    poy = np.zeros(n + 2)
    loss = np.zeros(2 * n + 2)
    U = np.zeros(build.n_edges)
    F = np.zeros(build.n_faces)
    B = np.zeros(build.n_faces)
    o = oracle_out()
    o.times, o.vals, o.cap_samples = _dp(times), _dp(vals), cap
    o.poynting, o.cap_poynting = _dp(poy), len(poy)
    o.losses, o.cap_losses = _dp(loss), len(loss)
    o.U, o.F, o.B = _dp(U), _dp(F), _dp(B)
    en = np.zeros(2 * cap) if energy else None
    o.energy = _dp(en)
    tol = build.cg_tol if cg_tol is None else cg_tol
    u0 = None if U0 is None else np.ascontiguousarray(U0, dtype=np.float64)
    f0 = None if F0 is None else np.ascontiguousarray(F0, dtype=np.float64)
    b0 = None if B0 is None else np.ascontiguousarray(B0, dtype=np.float64)
    rc = L.oracle_run(build.problem_ptr, sd, nsd, se, nse, pg, np_, 1 if (output and build.output_mode in (1, 2, 3)) else 0,
                      build.output_delay, build.output_period, int(first), int(n), 1 if faithful else 0,
                      float(tol), int(cg_maxiter), _dp(u0), _dp(f0), _dp(b0), C.byref(o))
    if rc != 0:
        raise RuntimeError(L.oracle_last_error().decode())
    ns = min(o.n_samples, cap)
    return dict(times=times[:ns], samples=vals[:ns * np_ * 6].reshape(ns, np_, 6), n_samples=o.n_samples,
                poynting=poy[:min(o.n_poynting, len(poy))], losses=loss[:min(o.n_losses, len(loss))], U=U, F=F, B=B,
                cg_iters_total=o.cg_iters_total, cg_iters_max=o.cg_iters_max,
                cg_not_converged=o.cg_not_converged, cg_max_relres=o.cg_max_relres, seconds=o.seconds,
                We=None if en is None else en[0:2 * ns:2].copy(), Wh=None if en is None else en[1:2 * ns:2].copy())


def energy(build, U, B):
    """(We, Wh) of a state as the oracle's energy series records it."""
    L = lib()
    we, wh = C.c_double(), C.c_double()
    u = np.ascontiguousarray(U, dtype=np.float64)
    b = np.ascontiguousarray(B, dtype=np.float64)
    L.oracle_energy(build.problem_ptr, _dp(u), _dp(b), C.byref(we), C.byref(wh))
    return we.value, wh.value


def source_compute(desc_struct, p, t):
    L = lib()
    out = np.zeros(3)
    pp = np.ascontiguousarray(p, dtype=np.float64)
    L.oracle_source_compute(C.byref(desc_struct), _dp(pp), float(t), _dp(out))
    return out


def cg_solve(rowptr, col, val, b, x0, tol, maxiter):
    L = lib()
    n = len(b)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    x = np.zeros(n)
    err = C.c_double()
    it = L.oracle_cg_solve(n, rowptr.ctypes.data_as(C.POINTER(C.c_int64)), col.ctypes.data_as(C.POINTER(C.c_int32)),
                           _dp(val), _dp(b), _dp(x0), float(tol), int(maxiter), _dp(x), C.byref(err))
    return x, it, err.value


def mlp(params, x, gy=None):
    """FP64 preconditioner MLP: returns (y, grad) for rows x [n, 16]."""
    L = lib()
    fp = C.POINTER(C.c_float)
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(x, dtype=np.float32)
    n = X.shape[0]
    G = np.zeros(n) if gy is None else np.ascontiguousarray(gy, dtype=np.float64)
    y = np.zeros(n)
    g = np.zeros(P.size)
    L.oracle_mlp(P.ctypes.data_as(fp), X.ctypes.data_as(fp), _dp(G), n, _dp(y), _dp(g))
    return y, g


def export(build, U, B):
    """Probe samples [probe, 6] of the state (U, B) by the checker's
    GetFit*Field / GetWhitneyElectricField / GetMagneticField."""
    from tetfit_amd import _ffi
    L = lib()
    H = _ffi.host_lib()
    pg = C.POINTER(_ffi.tf_probe_geom)()
    npg = C.c_int32()
    H.tf_build_probe_geom(build._h, C.byref(pg), C.byref(npg))
    u = np.ascontiguousarray(U, dtype=np.float64)
    b = np.ascontiguousarray(B, dtype=np.float64)
    six = np.zeros(max(npg.value, 1) * 6)
    L.oracle_export(1 if build.method == "fdtd" else 0, pg, npg.value, build.n_edges, build.n_faces, _dp(u), _dp(b),
                    _dp(six))
    return six[:npg.value * 6].reshape(npg.value, 6)
