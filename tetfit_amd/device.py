"""Python view of the stepping backend (libtetfit_hip.so): one device context
per simulation, created from a host Build (tf_create / tf_set_sources /
tf_set_probes / tf_set_output / tf_run / tf_read_*)."""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import hip_lib

PREC = {"fp64": 0, "fp32": 1}
PRECOND = {"jacobi": 0, "learned": 1}


class DeviceError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise DeviceError(f"tf status {rc}: {hip_lib().tf_last_error().decode()}")


def device_count():
    return hip_lib().tf_device_count()


class Simulation:
    """Device context for one assembled simulation."""

    def __init__(self, build, precision="fp64", output=True, energy=None):
        """energy: record (We, Wh) at every output sample (tf_set_energy_output);
        None = when the simulation's output mode is l2norm or maxerror (the
        modes in which the reference writes its energy files)."""
        L = hip_lib()
        self.build = build
        self.precision = precision
        h = C.c_void_p()
        _check(L.tf_create(build.problem_ptr, PREC[precision], C.byref(h)))
        self._h = h
        s, ns, t, nt = build.sources_raw()
        _check(L.tf_set_sources(h, s, ns, t, nt))
        p, _, npr = build.probes_raw()
        self.n_probes = npr
        _check(L.tf_set_probes(h, p, npr))
        if output and build.output_mode in (1, 2, 3):
            _check(L.tf_set_output(h, build.output_delay, build.output_period))
        if energy is None:
            energy = output and build.output_mode in (2, 3)
        if energy:
            _check(L.tf_set_energy_output(h, 1))
        self.next_step = build.first_step

    def close(self):
        if getattr(self, "_h", None):
            hip_lib().tf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run(self, n=None, first=None, cg_tol=None, cg_maxiter=100, precond="jacobi", learn_steps=0,
            train_iters=0, lr=0.0, guess=0):
        """Advances n steps from `first` (default: where the last run ended).
        precond "learned" trains the preconditioner during the context's first
        `learn_steps` solves (0: 200) and uses it afterwards. guess: start of
        each gfdtd solve, 0 the warm start (the reference), 1 / 2 linear /
        quadratic extrapolation from the previous steps (same stop test)."""
        L = hip_lib()
        first = self.next_step if first is None else first
        n = self.build.n_steps - (first - self.build.first_step) if n is None else n
        opts = _ffi.tf_cg_opts(self.build.cg_tol if cg_tol is None else cg_tol, int(cg_maxiter), PRECOND[precond],
                               int(learn_steps), int(train_iters), float(lr), int(guess))
        st = _ffi.tf_stats()
        _check(L.tf_run(self._h, int(first), int(n), C.byref(opts), C.byref(st)))
        self.next_step = first + n
        out = dict(steps=st.steps_done, cg_iters_total=st.cg_iters_total, cg_iters_max=st.cg_iters_max,
                   cg_not_converged=st.cg_not_converged, cg_max_relres=st.cg_max_relres, seconds=st.seconds)
        if precond == "learned":
            out.update(pc_state=st.pc_state, pc_pairs=st.pc_pairs, pc_fallbacks=st.pc_fallbacks,
                       pc_learned_solves=st.pc_learned_solves, pc_learned_iters=st.pc_learned_iters,
                       pc_jacobi_mean_iters=st.pc_jacobi_mean_iters, pc_train_seconds=st.pc_train_seconds,
                       pc_loss0=st.pc_loss0, pc_loss=st.pc_loss)
        return out

    def state(self):
        U = np.zeros(self.build.n_edges)
        F = np.zeros(self.build.n_faces)
        B = np.zeros(self.build.n_faces)
        dp = C.POINTER(C.c_double)
        _check(hip_lib().tf_read_state(self._h, U.ctypes.data_as(dp), F.ctypes.data_as(dp), B.ctypes.data_as(dp)))
        return U, F, B

    def set_state(self, U=None, F=None, B=None):
        dp = C.POINTER(C.c_double)
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(dp)
        keep = [U, F, B]
        _check(hip_lib().tf_write_state(self._h, conv(U), conv(F), conv(B)))
        del keep

    def probes(self):
        L = hip_lib()
        n = C.c_int64()
        _check(L.tf_read_probes(self._h, None, None, 0, C.byref(n)))
        k = n.value
        times = np.zeros(k)
        vals = np.zeros(max(k * self.n_probes * 6, 1))
        dp = C.POINTER(C.c_double)
        _check(L.tf_read_probes(self._h, times.ctypes.data_as(dp), vals.ctypes.data_as(dp), k, C.byref(n)))
        return times, vals[:k * self.n_probes * 6].reshape(k, self.n_probes, 6)

    def energy(self):
        """(We, Wh) of the current state, fp64 reductions on the device
        (fdtd: ep_vec / nu_vec weights; dga, gfdtd: U.E U / 2, B.N B / 2)."""
        we, wh = C.c_double(), C.c_double()
        _check(hip_lib().tf_energy(self._h, C.byref(we), C.byref(wh)))
        return we.value, wh.value

    def energy_series(self):
        """tf_read_energy: (times, We, Wh) of the output samples so far."""
        L = hip_lib()
        n = C.c_int64()
        _check(L.tf_read_energy(self._h, None, None, 0, C.byref(n)))
        k = n.value
        times = np.zeros(max(k, 1))
        vals = np.zeros(max(2 * k, 2))
        dp = C.POINTER(C.c_double)
        _check(L.tf_read_energy(self._h, times.ctypes.data_as(dp), vals.ctypes.data_as(dp), k, C.byref(n)))
        return times[:k], vals[0:2 * k:2].copy(), vals[1:2 * k:2].copy()

    def poynting(self):
        L = hip_lib()
        n = C.c_int64()
        _check(L.tf_read_poynting(self._h, None, 0, C.byref(n)))
        v = np.zeros(n.value)
        _check(L.tf_read_poynting(self._h, v.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return v

    def losses(self):
        """tf_read_losses: the losses column of poynting.dat, one per step."""
        L = hip_lib()
        n = C.c_int64()
        _check(L.tf_read_losses(self._h, None, 0, C.byref(n)))
        v = np.zeros(n.value)
        _check(L.tf_read_losses(self._h, v.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return v

    def set_profiling(self, on=True):
        _check(hip_lib().tf_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        L = hip_lib()
        names = C.create_string_buffer(4096)
        ms = (C.c_double * 64)()
        nl = (C.c_int64 * 64)()
        n = C.c_int32()
        _check(L.tf_kernel_times(self._h, names, 4096, ms, nl, 64, C.byref(n)))
        raw = names.raw.split(b"\0")
        return {raw[i].decode(): (ms[i], nl[i]) for i in range(n.value)}

    def device_bytes(self):
        b = C.c_int64()
        _check(hip_lib().tf_device_bytes(self._h, C.byref(b)))
        return b.value

    TUNED = ["sell_ct_kernel", "sell_h_kernel", "sell_n_kernel", "cg_w_kernel", "cg_pg_kernel", "cg_wu_kernel"]

    def tuning(self):
        """tf_get_tuning: {kernel: (blocks, unroll, band)} (blocks 0 = default grid)."""
        bl, un, bd, n = (C.c_int32 * 8)(), (C.c_int32 * 8)(), (C.c_int32 * 8)(), C.c_int32()
        _check(hip_lib().tf_get_tuning(self._h, bl, un, bd, 8, C.byref(n)))
        return {self.TUNED[k]: (bl[k], un[k], bd[k]) for k in range(min(n.value, len(self.TUNED)))}

    def autotune(self):
        """tf_autotune; returns {kernel: (blocks, unroll, band)}."""
        _check(hip_lib().tf_autotune(self._h))
        return self.tuning()

    def set_graphs(self, on):
        """tf_set_graphs: step fdtd / dga in HIP graphs when nothing is sampled (default on)."""
        _check(hip_lib().tf_set_graphs(self._h, int(bool(on))))

    def set_tuning(self, choice):
        """tf_set_tuning from {kernel: (blocks, unroll[, band])} (kernels as in
        autotune(); kernels not named keep their shape)."""
        cur = self.tuning()
        bl, un, bd = (C.c_int32 * 8)(), (C.c_int32 * 8)(), (C.c_int32 * 8)()
        for k, name in enumerate(self.TUNED):
            t = tuple(choice.get(name, cur[name]))
            if len(t) == 2:
                t = t + (cur[name][2],)
            bl[k], un[k], bd[k] = (int(v) for v in t)
        _check(hip_lib().tf_set_tuning(self._h, bl, un, bd, len(self.TUNED)))

    def fit_grid_order(self):
        """tf_fit_layout: True when the FIT fields sit in grid order."""
        v = C.c_int32()
        _check(hip_lib().tf_fit_layout(self._h, C.byref(v)))
        return bool(v.value)

    def sell_info(self, which):
        """tf_sell_info: storage of a device operator ("ct", "h", "n", "a",
        "k" (gfdtd residual operator E_f - [A 0]), "rhs1", "ctf") as
        dict(rows, slices, nnz, entries, col_bytes, col_bases) (col_bases:
        column bases per slice, 1 or 4 with 16-bit codes, 0 with int32)."""
        k = ["ct", "h", "n", "a", "k", "rhs1", "ctf"].index(which)
        info = (C.c_int64 * 6)()
        _check(hip_lib().tf_sell_info(self._h, k, info))
        return dict(rows=info[0], slices=info[1], nnz=info[2], entries=info[3], col_bytes=info[4],
                    col_bases=info[5])


class BatchSimulation:
    """One device context stepping many independent simulations together
    (tf_create_batch): operators concatenated block-diagonally, one launch per
    phase for the whole batch, each instance on its own dt and sources. dga
    instances step explicitly; gfdtd instances (lossless) solve one Jacobi CG
    per instance per step, each with its own scalars and stop test, at the
    instances' common CG tolerance."""

    def __init__(self, builds, precision="fp32", sample_every=0):
        L = hip_lib()
        self.builds = list(builds)
        arr = (C.POINTER(_ffi.tf_problem) * len(self.builds))(*[b.problem_ptr for b in self.builds])
        h = C.c_void_p()
        _check(L.tf_create_batch(arr, len(self.builds), PREC[precision], C.byref(h)))
        self._h = h
        self.n_probes = []
        for k, b in enumerate(self.builds):
            s, ns, t, nt = b.sources_raw()
            _check(L.tf_batch_set_sources(h, k, s, ns, t, nt))
            p, _, npr = b.probes_raw()
            _check(L.tf_batch_set_probes(h, k, p, npr))
            self.n_probes.append(npr)
        _check(L.tf_batch_set_output(h, int(sample_every)))
        self.next_step = 1

    def close(self):
        if getattr(self, "_h", None):
            hip_lib().tf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_graphs(self, on):
        """tf_set_graphs: step the batch in HIP graphs when nothing is sampled (default on)."""
        _check(hip_lib().tf_set_graphs(self._h, int(bool(on))))

    def run(self, n, first=None, cg_tol=None, cg_maxiter=100):
        first = self.next_step if first is None else first
        st = _ffi.tf_stats()
        opts = None
        if self.builds[0].method == "gfdtd":
            tols = {b.cg_tol for b in self.builds}
            if cg_tol is None and len(tols) != 1:
                raise ValueError("gfdtd batch instances need one CG tolerance")
            opts = C.byref(_ffi.tf_cg_opts(tols.pop() if cg_tol is None else cg_tol, int(cg_maxiter),
                                           PRECOND["jacobi"], 0, 0, 0.0, 0))
        _check(hip_lib().tf_run(self._h, int(first), int(n), opts, C.byref(st)))
        self.next_step = first + n
        return dict(steps=st.steps_done, seconds=st.seconds, cg_iters_total=st.cg_iters_total,
                    cg_iters_max=st.cg_iters_max, cg_not_converged=st.cg_not_converged,
                    cg_max_relres=st.cg_max_relres)

    def poynting(self):
        """tf_read_poynting: the batch's Poynting scalar per step (all instances summed)."""
        L = hip_lib()
        n = C.c_int64()
        _check(L.tf_read_poynting(self._h, None, 0, C.byref(n)))
        v = np.zeros(max(n.value, 1))
        _check(L.tf_read_poynting(self._h, v.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
        return v[:n.value]

    def state(self, k):
        b = self.builds[k]
        U, F, B = np.zeros(b.n_edges), np.zeros(b.n_faces), np.zeros(b.n_faces)
        dp = C.POINTER(C.c_double)
        _check(hip_lib().tf_batch_read_state(self._h, k, U.ctypes.data_as(dp), F.ctypes.data_as(dp),
                                             B.ctypes.data_as(dp)))
        return U, F, B

    def probes(self, k):
        L = hip_lib()
        n = C.c_int64()
        _check(L.tf_batch_read_probes(self._h, k, None, None, 0, C.byref(n)))
        m, npr = n.value, self.n_probes[k]
        times = np.zeros(m)
        vals = np.zeros(max(m * npr * 6, 1))
        dp = C.POINTER(C.c_double)
        _check(L.tf_batch_read_probes(self._h, k, times.ctypes.data_as(dp), vals.ctypes.data_as(dp), m,
                                      C.byref(n)))
        return times, vals[:m * npr * 6].reshape(m, npr, 6)

    def set_profiling(self, on=True):
        _check(hip_lib().tf_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        return Simulation.kernel_times(self)

    TUNED = Simulation.TUNED

    def tuning(self):
        return Simulation.tuning(self)

    def autotune(self):
        return Simulation.autotune(self)

    def device_bytes(self):
        return Simulation.device_bytes(self)


def mlp_grad(params, x, gy):
    """Device forward + backward of the preconditioner MLP (tf_mlp_grad):
    returns (y, grad) for feature rows x [n, 16], output gradients gy [n]."""
    f32 = C.POINTER(C.c_float)
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(x, dtype=np.float32)
    G = np.ascontiguousarray(gy, dtype=np.float32)
    n = len(G)
    assert P.size == _ffi.MLP_PARAMS and X.shape == (n, 16)
    y = np.zeros(n, np.float32)
    g = np.zeros(_ffi.MLP_PARAMS, np.float32)
    _check(hip_lib().tf_mlp_grad(P.ctypes.data_as(f32), X.ctypes.data_as(f32), G.ctypes.data_as(f32), n,
                                 y.ctypes.data_as(f32), g.ctypes.data_as(f32)))
    return y, g
