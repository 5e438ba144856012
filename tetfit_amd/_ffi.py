"""ctypes declarations of the C ABIs in include/tetfit_problem.h,
include/tetfit_host.h and include/tetfit_hip.h, and loaders for the two
in-tree shared libraries (libtetfit_host.so, libtetfit_hip.so).

The loaders never fall back to anything else: a missing library raises.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(PKG_DIR, "libtetfit_host.so")
HIP_LIB = os.path.join(PKG_DIR, "libtetfit_hip.so")

i8p = C.POINTER(C.c_int8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


class tf_csr(C.Structure):
    _fields_ = [("nrows", C.c_int32), ("ncols", C.c_int32), ("nnz", C.c_int64),
                ("rowptr", i64p), ("col", i32p), ("val", f64p), ("sgn", i8p)]


class tf_fit_ops(C.Structure):
    _fields_ = [("ct_ptr", i64p), ("ct_face", i32p), ("ct_sgn", i8p), ("c_edge", i32p),
                ("curl", i8p), ("dual_curl", i8p),
                ("M_h", f64p), ("M_q", f64p), ("upml", f64p),
                ("M_mu", f64p), ("M_nu", f64p), ("fpml", f64p),
                ("common_edges", i32p), ("n_common", C.c_int64),
                ("uncommon_edges", i32p), ("n_uncommon", C.c_int64),
                ("bc_edges", i32p), ("n_bc", C.c_int64), ("bc_value", f64p),
                ("src_edges", i32p), ("n_src", C.c_int64),
                ("tbc_faces", i32p), ("n_tbc", C.c_int64), ("ep_vec", f64p), ("nu_vec", f64p),
                ("edge_sid", i32p), ("face_sid", i32p), ("cells", C.c_int32 * 3)]


class tf_dga_frac(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("h_size", C.c_int32), ("p_size", C.c_int32), ("q_size", C.c_int32),
                ("b_size", C.c_int32), ("n_size", C.c_int32), ("r_size", C.c_int32), ("s_size", C.c_int32),
                ("Ku", tf_csr), ("Ru", tf_csr), ("Mu", tf_csr), ("Kf", tf_csr), ("Rf", tf_csr), ("Tf", tf_csr),
                ("slot", i32p)]


class tf_dga_ops(C.Structure):
    _fields_ = [("C", tf_csr), ("Ct", tf_csr), ("bnd_edge", i32p), ("n_bnd", C.c_int32),
                ("edge_class", i8p), ("fixed_value", f64p),
                ("H", tf_csr), ("h_row", i32p), ("N", tf_csr),
                ("E", tf_csr), ("A", tf_csr), ("RHS1", tf_csr),
                ("free_edges", i32p), ("n_free", C.c_int32), ("Sig", tf_csr), ("frac", tf_dga_frac),
                ("edge_rank", i32p), ("face_rank", i32p)]


class tf_problem(C.Structure):
    _fields_ = [("method", C.c_int32), ("n_edges", C.c_int32), ("n_faces", C.c_int32),
                ("n_vols", C.c_int32), ("dt", C.c_double), ("duration", C.c_double),
                ("fit", tf_fit_ops), ("dga", tf_dga_ops)]


class tf_source(C.Structure):
    _fields_ = [("carrier", C.c_int32), ("field", C.c_int32), ("amp", C.c_double),
                ("freq", C.c_double), ("delay", C.c_double), ("width", C.c_double),
                ("impedance", C.c_double)]


class tf_source_term(C.Structure):
    _fields_ = [("edge", C.c_int32), ("source", C.c_int32), ("weight", C.c_double)]


class tf_probe(C.Structure):
    _fields_ = [("ne", C.c_int32), ("nh", C.c_int32),
                ("e_idx", C.c_int32 * 12), ("e_w", (C.c_double * 3) * 12),
                ("h_idx", C.c_int32 * 6), ("h_w", (C.c_double * 3) * 6)]


class tf_source_desc(C.Structure):
    _fields_ = [("type", C.c_int32), ("carrier", C.c_int32), ("profile", C.c_int32),
                ("dir", C.c_int32), ("bfunc", C.c_int32 * 3),
                ("amp", C.c_double), ("freq", C.c_double), ("delay", C.c_double),
                ("width", C.c_double), ("impedance", C.c_double), ("radius", C.c_double),
                ("decay", C.c_double), ("kvec", C.c_double * 3), ("center", C.c_double * 3),
                ("border", C.c_int32 * 3)]


class tf_source_edge(C.Structure):
    _fields_ = [("edge", C.c_int32), ("field", C.c_int32), ("n_src", C.c_int32),
                ("src", C.c_int32 * 4), ("bary", C.c_double * 3), ("vec", C.c_double * 3)]


class tf_probe_geom(C.Structure):
    _fields_ = [("elem", C.c_int32), ("edges", C.c_int32 * 12), ("faces", C.c_int32 * 6),
                ("point", C.c_double * 3), ("nodes", (C.c_double * 3) * 4), ("L", C.c_double * 3),
                ("mu", C.c_double * 9), ("volume", C.c_double)]


class tf_cg_opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int32), ("precond", C.c_int32),
                ("learn_steps", C.c_int32), ("train_iters", C.c_int32), ("lr", C.c_double),
                ("guess", C.c_int32)]


class tf_stats(C.Structure):
    _fields_ = [("steps_done", C.c_int64), ("cg_iters_total", C.c_int64),
                ("cg_iters_max", C.c_int32), ("cg_not_converged", C.c_int32),
                ("cg_max_relres", C.c_double), ("seconds", C.c_double),
                ("pc_state", C.c_int32), ("pc_pairs", C.c_int32), ("pc_fallbacks", C.c_int64),
                ("pc_learned_solves", C.c_int64), ("pc_learned_iters", C.c_int64),
                ("pc_jacobi_mean_iters", C.c_double), ("pc_train_seconds", C.c_double),
                ("pc_loss0", C.c_double), ("pc_loss", C.c_double)]


MLP_PARAMS = 5313


_host = None
_hip = None


def host_lib():
    """libtetfit_host.so with argtypes declared."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise RuntimeError(f"{HOST_LIB} is not built: run __graft_entry__.build() "
                           "or make -C tetfit_amd/csrc")
    L = C.CDLL(HOST_LIB)
    vp = C.c_void_p
    L.tfh_last_error.restype = C.c_char_p
    L.tf_model_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.tf_model_load.restype = C.c_int32
    L.tf_model_n_simulations.argtypes = [vp]
    L.tf_model_simulation_label.argtypes = [vp, C.c_int32]
    L.tf_model_free.argtypes = [vp]
    L.tf_model_free.restype = None
    L.tf_build_simulation.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.tf_build_free.argtypes = [vp]
    L.tf_build_free.restype = None
    L.tf_build_problem.argtypes = [vp]
    L.tf_build_problem.restype = C.POINTER(tf_problem)
    for name, rt in [("tf_build_possibly_unstable", C.c_int32), ("tf_build_stable_dt", C.c_double),
                     ("tf_build_n_steps", C.c_int64), ("tf_build_first_step", C.c_int64),
                     ("tf_build_cg_tol", C.c_double), ("tf_build_output_mode", C.c_int32),
                     ("tf_build_output_delay", C.c_double), ("tf_build_output_period", C.c_double),
                     ("tf_build_output_name", C.c_char_p), ("tf_build_output_frame", C.c_int32)]:
        f = getattr(L, name)
        f.argtypes = [vp]
        f.restype = rt
    L.tf_build_sources.argtypes = [vp, C.POINTER(C.POINTER(tf_source)), i32p,
                                   C.POINTER(C.POINTER(tf_source_term)), i64p]
    L.tf_build_sources.restype = None
    L.tf_build_probes.argtypes = [vp, C.POINTER(C.POINTER(tf_probe)), C.POINTER(f64p), i32p]
    L.tf_build_probes.restype = None
    L.tf_build_source_desc.argtypes = [vp, C.POINTER(C.POINTER(tf_source_desc)), i32p,
                                       C.POINTER(C.POINTER(tf_source_edge)), i64p]
    L.tf_build_source_desc.restype = None
    L.tf_build_probe_geom.argtypes = [vp, C.POINTER(C.POINTER(tf_probe_geom)), i32p]
    L.tf_build_probe_geom.restype = None
    L.tf_build_fit_tables.argtypes = [vp, C.POINTER(i32p), C.POINTER(i32p), C.POINTER(i32p),
                                      C.POINTER(f64p), i32p, i32p, C.POINTER(i32p), f64p]
    L.tf_build_fit_tables.restype = None
    L.tf_build_tet_tables.argtypes = [vp, C.POINTER(i32p), C.POINTER(f64p), C.POINTER(i32p),
                                      C.POINTER(i32p), C.POINTER(i32p), i32p, i32p,
                                      C.POINTER(i32p), C.POINTER(i32p)]
    L.tf_build_tet_tables.restype = None
    L.tf_write_kuhn_mesh.argtypes = [C.c_char_p, C.c_int32, C.c_double, C.c_uint64, C.c_int32, f64p]
    L.tf_write_kuhn_mesh.restype = C.c_int32
    _host = L
    return L


def hip_lib():
    """libtetfit_hip.so with argtypes declared."""
    global _hip
    if _hip is not None:
        return _hip
    if not os.path.exists(HIP_LIB):
        raise RuntimeError(f"{HIP_LIB} is not built: run __graft_entry__.build() "
                           "or make -C tetfit_amd/csrc")
    L = C.CDLL(HIP_LIB)
    vp = C.c_void_p
    L.tf_last_error.restype = C.c_char_p
    L.tf_device_count.restype = C.c_int32
    L.tf_create.argtypes = [C.POINTER(tf_problem), C.c_int32, C.POINTER(vp)]
    L.tf_set_sources.argtypes = [vp, C.POINTER(tf_source), C.c_int32, C.POINTER(tf_source_term), C.c_int64]
    L.tf_set_probes.argtypes = [vp, C.POINTER(tf_probe), C.c_int32]
    L.tf_set_output.argtypes = [vp, C.c_double, C.c_double]
    L.tf_run.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(tf_cg_opts), C.POINTER(tf_stats)]
    L.tf_read_state.argtypes = [vp, f64p, f64p, f64p]
    L.tf_write_state.argtypes = [vp, f64p, f64p, f64p]
    L.tf_read_probes.argtypes = [vp, f64p, f64p, C.c_int64, i64p]
    L.tf_read_poynting.argtypes = [vp, f64p, C.c_int64, i64p]
    L.tf_read_losses.argtypes = [vp, f64p, C.c_int64, i64p]
    L.tf_energy.argtypes = [vp, f64p, f64p]
    L.tf_set_energy_output.argtypes = [vp, C.c_int32]
    L.tf_read_energy.argtypes = [vp, f64p, f64p, C.c_int64, i64p]
    L.tf_kernel_times.argtypes = [vp, C.c_char_p, C.c_int32, f64p, i64p, C.c_int32, i32p]
    L.tf_set_profiling.argtypes = [vp, C.c_int32]
    L.tf_device_bytes.argtypes = [vp, i64p]
    L.tf_destroy.argtypes = [vp]
    f32p = C.POINTER(C.c_float)
    L.tf_mlp_grad.argtypes = [f32p, f32p, f32p, C.c_int64, f32p, f32p]
    L.tf_create_batch.argtypes = [C.POINTER(C.POINTER(tf_problem)), C.c_int32, C.c_int32, C.POINTER(vp)]
    L.tf_batch_set_sources.argtypes = [vp, C.c_int32, C.POINTER(tf_source), C.c_int32,
                                       C.POINTER(tf_source_term), C.c_int64]
    L.tf_batch_set_probes.argtypes = [vp, C.c_int32, C.POINTER(tf_probe), C.c_int32]
    L.tf_batch_set_output.argtypes = [vp, C.c_int64]
    L.tf_batch_read_probes.argtypes = [vp, C.c_int32, f64p, f64p, C.c_int64, i64p]
    L.tf_batch_read_state.argtypes = [vp, C.c_int32, f64p, f64p, f64p]
    L.tf_autotune.argtypes = [vp]
    L.tf_get_tuning.argtypes = [vp, i32p, i32p, i32p, C.c_int32, i32p]
    L.tf_sell_info.argtypes = [vp, C.c_int32, i64p]
    L.tf_fit_layout.argtypes = [vp, i32p]
    L.tf_set_tuning.argtypes = [vp, i32p, i32p, i32p, C.c_int32]
    L.tf_set_graphs.argtypes = [vp, C.c_int32]
    # device pointers (integers such as torch's data_ptr()) and a hipStream_t
    L.tf_mlp_grad_noalloc.argtypes = [vp, vp, vp, C.c_int64, vp, vp, vp]
    for n in ["tf_create", "tf_set_sources", "tf_set_probes", "tf_set_output", "tf_run",
              "tf_read_state", "tf_write_state", "tf_read_probes", "tf_read_poynting", "tf_read_losses",
              "tf_energy", "tf_set_energy_output", "tf_read_energy", "tf_kernel_times", "tf_set_profiling", "tf_device_bytes", "tf_destroy",
              "tf_mlp_grad", "tf_create_batch", "tf_batch_set_sources", "tf_batch_set_probes",
              "tf_batch_set_output", "tf_batch_read_probes", "tf_batch_read_state", "tf_autotune",
              "tf_get_tuning", "tf_sell_info", "tf_set_tuning", "tf_set_graphs", "tf_fit_layout",
              "tf_mlp_grad_noalloc"]:
        getattr(L, n).restype = C.c_int32
    _hip = L
    return L
