"""Benchmark of the TetFIT time-stepping hot path on one MI355X.

Default workload (the metric's configuration): the 96^3-cube Kuhn
tetrahedral cavity (6,276,384 edges; vertices jittered +-0.2h with seed 5;
eps_r piecewise constant per 6^3-cube block, log-uniform in [1,80]; mu_r =
1; PEC on 5 walls; wave-profile h source on x-min at 2e8 Hz), stepped with
the CG-solved DGA leapfrog (method gfdtd, CG tolerance 1e-8, maxiter 100)
for 3000 steps from a zero field, once on the plain FP64 path (Jacobi CG)
and once on the learned+FP32 path (fp32 fields and operators, fp64 CG
reductions, the learned sparse-approximate-inverse preconditioner trained on
the device from the first 200 solves).

A step is one leapfrog step over the whole mesh. value = edge-DoF updates/s
of the learned+FP32 path over the whole job = n_edges * steps / wall seconds,
its one-time cost (preconditioner set-up, the 200 collection solves with
Jacobi CG, training) included; inputs are resident in HBM when the job
starts. Beside it: the plain FP64 rate over the same job, both steady-state
rates, CG iterations/step, the fp32 field error against the plain FP64 run,
the fp64 residual of the learned path's last solve, kernel rooflines, the
CPU baseline on the same mesh and the 48^3 legs. Other workloads: --workload
gfdtd (48^3 steady-state legs), contrast (96^3 steady-state legs with the
explicit dga path), batch, fit, fitprec, copy.

Prints one JSON line. --gpus N with N > 1 exits: this build drives one GPU.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "edge-DoF updates/s + CG iters/step: learned+FP32 vs plain FP64, 1 GPU"
# dt pinned with SET step: 0.9 x the host stability estimate of each method on
# these meshes (power iteration; see tetfit_amd/configs.py)
DT48 = {"dga": 2.281e-11, "gfdtd": 4.777e-11}
DT96_CONTRAST = {"gfdtd": 2.264e-11}
# CG start of the fast (learned+FP32) path's solves: quadratic extrapolation
# 3U_n - 3U_{n-1} + U_{n-2} (tf_cg_opts.guess); the stop test stays the
# reference's, relative to the warm-start residual. Plain FP64 starts warm.
FAST_GUESS = 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(d, method, n, steps, contrast=False):
    """The 48^3 workload (block 8, seed 2, eps_r U[1,10]); contrast=True is the
    high-contrast variant (6^3-cube blocks, seed 5, eps_r log-uniform [1,80])."""
    from tetfit_amd import Model, configs
    log(f"[bench] building {method} {n}^3{' (contrast)' if contrast else ''}")
    if contrast:
        # dga: dt from the host's stability estimate within the same build
        # (one host build: the explicit operator assembly at 96^3 takes minutes)
        dt = "courant" if method == "dga" else (DT96_CONTRAST.get(method) if n == 96 else None)
        path = configs.kuhn_cavity(d, n=n, method=method, n_steps=steps, n_probes=8, block=6, seed=5,
                                   eps_range=(1.0, 80.0), log_eps=True, name=f"kuhn{n}_{method}_contrast", dt=dt)
    else:
        path = configs.kuhn_cavity(d, n=n, method=method, n_steps=steps, n_probes=8, block=8, seed=2,
                                   dt=DT48.get(method) if n == 48 else None)
    t = time.time()
    b = Model(path).build(1)
    log(f"[bench] built {method} {n}^3: E={b.n_edges} F={b.n_faces} T={b.n_vols} dt={b.dt:.4g} "
        f"({time.time() - t:.1f}s host set-up)")
    return b


def slice_bytes(info):
    """Per-slice header bytes the SELL kernels read: the 8-byte slice
    pointer and 4 bytes per column base (one, four, or none with int32
    columns)."""
    return (8 + 4 * info["col_bases"]) * info["slices"]


def algorithmic_bytes(info, kernel, vb):
    """Bytes a CG kernel must move per launch: the operator's non-zeros once
    (values + stored column index; padding entries are not counted), the
    slice pointer and column bases of each 64-row slice, and each vector once
    per access (gathered vectors once). `info` is the device storage of A
    (Simulation.sell_info("a")), vb the value bytes of the vectors."""
    n, nnz, cb = info["rows"], info["nnz"], info["col_bytes"]
    per_slice = slice_bytes(info)
    if kernel == "cg_w_kernel":      # w = A u: u gathered, w written
        return nnz * (vb + cb) + per_slice + 2 * vb * n
    if kernel == "cg_wu_kernel":     # w = A u, p and s updated: u gathered, p, s read and written
        return nnz * (vb + cb) + per_slice + 5 * vb * n
    if kernel == "cg_pg_kernel":     # u = G r' (bf16 G): r, s gathered, x r/w, p read, r', u written
        return nnz * (2 + cb) + per_slice + 7 * vb * n
    raise KeyError(kernel)


def dga_bytes(info, kernel, vb, n_edges, n_faces):
    """Algorithmic bytes per launch of the explicit dga SELL kernels, same
    rules: `info` is the storage of the kernel's operator (sell_info "h" or
    "n"). sell_h: U += dt H curl_f (curl_f gathered over all edges, a row ->
    edge index, U read and written per row); sell_n: F -= dt N curl_u (curl_u
    gathered, F read and written)."""
    n, nnz, cb = info["rows"], info["nnz"], info["col_bytes"]
    per_slice = slice_bytes(info)
    if kernel == "sell_h_kernel":
        return nnz * (vb + cb) + per_slice + vb * n_edges + 4 * n + 2 * vb * n
    if kernel == "sell_n_kernel":
        return nnz * (vb + cb) + per_slice + vb * n_faces + 2 * vb * n
    raise KeyError(kernel)


def measured_traffic(kernel, vb, nbytes):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE scaled by a
    copy of known size), or None when no measurement exists for the same
    kernel, value type and storage (same algorithmic bytes)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            table = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    t = "double" if vb == 8 else "float"
    for e in table:
        if e["kernel"] == kernel and e["type"] == t and e["algorithmic_bytes"] == int(nbytes):
            return e["traffic_bytes"]
    return None


def roofline_of(info, kernels, names, vb):
    """Roofline entries of the listed CG kernels measured in a run; the
    first is the one with the largest total time."""
    out = []
    for k in names:
        t = kernels.get(k)
        if not t:
            continue
        nb = algorithmic_bytes(info, k, vb)
        ach = nb / (t["ms_avg"] * 1e-3) / 1e9
        out.append(dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(ach / HBM_PEAK_GBS, 4), traffic=measured_traffic(k, vb, nb), kernel=k,
                        bytes_per_launch=int(nb), ms_per_launch=t["ms_avg"], launches=t["launches"]))
    out.sort(key=lambda r: -r["ms_per_launch"] * r["launches"])
    return out


def time_path(b, precision, steps, warmup, precond="jacobi", label="", learn_steps=200):
    """Times `steps` steps after `warmup` untimed ones. For the learned
    preconditioner the context first runs its `learn_steps` Jacobi-PCG
    collection steps and the on-device training (untimed, reported apart);
    its solves start from the quadratic extrapolation (FAST_GUESS)."""
    from tetfit_amd.device import Simulation
    sim = Simulation(b, precision, output=False)
    learn = None
    g = FAST_GUESS if precond == "learned" else 0
    if precond == "learned":
        ls = sim.run(n=learn_steps, precond="learned", learn_steps=learn_steps, guess=g)
        learn = dict(learn_steps=learn_steps, collection_seconds=ls["seconds"] - ls["pc_train_seconds"],
                     train_seconds=ls["pc_train_seconds"], pairs=ls["pc_pairs"], loss0=ls["pc_loss0"],
                     loss=ls["pc_loss"], jacobi_iters_per_step=ls["pc_jacobi_mean_iters"], state=ls["pc_state"])
        log(f"[bench] {label}: trained on {ls['pc_pairs']} pairs in {ls['pc_train_seconds']:.2f} s, "
            f"loss {ls['pc_loss0']:.4g} -> {ls['pc_loss']:.4g}")
    tuned = sim.autotune()
    log(f"[bench] {label}: launch shapes {tuned}")
    if warmup:
        sim.run(n=warmup, precond=precond, guess=g)
    st = sim.run(n=steps, precond=precond, guess=g)
    sec = st["seconds"]
    out = dict(seconds=sec, ms_per_step=1e3 * sec / steps,
               edge_dof_per_s=b.n_edges * steps / sec,
               cg_iters_per_step=st["cg_iters_total"] / steps if b.method == "gfdtd" else None,
               cg_max_relres=st["cg_max_relres"] if b.method == "gfdtd" else None,
               cg_not_converged=st["cg_not_converged"] if b.method == "gfdtd" else None)
    out["tuning"] = tuned
    if b.method == "gfdtd":
        out["A"] = sim.sell_info("a")
    if b.method == "dga":
        out["H"], out["N"] = sim.sell_info("h"), sim.sell_info("n")
    if learn is not None:
        learn.update(state_after=st["pc_state"], fallbacks=st["pc_fallbacks"])
        out["learned"] = learn
    # per-kernel durations with HIP events on the context's stream
    sim.set_profiling(True)
    sim.run(n=min(steps, 50), precond=precond, guess=g)
    out["kernels"] = {k: {"ms_avg": v[0], "launches": v[1]} for k, v in sim.kernel_times().items()}
    sim.close()
    log(f"[bench] {label}: {out['ms_per_step']:.3f} ms/step, {out['edge_dof_per_s'] / 1e9:.3f} G edge-DoF/s"
        + (f", CG {out['cg_iters_per_step']:.2f} it/step" if out["cg_iters_per_step"] is not None else ""))
    return out


def cpu_baseline(b, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    r = pyoracle.run(b, n=steps, faithful=True, output=False)
    return dict(value=b.n_edges * steps / r["seconds"], unit="edge-DoF/s", cores=1, kind="port",
                sample=f"{steps} steps of method {b.method} on the same mesh, single-threaded fp64 CPU "
                       f"restatement of the reference loop (oracle/, -O3), {r['seconds']:.2f} s",
                cg_iters_per_step=(r["cg_iters_total"] / steps) if b.method == "gfdtd" else None,
                host=os.uname().nodename, nproc=os.cpu_count())


def batch_workload(d, count=256, steps=2000, seed=4, workers=16, method="dga"):
    """Batched workload: `count` Kuhn cavities with n ~ U{8..32}, jitter
    0.2h, one eps_r ~ U[1, 10] and one source frequency ~ U[1e8, 5e8] Hz per
    instance (numpy default_rng(seed)); dt = 0.9 x each instance's own
    stability estimate; method dga (explicit) or gfdtd (CG tol 1e-8). Host
    builds run in a thread pool."""
    from concurrent.futures import ThreadPoolExecutor
    from tetfit_amd import Model, configs
    rng = np.random.default_rng(seed)
    ns = rng.integers(8, 33, size=count)
    eps = rng.uniform(1.0, 10.0, size=count)
    freq = rng.uniform(1e8, 5e8, size=count)

    def one(k):
        p = configs.kuhn_cavity(d, n=int(ns[k]), method=method, n_steps=steps, n_probes=1, block=4,
                                seed=1000 + k, eps_range=(float(eps[k]), float(eps[k])), freq=float(freq[k]),
                                dt="courant", name=f"inst{k}")
        return Model(p).build(1)

    t = time.time()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        builds = list(ex.map(one, range(count)))
    log(f"[bench] built {count} {method} instances, n in [{ns.min()}, {ns.max()}], "
        f"sum E = {sum(b.n_edges for b in builds)}, sum F = {sum(b.n_faces for b in builds)} "
        f"({time.time() - t:.1f}s host set-up)")
    return builds


def fit_bytes(b, kernel, vb, grid_order=True):
    """Algorithmic bytes per launch of the FIT kernels: E phase reads U, ca,
    cb and the per-edge code, writes U, gathers F once; H phase reads B, cm,
    cn and the per-face code, writes B and F, gathers U once. The code is
    2 bytes (slot order and signs) in grid order, 17 bytes (packed faces or
    edges and a sign byte) in the index-table layout."""
    E, F = b.n_edges, b.n_faces
    cb = 2 if grid_order else 17
    if kernel == "fit_e_kernel":
        return E * (4 * vb + cb) + F * vb
    return F * (5 * vb + cb) + E * vb


def run_chunks(sim, steps, chunk, label, precond="jacobi", learn_steps=200, stop_before_last=False, guess=0):
    """Advances `steps` steps in tf_run calls of at most `chunk` steps (one
    progress line each), wall-clocked around all calls (each returns after
    its stream has drained). The learned path's first call covers exactly the
    collection solves, so the preconditioner set-up, collection and training
    fall inside the timed region; the launch shapes of the learned passes are
    then fetched (autotune, a process-cache hit after the warm-up context).
    With stop_before_last the last step is left for the caller. Returns
    (seconds, stats of the calls, seconds of the last full chunk, its steps)."""
    out = dict(cg_iters_total=0, cg_not_converged=0, cg_max_relres=0.0, pc=None, chunks=[])
    todo = steps - (1 if stop_before_last else 0)
    t0 = time.perf_counter()
    done = 0
    while done < todo:
        k = min(chunk, todo - done)
        if precond == "learned" and done == 0:
            k = min(learn_steps, todo)
        tc = time.perf_counter()
        st = sim.run(n=k, precond=precond, learn_steps=learn_steps, guess=guess)
        if precond == "learned" and done == 0:
            out["pc"] = st
            out["shapes"] = sim.autotune()
            log(f"[bench] {label}: launch shapes {out['shapes']}")
        dtc = time.perf_counter() - tc
        done += k
        out["cg_iters_total"] += st["cg_iters_total"]
        out["cg_not_converged"] += st["cg_not_converged"]
        out["cg_max_relres"] = max(out["cg_max_relres"], st["cg_max_relres"])
        out["chunks"].append((k, dtc, st["cg_iters_total"]))
        log(f"[bench] {label}: {done}/{steps} steps, {1e3 * dtc / k:.3f} ms/step over the last {k}"
            + (f", CG {st['cg_iters_total'] / k:.2f} it/step" if st["cg_iters_total"] else ""))
    return time.perf_counter() - t0, out


def true_residual(b, sim64, state_before, first_step, x_path):
    """fp64 residual of the last solve of a run: the state before that step is
    written into the fp64 context, which redoes the step with CG at tol 1e-13
    (x*); then, with the host-assembled fp64 A on the free edges,
    ||A (x* - x)|| / ||A (x* - x_old)|| = ||b - A x|| / ||b - A x_old||, the
    reference's stopping ratio evaluated in fp64 for the path's solution x."""
    import scipy.sparse as sp
    U0, F0, B0 = state_before
    sim64.set_state(U0, F0, B0)
    st = sim64.run(n=1, first=first_step, cg_tol=1e-13, cg_maxiter=5000)
    Ustar, _, _ = sim64.state()
    ops = b.dga_ops()
    A = ops["A"]
    free = ops["free_edges"]
    Am = sp.csr_matrix((A["val"], A["col"], A["rowptr"]), shape=(A["nrows"], A["ncols"]))
    r = Am @ (Ustar[free] - x_path[free])
    r0 = Am @ (Ustar[free] - U0[free])
    return dict(fp64_relres=float(np.linalg.norm(r) / max(np.linalg.norm(r0), 1e-300)),
                reference_solve_relres=st["cg_max_relres"], reference_solve_iters=st["cg_iters_total"])


def rel_l2(a, r):
    return float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300))


def bench_headline(args):
    """The metric on its configuration: the 96^3 Kuhn cavity with eps_r
    log-uniform in [1, 80] per 6^3 block (seed 5), method gfdtd, CG tol 1e-8,
    3000 steps. Plain FP64 (Jacobi CG) and learned+FP32 (200 collection
    solves, on-device training, then the learned preconditioner) each run the
    whole job from a zero field; value is the learned path's end-to-end
    edge-DoF/s over all steps, set-up of the preconditioner, collection and
    training included. A warm-up context of each path runs first (and fills
    the autotuner's process cache) and is discarded."""
    from tetfit_amd.device import Simulation
    n = args.n if args.n != 48 else 96
    # the job is the configuration's 3000 steps (the learned path's one-time
    # cost is amortised over it); a longer --steps lengthens it
    steps = max(args.steps, 3000) if args.steps_given else 3000
    chunk = 500
    with tempfile.TemporaryDirectory() as d:
        bg = build_workload(d, "gfdtd", n, steps + 60, contrast=True)
        E = bg.n_edges
        # warm-up contexts: tuning cache, code paths, first-touch allocations
        for prec, pc in (("fp64", "jacobi"), ("fp32", "learned")):
            sim = Simulation(bg, prec, output=False)
            sim.autotune()
            if pc == "learned":
                sim.run(n=args.learn_steps, precond="learned", learn_steps=args.learn_steps, guess=FAST_GUESS)
                sim.autotune()
            sim.run(n=max(args.warmup, 1), precond=pc, guess=FAST_GUESS if pc == "learned" else 0)
            sim.close()
            log(f"[bench] warm-up {prec} {pc} done")
        # plain FP64, whole job
        p64 = Simulation(bg, "fp64")
        log(f"[bench] plain fp64 launch shapes {p64.autotune()}")
        sec64, r64 = run_chunks(p64, steps, chunk, "plain fp64 (Jacobi)")
        U64, F64, B64 = p64.state()
        _, s64 = p64.probes()
        # learned + FP32, whole job (the last step taken apart for the residual check)
        l32 = Simulation(bg, "fp32")
        l32.autotune()
        sec32a, r32 = run_chunks(l32, steps, chunk, "learned fp32", precond="learned", guess=FAST_GUESS,
                                 learn_steps=args.learn_steps, stop_before_last=True)
        before = l32.state()
        t = time.perf_counter()
        st_last = l32.run(n=1, precond="learned", guess=FAST_GUESS)
        sec32 = sec32a + (time.perf_counter() - t)
        r32["cg_iters_total"] += st_last["cg_iters_total"]
        r32["cg_max_relres"] = max(r32["cg_max_relres"], st_last["cg_max_relres"])
        r32["cg_not_converged"] += st_last["cg_not_converged"]
        U32, F32, B32 = l32.state()
        _, s32 = l32.probes()
        pc = r32["pc"]
        # steady state: the learned path's chunks after the collection call
        ss = [c for c in r32["chunks"][1:]]
        ss_steps = sum(c[0] for c in ss)
        ss_sec = sum(c[1] for c in ss)
        ss_it = sum(c[2] for c in ss)
        p_steps = sum(c[0] for c in r64["chunks"][1:]) or steps
        p_sec = sum(c[1] for c in r64["chunks"][1:]) or sec64
        # per-kernel times (HIP events on the context's stream) over 30 more steps
        first_prof = l32.next_step
        l32.set_profiling(True)
        l32.run(n=30, precond="learned", guess=FAST_GUESS)
        kern32 = {k: {"ms_avg": v[0], "launches": v[1]} for k, v in l32.kernel_times().items()}
        info_a32 = l32.sell_info("a")
        l32.close()
        log("[bench] fp64 check of the learned path's last solve")
        resid = true_residual(bg, p64, before, bg.first_step + steps - 1, U32)
        p64.set_profiling(True)
        p64.run(n=10)
        kern64 = {k: {"ms_avg": v[0], "launches": v[1]} for k, v in p64.kernel_times().items()}
        info_a64 = p64.sell_info("a")
        p64.close()
        del first_prof
        nprobe = s64.shape[1]
        per_probe = [max(rel_l2(s32[:, p, :3], s64[:, p, :3]), rel_l2(s32[:, p, 3:], s64[:, p, 3:]))
                     for p in range(nprobe)]
        fp32_error = {"vs": "plain FP64 GPU path, same steps", "steps": steps,
                      "U": rel_l2(U32, U64), "F": rel_l2(F32, F64), "B": rel_l2(B32, B64),
                      "per_probe_max": max(per_probe), "per_probe": per_probe,
                      "last_solve_fp64_relres": resid["fp64_relres"],
                      "last_solve_device_relres": st_last["cg_max_relres"],
                      "reference_solve": {"tol": 1e-13, "relres": resid["reference_solve_relres"],
                                          "iters": resid["reference_solve_iters"]}}
        log(f"[bench] fp32 learned vs fp64 plain: U {fp32_error['U']:.3e} F {fp32_error['F']:.3e} "
            f"B {fp32_error['B']:.3e}, probes <= {fp32_error['per_probe_max']:.3e}; last solve fp64 "
            f"relres {resid['fp64_relres']:.3e}")
        value = E * steps / sec32
        plain_rate = E * steps / sec64
        roofs = roofline_of(info_a32, kern32, ["cg_wu_kernel", "cg_pg_kernel"], 4)
        plain_roofs = roofline_of(info_a64, kern64, ["cg_w_kernel"], 8)
        amortized = {
            "steps": steps, "learned_fp32_seconds": sec32, "plain_fp64_seconds": sec64,
            "learned_fp32_edge_dof_per_s": value, "plain_fp64_edge_dof_per_s": plain_rate,
            "speedup": value / plain_rate,
            "learned_one_time": {"collection_steps": args.learn_steps,
                                 "first_call_seconds": r32["chunks"][0][1],
                                 "train_seconds": pc["pc_train_seconds"],
                                 "collection_seconds": r32["chunks"][0][1] - pc["pc_train_seconds"],
                                 "jacobi_iters_per_step": pc["pc_jacobi_mean_iters"],
                                 "loss0": pc["pc_loss0"], "loss": pc["pc_loss"], "pairs": pc["pc_pairs"]},
            "learned_state_end": st_last.get("pc_state"), "learned_fallbacks": st_last.get("pc_fallbacks"),
        }
        steady = {
            "learned_fp32": {"steps": ss_steps, "ms_per_step": 1e3 * ss_sec / max(ss_steps, 1),
                             "edge_dof_per_s": E * ss_steps / ss_sec if ss_sec else None,
                             "cg_iters_per_step": ss_it / max(ss_steps, 1)},
            "plain_fp64": {"steps": p_steps, "ms_per_step": 1e3 * p_sec / p_steps,
                           "edge_dof_per_s": E * p_steps / p_sec,
                           "cg_iters_per_step": sum(c[2] for c in r64["chunks"][1:]) / p_steps
                           if len(r64["chunks"]) > 1 else r64["cg_iters_total"] / steps},
        }
        steady["speedup"] = (steady["learned_fp32"]["edge_dof_per_s"] / steady["plain_fp64"]["edge_dof_per_s"]
                             if steady["learned_fp32"]["edge_dof_per_s"] else None)
        cpu = None
        if not args.no_cpu:
            t = time.time()
            cpu = cpu_baseline(bg, args.cpu_steps)
            log(f"[bench] cpu baseline {cpu['value'] / 1e6:.2f} M edge-DoF/s ({time.time() - t:.1f}s)")
        legs48 = None
        if not args.no_48:
            b48 = build_workload(d, "gfdtd", 48, 400 + args.learn_steps, False)
            legs48 = {}
            for key, prec, pcn in (("plain_fp64", "fp64", "jacobi"), ("learned_fp32", "fp32", "learned")):
                r = time_path(b48, prec, 200, 20, pcn, f"48^3 gfdtd {prec} {pcn}", learn_steps=args.learn_steps)
                legs48[key] = {"edge_dof_per_s": r["edge_dof_per_s"], "ms_per_step": r["ms_per_step"],
                               "cg_iters_per_step": r["cg_iters_per_step"]}
                if r.get("learned"):
                    legs48[key]["learned"] = r["learned"]
            legs48["workload"] = "Kuhn 48^3, eps_r U[1,10] per 8^3 block (seed 2), gfdtd tol 1e-8, 200 steps steady state"
            legs48["speedup"] = legs48["learned_fp32"]["edge_dof_per_s"] / legs48["plain_fp64"]["edge_dof_per_s"]
        line = {
            "metric": METRIC, "value": value, "unit": "edge-DoF/s", "n_gpus": 1, "steps": steps,
            "requested_steps": args.steps if args.steps_given else None,
            "warmup": args.warmup, "ms_per_step": 1e3 * sec32 / steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"Kuhn {n}^3 jittered tet cavity (seed 5), eps_r log-uniform [1,80] per 6^3 "
                                   f"block, method gfdtd (CG tol 1e-8), {steps} steps end to end",
                       "path": "learned+FP32 (200 collection solves, training and set-up inside the timed job; "
                               "CG starts from the quadratic extrapolation of the previous steps, stop test relative "
                               "to the warm-start residual as in the reference)",
                       "cg_start": {"learned_fp32": "quadratic extrapolation (collection solves warm)",
                                    "plain_fp64": "warm start U_n (the reference)"},
                       "n_edges": E, "n_faces": bg.n_faces, "parallelism": "single GPU"},
            "cg_iters_per_step": r32["cg_iters_total"] / steps,
            "plain_fp64": {"edge_dof_per_s": plain_rate, "ms_per_step": 1e3 * sec64 / steps,
                           "cg_iters_per_step": r64["cg_iters_total"] / steps,
                           "cg_not_converged": r64["cg_not_converged"], "cg_max_relres": r64["cg_max_relres"]},
            "speedup_vs_plain_fp64": value / plain_rate,
            "amortized": amortized, "steady_state": steady, "fp32_error": fp32_error,
            "learned_cg": {"cg_not_converged": r32["cg_not_converged"], "cg_max_relres": r32["cg_max_relres"]},
            "roofline": roofs[0] if roofs else None, "roofline_kernels": roofs,
            "plain_fp64_roofline": plain_roofs[0] if plain_roofs else None,
            "A_storage": info_a32, "cpu_baseline": cpu, "legs_48": legs48, "kernels": kern32,
        }
        print(json.dumps(line), flush=True)


def bench_fit(args):
    """FIT/FDTD cavity: 64^3 cells of 0.0156 m (xgrid {0,0.0156,1}), PEC on
    5 walls, wave-profile h source on x-min, 32 probes; fp64 and fp32."""
    from tetfit_amd import Model, configs
    from tetfit_amd.device import Simulation
    steps = args.steps if args.steps_given else 2000
    with tempfile.TemporaryDirectory() as d:
        t = time.time()
        b = Model(configs.fit_cavity(d, n=64, step=0.0156, n_steps=steps + args.warmup + 60, n_probes=32)).build(1)
        log(f"[bench] built fdtd 64^3: E={b.n_edges} F={b.n_faces} dt={b.dt:.4g} ({time.time() - t:.1f}s)")
        res = {}
        for prec in ("fp64", "fp32"):
            sim = Simulation(b, prec, output=False)
            grid_order = sim.fit_grid_order()
            sim.run(n=args.warmup)
            st = sim.run(n=steps)
            sim.set_profiling(True)
            sim.run(n=50)
            kern = {k: {"ms_avg": v[0], "launches": v[1]} for k, v in sim.kernel_times().items()}
            sim.close()
            res[prec] = dict(ms_per_step=1e3 * st["seconds"] / steps, edge_dof_per_s=b.n_edges * steps / st["seconds"],
                             kernels=kern)
            log(f"[bench] fdtd {prec}: {res[prec]['ms_per_step']:.3f} ms/step, "
                f"{res[prec]['edge_dof_per_s'] / 1e9:.3f} G edge-DoF/s")
        roof = {}
        for kname in ("fit_e_kernel", "fit_h_kernel"):
            k = res["fp64"]["kernels"].get(kname)
            if k:
                nb = fit_bytes(b, kname, 8, grid_order)
                ach = nb / (k["ms_avg"] * 1e-3) / 1e9
                roof[kname] = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                                   frac=round(ach / HBM_PEAK_GBS, 4), traffic=measured_traffic(kname, 8, nb),
                                   kernel=kname, bytes_per_launch=int(nb), ms_per_launch=k["ms_avg"])
        cpu = None
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle
            r = pyoracle.run(b, n=args.cpu_steps * 4, faithful=True, output=False)
            cpu = dict(value=b.n_edges * args.cpu_steps * 4 / r["seconds"], unit="edge-DoF/s", cores=1, kind="port",
                       sample=f"{args.cpu_steps * 4} fdtd steps, loop-faithful single-thread fp64 restatement, "
                              f"{r['seconds']:.2f} s")
        line = {"metric": "edge-DoF updates/s, FIT/FDTD 64^3, 1 GPU", "value": res["fp64"]["edge_dof_per_s"],
                "unit": "edge-DoF/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
                "ms_per_step": res["fp64"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                "config": {"workload": "FIT 64^3 cavity, 0.0156 m cells, PEC on 5 walls, h source on x-min",
                           "n_edges": b.n_edges, "n_faces": b.n_faces},
                "fp32": {"edge_dof_per_s": res["fp32"]["edge_dof_per_s"], "ms_per_step": res["fp32"]["ms_per_step"]},
                "roofline": roof.get("fit_e_kernel"), "roofline_h": roof.get("fit_h_kernel"), "cpu_baseline": cpu,
                "kernels": res["fp64"]["kernels"]}
        print(json.dumps(line), flush=True)


def bench_fitprec(args):
    """Precision study: FIT n^3 random media (eps_r U[1,4) per cell, seed 3),
    sigma 0.05 S/m slab 0.4 <= x < 0.6, gaussmod carrier at 3e8 Hz, 16
    probes; the same steps in fp32 and fp64 on the GPU; errors of fp32 vs the
    fp64 run per probe (relative L2 over time), globally over U, F, B at the
    end, and of the fp64-reduced energy series."""
    from tetfit_amd import Model, configs
    from tetfit_amd.device import Simulation
    steps = args.steps if args.steps_given else 4000
    n = args.n if args.n != 48 else 256
    chunk = 100
    with tempfile.TemporaryDirectory() as d:
        t = time.time()
        log(f"[bench] writing the fdtd random media {n}^3 input")
        path = configs.fit_random_media(d, n=n, n_steps=steps, n_probes=16)
        log(f"[bench] input written ({time.time() - t:.1f}s); meshing and assembling on the host")
        b = Model(path).build(1)
        log(f"[bench] built fdtd random media {n}^3: E={b.n_edges} F={b.n_faces} dt={b.dt:.4g} "
            f"({time.time() - t:.1f}s host set-up)")
        out = {}
        for prec in ("fp64", "fp32"):
            sim = Simulation(b, prec)
            secs, energy = 0.0, []
            done = 0
            while done < steps:
                k = min(chunk, steps - done)
                st = sim.run(n=k)
                secs += st["seconds"]
                done += k
                energy.append(sim.energy())
            _, samples = sim.probes()
            U, F, B = sim.state()
            sim.close()
            out[prec] = dict(samples=samples, U=U, F=F, B=B, energy=np.array(energy), seconds=secs)
            log(f"[bench] fitprec {prec}: {1e3 * secs / steps:.3f} ms/step")
        ref, got = out["fp64"], out["fp32"]
        rl2 = lambda a, r: float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300))
        # probes the wave has not reached carry only rounding noise: a probe
        # counts when its fp64 signal (L2 over time) is at least 1e-6 of the
        # strongest probe's; the others are reported as null
        norms = [float(np.linalg.norm(ref["samples"][:, p])) for p in range(ref["samples"].shape[1])]
        floor = 1e-6 * max(norms + [0.0])
        per_probe = [rl2(got["samples"][:, p], ref["samples"][:, p]) if norms[p] > floor and norms[p] > 0 else None
                     for p in range(len(norms))]
        reached = [v for v in per_probe if v is not None]
        W64 = ref["energy"].sum(axis=1)
        W32 = got["energy"].sum(axis=1)
        line = {"metric": "fp32 vs fp64 relative L2 error, FIT random media", "value": max(reached) if reached else None,
                "unit": "relative L2 (max over the probes the wave reached)", "n_gpus": 1, "steps": steps, "warmup": 0,
                "ms_per_step": 1e3 * got["seconds"] / steps, "higher_is_better": False, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": f"FIT {n}^3 random media, eps_r U[1,4) per cell, sigma 0.05 slab, gaussmod 3e8 Hz",
                           "n_edges": b.n_edges, "n_faces": b.n_faces},
                "per_probe_rel_l2": per_probe, "probes_reached": len(reached), "probe_signal_floor": floor,
                "global_rel_l2": {"U": rl2(got["U"], ref["U"]), "F": rl2(got["F"], ref["F"]),
                                                                 "B": rl2(got["B"], ref["B"])},
                "energy_rel_diff_max": float(np.abs(W32 - W64).max() / max(np.abs(W64).max(), 1e-300)),
                "energy_fp64_final": float(W64[-1]), "energy_fp64_max": float(W64.max()),
                "edge_dof_per_s": {"fp32": b.n_edges * steps / got["seconds"], "fp64": b.n_edges * steps / ref["seconds"]},
                "roofline": None, "cpu_baseline": None}
        print(json.dumps(line), flush=True)


def bench_copy(args):
    """1 GiB device-to-device copy with torch (one elementwise copy kernel of
    known size): the reference for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({"workload": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "ms": ms,
                      "GBps": 8 * n / (ms * 1e-3) / 1e9}), flush=True)


def bench_batch(args):
    from tetfit_amd.device import BatchSimulation
    gf = args.batch_method == "gfdtd"
    steps = args.steps if args.steps_given else (500 if gf else 2000)
    with tempfile.TemporaryDirectory() as d:
        builds = batch_workload(d, count=args.batch, steps=steps, method=args.batch_method)
        E = sum(b.n_edges for b in builds)
        res = {}
        for prec in ("fp32", "fp64"):
            sim = BatchSimulation(builds, prec)
            tuned = sim.autotune()
            log(f"[bench] batch {prec} tuning: {tuned}")
            if args.warmup:
                sim.run(args.warmup)
            st = sim.run(steps)
            sim.set_profiling(True)
            sim.run(min(steps, 50))
            kern = {k: {"ms_avg": v[0], "launches": v[1]} for k, v in sim.kernel_times().items()}
            sim.close()
            res[prec] = dict(seconds=st["seconds"], ms_per_step=1e3 * st["seconds"] / steps,
                             edge_dof_per_s=E * steps / st["seconds"], kernels=kern, tuning=tuned)
            if gf:
                res[prec].update(cg_iters_per_step_per_instance=st["cg_iters_total"] / steps / len(builds),
                                 cg_iters_max=st["cg_iters_max"], cg_not_converged=st["cg_not_converged"],
                                 cg_max_relres=st["cg_max_relres"])
            log(f"[bench] batch {prec}: {res[prec]['ms_per_step']:.3f} ms/step, "
                f"{res[prec]['edge_dof_per_s'] / 1e9:.3f} G edge-DoF/s"
                + (f", CG {res[prec]['cg_iters_per_step_per_instance']:.2f} it/step/instance" if gf else ""))
        if gf:
            head = res["fp32"]
            line = {
                "metric": "edge-DoF updates/s, batched gfdtd (all instances), 1 GPU",
                "value": head["edge_dof_per_s"], "unit": "edge-DoF/s", "n_gpus": 1, "steps": steps,
                "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                "config": {"workload": f"{len(builds)} Kuhn gfdtd cavities, n~U{{8..32}}, eps~U[1,10], "
                                       f"f~U[1e8,5e8] Hz, CG tol 1e-8 (Jacobi, one solve per instance)",
                           "n_edges_total": E, "n_faces_total": sum(b.n_faces for b in builds),
                           "parallelism": "single GPU, one launch per phase and CG pass for the whole batch"},
                "fp64": {k: res["fp64"][k] for k in ("edge_dof_per_s", "ms_per_step", "cg_iters_per_step_per_instance",
                                                     "cg_not_converged", "cg_max_relres")},
                "cg_iters_per_step_per_instance": head["cg_iters_per_step_per_instance"],
                "cg_not_converged": head["cg_not_converged"], "cg_max_relres": head["cg_max_relres"],
                "roofline": None, "cpu_baseline": None, "kernels": head["kernels"], "tuning": head["tuning"],
            }
            print(json.dumps(line), flush=True)
            return
        head = res["fp32"]
        nnz_h = sum(int(b.dga_ops()["H"]["rowptr"][-1]) for b in builds)
        n_h = sum(int(b.dga_ops()["H"]["nrows"]) for b in builds)
        kh = head["kernels"].get("sell_h_kernel")
        roof = None
        if kh:
            # H values+cols, row->edge map, row dt, curl_f gathered, U read+write
            nbytes = nnz_h * 8 + 4 * n_h + 8 * n_h + 4 * E + 2 * 4 * n_h
            ach = nbytes / (kh["ms_avg"] * 1e-3) / 1e9
            roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel="sell_h_kernel",
                        bytes_per_launch=int(nbytes), ms_per_launch=kh["ms_avg"])
        line = {
            "metric": "edge-DoF updates/s, batched dga (all instances), 1 GPU",
            "value": head["edge_dof_per_s"], "unit": "edge-DoF/s", "n_gpus": 1, "steps": steps,
            "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{len(builds)} Kuhn dga cavities, n~U{{8..32}}, eps~U[1,10], f~U[1e8,5e8] Hz",
                       "n_edges_total": E, "n_faces_total": sum(b.n_faces for b in builds),
                       "parallelism": "single GPU, one launch per phase for the whole batch"},
            "fp64": {"edge_dof_per_s": res["fp64"]["edge_dof_per_s"], "ms_per_step": res["fp64"]["ms_per_step"]},
            "roofline": roof, "cpu_baseline": None, "kernels": head["kernels"], "tuning": head["tuning"],
        }
        print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=48, help="Kuhn cubes per side")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle steps of the CPU baseline (default 10 at 96^3, 60 at 48^3: 10-30 s of CPU work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dga", action="store_true")
    ap.add_argument("--no-48", action="store_true", help="headline: skip the 48^3 legs")
    ap.add_argument("--workload", choices=["headline", "gfdtd", "contrast", "batch", "fit", "fitprec", "copy"],
                    default="headline",
                    help="headline: the metric's configuration end to end (96^3 contrast mesh, 3000 steps, "
                         "learned+FP32 vs plain FP64, plus 48^3 legs); gfdtd: 48^3 steady-state legs; contrast: "
                         "96^3 steady-state legs with the explicit dga path (--n); batch: batched dga instances; "
                         "fit: FIT 64^3; fitprec: FIT random-media fp32 vs fp64 study (--n, default 256); copy: "
                         "a 1 GiB device copy (the reference kernel for HBM counter scaling).")
    ap.add_argument("--batch", type=int, default=256, help="instances of the batch workload")
    ap.add_argument("--batch-method", choices=["dga", "gfdtd"], default="dga", help="method of the batch workload")
    ap.add_argument("--paths", default="plain64,jacobi32,learned32,learned64",
                    help="gfdtd paths to run (profiling runs pick one)")
    ap.add_argument("--learn-steps", type=int, default=200, help="collection solves of the learned paths")
    args = ap.parse_args()
    args.steps_given = any(a.startswith("--steps") for a in sys.argv[1:])
    if args.cpu_steps is None:
        args.cpu_steps = 10 if args.workload == "headline" else 60
    if args.gpus != 1:
        print(f"bench.py drives exactly one GPU (asked for {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    from tetfit_amd.device import device_count
    if device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.workload == "headline":
        bench_headline(args)
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "copy":
        bench_copy(args)
        return
    if args.workload == "fit":
        bench_fit(args)
        return
    if args.workload == "fitprec":
        bench_fitprec(args)
        return
    contrast = args.workload == "contrast"
    if contrast and args.n == 48:
        args.n = 96
    paths = set(args.paths.split(","))
    total = args.warmup + args.steps + 260
    with tempfile.TemporaryDirectory() as d:
        bg = build_workload(d, "gfdtd", args.n, total + args.learn_steps, contrast)
        none = dict(edge_dof_per_s=None, ms_per_step=None, cg_iters_per_step=None, cg_max_relres=None,
                    cg_not_converged=None, kernels={})
        plain = f32j = fast = l64 = none
        if "plain64" in paths:
            plain = time_path(bg, "fp64", args.steps, args.warmup, "jacobi", "gfdtd fp64 jacobi (plain)")
        if "jacobi32" in paths:
            f32j = time_path(bg, "fp32", args.steps, args.warmup, "jacobi", "gfdtd fp32 jacobi")
        if "learned32" in paths:
            fast = time_path(bg, "fp32", args.steps, args.warmup, "learned", "gfdtd fp32 learned",
                             learn_steps=args.learn_steps)
        if "learned64" in paths:
            l64 = time_path(bg, "fp64", args.steps, args.warmup, "learned", "gfdtd fp64 learned",
                            learn_steps=args.learn_steps)
        head, headline_label = (fast, "gfdtd fp32 learned") if fast is not none else (plain, "gfdtd fp64 jacobi")
        dga = None
        if not args.no_dga:
            bd = build_workload(d, "dga", args.n, total, contrast)
            dga = time_path(bd, "fp64", args.steps, args.warmup, "jacobi", "dga fp64 (explicit)")
            dga32 = time_path(bd, "fp32", args.steps, args.warmup, "jacobi", "dga fp32 (explicit)")
        vb = 4 if head is fast else 8
        roofs = roofline_of(head["A"], head["kernels"], ["cg_wu_kernel", "cg_pg_kernel", "cg_w_kernel"], vb)
        roof = roofs[0] if roofs else None
        plain_roofs = roofline_of(plain["A"], plain["kernels"], ["cg_w_kernel"], 8) if plain is not none else []
        cpu = None
        if not args.no_cpu:
            t = time.time()
            cpu = cpu_baseline(bg, args.cpu_steps)
            log(f"[bench] cpu baseline {cpu['value'] / 1e6:.2f} M edge-DoF/s ({time.time() - t:.1f}s)")
        line = {
            "metric": METRIC,
            "value": head["edge_dof_per_s"],
            "unit": "edge-DoF/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": head["ms_per_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32" if head is fast else "f64",
            "data": "synthetic",
            "config": {"workload": f"Kuhn {args.n}^3 jittered tet cavity, method gfdtd (CG tol 1e-8)"
                                   + (", eps_r log-uniform [1,80] per 6^3 block" if contrast else ""),
                       "path": headline_label, "n_edges": bg.n_edges, "n_faces": bg.n_faces,
                       "parallelism": "single GPU"},
            "cg_iters_per_step": head["cg_iters_per_step"],
            "plain_fp64": {"edge_dof_per_s": plain["edge_dof_per_s"], "ms_per_step": plain["ms_per_step"],
                           "cg_iters_per_step": plain["cg_iters_per_step"]},
            "speedup_vs_plain_fp64": (head["edge_dof_per_s"] / plain["edge_dof_per_s"]
                                      if plain["edge_dof_per_s"] and head["edge_dof_per_s"] else None),
            "fp32_jacobi": {"edge_dof_per_s": f32j["edge_dof_per_s"], "cg_iters_per_step": f32j["cg_iters_per_step"],
                            "cg_max_relres": f32j["cg_max_relres"], "cg_not_converged": f32j["cg_not_converged"]},
            "fp64_learned": {"edge_dof_per_s": l64["edge_dof_per_s"], "cg_iters_per_step": l64["cg_iters_per_step"]},
            "learned": head.get("learned"),
            "roofline": roof,
            "roofline_kernels": roofs,
            "plain_fp64_roofline": plain_roofs[0] if plain_roofs else None,
            "A_storage": head.get("A"),
            "cpu_baseline": cpu,
        }
        if dga:
            line["dga_explicit"] = {"fp64_edge_dof_per_s": dga["edge_dof_per_s"], "fp64_ms_per_step": dga["ms_per_step"],
                                    "fp32_edge_dof_per_s": dga32["edge_dof_per_s"],
                                    "kernels_fp64": dga["kernels"]}
            droof = []
            for k, which in (("sell_h_kernel", "H"), ("sell_n_kernel", "N")):
                for r, vb_d in ((dga, 8), (dga32, 4)):
                    t = r["kernels"].get(k)
                    if not t:
                        continue
                    nb = dga_bytes(r[which], k, vb_d, bd.n_edges, bd.n_faces)
                    ach = nb / (t["ms_avg"] * 1e-3) / 1e9
                    droof.append(dict(kernel=k, dtype="f64" if vb_d == 8 else "f32", bound="hbm",
                                      achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                                      frac=round(ach / HBM_PEAK_GBS, 4), bytes_per_launch=int(nb),
                                      ms_per_launch=t["ms_avg"], traffic=measured_traffic(k, vb_d, nb)))
            line["dga_explicit"]["roofline_kernels"] = droof
        line["kernels"] = head["kernels"]
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
