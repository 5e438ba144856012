"""Writers of .fdtd inputs (and their mesh / solids files) for the synthetic
workloads this project measures and tests:

* fit_cavity   - FIT/FDTD unit box, PEC on 5 walls, wave-profile h source on
                 x-min (the 64^3 "0.0156 m step" workload and smaller ones)
* kuhn_cavity  - jittered Kuhn tetrahedral mesh of the unit cube with
                 piecewise-constant eps per block, PEC on 5 walls, the same
                 h source, method dga or gfdtd (the 48^3 / 96^3 workloads)

Every input pins the time step with `SET step` below the host's stability
estimate, so the reference and this build agree on dt.
"""
import math
import os

import numpy as np

from .host import Model, write_kuhn_mesh

C0 = 299792458.0

SOURCE_BLOCK = """DEFINE source 1
    SET type h
    SET surface 2
    SET profile wave
    SET direction y
    SET mode {{cos,cos,sin}}
    SET wavevector {{0,0,0.5}}
    SET carrier {carrier}
    SET frequency {freq}
    SET amplitude 1
{extra}END source 1
"""


def _probe_lines(points):
    return "".join("    SET probe {%.6g,%.6g,%.6g}\n" % tuple(p) for p in points)


def _duration(n_steps, dt, first_step):
    # fdtd loops i = 0..floor(T/dt) (n+1 trips), dga/gfdtd i = 1..floor(T/dt)
    k = n_steps - 1 if first_step == 0 else n_steps
    return (k + 0.5) * dt


def fit_cavity(directory, n=64, step=0.0156, n_steps=2000, n_probes=32, dt=None,
               freq=2e8, carrier="sin", name="fit", eps_r=1.0, sigma=0.0, extra_source="",
               pec_surfaces="1,3,4,5,6", source_surface=2, pml=None):
    """FIT box [0, n*step]^3 (xgrid {0,step,n*step+tiny}), one box solid of
    material 1; pml = {"+x": width, ...} adds pml boundary conditions. Returns
    the .fdtd path."""
    os.makedirs(directory, exist_ok=True)
    stop = 1.0 if abs(n * step - 0.9984) < 1e-9 else n * step + 0.25 * step
    solids = os.path.join(directory, f"{name}_solids.txt")
    with open(solids, "w") as f:
        f.write("DEFINE solid 1\n    SET type box\n    SET material 1\n"
                "    SET corner {0,0,0}\n    SET size {%r,%r,%r}\nEND solid 1\n" % (stop, stop, stop))
    if dt is None:
        # 0.95 x the Yee limit for vacuum (or the slowest-material-free bound)
        dt = 0.95 * step / (C0 * math.sqrt(3.0))
        dt = float("%.4g" % dt)
    h = n * step
    # probes on a line of cell centres along x through the middle of the box
    j = n // 2
    xs = [(2 * i * (n // (2 * n_probes)) + 0.5) * step if n >= 2 * n_probes else (i % n + 0.5) * step
          for i in range(n_probes)]
    pts = [(x, (j + 0.5) * step, (j + 0.5) * step) for x in xs]
    fdtd = os.path.join(directory, f"{name}.fdtd")
    mat = "DEFINE material 1\n    SET epsilon %r\n    SET mu 1\n" % eps_r
    if sigma:
        mat += "    SET sigma %r\n" % sigma
    mat += "END material 1\n"
    with open(fdtd, "w") as f:
        f.write(f"# FIT cavity {n}^3, step {step}\n")
        f.write("DEFINE mesh 1\n    SET type cartesian\n    SET file %s\n" % os.path.basename(solids))
        for ax in "xyz":
            f.write("    SET %sgrid {0,%r,%r}\n" % (ax, step, stop))
        f.write("END mesh 1\n")
        f.write(mat)
        f.write("DEFINE bc 1\n    SET type pec\n    SET surfaces {%s}\nEND bc 1\n" % pec_surfaces)
        for k, (side, w) in enumerate(sorted((pml or {}).items())):
            f.write("DEFINE bc %d\n    SET type pml\n    SET direction %s\n    SET width %r\nEND bc %d\n"
                    % (90 + k, side, w, 90 + k))
        f.write(SOURCE_BLOCK.format(carrier=carrier, freq=repr(freq), extra=extra_source)
                .replace("SET surface 2", "SET surface %d" % source_surface))
        f.write("DEFINE output 1\n    SET mode probepoint\n    SET name %s\n    SET period 0\n" % name)
        f.write(_probe_lines(pts))
        f.write("END output 1\n")
        f.write("DEFINE simulation 1\n    SET method fdtd\n    SET mesh 1\n    SET output 1\n"
                "    SET courant 0.98\n    SET step %r\n    SET duration %r\nEND simulation 1\n"
                % (dt, _duration(n_steps, dt, 0)))
    return fdtd


def fit_random_media(directory, n=512, n_steps=4000, n_probes=16, eps_range=(1.0, 4.0), seed=3, sigma=0.05,
                     slab=(0.4, 0.6), freq=3e8, name=None, dt=None):
    """FIT unit box with n^3 cells of step 1/n: every cell its own eps_r ~
    U[eps_range) (SET epsilon_random, seed), sigma S/m in the slab
    slab[0] <= x < slab[1], PEC on 5 walls, a wave-profile h source on x-min
    with the Gaussian-modulated sine carrier (gaussmod, f = freq, width
    1.5/f, delay 3/f), probes on a line through the centre. Returns the
    .fdtd path."""
    os.makedirs(directory, exist_ok=True)
    name = name or f"fitrand{n}"
    step = 1.0 / n
    stop = 1.0 + 0.25 * step
    lo, hi = eps_range
    solids = os.path.join(directory, f"{name}_solids.txt")
    with open(solids, "w") as f:
        f.write("DEFINE solid 1\n    SET type box\n    SET material 2\n"
                "    SET corner {%r,0,0}\n    SET size {%r,%r,%r}\nEND solid 1\n"
                % (slab[0], slab[1] - slab[0], stop, stop))
        f.write("DEFINE solid 2\n    SET type box\n    SET material 1\n"
                "    SET corner {0,0,0}\n    SET size {%r,%r,%r}\nEND solid 2\n" % (stop, stop, stop))
    if dt is None:
        # 0.95 x the Yee limit of the fastest (eps_r = lo) medium
        dt = float("%.4g" % (0.95 * step * math.sqrt(lo) / (C0 * math.sqrt(3.0))))
    j = n // 2
    pts = [(((i + 0.5) / n_probes), (j + 0.5) * step, (j + 0.5) * step) for i in range(n_probes)]
    extra = "    SET width %r\n    SET delay %r\n" % (1.5 / freq, 3.0 / freq)
    fdtd = os.path.join(directory, f"{name}.fdtd")
    with open(fdtd, "w") as f:
        f.write(f"# FIT {n}^3 random media, eps_r U[{lo},{hi}) per cell (seed {seed}), "
                f"sigma {sigma} in {slab[0]} <= x < {slab[1]}\n")
        f.write("DEFINE mesh 1\n    SET type cartesian\n    SET file %s\n" % os.path.basename(solids))
        for ax in "xyz":
            f.write("    SET %sgrid {0,%r,%r}\n" % (ax, step, stop))
        f.write("END mesh 1\n")
        f.write("DEFINE material 1\n    SET epsilon_random {%r,%r,%d}\n    SET mu 1\nEND material 1\n"
                % (lo, hi, seed))
        f.write("DEFINE material 2\n    SET epsilon_random {%r,%r,%d}\n    SET mu 1\n    SET sigma %r\n"
                "END material 2\n" % (lo, hi, seed, sigma))
        f.write("DEFINE bc 1\n    SET type pec\n    SET surfaces {1,3,4,5,6}\nEND bc 1\n")
        f.write(SOURCE_BLOCK.format(carrier="gaussmod", freq=repr(freq), extra=extra))
        f.write("DEFINE output 1\n    SET mode probepoint\n    SET name %s\n    SET period 0\n" % name)
        f.write(_probe_lines(pts))
        f.write("END output 1\n")
        f.write("DEFINE simulation 1\n    SET method fdtd\n    SET mesh 1\n    SET output 1\n"
                "    SET courant 0.98\n    SET step %r\n    SET duration %r\nEND simulation 1\n"
                % (dt, _duration(n_steps, dt, 0)))
    return fdtd


def kuhn_cavity(directory, n=48, method="dga", n_steps=1000, n_probes=8, jitter=0.2, seed=2,
                block=8, eps_range=(1.0, 10.0), log_eps=False, freq=2e8, tol=1e-8, dt=None,
                name=None, dt_safety=0.9, pec_surfaces="1,3,4,5,6", source_surface=2,
                sigma=0.0, sigma_frac=0.0, chi=0.0, chi_frac=0.0, mu_range=None, extra_source=""):
    """Kuhn mesh of the unit cube, eps_r per block drawn from eps_range
    (uniform, or log-uniform) with numpy default_rng(seed). Lossy media: a
    fraction sigma_frac of the blocks (drawn after eps) get conductivity
    sigma S/m, a fraction chi_frac magnetic loss chi; mu_range draws mu_r per
    block (last). Returns the .fdtd
    path. When dt is None the host's stability estimate is computed once and
    dt = dt_safety x estimate (rounded to 4 digits) is pinned with SET step.
    dt="courant" writes no step and `SET courant dt_safety` instead (one host
    build; dt = dt_safety x the estimate, the duration then only bounds the
    step count)."""
    os.makedirs(directory, exist_ok=True)
    name = name or f"kuhn{n}_{method}"
    mesh = os.path.join(directory, f"kuhn{n}_s{seed}_j{jitter}_b{block}.mesh")
    if not os.path.exists(mesh):
        max_label, _ = write_kuhn_mesh(mesh, n, jitter, seed, block)
    else:
        nb = -(-n // block)
        max_label = nb ** 3 + (1 if nb ** 3 >= 7 else 0)
    rng = np.random.default_rng(seed)
    labels = [l for l in range(1, max_label + 1) if l != 7]
    lo, hi = eps_range
    if log_eps:
        eps = np.exp(rng.uniform(math.log(lo), math.log(hi), size=len(labels)))
    else:
        eps = rng.uniform(lo, hi, size=len(labels))
    sig = np.where(rng.uniform(size=len(labels)) < sigma_frac, sigma, 0.0) if sigma_frac > 0 else np.zeros(len(labels))
    chis = np.where(rng.uniform(size=len(labels)) < chi_frac, chi, 0.0) if chi_frac > 0 else np.zeros(len(labels))
    mus = rng.uniform(mu_range[0], mu_range[1], size=len(labels)) if mu_range else np.ones(len(labels))
    h = 1.0 / n
    pts = [((i + 0.37) / n_probes, 0.5 + 0.11 * h, 0.5 + 0.07 * h) for i in range(n_probes)]

    def write(step_line):
        fdtd = os.path.join(directory, f"{name}.fdtd")
        with open(fdtd, "w") as f:
            f.write(f"# Kuhn {n}^3 cavity, jitter {jitter}h, seed {seed}, method {method}\n")
            f.write("DEFINE mesh 1\n    SET type tetrahedral\n    SET mesher netgen\n"
                    "    SET file %s\nEND mesh 1\n" % os.path.basename(mesh))
            for l, e, sg, ch, mu in zip(labels, eps, sig, chis, mus):
                f.write("DEFINE material %d\n    SET epsilon %r\n    SET mu %s\n"
                        % (l, float(e), "1" if mu == 1.0 else repr(float(mu))))
                if sg:
                    f.write("    SET sigma %r\n" % float(sg))
                if ch:
                    f.write("    SET chi %r\n" % float(ch))
                f.write("END material %d\n" % l)
            f.write("DEFINE bc 1\n    SET type pec\n    SET surfaces {%s}\nEND bc 1\n" % pec_surfaces)
            f.write(SOURCE_BLOCK.format(carrier="sin", freq=repr(freq), extra="")
                    .replace("SET surface 2", "SET surface %d" % source_surface))
            f.write(extra_source)
            f.write("DEFINE output 1\n    SET mode probepoint\n    SET name %s\n    SET period 0\n" % name)
            f.write(_probe_lines(pts))
            f.write("END output 1\n")
            f.write("DEFINE simulation 1\n    SET method %s\n    SET mesh 1\n    SET output 1\n"
                    "    SET courant %r\n%s    SET solver cg\n    SET tolerance %r\nEND simulation 1\n"
                    % (method, courant, step_line, tol))
        return fdtd

    courant = 1
    if dt == "courant":
        courant = dt_safety
        # an upper bound of the step count: dt >= 0.05 h / c0 on these meshes
        return write("    SET duration %r\n" % (n_steps * 0.05 / (n * C0)))
    if dt is None:
        probe = write("    SET duration 0\n")
        b = Model(probe).build(1)
        dt = float("%.4g" % (dt_safety * b.stable_dt))
        del b
    return write("    SET step %r\n    SET duration %r\n" % (dt, _duration(n_steps, dt, 1)))
