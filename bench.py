#!/usr/bin/env python3
"""Steps/s of device-resident semi-Lagrangian Navier-Stokes stepping on one
MI355X, at 1e-8 relative residual per CG solve.

Workload (default): lid-driven cavity on the unit square, 2048x2048 squares
(8,388,608 P2-P1 triangles, 16,785,409 velocity / 4,198,401 pressure nodes),
nu = 1e-3, dt = 1e-3, lid u = (1, 0).  A step = advect + two diffusion solves
+ pressure solve + two velocity-mass solves + right-hand sides, all on the
GPU (include/assembly.h:1516-1554 of the reference).

Two forms run on the same assembled simulation from the same initial state:
  * fp64          plain FP64 CG with the reference's initial guesses (the
                  previous pressure for the pressure solve)
  * fp32_refined  FP32 matrix values and r, p, q vectors, FP64 solution,
                  FP64 accumulation of products and dots, FP64 reliable
                  updates of the residual; pressure solves start from the
                  A-norm projection onto past pressures (--warm)
Both stop every solve at ||b - Ax|| <= 1e-8 ||b|| on the FP64 residual; the
faster of the two is the reported `value` (the chosen form).  fp32_refined
also runs once with the previous-pressure guess, reported beside them, to
separate the precision's share of the speed-up from the initial guess's.

Also reported: the roofline of the dominant kernel (the pressure CG kernel
of the chosen form: algorithmic bytes / measured kernel time, HIP events on
the kernel's stream) and the CPU oracle timed on this host on a bounded sample
of the same workload (cpu_baseline).

  python bench.py [--steps K] [--warmup W] [--n 2048] [--no-cpu] [--form both]
                  [--warm projection|previous|extrapolate] [--config NAME]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cg_bytes(nnz, n, iters, refinements, mixed, columns=1):
    """Algorithmic HBM bytes of one persistent-CG launch: the textbook CG
    iteration (SpMV + (p, Ap); x, r update + (r, r); p update) moves the
    matrix once (int32 column + value) plus row_start, and 11 vector words
    per row:
      FP64 iteration      (4+8)*nnz + (4 + 11*8)*n
      mixed iteration     (4+4)*nnz + (4 + 9*4 + 2*8)*n   (values, r, p, q
                          in FP32; the solution x read and written in FP64)
      residual b - A x    (4+8)*nnz + (4 + 5*8)*n  (FP64 matrix and vectors:
                          x gather, b, r write, p reset; at set-up and at
                          every reliable update)
    With 2 columns the matrix is read once and the vectors twice."""
    if mixed:
        it = (4 + 4) * nnz + 4 * n + (9 * 4 + 2 * 8) * n * columns
    else:
        it = (4 + 8) * nnz + 4 * n + 11 * 8 * n * columns
    res = (4 + 8) * nnz + 4 * n + 5 * 8 * n * columns
    return iters * it + (refinements + 1) * res


# HBM bytes per CG iteration (mixed?, launch class) measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE (scaled by a k_axpy calibration) on this
# workload, the launches' set-up and reliable-update residual passes
# included; source: profiles/k_cg_cavity2048_hbm_traffic.md (the mass row
# on the default command: k_cg_p2's narrow-gather form, ~10.6
# iterations and ~2.9 FP64 passes a launch;
# the others on the forms that still run those solves by the persistent
# CG).  roofline.traffic = this figure x the iterations of a launch.
MEASURED_BYTES_PER_ITERATION = {(False, "pressure"): 721.9e6, (True, "pressure"): 435.4e6,
                                (True, "diffusion v"): 2723.9e6, (True, "mass u+v"): 4898.0e6}


def spmv_bench(dev, sim, slot, n_rows, n_cols, nnz, reps=100):
    """Median-free throughput of `reps` back-to-back nsfem_sp_r_mult launches
    (x ~ U(-1, 1), seed 7): wall time between two device synchronisations,
    so launch gaps count against the kernel.  Algorithmic bytes per product:
    (4+8) nnz + 4 (rows+1) + 8 cols (x once) + 8 rows (y)."""
    import numpy as np

    x = dev.buffer(np.random.default_rng(7).uniform(-1, 1, n_cols))
    y = dev.buffer(np.zeros(n_rows))
    for _ in range(3):
        sim.sp_r_mult(slot, x, y)
    dev.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        sim.sp_r_mult(slot, x, y)
    dev.synchronize()
    sec = (time.perf_counter() - t) / reps
    x.close()
    y.close()
    byt = 12 * nnz + 4 * (n_rows + 1) + 8 * n_cols + 8 * n_rows
    return {"rows": n_rows, "nnz": nnz, "us": round(sec * 1e6, 1),
            "achieved_GBs": round(byt / sec / 1e9, 1), "frac": round(byt / sec / 1e9 / PEAK_HBM_GBS, 4)}


def run_form(sim, form, warmup, steps, n_p, nnz_p, warm="previous", coarse=None,
             coarse_rtol=1e-6, velocity_projection=0, coarse_levels=(), rtol=1e-8,
             pressure_multigrid=False, pair_velocity=True, velocity_jacobi=False,
             velocity_multigrid=0, mass_projection=0):
    """Steps `sim` in one form after `warmup` untimed steps.  A nonzero
    velocity_multigrid needs the simulation's velocity hierarchy set."""
    sim.reset()
    sim.configure(precision=form, rtol=rtol, pair_velocity=pair_velocity,
                  velocity_jacobi=velocity_jacobi)
    if velocity_multigrid or getattr(sim, "_velocity_sizes", None):
        sim.set_velocity_multigrid(None, mode=velocity_multigrid)
    sim.set_warm_start(warm)
    sim.set_coarse_space(coarse, rtol=coarse_rtol)
    if coarse is not None:
        for level in coarse_levels:
            sim.add_coarse_level(level)
    sim.set_pressure_multigrid(pressure_multigrid and coarse is not None)
    sim.set_velocity_projection(velocity_projection)
    sim.set_mass_projection(mass_projection)
    for i in range(warmup):
        sim.step(1)
        log(f"{form}: warmup step {i + 1}/{warmup} done")
    sim.device.synchronize()
    agg = None
    per_step = []
    t0 = time.perf_counter()
    for i in range(steps):
        st = sim.step(1)
        per_step.append(dict(st))
        if agg is None:
            agg = dict(st)
        else:
            for k in ("cg_iterations", "cg_seconds", "cg_refinements", "cg_launches",
                      "cg_launch_iterations", "vcycle_solves", "vcycle_iterations",
                      "vcycle_seconds"):
                agg[k] = [a + b for a, b in zip(agg[k], st[k])]
            agg["max_relative_residual"] = [max(a, b) for a, b in zip(agg["max_relative_residual"],
                                                                       st["max_relative_residual"])]
            agg["seconds"] += st["seconds"]
            agg["advect_seconds"] += st["advect_seconds"]
            agg["coarse_iterations"] += st["coarse_iterations"]
            agg["coarse_seconds"] += st["coarse_seconds"]
            agg["steps"] += 1
        log(f"{form}: step {i + 1}/{steps}: {st['seconds'] * 1e3:.1f} ms, "
            f"iterations {st['cg_iterations']}")
    sim.device.synchronize()
    wall = time.perf_counter() - t0
    agg["fields"] = sim.fields()
    agg["per_step"] = per_step
    return wall, agg


# CG-kernel launch classes of a step's stats: (name, stats index, columns,
# test of a step's stats for a launch of that class).  A velocity system's
# u entry holds its paired two-column launch when its v entry is empty; a
# solve by V-cycle preconditioned CG launches no CG kernel.
LAUNCH_CLASSES = [
    ("diffusion u+v", 0, 2, lambda s: s["cg_launches"][0] and not s["cg_launches"][1]),
    ("diffusion u", 0, 1, lambda s: s["cg_launches"][1] and not s["vcycle_solves"][0]),
    ("diffusion v", 1, 1, lambda s: s["cg_launches"][1] and not s["vcycle_solves"][1]),
    ("pressure", 2, 1, lambda s: s["cg_launches"][2] and not s["vcycle_solves"][2]),
    ("mass u+v", 3, 2, lambda s: s["cg_launches"][3] and not s["cg_launches"][4]),
    ("mass u", 3, 1, lambda s: s["cg_launches"][4]),
    ("mass v", 4, 1, lambda s: s["cg_launches"][4]),
]


def launch_classes(per_step):
    """Per class: [launches, iterations, seconds, reliable updates] summed
    over the steps whose stats hold a launch of that class."""
    out = {}
    for name, k, cols, test in LAUNCH_CLASSES:
        acc = [0, 0, 0.0, 0]
        for s in per_step:
            if test(s):
                acc[0] += s["cg_launches"][k]
                acc[1] += s["cg_launch_iterations"][k]
                acc[2] += s["cg_seconds"][k]
                acc[3] += s["cg_refinements"][k]
        out[name] = (k, cols, acc)
    return out


def coarse_levels(n, count):
    """Prolongations of `count` multigrid levels below the n x n cavity's P1
    coarse space, each halving the grid (n must stay even); count < 0 halves
    down to the 32 x 32 grid (33^2 nodes: the library's dense deepest
    level)."""
    from navier_stokes_fem_amd import mesh

    out = []
    if count < 0:
        count = 0
        k = n
        while k > 32 and k % 2 == 0:
            k //= 2
            count += 1
    for _ in range(count):
        if n % 2 or n < 4:
            raise ValueError(f"cannot halve the {n}^2 coarse grid again")
        out.append(mesh.cavity_prolongation(n, 2))
        n //= 2
    return out


def field_difference(fields, ref):
    """Per field (u, v, p) after the same steps: max |a - b|, ||a - b|| / ||b||
    and max |b|; for p also the relative difference once each field's mean
    is removed (a 1e-8 relative residual fixes the pressure only up to its
    smoothest, near-constant mode, whose eigenvalue is tiny)."""
    import numpy as np

    out = {}
    for name, a, b in zip(("u", "v", "p"), fields, ref):
        d = a - b
        nb = float(np.linalg.norm(b))
        out[name] = {"max_abs": float(np.abs(d).max()), "rel_l2": float(np.linalg.norm(d) / nb)
                     if nb > 0 else 0.0, "max_abs_reference": float(np.abs(b).max())}
        if name == "p":
            bc = b - b.mean()
            out[name]["rel_l2_mean_removed"] = float(np.linalg.norm(d - d.mean()) /
                                                     max(float(np.linalg.norm(bc)), 1e-300))
    return out


def cpu_baseline(m, nu, dt, gpu_iters_per_step, sample_iters):
    """The oracle (FP64 CG, KD advection; OpenMP over this host's cores)
    timed on a bounded sample of the same workload: one advection of every
    velocity node, the right-hand-side products, and `sample_iters` CG
    iterations of each matrix (timed as the difference between two runs so
    set-up costs cancel).  A step's CPU time is extrapolated with the
    per-step CG iteration counts of the GPU FP64 run.  The same sample runs
    once more on one thread (half the CG iterations), for the 1-thread
    figure the reference also publishes (tex/Chapters/04_implementation_
    details.tex:104)."""
    import numpy as np

    from navier_stokes_fem_amd import device as D
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if threads <= 0:
        threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)
    log(f"cpu baseline: assembling matrices for the oracle ({threads} threads)")
    a = D.assemble(m, nu, dt, threads=threads)
    tree = D.build_kdtree(m)
    ot = O.Tree(tree.nodes, tree.leaf_indices, tree.bbox, tree.max_depth)
    V = m.node_count
    csr = lambda k: a[k][:3]
    its = gpu_iters_per_step

    def sample(nthreads, iters):
        O.set_threads(nthreads)
        rng = np.random.default_rng(5)
        u = rng.uniform(-1, 1, V)
        v = rng.uniform(-1, 1, V)
        t = time.perf_counter()
        O.advect_f64(ot, m.xy(), m.elems(), u, v, dt)
        t_adv = time.perf_counter() - t
        log(f"cpu baseline ({nthreads} threads): advect {t_adv:.2f} s")
        t = time.perf_counter()
        O.spmv(csr("velocity_mass"), u)
        O.spmv(csr("velocity_mass"), v)
        O.spmv(csr("velocity_divergence"), np.concatenate([u, v]))
        O.spmv(csr("pressure_divergence"), rng.uniform(-1, 1, m.pressure_count))
        t_rhs = time.perf_counter() - t
        per_iter = {}
        for k in ("diffusion", "pressure_stiffness", "velocity_mass"):
            A = csr(k)
            n = len(A[0]) - 1
            b = rng.uniform(-1, 1, n)
            x0 = np.zeros(n)
            t = time.perf_counter()
            O.cg(A, b, x0, max_iter=2, rtol=0.0, atol=0.0)
            t2 = time.perf_counter() - t
            t = time.perf_counter()
            O.cg(A, b, x0, max_iter=2 + iters, rtol=0.0, atol=0.0)
            t22 = time.perf_counter() - t
            per_iter[k] = max(t22 - t2, 1e-9) / iters
            log(f"cpu baseline ({nthreads} threads): {k} {per_iter[k] * 1e3:.2f} ms/iteration")
        step = (t_adv + t_rhs + (its[0] + its[1]) * per_iter["diffusion"] +
                its[2] * per_iter["pressure_stiffness"] +
                (its[3] + its[4]) * per_iter["velocity_mass"])
        return t_adv, t_rhs, per_iter, step

    t_adv, t_rhs, per_iter, step = sample(threads, sample_iters)
    t_adv1, t_rhs1, per_iter1, step1 = sample(1, max(1, sample_iters // 2))
    O.set_threads(threads)
    return {
        "value": 1.0 / step,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle (FP64 CG + KD advection, OpenMP) on the same {m.name} matrices: "
                   f"1 advection of {V} nodes ({t_adv:.2f}s), RHS products ({t_rhs:.2f}s), "
                   f"{sample_iters} CG iterations per matrix (ms/iteration: "
                   + ", ".join(f"{k} {v * 1e3:.1f}" for k, v in per_iter.items())
                   + f"); step time extrapolated with the GPU FP64 run's iterations/step "
                   f"{[round(x, 1) for x in its]}"),
        "seconds_per_step": step,
        "one_thread": {"value": 1.0 / step1, "seconds_per_step": step1,
                       "advect_s": round(t_adv1, 3), "rhs_s": round(t_rhs1, 3),
                       "ms_per_iteration": {k: round(v * 1e3, 2) for k, v in per_iter1.items()},
                       "sample": f"the same sample on 1 thread, {max(1, sample_iters // 2)} CG "
                                 "iterations per matrix"},
    }


def _line(metric, value, unit, steps, ms, config, **extra):
    out = {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": steps,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "vs_baseline": None,
           "data": "synthetic", "config": config}
    out.update(extra)
    return out


# HBM bytes per launch of the product measured with rocprofv3 (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0, scaled by the saxpy calibration) on the 2048x2048
# cavity's matrices; keyed by matrix.
MEASURED_SPMV_BYTES = {"pressure_stiffness": 387_601_329, "diffusion": 2_880_960_586}


def bench_spmv(dev, args):
    """CSR product on the n x n cavity's pressure matrix (P1, 7 nonzeros per
    row, 16-bit column offsets) and diffusion matrix (P2) in the reference's
    node order (vertices first: 32-bit columns) and in the lattice order a
    simulation keeps it in (16-bit column offsets):
    mean device time per launch over 100 back-to-back launches, three rounds
    after 3 warm-up launches, timed with HIP events on the library's stream.
    Achieved = reference-CSR bytes per launch, (4+8) nnz + 4 (rows+1) +
    8 cols + 8 rows, over that time.  cpu_baseline: the oracle's C row loop
    (OpenMP, rows split statically over the threads) on the pressure matrix,
    whose result the GPU product must equal bit for bit."""
    import numpy as np

    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.device import assemble
    from oracle import oracle as O

    m = mesh.cavity(args.n)
    log(f"spmv: assembling the {args.n}x{args.n} cavity")
    a = assemble(m, 1e-3, 1e-3)
    rng = np.random.default_rng(7)
    res = {}
    check = None
    # the diffusion matrix also in the order a simulation keeps it in: the
    # velocity nodes sorted by (y, x)
    import scipy.sparse as sp

    order = np.lexsort((m.nodes[:, 0], m.nodes[:, 1]))
    rs_d, col_d, val_d, _ = a["diffusion"]
    D = sp.csr_matrix((val_d, col_d, rs_d), shape=(len(rs_d) - 1,) * 2)[order][:, order].tocsr()
    D.sort_indices()
    a = dict(a)
    a["diffusion_lattice_order"] = (D.indptr.astype(np.int32), D.indices.astype(np.int32), D.data,
                                    None)
    del D
    for slot, name in ((0, "pressure_stiffness"), (2, "diffusion"), (2, "diffusion_lattice_order")):
        rs, col, val, _ = a[name]
        dev.upload_csr(slot, rs, col, val)
        n_rows, nnz = len(rs) - 1, len(col)
        n_cols = n_rows  # square, as uploaded
        xh = rng.uniform(-1, 1, n_cols)
        x = dev.buffer(xh)
        y = dev.buffer(np.zeros(n_rows))
        for _ in range(3):
            dev.sp_r_mult(slot, x, y)
        rounds = [dev.time_sp_r_mult(slot, x, y, 100) for _ in range(3)]
        sec = float(np.mean(rounds))
        byt = 12 * nnz + 4 * (n_rows + 1) + 8 * n_cols + 8 * n_rows
        res[name] = {"rows": n_rows, "nnz": nnz, "us_per_launch": round(sec * 1e6, 2),
                     "us_per_round": [round(r * 1e6, 2) for r in rounds],
                     "algorithmic_bytes": byt, "achieved_GBs": round(byt / sec / 1e9, 1),
                     "frac": round(byt / sec / 1e9 / PEAK_HBM_GBS, 4),
                     "measured_hbm_bytes": MEASURED_SPMV_BYTES.get(name)}
        log(f"spmv {name}: {res[name]}")
        if slot == 0:
            yd = y.download()
            t = time.perf_counter()
            reps = 0
            while reps < 20 or time.perf_counter() - t < 10.0:
                yo = O.spmv((rs, col, val), xh)
                reps += 1
            cpu_sec = (time.perf_counter() - t) / reps
            check = {"value": round(byt / cpu_sec / 1e9, 2), "unit": "GB/s",
                     "cores": O.get_threads(),
                     "kind": "port", "sample": f"{reps} products of the pressure matrix by "
                                               f"the oracle's C row loop",
                     "seconds_per_product": round(cpu_sec, 5),
                     "gpu_equals_oracle_bitwise": bool(np.array_equal(yd, yo))}
        x.close()
        y.close()
    # a stream on the same chip for scale: nsfem_saxpy (16 B read + 8 B
    # written per element) on 2^25 doubles, 100 launches between two
    # synchronisations (launch gaps included)
    n = 1 << 25
    xs = dev.buffer(rng.uniform(-1, 1, n))
    ys = dev.buffer(np.zeros(n))
    for _ in range(3):
        dev.saxpy(n, 0.5, xs, ys)
    dev.synchronize()
    t = time.perf_counter()
    for _ in range(100):
        dev.saxpy(n, 0.5, xs, ys)
    dev.synchronize()
    sec = (time.perf_counter() - t) / 100
    stream = {"kernel": "k_axpy (nsfem_saxpy), 2^25 doubles", "us_per_launch": round(sec * 1e6, 1),
              "GBs": round(24 * n / sec / 1e9, 1)}
    log(f"spmv: stream reference {stream}")
    xs.close()
    ys.close()
    p = res["pressure_stiffness"]
    traffic = MEASURED_SPMV_BYTES.get("pressure_stiffness")
    roofline = {"bound": "hbm", "achieved": p["achieved_GBs"], "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": p["frac"], "traffic": traffic,
                "kernel": "k_spmv<0, short> (pressure matrix)"}
    return _line("CSR product achieved HBM GB/s (reference CSR bytes), n x n cavity pressure matrix",
                 p["achieved_GBs"], "GB/s", 300, p["us_per_launch"] * 1e-3,
                 {"workload": f"lid-driven cavity {args.n}x{args.n}: pressure matrix "
                              f"{p['rows']} rows / {p['nnz']} nonzeros, diffusion matrix "
                              f"{res['diffusion']['rows']} rows / {res['diffusion']['nnz']} "
                              f"nonzeros, x ~ U(-1, 1) seed 7"},
                 dtype="f64", roofline=roofline, matrices=res, stream_reference=stream,
                 cpu_baseline=check)


def bench_cavity32(dev, args):
    """32x32 cavity (2048 P2-P1 elements), nu = 0.01, dt = 0.01, 100 steps of
    the plain FP64 method on the GPU; the cpu_baseline leg runs the same 100
    steps in the CPU oracle and reports its speed and the field difference."""
    import numpy as np

    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation
    from oracle import oracle as O

    steps = 100
    m = mesh.cavity(32)
    nu, dt = 0.01, 0.01
    sim = Simulation(m, nu, dt, device=dev, rtol=1e-8, precision="fp64")
    dev.synchronize()
    t = time.perf_counter()
    st = sim.step(steps)
    wall = time.perf_counter() - t
    u, v, p = sim.fields()
    sim.close()
    asm = O.Assembly(m, nu, dt)
    tree = O.kd_build(m.xy(), m.elems())
    ou, ov, op = O.initial_fields(m)
    t = time.perf_counter()
    for _ in range(steps):
        ou, ov, op, _ = O.step(asm, tree, m, dt, ou, ov, op, rtol=1e-8)
    cpu = time.perf_counter() - t
    err = {k: float(np.abs(a - b).max()) for k, a, b in (("u", u, ou), ("v", v, ov), ("p", p, op))}
    scale = {k: float(np.abs(b).max()) for k, b in (("u", ou), ("v", ov), ("p", op))}
    log(f"cavity32: {steps / wall:.1f} steps/s on the GPU, {steps / cpu:.1f} on the CPU oracle; "
        f"max |gpu - oracle| {err}")
    cpu_baseline = {"value": round(steps / cpu, 3), "unit": "steps/s", "cores": O.get_threads(),
                    "kind": "port",
                    "sample": "the whole workload: 100 oracle steps (FP64 CG + KD advection) "
                              "from the same initial state",
                    "field_max_abs_difference_gpu_vs_oracle": err, "field_max_abs": scale}
    return _line("steps/sec, 32x32 cavity, FP64 CG at 1e-8 rel. residual", steps / wall,
                 "steps/s", steps, 1e3 * wall / steps,
                 {"workload": "lid-driven cavity 32x32 squares (2048 P2-P1 triangles, 4225 velocity "
                              "/ 1089 pressure nodes), nu=0.01, dt=0.01, 100 steps"},
                 dtype="f64", cg_iterations_per_step=[x / steps for x in st["cg_iterations"]],
                 cpu_baseline=cpu_baseline)


def bench_channel512(dev, args):
    """512x128 channel past the disk (jitter seed 2), nu = 1e-3, dt = 5e-4,
    200 steps in each form: plain FP64, FP32 + FP64 refinement with the
    previous-pressure guess, with the projection guess, with the
    projection guess and the pressure solved by CG preconditioned with a
    V-cycle over the channel's grid hierarchy (holes and outlet nodes
    kept out of the coarse spaces), and the headline's chain (that pressure
    solve, Jacobi-scaled mass solves with the projection guess, and the
    diffusion columns by V-cycle preconditioned CG over the channel's
    velocity hierarchy wherever that is measured faster: mode 1).  The
    chain leaves out the velocity projection of the diffusion guesses: on
    this channel it leaves 4.5x the plain method's field error against a
    1e-12 run at the same tolerance (u 1.0e-3, v 5.0e-3, p 9.8e-3 against
    2.2e-4, 1.1e-3, 2.2e-3) and makes the chain slower (180 against 188
    steps/s), where on the 2048^2 cavity it makes it faster and 3-4x more
    accurate."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation

    steps = 200
    m = mesh.channel(512, 128, seed=2)
    sim = Simulation(m, 1e-3, 5e-4, device=dev, rtol=1e-8, precision="fp64")
    res, fields = {}, {}
    hierarchy = mesh.grid_hierarchy(m, min_size=8)
    vhierarchy = mesh.velocity_hierarchy(m, min_size=8)
    for form, prec, warm in (("fp64", "fp64", "previous"),
                             ("fp32_refined", "fp32_refined", "previous"),
                             ("fp32_refined_projection", "fp32_refined", "projection"),
                             ("fp32_refined_multigrid", "fp32_refined", "projection"),
                             ("fp32_refined_chain", "fp32_refined", "projection")):
        chain = form == "fp32_refined_chain"
        sim.reset()
        sim.configure(precision=prec, rtol=1e-8, velocity_jacobi="mass" if chain else "none")
        sim.set_warm_start(warm)
        multigrid = form in ("fp32_refined_multigrid", "fp32_refined_chain")
        sim.set_coarse_space(hierarchy[0] if multigrid else None)
        if multigrid:
            for level in hierarchy[1:]:
                sim.add_coarse_level(level)
            sim.set_pressure_multigrid(True)
        sim.set_mass_projection(1 if chain else 0)
        sim.set_velocity_projection(0)
        if chain:
            sim.set_velocity_multigrid(vhierarchy, mode=1)
        dev.synchronize()
        t = time.perf_counter()
        st = sim.step(steps)
        wall = time.perf_counter() - t
        res[form] = {"steps_per_s": round(steps / wall, 2),
                     "cg_iterations_per_step": [x / steps for x in st["cg_iterations"]],
                     "max_relative_residual": st["max_relative_residual"]}
        fields[form] = sim.fields()
        log(f"channel512 {form}: {steps / wall:.1f} steps/s, iterations/step {res[form]['cg_iterations_per_step']}")
    # the same 200 steps in FP64 with every solve to 1e-12: each form's
    # distance to it per field (rel. L2), and the mixed forms' difference
    # from the FP64 run per field
    sim.reset()
    sim.set_coarse_space(None)
    sim.remove_velocity_multigrid()
    sim.set_mass_projection(0)
    sim.set_velocity_projection(0)
    sim.configure(precision="fp64", rtol=1e-12, velocity_jacobi="none")
    sim.set_warm_start("previous")
    sim.step(steps)
    tight = sim.fields()
    sim.close()
    for form in res:
        res[form]["error_vs_tight_solve"] = {k: v["rel_l2"] for k, v in
                                             field_difference(fields[form], tight).items()}
        if form != "fp64":
            res[form]["field_difference_vs_fp64"] = field_difference(fields[form], fields["fp64"])
        log(f"channel512 {form}: rel. L2 vs the 1e-12 run {res[form]['error_vs_tight_solve']}")
    best = max(res, key=lambda f: res[f]["steps_per_s"])
    return _line("steps/sec, 512x128 channel, 1e-8 rel. residual", res[best]["steps_per_s"],
                 "steps/s", steps, 1e3 / res[best]["steps_per_s"],
                 {"workload": f"channel [0,2.2]x[0,0.41] 512x128 jittered (seed 2) minus the disk: "
                              f"{m.element_count} P2-P1 triangles, {m.node_count} velocity / "
                              f"{m.pressure_count} pressure nodes, nu=1e-3, dt=5e-4, 200 steps",
                  "form": best},
                 dtype="f32+f64refine" if best.startswith("fp32") else "f64", forms=res)


def bench_cavity_batch64(dev, args):
    """64 cavities, side n ~ U{16..256}, lid ~ U(0.5, 2), nu ~ logU(1e-3,
    1e-2) (seed 3), dt = 0.005, 50 steps in FP32-refined form: all 64 in one
    batch (each solve kind one persistent launch for all), after the
    autotuner, against the same 64 stepped one after another."""
    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.simulation import Simulation, autotune_batch, step_batch

    steps = 50
    specs = mesh.cavity_batch(64, seed=3)
    total_p = sum(m.pressure_count for m, _ in specs)
    sims = [Simulation(m, nu, 0.005, device=dev, rtol=1e-8, precision="fp32_refined")
            for m, nu in specs]
    tune = autotune_batch(sims)
    log(f"cavity_batch64: autotune {tune}")
    for s in sims:
        s.reset()
    dev.synchronize()
    t = time.perf_counter()
    stats = step_batch(sims, steps)
    batched = time.perf_counter() - t
    for s in sims:
        s.reset()
    dev.synchronize()
    t = time.perf_counter()
    for s in sims:
        s.step(steps)
    serial = time.perf_counter() - t
    # the same batch with a coarse-grid correction (P1 from the n/2 cavity)
    # for every even n, the instances' coarse solves in one batched launch
    coarse_count = 0
    for (m, _), s in zip(specs, sims):
        n = m.info["n"]
        if n % 2 == 0:
            s.set_coarse_space(mesh.cavity_prolongation(n, 2))
            coarse_count += 1
        s.reset()
    dev.synchronize()
    t = time.perf_counter()
    stats_c = step_batch(sims, steps)
    batched_c = time.perf_counter() - t
    log(f"cavity_batch64: batched {steps / batched:.2f} batch-steps/s, one by one "
        f"{steps / serial:.2f}; batched with coarse corrections ({coarse_count} instances) "
        f"{steps / batched_c:.2f}")
    its = [sum(st["cg_iterations"][k] for st in stats) / steps for k in range(5)]
    its_c = [sum(st["cg_iterations"][k] for st in stats_c) / steps for k in range(5)]
    for s in sims:
        s.close()
    return _line("batch steps/sec (all 64 instances advanced one step), 1e-8 rel. residual",
                 steps / batched, "steps/s", steps, 1e3 * batched / steps,
                 {"workload": f"64 lid-driven cavities, n ~ U{{16..256}} (seed 3), "
                              f"{total_p} pressure nodes in total, lid ~ U(0.5,2), "
                              f"nu ~ logU(1e-3,1e-2), dt=0.005, 50 steps", "form": "fp32_refined"},
                 dtype="f32+f64refine", autotune=tune,
                 one_by_one_steps_per_s=round(steps / serial, 3),
                 batching_speedup=round(serial / batched, 3),
                 cg_iterations_per_batch_step=its,
                 with_coarse_correction={
                     "steps_per_s": round(steps / batched_c, 3),
                     "instances": coarse_count,
                     "cg_iterations_per_batch_step": its_c,
                     "coarse_iterations_per_batch_step":
                         sum(st["coarse_iterations"] for st in stats_c) / steps})


def bench_learned_channels(dev, args):
    """Learned warm start: trained on (rhs, pressure) pairs of 512x128
    channels with inflow max ~ U(0.5, 2) and reseeded jitter; evaluated on
    held-out seeds at 512x128 and 1024x256 (pressure CG iterations/step and
    steps/s per initial guess)."""
    import numpy as np

    from navier_stokes_fem_amd import learned as L
    from navier_stokes_fem_amd import mesh

    rng = np.random.default_rng(2024)
    n_train, per = 50, 400  # 50 meshes x 400 steps = 20000 pairs per target
    train = [mesh.channel(512, 128, seed=1000 + k, umax=float(rng.uniform(0.5, 2.0)))
             for k in range(n_train)]
    t = time.perf_counter()
    sets = L.collect_pair_sets(dev, train, dt=5e-4, steps=per, warmup=10, every=1,
                               targets=("solution", "correction"), log=log)
    log(f"learned_channels: {len(sets['solution'])} solution and {len(sets['correction'])} "
        f"correction pairs in {time.perf_counter() - t:.1f} s")
    nets = {}
    for target in ("solution", "correction"):
        pairs = sets.pop(target)
        t = time.perf_counter()
        net = L.WarmStartNet(dev, seed=1)
        hist = L.train(net, pairs, epochs=4, lr=5e-4, lr_final=2e-5, log=log, log_every=20000)
        log(f"learned_channels: {target} training loss per epoch {[round(h, 4) for h in hist]} "
            f"({time.perf_counter() - t:.1f} s)")
        nets[target] = (net, hist, len(pairs))
        for g, b, p in pairs:
            b.close()
            p.close()
    held_sets = L.collect_pair_sets(dev, [mesh.channel(512, 128, seed=900001, umax=1.25)],
                                    dt=5e-4, steps=100, warmup=10, every=5,
                                    targets=("solution", "correction"))
    held_loss = {}
    for target, hp in held_sets.items():
        held_loss[target] = L.mean_loss(nets[target][0], hp)
        log(f"learned_channels: {target} held-out loss {held_loss[target]:.4f} ({len(hp)} pairs)")
        for g, b, p in hp:
            b.close()
            p.close()
    evals = {}
    for nx, ny, seed in ((512, 128, 900001), (1024, 256, 900002)):
        held = mesh.channel(nx, ny, seed=seed, umax=1.25)
        r = L.evaluate_warm_start(dev, held, nets["solution"][0], steps=20, warmup=10, dt=5e-4,
                                  modes=("previous", "zero", "learned", "learned_or_previous",
                                         "learned_correction", "learned_search", "projection",
                                         "learned_projection"),
                                  correction_net=nets["correction"][0])
        evals[f"{nx}x{ny}"] = r
        for mode, v in r.items():
            log(f"learned_channels {nx}x{ny} {mode}: {v['pressure_iterations_per_step']:.1f} its/step, "
                f"{v['steps_per_s']:.1f} steps/s")
    base = evals["1024x256"]["previous"]
    best_mode = min(evals["1024x256"], key=lambda k: evals["1024x256"][k]["pressure_iterations_per_step"])
    for net, _, _ in nets.values():
        net.close()
    ev = evals["1024x256"]
    gains = {size: {"learned_search_vs_previous":
                    1 - r["learned_search"]["pressure_iterations_per_step"]
                    / r["previous"]["pressure_iterations_per_step"],
                    "learned_projection_vs_projection":
                    1 - r["learned_projection"]["pressure_iterations_per_step"]
                    / r["projection"]["pressure_iterations_per_step"]}
             for size, r in evals.items()}
    return _line("pressure CG iterations/step with the learned warm start (A-norm-optimal step "
                 "from the previous pressure) vs previous-step guess (1024x256 held-out channel)",
                 ev["learned_search"]["pressure_iterations_per_step"],
                 "iterations/step", 20, 1e3 / ev["learned_search"]["steps_per_s"],
                 {"workload": "channel 512x128 training meshes (50 seeds x 400 steps = 20000 pairs, "
                              "inflow max ~ U(0.5,2)), "
                              "held-out seeds at 512x128 and 1024x256, nu=1e-3, dt=5e-4",
                  "form": "fp64"},
                 higher_is_better=False, dtype="bf16 net (fp32 accumulate) + f64 CG",
                 training={k: {"pairs": v[2], "loss_per_epoch": v[1],
                               "held_out_loss": held_loss[k]} for k, v in nets.items()},
                 evaluation=evals, previous_guess_iterations=base["pressure_iterations_per_step"],
                 network_gain=gains, best_mode=best_mode)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2048, help="cavity squares per side")
    ap.add_argument("--form", choices=["both", "fp32_refined", "fp64"], default="both")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tight", action="store_true",
                    help="skip the run solved to 1e-12 that the forms' accuracy is measured against")
    ap.add_argument("--no-attribution", action="store_true",
                    help="skip the attribution runs (the chosen form without the coarse "
                         "correction, and with the previous-pressure guess)")
    ap.add_argument("--cpu-sample-iters", type=int, default=10)
    ap.add_argument("--warm", choices=["previous", "extrapolate", "projection"],
                    default="previous",
                    help="initial guess of the fp32_refined form's pressure solves (default: the "
                         "previous pressure, the reference's; with the V-cycle preconditioned "
                         "pressure the A-norm projection onto past pressures saves about one "
                         "iteration a step and costs more than that: 18.0 against 18.2 steps/s, "
                         "fields identical to 1e-8); the fp64 form always starts from the "
                         "previous pressure, as the reference")
    ap.add_argument("--coarse-ratio", type=int, default=2,
                    help="coarse-grid correction of the fp32_refined form's pressure guesses: "
                         "P1 interpolation from the (n/ratio)^2 cavity (0 = off)")
    ap.add_argument("--coarse-rtol", type=float, default=1e-5,
                    help="relative tolerance of the coarse solves")
    ap.add_argument("--pair-velocity", type=int, default=1, choices=[0, 1],
                    help="fp32_refined form: solve the u and v columns of a velocity system as "
                         "one two-column CG while their iteration counts stay balanced (1) or "
                         "one after the other (0)")
    ap.add_argument("--velocity-jacobi", choices=["none", "mass", "diffusion", "both"],
                    default="mass",
                    help="fp32_refined form: these velocity solves as Jacobi-preconditioned CG "
                         "(same unscaled stopping rule; default mass: at 2048^2 its iterations "
                         "fall 43%%, the diffusion's 2%%)")
    ap.add_argument("--pressure-multigrid", action=argparse.BooleanOptionalAction, default=True,
                    help="fp32_refined form (default on): solve the pressure by FP64 CG "
                         "preconditioned by one V-cycle over the fine grid, the coarse space and "
                         "its levels; off: CG on the pressure with the coarse correction of its "
                         "guess")
    ap.add_argument("--velocity-multigrid", type=int, default=2, choices=[0, 1, 2],
                    help="fp32_refined form: the diffusion columns by FP64 CG preconditioned with "
                         "one V-cycle over the velocity nodes, the P1 vertex space and its grid "
                         "halvings (2, default: both columns -- at 2048^2 the v column takes 6-9 "
                         "iterations a step instead of ~90; 1: each column the way measured "
                         "faster; 0: off)")
    ap.add_argument("--mass-projection", type=int, default=1,
                    help="fp32_refined form: mass solves started from the A-norm projection of "
                         "the velocity correction onto this many past ones (default 1: at "
                         "2048^2 the mass iterations fall from 16 to 11 a step, 2 and 4 gain no "
                         "more; 0 = the reference's guess, the velocity)")
    ap.add_argument("--coarse-levels", type=int, default=-1,
                    help="multigrid levels below the coarse space (each halves the grid; -1 = "
                         "halve down to the 32x32 grid); with levels the coarse system is solved "
                         "by V-cycles instead of CG")
    ap.add_argument("--velocity-projection", type=int, default=2,
                    help="fp32_refined form: diffusion guesses from the advected velocity plus the "
                         "projection of the correction onto this many past ones (0 = off: the "
                         "reference's guess; on, the fields are closer to a tightly solved run and "
                         "the diffusion solves shorter, the pressure solves that follow longer)")
    ap.add_argument("--config", default="cavity2048",
                    choices=["cavity2048", "cavity32", "channel512", "cavity_batch64",
                             "learned_channels", "spmv"],
                    help="cavity2048 (default): the 2048^2 cavity headline; cavity32: 32^2 "
                         "cavity, fields vs the CPU oracle; channel512: 512x128 channel past a "
                         "disk; cavity_batch64: 64 cavities in one launch; learned_channels: "
                         "learned warm start trained and evaluated on channels; spmv: the "
                         "CSR product on the --n cavity's pressure and diffusion matrices")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only nsfem_saxpy on 2^26 doubles x 5 (known HBM bytes: "
                         "16 B read + 8 B written per element) to scale rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the multi-GPU design "
                                   "(row-strip halo exchange + 2 scalar all-reduces per CG "
                                   "iteration) is documented, not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    from navier_stokes_fem_amd import mesh
    from navier_stokes_fem_amd.device import GPUSimulationDevice

    if args.config != "cavity2048":
        runner = {"cavity32": bench_cavity32, "channel512": bench_channel512,
                  "cavity_batch64": bench_cavity_batch64,
                  "learned_channels": bench_learned_channels, "spmv": bench_spmv}[args.config]
        dev = GPUSimulationDevice(0)
        out = runner(dev, args)
        dev.close()
        print(json.dumps(out), flush=True)
        return

    if args.calibrate:
        dev = GPUSimulationDevice(0)
        n = 1 << 26
        x = dev.buffer(np.ones(n))
        y = dev.buffer(np.zeros(n))
        dev.saxpy(n, 0.5, x, y)  # first touch
        dev.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            dev.saxpy(n, 0.5, x, y)
        dev.synchronize()
        wall = time.perf_counter() - t
        got = y.download(np.float64, 4)
        print(json.dumps({"calibration": "saxpy", "elements": n, "launches": 5,
                          "fetch_bytes_per_launch": 16 * n, "write_bytes_per_launch": 8 * n,
                          "stream_GBs": round(5 * 24 * n / wall / 1e9, 1),
                          "check": float(got[0])}), flush=True)
        x.close()
        y.close()
        dev.close()
        return
    from navier_stokes_fem_amd.simulation import Simulation

    nu, dt = 1e-3, 1e-3
    log(f"building cavity {args.n}x{args.n} mesh")
    m = mesh.cavity(args.n)
    dev = GPUSimulationDevice(0)
    info = dev.info()
    log(f"device {info}")
    t = time.perf_counter()
    sim = Simulation(m, nu, dt, device=dev, rtol=1e-8, precision="fp64", pair_velocity=True)
    setup_s = time.perf_counter() - t
    log(f"simulation set-up (velocity order, device assembly, KD tree, tiles) {setup_s:.1f} s")
    n_p = m.pressure_count
    nnz_p = sim.matrix_info(0)[2]

    coarse = None
    if args.coarse_ratio:
        t = time.perf_counter()
        coarse = mesh.cavity_prolongation(args.n, args.coarse_ratio)
        log(f"prolongation from the {args.n // args.coarse_ratio}^2 cavity "
            f"({time.perf_counter() - t:.1f} s)")
    levels = coarse_levels(args.n // args.coarse_ratio, args.coarse_levels) if coarse else []
    if args.velocity_multigrid:
        t = time.perf_counter()
        vhier = mesh.velocity_hierarchy(m)
        t1 = time.perf_counter()
        sim.set_velocity_multigrid(vhier, mode=0)
        log(f"velocity hierarchy {[h[3] for h in vhier]}: prolongations {t1 - t:.1f} s, "
            f"Galerkin products {time.perf_counter() - t1:.1f} s")
    results = {}
    forms = ["fp64", "fp32_refined"] if args.form == "both" else [args.form]
    for form in forms:
        mixed_form = form == "fp32_refined"
        wall, st = run_form(sim, form, args.warmup, args.steps, n_p, nnz_p,
                            warm=args.warm if mixed_form else "previous",
                            coarse=coarse if mixed_form else None, coarse_rtol=args.coarse_rtol,
                            velocity_projection=args.velocity_projection if mixed_form else 0,
                            coarse_levels=levels,
                            pressure_multigrid=args.pressure_multigrid and mixed_form,
                            pair_velocity=bool(args.pair_velocity) or not mixed_form,
                            velocity_jacobi=args.velocity_jacobi if mixed_form else "none",
                            velocity_multigrid=args.velocity_multigrid if mixed_form else 0,
                            mass_projection=args.mass_projection if mixed_form else 0)
        results[form] = (wall, st)
        log(f"{form}: {args.steps / wall:.3f} steps/s")
    # attribution: the same mixed solver without the coarse correction, and
    # with the reference's previous-pressure guess alone
    # attribution: the mixed CG alone on every solve (no preconditioner, no
    # velocity guesses), from the A-norm projection of the pressure and from
    # the reference's previous pressure: the precision's share of the gain
    attribution = {}
    if not args.no_attribution and "fp32_refined" in forms:
        for warm in ("projection", "previous"):
            wall_a, st_a = run_form(sim, "fp32_refined", args.warmup, args.steps, n_p, nnz_p,
                                    warm=warm)
            attribution[f"fp32_refined_{warm}_guess"] = (wall_a, st_a)
            log(f"fp32_refined ({warm} guess, mixed CG alone): "
                f"{args.steps / wall_a:.3f} steps/s")

    # both forms stop at the same FP64 residual; the faster one is reported
    chosen = min(results, key=lambda f: results[f][0])
    wall, st = results[chosen]
    value = args.steps / wall
    mixed = chosen == "fp32_refined"
    # dominant kernel: the persistent CG kernel of the launch class (solve
    # kind, paired or not) with the most device time in the chosen form's
    # timed steps; the V-cycle preconditioned solves launch no CG kernel
    # (their products and sweeps are reported beside it).  At 2048^2 the
    # paired mass launch is the single kernel with the most time per step,
    # ahead of the diffusion's FP32 fine-level product inside the V-cycles.
    # achieved = the textbook iteration's bytes
    # (matrix once with int32 columns, row_start, 11 vector words per row
    # and column; w = 4 for the FP32 inner iteration, 8 for FP64) x the
    # launches' iterations / their HIP-event time; the build's own count,
    # with the FP64 solution and the reliable updates' residual passes, is
    # reported beside it
    kinds = [("diffusion u", 0, 1), ("diffusion v", 1, 1), ("pressure", 2, 1),
             ("mass u+v", 3, 2), ("mass v", 4, 1)]
    n_v, nnz_v = sim.matrix_info(2)[0], sim.matrix_info(2)[2]
    classes = launch_classes(st["per_step"])
    name = max((c for c in classes if classes[c][2][0] > 0), key=lambda c: classes[c][2][2])
    k, cols, (launches, k_iters, k_secs, k_refines) = classes[name]
    n_k, nnz_k = (n_p, nnz_p) if k == 2 else (n_v, nnz_v)
    w = 4 if mixed else 8
    textbook = (4 + w) * nnz_k + (4 + 11 * w * cols) * n_k
    bytes_textbook = textbook * k_iters
    bytes_k = cg_bytes(nnz_k, n_k, k_iters, k_refines + launches - 1, mixed,
                       columns=cols)
    achieved = bytes_textbook / max(k_secs, 1e-12) / 1e9
    measured = MEASURED_BYTES_PER_ITERATION.get((mixed, name)) if args.n == 2048 else None
    traffic = measured * k_iters / launches if measured else None
    if cols == 2 and mixed and os.environ.get("NSFEM_CG_PAIRED") != "twophase":
        # the single-reduction two-column iteration: form, block size and
        # loads in flight as the library picks them
        form = "packed" if os.environ.get("NSFEM_CG_PAIRED") == "packed" else "narrow-gather"
        bs = 768 if os.environ.get("NSFEM_CG_P2_BLOCK") == "768" else 1024
        ib = os.environ.get("NSFEM_CG_P2_LOADS")
        ib = int(ib) if ib in ("4", "8") else (8 if form == "narrow-gather" else 4)
        kernel = "k_cg_p2<%d, short, %d>, %s form" % (bs, ib, form)
    elif cols == 2:
        kernel = "k_cg<2, 1024, %s, short, 8>" % ("true" if mixed else "false")
    elif mixed:
        kernel = "k_cg<1, 768, true, short, 1>"
    else:
        kernel = "k_cg<1, 1024, false, short, 8>"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": f"{kernel} ({name} CG, persistent)", "solve": name,
                "launches": launches, "iterations_per_launch": k_iters / launches,
                "bytes_per_iteration": textbook,
                "bytes_per_launch": bytes_textbook / launches,
                "seconds_per_launch": k_secs / launches,
                "with_refinement_passes": {
                    "bytes_per_launch": bytes_k / launches,
                    "achieved": round(bytes_k / max(k_secs, 1e-12) / 1e9, 1),
                    "frac": round(bytes_k / max(k_secs, 1e-12) / 1e9 / PEAK_HBM_GBS, 4)},
                "seconds_per_step_by_solve": {
                    kn: round(st["cg_seconds"][kk] / args.steps, 5) for kn, kk, _ in kinds},
                "vcycle_seconds_per_step_by_solve": {
                    kn: round(st["vcycle_seconds"][kk] / args.steps, 5) for kn, kk, _ in kinds
                    if st["vcycle_solves"][kk]}}
    out = {
        "metric": "steps/sec at 1e-8 rel. residual, chosen vs plain FP64 CG; achieved HBM GB/s",
        "value": round(value, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64refine" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {args.n}x{args.n} squares "
                               f"({m.element_count} P2-P1 triangles, {m.node_count} velocity / "
                               f"{m.pressure_count} pressure nodes), nu=1e-3, dt=1e-3, "
                               f"rel tol 1e-8 per solve",
                   "form": chosen,
                   "pair_velocity": bool(args.pair_velocity) if mixed else True,
                   "parallelism": "single GPU",
                   "diffusion_initial_guess": (
                       f"advected velocity + projection of the correction (depth "
                       f"{args.velocity_projection})" if chosen == "fp32_refined" and
                       args.velocity_projection else "advected velocity"),
                   "pressure_initial_guess": (
                       (args.warm + (f" + coarse-grid correction (P1 from the "
                                     f"{args.n // args.coarse_ratio}^2 cavity, coarse system "
                                     + (f"by V-cycles over {len(levels)} more levels"
                                        if levels else "by CG")
                                     + f" to {args.coarse_rtol:g})"
                                     if coarse is not None and not args.pressure_multigrid
                                     else ""))
                       if chosen == "fp32_refined" else "previous"),
                   "pressure_solve": (
                       f"FP64 CG preconditioned by one FP32 V-cycle (damped Jacobi 2+2 sweeps "
                       f"on the fine grid, the P1 space of the {args.n // args.coarse_ratio}^2 "
                       f"cavity and {len(levels)} more levels, the deepest by its dense "
                       f"inverse), to the same 1e-8 relative residual"
                       if chosen == "fp32_refined" and args.pressure_multigrid and coarse is not None
                       else ("FP32 + FP64 refinement CG" if chosen == "fp32_refined"
                             else "FP64 CG")),
                   "mass_initial_guess": (
                       f"projection of the velocity correction onto its last "
                       f"{args.mass_projection} (A-norm, mass matrix)"
                       if chosen == "fp32_refined" and args.mass_projection
                       else "the velocity (the reference's)"),
                   "diffusion_solve": (
                       ("FP64 CG preconditioned by one FP32 V-cycle (damped Jacobi 2+2 sweeps "
                        "on the P2 nodes, the P1 vertex space and its grid halvings, the deepest "
                        "by its dense inverse) for "
                        + ("each column when measured faster than the FP32 + FP64 refinement CG"
                           if args.velocity_multigrid == 1 else "both columns"))
                       if chosen == "fp32_refined" and args.velocity_multigrid
                       else ("FP32 + FP64 refinement CG" if chosen == "fp32_refined"
                             else "FP64 CG")),
                   "mass_solve": (
                       ("FP32 + FP64 refinement CG"
                        + (" on D^-1/2 M D^-1/2 (Jacobi), the same unscaled stopping rule"
                           if args.velocity_jacobi in ("mass", "both") else "")
                        + ", u and v as one two-column solve while that is measured faster")
                       if chosen == "fp32_refined" else "FP64 CG")},
        "roofline": roofline,
    }
    for form, (w, s) in list(results.items()) + list(attribution.items()):
        out[form] = {"steps_per_s": round(args.steps / w, 4), "ms_per_step": round(1e3 * w / args.steps, 3),
                     "cg_iterations": s["cg_iterations"], "max_relative_residual": s["max_relative_residual"],
                     "cg_seconds": [round(x, 4) for x in s["cg_seconds"]],
                     "cg_refinements": s["cg_refinements"],
                     "coarse_iterations": s["coarse_iterations"],
                     "coarse_seconds": round(s["coarse_seconds"], 4),
                     "vcycle_solves": s["vcycle_solves"],
                     "vcycle_iterations": s["vcycle_iterations"],
                     "advect_seconds": round(s["advect_seconds"], 4)}
    # SpMV micro-benchmark on the assembled pressure and diffusion matrices
    out["spmv"] = {}
    for name, slot in (("pressure_stiffness", 0), ("diffusion", 2)):
        r_, c_, z_ = sim.matrix_info(slot)
        out["spmv"][name] = spmv_bench(dev, sim, slot, r_, c_, z_)
        log(f"spmv {name}: {out['spmv'][name]}")
    out["setup_seconds"] = round(setup_s, 2)
    # accuracy at equal tolerance: every form's distance to the same steps
    # solved to 1e-12 (the chosen form's solver chain at 1e-12: its FP64
    # residual makes it as exact as an FP64 solve, at a fraction of the time)
    if not args.no_tight and "fp32_refined" in forms:
        t = time.perf_counter()
        w_t, st_t = run_form(sim, "fp32_refined", args.warmup, args.steps, n_p, nnz_p,
                             warm=args.warm, coarse=coarse, coarse_rtol=min(args.coarse_rtol, 1e-8),
                             velocity_projection=args.velocity_projection, coarse_levels=levels,
                             rtol=1e-12, pressure_multigrid=args.pressure_multigrid,
                             pair_velocity=bool(args.pair_velocity),
                             velocity_jacobi=args.velocity_jacobi,
                             velocity_multigrid=args.velocity_multigrid,
                             mass_projection=args.mass_projection)
        log(f"tight reference (1e-12) {time.perf_counter() - t:.1f} s, max residuals "
            f"{st_t['max_relative_residual']}")
        tight = st_t["fields"]
        out["error_vs_tight_solve"] = {
            "reference": f"the chosen form's chain solved to 1e-12 (coarse solves to "
                         f"{min(args.coarse_rtol, 1e-8):g}), same {args.warmup + args.steps} "
                         f"steps from rest; relative L2 per field",
            **{form: {k: v["rel_l2"] for k, v in field_difference(s_["fields"], tight).items()}
               for form, (w_, s_) in list(results.items()) + list(attribution.items())}}
    if "fp64" in results and "fp32_refined" in results:
        out["speedup_vs_fp64"] = round((args.steps / results["fp32_refined"][0]) /
                                       (args.steps / results["fp64"][0]), 3)
        out["field_difference_vs_fp64"] = field_difference(results["fp32_refined"][1]["fields"],
                                                           results["fp64"][1]["fields"])
        for name, (w, st_a) in attribution.items():
            out[name]["field_difference_vs_fp64"] = field_difference(st_a["fields"],
                                                                      results["fp64"][1]["fields"])
    sim.close()
    dev.close()  # release the device before interpreter teardown
    if not args.no_cpu:
        ref = results.get("fp64", results[chosen])[1]
        its = [x / args.steps for x in ref["cg_iterations"]]
        out["cpu_baseline"] = cpu_baseline(m, nu, dt, its, args.cpu_sample_iters)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
