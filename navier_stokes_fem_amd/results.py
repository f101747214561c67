"""Result files of a simulation and the time-step convergence estimate read
from them.

The reference writes one JSON file per exported time step,
``<folder>/out_<step>.json`` with the velocity components ``u``, ``v`` (one
value per velocity node), the pressure ``p`` (one per pressure node) and
``timeStep`` (include/assembly.h:953-979; its SVG plot beside it is not
written here).  Its driver reads those files back from three runs whose time
steps shrink tenfold (dt = 0.01, 0.001, 0.0001 over the same total time) and
prints a Runge estimate of the convergence order and of the error at every
common time (cpp/main.cpp:22-84).

    export_solution("run_0", 12, u, v, p)
    sol = read_solution("run_0/out_12.json")
    est = runge_estimates(["run_0", "run_1", "run_2"], total_steps=100)
"""
from __future__ import annotations

import json
import os

import numpy as np

__all__ = ["export_solution", "read_solution", "solution_path", "runge_estimates",
           "export_simulation"]


def solution_path(folder: str, time_step: int) -> str:
    """``<folder>/out_<time_step>.json`` (include/assembly.h:973)."""
    return os.path.join(folder, f"out_{int(time_step)}.json")


def export_solution(folder: str, time_step: int, u, v, p) -> str:
    """Write one time step as the reference does (include/assembly.h:
    958-972): arrays ``u``, ``v`` (velocity nodes), ``p`` (pressure nodes)
    and the integer ``timeStep``; values are written with full double
    precision.  Returns the path.  An empty folder name exports nothing
    (the reference's ``outFolder.empty()`` early return) and returns None."""
    if not folder:
        return None
    u = np.asarray(u, np.float64).reshape(-1)
    v = np.asarray(v, np.float64).reshape(-1)
    p = np.asarray(p, np.float64).reshape(-1)
    if u.shape != v.shape:
        raise ValueError("u and v must have one value per velocity node each")
    doc = {"u": u.tolist(), "v": v.tolist(), "timeStep": int(time_step), "p": p.tolist()}
    path = solution_path(folder, time_step)
    with open(path, "w") as f:
        json.dump(doc, f, indent=4)
        f.write("\n")
    return path


def read_solution(path: str) -> dict:
    """Read an ``out_<step>.json`` file (ours or the reference's): ``u``,
    ``v``, ``p`` as float64 arrays (``p`` empty if absent) and ``timeStep``."""
    with open(path) as f:
        doc = json.load(f)
    if "u" not in doc or "v" not in doc:
        raise ValueError(f"{path}: not a solution file (no 'u'/'v' arrays)")
    u = np.asarray(doc["u"], np.float64)
    v = np.asarray(doc["v"], np.float64)
    if u.shape != v.shape:
        raise ValueError(f"{path}: 'u' and 'v' differ in length")
    return {"u": u, "v": v, "p": np.asarray(doc.get("p", []), np.float64),
            "timeStep": int(doc.get("timeStep", -1))}


def export_simulation(sim, folder: str, time_step: int) -> str:
    """Export the current fields of a ``Simulation`` (one device download)."""
    u, v, p = sim.fields()
    return export_solution(folder, time_step, u, v, p)


def runge_estimates(folders, total_steps: int, refinement: int = 10):
    """Runge estimate of the time-discretisation error from three runs whose
    time steps shrink by ``refinement`` from one folder to the next
    (cpp/main.cpp:48-83).  For each of the coarse run's ``total_steps``
    steps k, the runs' velocities at the common time (steps k,
    refinement*k, refinement^2*k) give

        a = ||w_0 - w_1||,  b = ||w_1 - w_2||   (w = [u, v], 2-norm),
        order = log10(|a / b|),
        error = 10^order * a / (10^order - 1);

    steps where a is 0 are skipped (order and error NaN).  The reference
    tests a twice and never b (cpp/main.cpp:77), so b = 0 with a > 0 gives an
    infinite order and a NaN error here too.  Returns (order, error) arrays of
    length total_steps."""
    if len(folders) != 3:
        raise ValueError("three runs are needed")
    order = np.full(total_steps, np.nan)
    error = np.full(total_steps, np.nan)
    for k in range(total_steps):
        w = []
        for i, folder in enumerate(folders):
            sol = read_solution(solution_path(folder, k * refinement ** i))
            w.append(np.concatenate([sol["u"], sol["v"]]))
        if not (w[0].shape == w[1].shape == w[2].shape):
            raise ValueError("the three runs must be on the same mesh")
        a = float(np.sqrt(np.sum((w[0] - w[1]) ** 2)))
        b = float(np.sqrt(np.sum((w[1] - w[2]) ** 2)))
        if a == 0.0:
            continue
        with np.errstate(divide="ignore", invalid="ignore"):
            q = np.log10(np.abs(np.float64(a) / np.float64(b)))
            t = np.float64(10.0) ** q
            order[k] = q
            error[k] = t * a / (t - 1.0)
    return order, error
