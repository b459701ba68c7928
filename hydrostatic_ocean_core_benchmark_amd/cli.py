"""Experiment driver with the reference's CLI contract, routed to the v0.5
step on the GPU (main.py:295-359: `--config DIR`, repeated `--set KEY=VALUE`,
subcommands with `--case/--nx/--cfl/--steps`).

    python main.py [--config config] [--set key=value ...] run    [--case C] [--nx N] [--cfl X] [--steps K]
    python main.py ... verify --against PREFIX_state3d5.bin [--steps K]
    python main.py ... verify --case igw [--study space|time]
    python main.py ... bench  [--sweep 128,256,512] [--steps K] [--repeat R]
    python main.py ... verify3d [--case C] [--sweep 16,32,64] [--nz N] [--scale-nz] [--dim nx|nz]
    python main.py ... bench3d  [--case C] [--sweep 64,100,128] [--nz N] [--repeat R]

With --case igw | igw_broadband | geo_balance the subcommands drive the 2D
shallow-water model instead, as the reference's own run|verify|bench do
(main.py:45-275): run integrates one case and reports the error against its
exact solution, mass and energy drift; verify --study space|time is the
convergence study (observed order between refinements); bench is the
error-vs-wallclock frontier over the CFL sweep. verify3d and bench3d drive
the v0.2 3D model (main.py:199-264, libs/core/driver3d.py) over its five
cases: the convergence table against the exact solutions and the per-cell
cost across grid sizes.

run     integrate one case (lock_exchange or basin) and write the final state
        (the bundle's `_state3d5.bin` layout) and a metrics JSON into
        `<paths.output_dir>/run_<case>_<scheme>_<stamp>/`;
verify  the reference's backend gate (tools/compare_backends3d5.py:28-29,
        tools/compare_backends3d.py:52-67): integrate the case and compare
        every field with a state file written by another backend on the same
        problem, relative L2 < 1e-12 for one step and < 1e-9 for a run;
bench   per-cell cost of the step over grid sizes (ns/cell/step, median of
        `--repeat` timed runs after one warm-up run).

Without a 2D case, run|verify|bench drive the v0.5 step with the same
contract (flags, config tree, overrides, output directory). The lock_exchange
and basin cases are fixed workloads (their physics and scheme are those of
the benchmark); of the config tree the v0.5 route reads `[run] backend`,
`[run] precision`, `basin.nz`, `basin.solver` and `paths.output_dir`.

[run] keys (config/settings.toml; the reference declares the same keys):
  backend    "hip" -- the only backend of this package; any other value
             (the reference's ref | fortran_cpu | openacc | cuda | py_gpu)
             exits with an error instead of running something else;
  precision  "float64" (the plain step, the reference's arithmetic) |
             "mixed" | "mixed_fp32_momentum" (theta + pcg_mg only; see
             HipStepper); the reference's whole-core "float32" is not
             offered and exits with an error. The 2D and v0.2 models run
             float64 only.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from datetime import datetime
from pathlib import Path

import numpy as np

from . import cases as C
from .config import apply_overrides, load_config

TOL_ONE_STEP = 1e-12
TOL_FULL_RUN = 1e-9


def barotropic_dt_max(domain, g: float) -> float:
    """Surface-gravity CFL limit of the deepest wet column
    (libs/core/driver3d_v05.py:97-100)."""
    c = np.sqrt(g * float(np.max(domain.H[domain.mask > 0])))
    return 1.0 / (c * np.sqrt(1.0 / domain.grid.dx ** 2 + 1.0 / domain.grid.dy ** 2))


def field_error(a: np.ndarray, b: np.ndarray, state_scale: float) -> float:
    """Relative L2 of a against b; below 1e-11 of the state scale the error
    is normalised by the state scale instead (tools/compare_backends3d.py:52-67)."""
    num = float(np.sqrt(np.sum((a - b) ** 2)))
    den = float(np.sqrt(np.sum(b * b)))
    if den > 1e-11 * state_scale:
        return num / den
    return num / state_scale if state_scale > 0.0 else num


PRECISIONS = {"float64": "fp64", "fp64": "fp64", "mixed": "mixed",
              "mixed_fp32_momentum": "mixed_fp32_momentum"}


def run_settings(config) -> tuple[str, str]:
    """([run] backend, the HipStepper precision for [run] precision); an
    unsupported value exits with an error naming it."""
    backend = str(config.get("run.backend", "hip"))
    if backend != "hip":
        raise SystemExit(f"run.backend = {backend!r} is not this package's backend: it runs "
                         "'hip' (the MI355X kernels); set run.backend = \"hip\"")
    precision = str(config.get("run.precision", "float64"))
    if precision not in PRECISIONS:
        raise SystemExit(f"run.precision = {precision!r} is not offered "
                         "(float64 | mixed | mixed_fp32_momentum)")
    return backend, PRECISIONS[precision]


def build_problem(config, case: str | None, nx: int | None, cfl: float | None,
                  steps: int | None) -> C.Problem:
    name = case or str(config.get("case.name", "lock_exchange"))
    if name == "lock_exchange":
        prob = C.lock_exchange(nx=nx or 128)
    elif name == "basin":
        prob = C.basin(nx=nx or 512, nz=int(config.get("basin.nz", 40)),
                       solver=str(config.get("basin.solver", "pcg_mg")))
    else:
        raise SystemExit(f"unknown case '{name}' (this backend runs: lock_exchange, basin)")
    if cfl is not None:
        prob.dt = cfl * barotropic_dt_max(prob.domain, prob.physics.g)
    if steps is not None:
        prob.n_steps = int(steps)
    prob.meta["precision"] = run_settings(config)[1]
    return prob


CASES_2D = ("igw", "igw_broadband", "geo_balance")


def simulate2d(config, nx: int, cfl: float | None, case_name: str,
               n_steps_override: int | None = None) -> dict:
    """One 2D run (main.py:45-98): dt from the CFL of the forward-backward
    limit (or pinned by time.dt), n_steps to land on the case's t_final,
    metrics against the exact solution."""
    from . import cases2d as C2
    from .domain import CGrid
    from .model2d import SWStepper, State, explicit_dt_max, physics_from_config
    from .types import SchemeParams
    grid = CGrid(nx=nx, ny=nx, Lx=float(config.get("grid.Lx")), Ly=float(config.get("grid.Ly")))
    physics = physics_from_config(config)
    params = SchemeParams(name=str(config.get("scheme.name")),
                          theta=float(config.get("scheme.theta")),
                          theta_cor=float(config.get("scheme.theta_cor")),
                          n_picard=int(config.get("scheme.n_picard")),
                          solver=str(config.get("solver.kind")),
                          rtol=float(config.get("solver.rtol")),
                          max_iter=int(config.get("solver.max_iter")))
    case, spec = C2.build_case(case_name, grid, physics, config)
    dt_max = explicit_dt_max(grid, physics)
    pinned = float(config.get("time.dt", 0.0))
    dt_req = pinned if pinned > 0.0 else (cfl if cfl is not None
                                          else float(config.get("time.cfl_factor"))) * dt_max
    t_final = spec.t_final
    if n_steps_override is not None:
        n_steps, dt = int(n_steps_override), dt_req
        t_final = dt * n_steps
    else:
        n_steps = max(1, int(round(t_final / dt_req)))
        dt = t_final / n_steps
    import torch
    st = SWStepper(grid, physics, params, dt)
    s0 = case.initial()
    m0, e0 = C2.total_mass(s0, grid), C2.total_energy(s0, grid, physics)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fin = st.integrate(s0, n_steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    final = State(u=fin.u.cpu().numpy(), v=fin.v.cpu().numpy(), eta=fin.eta.cpu().numpy(),
                  t=fin.t)
    exact = case.exact(t_final)
    metrics = C2.summarize(final, exact, grid, physics, m0, e0)
    if case_name == "igw":
        metrics.update(C2.wave_errors(final, exact, grid, case.k, case.l))
    metrics.update({"nx": grid.nx, "ny": grid.ny, "dx": grid.dx, "dt": dt,
                    "dt_max_explicit": dt_max, "cfl": dt / dt_max, "n_steps": n_steps,
                    "t_final": t_final, "scheme": params.name, "theta": params.theta,
                    "n_picard": params.n_picard, "solver": params.solver,
                    "solver_iterations": st.solver_iterations,
                    "solver_failures": st.solver_failures, "wall_s": wall})
    return {"state": final, "exact": exact, "metrics": metrics, "params": params}


def _order(errors, resolution):
    """Observed order between successive refinements (main.py:101-118)."""
    out = [float("nan")]
    for e0, e1, r0, r1 in zip(errors[:-1], errors[1:], resolution[:-1], resolution[1:]):
        ok = e0 > 0.0 and e1 > 0.0 and r0 > 0.0 and r1 > 0.0 and r1 != r0
        out.append(float(np.log(e0 / e1) / np.log(r1 / r0)) if ok else float("nan"))
    return out


def _run_dir(config, tag: str) -> Path:
    stamp = datetime.now().strftime("%Y%m%d-%H%M%S")
    path = Path(str(config.get("paths.output_dir", "output"))) / f"{tag}_{stamp}"
    path.mkdir(parents=True, exist_ok=True)
    return path


def _integrate(prob: C.Problem, n_warmup: int = 0, n_repeat: int = 1):
    import torch

    from .stepper import HipStepper
    from .types import State3D
    g = prob.domain.grid
    try:
        st = HipStepper(prob.domain, prob.physics, prob.params, prob.dt,
                        precision=prob.meta.get("precision", "fp64"))
    except ValueError as exc:
        raise SystemExit(f"{prob.name}: {exc}") from exc
    b = st.buoyancy_of(prob.initial_T, prob.initial_S) * torch.as_tensor(
        prob.domain.mask3, device=st.device)
    s0 = State3D(u=np.zeros(g.shape3d), v=np.zeros(g.shape3d), b=b, eta=np.zeros(g.shape),
                 T=prob.initial_T, S=prob.initial_S)
    for _ in range(n_warmup):
        st.reset_mixing()
        st.integrate(s0, min(prob.n_steps, 3))
    walls, final, it0 = [], s0, st.solver_iterations
    for _ in range(max(1, n_repeat)):
        st.reset_mixing()
        it0 = st.solver_iterations
        torch.cuda.synchronize(st.device)
        t0 = time.perf_counter()
        final = st.integrate(s0, prob.n_steps)
        torch.cuda.synchronize(st.device)
        walls.append(time.perf_counter() - t0)
    return st, final, float(np.median(walls)), st.solver_iterations - it0


def _host_state(final) -> dict:
    out = {}
    for k in ("eta", "u", "v", "b", "T", "S"):
        a = getattr(final, k)
        if a is not None:
            out[k] = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    return out


def cmd_run2d(config, args, case_name) -> int:
    res = simulate2d(config, args.nx or int(config.get("grid.nx")), args.cfl, case_name,
                     n_steps_override=args.steps)
    m = res["metrics"]
    out = _run_dir(config, f"run_{case_name}_{m['scheme']}")
    np.savez(out / "state2d.npz", u=res["state"].u, v=res["state"].v, eta=res["state"].eta)
    (out / "metrics.json").write_text(json.dumps(m, indent=2) + "\n", encoding="utf-8")
    print(f"case={case_name} nx={m['nx']} scheme={m['scheme']} theta={m['theta']} "
          f"dt={m['dt']:.2f}s steps={m['n_steps']} solver={m['solver']} "
          f"iters={m['solver_iterations']}")
    print(f"L2(eta)={m['l2_rel_eta']:.4e}  mass drift={m['mass_drift']:+.3e}  "
          f"energy drift={m['energy_drift']:+.3e}  -> {out}")
    return 0


def cmd_verify2d(config, args, case_name) -> int:
    if args.study == "space":
        sizes = list(config.get("grid.sweep.nx_list"))
        rows = [simulate2d(config, int(n), args.cfl, case_name)["metrics"] for n in sizes]
        errors = [r["l2_rel_eta"] for r in rows]
        resolution = [float(r["nx"]) for r in rows]
        label, axis = "nx", [float(r["nx"]) for r in rows]
    else:
        nx = args.nx or int(config.get("grid.nx"))
        cfls = sorted(config.get("time.sweep.cfl_list"), reverse=True)
        ref = simulate2d(config, nx, min(cfls) / 8.0, case_name)["state"]
        rows, errors = [], []
        for cfl in cfls:
            res = simulate2d(config, nx, cfl, case_name)
            err = float(np.sqrt(np.sum((res["state"].eta - ref.eta) ** 2))
                        / np.sqrt(np.sum(ref.eta ** 2)))
            res["metrics"]["l2_rel_eta_vs_reference"] = err
            rows.append(res["metrics"])
            errors.append(err)
        resolution = [1.0 / r["dt"] for r in rows]
        label, axis = "1/cfl", [1.0 / r["cfl"] for r in rows]
    orders = _order(errors, resolution)
    print(f"--- convergence study: {args.study}, case={case_name}, scheme={rows[0]['scheme']}, "
          f"theta={rows[0]['theta']}, n_picard={rows[0]['n_picard']} ---")
    print(f"{label:>8} {'dx [km]':>10} {'dt [s]':>10} {'cfl':>7} {'L2(eta)':>12} {'order':>7} "
          f"{'amp_ratio':>10} {'phase[rad]':>11}")
    for row, err, order, ax in zip(rows, errors, orders, axis):
        row["observed_order"] = order
        print(f"{ax:>8.3g} {row['dx'] / 1000:>10.2f} {row['dt']:>10.2f} {row['cfl']:>7.3f} "
              f"{err:>12.4e} {order:>7.2f} {row.get('amp_ratio', float('nan')):>10.6f} "
              f"{row.get('phase_err_rad', float('nan')):>11.3e}")
    out = _run_dir(config, f"verify_{args.study}_{case_name}")
    (out / "metrics.json").write_text(json.dumps({"study": args.study, "case": case_name,
                                                  "rows": rows}, indent=2) + "\n",
                                      encoding="utf-8")
    return 0


def cmd_bench2d(config, args, case_name) -> int:
    nx = args.nx or int(config.get("grid.nx"))
    rows = []
    print(f"{'cfl':>7} {'dt [s]':>10} {'steps':>7} {'L2(eta)':>12} {'wall [s]':>10} {'us/step':>9}")
    for cfl in config.get("time.sweep.cfl_list"):
        m = simulate2d(config, nx, float(cfl), case_name, n_steps_override=args.steps)["metrics"]
        m["us_per_step"] = 1e6 * m["wall_s"] / m["n_steps"]
        rows.append(m)
        print(f"{cfl:>7.3f} {m['dt']:>10.2f} {m['n_steps']:>7d} {m['l2_rel_eta']:>12.4e} "
              f"{m['wall_s']:>10.4f} {m['us_per_step']:>9.1f}")
    out = _run_dir(config, f"bench_{case_name}")
    (out / "metrics.json").write_text(json.dumps({"mode": "bench", "case": case_name, "nx": nx,
                                                  "rows": rows}, indent=2) + "\n",
                                      encoding="utf-8")
    return 0


def _rel_l2(a, b) -> float:
    den = float(np.sqrt(np.sum(b * b)))
    num = float(np.sqrt(np.sum((a - b) ** 2)))
    return num / den if den > 0.0 else num


def simulate3d(config, nx: int, nz: int, cfl: float | None, case_name: str,
               n_steps_override: int | None = None, n_repeat: int = 1) -> dict:
    """One v0.2 3D run (libs/core/driver3d.py:53-126): dt from the CFL of
    the forward-backward limit (or pinned by time.dt), n_steps to land on the
    case's t_final (or pinned), the device Stepper3D integrated n_repeat
    times (median wall time), metrics against the exact solution."""
    import torch
    from . import cases3d_v02 as C3
    from .domain import CGrid
    from .model2d import explicit_dt_max
    from .model3d_v02 import Stepper3D
    from .types import SchemeParams, State3D
    grid = CGrid(nx=nx, ny=nx, Lx=float(config.get("grid.Lx")), Ly=float(config.get("grid.Ly")),
                 bc_x=str(config.get("grid.bc_x")), bc_y=str(config.get("grid.bc_y")),
                 nz=nz, depth=float(config.get("physics.H")))
    physics = C3.physics3d_from_config(config)
    params = SchemeParams(name=str(config.get("scheme.name")),
                          theta=float(config.get("scheme.theta")),
                          theta_cor=float(config.get("scheme.theta_cor")),
                          theta_v=float(config.get("scheme.theta_v", 0.5)),
                          advection=str(config.get("scheme.advection", "none")),
                          n_split=int(config.get("scheme.n_split", 0)),
                          barotropic_average=str(config.get("scheme.barotropic_average", "none")),
                          n_picard=int(config.get("scheme.n_picard")),
                          solver=str(config.get("solver.kind")),
                          rtol=float(config.get("solver.rtol")),
                          max_iter=int(config.get("solver.max_iter")))
    params.validate()
    case, t_final = C3.build_case3d(case_name, grid, physics, config)
    dt_max = explicit_dt_max(grid, physics)
    pinned = float(config.get("time.dt", 0.0))
    factor = cfl if cfl is not None else float(config.get("time.cfl_factor"))
    dt_req = pinned if pinned > 0.0 else factor * dt_max
    if n_steps_override is not None:
        n_steps, dt = int(n_steps_override), dt_req
        t_final = dt * n_steps
    else:
        n_steps = max(1, int(round(t_final / dt_req)))
        dt = t_final / n_steps
    st = Stepper3D(grid, physics, params, dt)
    initial = case.initial()
    walls, fin = [], None
    for _ in range(max(1, n_repeat)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fin = st.integrate(initial, n_steps)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    final = State3D(u=fin.u.cpu().numpy(), v=fin.v.cpu().numpy(), b=fin.b.cpu().numpy(),
                    eta=fin.eta.cpu().numpy(), t=fin.t,
                    T=None if fin.T is None else fin.T.cpu().numpy(),
                    S=None if fin.S is None else fin.S.cpu().numpy())
    exact = case.exact(t_final)
    ordered = sorted(walls)
    median = ordered[len(ordered) // 2] if len(ordered) % 2 else \
        0.5 * (ordered[len(ordered) // 2 - 1] + ordered[len(ordered) // 2])
    cells = grid.nx * grid.ny * grid.nz
    solves = max(1, n_steps * params.n_picard)
    lim = C3.time_step_limits(grid, physics, u_scale=float(np.max(np.abs(initial.u))) or 1.0)
    dt_allowed = lim.for_scheme(params.name, params.advection, physics.A_h)
    mem = C3.memory_model(grid, params.name)
    per_step = median / n_steps
    sdpd = dt_allowed / per_step if per_step > 0 else float("nan")
    metrics = {
        "case": case_name, "nx": grid.nx, "ny": grid.ny, "nz": grid.nz, "cells": cells,
        "dx": grid.dx, "dz": grid.dz, "dt": dt, "dt_max_explicit": dt_max, "cfl": dt / dt_max,
        "n_steps": n_steps, "t_final": t_final, "scheme": params.name, "theta": params.theta,
        "theta_v": params.theta_v, "n_picard": params.n_picard, "solver": params.solver,
        "h_eff": st.h_eff, "n_split": st.n_split,
        "barotropic_substeps": st.barotropic_substeps,
        "solver_iterations": st.solver_iterations,
        "pcg_per_solve": st.solver_iterations / solves,
        "tridiagonal_solves": st.tridiagonal_solves,
        "l2_rel_u": _rel_l2(final.u, exact.u),
        "l2_rel_T": (_rel_l2(final.T - physics.T0, exact.T - physics.T0)
                     if final.T is not None else float("nan")),
        "l2_rel_b": _rel_l2(final.b, exact.b) if np.any(exact.b) else float("nan"),
        "l2_rel_eta": _rel_l2(final.eta, exact.eta) if np.any(exact.eta) else float("nan"),
        "max_abs_eta": float(np.max(np.abs(final.eta))),
        "wall_s": median, "us_per_step": 1e6 * median / n_steps,
        "ns_per_cell_step": 1e9 * median / (n_steps * cells),
        "dt_limit_scheme": dt_allowed, "dt_limits": lim.to_dict(),
        "sim_seconds_per_wall_second": sdpd, "sim_days_per_wall_day": sdpd,
        "bytes": mem["bytes"], "bytes_per_cell": mem["bytes_per_cell"], "gib": mem["gib"],
        "max_nx_32gib_nz30": mem["max_nx_32gib_nz30"],
    }
    return {"grid": grid, "physics": physics, "params": params, "case": case,
            "state": final, "exact": exact, "metrics": metrics}


def cmd_verify3d(config, args) -> int:
    """Convergence study of the v0.2 3D model (main.py:199-237)."""
    case_name = args.case or "barotropic3d"
    sizes = [int(n) for n in (args.sweep or "16,32,64").split(",")]
    if args.dim == "nz":
        pairs = [(args.nx or int(config.get("grid.nx")), n) for n in sizes]
    elif args.scale_nz:
        pairs = [(n, max(1, n // 2)) for n in sizes]
    else:
        pairs = [(n, args.nz) for n in sizes]
    rows, errors = [], []
    for n, (nx_i, nz_i) in zip(sizes, pairs):
        m = simulate3d(config, nx_i, nz_i, args.cfl, case_name,
                       n_steps_override=(args.steps or 4 * n))["metrics"]
        field = {"baroclinic_igw": "l2_rel_b", "tracer_advect": "l2_rel_T"}.get(case_name,
                                                                                  "l2_rel_u")
        rows.append(m)
        errors.append(m[field])
    orders = _order(errors, [float(n) for n in sizes])
    print(f"--- 3D convergence: case={case_name}, scheme={rows[0]['scheme']}, "
          f"theta={rows[0]['theta']}, theta_v={rows[0]['theta_v']} ---")
    print(f"{'nx':>6}{'nz':>5}{'cells':>10}{'dt [s]':>10}{'L2':>13}{'order':>7}{'pcg/solve':>11}")
    for m, e, o in zip(rows, errors, orders):
        m["observed_order"] = o
        print(f"{m['nx']:>6}{m['nz']:>5}{m['cells']:>10}{m['dt']:>10.2f}"
              f"{e:>13.4e}{o:>7.2f}{m['pcg_per_solve']:>11.1f}")
    out = _run_dir(config, f"verify3d_{case_name}")
    (out / "metrics.json").write_text(json.dumps({"study": "3d", "case": case_name,
                                                  "rows": rows}, indent=2) + "\n",
                                      encoding="utf-8")
    return 0


def cmd_bench3d(config, args) -> int:
    """Per-cell cost of the v0.2 3D model across grid sizes (main.py:240-264)."""
    case_name = args.case or "baroclinic_igw"
    sizes = [int(n) for n in (args.sweep or "64,100,128").split(",")]
    rows = []
    print(f"{'nx':>6}{'nz':>5}{'cells':>10}{'steps':>7}{'wall [s]':>11}"
          f"{'ms/step':>10}{'ns/cell/step':>14}{'pcg/solve':>11}")
    for n in sizes:
        m = simulate3d(config, n, args.nz, args.cfl, case_name,
                       n_steps_override=(args.steps or 20), n_repeat=args.repeat)["metrics"]
        rows.append(m)
        print(f"{m['nx']:>6}{m['nz']:>5}{m['cells']:>10}{m['n_steps']:>7}"
              f"{m['wall_s']:>11.4f}{m['us_per_step'] / 1000:>10.2f}"
              f"{m['ns_per_cell_step']:>14.2f}{m['pcg_per_solve']:>11.1f}")
    out = _run_dir(config, f"bench3d_{case_name}")
    (out / "metrics.json").write_text(json.dumps({"mode": "bench3d", "case": case_name,
                                                  "rows": rows}, indent=2) + "\n",
                                      encoding="utf-8")
    return 0


def _case_2d(config, args):
    name = args.case or str(config.get("case.name", "lock_exchange"))
    return name if name in CASES_2D else None


def cmd_run(config, args) -> int:
    c2 = _case_2d(config, args)
    if c2:
        return cmd_run2d(config, args, c2)
    from .bundle import write_state_file
    prob = build_problem(config, args.case, args.nx, args.cfl, args.steps)
    st, final, wall, iters = _integrate(prob)
    out = _run_dir(config, f"run_{prob.name}_{prob.params.name}")
    host = _host_state(final)
    ts = prob.physics.tracers == "TS"
    write_state_file(out / "state3d5.bin", host, ts)
    g = prob.domain.grid
    metrics = {"case": prob.name, "nx": g.nx, "ny": g.ny, "nz": g.nz,
               "cells": g.nx * g.ny * g.nz, "dt": prob.dt, "n_steps": prob.n_steps,
               "scheme": prob.params.name, "solver": prob.params.solver,
               "backend": "hip", "precision": st.precision,
               "solver_iterations": iters, "solver_failures": st.solver_failures,
               "wall_s": wall, "eta_max": float(np.max(np.abs(host["eta"]))),
               "finite": bool(all(np.isfinite(a).all() for a in host.values()))}
    (out / "metrics.json").write_text(json.dumps(metrics, indent=2) + "\n", encoding="utf-8")
    print(f"case={prob.name} {g.nx}x{g.ny}x{g.nz} scheme={prob.params.name} "
          f"solver={prob.params.solver} precision={st.precision} dt={prob.dt:.3f}s "
          f"steps={prob.n_steps} "
          f"wall={wall:.3f}s iters={iters} -> {out}")
    return 0 if metrics["finite"] else 1


def cmd_verify(config, args) -> int:
    c2 = _case_2d(config, args)
    if c2:
        return cmd_verify2d(config, args, c2)
    if not args.against:
        raise SystemExit("verify of a 3D case needs --against PREFIX_state3d5.bin")
    from .bundle import read_state_file
    prob = build_problem(config, args.case, args.nx, args.cfl, args.steps)
    _, final, _, _ = _integrate(prob)
    g = prob.domain.grid
    ts = prob.physics.tracers == "TS"
    got = _host_state(final)
    want = read_state_file(args.against, g.nx, g.ny, g.nz, ts)
    tol = TOL_ONE_STEP if prob.n_steps == 1 else TOL_FULL_RUN
    scale = max(float(np.max(np.abs(w))) for w in want.values()) or 1.0
    worst = 0.0
    for k, w in want.items():
        e = field_error(got[k], w, scale)
        worst = max(worst, e)
        print(f"  {k:>4}: rel L2 {e:.3e}")
    ok = worst < tol
    print(f"{'PASS' if ok else 'FAIL'}: worst {worst:.3e} (gate {tol:g}, {prob.n_steps} steps)")
    return 0 if ok else 1


def cmd_bench(config, args) -> int:
    c2 = _case_2d(config, args)
    if c2:
        return cmd_bench2d(config, args, c2)
    sizes = [int(n) for n in (args.sweep or "128,256,512").split(",")]
    print(f"{'nx':>6}{'nz':>5}{'cells':>11}{'steps':>7}{'wall [s]':>11}{'ms/step':>10}"
          f"{'ns/cell/step':>14}{'its/step':>10}")
    rows = []
    for n in sizes:
        prob = build_problem(config, args.case or "basin", n, args.cfl, args.steps or 20)
        _, _, wall, iters = _integrate(prob, n_warmup=1, n_repeat=args.repeat)
        g = prob.domain.grid
        cells = g.nx * g.ny * g.nz
        row = {"nx": g.nx, "nz": g.nz, "cells": cells, "n_steps": prob.n_steps, "wall_s": wall,
               "ms_per_step": 1e3 * wall / prob.n_steps,
               "ns_per_cell_step": 1e9 * wall / (prob.n_steps * cells),
               "iters_per_step": iters / prob.n_steps}
        rows.append(row)
        print(f"{g.nx:>6}{g.nz:>5}{cells:>11}{prob.n_steps:>7}{wall:>11.4f}"
              f"{row['ms_per_step']:>10.3f}{row['ns_per_cell_step']:>14.3f}"
              f"{row['iters_per_step']:>10.1f}")
    out = _run_dir(config, f"bench_{args.case or 'basin'}")
    (out / "metrics.json").write_text(json.dumps({"mode": "bench", "rows": rows}, indent=2)
                                      + "\n", encoding="utf-8")
    return 0


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="hoc", description="MI355X ocean-core experiment driver")
    ap.add_argument("--config", default=str(Path(__file__).resolve().parents[1] / "config"),
                    help="configuration directory")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="override a dotted config key, e.g. --set scheme.theta=0.6")
    sub = ap.add_subparsers(dest="command", required=True)
    p_run = sub.add_parser("run", help="single simulation")
    p_ver = sub.add_parser("verify", help="backend gate against another backend's state file")
    p_ver.add_argument("--against", default=None,
                       help="3D cases: final state file (bundle layout) of another backend")
    p_ver.add_argument("--study", choices=("space", "time"), default="space",
                       help="2D cases: convergence in space (nx sweep) or time (cfl sweep)")
    p_ben = sub.add_parser("bench", help="step cost over grid sizes")
    p_ben.add_argument("--sweep", default=None, help="comma-separated nx list")
    p_ben.add_argument("--repeat", type=int, default=3)
    p_v3d = sub.add_parser("verify3d", help="3D convergence study (v0.2 model)")
    p_v3d.add_argument("--sweep", default=None, help="comma-separated nx list")
    p_v3d.add_argument("--scale-nz", action="store_true",
                       help="scale nz with nx (nz = nx/2) for a full 3D refinement")
    p_v3d.add_argument("--dim", choices=("nx", "nz"), default="nx",
                       help="which dimension the sweep refines (default nx)")
    p_b3d = sub.add_parser("bench3d", help="3D cost across grid sizes (v0.2 model)")
    p_b3d.add_argument("--sweep", default=None, help="comma-separated nx list")
    p_b3d.add_argument("--repeat", type=int, default=1)
    for sp in (p_v3d, p_b3d):
        sp.add_argument("--nz", type=int, default=30)
    for sp in (p_run, p_ver, p_ben, p_v3d, p_b3d):
        sp.add_argument("--case", default=None,
                        help="lock_exchange | basin (3D) | igw | igw_broadband | geo_balance (2D)"
                             " | barotropic3d | vdiffusion | baroclinic_igw | tracer_advect |"
                             " ekman (verify3d, bench3d)")
        sp.add_argument("--nx", type=int, default=None)
        sp.add_argument("--cfl", type=float, default=None,
                        help="dt as a fraction of the surface-gravity limit")
        sp.add_argument("--steps", type=int, default=None)
    return ap


def main(argv: list[str] | None = None) -> int:
    args = build_parser().parse_args(argv)
    try:
        config = apply_overrides(load_config(args.config), args.set)
    except ValueError as exc:
        raise SystemExit(str(exc)) from exc
    _, precision = run_settings(config)
    two_d = args.command in ("run", "verify", "bench") and _case_2d(config, args)
    if precision != "fp64" and (two_d or args.command in ("verify3d", "bench3d")):
        raise SystemExit(f"run.precision = {config.get('run.precision')!r}: the 2D and v0.2 "
                         "models run float64 only")
    return {"run": cmd_run, "verify": cmd_verify, "bench": cmd_bench, "verify3d": cmd_verify3d,
            "bench3d": cmd_bench3d}[args.command](config, args)


if __name__ == "__main__":
    sys.exit(main())
