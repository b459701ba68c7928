"""Experiment driver with the reference's CLI contract, routed to the v0.5
step on the GPU (main.py:295-359: `--config DIR`, repeated `--set KEY=VALUE`,
subcommands with `--case/--nx/--cfl/--steps`).

    python main.py [--config config] [--set key=value ...] run    [--case C] [--nx N] [--cfl X] [--steps K]
    python main.py ... verify --against PREFIX_state3d5.bin [--steps K]
    python main.py ... bench  [--sweep 128,256,512] [--steps K] [--repeat R]

run     integrate one case (lock_exchange or basin) and write the final state
        (the bundle's `_state3d5.bin` layout) and a metrics JSON into
        `<paths.output_dir>/run_<case>_<scheme>_<stamp>/`;
verify  the reference's backend gate (tools/compare_backends3d5.py:28-29,
        tools/compare_backends3d.py:52-67): integrate the case and compare
        every field with a state file written by another backend on the same
        problem, relative L2 < 1e-12 for one step and < 1e-9 for a run;
bench   per-cell cost of the step over grid sizes (ns/cell/step, median of
        `--repeat` timed runs after one warm-up run).

The reference's own `run|verify|bench` drive its 2D barotropic model, which is
not on this backend's path; the contract (flags, config tree, overrides,
output directory) is what is kept.
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
    return prob


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
    st = HipStepper(prob.domain, prob.physics, prob.params, prob.dt)
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


def cmd_run(config, args) -> int:
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
               "solver_iterations": iters, "solver_failures": st.solver_failures,
               "wall_s": wall, "eta_max": float(np.max(np.abs(host["eta"]))),
               "finite": bool(all(np.isfinite(a).all() for a in host.values()))}
    (out / "metrics.json").write_text(json.dumps(metrics, indent=2) + "\n", encoding="utf-8")
    print(f"case={prob.name} {g.nx}x{g.ny}x{g.nz} scheme={prob.params.name} "
          f"solver={prob.params.solver} dt={prob.dt:.3f}s steps={prob.n_steps} "
          f"wall={wall:.3f}s iters={iters} -> {out}")
    return 0 if metrics["finite"] else 1


def cmd_verify(config, args) -> int:
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
    p_ver.add_argument("--against", required=True, help="final state file (bundle layout)")
    p_ben = sub.add_parser("bench", help="step cost over grid sizes")
    p_ben.add_argument("--sweep", default=None, help="comma-separated nx list")
    p_ben.add_argument("--repeat", type=int, default=3)
    for sp in (p_run, p_ver, p_ben):
        sp.add_argument("--case", default=None, help="lock_exchange | basin")
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
    return {"run": cmd_run, "verify": cmd_verify, "bench": cmd_bench}[args.command](config, args)


if __name__ == "__main__":
    sys.exit(main())
