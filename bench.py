"""Benchmark: FP64 cell-steps/s of the ocean-core time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--kernel NAME]

Workload: the closed basin at 512 x 512 x 40 (dx = 10 km, dt = 600 s),
Gaussian-random-field bathymetry (correlation 200 km, 500-4000 m,
z-level partial cells), T(z) = 2 + 18 exp(z/800) + N(0, 0.05), S with a weak
linear gradient, tau_x = 0.1 cos(pi y / Ly), polyTEOS10, up3_tvd, TKE
(integral mixing length), theta = 0.5 (2 Picard passes, so 2 barotropic
solves per step) with the MG-preconditioned PCG (5 levels, 2 + 2 damped-Jacobi
sweeps at omega 0.8) at rtol 1e-12. Inputs are resident in HBM before the
timed region; W untimed steps, then exactly K timed steps bracketed by device
synchronisation.
A "step" is one full model time step. cells = nx*ny*nz (the reference's
convention, dry cells included).

The JSON line also carries
  roofline      the tracer kernel's algorithmic bytes per launch (stage E of
                the step's byte model, 72 B per layer cell) over its average
                launch time (HIP events on the launch stream, inside the timed
                region), against 8 TB/s; roofline_predictor the same for the
                predictor (the next largest 3D kernel) with the bytes of its
                stage of the model, stage B (56 B per layer cell; stage C also
                holds back-substitution 1, a kernel of its own), over 10
                further steps;
  step_model    the whole step against the stage model of the step's
                compulsory traffic: 336 B per layer cell plus, per barotropic
                solve, 64 B per column and B_it per column and iteration
                (B_it = 520 for MG-preconditioned PCG, 120 for PCG-Jacobi),
                with the measured iteration counts;
  mixed         the plain fp64 step and the mixed-precision step on the
                1024 x 1024 x 60 basin (dx 5 km, dt 300 s, 50 steps): cell-steps/s
                of each, their ratio, and the per-field relative L2 of the mixed
                final state against the fp64 one;
  ensemble      aggregate cell-steps/s of the 64-member ensemble of uneven
                grids (16 each at 128^2 .. 384^2, 30 levels) replayed from one
                HIP graph per two steps, beside the members run one by one;
  adjoint       the adjoint:forward wall-time ratio of dJ/dT0 and dJ/dkappa_b
                on the 256 x 256 x 30 basin (60 steps, checkpoints every 10);
  ladder        the reference's own published workload, lock_exchange_v06
                (TKE + up3_tvd + TEOS-10 over rough topography, tier-2 case of
                tools/tier2_sweep.sh:85) at the grids of its hardware ladder,
                timed as the reference times it (50 steps from the initial
                state, median of repeats) beside the H100 CUDA ms/step the
                reference publishes for the same grid and scheme
                (docs/32_hardware_ladder.md); context, not vs_baseline;
  cpu_baseline  the NumPy oracle of the same step on one host core, on the
                benchmark grid itself (512 x 512 x 40, same basin and initial
                state), extrapolated from --cpu-baseline-steps steps (default 1,
                about a minute).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-steps/s FP64 plain vs mixed at stated L2 error; adjoint:forward ratio"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) HBM bytes of one launch per layer cell of each
# kernel: every array it must read or write touched once, fp64 = 8 B, masks
# 1 B; interface arrays counted per layer cell (the (nz+1)/nz excess is
# ignored). Stage figures follow the stage model of the step:
#   k_tracers      stage E: T, S, u, v, K_h, dz3 read; T, S, b written  (72 B)
#   k_closure      e, K_m, K_h, u, v, b, dz3 read; e, K_m, K_h written    (80 B)
#   k_predictor    stage B: u, v, b, K_m, dz3 read; gh_u, gh_v written   (56 B)
#   k_momentum_explicit  u, v, b, w, K_m read; 6 explicit terms written  (88 B)
ALG_BYTES_PER_CELL = {"k_tracers": 72.0, "k_closure": 80.0, "k_predictor": 56.0,
                      "k_momentum_explicit": 88.0}
# Measured HBM bytes per launch at 512 x 512 x 40 (rocprofv3 FETCH_SIZE and
# WRITE_SIZE, scaled by a 1 GiB copy of known size: fetch x 2.0, write x 1.0;
# the summary is profiles/basin512x512x40_hbm_traffic_pmc.md). Re-measured
# whenever the kernel changes.
MEASURED_TRAFFIC_512x40 = {"k_tracers": 1.305e9, "k_momentum_explicit": 1.121e9,
                           "k_predictor": 1.828e9, "k_closure": 1.195e9}
STEP_BYTES_PER_CELL = 336.0
SOLVE_BYTES_PER_COLUMN = 64.0
ITER_BYTES_PER_COLUMN = {"pcg_mg": 520.0, "pcg_jacobi": 120.0, "pcg_rbgs": 520.0,
                         "multigrid": 400.0, "rbgs": 48.0}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--dx", type=float, default=1.0e4)
    ap.add_argument("--dt", type=float, default=600.0)
    ap.add_argument("--kernel", default="k_tracers",
                    help="kernel timed with HIP events for the roofline figure (the tracer "
                         "kernel: one whole stage of the step's byte model)")
    ap.add_argument("--kernel2", default="k_predictor",
                    help="second kernel timed over 10 extra steps for roofline_<name> "
                         "('none' to skip)")
    ap.add_argument("--solver-path", choices=("multikernel", "persistent"), default="multikernel",
                    help="Helmholtz launch strategy (bitwise the same iterates)")
    ap.add_argument("--solver", default=None,
                    choices=("pcg_jacobi", "pcg_rbgs", "rbgs", "multigrid", "pcg_mg"),
                    help="Helmholtz solver kind (default: the workload's)")
    ap.add_argument("--mg-levels", type=int, default=None, help="pcg_mg levels (default: the workload's)")
    ap.add_argument("--mg-sweeps", type=int, default=None, help="pcg_mg sweeps per side")
    ap.add_argument("--mg-smoother", choices=("jacobi", "rbgs"), default="rbgs",
                    help="pcg_mg V-cycle smoother of the headline step (rbgs: red-black "
                         "Gauss-Seidel, 16 PCG iterations per step at 512 x 512 x 40 against "
                         "20 with damped Jacobi)")
    ap.add_argument("--mg-tile", type=int, default=None,
                    help="fused V-cycle leg tile shape 0..3 (32x16, 64x8, 16x16, 64x16); "
                         "default: the library's")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="step precision timed for the headline value (fp64: the plain step)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-baseline-steps", type=int, default=1,
                    help="oracle steps timed on the benchmark grid (about a minute each at "
                         "512 x 512 x 40 on one host core)")
    ap.add_argument("--ladder-only", action="store_true",
                    help="time only the ladder workloads and print {'ladder': ...} (profiling)")
    ap.add_argument("--ladder", default=",".join(LADDER_DEFAULT),
                    help="reference hardware-ladder workloads to time: comma list of "
                         f"{', '.join(LADDER)}, 'all' or 'none'")
    ap.add_argument("--ensemble-steps", type=int, default=100,
                    help="steps of the 64-member ensemble throughput run (0: skip it)")
    ap.add_argument("--adjoint-steps", type=int, default=60,
                    help="steps of the adjoint:forward measurement on the 256 x 256 x 30 "
                         "basin (0: skip it)")
    ap.add_argument("--adjoint-fd-directions", type=int, default=16,
                    help="directions of the central-difference gradient check after the "
                         "adjoint timing (0: skip the check)")
    ap.add_argument("--adjoint-mg-smoother", choices=("jacobi", "rbgs"), default="jacobi",
                    help="pcg_mg V-cycle smoother of the adjoint measurement's forward and "
                         "reverse solves")
    ap.add_argument("--mixed-steps", type=int, default=50,
                    help="steps of the plain-vs-mixed comparison on the 1024 x 1024 x 60 "
                         "basin (0: skip it)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the workload, copy this many MiB device to device once "
                         "(a kernel of known traffic, to scale HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(nx: int, nz: int, dx: float, dt: float, steps: int,
                 mg_smoother: str = "rbgs") -> dict:
    """The oracle on one core on the benchmark's own grid: the same basin
    (same generator, seed and initial state) and the same step, `steps` steps
    from the initial state after set-up. NumPy's element-wise kernels are
    single-threaded and nothing on the path calls BLAS, so one core does
    the work; the rate is cells * steps / wall, labelled as taken from that
    many steps."""
    import numpy as np

    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    from oracle.stepper import OracleStepper, OState
    p = basin(nx=nx, nz=nz, dx=dx, dt=dt)
    p.params.mg_smoother = mg_smoother
    st = OracleStepper(p.domain, p.physics, p.params, p.dt)
    g = p.domain.grid
    b = st.buoyancy_of(p.initial_T, p.initial_S) * p.domain.mask3
    s = OState(u=np.zeros(g.shape3d), v=np.zeros(g.shape3d), b=b, eta=np.zeros(g.shape),
               T=p.initial_T, S=p.initial_S)
    t0 = time.perf_counter()
    for _ in range(steps):
        s = st.step(s)
    wall = time.perf_counter() - t0
    cells = g.nx * g.ny * g.nz
    return {"value": cells * steps / wall, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"NumPy oracle of the same step on the benchmark grid itself "
                      f"({g.nx}x{g.ny}x{g.nz} basin, same generator and initial state, "
                      f"{p.params.solver}, {mg_smoother} smoother): extrapolated from {steps} "
                      f"step(s) from the initial "
                      f"state, {wall:.1f} s; host {platform.processor() or 'cpu'}, "
                      f"nproc {os.cpu_count()}, numpy {np.__version__}, OMP_NUM_THREADS="
                      f"{os.environ.get('OMP_NUM_THREADS', 'unset')}, OPENBLAS_NUM_THREADS="
                      f"{os.environ.get('OPENBLAS_NUM_THREADS', 'unset')}"}


def _initial_fields(prob, st, torch):
    g = prob.domain.grid
    T = torch.as_tensor(prob.initial_T, device=st.device)
    S = torch.as_tensor(prob.initial_S, device=st.device)
    b = st.buoyancy_of(T, S) * torch.as_tensor(prob.domain.mask3, device=st.device)
    zero3 = torch.zeros(g.shape3d, dtype=torch.float64, device=st.device)
    return {"u": zero3.clone(), "v": zero3.clone(), "b": b, "T": T, "S": S,
            "eta": torch.zeros(g.shape, dtype=torch.float64, device=st.device),
            "e": torch.full((g.nz + 1,) + g.shape, float(prob.physics.e_min),
                            dtype=torch.float64, device=st.device)}


FP32_PARTS = {"vcycle": 1, "tracer_flux": 2, "eos": 4, "momentum_flux": 8}


def mixed_comparison(steps: int, nx: int = 1024, nz: int = 60, per_part: bool = True) -> dict:
    """The plain fp64 step and the mixed step on the 1024 x 1024 x 60 basin
    (dx 5 km, dt 300 s, same generator), each timed over `steps` steps after
    one untimed warm-up step from the same initial state; per-field relative
    L2 of the mixed final state against the fp64 one (the fp64 GPU step is
    the oracle's bit for bit on the parity cases). With per_part, each fp32
    part alone as well (hoc_stepper_set_fp32_parts), so the error and the
    time of the mixed step can be attributed to its parts."""
    import torch

    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    from hydrostatic_ocean_core_benchmark_amd.stepper import HipStepper
    prob = basin(nx=nx, nz=nz, dx=5.0e3, dt=300.0)
    # the fp32 V-cycle smooths with damped Jacobi, so both arms use that smoother
    prob.params.mg_smoother = "jacobi"
    g = prob.domain.grid
    cells = g.nx * g.ny * g.nz
    final, rows = {}, {}
    runs = [("fp64", "fp64", None), ("mixed", "mixed", None),
            ("mixed_fp32_momentum", "mixed_fp32_momentum", None)]
    if per_part:
        runs += [(f"only_{name}", "mixed", bit) for name, bit in FP32_PARTS.items()]
    for label, prec, parts in runs:
        st = HipStepper(prob.domain, prob.physics, prob.params, prob.dt, precision=prec)
        if parts is not None:
            st.set_fp32_parts(parts)
        bufs = [st._new_state_buffers(), st._new_state_buffers()]
        st._step_into(_initial_fields(prob, st, torch), bufs[1])  # warm-up
        st.reset_mixing()
        cur = _initial_fields(prob, st, torch)
        torch.cuda.synchronize()
        it0 = st.solver_iterations
        t0 = time.perf_counter()
        for n in range(steps):
            st._step_into(cur, bufs[n % 2])
            cur = bufs[n % 2]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        final[label] = {k: cur[k].clone() for k in ("eta", "T", "S", "u", "v", "b")}
        rows[label] = {"value": cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
                       "solver_iterations": st.solver_iterations - it0,
                       "solver_failures": st.solver_failures}
        st.close()
        del st, bufs, cur
        torch.cuda.empty_cache()

    def err_of(label):
        return {k: float(torch.linalg.vector_norm(final[label][k] - final["fp64"][k])
                         / torch.linalg.vector_norm(final["fp64"][k])) for k in final["fp64"]}

    out = {"workload": f"closed GRF basin {g.nx}x{g.ny}x{g.nz}, dx 5 km, dt 300 s, "
                       f"{steps} steps, polyTEOS10, up3_tvd, TKE, theta 0.5, pcg_mg "
                       "(5 levels, 2 + 2 damped-Jacobi sweeps, omega 0.8) rtol 1e-12",
           "unit": "cell-steps/s", "fp64": rows["fp64"], "mixed": rows["mixed"],
           "speedup": rows["mixed"]["value"] / rows["fp64"]["value"],
           "rel_l2_vs_fp64": err_of("mixed"),
           "fp32_parts": "pcg_mg V-cycle preconditioner, tracer advective fluxes, EOS "
                         "(anomaly form)",
           "fp64_parts": "momentum advective fluxes, state updates, column (Thomas) solves, "
                         "closure, outer PCG residual / dot products / update",
           "mixed_fp32_momentum": {**rows["mixed_fp32_momentum"],
                                   "speedup": rows["mixed_fp32_momentum"]["value"]
                                   / rows["fp64"]["value"],
                                   "rel_l2_vs_fp64": err_of("mixed_fp32_momentum"),
                                   "note": "as mixed, with the momentum advective fluxes in "
                                           "fp32 too"}}
    if per_part:
        out["per_part"] = {name: {"ms_per_step": round(rows[f"only_{name}"]["ms_per_step"], 3),
                                  "speedup": rows[f"only_{name}"]["value"] / rows["fp64"]["value"],
                                  "solver_iterations": rows[f"only_{name}"]["solver_iterations"],
                                  "rel_l2_vs_fp64": err_of(f"only_{name}")}
                           for name in FP32_PARTS}
    return out


def adjoint_ratio(steps: int = 60, every: int = 10, nx: int = 256, nz: int = 30,
                  fd_directions: int = 16, mg_smoother: str = "jacobi") -> dict:
    """dJ/dT0 and dJ/dkappa_b after `steps` steps of the 256 x 256 x 30 basin
    (same generator, dx 10 km, dt 600 s; kappa_b = 1e-5 (1 + 0.2 xi), xi
    seeded N(0, 1) per column), checkpoints every `every` steps: wall time of
    the gradient (forward sweep with checkpoints + reverse sweep) over the
    wall time of the plain forward of the same steps, after one warm-up.
    Then the stated gradient check: central differences with h = 1e-6 along
    `fd_directions` seeded random directions of T0, each J from the plain
    fp64 HIP forward of the same steps, against the adjoint's directional
    derivative, for the workload's up3_tvd and for the unlimited up3."""
    import numpy as np
    import torch

    from hydrostatic_ocean_core_benchmark_amd.adjoint import Adjoint, directional_check
    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    prob = basin(nx=nx, nz=nz)
    prob.params.mg_smoother = mg_smoother
    g = prob.domain.grid
    xi = np.random.default_rng(20261203).standard_normal(g.shape)
    kb = torch.as_tensor(1e-5 * (1.0 + 0.2 * xi), device="cuda")
    T0 = torch.as_tensor(prob.initial_T, device="cuda")
    S0 = torch.as_tensor(prob.initial_S, device="cuda")
    adj = Adjoint(prob.domain, prob.physics, prob.params, prob.dt)
    adj.gradient(T0, S0, kb, n_steps=2, every=every)
    torch.cuda.synchronize()
    r = adj.gradient(T0, S0, kb, n_steps=steps, every=every)
    out = {"workload": f"closed GRF basin {g.nx}x{g.ny}x{g.nz}, dx 10 km, dt 600 s, "
                       f"pcg_mg ({mg_smoother} smoother), {steps} steps, checkpoints every "
                       f"{every}, J = depth-weighted mean "
                       "SSH^2 + 0.5 volume-mean face-weighted speed^2; gradients w.r.t. T0 "
                       "and kappa_b(x, y)",
           "forward_ms": round(r.forward_s * 1e3, 2), "gradient_ms": round(r.gradient_s * 1e3, 1),
           "adjoint_forward_ratio": round(r.gradient_s / r.forward_s, 2), "J": r.J,
           "reverse": "window recompute (hoc_stepper_step_tape) on a second stream under the "
                      "reverse of the next window; sync-free solves: recompute "
                      f"{r.solve_budget} iterations, adjoint (pcg_mg with the rbgs smoother) "
                      f"{r.adjoint_budget}"
                      + (" (a solve outran it: timed again polled)" if r.redone_polled else ""),
           "hand_written": "the whole step: hoc_stepper_step_adjoint (tracers with the EOS, "
                           "advection, Picard passes with adjoint Helmholtz PCG solves, "
                           "Coriolis, pressure gradients, explicit diffusion, q and the face "
                           "coefficients) and hoc_stepper_closure_adjoint (TKE closure with "
                           "the reverse Gaspar march) on the tapes of hoc_stepper_step_tape; "
                           "no autograd"}
    if fd_directions > 0:
        checks = {}
        for adv in ("up3_tvd", "up3"):
            if adv != prob.params.advection:
                prob.params.advection = adv
                adj = Adjoint(prob.domain, prob.physics, prob.params, prob.dt)
                r = adj.gradient(T0, S0, kb, n_steps=steps, every=every)
            c = directional_check(adj, T0, S0, kb, r, n_dirs=fd_directions, h=1e-6)
            checks[adv] = {k: c[k] for k in ("rel_diff_max", "rel_diff_median",
                                             "grad_scaled_diff_max", "grad_scaled_diff_median",
                                             "rel_diff")}
        out["fd_check"] = {"directions": fd_directions, "h": 1e-6, "steps": steps,
                           "metric": "rel_diff: |fd - ad| / |ad| of the directional derivative "
                                     "along seeded N(0, 1) directions d of T0; grad_scaled_diff: "
                                     "|fd - ad| / (|dJ/dT0| |d|)", **checks}
    return out


def ensemble_throughput(steps: int = 100, mg_smoother: str = "rbgs") -> dict:
    """The 64-member ensemble (16 each at 128^2, 192^2, 256^2, 384^2, 30
    levels, per-member seeds, T noise N(0, 0.01)): leg tile shapes by the
    autotuner, solve budgets calibrated, one HIP graph per two steps;
    aggregate cell-steps/s over `steps` steps, beside the same steps run
    member after member with polled solves."""
    from hydrostatic_ocean_core_benchmark_amd.ensemble import Ensemble, benchmark_members
    t0 = time.perf_counter()
    e = Ensemble(benchmark_members(mg_smoother=mg_smoother))
    setup = time.perf_counter() - t0
    tiles = e.autotune()
    budgets = e.calibrate()
    e.capture()
    e.run(2)
    r = e.run(steps)
    eager = e.run_eager(steps)
    return {"workload": "64 members: 16 each at 128x128, 192x192, 256x256, 384x384, 30 levels, "
                        "GRF basins from per-member seeds, T noise N(0, 0.01 K), fp64, "
                        f"pcg_mg ({mg_smoother} smoother), {steps} steps, one HIP graph "
                        "replay per 2 steps",
            "unit": "cell-steps/s", "value": r.cell_steps_per_s, "cells": r.cells,
            "wall_s": round(r.wall_s, 3), "solver_failures": r.solver_failures,
            "eager_sequential_value": r.cells * steps / eager,
            "graph_chains": len(e.chains), "tile_choice": {k: v["tile"] for k, v in tiles.items()},
            "solve_budget_range": [min(budgets.values()), max(budgets.values())],
            "setup_s": round(setup, 1)}


# name: (nx, scheme, solver, cfl, published H100 CUDA ms/step, source line)
LADDER = {
    "fb_400": (400, "fb", "multigrid", 0.5, 7.59, "docs/32_hardware_ladder.md:137"),
    "fb_1000": (1000, "fb", "multigrid", 0.5, 48.91, "docs/32_hardware_ladder.md:140"),
    "fb_2000": (2000, "fb", "multigrid", 0.5, 185.50, "docs/32_hardware_ladder.md:143"),
    "theta_mg_2000": (2000, "theta", "multigrid", 2.0, 214.16, "docs/32_hardware_ladder.md:109"),
    "theta_pcgj_2000": (2000, "theta", "pcg_jacobi", 2.0, 206.77,
                        "docs/32_hardware_ladder.md:117"),
}
LADDER_DEFAULT = tuple(LADDER)


def reference_ladder(names, steps: int = 50, repeats: int = 5) -> dict:
    """The reference's hardware-ladder workloads on this GPU: per workload, a
    5-step warm-up, then `repeats` runs of `steps` steps from the initial
    state (mixing coefficients reset too), median ms/step and its MAD, beside
    the published H100 CUDA figure (the reference's protocol, RULES.md:61-98:
    at least 5 repeats, median and MAD)."""
    import numpy as np
    import torch
    from hydrostatic_ocean_core_benchmark_amd.cases import ladder_lock_exchange
    from hydrostatic_ocean_core_benchmark_amd.stepper import HipStepper
    out = {"case": "lock_exchange_v06 (tools/tier2_sweep.sh:85), 30 levels, fp64",
           "protocol": f"5-step warm-up, median of {repeats} runs of {steps} steps from the "
                       "initial state", "rows": []}
    for name in names:
        nx, scheme, solver, cfl, h100, src = LADDER[name]
        t_setup = time.perf_counter()
        prob = ladder_lock_exchange(nx, 30, scheme, solver, cfl, steps)
        st = HipStepper(prob.domain, prob.physics, prob.params, prob.dt)
        init = _initial_fields(prob, st, torch)
        bufs = [st._new_state_buffers(), st._new_state_buffers()]

        def run(n):
            cur = init
            for k in range(n):
                st._step_into(cur, bufs[k % 2])
                cur = bufs[k % 2]
            return cur

        run(5)
        st.reset_mixing()
        torch.cuda.synchronize()
        setup_s = time.perf_counter() - t_setup
        walls, it0 = [], st.solver_iterations
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fin = run(steps)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
            st.reset_mixing()
        ms = 1e3 * float(np.median(walls)) / steps
        mad = 1e3 * float(np.median(np.abs(np.asarray(walls) - np.median(walls)))) / steps
        cells = nx * nx * 30
        out["rows"].append({
            "workload": name, "grid": f"{nx}x{nx}x30", "scheme": prob.meta["scheme"],
            "cfl": cfl, "dt": prob.dt, "ms_per_step": round(ms, 3),
            "mad_ms_per_step": round(mad, 4),
            "cell_steps_per_s": cells / (ms / 1e3),
            "solver_iterations_per_step": (st.solver_iterations - it0) / (repeats * steps),
            "solver_failures": st.solver_failures,
            "state_finite": bool(torch.isfinite(fin["T"]).all()),
            "h100_cuda_ms_per_step": h100, "h100_source": src,
            "speedup_vs_h100_cuda": round(h100 / ms, 2), "setup_s": round(setup_s, 1)})
        del st, bufs, init, fin
        torch.cuda.empty_cache()
    return out


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU slab decomposition is design-only)"}))
        return 2
    import numpy as np
    import torch

    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    from hydrostatic_ocean_core_benchmark_amd.stepper import HipStepper

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 2
    if args.ladder_only:
        names = list(LADDER) if args.ladder in ("all", "none") else args.ladder.split(",")
        print(json.dumps({"ladder": reference_ladder(names)}), flush=True)
        return 0
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.arange(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    prob = basin(nx=args.nx, nz=args.nz, dx=args.dx, dt=args.dt)
    if args.solver:
        prob.params.solver = args.solver
    if args.mg_levels is not None:
        prob.params.mg_levels = args.mg_levels
    if args.mg_sweeps is not None:
        prob.params.mg_sweeps = args.mg_sweeps
    if prob.params.solver == "pcg_mg":
        prob.params.mg_smoother = args.mg_smoother
    g = prob.domain.grid
    st = HipStepper(prob.domain, prob.physics, prob.params, prob.dt, precision=args.precision)
    st.solver_path = args.solver_path
    if args.mg_tile is not None:
        from hydrostatic_ocean_core_benchmark_amd import _lib
        _lib.check(st.lib.hoc_stepper_set_mg_tile(st._h, args.mg_tile))
    cur = _initial_fields(prob, st, torch)
    bufs = [st._new_state_buffers(), st._new_state_buffers()]
    n = 0

    def one_step():
        nonlocal cur, n
        out = bufs[n % 2]
        st._step_into(cur, out)
        cur = out
        n += 1

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    it0 = st.solver_iterations
    st.profile_kernel(args.kernel)
    torch.cuda.synchronize()
    # the timed steps in one library call (Stepper3DV05.integrate's loop)
    a, b = cur, bufs[n % 2]
    t0 = time.perf_counter()
    cur = st._step_n_into(a, b, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n += args.steps
    launches, kms = st.profile_read()
    st.profile_kernel(None)
    iters = st.solver_iterations - it0
    fails = st.solver_failures
    finite = bool(torch.isfinite(cur["eta"]).all() and torch.isfinite(cur["T"]).all())

    cells = g.nx * g.ny * g.nz
    value = cells * args.steps / wall
    def roofline_of(kernel, launches, kms):
        if launches <= 0 or kernel not in ALG_BYTES_PER_CELL:
            return None
        avg_s = kms / launches / 1e3
        alg = ALG_BYTES_PER_CELL[kernel] * cells
        achieved = alg / avg_s / 1e9
        traffic = (MEASURED_TRAFFIC_512x40.get(kernel)
                   if (g.nx, g.ny, g.nz) == (512, 512, 40) else None)
        return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "alg_bytes_per_launch": alg,
                "avg_launch_ms": round(avg_s * 1e3, 4), "launches": launches}

    roofline = roofline_of(args.kernel, launches, kms)
    roofline2 = None
    if args.kernel2 != "none":
        st.profile_kernel(args.kernel2)
        for _ in range(10):
            one_step()
        torch.cuda.synchronize()
        roofline2 = roofline_of(args.kernel2, *st.profile_read())
        st.profile_kernel(None)
    n_solves = prob.params.n_picard if prob.params.name == "theta" else 0
    # the Helmholtz solves' share of the step: HIP events around each solve
    # (its span on the device timeline, host polls included) over 10 steps
    solver_share = None
    if n_solves:
        st.profile_kernel("helmholtz_solve")
        i1 = st.solver_iterations
        for _ in range(10):
            one_step()
        torch.cuda.synchronize()
        n_sv, sv_ms = st.profile_read()
        i2 = st.solver_iterations
        st.profile_kernel(None)
        if n_sv > 0:
            per_step = sv_ms / 10
            solver_share = {"ms_per_step": round(per_step, 4),
                            "share_of_step": round(per_step / (1e3 * wall / args.steps), 3),
                            "solves_per_step": n_sv / 10, "iterations_per_step": (i2 - i1) / 10}
    solve_bytes = g.nx * g.ny * (SOLVE_BYTES_PER_COLUMN * n_solves
                                 + ITER_BYTES_PER_COLUMN[prob.params.solver] * iters / args.steps)
    step_bytes = STEP_BYTES_PER_CELL * cells + solve_bytes
    step_gbs = step_bytes / (wall / args.steps) / 1e9
    step_model = {"bytes_per_step": step_bytes, "achieved": round(step_gbs, 1),
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4)}
    rec = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": ("fp64 plain" if args.precision == "fp64" else "mixed")
                               + f" step, closed GRF basin {g.nx}x{g.ny}x{g.nz}, "
                               f"dx {args.dx / 1e3:g} km, dt {args.dt:g} s, up3_tvd, TKE, "
                               "polyTEOS10, theta 0.5, "
                               f"{prob.params.solver}"
                               + (f" ({prob.params.mg_smoother} smoother)"
                                  if prob.params.solver == "pcg_mg" else "")
                               + f" rtol {prob.params.rtol:g}",
                   "cells": cells, "wet_fraction": round(float(np.mean(prob.domain.mask3)), 4),
                   "solver": prob.params.solver, "mg_smoother": prob.params.mg_smoother,
                   "solver_path": st.solver_path,
                   "solver_iterations_per_step": iters / args.steps,
                   "solver_failures": fails, "state_finite": finite},
        "roofline": roofline,
        "step_model": step_model,
        "helmholtz_solves": solver_share,
    }
    if roofline2 is not None:
        rec[f"roofline_{args.kernel2.removeprefix('k_')}"] = roofline2
    if args.mixed_steps > 0:
        del st, bufs, cur
        torch.cuda.empty_cache()
        rec["mixed"] = mixed_comparison(args.mixed_steps)
    if args.ensemble_steps > 0:
        rec["ensemble"] = ensemble_throughput(args.ensemble_steps)
    if args.adjoint_steps > 0:
        rec["adjoint"] = adjoint_ratio(args.adjoint_steps,
                                       fd_directions=args.adjoint_fd_directions,
                                       mg_smoother=args.adjoint_mg_smoother)
    if args.ladder != "none":
        names = list(LADDER) if args.ladder == "all" else args.ladder.split(",")
        rec["ladder"] = reference_ladder(names)
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(args.nx, args.nz, args.dx, args.dt,
                                           args.cpu_baseline_steps, prob.params.mg_smoother)
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
