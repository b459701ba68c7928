"""Batched ensembles of the step: many independent members, of uneven grid
sizes, advanced together by replaying one HIP graph per two steps.

Each member is a HipStepper of its own (own geometry, workspaces and
multigrid hierarchy) with two state buffer sets. For capture the members'
Helmholtz solves run sync-free (hoc_stepper_set_sync_free): a fixed number
of PCG iterations per solve, calibrated per member from polled steps with a
margin, so that the step is one host-independent launch sequence; phases
after convergence return at once, so the results are bitwise those of the
polled step, and a solve that would have needed more than its budget is
counted as a failure and reported. The graph holds, for every member, the
step a -> b followed by the step b -> a, the members dealt to a few
independent stream chains (forked from and joined to the capture stream)
that the device runs concurrently; one replay advances the whole ensemble by
two steps.

The autotuner picks, per grid size, the tile shape of the multigrid V-cycle
legs (hoc_stepper_set_mg_tile: 32x16, 64x8, 16x16 or 64x16 cells) by timing
polled steps of one member of that size; every shape gives the same bits.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .stepper import HipStepper

TILE_SHAPES = ("32x16", "64x8", "16x16", "64x16")


@dataclass(eq=False)
class Member:
    problem: object
    stepper: HipStepper
    initial: dict
    bufs: list
    budget: int = 0
    cells: int = 0


@dataclass
class EnsembleReport:
    n_members: int
    n_steps: int
    cells: int
    wall_s: float
    cell_steps_per_s: float
    tile_choice: dict = field(default_factory=dict)
    budgets: dict = field(default_factory=dict)
    solver_failures: int = 0


def _initial_fields(prob, st) -> dict:
    g = prob.domain.grid
    dev = st.device
    T = torch.as_tensor(prob.initial_T, device=dev)
    S = torch.as_tensor(prob.initial_S, device=dev)
    b = st.buoyancy_of(T, S) * torch.as_tensor(prob.domain.mask3, device=dev)
    z3 = torch.zeros(g.shape3d, dtype=torch.float64, device=dev)
    out = {"u": z3, "v": z3.clone(), "b": b, "T": T, "S": S,
           "eta": torch.zeros(g.shape, dtype=torch.float64, device=dev)}
    if prob.physics.closure == "tke":
        out["e"] = torch.full((g.nz + 1,) + g.shape, float(prob.physics.e_min),
                              dtype=torch.float64, device=dev)
    return out


class Ensemble:
    """Members from a list of problems (cases.Problem); theta scheme with
    pcg_mg or pcg_rbgs."""

    def __init__(self, problems, device="cuda:0"):
        self.device = torch.device(device)
        self.members: list[Member] = []
        for prob in problems:
            st = HipStepper(prob.domain, prob.physics, prob.params, prob.dt, device=self.device)
            g = prob.domain.grid
            self.members.append(Member(problem=prob, stepper=st,
                                       initial=_initial_fields(prob, st),
                                       bufs=[st._new_state_buffers(), st._new_state_buffers()],
                                       cells=g.nx * g.ny * g.nz))
        self.graph = None
        self.tile_choice: dict = {}

    # ----------------------------------------------------------- helpers
    def _restart(self, m: Member) -> None:
        """Back to the initial state (mixing coefficients at their
        backgrounds), in the member's buffer a."""
        m.stepper.reset_mixing()
        for k, v in m.initial.items():
            m.bufs[0][k].copy_(v)

    def _step(self, m: Member, src: int) -> None:
        m.stepper._step_into(m.bufs[src], m.bufs[1 - src])

    @staticmethod
    def _size(m: Member) -> tuple:
        g = m.problem.domain.grid
        return (g.nx, g.ny, g.nz)

    # ---------------------------------------------------------- autotune
    def autotune(self, steps: int = 3, repeats: int = 2) -> dict:
        """Per grid size, the V-cycle leg tile shape with the least time per
        polled step (one member of that size), applied to every member of
        that size."""
        by_size: dict = {}
        for m in self.members:
            by_size.setdefault(self._size(m), []).append(m)
        choice = {}
        for size, group in by_size.items():
            m = group[0]
            times = []
            for shape in range(len(TILE_SHAPES)):
                _lib.check(m.stepper.lib.hoc_stepper_set_mg_tile(m.stepper._h, shape))
                best = math.inf
                for _ in range(repeats):
                    self._restart(m)
                    torch.cuda.synchronize(self.device)
                    t0 = time.perf_counter()
                    for n in range(steps):
                        self._step(m, n % 2)
                    torch.cuda.synchronize(self.device)
                    best = min(best, (time.perf_counter() - t0) / steps)
                times.append(best)
            pick = int(np.argmin(times))
            for mm in group:
                _lib.check(mm.stepper.lib.hoc_stepper_set_mg_tile(mm.stepper._h, pick))
            choice["x".join(map(str, size))] = {
                "tile": TILE_SHAPES[pick],
                "ms_per_step": {TILE_SHAPES[q]: round(1e3 * t, 3) for q, t in enumerate(times)}}
        self.tile_choice = choice
        return choice

    # --------------------------------------------------------- calibrate
    def calibrate(self, steps: int = 2, margin: float = 1.0, extra: int = 2) -> dict:
        """Iterations per solve from `steps` polled steps of every member;
        budget = ceil(margin * max) + extra. Measured on the benchmark
        ensemble (rbgs smoother), 100 polled steps per member: 16-18
        iterations per step (two solves) throughout, against 18 in the first
        two steps, so two iterations of headroom cover every solve; an
        unconverged solve is counted as a failure, never passed silently."""
        budgets = {}
        for i, m in enumerate(self.members):
            self._restart(m)
            st = m.stepper
            it0, n_solves = st.solver_iterations, 0
            worst = 0
            for n in range(steps):
                before = st.solver_iterations
                self._step(m, n % 2)
                per = st.solver_iterations - before
                n_solves = st.params.n_picard
                worst = max(worst, math.ceil(per / n_solves))
            del it0
            m.budget = int(math.ceil(margin * worst)) + extra
            budgets[i] = m.budget
        return budgets

    # ----------------------------------------------------------- capture
    def capture(self, branches: int = 0) -> None:
        """One HIP graph advancing every member by two steps (a -> b -> a),
        with sync-free solves. The members are dealt, largest first, to
        `branches` streams (0: one per member) forked from the capture
        stream and joined at the end, so the graph holds that many
        independent chains the device may run concurrently (measured on the
        64-member benchmark ensemble, 40 steps: 1 chain 7.2e8 cell-steps/s,
        4 chains 7.8e8, 8 chains 9.3e8, 16 chains 1.40e9)."""
        for m in self.members:
            if m.budget <= 0:
                raise RuntimeError("calibrate() before capture()")
            _lib.check(m.stepper.lib.hoc_stepper_set_sync_free(m.stepper._h, m.budget))
            self._restart(m)
        order = sorted(range(len(self.members)), key=lambda i: -self.members[i].cells)
        nb = len(order) if branches <= 0 else max(1, min(branches, len(order)))
        self.chains = [[] for _ in range(nb)]
        load = [0] * nb
        for i in order:  # greedy: the least-loaded chain takes the next largest member
            q = int(np.argmin(load))
            self.chains[q].append(i)
            load[q] += self.members[i].cells
        chains = [[self.members[i] for i in c] for c in self.chains]
        streams = [torch.cuda.Stream(device=self.device) for _ in range(nb)]
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cap = torch.cuda.current_stream(self.device)
            for st in streams:
                st.wait_stream(cap)
            for st, chain in zip(streams, chains):
                with torch.cuda.stream(st):
                    for m in chain:
                        self._step(m, 0)
                        self._step(m, 1)
            for st in streams:
                cap.wait_stream(st)
        torch.cuda.synchronize(self.device)
        self.graph = g

    def restart(self) -> None:
        for m in self.members:
            self._restart(m)
        torch.cuda.synchronize(self.device)

    def replay(self, n_pairs: int) -> None:
        for _ in range(n_pairs):
            self.graph.replay()

    def failures(self) -> int:
        return sum(m.stepper.solver_failures for m in self.members)

    def state(self, i: int) -> dict:
        """Member i's current state (buffer a: after whole replays)."""
        return self.members[i].bufs[0]

    # --------------------------------------------------------------- run
    def run(self, n_steps: int) -> EnsembleReport:
        """n_steps (even) of every member from the initial states by graph
        replays, timed between device synchronisations."""
        if self.graph is None:
            raise RuntimeError("capture() before run()")
        if n_steps % 2:
            raise ValueError("the graph advances two steps per replay: n_steps must be even")
        f0 = self.failures()
        self.restart()
        t0 = time.perf_counter()
        self.replay(n_steps // 2)
        torch.cuda.synchronize(self.device)
        wall = time.perf_counter() - t0
        cells = sum(m.cells for m in self.members)
        fails = self.failures() - f0
        return EnsembleReport(n_members=len(self.members), n_steps=n_steps, cells=cells,
                              wall_s=wall, cell_steps_per_s=cells * n_steps / wall,
                              tile_choice=self.tile_choice,
                              budgets={i: m.budget for i, m in enumerate(self.members)},
                              solver_failures=fails)

    def run_eager(self, n_steps: int) -> float:
        """The same steps member after member with polled solves (the
        sequential baseline); returns the wall time."""
        for m in self.members:
            _lib.check(m.stepper.lib.hoc_stepper_set_sync_free(m.stepper._h, 0))
        self.restart()
        t0 = time.perf_counter()
        for m in self.members:
            for n in range(n_steps):
                self._step(m, n % 2)
        torch.cuda.synchronize(self.device)
        wall = time.perf_counter() - t0
        for m in self.members:
            _lib.check(m.stepper.lib.hoc_stepper_set_sync_free(m.stepper._h, m.budget))
        return wall


def benchmark_members(sizes=(128, 192, 256, 384), per_size: int = 16, nz: int = 30,
                      base_seed: int = 20261204, sigma_T: float = 0.01,
                      mg_smoother: str = "rbgs"):
    """The ensemble of the benchmark: `per_size` members at each size, 30
    levels, bathymetry and stratification generated like the 512 x 512
    basin (dx 10 km, dt 600 s) from per-member seeds, T perturbed by
    N(0, sigma_T); pcg_mg with the given V-cycle smoother."""
    from .cases import basin
    out = []
    idx = 0
    for n in sizes:
        for _ in range(per_size):
            prob = basin(nx=n, nz=nz, seed=base_seed + idx, sigma_T=sigma_T)
            prob.params.mg_smoother = mg_smoother
            out.append(prob)
            idx += 1
    return out
