"""The v0.5 verification cases and the one-case driver, run on the GPU step.

Each case builds an initial state on the host, the HipStepper integrates it on
the device, and the case's diagnostic is evaluated on the host from the final
state. Most cases have no closed-form solution; their diagnostic is zero for
the exact equations, so the measured number is the error.

    SeamountRest       libs/core/cases3d_v05.py:32-69    max|u| over a resting seamount [cm/s]
    ZStarConstancy     libs/core/cases3d_v05.py:72-146   spread of a uniform tracer
    FlatherRadiation   libs/core/cases3d_v05.py:149-180  surface energy left after radiation
    LockExchange       libs/core/cases3d_v05.py:183-264  rise of reference potential energy
    BasinSeiche        libs/core/cases3d_v05.py:267-321  energy drift of a free seiche
    BaroclinicIGWTopo  libs/core/cases3d_v05.py:324-375  an internal wave launched over topography
    build_case3d_v05   libs/core/cases3d_v05.py:378-400
    build_domain3d     libs/core/driver3d_v05.py:25-54
    barotropic_dt_max  libs/core/driver3d_v05.py:97-100
    simulate3d_v05     libs/core/driver3d_v05.py:103-146
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .domain import CGrid, Domain, bathymetry_from_config, build_bathymetry
from .types import PhysicsV05, State3D, scheme_params_from_config


@dataclass
class CaseResult:
    """`metric` is the number under test; `extra` carries the side diagnostics."""

    name: str
    metric: float
    metric_name: str
    extra: dict


class SeamountRest:
    """A depth-only stratification over a seamount is an exact rest state:
    every velocity the model develops is pressure-gradient error
    (Beckmann and Haidvogel 1993). Metric: max |u| in cm/s."""

    name = "seamount_rest"
    metric_name = "max_abs_u_cm_s"

    def __init__(self, domain: Domain, physics: PhysicsV05, delta_rho: float,
                 n_days: float) -> None:
        self.domain, self.physics = domain, physics
        self.delta_rho = float(delta_rho)
        self.t_final = float(n_days) * 86400.0

    def initial(self) -> State3D:
        d, p = self.domain, self.physics
        depth = d.depth_at_centres()
        anomaly = self.delta_rho * np.exp(-depth / 500.0)
        b = -p.g * anomaly / p.rho0 * d.mask3
        rest = np.zeros(d.grid.shape3d)
        return State3D(u=rest, v=rest, b=b, eta=np.zeros(d.grid.shape), t=0.0)

    def evaluate(self, s: State3D) -> CaseResult:
        speed = np.sqrt(s.u ** 2 + s.v ** 2)
        return CaseResult(self.name, float(np.max(speed) * 100.0), self.metric_name,
                          {"rms_speed_cm_s": float(np.sqrt(np.mean(speed ** 2)) * 100.0),
                           "rx0": self.domain.roughness.rx0, "vcoord": self.domain.vcoord})


class ZStarConstancy:
    """A uniform tracer under a moving free surface must stay uniform.
    Metric: max(T) - min(T) over wet cells. Conservation is reported on the
    resting thickness and on the live thickness (which holds the surface
    volume in the top cell for z-level)."""

    name = "zstar_constancy"
    metric_name = "tracer_spread"

    def __init__(self, domain: Domain, physics: PhysicsV05, eta0: float,
                 n_periods: float) -> None:
        self.domain, self.physics = domain, physics
        self.eta0 = float(eta0)
        speed = np.sqrt(physics.g * float(np.max(domain.H)))
        self.t_final = float(n_periods) * domain.grid.Lx / speed
        self._inv0 = self._inv0_live = None

    def _live_thickness(self, eta: np.ndarray) -> np.ndarray:
        d = self.domain
        if d.vcoord != "zlevel":
            return d.layer_thickness(eta)
        dz = d.dz3_ref.copy()
        dz[0] = dz[0] + eta * (d.mask3[0] > 0)
        return dz

    def _inventories(self, T: np.ndarray, eta: np.ndarray) -> tuple[float, float]:
        d = self.domain
        wet = d.mask3 > 0
        area = d.grid.dx * d.grid.dy
        resting = float(np.sum(T[wet] * (d.dz3_ref * area)[wet]))
        live = float(np.sum(T[wet] * (self._live_thickness(eta) * area)[wet]))
        return resting, live

    def initial(self) -> State3D:
        d, p = self.domain, self.physics
        x, _ = d.grid.coords("eta")
        eta = self.eta0 * np.cos(2.0 * np.pi * x / d.grid.Lx) * d.mask
        T = np.full(d.grid.shape3d, p.T0) * d.mask3
        S = np.full(d.grid.shape3d, p.S0) * d.mask3
        self._inv0, self._inv0_live = self._inventories(T, eta)
        rest = np.zeros(d.grid.shape3d)
        return State3D(u=rest, v=rest, b=np.zeros(d.grid.shape3d), eta=eta, t=0.0, T=T, S=S)

    def evaluate(self, s: State3D) -> CaseResult:
        d = self.domain
        T = s.T[d.mask3 > 0]
        inv, inv_live = self._inventories(s.T, s.eta)
        drift = 0.0 if not self._inv0 else (inv - self._inv0) / abs(self._inv0)
        drift_live = (0.0 if not self._inv0_live
                      else (inv_live - self._inv0_live) / abs(self._inv0_live))
        return CaseResult(self.name, float(np.max(T) - np.min(T)), self.metric_name,
                          {"T_mean": float(np.mean(T)), "tracer_integral_drift": drift,
                           "tracer_integral_drift_live": drift_live, "vcoord": d.vcoord})


class FlatherRadiation:
    """A Gaussian surface bump in a channel open in x: what surface energy is
    left after the two waves had time to leave is reflection.
    Metric: sum(eta^2) at the end over sum(eta^2) at the start."""

    name = "flather_radiate"
    metric_name = "reflected_energy_fraction"

    def __init__(self, domain: Domain, physics: PhysicsV05, eta0: float, width: float,
                 n_transits: float) -> None:
        self.domain, self.physics = domain, physics
        self.eta0, self.width = float(eta0), float(width)
        speed = np.sqrt(physics.g * float(np.mean(domain.H)))
        self.t_final = float(n_transits) * domain.grid.Lx / speed
        self._e0 = None

    def initial(self) -> State3D:
        d = self.domain
        x, _ = d.grid.coords("eta")
        bump = self.eta0 * np.exp(-((x - 0.5 * d.grid.Lx) / self.width) ** 2)
        # the reference normalises by the unmasked bump
        self._e0 = float(np.sum(bump ** 2))
        rest = np.zeros(d.grid.shape3d)
        return State3D(u=rest, v=rest, b=rest, eta=bump * d.mask, t=0.0)

    def evaluate(self, s: State3D) -> CaseResult:
        return CaseResult(self.name, float(np.sum(s.eta ** 2)) / self._e0, self.metric_name,
                          {"eta_max_final": float(np.max(np.abs(s.eta))),
                           "eta_max_initial": self.eta0})


class LockExchange:
    """A gravity current released over topography. Metric: the relative rise
    of the reference potential energy, i.e. the model's spurious mixing. The
    reference state stacks the water densest-first into the basin's own
    volume-below-height curve."""

    name = "lock_exchange"
    metric_name = "rpe_drift_rel"

    def __init__(self, domain: Domain, physics: PhysicsV05, delta_T: float,
                 n_hours: float) -> None:
        self.domain, self.physics = domain, physics
        self.delta_T = float(delta_T)
        self.t_final = float(n_hours) * 3600.0
        self._hyps = None
        self._rpe0 = None
        self.device = "cuda:0"      # where the initial buoyancy is evaluated

    def _hypsometry(self) -> tuple[np.ndarray, np.ndarray]:
        """(volume below height z, z) over the distinct cell-edge heights,
        measured from the deepest point. The volume below z is
        sum_cells clip(z - z_bottom, 0, dz), evaluated for every z at once
        with sorted prefix sums."""
        if self._hyps is None:
            d = self.domain
            dz = d.dz3_ref
            wet = d.mask3 > 0
            edges = np.concatenate([np.zeros((1,) + dz.shape[1:]), np.cumsum(dz, axis=0)])
            deepest = float(np.max(edges[-1]))
            z_top, z_bot = (deepest - edges[:-1])[wet], (deepest - edges[1:])[wet]
            heights = np.unique(np.concatenate([z_bot, z_top, [0.0, deepest]]))

            def ramp(start: np.ndarray) -> np.ndarray:
                # sum over cells of max(z - start, 0) for every z in heights
                srt = np.sort(start)
                prefix = np.concatenate([[0.0], np.cumsum(srt)])
                k = np.searchsorted(srt, heights, side="left")
                return k * heights - prefix[k]

            area = d.grid.dx * d.grid.dy
            self._hyps = (area * (ramp(z_bot) - ramp(z_top)), heights)
        return self._hyps

    def _rpe(self, b: np.ndarray, thickness: np.ndarray | None = None) -> float:
        """Sort the buoyancy densest-first, stack the cells' volumes (from
        `thickness`, default the resting thickness) into the basin's
        hypsometric curve and integrate b z dV."""
        d = self.domain
        wet = d.mask3 > 0
        dz = d.dz3_ref if thickness is None else thickness
        vol = (d.grid.dx * d.grid.dy * dz)[wet]
        bw = b[wet]
        order = np.argsort(bw)
        v = vol[order]
        vol_below, heights = self._hypsometry()
        z = np.interp(np.cumsum(v), vol_below, heights)
        return float(np.sum(bw[order] * z * v))

    def initial(self) -> State3D:
        import torch

        from .stepper import eos_buoyancy
        d, p = self.domain, self.physics
        x, _ = d.grid.coords("eta")
        warm = (x > 0.5 * d.grid.Lx).astype(np.float64)
        T = (p.T0 + self.delta_T * warm)[None, :, :] * np.ones(d.grid.shape3d)
        S = np.full(d.grid.shape3d, p.S0)
        dev = dict(dtype=torch.float64, device=self.device)
        b = eos_buoyancy(p.eos, torch.as_tensor(T, **dev), torch.as_tensor(S, **dev),
                         torch.as_tensor(d.depth_at_centres(), **dev), g=p.g, rho0=p.rho0,
                         alpha_T=p.alpha_T, beta_S=p.beta_S, T0=p.T0, S0=p.S0)
        b = b.cpu().numpy() * d.mask3
        self._rpe0 = self._rpe(b)
        rest = np.zeros(d.grid.shape3d)
        return State3D(u=rest, v=rest, b=b, eta=np.zeros(d.grid.shape), t=0.0,
                       T=T * d.mask3, S=S * d.mask3)

    def evaluate(self, s: State3D) -> CaseResult:
        rpe = self._rpe(s.b)
        drift = (rpe - self._rpe0) / abs(self._rpe0) if self._rpe0 else 0.0
        return CaseResult(self.name, float(drift), self.metric_name,
                          {"rpe_initial": self._rpe0, "rpe_final": rpe})


class BasinSeiche:
    """A free barotropic seiche in a closed basin over topography.
    Metric: relative drift of the linear free-surface energy, with the
    kinetic part weighted by the face thickness (u, v live on faces)."""

    name = "basin_seiche"
    metric_name = "energy_drift_rel"

    def __init__(self, domain: Domain, physics: PhysicsV05, eta0: float,
                 n_periods: float) -> None:
        self.domain, self.physics = domain, physics
        self.eta0 = float(eta0)
        h_mean = float(np.mean(domain.H[domain.mask > 0]))
        self.period = 2.0 * domain.grid.Lx / np.sqrt(physics.g * h_mean)
        self.t_final = float(n_periods) * self.period
        self._e0 = self._m0 = None

    def _energy(self, s: State3D) -> float:
        d = self.domain
        dzu, dzv = d.face_thickness(d.layer_thickness(s.eta))
        kinetic = 0.5 * float(np.sum(dzu * s.u ** 2 + dzv * s.v ** 2))
        potential = 0.5 * self.physics.g * float(np.sum(d.mask * s.eta ** 2))
        return (kinetic + potential) * d.grid.dx * d.grid.dy

    def initial(self) -> State3D:
        d = self.domain
        x, _ = d.grid.coords("eta")
        eta = self.eta0 * np.cos(np.pi * x / d.grid.Lx) * d.mask
        eta -= np.sum(eta) / max(1.0, float(np.sum(d.mask)))
        eta *= d.mask
        rest = np.zeros(d.grid.shape3d)
        s0 = State3D(u=rest, v=rest, b=rest, eta=eta, t=0.0)
        self._e0 = self._energy(s0)
        self._m0 = float(np.sum(eta))
        return s0

    def evaluate(self, s: State3D) -> CaseResult:
        e = self._energy(s)
        return CaseResult(self.name, (e - self._e0) / abs(self._e0), self.metric_name,
                          {"eta_max": float(np.max(np.abs(s.eta))),
                           "eta_mean_drift": float(np.sum(s.eta)) - self._m0,
                           "period_s": self.period})


class BaroclinicIGWTopo:
    """The flat-bottom internal-wave mode (vertical mode mode_z) launched
    on the nominal z-levels and masked to the wet cells. Over topography it
    is no eigenmode; the metric is the rms of eta over wet columns."""

    name = "baroclinic_igw"
    metric_name = "eta_rms"

    def __init__(self, domain: Domain, physics: PhysicsV05, mode_x: int, mode_y: int,
                 mode_z: int, eta0: float, n_periods: float) -> None:
        if physics.N2 <= 0.0:
            raise ValueError("baroclinic_igw requires physics3d.N2 > 0")
        self.domain, self.physics, self.eta0 = domain, physics, float(eta0)
        g = domain.grid
        self.k = 2.0 * np.pi * mode_x / g.Lx
        self.l = 2.0 * np.pi * mode_y / g.Ly
        self.kappa2 = self.k ** 2 + self.l ** 2
        self.m = np.pi * mode_z / physics.H
        self.c2 = physics.N2 / self.m ** 2
        self.omega = float(np.sqrt(physics.f ** 2 + self.c2 * self.kappa2))
        self.period = 2.0 * np.pi / self.omega
        self.t_final = float(n_periods) * self.period

    def initial(self) -> State3D:
        d, grav, f = self.domain, self.physics.g, self.physics.f
        amp = grav * self.eta0 / (self.c2 * self.kappa2)
        zc = d.grid.z_centre()[:, None, None]
        cz, sz = np.cos(self.m * zc), np.sin(self.m * zc)

        def phase(where: str) -> np.ndarray:
            x, y = d.grid.coords(where)
            return self.k * x + self.l * y

        pu, pv, pe = phase("u"), phase("v"), phase("eta")
        u = amp * (self.omega * self.k * np.cos(pu) - f * self.l * np.sin(pu)) * cz * d.mask3u
        v = amp * (self.omega * self.l * np.cos(pv) + f * self.k * np.sin(pv)) * cz * d.mask3v
        b = -grav * self.eta0 * self.m * np.cos(pe) * sz * d.mask3
        return State3D(u=u, v=v, b=b, eta=np.zeros(d.grid.shape), t=0.0)

    def evaluate(self, s: State3D) -> CaseResult:
        d = self.domain
        wet = max(1.0, float(np.sum(d.mask)))
        rms = float(np.sqrt(np.sum((s.eta * d.mask) ** 2) / wet))
        return CaseResult(self.name, rms, self.metric_name,
                          {"eta_max": float(np.max(np.abs(s.eta))), "period_s": self.period,
                           "u_max": float(np.max(np.abs(s.u)))})


CASES_V05 = ("seamount_rest", "zstar_constancy", "flather_radiate", "lock_exchange",
             "basin_seiche", "baroclinic_igw")


def build_case3d_v05(name: str, domain: Domain, physics: PhysicsV05, config):
    """A v0.5 case from its `[case.<name>]` config section."""
    if name not in CASES_V05:
        raise ValueError(f"unknown v0.5 case '{name}' (implemented: {'|'.join(CASES_V05)})")
    s = config.section(f"case.{name}")
    if name == "seamount_rest":
        return SeamountRest(domain, physics, float(s["delta_rho"]), float(s["n_days"]))
    if name == "zstar_constancy":
        return ZStarConstancy(domain, physics, float(s["eta0"]), float(s["n_periods"]))
    if name == "flather_radiate":
        return FlatherRadiation(domain, physics, float(s["eta0"]), float(s["width"]),
                                float(s["n_transits"]))
    if name == "lock_exchange":
        return LockExchange(domain, physics, float(s["delta_T"]), float(s["n_hours"]))
    if name == "basin_seiche":
        return BasinSeiche(domain, physics, float(s["eta0"]), float(s["n_periods"]))
    if name == "baroclinic_igw":
        return BaroclinicIGWTopo(domain, physics, int(s["mode_x"]), int(s["mode_y"]),
                                 int(s["mode_z"]), float(s["eta0"]), float(s["n_periods"]))
    raise AssertionError(name)


def build_domain3d(config, nx: int, nz: int, *, vcoord: str | None = None,
                   bathymetry: str | None = None, bc: str | None = None,
                   **bath_kw) -> Domain:
    """A square nx x nx x nz domain from config, with the vertical
    coordinate, the topography (and its keywords) and both boundaries
    overridable per run."""
    grid = CGrid(nx=nx, ny=nx, Lx=float(config.get("grid.Lx")),
                 Ly=float(config.get("grid.Ly")), nz=nz,
                 depth=float(config.get("physics.H")))
    if bathymetry is None and not bath_kw:
        H, mask = bathymetry_from_config(grid, config)
    else:
        get = config.get
        kw = dict(h_rel=float(get("bathymetry.h_rel", 0.9)),
                  length=float(get("bathymetry.length", 0.0)),
                  slope=float(get("bathymetry.slope", 0.5)),
                  spectrum_slope=float(get("bathymetry.spectrum_slope", 1.5)),
                  r_target=float(get("bathymetry.r_target", 0.1)),
                  seed=int(get("bathymetry.seed", 20260911)),
                  island=bool(get("bathymetry.island", False)))
        kw.update(bath_kw)
        H, mask = build_bathymetry(grid, bathymetry or str(get("bathymetry.kind")),
                                   float(get("physics.H")), **kw)
    return Domain(grid=grid, H=H, mask=mask,
                  vcoord=vcoord or str(config.get("domain.vcoord")),
                  bc_x=bc or str(config.get("domain.bc_x")),
                  bc_y=bc or str(config.get("domain.bc_y")),
                  min_partial=float(config.get("domain.min_partial", 0.1)),
                  face_rule=str(config.get("domain.face_rule", "min")))


def barotropic_dt_max(domain: Domain, g: float) -> float:
    """Surface-gravity CFL limit of the deepest wet column."""
    c = np.sqrt(g * float(np.max(domain.H[domain.mask > 0])))
    return 1.0 / (c * np.sqrt(1.0 / domain.grid.dx ** 2 + 1.0 / domain.grid.dy ** 2))


def simulate3d_v05(config, case_name: str, domain: Domain, physics: PhysicsV05, cfl: float,
                   *, n_steps_override: int | None = None, n_repeat: int = 1,
                   n_warmup: int = 0, device: str = "cuda:0",
                   precision: str = "fp64") -> dict:
    """Integrate one case on the GPU and return its metric with the cost and
    the run's description: dt = cfl x the surface-gravity limit; n_steps
    lands on the case's t_final unless overridden; the wall time is the
    median over `n_repeat` runs from the initial state."""
    import torch

    from .stepper import FLOPS_PER_CELL, HipStepper, to_numpy
    params = scheme_params_from_config(config)
    dt = cfl * barotropic_dt_max(domain, physics.g)
    stepper = HipStepper(domain, physics, params, dt, device=device, precision=precision)
    case = build_case3d_v05(case_name, domain, physics, config)
    case.device = device
    n_steps = (int(n_steps_override) if n_steps_override is not None
               else max(1, int(round(case.t_final / dt))))
    s0 = case.initial()
    for _ in range(n_warmup):
        stepper.integrate(s0, min(n_steps, 3))
    walls, final, counts = [], s0, None
    for _ in range(max(1, n_repeat)):
        before = stepper.counters()
        torch.cuda.synchronize(stepper.device)
        t0 = time.perf_counter()
        final = stepper.integrate(s0, n_steps)
        torch.cuda.synchronize(stepper.device)
        walls.append(time.perf_counter() - t0)
        after = stepper.counters()
        counts = {k: after[k] - before[k] for k in after}
    wall = float(np.median(walls))
    result = case.evaluate(to_numpy(final))
    return {
        "case": case_name, "metric_name": result.metric_name, "metric": result.metric,
        "extra": result.extra, "nx": domain.grid.nx, "nz": domain.grid.nz,
        "cells": domain.grid.nx * domain.grid.ny * domain.grid.nz,
        "n_steps": n_steps, "dt": dt, "cfl": cfl, "scheme": params.name,
        "solver": params.solver, "eos": physics.eos, "pgf": physics.pgf,
        "pgf_correction": physics.pgf_correction, "precision": precision,
        "wall_s": wall,
        "wall_mad_s": float(np.median(np.abs(np.array(walls) - wall))),
        "solver_iters_per_step": counts["solver_iterations"] / max(1, n_steps),
        "barotropic_substeps_per_step": counts["barotropic_substeps"] / max(1, n_steps),
        "tridiagonal_solves": counts["tridiagonal_solves"],
        "eos_flops": counts["eos_cells"] * FLOPS_PER_CELL[physics.eos],
        "moving_coordinate": stepper.moving_coordinate,
        **domain.summary(),
    }
