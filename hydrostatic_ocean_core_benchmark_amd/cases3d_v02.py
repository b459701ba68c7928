"""Verification cases, stability limits and the memory model of the v0.2 3D
model (libs/core/cases3d.py:21-280, libs/core/limits.py:16-101): the
barotropic wave over nz layers, the vertical-diffusion eigenmode, the exact
baroclinic internal wave, uniform-flow tracer advection and the Ekman
spiral, each with its exact solution. Initial conditions, exact solutions
and limits are host-side NumPy, as in the reference; the step runs on the
device (model3d_v02.Stepper3D)."""

from __future__ import annotations

from dataclasses import asdict, dataclass
from typing import Any

import numpy as np

from .types import Physics3D, State3D


def physics3d_from_config(config) -> Physics3D:
    """libs/core/driver3d.py:19-42"""
    return Physics3D(
        g=float(config.get("physics.g")), f=float(config.get("physics.f0")),
        H=float(config.get("physics.H")), N2=float(config.get("physics3d.N2", 0.0)),
        nu=float(config.get("physics3d.nu", 0.0)),
        kappa=float(config.get("physics3d.kappa", 0.0)),
        rho0=float(config.get("physics.rho0", 1025.0)),
        tau_x=float(config.get("physics3d.tau_x", 0.0)),
        tau_y=float(config.get("physics3d.tau_y", 0.0)),
        bottom_drag=float(config.get("physics3d.bottom_drag", 0.0)),
        tracers=str(config.get("physics3d_v04.tracers", "buoyancy")),
        alpha_T=float(config.get("physics3d_v04.alpha_T", 2.0e-4)),
        beta_S=float(config.get("physics3d_v04.beta_S", 7.4e-4)),
        T0=float(config.get("physics3d_v04.T0", 10.0)),
        S0=float(config.get("physics3d_v04.S0", 35.0)),
        cp=float(config.get("physics3d_v04.cp", 3990.0)),
        A_h=float(config.get("physics3d_v04.A_h", 0.0)),
        K_h=float(config.get("physics3d_v04.K_h", 0.0)),
        q_heat=float(config.get("physics3d_v04.q_heat", 0.0)),
        q_salt=float(config.get("physics3d_v04.q_salt", 0.0)))


class Barotropic3D:
    """The 2D inertia-gravity wave replicated over the layers (cases3d.py:21-62)."""
    name = "barotropic3d"

    def __init__(self, grid, physics: Physics3D, mode_x: int, mode_y: int, eta0: float):
        self.grid, self.physics, self.eta0 = grid, physics, eta0
        self.k = 2.0 * np.pi * mode_x / grid.Lx
        self.l = 2.0 * np.pi * mode_y / grid.Ly
        self.kappa2 = self.k ** 2 + self.l ** 2
        self.omega = float(np.sqrt(physics.f ** 2 + physics.g * physics.H * self.kappa2))
        self.period = 2.0 * np.pi / self.omega

    def exact(self, t: float) -> State3D:
        f, H = self.physics.f, self.physics.H
        amp = self.eta0 / (H * self.kappa2)
        (xe, ye), (xu, yu), (xv, yv) = (self.grid.coords(n) for n in ("eta", "u", "v"))
        pe = self.k * xe + self.l * ye - self.omega * t
        pu = self.k * xu + self.l * yu - self.omega * t
        pv = self.k * xv + self.l * yv - self.omega * t
        eta = self.eta0 * np.cos(pe)
        u2 = amp * (self.omega * self.k * np.cos(pu) - f * self.l * np.sin(pu))
        v2 = amp * (self.omega * self.l * np.cos(pv) + f * self.k * np.sin(pv))
        ones = np.ones((self.grid.nz, 1, 1))
        return State3D(u=ones * u2, v=ones * v2, b=np.zeros(self.grid.shape3d), eta=eta, t=t)

    def initial(self) -> State3D:
        return self.exact(0.0)


class VerticalDiffusion:
    """Decay of the discrete vertical eigenmode mode_z (cases3d.py:65-107)."""
    name = "vdiffusion"

    def __init__(self, grid, physics: Physics3D, mode_z: int, u0: float):
        if physics.nu <= 0.0:
            raise ValueError("vdiffusion requires physics.nu > 0")
        if abs(physics.f) > 0.0:
            raise ValueError("VerticalDiffusion needs f = 0 (its exact solution has no "
                             "Coriolis); pass --set physics.f0=0")
        self.grid, self.physics, self.u0 = grid, physics, u0
        self.m = np.pi * mode_z / physics.H
        self.lam = (2.0 * np.cos(np.pi * mode_z / grid.nz) - 2.0) / grid.dz ** 2
        self.rate = physics.nu * self.lam
        self.period = -1.0 / self.rate if self.rate < 0.0 else np.inf
        self._profile = np.cos(self.m * grid.z_centre())[:, None, None]

    def exact(self, t: float) -> State3D:
        shape = self.grid.shape3d
        u = self.u0 * self._profile * np.exp(self.rate * t) * np.ones(shape)
        return State3D(u=u, v=np.zeros(shape), b=np.zeros(shape),
                       eta=np.zeros(self.grid.shape), t=t)

    def initial(self) -> State3D:
        return self.exact(0.0)


class BaroclinicIGW:
    """Exact internal gravity wave of vertical mode mode_z (cases3d.py:110-166)."""
    name = "baroclinic_igw"

    def __init__(self, grid, physics: Physics3D, mode_x: int, mode_y: int, mode_z: int,
                 eta0: float):
        if physics.N2 <= 0.0:
            raise ValueError("baroclinic_igw requires physics.N2 > 0")
        self.grid, self.physics, self.eta0 = grid, physics, eta0
        self.k = 2.0 * np.pi * mode_x / grid.Lx
        self.l = 2.0 * np.pi * mode_y / grid.Ly
        self.kappa2 = self.k ** 2 + self.l ** 2
        self.m = np.pi * mode_z / physics.H
        self.c2 = physics.N2 / self.m ** 2
        self.omega = float(np.sqrt(physics.f ** 2 + self.c2 * self.kappa2))
        self.period = 2.0 * np.pi / self.omega
        zt = grid.z_centre()[:, None, None]
        self._cos_z = np.cos(self.m * zt)
        self._sin_z = np.sin(self.m * zt)

    def exact(self, t: float) -> State3D:
        g_, f = self.physics.g, self.physics.f
        amp = g_ * self.eta0 / (self.c2 * self.kappa2)
        (xu, yu), (xv, yv), (xe, ye) = (self.grid.coords(n) for n in ("u", "v", "eta"))
        pu = self.k * xu + self.l * yu - self.omega * t
        pv = self.k * xv + self.l * yv - self.omega * t
        pe = self.k * xe + self.l * ye - self.omega * t
        u = amp * (self.omega * self.k * np.cos(pu) - f * self.l * np.sin(pu)) * self._cos_z
        v = amp * (self.omega * self.l * np.cos(pv) + f * self.k * np.sin(pv)) * self._cos_z
        b = -g_ * self.eta0 * self.m * np.cos(pe) * self._sin_z
        return State3D(u=u, v=v, b=b, eta=np.zeros(self.grid.shape), t=t)

    def initial(self) -> State3D:
        return self.exact(0.0)

    def exact_w(self, t: float) -> np.ndarray:
        xe, ye = self.grid.coords("eta")
        pe = self.k * xe + self.l * ye - self.omega * t
        zt = self.grid.z_interface()[:, None, None]
        scale = self.physics.g * self.eta0 * self.omega / (self.c2 * self.m)
        return scale * np.sin(pe) * np.sin(self.m * zt)


class TracerAdvection:
    """A tracer sinusoid carried by a uniform flow (cases3d.py:169-206)."""
    name = "tracer_advect"

    def __init__(self, grid, physics: Physics3D, mode_x: int, mode_y: int, amp: float,
                 u0: float, v0: float):
        if physics.tracers != "TS":
            raise ValueError("tracer_advect requires physics3d_v04.tracers = 'TS'")
        if physics.f != 0.0:
            raise ValueError("tracer_advect requires f = 0; Coriolis would turn "
                             "the uniform flow and break the exact solution")
        self.grid, self.physics = grid, physics
        self.amp, self.u0, self.v0 = amp, u0, v0
        self.k = 2.0 * np.pi * mode_x / grid.Lx
        self.l = 2.0 * np.pi * mode_y / grid.Ly
        speed = np.hypot(u0, v0)
        self.period = (grid.Lx / speed) if speed > 0.0 else np.inf

    def exact(self, t: float) -> State3D:
        shape = self.grid.shape3d
        x, y = self.grid.coords("eta")
        phase = self.k * (x - self.u0 * t) + self.l * (y - self.v0 * t)
        T = self.physics.T0 + self.amp * np.cos(phase)[None, :, :] * np.ones(shape)
        S = np.full(shape, self.physics.S0)
        return State3D(u=np.full(shape, self.u0), v=np.full(shape, self.v0),
                       b=self.physics.buoyancy(T, S), eta=np.zeros(self.grid.shape), t=t,
                       T=T, S=S)

    def initial(self) -> State3D:
        return self.exact(0.0)


class EkmanSpiral:
    """Steady wind-driven Ekman spiral of the finite-depth, free-slip-bottom
    problem (cases3d.py:209-250)."""
    name = "ekman"

    def __init__(self, grid, physics: Physics3D, n_inertial: float):
        if physics.nu <= 0.0 or physics.f == 0.0:
            raise ValueError("ekman requires nu > 0 and f != 0")
        if physics.tau_x == 0.0 and physics.tau_y == 0.0:
            raise ValueError("ekman requires a non-zero surface stress")
        self.grid, self.physics = grid, physics
        self.D = np.sqrt(2.0 * physics.nu / abs(physics.f))
        a = (1.0 + 1j) / self.D
        tau = complex(physics.tau_x, physics.tau_y)
        amp = tau / (physics.rho0 * physics.nu * a * np.sinh(a * physics.H))
        zt = grid.z_centre()[:, None, None]
        self._W = amp * np.cosh(a * zt)
        self.period = 2.0 * np.pi / abs(physics.f)
        self.n_inertial = n_inertial

    def exact(self, t: float) -> State3D:
        shape = self.grid.shape3d
        return State3D(u=np.real(self._W) * np.ones(shape), v=np.imag(self._W) * np.ones(shape),
                       b=np.zeros(shape), eta=np.zeros(self.grid.shape), t=t)

    def initial(self) -> State3D:
        return self.exact(0.0)

    def surface_angle(self) -> float:
        return float(np.degrees(np.angle(self._W[0, 0, 0])))


def build_case3d(name: str, grid, physics: Physics3D, config):
    """cases3d.py:253-280, parameters from config/cases.toml."""
    sec = config.section(f"case.{name}")
    if name == "barotropic3d":
        case = Barotropic3D(grid, physics, int(sec["mode_x"]), int(sec["mode_y"]),
                            float(sec["eta0"]))
        t_final = float(sec["n_periods"]) * case.period
    elif name == "vdiffusion":
        case = VerticalDiffusion(grid, physics, int(sec["mode_z"]), float(sec["u0"]))
        t_final = float(sec["n_efolds"]) * case.period
    elif name == "tracer_advect":
        case = TracerAdvection(grid, physics, int(sec["mode_x"]), int(sec["mode_y"]),
                               float(sec["amp"]), float(sec["u0"]), float(sec["v0"]))
        t_final = float(sec["n_transits"]) * case.period
    elif name == "ekman":
        case = EkmanSpiral(grid, physics, float(sec["n_inertial"]))
        t_final = float(sec["n_inertial"]) * case.period
    elif name == "baroclinic_igw":
        case = BaroclinicIGW(grid, physics, int(sec["mode_x"]), int(sec["mode_y"]),
                             int(sec["mode_z"]), float(sec["eta0"]))
        t_final = float(sec["n_periods"]) * case.period
    else:
        raise ValueError(f"unknown 3D case '{name}' (implemented: barotropic3d|"
                         f"vdiffusion|baroclinic_igw|tracer_advect|ekman)")
    return case, t_final


CASES_3D = ("barotropic3d", "vdiffusion", "baroclinic_igw", "tracer_advect", "ekman")


# ------------------------------------------------------------------ limits
@dataclass(frozen=True)
class TimeStepLimits:
    """Every explicit stability limit in seconds (limits.py:16-44)."""
    surface_gravity: float
    internal_wave: float
    advective: float
    horizontal_diffusion: float
    vertical_diffusion: float
    coriolis: float

    def for_scheme(self, name: str, advection: str, A_h: float) -> float:
        caps = [self.coriolis]
        caps.append(self.surface_gravity if name == "fb" else self.internal_wave)
        if advection != "none":
            caps.append(self.advective)
        if A_h > 0.0:
            caps.append(self.horizontal_diffusion)
        return float(min(caps))

    def to_dict(self) -> dict[str, Any]:
        return asdict(self)


def time_step_limits(grid, physics: Physics3D, u_scale: float = 1.0) -> TimeStepLimits:
    """limits.py:47-69"""
    s_max = np.sqrt(4.0 / grid.dx ** 2 + 4.0 / grid.dy ** 2)
    c_ext = np.sqrt(physics.g * physics.H)
    c_int = (np.sqrt(physics.N2) * physics.H / np.pi) if physics.N2 > 0.0 else c_ext
    inf = float("inf")
    return TimeStepLimits(
        surface_gravity=float(2.0 / (c_ext * s_max)),
        internal_wave=float(2.0 / (c_int * s_max)) if c_int > 0 else inf,
        advective=float(grid.dx / u_scale) if u_scale > 0 else inf,
        horizontal_diffusion=(float(grid.dx ** 2 / (4.0 * physics.A_h))
                              if physics.A_h > 0 else inf),
        vertical_diffusion=(float(grid.dz ** 2 / (2.0 * physics.nu))
                            if physics.nu > 0 else inf),
        coriolis=float(2.0 / abs(physics.f)) if physics.f else inf)


FIELDS_3D_THETA = 28
FIELDS_3D_SPLIT = 30
FIELDS_2D = 12


def memory_model(grid, scheme: str, itemsize: int = 8) -> dict[str, Any]:
    """The reference's working-set model of the NumPy step (limits.py:72-101);
    reported as the reference reports it, not a measurement of this backend."""
    cells = grid.nx * grid.ny * grid.nz
    n3 = FIELDS_3D_SPLIT if scheme == "split_explicit" else FIELDS_3D_THETA
    total = (n3 * cells + FIELDS_2D * grid.nx * grid.ny) * itemsize
    per_cell = total / cells
    return {"cells": cells, "fields_3d": n3, "bytes": int(total), "bytes_per_cell": per_cell,
            "gib": total / 2 ** 30, "max_cells_32gib": int(32 * 2 ** 30 / per_cell),
            "max_nx_32gib_nz30": int(np.sqrt(32 * 2 ** 30 / per_cell / 30))}
