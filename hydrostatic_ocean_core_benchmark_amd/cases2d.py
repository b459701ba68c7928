"""Verification cases and diagnostics of the 2D model
(libs/core/cases.py:29-239, libs/core/diagnostics.py:19-100): the
inertia-gravity wave (single mode and broadband) and the discrete
geostrophic balance, with their exact solutions, and the metrics block of
the reference's run/verify/bench (relative L2/Linf errors, mass and energy
drift, modal amplitude ratio and phase error). Initial conditions and
metrics are host-side NumPy, as in the reference; the step runs on the
device (model2d.SWStepper)."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .model2d import Physics, State


@dataclass(frozen=True)
class CaseSpec:
    name: str
    mode_x: int
    mode_y: int
    eta0: float
    t_final: float
    period: float
    has_exact: bool


def _shift(a, di=0, dj=0):
    """a[j + dj, i + di], periodic (the reference's np.roll(a, -d))."""
    return np.roll(np.roll(a, -di, axis=-1), -dj, axis=-2)


def gradx_u(eta, grid):
    return (_shift(eta, di=1) - eta) / grid.dx


def grady_v(eta, grid):
    return (_shift(eta, dj=1) - eta) / grid.dy


def div_eta(u, v, grid):
    return (u - _shift(u, di=-1)) / grid.dx + (v - _shift(v, dj=-1)) / grid.dy


def avg_v_to_u(v):
    vs = _shift(v, dj=-1)
    return 0.25 * (v + vs + _shift(v, di=1) + _shift(vs, di=1))


def avg_u_to_v(u):
    uw = _shift(u, di=-1)
    return 0.25 * (u + uw + _shift(u, dj=1) + _shift(uw, dj=1))


class InertiaGravityWave:
    """Linear inertia-gravity wave on an f-plane (cases.py:41-79)."""
    name = "igw"
    has_exact = True

    def __init__(self, grid, physics: Physics, mode_x: int, mode_y: int, eta0: float):
        self.grid, self.physics, self.eta0 = grid, physics, eta0
        self.k = 2.0 * np.pi * mode_x / grid.Lx
        self.l = 2.0 * np.pi * mode_y / grid.Ly
        self.kappa2 = self.k ** 2 + self.l ** 2
        self.omega = float(np.sqrt(physics.f ** 2 + physics.g * physics.H * self.kappa2))
        self.period = 2.0 * np.pi / self.omega

    def exact(self, t: float) -> State:
        f, H = self.physics.f, self.physics.H
        amp = self.eta0 / (H * self.kappa2)
        x_e, y_e = self.grid.coords("eta")
        x_u, y_u = self.grid.coords("u")
        x_v, y_v = self.grid.coords("v")
        ph_e = self.k * x_e + self.l * y_e - self.omega * t
        ph_u = self.k * x_u + self.l * y_u - self.omega * t
        ph_v = self.k * x_v + self.l * y_v - self.omega * t
        eta = self.eta0 * np.cos(ph_e)
        u = amp * (self.omega * self.k * np.cos(ph_u) - f * self.l * np.sin(ph_u))
        v = amp * (self.omega * self.l * np.cos(ph_v) + f * self.k * np.sin(ph_v))
        return State(u=u, v=v, eta=eta, t=t)

    def initial(self) -> State:
        return self.exact(0.0)


class InertiaGravityWaveBroadband:
    """Superposition of n_modes^2 exact modes with a red spectrum and
    integer-arithmetic phases (cases.py:82-145)."""
    name = "igw_broadband"
    has_exact = True

    def __init__(self, grid, physics: Physics, n_modes: int, eta0: float, slope: float):
        self.grid, self.physics, self.eta0 = grid, physics, eta0
        modes = [(m, n) for m in range(1, n_modes + 1) for n in range(1, n_modes + 1)]
        k = np.array([2.0 * np.pi * m / grid.Lx for m, _ in modes])
        l = np.array([2.0 * np.pi * n / grid.Ly for _, n in modes])
        kappa2 = k ** 2 + l ** 2
        omega = np.sqrt(physics.f ** 2 + physics.g * physics.H * kappa2)
        raw = np.array([(m * m + n * n) ** (-0.5 * slope) for m, n in modes])
        amp = raw * eta0 / np.sqrt(np.sum(raw ** 2) / 2.0)
        phase = np.array([2.0 * np.pi * ((m * 37 + n * 17) % 101) / 101.0 for m, n in modes])
        self.modes = modes
        self.k, self.l, self.kappa2 = k, l, kappa2
        self.omega, self.amp, self.phase = omega, amp, phase
        self.period = 2.0 * np.pi / float(np.min(omega))

    def exact(self, t: float) -> State:
        f, H = self.physics.f, self.physics.H
        x_e, y_e = self.grid.coords("eta")
        x_u, y_u = self.grid.coords("u")
        x_v, y_v = self.grid.coords("v")
        eta = np.zeros_like(x_e)
        u = np.zeros_like(x_u)
        v = np.zeros_like(x_v)
        for idx in range(len(self.modes)):
            k, l, om = self.k[idx], self.l[idx], self.omega[idx]
            a, ph = self.amp[idx], self.phase[idx]
            scale = a / (H * self.kappa2[idx])
            pe = k * x_e + l * y_e - om * t + ph
            pu = k * x_u + l * y_u - om * t + ph
            pv = k * x_v + l * y_v - om * t + ph
            eta += a * np.cos(pe)
            u += scale * (om * k * np.cos(pu) - f * l * np.sin(pu))
            v += scale * (om * l * np.cos(pv) + f * k * np.sin(pv))
        return State(u=u, v=v, eta=eta, t=t)

    def initial(self) -> State:
        return self.exact(0.0)


class GeostrophicBalance:
    """One Fourier mode in exact discrete geostrophic balance (cases.py:148-213)."""
    name = "geo_balance"
    has_exact = True

    def __init__(self, grid, physics: Physics, mode_x: int, mode_y: int, eta0: float):
        if physics.f == 0.0:
            raise ValueError("geo_balance requires a non-zero Coriolis parameter")
        self.grid, self.physics, self.eta0 = grid, physics, eta0
        self.period = 2.0 * np.pi / abs(physics.f)
        tx = 2.0 * np.pi * mode_x / grid.nx
        ty = 2.0 * np.pi * mode_y / grid.ny
        if np.isclose(abs(np.cos(tx / 2.0)), 0.0) or np.isclose(abs(np.cos(ty / 2.0)), 0.0):
            raise ValueError("geo_balance mode hits the grid Nyquist limit; "
                             "the averaging operator is singular there")
        ex, ey = np.exp(1j * tx), np.exp(1j * ty)
        fac = 4.0 * physics.g / physics.f
        self._v_hat = fac * (ex - 1.0) / (grid.dx * (1.0 + np.conj(ey)) * (1.0 + ex)) * eta0
        self._u_hat = -fac * (ey - 1.0) / (grid.dy * (1.0 + np.conj(ex)) * (1.0 + ey)) * eta0
        ii, jj = np.meshgrid(np.arange(grid.nx), np.arange(grid.ny), indexing="xy")
        self._phase = np.exp(1j * (tx * ii + ty * jj))

    def initial(self) -> State:
        s = self.exact(0.0)
        g_, f = self.physics.g, self.physics.f
        res_u = f * avg_v_to_u(s.v) - g_ * gradx_u(s.eta, self.grid)
        res_v = -f * avg_u_to_v(s.u) - g_ * grady_v(s.eta, self.grid)
        scale = g_ * np.max(np.abs(gradx_u(s.eta, self.grid))) + 1e-300
        worst = max(np.max(np.abs(res_u)), np.max(np.abs(res_v))) / scale
        if worst > 1e-10:
            raise AssertionError(f"geo_balance is not an exact discrete steady state "
                                 f"(relative momentum residual {worst:.3e})")
        return s

    def exact(self, t: float) -> State:
        return State(u=np.real(self._u_hat * self._phase), v=np.real(self._v_hat * self._phase),
                     eta=self.eta0 * np.real(self._phase), t=t)


def build_case(name: str, grid, physics: Physics, config):
    """cases.py:216-239, parameters from config/cases.toml."""
    section = config.section(f"case.{name}")
    if name == "igw":
        case = InertiaGravityWave(grid, physics, int(section["mode_x"]), int(section["mode_y"]),
                                  float(section["eta0"]))
        t_final = float(section["n_periods"]) * case.period
    elif name == "igw_broadband":
        case = InertiaGravityWaveBroadband(grid, physics, int(section["n_modes"]),
                                           float(section["eta0"]), float(section["slope"]))
        t_final = float(section["n_periods"]) * case.period
    elif name == "geo_balance":
        case = GeostrophicBalance(grid, physics, int(section["mode_x"]), int(section["mode_y"]),
                                  float(section["eta0"]))
        t_final = float(section["n_days"]) * 86400.0
    else:
        raise ValueError(f"unknown case '{name}' (implemented: igw|igw_broadband|geo_balance)")
    spec = CaseSpec(name=name, mode_x=int(section.get("mode_x", 0)),
                    mode_y=int(section.get("mode_y", 0)), eta0=float(section["eta0"]),
                    t_final=t_final, period=case.period, has_exact=case.has_exact)
    return case, spec


# ------------------------------------------------------------- diagnostics
def l2_rel(numeric, exact) -> float:
    denom = float(np.sqrt(np.sum(exact * exact)))
    num = float(np.sqrt(np.sum((numeric - exact) ** 2)))
    return num / denom if denom > 0.0 else num


def linf_rel(numeric, exact) -> float:
    denom = float(np.max(np.abs(exact)))
    num = float(np.max(np.abs(numeric - exact)))
    return num / denom if denom > 0.0 else num


def total_mass(s: State, grid) -> float:
    return float(np.sum(s.eta)) * grid.cell_area


def total_energy(s: State, grid, physics: Physics) -> float:
    u2 = 0.5 * (s.u ** 2 + np.roll(s.u, 1, axis=1) ** 2)
    v2 = 0.5 * (s.v ** 2 + np.roll(s.v, 1, axis=0) ** 2)
    density = physics.H * (u2 + v2) + physics.g * s.eta ** 2
    return 0.5 * float(np.sum(density)) * grid.cell_area


def modal_amplitude(field, grid, k: float, l: float, variable: str = "eta") -> complex:
    x, y = grid.coords(variable)
    basis = np.exp(-1j * (k * x + l * y))
    return complex(2.0 * np.sum(field * basis) / field.size)


def wave_errors(numeric: State, exact: State, grid, k: float, l: float) -> dict:
    a_num = modal_amplitude(numeric.eta, grid, k, l, "eta")
    a_exa = modal_amplitude(exact.eta, grid, k, l, "eta")
    if abs(a_exa) == 0.0:
        return {"amp_ratio": float("nan"), "phase_err_rad": float("nan")}
    ratio = a_num / a_exa
    return {"amp_ratio": float(abs(ratio)), "phase_err_rad": float(np.angle(ratio))}


def summarize(numeric: State, exact: State, grid, physics: Physics, m0: float,
              e0: float) -> dict:
    """diagnostics.py:69-100"""
    mass = total_mass(numeric, grid)
    energy = total_energy(numeric, grid, physics)
    domain_area = grid.cell_area * grid.nx * grid.ny
    mass_scale = float(np.sqrt(np.mean(exact.eta ** 2))) * domain_area
    if mass_scale == 0.0:
        mass_scale = 1.0
    return {
        "mass_scale": mass_scale,
        "l2_rel_eta": l2_rel(numeric.eta, exact.eta),
        "l2_rel_u": l2_rel(numeric.u, exact.u),
        "l2_rel_v": l2_rel(numeric.v, exact.v),
        "linf_rel_eta": linf_rel(numeric.eta, exact.eta),
        "mass": mass,
        "energy": energy,
        "mass_drift": (mass - m0) / mass_scale,
        "energy_drift": (energy - e0) / abs(e0) if e0 != 0.0 else energy - e0,
    }
