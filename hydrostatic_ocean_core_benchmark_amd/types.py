"""Parameter and state containers of the step API.

Field names, defaults and validation errors are those of the reference types
so a caller can hand either set to the HIP stepper:
  State3D    libs/core/model3d.py:34-54
  Physics3D  libs/core/model3d.py:57-90
  PhysicsV05 libs/core/model3d_v05.py:40-75 (delegates unknown names to base)
  SchemeParams + validate  libs/core/schemes.py:54-82
  SolveReport  libs/core/solvers.py:25-30

Two extensions, each reducing to the reference when uniform:
  * ``Physics3D.tau_x`` / ``tau_y`` may be a [ny, nx] array (wind stress per
    column); a scalar is the reference behaviour.
  * ``Physics3D.kappa`` may be a [ny, nx] array (background vertical
    diffusivity per column, constant with depth); a scalar is the reference.
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Any

import numpy as np

SCHEMES = ("fb", "theta", "split_explicit")
ADVECTION = ("none", "centered2", "upwind1", "up3", "up3_tvd")
SOLVERS = ("pcg_jacobi", "pcg_rbgs", "rbgs", "multigrid", "pcg_mg")
EOS_KINDS = ("linear", "seos", "teos10")


@dataclass(frozen=True)
class State3D:
    u: Any                 # [nz, ny, nx] east faces
    v: Any                 # [nz, ny, nx] north faces
    b: Any                 # [nz, ny, nx] buoyancy at centres
    eta: Any               # [ny, nx]
    t: float = 0.0
    T: Any = None          # [nz, ny, nx] when tracers == "TS"
    S: Any = None
    e: Any = None          # [nz+1, ny, nx] turbulent kinetic energy

    def astype(self, dtype) -> "State3D":
        out = replace(self, u=self.u.astype(dtype), v=self.v.astype(dtype),
                      b=self.b.astype(dtype), eta=self.eta.astype(dtype))
        if self.T is not None:
            out = replace(out, T=self.T.astype(dtype), S=self.S.astype(dtype))
        return out


@dataclass(frozen=True)
class Physics3D:
    g: float
    f: float
    H: float
    N2: float = 0.0
    nu: float = 0.0
    kappa: Any = 0.0
    rho0: float = 1025.0
    tau_x: Any = 0.0
    tau_y: Any = 0.0
    bottom_drag: float = 0.0
    tracers: str = "buoyancy"
    alpha_T: float = 2.0e-4
    beta_S: float = 7.4e-4
    T0: float = 10.0
    S0: float = 35.0
    cp: float = 3990.0
    A_h: float = 0.0
    K_h: float = 0.0
    q_heat: float = 0.0
    q_salt: float = 0.0

    def buoyancy(self, T, S):
        """Linear equation of state (libs/core/model3d.py:84-88)."""
        return self.g * (self.alpha_T * (T - self.T0) - self.beta_S * (S - self.S0))

    @property
    def c(self) -> float:
        return float(np.sqrt(self.g * self.H))


@dataclass(frozen=True)
class PhysicsV05:
    base: Physics3D
    eos: str = "linear"
    eta_ext: float = 0.0
    u_ext: float = 0.0
    n_relax: int = 0
    tau_relax: float = 3600.0
    pgf: str = "remove"
    pgf_correction: bool = True
    closure: str = "none"
    c_k: float = 0.1
    c_eps: float = 0.7
    pr_t: float = 1.0
    kappa_vk: float = 0.4
    z_0: float = 0.1
    e_min: float = 1.0e-6
    n2_min: float = 1.0e-8
    l_min: float = 0.01
    e_bb: float = 3.75
    mxl: str = "integral"

    def __getattr__(self, name):
        # Only reached for names not defined here: forward to the v0.4 set.
        if name == "base":
            raise AttributeError(name)
        return getattr(self.base, name)


@dataclass
class SchemeParams:
    name: str = "theta"
    theta: float = 0.5
    theta_cor: float = 0.5
    n_picard: int = 2
    theta_v: float = 0.5
    advection: str = "none"
    n_split: int = 0
    barotropic_average: str = "none"
    solver: str = "fft"
    rtol: float = 1e-12
    max_iter: int = 2000
    # pcg_mg only: multigrid preconditioner shape (no reference counterpart)
    mg_levels: int = 5
    mg_sweeps: int = 2
    mg_omega: float = 0.8
    mg_smoother: str = "jacobi"   # "jacobi" (damped) or "rbgs" (red-black Gauss-Seidel)

    def validate(self) -> None:
        if self.mg_smoother not in ("jacobi", "rbgs"):
            raise ValueError(f"unknown pcg_mg smoother '{self.mg_smoother}' (expected jacobi|rbgs)")
        if self.name not in SCHEMES:
            raise ValueError(f"unknown scheme '{self.name}' "
                             f"(expected fb|theta|split_explicit)")
        if self.name == "theta" and not (0.0 < self.theta <= 1.0):
            raise ValueError(f"theta must lie in (0, 1]; got {self.theta}")
        if self.n_picard < 1:
            raise ValueError("n_picard must be >= 1")
        if self.advection not in ADVECTION:
            raise ValueError(f"unknown advection '{self.advection}' "
                             f"(expected none|centered2|upwind1)")


@dataclass
class SolveReport:
    iterations: int = 0
    residual: float = 0.0
    converged: bool = True


def physics_v05_from_config(config, **over) -> PhysicsV05:
    """Same keys and fallbacks as libs/core/driver3d_v05.py:57-94."""
    g = config.get
    base = Physics3D(
        g=float(g("physics.g")), f=float(g("physics.f0")), H=float(g("physics.H")),
        N2=float(g("physics3d.N2")), nu=float(g("physics3d.nu")),
        kappa=float(g("physics3d.kappa")), rho0=float(g("physics.rho0")),
        tau_x=float(g("physics3d.tau_x")), tau_y=float(g("physics3d.tau_y")),
        bottom_drag=float(g("physics3d.bottom_drag")),
        tracers=str(g("physics3d_v04.tracers")),
        alpha_T=float(g("physics3d_v04.alpha_T")), beta_S=float(g("physics3d_v04.beta_S")),
        T0=float(g("physics3d_v04.T0")), S0=float(g("physics3d_v04.S0")),
        cp=float(g("physics3d_v04.cp")), A_h=float(g("physics3d_v04.A_h")),
        K_h=float(g("physics3d_v04.K_h")), q_heat=float(g("physics3d_v04.q_heat")),
        q_salt=float(g("physics3d_v04.q_salt")))
    kw: dict[str, Any] = dict(
        eos=str(g("physics3d_v05.eos")), pgf=str(g("physics3d_v05.pgf")),
        pgf_correction=bool(g("physics3d_v05.pgf_correction", True)),
        eta_ext=float(g("physics3d_v05.eta_ext")), u_ext=float(g("physics3d_v05.u_ext")),
        n_relax=int(g("physics3d_v05.n_relax")), tau_relax=float(g("physics3d_v05.tau_relax")))
    v6 = config.get("physics3d_v06", {}) or {}
    kw["closure"] = str(v6.get("closure", "none"))
    kw["mxl"] = str(v6.get("mxl", "integral"))
    for key in ("c_k", "c_eps", "pr_t", "kappa_vk", "z_0", "e_min", "n2_min", "l_min", "e_bb"):
        if key in v6:
            kw[key] = float(v6[key])
    kw.update(over)
    return PhysicsV05(base=base, **kw)


def scheme_params_from_config(config) -> SchemeParams:
    """Same keys as libs/core/schemes.py:180-196, plus the pcg_mg knobs."""
    g = config.get
    params = SchemeParams(
        name=str(g("scheme.name")), theta=float(g("scheme.theta")),
        theta_cor=float(g("scheme.theta_cor")), theta_v=float(g("scheme.theta_v", 0.5)),
        advection=str(g("scheme.advection", "none")), n_split=int(g("scheme.n_split", 0)),
        barotropic_average=str(g("scheme.barotropic_average", "none")),
        n_picard=int(g("scheme.n_picard")), solver=str(g("solver.kind")),
        rtol=float(g("solver.rtol")), max_iter=int(g("solver.max_iter")),
        mg_levels=int(g("solver.mg_levels", 5)), mg_sweeps=int(g("solver.mg_sweeps", 2)),
        mg_omega=float(g("solver.mg_omega", 0.8)),
        mg_smoother=str(g("solver.mg_smoother", "jacobi")))
    params.validate()
    return params
