"""The 2D rotating shallow-water model of the reference's original driver
(libs/core/schemes.py:15-202), on the GPU.

`SWStepper(grid, physics, params, dt)` has the constructor, `step`,
`integrate` and counters of the reference's `Stepper`; the state is
`State(u, v, eta, t)` on the doubly periodic C-grid (u at east faces, v at
north faces, eta at centres; [ny, nx]). Both schemes run their Picard loop
here with the HIP operators of the C-ABI (hoc_sw2d_*): forward-backward, and
the theta scheme with a Helmholtz solve per Picard pass (`Helmholtz2D`,
build_solver of libs/core/solvers.py:261-281) by
  fft         the spectral solve of HelmholtzFFT (solvers.py:33-52) with the
              device FFT (rocFFT through torch.fft): the same symbol, agreement
              with NumPy's FFT to round-off (not bitwise);
  pcg_jacobi, pcg_rbgs, rbgs, multigrid
              HelmholtzPCG (Jacobi or symmetric red-black Gauss-Seidel
              preconditioner), HelmholtzRBGS and HelmholtzMultigrid on the
              device (hoc_helm2d_*), the dot products in NumPy's pairwise
              order, so iterates and iteration counts are bitwise the
              reference's.
"""

from __future__ import annotations

import ctypes as C
import sys
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib


@dataclass(frozen=True)
class State:
    """Prognostic state on the C-grid (libs/core/schemes.py:15-25)."""
    u: object
    v: object
    eta: object
    t: float = 0.0

    def astype(self, dtype) -> "State":
        return replace(self, u=self.u.astype(dtype), v=self.v.astype(dtype),
                       eta=self.eta.astype(dtype))


@dataclass(frozen=True)
class Physics:
    """g, f, H (libs/core/schemes.py:28-38)."""
    g: float
    f: float
    H: float

    @property
    def c(self) -> float:
        return float(np.sqrt(self.g * self.H))


def explicit_dt_max(grid, physics: Physics) -> float:
    """Forward-backward stability limit, also the CFL yardstick
    (libs/core/schemes.py:80-93)."""
    s_max = np.sqrt(4.0 / grid.dx ** 2 + 4.0 / grid.dy ** 2)
    dt_gravity = 2.0 / (physics.c * s_max)
    dt_coriolis = 2.0 / abs(physics.f) if physics.f != 0.0 else np.inf
    return float(min(dt_gravity, dt_coriolis))


def physics_from_config(config) -> Physics:
    """libs/core/schemes.py:197-200"""
    return Physics(g=float(config.get("physics.g")), f=float(config.get("physics.f0")),
                   H=float(config.get("physics.H")))


SOLVERS = ("fft", "pcg_jacobi", "pcg_rbgs", "rbgs", "multigrid")
_KIND = {"pcg_jacobi": 0, "pcg_rbgs": 1, "rbgs": 2, "multigrid": 3}


@dataclass
class SolveReport:
    """Per-solve diagnostics (libs/core/solvers.py:24-30)."""
    iterations: int = 0
    residual: float = 0.0
    converged: bool = True


class Helmholtz2D:
    """The free-surface solve (I - coef L) x = rhs on a doubly periodic grid,
    as build_solver(kind, grid, coef, rtol, max_iter) builds it
    (libs/core/solvers.py:261-281); `solve(rhs)` returns (x, SolveReport)
    with x a device tensor. Invalid setups raise the reference's ValueError."""

    def __init__(self, kind: str, grid, coef: float, rtol: float, max_iter: int,
                 device="cuda:0"):
        if kind not in SOLVERS:
            raise ValueError(f"unknown solver kind '{kind}' "
                             "(expected fft|pcg_jacobi|pcg_rbgs|rbgs|multigrid)")
        if kind == "fft" and (getattr(grid, "bc_x", "periodic") != "periodic"
                              or getattr(grid, "bc_y", "periodic") != "periodic"):
            raise ValueError("the fft solver requires doubly periodic boundaries")
        if kind in ("rbgs", "multigrid", "pcg_rbgs") and (grid.nx % 2 or grid.ny % 2):
            raise ValueError(f"{kind} needs even nx and ny for a consistent red-black "
                             f"colouring under periodic wrap; got {grid.nx}x{grid.ny}")
        self.kind = kind
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("Helmholtz2D needs a HIP device (torch.cuda)")
        self.shape = (grid.ny, grid.nx)
        self._h = None
        self._symbol = None
        if kind == "fft":
            # HelmholtzFFT's symbol 1 - coef * lambda (solvers.py:38-47)
            theta_x = 2.0 * np.pi * np.arange(grid.nx // 2 + 1) / grid.nx
            theta_y = 2.0 * np.pi * np.arange(grid.ny) / grid.ny
            lam_x = 2.0 * (np.cos(theta_x) - 1.0) / grid.dx ** 2
            lam_y = 2.0 * (np.cos(theta_y) - 1.0) / grid.dy ** 2
            lam = lam_y[:, None] + lam_x[None, :]
            self._symbol = torch.as_tensor(1.0 - coef * lam, device=self.device)
            self.n_levels = 1
            return
        cfg = _lib.HocHelm2dConfig()
        cfg.nx, cfg.ny, cfg.dx, cfg.dy = grid.nx, grid.ny, grid.dx, grid.dy
        cfg.coef, cfg.kind, cfg.rtol, cfg.max_iter = coef, _KIND[kind], rtol, max_iter
        h = C.c_void_p()
        _lib.check(self.lib.hoc_helm2d_create(C.byref(cfg), C.byref(h)))
        self._h = h
        _lib.register_handle(self)
        nl = C.c_int32(0)
        _lib.check(self.lib.hoc_helm2d_levels(h, C.byref(nl)))
        self.n_levels = nl.value

    def close(self) -> None:
        """Destroy the library handle. Idempotent; also runs at interpreter exit."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            self.lib.hoc_helm2d_destroy(h)

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at class creation: module globals are None during teardown
        if not _finalizing():
            self.close()

    def solve(self, rhs: torch.Tensor):
        rhs = rhs.to(device=self.device, dtype=torch.float64).contiguous()
        if tuple(rhs.shape) != self.shape:
            raise ValueError(f"rhs shape {tuple(rhs.shape)} does not match the grid {self.shape}")
        if self._symbol is not None:
            spec = torch.fft.rfft2(rhs)
            return (torch.fft.irfft2(spec / self._symbol, s=self.shape).contiguous(),
                    SolveReport())
        x = torch.empty_like(rhs)
        it, res, cv = C.c_int32(0), C.c_double(0.0), C.c_int32(0)
        st = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(self.lib.hoc_helm2d_solve(self._h, _lib.ptr(rhs), _lib.ptr(x), C.byref(it),
                                             C.byref(res), C.byref(cv), st))
        return x, SolveReport(it.value, res.value, bool(cv.value))


class SWStepper:
    """One-step integrator for a fixed (grid, physics, params, dt)."""

    def __init__(self, grid, physics: Physics, params, dt: float, device="cuda:0"):
        params.validate()
        if params.name not in ("fb", "theta"):
            raise ValueError(f"scheme '{params.name}' is not a 2D scheme (expected fb|theta)")
        if getattr(grid, "bc_x", "periodic") != "periodic" or \
                getattr(grid, "bc_y", "periodic") != "periodic":
            raise NotImplementedError("the 2D model is doubly periodic")
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("SWStepper needs a HIP device (torch.cuda)")
        self.grid, self.physics, self.params = grid, physics, params
        self.dt = float(dt)
        self.solver_iterations = 0
        self.solver_failures = 0
        cfg = _lib.HocSw2dConfig()
        cfg.nx, cfg.ny = grid.nx, grid.ny
        cfg.dx, cfg.dy = grid.dx, grid.dy
        cfg.g, cfg.f, cfg.H, cfg.dt = physics.g, physics.f, physics.H, self.dt
        cfg.scheme = 0 if params.name == "theta" else 1
        cfg.theta, cfg.theta_cor = params.theta, params.theta_cor
        h = C.c_void_p()
        _lib.check(self.lib.hoc_sw2d_create(C.byref(cfg), C.byref(h)))
        self._h = h
        _lib.register_handle(self)
        self.solver = None
        if params.name == "theta":
            # schemes.py:110-113
            coef = physics.g * physics.H * params.theta ** 2 * self.dt ** 2
            self.solver = Helmholtz2D(params.solver, grid, coef, params.rtol, params.max_iter,
                                      device=self.device)

    def close(self) -> None:
        """Destroy the library handle. Idempotent; also runs at interpreter exit."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            self.lib.hoc_sw2d_destroy(h)

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at class creation: module globals are None during teardown
        if not _finalizing():
            self.close()

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, a) -> torch.Tensor:
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=torch.float64).contiguous()
        return torch.as_tensor(np.array(a, dtype=np.float64, order="C"), device=self.device)

    def _solve(self, rhs: torch.Tensor) -> torch.Tensor:
        x, rep = self.solver.solve(rhs)
        self.solver_iterations += rep.iterations
        if not rep.converged:
            self.solver_failures += 1
        return x

    def step(self, s: State) -> State:
        """libs/core/schemes.py:169-172"""
        u0, v0, e0 = self._dev(s.u), self._dev(s.v), self._dev(s.eta)
        p = _lib.ptr
        st = self._stream()
        u_it, v_it, eta_new = u0, v0, e0
        for _ in range(self.params.n_picard):
            nu, nv = torch.empty_like(u0), torch.empty_like(v0)
            _lib.check(self.lib.hoc_sw2d_momentum(self._h, p(u0), p(v0), p(e0), p(u_it),
                                                  p(v_it), p(nu), p(nv), st))
            if self.params.name == "fb":
                u_it, v_it = nu, nv
                continue
            rhs = torch.empty_like(e0)
            _lib.check(self.lib.hoc_sw2d_rhs(self._h, p(e0), p(u0), p(v0), p(nu), p(nv), p(rhs),
                                             st))
            eta_new = self._solve(rhs)
            u_it, v_it = torch.empty_like(u0), torch.empty_like(v0)
            _lib.check(self.lib.hoc_sw2d_backsub(self._h, p(nu), p(nv), p(eta_new), p(u_it),
                                                 p(v_it), st))
        if self.params.name == "fb":
            eta_new = torch.empty_like(e0)
            _lib.check(self.lib.hoc_sw2d_fb_eta(self._h, p(e0), p(u_it), p(v_it), p(eta_new),
                                                st))
        return State(u=u_it, v=v_it, eta=eta_new, t=s.t + self.dt)

    def integrate(self, state: State, n_steps: int) -> State:
        for _ in range(n_steps):
            state = self.step(state)
        return state
