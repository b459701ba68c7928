"""The reference's v0.2 3D hydrostatic model (libs/core/model3d.py:93-391,
Stepper3D) on the GPU: flat bottom, nz uniform layers, doubly periodic.

`Stepper3D(grid, physics, params, dt)` has the constructor, `step`,
`integrate`, `diagnose_w`, the counters (solver_iterations,
solver_failures, tridiagonal_solves, barotropic_substeps), `q`, `h_eff` and
`n_split` of the reference's class. States are `types.State3D` with device
tensors ([nz, ny, nx]; eta [ny, nx]); NumPy inputs are copied in. The step
runs the reference's sequence with the HIP operators of the C-ABI
(hoc_m3_*): the prelude (buoyancy potential, depth sums, w, explicit
stress), the Picard passes of the predictor with their tridiagonal solves,
the free-surface solve (`model2d.Helmholtz2D`: the spectral solve through
the device FFT, or one of the four iterative kinds bitwise), the
back-substitution or the barotropic substeps, then w and the tracers.
Every kernel keeps the NumPy operation order, so with any solver but fft
the states are bitwise the reference's.
"""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import torch

from . import _lib
from .model2d import Helmholtz2D, explicit_dt_max
from .types import Physics3D, SchemeParams, State3D

_SCHEME = {"theta": 0, "fb": 1, "split_explicit": 2}
_ADVECTION = {"none": 0, "centered2": 1, "upwind1": 2}


class Stepper3D:
    """One-step v0.2 3D integrator for a fixed (grid, physics, params, dt)."""

    def __init__(self, grid, physics: Physics3D, params: SchemeParams, dt: float,
                 device="cuda:0"):
        params.validate()
        if grid.nz < 1 or grid.depth <= 0.0:
            raise ValueError("the 3D model needs grid.nz >= 1 and physics.H > 0")
        if getattr(grid, "bc_x", "periodic") != "periodic" or \
                getattr(grid, "bc_y", "periodic") != "periodic":
            raise NotImplementedError("the v0.2 3D model is doubly periodic")
        if params.advection not in _ADVECTION:
            # advection.py:_face raises this on the first step
            raise ValueError(f"unknown advection scheme '{params.advection}'")
        for name in ("kappa", "tau_x", "tau_y"):
            if np.ndim(getattr(physics, name)) != 0:
                raise ValueError(f"the v0.2 model takes a scalar physics.{name}")
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("Stepper3D needs a HIP device (torch.cuda)")
        self.grid, self.physics, self.params = grid, physics, params
        self.dt = float(dt)
        self.solver_iterations = 0
        self.solver_failures = 0
        self.tridiagonal_solves = 0
        self.barotropic_substeps = 0

        # Barotropic substeps (model3d.py:141-151)
        self.n_split = 0
        if params.name == "split_explicit":
            dt_baro = explicit_dt_max(grid, physics)
            self.n_split = (params.n_split if params.n_split > 0
                            else max(1, int(np.ceil(1.2 * self.dt / dt_baro))))
            if self.dt / self.n_split > dt_baro:
                raise ValueError(
                    f"n_split={self.n_split} leaves a barotropic substep of "
                    f"{self.dt / self.n_split:.1f}s above the stability limit "
                    f"{dt_baro:.1f}s")

        cfg = _lib.HocM3Config()
        cfg.nx, cfg.ny, cfg.nz = grid.nx, grid.ny, grid.nz
        cfg.dx, cfg.dy, cfg.dz, cfg.dt = grid.dx, grid.dy, grid.dz, self.dt
        for name in ("g", "f", "H", "N2", "nu", "kappa", "rho0", "tau_x", "tau_y",
                     "bottom_drag", "alpha_T", "beta_S", "T0", "S0", "cp", "A_h", "K_h",
                     "q_heat", "q_salt"):
            setattr(cfg, name, float(getattr(physics, name)))
        cfg.tracers = 1 if physics.tracers == "TS" else 0
        cfg.scheme = _SCHEME[params.name]
        cfg.theta, cfg.theta_cor, cfg.theta_v = params.theta, params.theta_cor, params.theta_v
        cfg.advection = _ADVECTION[params.advection]
        cfg.n_split = self.n_split
        h = C.c_void_p()
        _lib.check(self.lib.hoc_m3_create(C.byref(cfg), C.byref(h)))
        self._h = h
        _lib.register_handle(self)
        self.tridiagonal_solves += 1

        # q = A^-1 1 and the effective depth (model3d.py:125-134): the column
        # sum accumulates over layers, the mean is NumPy's over the 2D field
        nz, ny, nx = grid.shape3d
        q = np.empty(nz)
        _lib.check(self.lib.hoc_m3_q(h, q.ctypes.data_as(C.c_void_p)))
        self._q1 = q
        col = q[0]
        for k in range(1, nz):
            col = col + q[k]
        h_eff = np.full((ny, nx), grid.dz * col)
        self.h_eff = float(np.mean(h_eff))
        self.nu = np.full((nz + 1, ny, nx), physics.nu, dtype=np.float64)
        self.kappa = np.full((nz + 1, ny, nx), physics.kappa, dtype=np.float64)

        coef = physics.g * self.h_eff * params.theta ** 2 * self.dt ** 2
        self.solver = (Helmholtz2D(params.solver, grid, coef, params.rtol, params.max_iter,
                                   device=self.device)
                       if params.name == "theta" else None)

    @property
    def q(self) -> np.ndarray:
        """The effective-depth profile [nz, ny, nx] (every column alike)."""
        return np.broadcast_to(self._q1[:, None, None], self.grid.shape3d)

    def close(self) -> None:
        """Destroy the library handle. Idempotent; also runs at interpreter exit."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            self.lib.hoc_m3_destroy(h)

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at class creation: module globals are None during teardown
        if not _finalizing():
            self.close()

    # ---------------------------------------------------------------- helpers
    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, a) -> torch.Tensor:
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=torch.float64).contiguous()
        return torch.as_tensor(np.array(a, dtype=np.float64, order="C"), device=self.device)

    def _e3(self):
        return torch.empty(self.grid.shape3d, dtype=torch.float64, device=self.device)

    def _e2(self):
        return torch.empty(self.grid.shape, dtype=torch.float64, device=self.device)

    def _w(self, u, v) -> torch.Tensor:
        nz, ny, nx = self.grid.shape3d
        w = torch.empty((nz + 1, ny, nx), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.hoc_m3_w(self._h, _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
                                     self._stream()))
        return w

    def _solve(self, rhs):
        x, rep = self.solver.solve(rhs)
        self.solver_iterations += rep.iterations
        if not rep.converged:
            self.solver_failures += 1
        return x

    # ------------------------------------------------------------------- step
    def step(self, s: State3D) -> State3D:
        """model3d.py:326-386 (theta, fb) and 249-301 (split_explicit)."""
        p = _lib.ptr
        st = self._stream()
        u0, v0, eta0 = self._dev(s.u), self._dev(s.v), self._dev(s.eta)
        b0 = self._dev(s.b)
        phi, st_u, st_v = self._e3(), self._e3(), self._e3()
        U0, V0 = self._e2(), self._e2()
        nz, ny, nx = self.grid.shape3d
        w_old = torch.empty((nz + 1, ny, nx), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.hoc_m3_prelude(self._h, p(u0), p(v0), p(b0), p(phi), p(U0), p(V0),
                                           p(w_old), p(st_u), p(st_v), st))
        split = self.params.name == "split_explicit"
        cor_x, cor_y = (self._e2(), self._e2()) if split else (None, None)
        u_it, v_it, eta_new = u0, v0, eta0
        gh_u = gh_v = GU = GV = None
        for _ in range(self.params.n_picard):
            gh_u, gh_v, GU, GV = self._e3(), self._e3(), self._e2(), self._e2()
            _lib.check(self.lib.hoc_m3_predict(
                self._h, p(u0), p(v0), p(phi), p(eta0), p(u_it), p(v_it), p(st_u), p(st_v),
                p(gh_u), p(gh_v), p(GU), p(GV), p(cor_x) if split else None,
                p(cor_y) if split else None, st))
            self.tridiagonal_solves += 2
            if self.params.name == "theta":
                rhs = self._e2()
                _lib.check(self.lib.hoc_m3_rhs(self._h, p(eta0), p(U0), p(V0), p(GU), p(GV),
                                               p(rhs), st))
                eta_new = self._solve(rhs)
                u_it, v_it = self._e3(), self._e3()
                _lib.check(self.lib.hoc_m3_backsub(self._h, p(gh_u), p(gh_v), p(eta_new),
                                                   p(u_it), p(v_it), st))
            else:
                u_it, v_it = gh_u, gh_v
        if split:
            u_it, v_it, eta_new = self._split_barotropic(u0, v0, eta0, U0, V0, gh_u, gh_v,
                                                         GU, GV, cor_x, cor_y)
        elif self.params.name == "fb":
            eta_new = self._e2()
            _lib.check(self.lib.hoc_m3_eta_div(self._h, self.dt, p(eta0), p(GU), p(GV),
                                               p(eta_new), st))
        w = self._w(u_it, v_it)
        return self._tracers(s, u_it, v_it, w, eta_new)

    def _split_barotropic(self, u0, v0, eta0, U0, V0, gh_u, gh_v, GU, GV, cor_x, cor_y):
        """Forcing, barotropic substeps and the depth-mean replacement
        (model3d.py:206-247, 290-300)."""
        p = _lib.ptr
        st = self._stream()
        fx, fy = self._e2(), self._e2()
        _lib.check(self.lib.hoc_m3_forcing(self._h, p(GU), p(GV), p(U0), p(V0), p(cor_x),
                                           p(cor_y), p(fx), p(fy), st))
        ddt = self.dt / self.n_split
        u_i, v_i, e = U0, V0, eta0
        mean = self.params.barotropic_average == "mean"
        acc_u = torch.zeros_like(U0) if mean else None
        acc_v = torch.zeros_like(V0) if mean else None
        for _ in range(self.n_split):
            ui, vi = u_i, v_i
            for _ in range(self.params.n_picard):
                nu_, nv_ = self._e2(), self._e2()
                _lib.check(self.lib.hoc_m3_bt_momentum(self._h, p(u_i), p(v_i), p(e), p(ui),
                                                       p(vi), p(fx), p(fy), p(nu_), p(nv_), st))
                ui, vi = nu_, nv_
            u_i, v_i = ui, vi
            e_new = self._e2()
            _lib.check(self.lib.hoc_m3_eta_div(self._h, ddt, p(e), p(u_i), p(v_i), p(e_new), st))
            e = e_new
            if mean:
                _lib.check(self.lib.hoc_m3_bt_acc(self._h, p(u_i), p(v_i), p(acc_u), p(acc_v), st))
        self.barotropic_substeps += self.n_split
        Ub, Vb = (acc_u, acc_v) if mean else (u_i, v_i)
        u_new, v_new = self._e3(), self._e3()
        _lib.check(self.lib.hoc_m3_split_final(self._h, self.n_split if mean else 0, p(gh_u),
                                               p(gh_v), p(Ub), p(Vb), p(u_new), p(v_new), st))
        return u_new, v_new, e

    def _tracers(self, s: State3D, u, v, w, eta_new) -> State3D:
        """model3d.py:303-324"""
        p = _lib.ptr
        st = self._stream()
        t_new = s.t + self.dt
        if self.physics.tracers == "TS":
            T0, S0 = self._dev(s.T), self._dev(s.S)
            T, S, b = self._e3(), self._e3(), self._e3()
            _lib.check(self.lib.hoc_m3_tracer(self._h, 1, p(T0), p(u), p(v), p(w), p(T), st))
            _lib.check(self.lib.hoc_m3_tracer(self._h, 2, p(S0), p(u), p(v), p(w), p(S), st))
            _lib.check(self.lib.hoc_m3_eos(self._h, p(T), p(S), p(b), st))
            self.tridiagonal_solves += 2
            return State3D(u=u, v=v, b=b, eta=eta_new, t=t_new, T=T, S=S)
        b0 = self._dev(s.b)
        b = self._e3()
        _lib.check(self.lib.hoc_m3_tracer(self._h, 0, p(b0), p(u), p(v), p(w), p(b), st))
        self.tridiagonal_solves += 1
        return State3D(u=u, v=v, b=b, eta=eta_new, t=t_new)

    def integrate(self, state: State3D, n_steps: int) -> State3D:
        for _ in range(n_steps):
            state = self.step(state)
        return state

    def diagnose_w(self, state: State3D) -> torch.Tensor:
        """model3d.py:389-391"""
        return self._w(self._dev(state.u), self._dev(state.v))
