"""HipStepper: the MI355X drop-in for the reference's `Stepper3DV05`.

Same constructor, `step`, `integrate`, `diagnose_w`, `buoyancy_of` and
counters as libs/core/model3d_v05.py:78-545; the step runs as HIP kernels
through libhoc.so. Arrays come back as torch tensors on the device (call
`.cpu().numpy()` for NumPy). The stepper keeps the eddy viscosity and
diffusivity between steps as `nu` and `kappa` ([nz+1, ny, nx] device
tensors), as the reference does (model3d_v05.py:97-98,179).

Supported: zlevel|zstar|sigma coordinates, periodic, closed and open
(Flather) boundaries, the theta (pcg_jacobi|pcg_rbgs|rbgs|multigrid|pcg_mg),
fb and split_explicit schemes, closure none|tke (integral|recursive), every
advection scheme, linear|seos|teos10, tracers TS|buoyancy. Other settings
raise an error, never run a fallback.
"""

from __future__ import annotations

import ctypes as C
import sys
from dataclasses import replace

import numpy as np
import torch

from . import _lib
from .types import State3D

ADVECTION = {"none": 0, "centered2": 1, "upwind1": 2, "up3": 3, "up3_tvd": 4}
EOS = {"linear": 0, "seos": 1, "teos10": 2}
SOLVER = {"pcg_jacobi": 0, "pcg_rbgs": 1, "rbgs": 2, "multigrid": 3, "pcg_mg": 4}
FLOPS_PER_CELL = {"linear": 4, "seos": 15, "teos10": 71}


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a) > 0, dtype=np.uint8)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _field2d(value, shape) -> np.ndarray:
    return _f64(np.broadcast_to(np.asarray(value, dtype=np.float64), shape))


PRECISION = {"fp64": 0, "mixed": 1, "mixed_fp32_momentum": 2}


class HipStepper:
    """One-step integrator over a Domain, on the GPU."""

    def __init__(self, domain, physics, params, dt, device: str | torch.device = "cuda:0", *,
                 ustar2=None, precision: str = "fp64"):
        """`ustar2` optionally replaces the surface friction-velocity field
        u*^2 [ny, nx] (default hypot(tau_x, tau_y) / rho0, as the reference).
        `precision`: "fp64" is the plain step (the reference's arithmetic);
        "mixed" runs the pcg_mg V-cycle preconditioner, the tracer advective
        fluxes and the equation of state in fp32 and keeps the momentum
        fluxes, the state updates, the column solves and the outer PCG
        (residual, dot products, update) in fp64; "mixed_fp32_momentum" also
        runs the momentum advective fluxes in fp32 (on MI355X that kernel's
        fp32 form is slower than its fp64 form, and the step's error is the
        same; kept for comparison) -- theta scheme with pcg_mg only."""
        params.validate()
        if precision not in PRECISION:
            raise ValueError(f"unknown precision {precision!r} "
                             "(expected fp64|mixed|mixed_fp32_momentum)")
        if precision != "fp64" and params.name == "theta" and params.solver != "pcg_mg":
            raise ValueError("mixed precision runs the theta scheme with pcg_mg only")
        if (getattr(params, "mg_smoother", "jacobi") == "rbgs" and params.solver == "pcg_mg"
                and (precision != "fp64" or int(getattr(params, "mg_sweeps", 2)) > 3)):
            raise ValueError("the rbgs pcg_mg smoother runs in fp64 with mg_sweeps <= 3")
        self.precision = precision
        if params.name == "theta" and params.solver == "fft" and not domain.is_flat_periodic:
            raise ValueError(
                "solver 'fft' needs a flat bottom and periodic boundaries; "
                "with topography the Helmholtz operator is variable-coefficient "
                "(spec S10.5). Use pcg_jacobi | pcg_rbgs | rbgs | multigrid.")
        if params.name == "theta" and params.solver not in SOLVER:
            raise ValueError(f"unknown variable-coefficient solver '{params.solver}' "
                             "(expected pcg_jacobi|pcg_rbgs|rbgs|multigrid|pcg_mg)")
        if physics.eos not in EOS:
            raise ValueError(f"unknown eos '{physics.eos}' (expected linear|seos|teos10)")
        if physics.closure == "tke" and physics.mxl not in ("integral", "recursive"):
            raise ValueError(f"unknown mixing length form {physics.mxl!r} (integral|recursive)")
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("HipStepper needs a HIP device (torch.cuda)")
        self.domain, self.grid = domain, domain.grid
        self.physics, self.params = physics, params
        self.dt = float(dt)
        self.n_split = 0
        if params.name == "split_explicit":
            # libs/core/model3d_v05.py:111-123: the deepest column sets the
            # fastest surface gravity wave
            h_max = float(np.max(domain.H[domain.mask > 0]))
            c = np.sqrt(physics.g * h_max)
            gr = domain.grid
            dt_baro = 1.0 / (c * np.sqrt(1.0 / gr.dx ** 2 + 1.0 / gr.dy ** 2))
            self.n_split = (params.n_split if params.n_split > 0
                            else max(1, int(np.ceil(1.2 * self.dt / dt_baro))))
            if self.dt / self.n_split > dt_baro:
                raise ValueError(
                    f"n_split={self.n_split} leaves a barotropic substep of "
                    f"{self.dt / self.n_split:.1f}s above the limit {dt_baro:.1f}s")
        self._eos_host = 0
        g = self.grid
        nz, ny, nx = g.nz, g.ny, g.nx
        shape2 = (ny, nx)

        p = physics
        cfg = _lib.HocConfig()
        cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
        cfg.dx, cfg.dy, cfg.dz_nominal, cfg.dt = g.dx, g.dy, g.dz, self.dt
        cfg.scheme = {"theta": 0, "fb": 1, "split_explicit": 2}[params.name]
        cfg.n_split = self.n_split
        cfg.barotropic_mean = 1 if getattr(params, "barotropic_average", "none") == "mean" else 0
        cfg.vcoord = {"zlevel": 0, "zstar": 1, "sigma": 2}[domain.vcoord]
        cfg.face_rule = {"min": 0, "mean": 1, "harmonic": 2}[domain.face_rule]
        cfg.open_x = 1 if domain.bc_x == "open" else 0
        cfg.open_y = 1 if domain.bc_y == "open" else 0
        cfg.eta_ext = float(getattr(physics, "eta_ext", 0.0))
        cfg.u_ext = float(getattr(physics, "u_ext", 0.0))
        cfg.theta, cfg.theta_cor, cfg.theta_v = params.theta, params.theta_cor, params.theta_v
        cfg.n_picard = params.n_picard
        cfg.advection = ADVECTION[params.advection]
        cfg.solver = SOLVER.get(params.solver, 0)
        cfg.rtol, cfg.max_iter = params.rtol, params.max_iter
        cfg.mg_levels = int(getattr(params, "mg_levels", 5))
        cfg.mg_sweeps = int(getattr(params, "mg_sweeps", 2))
        cfg.mg_omega = float(getattr(params, "mg_omega", 0.8))
        cfg.mg_smoother = {"jacobi": 0, "rbgs": 1}[getattr(params, "mg_smoother", "jacobi")]
        cfg.precision = PRECISION[precision]
        cfg.g, cfg.f, cfg.rho0 = p.g, p.f, p.rho0
        cfg.nu_b, cfg.bottom_drag, cfg.N2 = p.nu, p.bottom_drag, p.N2
        cfg.tracers_ts = 1 if p.tracers == "TS" else 0
        cfg.alpha_T, cfg.beta_S, cfg.T0, cfg.S0, cfg.cp = p.alpha_T, p.beta_S, p.T0, p.S0, p.cp
        cfg.A_h, cfg.K_h, cfg.q_heat, cfg.q_salt = p.A_h, p.K_h, p.q_heat, p.q_salt
        cfg.eos = EOS[p.eos]
        cfg.pgf_remove = 1 if p.pgf == "remove" else 0
        cfg.pgf_correction = 1 if p.pgf_correction else 0
        cfg.closure_tke = 1 if p.closure == "tke" else 0
        cfg.mxl_recursive = 1 if p.mxl == "recursive" else 0
        for name in ("c_k", "c_eps", "pr_t", "kappa_vk", "z_0", "e_min", "n2_min", "l_min",
                     "e_bb"):
            setattr(cfg, name, float(getattr(p, name)))

        d = domain
        dz3 = d.dz3_ref if hasattr(d, "dz3_ref") else d.layer_thickness(np.zeros(shape2))
        dzu, dzv = d.face_thickness(dz3)
        self._host = dict(
            mask=_u8(d.mask), mask_u=_u8(d.mask_u), mask_v=_u8(d.mask_v),
            mask3=_u8(d.mask3), mask3u=_u8(d.mask3u), mask3v=_u8(d.mask3v),
            dz3=_f64(dz3), dz3u=_f64(dzu), dz3v=_f64(dzv), z_centre=_f64(d.depth_at_centres()),
            tau_x=_field2d(p.tau_x, shape2), tau_y=_field2d(p.tau_y, shape2),
            ustar2=_field2d(np.hypot(p.tau_x, p.tau_y) / p.rho0 if ustar2 is None else ustar2,
                            shape2),
            kappa_b=_field2d(p.kappa, shape2), H=_f64(d.H))
        geo = _lib.HocGeometry()
        for name, arr in self._host.items():
            ctype = _lib.P_u8 if arr.dtype == np.uint8 else _lib.P_f64
            setattr(geo, name, arr.ctypes.data_as(ctype))
        handle = C.c_void_p()
        _lib.check(self.lib.hoc_stepper_create(C.byref(cfg), C.byref(geo), C.byref(handle)))
        self._h = handle
        _lib.register_handle(self)
        self.cfg = cfg

        opts = dict(dtype=torch.float64, device=self.device)
        self.nu = torch.full((nz + 1, ny, nx), float(p.nu), **opts)
        self.kappa = torch.as_tensor(
            np.ascontiguousarray(np.broadcast_to(self._host["kappa_b"], (nz + 1, ny, nx))),
            **opts)
        self.z_centre = torch.as_tensor(self._host["z_centre"], **opts)
        _lib.check(self.lib.hoc_stepper_bind_mixing(self._h, _lib.ptr(self.nu),
                                                    _lib.ptr(self.kappa)))
        _lib.check(self.lib.hoc_stepper_build_coefficients(self._h, self._stream()))
        self._bufs: list[dict] = []

    def reset_mixing(self) -> None:
        """Restore nu / kappa to their backgrounds and rebuild the implicit
        coefficients, as at construction (a run restarting from its initial
        state; cfd_exp3d5_cuda.cu:1163-1175)."""
        p = self.physics
        self.nu.fill_(float(p.nu))
        self.kappa.copy_(torch.as_tensor(
            np.ascontiguousarray(np.broadcast_to(self._host["kappa_b"], self.kappa.shape)),
            device=self.device))
        _lib.check(self.lib.hoc_stepper_build_coefficients(self._h, self._stream()))

    def set_fp32_parts(self, parts: int) -> None:
        """Select the parts of a mixed step that run in fp32 (bits: 1 the
        pcg_mg V-cycle, 2 the tracer advective fluxes, 4 the EOS, 8 the
        momentum advective fluxes); needs precision "mixed" or
        "mixed_fp32_momentum", whose defaults are 7 and 15."""
        _lib.check(self.lib.hoc_stepper_set_fp32_parts(self._h, int(parts)))

    def geometry_form(self, column: int | None = None) -> int:
        """1 when the kernels form the static z-level geometry from per-column
        wet-level counts, 0 when they read the stored arrays (same results).
        column = 1 or 0 selects a form (1 raises NotImplementedError where the
        column form does not exist: moving coordinates); None only queries."""
        active = C.c_int32(-1)
        _lib.check(self.lib.hoc_stepper_geometry_form(
            self._h, -1 if column is None else int(column), C.byref(active)))
        return int(active.value)

    # ------------------------------------------------------------ plumbing
    def close(self) -> None:
        """Destroy the library handle (its workspaces and multigrid levels)
        after the device has finished queued work. Idempotent; the stepper
        is unusable afterwards. Also runs at interpreter exit."""
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._h = None
            self.lib.hoc_stepper_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the exit hook has already closed live
        # handles; never call into the runtime from the finaliser then.
        # is_finalizing is bound at class creation because module teardown
        # sets this module's globals (sys among them) to None first.
        if not _finalizing():
            self.close()

    def _stream(self) -> C.c_void_p:
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _dev(self, a):
        if a is None:
            return None
        if isinstance(a, torch.Tensor):
            return a.to(device=self.device, dtype=torch.float64).contiguous()
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.device)

    def _new_state_buffers(self) -> dict:
        g = self.grid
        opts = dict(dtype=torch.float64, device=self.device)
        out = {k: torch.empty(g.shape3d, **opts) for k in ("u", "v", "b")}
        out["eta"] = torch.empty(g.shape, **opts)
        if self.physics.tracers == "TS":
            out["T"] = torch.empty(g.shape3d, **opts)
            out["S"] = torch.empty(g.shape3d, **opts)
        if self.physics.closure == "tke":
            out["e"] = torch.empty((g.nz + 1,) + g.shape, **opts)
        return out

    @staticmethod
    def _hoc_state(fields: dict) -> _lib.HocState:
        st = _lib.HocState()
        for name in ("u", "v", "b", "eta", "T", "S", "e"):
            t = fields.get(name)
            setattr(st, name, None if t is None else t.data_ptr())
        return st

    def _inputs(self, s) -> dict:
        g = self.grid
        f = {k: self._dev(getattr(s, k)) for k in ("u", "v", "b", "eta", "T", "S", "e")}
        if self.physics.closure == "tke" and f["e"] is None:
            f["e"] = torch.full((g.nz + 1,) + g.shape, float(self.physics.e_min),
                                dtype=torch.float64, device=self.device)
        if self.physics.tracers == "TS" and (f["T"] is None or f["S"] is None):
            raise ValueError("tracers='TS' needs T and S in the state")
        return f

    def _step_into(self, fin: dict, fout: dict) -> None:
        sin, sout = self._hoc_state(fin), self._hoc_state(fout)
        _lib.check(self.lib.hoc_stepper_step(self._h, C.byref(sin), C.byref(sout),
                                             self._stream()))
        if self.physics.closure != "tke":
            fout["e"] = fin.get("e")

    def _new_tape(self) -> dict:
        """Device arrays of one step's tape (hoc_step_tape)."""
        g = self.grid
        opts = dict(dtype=torch.float64, device=self.device)
        shape_i = (g.nz + 1,) + g.shape
        n = self.params.n_picard
        return {"q": torch.empty(g.shape3d, **opts), "Ku": torch.empty(g.shape, **opts),
                "Kv": torch.empty(g.shape, **opts), "w_old": torch.empty(shape_i, **opts),
                "w": torch.empty(shape_i, **opts),
                "gh_u": [torch.empty(g.shape3d, **opts) for _ in range(n)],
                "gh_v": [torch.empty(g.shape3d, **opts) for _ in range(n)],
                "eta": [torch.empty(g.shape, **opts) for _ in range(n)]}

    def _step_tape_into(self, fin: dict, fout: dict, tape: dict) -> None:
        """One step fin -> fout (hoc_stepper_step_tape) writing `tape`."""
        if self.params.n_picard > _lib.TAPE_MAX_PICARD:
            raise ValueError(f"the step tape holds at most {_lib.TAPE_MAX_PICARD} Picard passes")
        t = _lib.HocStepTape()
        for k in ("q", "Ku", "Kv", "w_old", "w"):
            setattr(t, k, tape[k].data_ptr())
        for k in ("gh_u", "gh_v", "eta"):
            arr = getattr(t, k)
            for i, buf in enumerate(tape[k]):
                arr[i] = buf.data_ptr()
        sin, sout = self._hoc_state(fin), self._hoc_state(fout)
        _lib.check(self.lib.hoc_stepper_step_tape(self._h, C.byref(sin), C.byref(sout),
                                                  C.byref(t), self._stream()))
        if self.physics.closure != "tke":
            fout["e"] = fin.get("e")

    def _step_n_into(self, a: dict, b: dict, n: int) -> dict:
        """n steps from a, ping-ponging between the buffer sets a and b in one
        library call (hoc_stepper_step_n: from the second step on, w of the
        input is the one the previous step formed); returns the set holding
        the final state."""
        sa, sb = self._hoc_state(a), self._hoc_state(b)
        _lib.check(self.lib.hoc_stepper_step_n(self._h, int(n), C.byref(sa), C.byref(sb),
                                               self._stream()))
        return a if n % 2 == 0 else b

    # ---------------------------------------------------------------- API
    def step(self, s) -> State3D:
        """libs/core/model3d_v05.py:337-415"""
        fin = self._inputs(s)
        fout = self._new_state_buffers()
        self._step_into(fin, fout)
        return State3D(u=fout["u"], v=fout["v"], b=fout["b"], eta=fout["eta"],
                       t=s.t + self.dt, T=fout.get("T"), S=fout.get("S"), e=fout.get("e"))

    def integrate(self, state, n_steps: int) -> State3D:
        """libs/core/model3d_v05.py:535-538: the input copied into the first of
        two buffer sets, then hoc_stepper_step_n ping-ponging between them."""
        if n_steps <= 0:
            return state
        fin = self._inputs(state)
        if not self._bufs:
            self._bufs = [self._new_state_buffers(), self._new_state_buffers()]
        a, b = self._bufs
        for k, buf in a.items():
            buf.copy_(fin[k])
        cur = dict(self._step_n_into(a, b, n_steps))
        if self.physics.closure != "tke":
            cur["e"] = fin.get("e")
        t = state.t
        for _ in range(n_steps):
            t = t + self.dt
        res = {k: (v.clone() if isinstance(v, torch.Tensor) else v) for k, v in cur.items()}
        return State3D(u=res["u"], v=res["v"], b=res["b"], eta=res["eta"], t=t,
                       T=res.get("T"), S=res.get("S"), e=res.get("e"))

    def buoyancy_of(self, T, S):
        """libs/core/model3d_v05.py:189-193 (EOS on the device)."""
        p = self.physics
        Td, Sd = self._dev(T), self._dev(S)
        b = torch.empty_like(Td)
        if self.moving_coordinate:   # centre depths follow the last step's eta
            _lib.check(self.lib.hoc_stepper_z_centre(self._h, _lib.ptr(self.z_centre),
                                                     self._stream()))
        _lib.check(self.lib.hoc_eos_buoyancy(EOS[p.eos], p.g, p.rho0, p.alpha_T, p.beta_S,
                                             p.T0, p.S0, _lib.ptr(Td), _lib.ptr(Sd),
                                             _lib.ptr(self.z_centre), _lib.ptr(b), Td.numel(),
                                             self._stream()))
        self._eos_host += Td.numel()
        return b

    def diagnose_w(self, state):
        """libs/core/model3d_v05.py:540-541"""
        u, v = self._dev(state.u), self._dev(state.v)
        g = self.grid
        w = torch.empty((g.nz + 1,) + g.shape, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.hoc_stepper_w_diag(self._h, _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
                                               self._stream()))
        return w

    # ------------------------------------------------------------ counters
    def _counters(self) -> _lib.HocCounters:
        c = _lib.HocCounters()
        _lib.check(self.lib.hoc_stepper_counters(self._h, C.byref(c)))
        return c

    @property
    def solver_iterations(self) -> int:
        return int(self._counters().solver_iterations)

    @property
    def solver_failures(self) -> int:
        return int(self._counters().solver_failures)

    @property
    def solver_rebuilds(self) -> int:
        return int(self._counters().solver_rebuilds)

    @property
    def tridiagonal_solves(self) -> int:
        return int(self._counters().tridiagonal_solves)

    @property
    def barotropic_substeps(self) -> int:
        return int(self._counters().barotropic_substeps)

    @property
    def eos_cells(self) -> int:
        return int(self._counters().eos_cells) + self._eos_host

    @property
    def eos_flops(self) -> int:
        return self.eos_cells * FLOPS_PER_CELL[self.physics.eos]

    @property
    def moving_coordinate(self) -> bool:
        return self.domain.vcoord != "zlevel"

    def counters(self) -> dict:
        c = self._counters()
        return {k: int(getattr(c, k)) for k, _ in _lib.HocCounters._fields_} | {
            "eos_cells": int(c.eos_cells) + self._eos_host}

    @property
    def solver_path(self) -> str:
        """'persistent' (one cooperative launch per solve) or 'multikernel'."""
        return "persistent" if self.lib.hoc_stepper_solver_path(self._h) == 0 else "multikernel"

    @solver_path.setter
    def solver_path(self, path: str) -> None:
        _lib.check(self.lib.hoc_stepper_set_solver_path(self._h, 0 if path == "persistent"
                                                        else 1))

    # --------------------------------------------------------- measurement
    def profile_kernel(self, name: str | None) -> None:
        """Time every launch of one step kernel with HIP events on its stream."""
        _lib.check(self.lib.hoc_stepper_profile_kernel(self._h, (name or "").encode()))

    def profile_read(self) -> tuple[int, float]:
        """(launches, total ms) of the timed kernel since the last read."""
        n, ms = C.c_int64(), C.c_double()
        _lib.check(self.lib.hoc_stepper_profile_read(self._h, C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    # ----------------------------------------------------- single operators
    def tke_step(self, e, u, v, b):
        """closure.tke_step with this stepper's nu/kappa as K_m_old/K_h_old;
        updates them in place and returns e_new (libs/core/closure.py:141-205)."""
        e, u, v, b = (self._dev(x) for x in (e, u, v, b))
        out = torch.empty_like(e)
        _lib.check(self.lib.hoc_stepper_tke_step(self._h, _lib.ptr(e), _lib.ptr(u), _lib.ptr(v),
                                                 _lib.ptr(b), _lib.ptr(out), self._stream()))
        return out

    def momentum_advection(self, u, v, w):
        u, v, w = (self._dev(x) for x in (u, v, w))
        au, av = torch.empty_like(u), torch.empty_like(v)
        _lib.check(self.lib.hoc_stepper_momentum_advection(
            self._h, _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), _lib.ptr(au), _lib.ptr(av),
            self._stream()))
        return au, av

    def tracer_advection(self, c, u, v, w):
        c, u, v, w = (self._dev(x) for x in (c, u, v, w))
        out = torch.empty_like(c)
        _lib.check(self.lib.hoc_stepper_tracer_advection(
            self._h, _lib.ptr(c), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w), _lib.ptr(out),
            self._stream()))
        return out

    def pressure_gradients(self, b):
        b = self._dev(b)
        px, py = torch.empty_like(b), torch.empty_like(b)
        _lib.check(self.lib.hoc_stepper_pressure_gradients(self._h, _lib.ptr(b), _lib.ptr(px),
                                                           _lib.ptr(py), self._stream()))
        return px, py

    def helmholtz_solve(self, Ku, Kv, rhs):
        """HelmholtzVar(kind=params.solver).solve on the given face coefficients."""
        Ku, Kv, rhs = (self._dev(x) for x in (Ku, Kv, rhs))
        x = torch.empty_like(rhs)
        it, cv = C.c_int32(), C.c_int32()
        _lib.check(self.lib.hoc_stepper_helmholtz_solve(
            self._h, _lib.ptr(Ku), _lib.ptr(Kv), _lib.ptr(rhs), _lib.ptr(x), C.byref(it),
            C.byref(cv), self._stream()))
        return x, int(it.value), bool(cv.value)


def to_numpy(state: State3D) -> State3D:
    """Copy a device state to host NumPy arrays."""
    def cv(a):
        return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return replace(state, u=cv(state.u), v=cv(state.v), b=cv(state.b), eta=cv(state.eta),
                   T=cv(state.T), S=cv(state.S), e=cv(state.e))


def thomas_batched(sub, diag, sup, rhs):
    """vertical.thomas_batched on the device for [n, ...] torch tensors."""
    lib = _lib.load()
    n = diag.shape[0]
    ncol = diag.numel() // n
    args = [t.contiguous() for t in (sub, diag, sup, rhs)]
    x = torch.empty_like(args[3])
    scratch = torch.empty_like(args[3])
    stream = C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(lib.hoc_thomas_batched(*(_lib.ptr(a) for a in args), _lib.ptr(x),
                                      _lib.ptr(scratch), n, ncol, stream))
    return x


def pairwise_dot(a, b):
    """numpy.sum(a * b) for float64 device tensors, bit for bit."""
    lib = _lib.load()
    a, b = a.contiguous(), b.contiguous()
    n = a.numel()
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    scratch = torch.empty(max(1, (n + 8191) // 8192), dtype=torch.float64, device=a.device)
    stream = C.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)
    _lib.check(lib.hoc_pairwise_dot(_lib.ptr(a), _lib.ptr(b), n, _lib.ptr(out), _lib.ptr(scratch),
                                    stream))
    return float(out.item())


def eos_buoyancy(kind, T, S, z, *, g, rho0, alpha_T, beta_S, T0, S0):
    """eos.buoyancy on the device."""
    lib = _lib.load()
    T, S, z = (t.contiguous() for t in (T, S, z))
    b = torch.empty_like(T)
    stream = C.c_void_p(torch.cuda.current_stream(T.device).cuda_stream)
    _lib.check(lib.hoc_eos_buoyancy(EOS[kind], g, rho0, alpha_T, beta_S, T0, S0, _lib.ptr(T),
                                    _lib.ptr(S), _lib.ptr(z), _lib.ptr(b), T.numel(), stream))
    return b
