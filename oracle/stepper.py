"""The oracle's time step (libs/core/model3d_v05.py:78-545, z-level only).

`OracleStepper(domain, physics, params, dt).step(state)` returns a new
state exactly as `Stepper3DV05.step` does, for the theta, fb and
split_explicit schemes, closure none|tke, tracers buoyancy|TS, every
advection scheme and every EOS, periodic, closed or Flather open
boundaries, on z-level or moving (z*, sigma) vertical coordinates.

Inputs are duck-typed: `domain` needs grid (nx, ny, nz, dx, dy, dz), mask,
mask_u, mask_v, mask3, mask3u, mask3v, dz3_ref, face_thickness(),
depth_at_centres(), vcoord, bc_x, bc_y; `physics` and `params` need the
reference attribute names. Array-valued tau_x/tau_y and kappa ([ny, nx])
are accepted and reduce to the reference when uniform.
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from . import column as col
from . import horizontal as hz
from .closure import TKEParams, tke_step
from .eos import buoyancy
from .helmholtz import HelmholtzVar

FLOPS_PER_CELL = {"linear": 4, "seos": 15, "teos10": 71}


@dataclass(frozen=True)
class OState:
    u: np.ndarray
    v: np.ndarray
    b: np.ndarray
    eta: np.ndarray
    t: float = 0.0
    T: np.ndarray | None = None
    S: np.ndarray | None = None
    e: np.ndarray | None = None


class OracleStepper:
    def __init__(self, domain, physics, params, dt):
        params.validate()
        self.domain, self.grid = domain, domain.grid
        self.physics, self.params = physics, params
        self.dt = float(dt)
        self.solver_iterations = 0
        self.solver_failures = 0
        self.solver_rebuilds = 0
        self.tridiagonal_solves = 0
        self.barotropic_substeps = 0
        self.eos_cells = 0
        self.n_split = 0
        g = self.grid
        shape3i = (g.nz + 1, g.ny, g.nx)
        self.nu = np.full(shape3i, physics.nu)
        self.kappa = np.broadcast_to(np.asarray(physics.kappa, dtype=np.float64),
                                     shape3i).copy()
        self._build_geometry(np.zeros(self.grid.shape))
        if params.name == "split_explicit":
            # libs/core/model3d_v05.py:111-123: the deepest column sets the
            # barotropic CFL
            h_max = float(np.max(domain.H[domain.mask > 0]))
            c = np.sqrt(physics.g * h_max)
            dt_baro = 1.0 / (c * np.sqrt(1.0 / self.grid.dx ** 2 + 1.0 / self.grid.dy ** 2))
            self.n_split = (params.n_split if params.n_split > 0
                            else max(1, int(np.ceil(1.2 * self.dt / dt_baro))))
            if self.dt / self.n_split > dt_baro:
                raise ValueError(
                    f"n_split={self.n_split} leaves a barotropic substep of "
                    f"{self.dt / self.n_split:.1f}s above the limit {dt_baro:.1f}s")

    @property
    def moving_coordinate(self):
        return self.domain.vcoord != "zlevel"

    def _build_geometry(self, eta):
        """Layer and face thickness, centre depths and the implicit/Helmholtz
        coefficients for the free surface eta (libs/core/model3d_v05.py:128-139);
        once for z-level, every step for z* and sigma."""
        d = self.domain
        self.dz3 = d.layer_thickness(eta)
        self.dz3u, self.dz3v = d.face_thickness(self.dz3)
        self.z_centre = d.depth_at_centres(eta)
        self.solver = None
        self._build_coefficients()

    # ---------------------------------------------------------- coefficients
    def _build_coefficients(self):
        """libs/core/model3d_v05.py:141-167"""
        d, p, thv = self.domain, self.physics, self.params.theta_v
        self.tri_m = col.diffusion_coeffs(self.nu, self.dz3, self.dt, thv, p.bottom_drag, d.mask3)
        self.tri_b = col.diffusion_coeffs(self.kappa, self.dz3, self.dt, thv, 0.0, d.mask3)
        self.q = col.thomas(*self.tri_m, np.ones(self.grid.shape3d))
        self.tridiagonal_solves += 1
        qu = np.minimum(self.q, np.roll(self.q, -1, axis=-1))
        qv = np.minimum(self.q, np.roll(self.q, -1, axis=-2))
        self.Ku = np.sum(self.dz3u * qu, axis=0) * d.mask_u
        self.Kv = np.sum(self.dz3v * qv, axis=0) * d.mask_v
        if self.params.name == "theta":
            if self.solver is None:
                g = self.grid
                coef = p.g * self.params.theta ** 2 * self.dt ** 2
                pr = self.params
                self.solver = HelmholtzVar(
                    g.nx, g.ny, g.dx, g.dy, coef, self.Ku, self.Kv, d.mask, kind=pr.solver,
                    rtol=pr.rtol, max_iter=pr.max_iter,
                    mg_levels=getattr(pr, "mg_levels", 5), mg_sweeps=getattr(pr, "mg_sweeps", 2),
                    mg_omega=getattr(pr, "mg_omega", 0.8),
                    mg_smoother=getattr(pr, "mg_smoother", "jacobi"))
            else:
                self.solver.update(self.Ku, self.Kv)
                self.solver_rebuilds += 1

    def _ustar2(self):
        p = self.physics
        return np.broadcast_to(np.hypot(p.tau_x, p.tau_y) / p.rho0, self.grid.shape).copy()

    def _closure_step(self, s):
        """libs/core/model3d_v05.py:169-182"""
        p, d = self.physics, self.domain
        tke = TKEParams(p.c_k, p.c_eps, p.pr_t, p.kappa_vk, p.z_0, p.e_min, p.n2_min,
                        p.l_min, p.e_bb, p.mxl)
        e = s.e if s.e is not None else np.full((self.grid.nz + 1,) + self.grid.shape, p.e_min)
        e_new, K_m, K_h = tke_step(e, s.u, s.v, s.b, self.dz3, d.mask3, self._ustar2(), self.dt,
                                   tke, p.nu, p.kappa, self.nu, self.kappa)
        self.nu, self.kappa = K_m, K_h
        self.tridiagonal_solves += 1
        self._build_coefficients()
        return e_new

    # ---------------------------------------------------------------- physics
    def buoyancy_of(self, T, S):
        p = self.physics
        self.eos_cells += T.size
        return buoyancy(p.eos, T, S, self.z_centre, g=p.g, rho0=p.rho0, alpha_T=p.alpha_T,
                        beta_S=p.beta_S, T0=p.T0, S0=p.S0)

    def pressure_gradients(self, b):
        """libs/core/model3d_v05.py:195-230"""
        d, g = self.domain, self.grid
        phi = col.buoyancy_potential(b, self.dz3)
        if self.physics.pgf == "remove":
            phi = col.remove_depth_mean(phi, self.dz3)
        if not self.physics.pgf_correction:
            return (hz.gradx(phi, g.dx, d.mask3u) * d.mask3u,
                    hz.grady(phi, g.dy, d.mask3v) * d.mask3v)
        zc = self.z_centre
        out = []
        for axis, fmask, delta in ((-1, d.mask3u, g.dx), (-2, d.mask3v, g.dy)):
            phi_r = np.roll(phi, -1, axis=axis)
            b_r = np.roll(b, -1, axis=axis)
            z_r = np.roll(zc, -1, axis=axis)
            z_face = 0.5 * (zc + z_r)
            left = phi + b * (z_face - zc)
            right = phi_r + b_r * (z_face - z_r)
            out.append(fmask * (right - left) / delta)
        return out[0], out[1]

    def w_diag(self, u, v):
        """libs/core/model3d_v05.py:232-235"""
        g = self.grid
        return col.w_from_transport_divergence(
            hz.div(self.dz3u * u, self.dz3v * v, g.dx, g.dy, self.domain.mask3))

    def momentum_extras(self, u, v, w):
        """libs/core/model3d_v05.py:241-260"""
        d, g, adv = self.domain, self.grid, self.params.advection
        ex_u = np.zeros_like(u)
        ex_v = np.zeros_like(v)
        if adv in ("up3", "up3_tvd"):
            au, av = hz.momentum_advection_up3(u, v, w, g.dx, g.dy, g.dz, d.mask3u, d.mask3v,
                                               adv == "up3_tvd")
            ex_u -= au * d.mask3u
            ex_v -= av * d.mask3v
        elif adv != "none":
            au, av = hz.momentum_advection_centred(u, v, w, g.dx, g.dy, g.dz)
            ex_u -= au * d.mask3u
            ex_v -= av * d.mask3v
        if self.physics.A_h > 0.0:
            ex_u += self.physics.A_h * hz.laplacian(u, g.dx, g.dy, d.mask3u, d.mask3v, d.mask3u)
            ex_v += self.physics.A_h * hz.laplacian(v, g.dx, g.dy, d.mask3u, d.mask3v, d.mask3v)
        return ex_u, ex_v

    def _flather(self, eta):
        """Flather radiation transports on the open edges
        (libs/core/model3d_v05.py:262-288)."""
        d, p = self.domain, self.physics
        bnd = {}
        if d.bc_x == "open":
            Hw = np.maximum(d.H[:, 0], 1.0)
            He = np.maximum(d.H[:, -1], 1.0)
            bnd["west"] = p.u_ext - np.sqrt(p.g * Hw) * (eta[:, 0] - p.eta_ext)
            bnd["east"] = p.u_ext + np.sqrt(p.g * He) * (eta[:, -1] - p.eta_ext)
        if d.bc_y == "open":
            Hs = np.maximum(d.H[0, :], 1.0)
            Hn = np.maximum(d.H[-1, :], 1.0)
            bnd["south"] = p.u_ext - np.sqrt(p.g * Hs) * (eta[0, :] - p.eta_ext)
            bnd["north"] = p.u_ext + np.sqrt(p.g * Hn) * (eta[-1, :] - p.eta_ext)
        return bnd

    def _div_bt(self, U, V, eta):
        """Barotropic divergence with the open-edge fluxes
        (libs/core/model3d_v05.py:290-309)."""
        d, g = self.domain, self.grid
        div = hz.div(U, V, g.dx, g.dy, d.mask)
        if d.bc_x != "open" and d.bc_y != "open":
            return div
        b = self._flather(eta)
        div = div.copy()
        if "west" in b:
            div[:, 0] += (U[:, 0] - b["west"]) / g.dx - ((U[:, 0] - U[:, -1]) / g.dx)
            div[:, -1] += (b["east"] - U[:, -2]) / g.dx - ((U[:, -1] - U[:, -2]) / g.dx)
        if "south" in b:
            div[0, :] += (V[0, :] - b["south"]) / g.dy - ((V[0, :] - V[-1, :]) / g.dy)
            div[-1, :] += (b["north"] - V[-2, :]) / g.dy - ((V[-1, :] - V[-2, :]) / g.dy)
        return div * d.mask

    def _solve_column(self, coeffs, rhs):
        self.tridiagonal_solves += 1
        return col.thomas(*coeffs, rhs)

    # -------------------------------------------------------------- stepping
    def step(self, s):
        """libs/core/model3d_v05.py:337-415 (theta and fb branches)."""
        d, p, g = self.domain, self.physics, self.grid
        pr = self.params
        dt, th, tc, thv = self.dt, pr.theta, pr.theta_cor, pr.theta_v
        if pr.name != "theta":
            th = 0.0
        if self.moving_coordinate:
            self._build_geometry(s.eta)
        e_new = self._closure_step(s) if p.closure == "tke" else s.e
        self._e_new = e_new
        mask3 = d.mask3

        pgx, pgy = self.pressure_gradients(s.b)
        av_prev_v = hz.avg_v_to_u(s.v, d.mask3v, d.mask3u)
        av_prev_u = hz.avg_u_to_v(s.u, d.mask3u, d.mask3v)
        tau_u = np.broadcast_to(np.asarray(p.tau_x) / p.rho0, g.shape) * d.mask3u[0]
        tau_v = np.broadcast_to(np.asarray(p.tau_y) / p.rho0, g.shape) * d.mask3v[0]
        dzu_diff = col.apply_diffusion(s.u, self.nu, self.dz3u, tau_u, p.bottom_drag, mask3)
        dzv_diff = col.apply_diffusion(s.v, self.nu, self.dz3v, tau_v, p.bottom_drag, mask3)
        w_old = self.w_diag(s.u, s.v)
        ex_u, ex_v = self.momentum_extras(s.u, s.v, w_old)
        stress_u = np.zeros(g.shape3d)
        stress_v = np.zeros(g.shape3d)
        stress_u[0] = thv * dt * tau_u * col.safe_inv(self.dz3u[0])
        stress_v[0] = thv * dt * tau_v * col.safe_inv(self.dz3v[0])
        stress_u = stress_u + dt * ex_u
        stress_v = stress_v + dt * ex_v
        if pr.name == "split_explicit":
            return self._step_split(s, pgx, pgy, av_prev_u, av_prev_v, dzu_diff, dzv_diff,
                                    stress_u, stress_v)

        gx_eta = hz.gradx(s.eta, g.dx, d.mask_u)[None] * d.mask3u
        gy_eta = hz.grady(s.eta, g.dy, d.mask_v)[None] * d.mask3v
        U_old = np.sum(self.dz3u * s.u, axis=0)
        V_old = np.sum(self.dz3v * s.v, axis=0)
        div_old = self._div_bt(U_old, V_old, s.eta)

        u_it, v_it, eta_new = s.u, s.v, s.eta
        for _ in range(pr.n_picard):
            v_cor = (1.0 - tc) * av_prev_v + tc * hz.avg_v_to_u(v_it, d.mask3v, d.mask3u)
            u_cor = (1.0 - tc) * av_prev_u + tc * hz.avg_u_to_v(u_it, d.mask3u, d.mask3v)
            gu = (s.u + dt * (p.f * v_cor + pgx) - p.g * dt * (1.0 - th) * gx_eta
                  + (1.0 - thv) * dt * dzu_diff)
            gv = (s.v + dt * (-p.f * u_cor + pgy) - p.g * dt * (1.0 - th) * gy_eta
                  + (1.0 - thv) * dt * dzv_diff)
            gh_u = self._solve_column(self.tri_m, gu + stress_u) * mask3
            gh_v = self._solve_column(self.tri_m, gv + stress_v) * mask3
            if pr.name == "theta":
                div_g = self._div_bt(np.sum(self.dz3u * gh_u, axis=0),
                                     np.sum(self.dz3v * gh_v, axis=0), s.eta)
                rhs = s.eta - dt * ((1.0 - th) * div_old + th * div_g)
                eta_new, rep = self.solver.solve(rhs)
                self.solver_iterations += rep.iterations
                self.solver_failures += 0 if rep.converged else 1
                u_it = gh_u - p.g * dt * th * self.q * hz.gradx(eta_new, g.dx, d.mask_u)[None] * d.mask3u
                v_it = gh_v - p.g * dt * th * self.q * hz.grady(eta_new, g.dy, d.mask_v)[None] * d.mask3v
                u_it, v_it = u_it * mask3, v_it * mask3
            else:
                u_it, v_it = gh_u, gh_v
        if pr.name == "fb":
            eta_new = s.eta - dt * self._div_bt(np.sum(self.dz3u * u_it, axis=0),
                                                np.sum(self.dz3v * v_it, axis=0), s.eta)
        w = self.w_diag(u_it, v_it)
        return self._advance_tracers(s, u_it, v_it, w, eta_new)

    def _barotropic(self, U0, V0, eta, fx, fy):
        """n_split forward-backward substeps of the depth-integrated flow
        with Picard-iterated Coriolis (libs/core/model3d_v05.py:312-335)."""
        d, p, g = self.domain, self.physics, self.grid
        ddt = self.dt / self.n_split
        U, V, e = U0, V0, eta
        acc_u, acc_v = np.zeros_like(U0), np.zeros_like(V0)
        tc, npic = self.params.theta_cor, self.params.n_picard
        for _ in range(self.n_split):
            av_v = hz.avg_v_to_u(V, d.mask_v, d.mask_u)
            av_u = hz.avg_u_to_v(U, d.mask_u, d.mask_v)
            gxe = hz.gradx(e, g.dx, d.mask_u)
            gye = hz.grady(e, g.dy, d.mask_v)
            Ui, Vi = U, V
            for _ in range(npic):
                v_cor = (1.0 - tc) * av_v + tc * hz.avg_v_to_u(Vi, d.mask_v, d.mask_u)
                u_cor = (1.0 - tc) * av_u + tc * hz.avg_u_to_v(Ui, d.mask_u, d.mask_v)
                Ui = (U + ddt * (p.f * v_cor - p.g * self.Ku * gxe + fx)) * d.mask_u
                Vi = (V + ddt * (-p.f * u_cor - p.g * self.Kv * gye + fy)) * d.mask_v
            U, V = Ui, Vi
            e = e - ddt * self._div_bt(U, V, e)
            acc_u += U
            acc_v += V
        self.barotropic_substeps += self.n_split
        if self.params.barotropic_average == "mean":
            return acc_u / self.n_split, acc_v / self.n_split, e
        return U, V, e

    def _step_split(self, s, pgx, pgy, av_prev_u, av_prev_v, dzu_diff, dzv_diff, stress_u,
                    stress_v):
        """Baroclinic predictor, barotropic substeps, depth-mean replacement
        (libs/core/model3d_v05.py:417-463)."""
        d, p = self.domain, self.physics
        dt, tc, thv = self.dt, self.params.theta_cor, self.params.theta_v
        mask3 = d.mask3
        U0 = np.sum(self.dz3u * s.u, axis=0)
        V0 = np.sum(self.dz3v * s.v, axis=0)
        u_it, v_it = s.u, s.v
        gh_u = gh_v = None
        for _ in range(self.params.n_picard):
            v_cor = (1.0 - tc) * av_prev_v + tc * hz.avg_v_to_u(v_it, d.mask3v, d.mask3u)
            u_cor = (1.0 - tc) * av_prev_u + tc * hz.avg_u_to_v(u_it, d.mask3u, d.mask3v)
            gu = s.u + dt * (p.f * v_cor + pgx) + (1.0 - thv) * dt * dzu_diff
            gv = s.v + dt * (-p.f * u_cor + pgy) + (1.0 - thv) * dt * dzv_diff
            gh_u = self._solve_column(self.tri_m, gu + stress_u) * mask3
            gh_v = self._solve_column(self.tri_m, gv + stress_v) * mask3
            u_it, v_it = gh_u, gh_v
        Um, Vm = np.sum(self.dz3u * gh_u, axis=0), np.sum(self.dz3v * gh_v, axis=0)
        cor_x = p.f * ((1.0 - tc) * hz.avg_v_to_u(V0, d.mask_v, d.mask_u)
                       + tc * hz.avg_v_to_u(Vm, d.mask_v, d.mask_u))
        cor_y = -p.f * ((1.0 - tc) * hz.avg_u_to_v(U0, d.mask_u, d.mask_v)
                        + tc * hz.avg_u_to_v(Um, d.mask_u, d.mask_v))
        fx = (Um - U0) / dt - cor_x
        fy = (Vm - V0) / dt - cor_y
        U, V, eta_new = self._barotropic(U0, V0, s.eta, fx, fy)
        Hu = np.sum(self.dz3u, axis=0)
        Hv = np.sum(self.dz3v, axis=0)
        inv_hu = col.safe_inv(Hu)
        inv_hv = col.safe_inv(Hv)
        u_new = (gh_u - np.sum(self.dz3u * gh_u, axis=0) * inv_hu + U * inv_hu) * d.mask3u
        v_new = (gh_v - np.sum(self.dz3v * gh_v, axis=0) * inv_hv + V * inv_hv) * d.mask3v
        w = self.w_diag(u_new, v_new)
        return self._advance_tracers(s, u_new, v_new, w, eta_new)

    # ---------------------------------------------------------------- tracers
    def tracer_advection(self, c, u, v, w):
        """libs/core/model3d_v05.py:482-512"""
        d, g = self.domain, self.grid
        sc = self.params.advection
        if sc == "centered2":
            cx = 0.5 * (c + np.roll(c, -1, axis=-1))
            cy = 0.5 * (c + np.roll(c, -1, axis=-2))
        elif sc in ("up3", "up3_tvd"):
            cx = hz.kappa_face(c, u, -1, d.mask3, sc == "up3_tvd")
            cy = hz.kappa_face(c, v, -2, d.mask3, sc == "up3_tvd")
        else:
            cx = np.where(u > 0, c, np.roll(c, -1, axis=-1))
            cy = np.where(v > 0, c, np.roll(c, -1, axis=-2))
        fx = self.dz3u * u * cx
        fy = self.dz3v * v * cy
        inv = col.safe_inv(self.dz3)
        hdiv = ((fx - np.roll(fx, 1, axis=-1)) / g.dx
                + (fy - np.roll(fy, 1, axis=-2)) / g.dy) * inv
        nz = c.shape[0]
        cf = np.zeros((nz + 1,) + c.shape[1:])
        if sc in ("up3", "up3_tvd"):
            cf[1:nz] = hz.kappa_face_z(c, w, d.mask3, sc == "up3_tvd")
        else:
            cf[1:nz] = 0.5 * (c[:nz - 1] + c[1:nz])
        cf[0] = c[0]
        vdiv = (w[:nz] * cf[:nz] - w[1:] * cf[1:]) * inv
        return (hdiv + vdiv) * d.mask3

    def _advance_tracer(self, c, u, v, w, surf_flux, source=None):
        """libs/core/model3d_v05.py:466-480"""
        d, g, dt, thv = self.domain, self.grid, self.dt, self.params.theta_v
        tend = np.zeros_like(c) if source is None else source
        if self.params.advection != "none":
            tend = tend - self.tracer_advection(c, u, v, w)
        if self.physics.K_h > 0.0:
            tend = tend + self.physics.K_h * hz.laplacian(c, g.dx, g.dy, d.mask3u, d.mask3v,
                                                          d.mask3)
        rhs = (c + dt * tend
               + (1.0 - thv) * dt * col.apply_diffusion(c, self.kappa, self.dz3, surf_flux,
                                                        0.0, d.mask3))
        rhs[0] = rhs[0] + thv * dt * surf_flux * col.safe_inv(self.dz3[0])
        return self._solve_column(self.tri_b, rhs) * d.mask3

    def _advance_tracers(self, s, u, v, w, eta_new):
        """libs/core/model3d_v05.py:514-533"""
        d, p, dt, g = self.domain, self.physics, self.dt, self.grid
        if p.tracers == "TS":
            flux_T = np.full(g.shape, p.q_heat / (p.rho0 * p.cp))
            flux_S = np.full(g.shape, p.q_salt)
            T_new = self._advance_tracer(s.T, u, v, w, flux_T)
            S_new = self._advance_tracer(s.S, u, v, w, flux_S)
            b_new = self.buoyancy_of(T_new, S_new) * d.mask3
            return OState(u=u, v=v, b=b_new, eta=eta_new, t=s.t + dt, T=T_new, S=S_new,
                          e=self._e_new)
        source = -p.N2 * col.w_at_centres(w)
        b_new = self._advance_tracer(s.b, u, v, w, np.zeros(g.shape), source=source)
        return OState(u=u, v=v, b=b_new, eta=eta_new, t=s.t + dt, e=self._e_new)

    def integrate(self, state, n_steps):
        for _ in range(n_steps):
            state = self.step(state)
        return state

    def diagnose_w(self, state):
        return self.w_diag(state.u, state.v)

    @property
    def eos_flops(self):
        return self.eos_cells * FLOPS_PER_CELL[self.physics.eos]
