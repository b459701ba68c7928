"""The reference's v0.2 3D hydrostatic model, restated for the oracle
(libs/core/model3d.py:93-391, libs/core/vertical.py, libs/core/advection.py,
libs/core/operators.py): flat bottom, uniform layers, doubly periodic;
theta, fb and split_explicit steps with buoyancy or T/S tracers. Test
infrastructure only.

Every expression keeps the reference's operation order (Python left to
right, scalars first), and axis-0 sums are NumPy's sequential accumulation
over layers, so states are bitwise the reference's."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle.model2d import build_solver


def _e(a):
    return np.roll(a, -1, axis=-1)


def _w(a):
    return np.roll(a, 1, axis=-1)


def _n(a):
    return np.roll(a, -1, axis=-2)


def _s(a):
    return np.roll(a, 1, axis=-2)


def gradx(a, dx):
    return (_e(a) - a) / dx


def grady(a, dy):
    return (_n(a) - a) / dy


def div(u, v, dx, dy):
    return (u - _w(u)) / dx + (v - _s(v)) / dy


def avg_v_to_u(v):
    vs = _s(v)
    return 0.25 * (v + vs + _e(v) + _e(vs))


def avg_u_to_v(u):
    uw = _w(u)
    return 0.25 * (u + uw + _n(u) + _n(uw))


def laplacian_h(a, dx, dy):
    """advection.py:97-106"""
    return (_e(a) - 2.0 * a + _w(a)) / dx ** 2 + (_n(a) - 2.0 * a + _s(a)) / dy ** 2


# ------------------------------------------------------------ vertical.py
def thomas(sub, diag, sup, rhs):
    """vertical.py:17-44"""
    nz = diag.shape[0]
    c = np.empty_like(sup)
    d = np.empty_like(rhs)
    c[0] = sup[0] / diag[0]
    d[0] = rhs[0] / diag[0]
    for k in range(1, nz):
        den = diag[k] - sub[k] * c[k - 1]
        c[k] = sup[k] / den
        d[k] = (rhs[k] - sub[k] * d[k - 1]) / den
    x = np.empty_like(rhs)
    x[nz - 1] = d[nz - 1]
    for k in range(nz - 2, -1, -1):
        x[k] = d[k] - c[k] * x[k + 1]
    return x


def diffusion_coeffs(nu, dz, dt, theta_v, bottom_drag=0.0):
    """vertical.py:47-72"""
    nz = nu.shape[0] - 1
    fac = theta_v * dt / dz ** 2
    a_top = fac * nu[:nz]
    a_bot = fac * nu[1:nz + 1]
    sub = -a_top.copy()
    sup = -a_bot.copy()
    diag = 1.0 + a_top + a_bot
    diag[0] = 1.0 + a_bot[0]
    sub[0] = 0.0
    diag[nz - 1] = 1.0 + a_top[nz - 1] + theta_v * dt * bottom_drag / dz
    sup[nz - 1] = 0.0
    return sub, diag, sup


def apply_diffusion(field, nu, dz, surface_flux=None, bottom_drag=0.0):
    """vertical.py:75-86"""
    nz = field.shape[0]
    flux = np.zeros((nz + 1,) + field.shape[1:])
    flux[1:nz] = nu[1:nz] * (field[:nz - 1] - field[1:nz]) / dz
    flux[0] = 0.0 if surface_flux is None else surface_flux
    flux[nz] = bottom_drag * field[nz - 1]
    return (flux[:nz] - flux[1:nz + 1]) / dz


def w_from_divergence(dv, dz):
    """vertical.py:89-96"""
    nz = dv.shape[0]
    w = np.zeros((nz + 1,) + dv.shape[1:])
    w[:nz] = -dz * np.cumsum(dv[::-1], axis=0)[::-1]
    return w


def buoyancy_potential(b, dz):
    """vertical.py:103-107"""
    above = np.concatenate([np.zeros((1,) + b.shape[1:]), np.cumsum(b[:-1], axis=0)], axis=0)
    return dz * (0.5 * b + above)


# ----------------------------------------------------------- advection.py
def _face(c, vel, axis, scheme):
    c_next = np.roll(c, -1, axis=axis)
    if scheme == "centered2":
        return 0.5 * (c + c_next)
    if scheme == "upwind1":
        return np.where(vel > 0.0, c, c_next)
    raise ValueError(f"unknown advection scheme '{scheme}'")


def tracer_advection(c, u, v, w, dx, dy, dz, scheme):
    """advection.py:39-60"""
    nz = c.shape[0]
    fx = u * _face(c, u, -1, scheme)
    fy = v * _face(c, v, -2, scheme)
    fz = np.zeros((nz + 1,) + c.shape[1:])
    if nz > 1:
        above, below = c[:-1], c[1:]
        cz = 0.5 * (above + below) if scheme == "centered2" else np.where(w[1:-1] > 0.0, below, above)
        fz[1:nz] = w[1:nz] * cz
    return (fx - _w(fx)) / dx + (fy - _s(fy)) / dy + (fz[:nz] - fz[1:nz + 1]) / dz


def momentum_advection(u, v, w, dx, dy, dz):
    """advection.py:63-94"""
    nz = u.shape[0]
    ubar_c = 0.5 * (_w(u) + u)
    adv_ux = (_e(ubar_c) ** 2 - ubar_c ** 2) / dx
    vbar = 0.5 * (v + _e(v))
    fy_u = vbar * (0.5 * (u + _n(u)))
    adv_uy = (fy_u - _s(fy_u)) / dy
    wbar_u = 0.5 * (w + _e(w))
    fz_u = np.zeros_like(wbar_u)
    if nz > 1:
        fz_u[1:nz] = wbar_u[1:nz] * 0.5 * (u[:-1] + u[1:])
    adv_uz = (fz_u[:nz] - fz_u[1:nz + 1]) / dz
    vbar_c = 0.5 * (_s(v) + v)
    adv_vy = (_n(vbar_c) ** 2 - vbar_c ** 2) / dy
    ubar = 0.5 * (u + _n(u))
    fx_v = ubar * (0.5 * (v + _e(v)))
    adv_vx = (fx_v - _w(fx_v)) / dx
    wbar_v = 0.5 * (w + _n(w))
    fz_v = np.zeros_like(wbar_v)
    if nz > 1:
        fz_v[1:nz] = wbar_v[1:nz] * 0.5 * (v[:-1] + v[1:])
    adv_vz = (fz_v[:nz] - fz_v[1:nz + 1]) / dz
    return adv_ux + adv_uy + adv_uz, adv_vx + adv_vy + adv_vz


# -------------------------------------------------------------- the model
@dataclass
class State3:
    u: np.ndarray
    v: np.ndarray
    b: np.ndarray
    eta: np.ndarray
    t: float = 0.0
    T: np.ndarray | None = None
    S: np.ndarray | None = None


@dataclass
class Phys3:
    """Physics3D (model3d.py:53-90)."""
    g: float
    f: float
    H: float
    N2: float = 0.0
    nu: float = 0.0
    kappa: float = 0.0
    rho0: float = 1025.0
    tau_x: float = 0.0
    tau_y: float = 0.0
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
        return self.g * (self.alpha_T * (T - self.T0) - self.beta_S * (S - self.S0))


@dataclass
class Params3:
    name: str = "theta"
    theta: float = 0.5
    theta_cor: float = 0.5
    theta_v: float = 0.5
    n_picard: int = 2
    solver: str = "pcg_jacobi"
    rtol: float = 1e-12
    max_iter: int = 2000
    advection: str = "none"
    n_split: int = 0
    barotropic_average: str = "none"


class Stepper3:
    """model3d.py:93-391"""

    def __init__(self, nx, ny, nz, dx, dy, H, phys: Phys3, params: Params3, dt):
        self.nx, self.ny, self.nz, self.dx, self.dy = nx, ny, nz, dx, dy
        self.dz = H / nz
        self.phys, self.params, self.dt = phys, params, float(dt)
        self.solver_iterations = 0
        self.solver_failures = 0
        self.barotropic_substeps = 0
        shape3 = (nz, ny, nx)
        self.nu = np.full((nz + 1, ny, nx), phys.nu)
        self.kappa = np.full((nz + 1, ny, nx), phys.kappa)
        thv = params.theta_v
        self.tri_m = diffusion_coeffs(self.nu, self.dz, self.dt, thv, phys.bottom_drag)
        self.tri_b = diffusion_coeffs(self.kappa, self.dz, self.dt, thv, 0.0)
        self.q = thomas(*self.tri_m, np.ones(shape3))
        h_eff = self.dz * np.sum(self.q, axis=0)
        self.h_eff = float(np.mean(h_eff))
        coef = phys.g * self.h_eff * params.theta ** 2 * self.dt ** 2
        self.solver = (build_solver(params.solver, nx, ny, dx, dy, coef, params.rtol,
                                    params.max_iter) if params.name == "theta" else None)
        self.n_split = 0
        if params.name == "split_explicit":
            s_max = np.sqrt(4.0 / dx ** 2 + 4.0 / dy ** 2)
            dt_g = 2.0 / (np.sqrt(phys.g * phys.H) * s_max)
            dt_c = 2.0 / abs(phys.f) if phys.f != 0.0 else np.inf
            dt_baro = float(min(dt_g, dt_c))
            self.n_split = (params.n_split if params.n_split > 0
                            else max(1, int(np.ceil(1.2 * self.dt / dt_baro))))

    def _pg(self, b):
        phi = buoyancy_potential(b, self.dz)
        phi = phi - np.mean(phi, axis=0, keepdims=True)
        return gradx(phi, self.dx), grady(phi, self.dy)

    def _extras(self, u, v, w):
        ex_u, ex_v = np.zeros_like(u), np.zeros_like(v)
        if self.params.advection != "none":
            au, av = momentum_advection(u, v, w, self.dx, self.dy, self.dz)
            ex_u -= au
            ex_v -= av
        if self.phys.A_h > 0.0:
            ex_u += self.phys.A_h * laplacian_h(u, self.dx, self.dy)
            ex_v += self.phys.A_h * laplacian_h(v, self.dx, self.dy)
        return ex_u, ex_v

    def _tracer(self, c, u, v, w, flux, source=None):
        dt, thv, dz = self.dt, self.params.theta_v, self.dz
        tend = np.zeros_like(c) if source is None else source
        if self.params.advection != "none":
            tend = tend - tracer_advection(c, u, v, w, self.dx, self.dy, dz, self.params.advection)
        if self.phys.K_h > 0.0:
            tend = tend + self.phys.K_h * laplacian_h(c, self.dx, self.dy)
        rhs = c + dt * tend + (1.0 - thv) * dt * apply_diffusion(c, self.kappa, dz, flux)
        rhs[0] = rhs[0] + thv * dt * flux / dz
        return thomas(*self.tri_b, rhs)

    def _tracers(self, s, u, v, w, eta_new):
        p, dt = self.phys, self.dt
        shape = (self.ny, self.nx)
        if p.tracers == "TS":
            T = self._tracer(s.T, u, v, w, np.full(shape, p.q_heat / (p.rho0 * p.cp)))
            S = self._tracer(s.S, u, v, w, np.full(shape, p.q_salt))
            return State3(u, v, p.buoyancy(T, S), eta_new, s.t + dt, T, S)
        src = -p.N2 * (0.5 * (w[:-1] + w[1:]))
        b = self._tracer(s.b, u, v, w, np.zeros(shape), source=src)
        return State3(u, v, b, eta_new, s.t + dt)

    def _common(self, s):
        p, dt, dz = self.phys, self.dt, self.dz
        thv = self.params.theta_v
        shape = (self.ny, self.nx)
        tau_u = np.full(shape, p.tau_x / p.rho0)
        tau_v = np.full(shape, p.tau_y / p.rho0)
        dz_u = apply_diffusion(s.u, self.nu, dz, tau_u, p.bottom_drag)
        dz_v = apply_diffusion(s.v, self.nu, dz, tau_v, p.bottom_drag)
        w_old = w_from_divergence(div(s.u, s.v, self.dx, self.dy), dz)
        ex_u, ex_v = self._extras(s.u, s.v, w_old)
        st_u = np.zeros((self.nz,) + shape)
        st_v = np.zeros((self.nz,) + shape)
        st_u[0] = thv * dt * tau_u / dz
        st_u = st_u + dt * ex_u
        st_v[0] = thv * dt * tau_v / dz
        st_v = st_v + dt * ex_v
        return dz_u, dz_v, st_u, st_v

    def step(self, s: State3) -> State3:
        if self.params.name == "split_explicit":
            return self._step_split(s)
        p, dt, dz = self.phys, self.dt, self.dz
        g, f = p.g, p.f
        th, tc, thv = self.params.theta, self.params.theta_cor, self.params.theta_v
        if self.params.name != "theta":
            th = 0.0
        pgx, pgy = self._pg(s.b)
        gx_eta, gy_eta = gradx(s.eta, self.dx), grady(s.eta, self.dy)
        av_v, av_u = avg_v_to_u(s.v), avg_u_to_v(s.u)
        div_old = div(dz * np.sum(s.u, axis=0), dz * np.sum(s.v, axis=0), self.dx, self.dy)
        dz_u, dz_v, st_u, st_v = self._common(s)
        u_it, v_it, eta_new = s.u, s.v, s.eta
        for _ in range(self.params.n_picard):
            v_cor = (1.0 - tc) * av_v + tc * avg_v_to_u(v_it)
            u_cor = (1.0 - tc) * av_u + tc * avg_u_to_v(u_it)
            gu = (s.u + dt * (f * v_cor + pgx) - g * dt * (1.0 - th) * gx_eta
                  + (1.0 - thv) * dt * dz_u)
            gv = (s.v + dt * (-f * u_cor + pgy) - g * dt * (1.0 - th) * gy_eta
                  + (1.0 - thv) * dt * dz_v)
            gh_u = thomas(*self.tri_m, gu + st_u)
            gh_v = thomas(*self.tri_m, gv + st_v)
            if self.params.name == "theta":
                div_g = div(dz * np.sum(gh_u, axis=0), dz * np.sum(gh_v, axis=0), self.dx, self.dy)
                rhs = s.eta - dt * ((1.0 - th) * div_old + th * div_g)
                eta_new, rep = self.solver.solve(rhs)
                self.solver_iterations += rep.iterations
                if not rep.converged:
                    self.solver_failures += 1
                u_it = gh_u - g * dt * th * self.q * gradx(eta_new, self.dx)
                v_it = gh_v - g * dt * th * self.q * grady(eta_new, self.dy)
            else:
                u_it, v_it = gh_u, gh_v
        if self.params.name == "fb":
            eta_new = s.eta - dt * div(dz * np.sum(u_it, axis=0), dz * np.sum(v_it, axis=0),
                                       self.dx, self.dy)
        w = w_from_divergence(div(u_it, v_it, self.dx, self.dy), dz)
        return self._tracers(s, u_it, v_it, w, eta_new)

    def _barotropic(self, u_int, v_int, eta, fx, fy):
        """model3d.py:206-247"""
        g, f, h = self.phys.g, self.phys.f, self.phys.H
        ddt = self.dt / self.n_split
        u_i, v_i, e = u_int, v_int, eta
        acc_u, acc_v = np.zeros_like(u_int), np.zeros_like(v_int)
        tc, npic = self.params.theta_cor, self.params.n_picard
        for _ in range(self.n_split):
            av_v, av_u = avg_v_to_u(v_i), avg_u_to_v(u_i)
            gxe, gye = gradx(e, self.dx), grady(e, self.dy)
            ui, vi = u_i, v_i
            for _ in range(npic):
                v_cor = (1.0 - tc) * av_v + tc * avg_v_to_u(vi)
                u_cor = (1.0 - tc) * av_u + tc * avg_u_to_v(ui)
                ui = u_i + ddt * (f * v_cor - g * h * gxe + fx)
                vi = v_i + ddt * (-f * u_cor - g * h * gye + fy)
            u_i, v_i = ui, vi
            e = e - ddt * div(u_i, v_i, self.dx, self.dy)
            acc_u += u_i
            acc_v += v_i
        self.barotropic_substeps += self.n_split
        if self.params.barotropic_average == "mean":
            return acc_u / self.n_split, acc_v / self.n_split, e
        return u_i, v_i, e

    def _step_split(self, s):
        """model3d.py:249-301"""
        p, dt, dz = self.phys, self.dt, self.dz
        f, tc, thv, h = p.f, self.params.theta_cor, self.params.theta_v, p.H
        pgx, pgy = self._pg(s.b)
        av_v, av_u = avg_v_to_u(s.v), avg_u_to_v(s.u)
        dz_u, dz_v, st_u, st_v = self._common(s)
        u_int0 = dz * np.sum(s.u, axis=0)
        v_int0 = dz * np.sum(s.v, axis=0)
        u_it, v_it = s.u, s.v
        gh_u = gh_v = cor_x = cor_y = None
        for _ in range(self.params.n_picard):
            v_cor = (1.0 - tc) * av_v + tc * avg_v_to_u(v_it)
            u_cor = (1.0 - tc) * av_u + tc * avg_u_to_v(u_it)
            gu = s.u + dt * (f * v_cor + pgx) + (1.0 - thv) * dt * dz_u
            gv = s.v + dt * (-f * u_cor + pgy) + (1.0 - thv) * dt * dz_v
            gh_u = thomas(*self.tri_m, gu + st_u)
            gh_v = thomas(*self.tri_m, gv + st_v)
            cor_x = dz * np.sum(f * v_cor, axis=0)
            cor_y = -dz * np.sum(f * u_cor, axis=0)
            u_it, v_it = gh_u, gh_v
        fx = (dz * np.sum(gh_u, axis=0) - u_int0) / dt - cor_x
        fy = (dz * np.sum(gh_v, axis=0) - v_int0) / dt - cor_y
        u_int, v_int, eta_new = self._barotropic(u_int0, v_int0, s.eta, fx, fy)
        u_new = gh_u - np.mean(gh_u, axis=0, keepdims=True) + u_int / h
        v_new = gh_v - np.mean(gh_v, axis=0, keepdims=True) + v_int / h
        w = w_from_divergence(div(u_new, v_new, self.dx, self.dy), dz)
        return self._tracers(s, u_new, v_new, w, eta_new)
