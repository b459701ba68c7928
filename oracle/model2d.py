"""2D rotating shallow-water model of the oracle (libs/core/schemes.py:97-177,
libs/core/solvers.py:33-281, libs/core/operators.py): forward-backward and
theta schemes with the Picard Coriolis loop, and the five Helmholtz solves of
the reference (fft, pcg_jacobi, pcg_rbgs, rbgs, multigrid). Test
infrastructure only.

Every expression keeps the reference's operation order (Python left to right,
scalars first), so the PCG iterates and the states are bitwise the
reference's; `np.sum` is NumPy's pairwise sum, which the device solve
reproduces."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def _e(a):
    return np.roll(a, -1, axis=-1)


def _w(a):
    return np.roll(a, 1, axis=-1)


def _n(a):
    return np.roll(a, -1, axis=-2)


def _s(a):
    return np.roll(a, 1, axis=-2)


def gradx_u(eta, dx):
    """operators.py:25-27"""
    return (_e(eta) - eta) / dx


def grady_v(eta, dy):
    """operators.py:29-31"""
    return (_n(eta) - eta) / dy


def div_eta(u, v, dx, dy):
    """operators.py:32-34"""
    return (u - _w(u)) / dx + (v - _s(v)) / dy


def avg_v_to_u(v):
    """operators.py:37-40"""
    vs = _s(v)
    return 0.25 * (v + vs + _e(v) + _e(vs))


def avg_u_to_v(u):
    """operators.py:43-46"""
    uw = _w(u)
    return 0.25 * (u + uw + _n(u) + _n(uw))


def laplacian(x, dx, dy):
    """operators.py:49-53: D o G"""
    return div_eta(gradx_u(x, dx), grady_v(x, dy), dx, dy)


@dataclass
class Report:
    iterations: int = 0
    residual: float = 0.0
    converged: bool = True


class FFTSolver:
    """solvers.py:33-52"""

    def __init__(self, nx, ny, dx, dy, coef):
        theta_x = 2.0 * np.pi * np.arange(nx // 2 + 1) / nx
        theta_y = 2.0 * np.pi * np.arange(ny) / ny
        lam_x = 2.0 * (np.cos(theta_x) - 1.0) / dx ** 2
        lam_y = 2.0 * (np.cos(theta_y) - 1.0) / dy ** 2
        self.symbol = 1.0 - coef * (lam_y[:, None] + lam_x[None, :])
        self.shape = (ny, nx)

    def solve(self, rhs):
        return np.fft.irfft2(np.fft.rfft2(rhs) / self.symbol, s=self.shape), Report()


def colour_masks(ny, nx):
    """solvers.py:146-149: red where (i + j) is even."""
    j, i = np.indices((ny, nx))
    red = (i + j) % 2 == 0
    return red, ~red


def rbgs_sweep(x, b, coef, dx, dy, masks, reverse=False):
    """solvers.py:120-136: one red-black sweep in place; each colour's update
    reads the whole array as it stood before that colour."""
    cx, cy = coef / dx ** 2, coef / dy ** 2
    diag = 1.0 + 2.0 * cx + 2.0 * cy
    for m in (masks[::-1] if reverse else masks):
        nb = cx * (_e(x) + _w(x)) + cy * (_n(x) + _s(x))
        x[m] = ((b + nb) / diag)[m]


def apply_5pt(x, coef, dx, dy):
    """solvers.py:139-143: A x with the 5-point Laplacian written out."""
    lap = (_e(x) - 2.0 * x + _w(x)) / dx ** 2 + (_n(x) - 2.0 * x + _s(x)) / dy ** 2
    return x - coef * lap


class PCGSolver:
    """solvers.py:55-117, Jacobi or symmetric red-black Gauss-Seidel
    preconditioner."""

    def __init__(self, dx, dy, coef, rtol, max_iter, precond="jacobi", shape=None):
        self.dx, self.dy, self.coef, self.rtol, self.max_iter = dx, dy, coef, rtol, max_iter
        self.precond = precond
        self.inv_diag = 1.0 / (1.0 + coef * (2.0 / dx ** 2 + 2.0 / dy ** 2))
        self.masks = colour_masks(*shape) if precond == "rbgs" else None

    def apply(self, x):
        return x - self.coef * laplacian(x, self.dx, self.dy)

    def precondition(self, r):
        if self.precond == "jacobi":
            return self.inv_diag * r
        z = np.zeros_like(r)
        rbgs_sweep(z, r, self.coef, self.dx, self.dy, self.masks)
        rbgs_sweep(z, r, self.coef, self.dx, self.dy, self.masks, reverse=True)
        return z

    def solve(self, rhs):
        x = np.zeros_like(rhs)
        r = rhs - self.apply(x)
        norm_b = float(np.sqrt(np.sum(rhs * rhs)))
        if norm_b == 0.0:
            return x, Report(0, 0.0, True)
        z = self.precondition(r)
        p = z.copy()
        rz = float(np.sum(r * z))
        residual = 0.0
        for it in range(1, self.max_iter + 1):
            ap = self.apply(p)
            alpha = rz / float(np.sum(p * ap))
            x = x + alpha * p
            r = r - alpha * ap
            residual = float(np.sqrt(np.sum(r * r))) / norm_b
            if residual < self.rtol:
                return x, Report(it, residual, True)
            z = self.precondition(r)
            rz_new = float(np.sum(r * z))
            p = z + (rz_new / rz) * p
            rz = rz_new
        return x, Report(self.max_iter, residual, False)


class RBGSSolver:
    """solvers.py:152-184: red-black Gauss-Seidel as the solver, residual
    checked every 10 sweeps and at the last."""

    def __init__(self, nx, ny, dx, dy, coef, rtol, max_iter, check_every=10):
        self.dx, self.dy, self.coef, self.rtol = dx, dy, coef, rtol
        self.max_iter, self.check_every = max_iter, check_every
        self.masks = colour_masks(ny, nx)

    def solve(self, rhs):
        x = np.zeros_like(rhs)
        norm_b = float(np.sqrt(np.sum(rhs * rhs))) or 1.0
        residual = 0.0
        for it in range(1, self.max_iter + 1):
            rbgs_sweep(x, rhs, self.coef, self.dx, self.dy, self.masks)
            if it % self.check_every == 0 or it == self.max_iter:
                r = rhs - apply_5pt(x, self.coef, self.dx, self.dy)
                residual = float(np.sqrt(np.sum(r * r))) / norm_b
                if residual < self.rtol:
                    return x, Report(it, residual, True)
        return x, Report(self.max_iter, residual, False)


class MultigridSolver:
    """solvers.py:187-258: V-cycles with red-black smoothing (2 pre, 2 post,
    40 on the coarsest level), 2x2 mean restriction and piecewise-constant
    prolongation; levels halve while both sizes are even and above 4."""

    def __init__(self, nx, ny, dx, dy, coef, rtol, max_iter, n_pre=2, n_post=2, n_coarse=40):
        self.dx, self.dy, self.coef, self.rtol, self.max_iter = dx, dy, coef, rtol, max_iter
        self.n_pre, self.n_post, self.n_coarse = n_pre, n_post, n_coarse
        self.levels = []
        while True:
            self.levels.append((nx, ny, dx, dy, colour_masks(ny, nx)))
            if nx % 2 or ny % 2 or min(nx, ny) <= 4:
                break
            nx, ny, dx, dy = nx // 2, ny // 2, dx * 2, dy * 2

    def vcycle(self, x, b, lvl):
        nx, ny, dx, dy, masks = self.levels[lvl]
        if lvl == len(self.levels) - 1:
            for _ in range(self.n_coarse):
                rbgs_sweep(x, b, self.coef, dx, dy, masks)
            return x
        for k in range(self.n_pre):
            rbgs_sweep(x, b, self.coef, dx, dy, masks, reverse=bool(k % 2))
        r = b - apply_5pt(x, self.coef, dx, dy)
        coarse = r.reshape(ny // 2, 2, nx // 2, 2).mean(axis=(1, 3))
        ec = self.vcycle(np.zeros((ny // 2, nx // 2)), coarse, lvl + 1)
        x = x + np.repeat(np.repeat(ec, 2, axis=0), 2, axis=1)
        for k in range(self.n_post):
            rbgs_sweep(x, b, self.coef, dx, dy, masks, reverse=not bool(k % 2))
        return x

    def solve(self, rhs):
        x = np.zeros_like(rhs)
        norm_b = float(np.sqrt(np.sum(rhs * rhs))) or 1.0
        residual = 0.0
        for it in range(1, self.max_iter + 1):
            x = self.vcycle(x, rhs, 0)
            r = rhs - apply_5pt(x, self.coef, self.dx, self.dy)
            residual = float(np.sqrt(np.sum(r * r))) / norm_b
            if residual < self.rtol:
                return x, Report(it, residual, True)
        return x, Report(self.max_iter, residual, False)


def build_solver(kind, nx, ny, dx, dy, coef, rtol, max_iter):
    """solvers.py:261-281"""
    if kind == "fft":
        return FFTSolver(nx, ny, dx, dy, coef)
    if kind == "pcg_jacobi":
        return PCGSolver(dx, dy, coef, rtol, max_iter)
    if kind in ("rbgs", "multigrid", "pcg_rbgs"):
        if nx % 2 or ny % 2:
            raise ValueError(f"{kind} needs even nx and ny")
        if kind == "rbgs":
            return RBGSSolver(nx, ny, dx, dy, coef, rtol, max_iter)
        if kind == "multigrid":
            return MultigridSolver(nx, ny, dx, dy, coef, rtol, max_iter)
        return PCGSolver(dx, dy, coef, rtol, max_iter, precond="rbgs", shape=(ny, nx))
    raise ValueError(f"unknown solver kind '{kind}'")


class Stepper2D:
    """schemes.py:97-177 (fb and theta)."""

    def __init__(self, nx, ny, dx, dy, g, f, H, scheme, theta, theta_cor, n_picard, solver,
                 rtol, max_iter, dt):
        self.dx, self.dy, self.g, self.f, self.H = dx, dy, g, f, H
        self.scheme, self.theta, self.theta_cor, self.n_picard = scheme, theta, theta_cor, n_picard
        self.dt = float(dt)
        self.solver_iterations = 0
        self.solver_failures = 0
        self.solver = None
        if scheme == "theta":
            coef = g * H * theta ** 2 * self.dt ** 2
            self.solver = build_solver(solver, nx, ny, dx, dy, coef, rtol, max_iter)

    def _coriolis(self, u0, v0, u_it, v_it):
        tc = self.theta_cor
        v_cor = (1.0 - tc) * avg_v_to_u(v0) + tc * avg_v_to_u(v_it)
        u_cor = (1.0 - tc) * avg_u_to_v(u0) + tc * avg_u_to_v(u_it)
        return u_cor, v_cor

    def step(self, u, v, eta):
        g, f, H, dt = self.g, self.f, self.H, self.dt
        gx_eta = gradx_u(eta, self.dx)
        gy_eta = grady_v(eta, self.dy)
        if self.scheme == "fb":
            u_it, v_it = u, v
            for _ in range(self.n_picard):
                u_cor, v_cor = self._coriolis(u, v, u_it, v_it)
                u_it = u + dt * (f * v_cor - g * gx_eta)
                v_it = v + dt * (-f * u_cor - g * gy_eta)
            eta_new = eta - dt * H * div_eta(u_it, v_it, self.dx, self.dy)
            return u_it, v_it, eta_new
        th = self.theta
        div_old = div_eta(u, v, self.dx, self.dy)
        u_it, v_it, eta_new = u, v, eta
        for _ in range(self.n_picard):
            u_cor, v_cor = self._coriolis(u, v, u_it, v_it)
            gu = u + dt * f * v_cor - g * dt * (1.0 - th) * gx_eta
            gv = v - dt * f * u_cor - g * dt * (1.0 - th) * gy_eta
            rhs = eta - dt * H * ((1.0 - th) * div_old + th * div_eta(gu, gv, self.dx, self.dy))
            eta_new, rep = self.solver.solve(rhs)
            self.solver_iterations += rep.iterations
            if not rep.converged:
                self.solver_failures += 1
            u_it = gu - g * dt * th * gradx_u(eta_new, self.dx)
            v_it = gv - g * dt * th * grady_v(eta_new, self.dy)
        return u_it, v_it, eta_new
