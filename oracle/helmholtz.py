"""Variable-coefficient free-surface Helmholtz solvers of the oracle:
    eta - coef * D[K G[eta]] = rhs   on the masked C-grid
(libs/core/solvers_var.py:22-197; SolveReport libs/core/solvers.py:25-30).

Kinds: pcg_jacobi, pcg_rbgs, rbgs, multigrid (the reference's RBGS V-cycle
iterated as a solver) and pcg_mg (no reference counterpart: PCG whose
preconditioner is one V-cycle on `mg_levels` levels from a zero guess, with
`mg_sweeps` damped-Jacobi sweeps per side or, with mg_smoother="rbgs",
red-black Gauss-Seidel sweeps: red then black on the way down, black then
red on the way up, so that the preconditioner stays symmetric). The pcg_mg
V-cycle is pinned in tests/test_oracle_pins.py to a dense matrix recursion
assembled independently of this code.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SolveReport:
    iterations: int = 0
    residual: float = 0.0
    converged: bool = True


def colour_masks(ny, nx):
    """Red = (i + j) even (libs/core/solvers.py:145-148)."""
    j, i = np.indices((ny, nx))
    red = (i + j) % 2 == 0
    return red, ~red


class VarOperator:
    """A = I - coef D[K G .] (libs/core/solvers_var.py:22-60)."""

    def __init__(self, nx, ny, dx, dy, coef, Ku, Kv, mask):
        self.nx, self.ny, self.dx, self.dy = nx, ny, float(dx), float(dy)
        self.coef = float(coef)
        self.Ku = np.ascontiguousarray(Ku)
        self.Kv = np.ascontiguousarray(Kv)
        self.mask = np.ascontiguousarray(mask)
        self.cx = self.coef * (1.0 / self.dx ** 2)
        self.cy = self.coef * (1.0 / self.dy ** 2)
        diag = (1.0 + self.cx * (self.Ku + np.roll(self.Ku, 1, axis=-1))
                + self.cy * (self.Kv + np.roll(self.Kv, 1, axis=-2)))
        self.diag = np.where(self.mask > 0, diag, 1.0)
        self.inv_diag = 1.0 / self.diag

    def apply(self, x):
        x = x * self.mask
        fx = self.Ku * (np.roll(x, -1, axis=-1) - x)
        fy = self.Kv * (np.roll(x, -1, axis=-2) - x)
        lap = ((fx - np.roll(fx, 1, axis=-1)) / self.dx ** 2
               + (fy - np.roll(fy, 1, axis=-2)) / self.dy ** 2)
        return (x - self.coef * lap) * self.mask + x * (1.0 - self.mask)

    def neighbour_sum(self, x):
        return (self.cx * (self.Ku * np.roll(x, -1, axis=-1)
                           + np.roll(self.Ku, 1, axis=-1) * np.roll(x, 1, axis=-1))
                + self.cy * (self.Kv * np.roll(x, -1, axis=-2)
                             + np.roll(self.Kv, 1, axis=-2) * np.roll(x, 1, axis=-2)))


def rbgs_sweep(x, b, op, masks, reverse=False):
    """One red-black Gauss-Seidel sweep in place (libs/core/solvers_var.py:63-69)."""
    for colour in (masks[::-1] if reverse else masks):
        upd = op.inv_diag * (b + op.neighbour_sum(x))
        x[colour] = upd[colour]
    x *= op.mask


def coarsen(op):
    """Agglomerated coarse operator: face = mean of the two fine faces on the
    coarse cell's east/north edge, wet if any child is wet
    (libs/core/solvers_var.py:118-127)."""
    cKu = 0.5 * (op.Ku[0::2, 1::2] + op.Ku[1::2, 1::2])
    cKv = 0.5 * (op.Kv[1::2, 0::2] + op.Kv[1::2, 1::2])
    cm = np.maximum.reduce([op.mask[0::2, 0::2], op.mask[0::2, 1::2],
                            op.mask[1::2, 0::2], op.mask[1::2, 1::2]])
    return VarOperator(op.nx // 2, op.ny // 2, op.dx * 2.0, op.dy * 2.0, op.coef, cKu, cKv, cm)


def restrict(r):
    return 0.25 * (r[0::2, 0::2] + r[0::2, 1::2] + r[1::2, 0::2] + r[1::2, 1::2])


def prolong(e):
    return np.repeat(np.repeat(e, 2, axis=0), 2, axis=1)


class HelmholtzVar:
    """libs/core/solvers_var.py:72-197 plus the pcg_mg kind."""

    KINDS = ("pcg_jacobi", "pcg_rbgs", "rbgs", "multigrid", "pcg_mg")

    def __init__(self, nx, ny, dx, dy, coef, Ku, Kv, mask, kind="pcg_jacobi",
                 rtol=1e-12, max_iter=2000, n_pre=2, n_post=2, min_size=8,
                 mg_levels=5, mg_sweeps=2, mg_omega=0.8, mg_smoother="jacobi"):
        if kind not in self.KINDS:
            raise ValueError(f"unknown variable-coefficient solver '{kind}' "
                             "(expected pcg_jacobi|pcg_rbgs|rbgs|multigrid|pcg_mg)")
        self.kind = kind
        self.rtol, self.max_iter = rtol, max_iter
        self.n_pre, self.n_post, self.min_size = n_pre, n_post, min_size
        self.mg_levels, self.mg_sweeps, self.mg_omega = mg_levels, mg_sweeps, mg_omega
        if mg_smoother not in ("jacobi", "rbgs"):
            raise ValueError(f"unknown pcg_mg smoother '{mg_smoother}' (expected jacobi|rbgs)")
        self.mg_smoother = mg_smoother
        self.masks = colour_masks(ny, nx)
        self.op = VarOperator(nx, ny, dx, dy, coef, Ku, Kv, mask)
        self._build_levels()

    def update(self, Ku, Kv):
        op = self.op
        self.op = VarOperator(op.nx, op.ny, op.dx, op.dy, op.coef, Ku, Kv, op.mask)
        self._build_levels()

    def _build_levels(self):
        self.levels = [self.op]
        if self.kind == "multigrid":
            # halve while both sides are even and stay >= min_size
            while (self.levels[-1].nx % 2 == 0 and self.levels[-1].ny % 2 == 0
                   and self.levels[-1].nx // 2 >= self.min_size
                   and self.levels[-1].ny // 2 >= self.min_size):
                self.levels.append(coarsen(self.levels[-1]))
        elif self.kind == "pcg_mg":
            while (len(self.levels) < self.mg_levels and self.levels[-1].nx % 2 == 0
                   and self.levels[-1].ny % 2 == 0 and self.levels[-1].nx >= 4
                   and self.levels[-1].ny >= 4):
                self.levels.append(coarsen(self.levels[-1]))
        self.level_masks = [colour_masks(op.ny, op.nx) for op in self.levels]

    # reference multigrid: RBGS V-cycle (libs/core/solvers_var.py:139-151)
    def _vcycle(self, x, b, lvl):
        op, masks = self.levels[lvl], self.level_masks[lvl]
        for _ in range(self.n_pre):
            rbgs_sweep(x, b, op, masks)
        if lvl + 1 < len(self.levels):
            r = (b - op.apply(x)) * op.mask
            e = np.zeros((self.levels[lvl + 1].ny, self.levels[lvl + 1].nx))
            e = self._vcycle(e, restrict(r), lvl + 1)
            x = x + prolong(e) * op.mask
        for _ in range(self.n_post):
            rbgs_sweep(x, b, op, masks, reverse=True)
        return x

    # pcg_mg preconditioner: damped-Jacobi V-cycle from a zero guess
    def _jacobi(self, x, b, op):
        return x + self.mg_omega * (op.inv_diag * (b - op.apply(x))) * op.mask

    def _mg_precond(self, b, lvl=0):
        op = self.levels[lvl]
        rb = self.mg_smoother == "rbgs"
        x = np.zeros_like(b)
        for _ in range(self.mg_sweeps):
            if rb:
                rbgs_sweep(x, b, op, self.level_masks[lvl])
            else:
                x = self._jacobi(x, b, op)
        if lvl + 1 < len(self.levels):
            r = (b - op.apply(x)) * op.mask
            e = self._mg_precond(restrict(r), lvl + 1)
            x = x + prolong(e) * op.mask
        for _ in range(self.mg_sweeps):
            if rb:
                rbgs_sweep(x, b, op, self.level_masks[lvl], reverse=True)
            else:
                x = self._jacobi(x, b, op)
        return x

    def _precondition(self, r):
        if self.kind == "pcg_jacobi":
            return self.op.inv_diag * r
        if self.kind == "pcg_mg":
            return self._mg_precond(r)
        z = np.zeros_like(r)
        rbgs_sweep(z, r, self.op, self.masks)
        rbgs_sweep(z, r, self.op, self.masks, reverse=True)
        return z

    def solve(self, rhs):
        """Zero initial guess; stop when ||r|| <= rtol ||rhs * mask||
        (libs/core/solvers_var.py:162-197)."""
        rhs = rhs * self.op.mask
        norm_b = float(np.sqrt(np.sum(rhs * rhs)))
        x = np.zeros_like(rhs)
        if norm_b == 0.0:
            return x, SolveReport(0, 0.0, True)
        tol = self.rtol * norm_b
        if self.kind in ("rbgs", "multigrid"):
            res = np.inf
            for it in range(1, self.max_iter + 1):
                if self.kind == "multigrid":
                    x = self._vcycle(x, rhs, 0)
                else:
                    rbgs_sweep(x, rhs, self.op, self.masks)
                res = float(np.sqrt(np.sum((rhs - self.op.apply(x)) ** 2)))
                if res <= tol:
                    return x, SolveReport(it, res / norm_b, True)
            return x, SolveReport(self.max_iter, res / norm_b, False)
        r = rhs.copy()
        z = self._precondition(r)
        p = z.copy()
        rz = float(np.sum(r * z))
        res = np.inf
        for it in range(1, self.max_iter + 1):
            Ap = self.op.apply(p)
            alpha = rz / float(np.sum(p * Ap))
            x += alpha * p
            r -= alpha * Ap
            res = float(np.sqrt(np.sum(r * r)))
            if res <= tol:
                return x, SolveReport(it, res / norm_b, True)
            z = self._precondition(r)
            rz_new = float(np.sum(r * z))
            p = z + (rz_new / rz) * p
            rz = rz_new
        return x, SolveReport(self.max_iter, res / norm_b, False)
