"""Adjoint of the fp64 step (Stepper3DV05.step, libs/core/model3d_v05.py:
337-415, theta scheme) and the checkpointed gradient of a scalar of the
state after n steps.

The reverse step runs in hand-written HIP kernels (hoc_adjoint.hip) behind
three C-ABI entries: hoc_stepper_step_tape (the step, bit for bit, writing
the intermediates the reverse reads), hoc_stepper_step_adjoint (everything
after the closure: tracers with the EOS, advection, the Picard passes with
adjoint Helmholtz PCG solves, Coriolis, pressure gradients, explicit
diffusion, q and the face coefficients) and hoc_stepper_closure_adjoint (the
TKE closure with integral lengths). DiffStep.reverse drives them
(fused=True, the default with kernels="hip").

DiffStep also restates the step in torch operations, in the oracle's
operator order; DiffStep.reverse(..., fused=False) is the same reverse
written with transposed torch operators (the reference the fused kernels
are checked against). Six operators of the restatement sit behind
torch.autograd Functions with hand-written adjoints, for torch.autograd.grad
through DiffStep.step:
  equation of state      hoc_eos_buoyancy / hoc_eos_buoyancy_adjoint (dual
                         numbers through the same expression);
  mixing lengths         hoc_gaspar_lengths / hoc_gaspar_lengths_adjoint (the
                         reverse march per interface);
  momentum advection     hoc_stepper_momentum_advection(_adjoint);
  Thomas solves          hoc_thomas_batched / hoc_thomas_adjoint;
  Helmholtz solve        the stepper's PCG on the right-hand side and on the
                         adjoint one (the masked operator is symmetric) plus the
                         face coefficient gradients of -lambda^T (dA/dK) x;
  tracer advection       hoc_stepper_tracer_advection(_adjoint).
kernels="torch" replaces them by torch operations (the formulas, a
sequential Thomas sweep, a dense Helmholtz solve), so autograd
differentiates the whole step: the reference for the hand-written adjoints
on grids small enough for the dense solve.

Adjoint.gradient: the plain HIP step forward with checkpoints, then the
windows reversed last to first, each earlier window recomputed with its
tapes on a second stream while the next is reversed (see _reverse_sweep).

Derivative conventions: NumPy's maximum/minimum keep the first operand on a
tie, and the branch kept is the one differentiated (nmax, nmin below and in
the kernels); sqrt has derivative 0 at 0 (ssqrt), where the reference only
ever uses its value through a selected other branch.

Scope: theta scheme, z-level, closed or periodic edges, advection up3 or
up3_tvd, closure tke (integral or recursive mixing length; the recursive
form's closure reverse by autograd through tke_step) or none, any EOS.
"""

from __future__ import annotations

import ctypes as C
import math
import dataclasses
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

F64 = torch.float64
KAPPA = 1.0 / 3.0


# ------------------------------------------------------------ elementwise
def nmax(a, b):
    """NumPy's maximum (NaN propagates, a tie keeps a); a or b may be a scalar."""
    keep = (a >= b) | (a != a) if isinstance(a, torch.Tensor) else (b <= a)
    return torch.where(keep, a, b)


def nmin(a, b):
    """NumPy's minimum (NaN propagates, a tie keeps a); a or b may be a scalar."""
    keep = (a <= b) | (a != a) if isinstance(a, torch.Tensor) else (b >= a)
    return torch.where(keep, a, b)


def ssqrt(x):
    pos = x > 0.0
    return torch.where(pos, torch.sqrt(torch.where(pos, x, 1.0)), 0.0)


def safe_inv(x):
    pos = x > 0.0
    return torch.where(pos, 1.0 / torch.where(pos, x, 1.0), 0.0)


def east(a, n=1):
    return torch.roll(a, -n, dims=-1)


def west(a, n=1):
    return torch.roll(a, n, dims=-1)


def north(a, n=1):
    return torch.roll(a, -n, dims=-2)


def south(a, n=1):
    return torch.roll(a, n, dims=-2)


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


# ------------------------------------------------ masked C-grid operators
def gradx(eta, dx, mask_u):
    """libs/core/operators_masked.py:18-20"""
    return mask_u * (east(eta) - eta) / dx


def grady(eta, dy, mask_v):
    """libs/core/operators_masked.py:23-24"""
    return mask_v * (north(eta) - eta) / dy


def div(u, v, dx, dy, mask):
    """libs/core/operators_masked.py:27-32"""
    return mask * ((u - west(u)) / dx + (v - south(v)) / dy)


def _wet_mean(pairs):
    """libs/core/operators_masked.py:35-42"""
    num = 0.0
    den = 0.0
    for f, w in pairs:
        num = num + w * f
    for _, w in pairs:
        den = den + w
    pos = den > 0.0
    return torch.where(pos, num / torch.where(pos, den, 1.0), 0.0)


def avg_v_to_u(v, mask_v, mask_u):
    """libs/core/operators_masked.py:45-52"""
    vs, ms = south(v), south(mask_v)
    return _wet_mean([(v, mask_v), (vs, ms), (east(v), east(mask_v)),
                      (east(vs), east(ms))]) * mask_u


def avg_u_to_v(u, mask_u, mask_v):
    """libs/core/operators_masked.py:55-61"""
    uw, mw = west(u), west(mask_u)
    return _wet_mean([(u, mask_u), (uw, mw), (north(u), north(mask_u)),
                      (north(uw), north(mw))]) * mask_v


def laplacian(a, dx, dy, mask_u, mask_v, mask):
    """libs/core/operators_masked.py:64-68"""
    return div(gradx(a, dx, mask_u), grady(a, dy, mask_v), dx, dy, mask)


# ------------------------------------- transposes of the C-grid operators
# east and north are rolls by -1, west and south by +1: east^T = west and
# north^T = south. Each function below is the adjoint of the operator of the
# same name without the suffix, at the same masks.
def gradx_T(X, dx, mask_u):
    m = mask_u * X
    return (west(m) - m) / dx


def grady_T(Y, dy, mask_v):
    m = mask_v * Y
    return (south(m) - m) / dy


def div_T(D, dx, dy, mask):
    m = mask * D
    return (m - east(m)) / dx, (m - north(m)) / dy


def _mean_weight(G, den, mask_out):
    pos = den > 0.0
    return torch.where(pos, G * mask_out / torch.where(pos, den, 1.0), 0.0)


def avg_v_to_u_T(G, mask_v, mask_u):
    ms = south(mask_v)
    me, mes = east(mask_v), east(ms)
    H = _mean_weight(G, mask_v + ms + me + mes, mask_u)
    return mask_v * H + north(ms * H) + west(me * H) + north(west(mes * H))


def avg_u_to_v_T(G, mask_u, mask_v):
    mw = west(mask_u)
    mn, mnw = north(mask_u), north(mw)
    H = _mean_weight(G, mask_u + mw + mn + mnw, mask_v)
    return mask_u * H + east(mw * H) + south(mn * H) + south(east(mnw * H))


def laplacian_T(G, dx, dy, mask_u, mask_v, mask):
    gu, gv = div_T(G, dx, dy, mask)
    return gradx_T(gu, dx, mask_u) + grady_T(gv, dy, mask_v)


# ---------------------------------------------------- kappa = 1/3 faces
def superbee(r):
    return nmax(0.0, nmax(nmin(2.0 * r, 1.0), nmin(r, 2.0)))


def kappa_face(c, vel, dim, wet, limited):
    """libs/core/advection_v06.py:21-47 (upwind1 where the far upstream
    cell is dry)."""
    cm = torch.roll(c, 1, dims=dim)
    cp = torch.roll(c, -1, dims=dim)
    cpp = torch.roll(c, -2, dims=dim)
    if limited:
        dc = cp - c
        zero = dc == 0.0
        eps = torch.where(zero, 1.0, dc)
        rp = torch.where(zero, 0.0, (c - cm) / eps)
        rn = torch.where(zero, 0.0, (cpp - cp) / eps)
        fp = c + 0.5 * superbee(rp) * dc
        fn = cp - 0.5 * superbee(rn) * dc
    else:
        fp = c + 0.25 * ((1.0 - KAPPA) * (c - cm) + (1.0 + KAPPA) * (cp - c))
        fn = cp - 0.25 * ((1.0 - KAPPA) * (cpp - cp) + (1.0 + KAPPA) * (cp - c))
    fp = torch.where(torch.roll(wet, 1, dims=dim) > 0.0, fp, c)
    fn = torch.where(torch.roll(wet, -2, dims=dim) > 0.0, fn, cp)
    return torch.where(vel > 0.0, fp, fn)


def kappa_face_z(c, w, wet3, limited):
    """Interior interfaces 1..nz-1 (libs/core/advection_v06.py:50-74)."""
    above, below = c[:-1], c[1:]
    zc = torch.zeros_like(c[:1])
    far_up = torch.cat([zc, c[:-2]], dim=0)
    far_dn = torch.cat([c[2:], zc], dim=0)
    zw = torch.zeros_like(wet3[:1])
    has_up = torch.cat([zw, wet3[:-2]], dim=0) > 0.0
    has_dn = torch.cat([wet3[2:], zw], dim=0) > 0.0
    d = above - below
    if limited:
        zero = d == 0.0
        eps = torch.where(zero, 1.0, d)
        r_up = torch.where(zero, 0.0, (below - far_dn) / eps)
        r_dn = torch.where(zero, 0.0, (far_up - above) / eps)
        f_up = below + 0.5 * superbee(r_up) * d
        f_dn = above - 0.5 * superbee(r_dn) * d
    else:
        f_up = below + 0.25 * ((1.0 - KAPPA) * (below - far_dn) + (1.0 + KAPPA) * (above - below))
        f_dn = above - 0.25 * ((1.0 - KAPPA) * (far_up - above) + (1.0 + KAPPA) * (above - below))
    f_up = torch.where(has_dn, f_up, below)
    f_dn = torch.where(has_up, f_dn, above)
    nz = c.shape[0]
    return torch.where(w[1:nz] > 0.0, f_up, f_dn)


def momentum_advection_up3(u, v, w, dx, dy, dz, mask3u, mask3v, limited):
    """libs/core/advection_v06.py:77-108 (vertical flux over the nominal dz)."""
    nz = u.shape[0]
    ub = 0.5 * (west(u) + u)
    u_c = west(kappa_face(u, east(ub), -1, mask3u, limited))
    ax = (east(ub * u_c) - ub * u_c) / dx
    vb = 0.5 * (v + east(v))
    fy = vb * kappa_face(u, vb, -2, mask3u, limited)
    ay = (fy - south(fy)) / dy
    wb = 0.5 * (w + east(w))
    zero = torch.zeros_like(wb[:1])
    fz = torch.cat([zero, wb[1:nz] * kappa_face_z(u, wb, mask3u, limited), zero], dim=0)
    az = (fz[:nz] - fz[1:]) / dz
    vbc = 0.5 * (south(v) + v)
    v_c = south(kappa_face(v, north(vbc), -2, mask3v, limited))
    by = (north(vbc * v_c) - vbc * v_c) / dy
    ubr = 0.5 * (u + north(u))
    fx = ubr * kappa_face(v, ubr, -1, mask3v, limited)
    bx = (fx - west(fx)) / dx
    wbv = 0.5 * (w + north(w))
    gz = torch.cat([zero, wbv[1:nz] * kappa_face_z(v, wbv, mask3v, limited), zero], dim=0)
    bz = (gz[:nz] - gz[1:]) / dz
    return ax + ay + az, bx + by + bz


# ------------------------------------------------------------------ EOS
_SEOS = dict(a0=1.6550e-1, b0=7.6554e-1, lam1=5.9520e-2, lam2=7.4914e-4,
             mu1=1.4970e-4, mu2=1.1090e-5, nu_ts=2.4341e-3)
_R0 = (4.6494977072e+01, -5.2099962525e+00, 2.2601900708e-01, 6.4326772569e-02,
       1.5616995503e-02, -1.7243708991e-03)
# polyTEOS10-bsq coefficients of ss^a tt^b zh^c (Roquet et al. 2015, Table 4)
_E = {
    (0, 0, 0): 8.0189615746e+02, (1, 0, 0): 8.6672408165e+02, (2, 0, 0): -1.7864682637e+03,
    (3, 0, 0): 2.0375295546e+03, (4, 0, 0): -1.2849161071e+03, (5, 0, 0): 4.3227585684e+02,
    (6, 0, 0): -6.0579916612e+01, (0, 1, 0): 2.6010145068e+01, (1, 1, 0): -6.5281885265e+01,
    (2, 1, 0): 8.1770425108e+01, (3, 1, 0): -5.6888046321e+01, (4, 1, 0): 1.7681814114e+01,
    (5, 1, 0): -1.9193502195e+00, (0, 2, 0): -3.7074170417e+01, (1, 2, 0): 6.1548258127e+01,
    (2, 2, 0): -6.0362551501e+01, (3, 2, 0): 2.9130021253e+01, (4, 2, 0): -5.4723692739e+00,
    (0, 3, 0): 2.1661789529e+01, (1, 3, 0): -3.3449108469e+01, (2, 3, 0): 1.9717078466e+01,
    (3, 3, 0): -3.1742946532e+00, (0, 4, 0): -8.3627885467e+00, (1, 4, 0): 1.1311538584e+01,
    (2, 4, 0): -5.3563304045e+00, (0, 5, 0): 5.4048723791e-01, (1, 5, 0): 4.5111434961e-01,
    (0, 6, 0): -1.9098268277e-01, (0, 0, 1): 1.9681925209e+01, (1, 0, 1): -4.2549998214e+01,
    (2, 0, 1): 5.0774768218e+01, (3, 0, 1): -3.0938076334e+01, (4, 0, 1): 6.6051753097e+00,
    (0, 1, 1): -1.3336301113e+01, (1, 1, 1): -4.4870114575e+00, (2, 1, 1): 5.0042598061e+00,
    (3, 1, 1): -6.5399043664e-01, (0, 2, 1): 6.7080479603e+00, (1, 2, 1): 3.5063081279e+00,
    (2, 2, 1): -1.8795372996e+00, (0, 3, 1): -2.4649669534e+00, (1, 3, 1): -5.5077101279e-01,
    (0, 4, 1): 5.5927935970e-01, (0, 0, 2): 2.0660924175e+00, (1, 0, 2): -4.9527603989e+00,
    (2, 0, 2): 2.5019633803e+00, (0, 1, 2): 2.0564311499e+00, (1, 1, 2): -2.1311365518e-01,
    (0, 2, 2): -1.2419983026e+00, (0, 0, 3): -2.3342758797e-02, (1, 0, 3): -1.8507636718e-02,
    (0, 1, 3): 3.7969820455e-01,
}


def _horner_ss(tt_pow: int, zpow: int, ss):
    """sum_a E[a, tt_pow, zpow] ss^a in Horner form."""
    deg = max(a for (a, b, c) in _E if b == tt_pow and c == zpow)
    acc = _E[(deg, tt_pow, zpow)]
    for a in range(deg - 1, -1, -1):
        acc = acc * ss + _E[(a, tt_pow, zpow)]
    return acc


def rho_teos10(CT, SA, z):
    """libs/core/eos.py:68-100 (Horner in tt over polynomials in ss)."""
    zh = z * 1.0e-4
    ss = torch.sqrt((SA + 32.0) * (0.875 / 35.16504))
    tt = CT * (1.0 / 40.0)
    r0 = zh * 0.0 + _R0[5]
    for c in _R0[4::-1]:
        r0 = r0 * zh + c
    r0 = r0 * zh
    out = 0.0
    for zpow in range(3, -1, -1):
        degt = max(b for (a, b, c) in _E if c == zpow)
        rz = _horner_ss(degt, zpow, ss)
        for b in range(degt - 1, -1, -1):
            rz = rz * tt + _horner_ss(b, zpow, ss)
        out = out * zh + rz if zpow < 3 else rz
    return out + r0


def buoyancy(kind, T, S, z, *, g, rho0, alpha_T, beta_S, T0, S0):
    """b = -g rho'/rho0 (libs/core/eos.py:103-133)."""
    if kind == "linear":
        rp = rho0 * (-alpha_T * (T - T0) + beta_S * (S - S0))
    elif kind == "seos":
        c = _SEOS
        Ta, Sa = T - 10.0, S - 35.0
        rp = (-c["a0"] * (1.0 + 0.5 * c["lam1"] * Ta + c["mu1"] * z) * Ta
              + c["b0"] * (1.0 - 0.5 * c["lam2"] * Sa - c["mu2"] * z) * Sa
              - c["nu_ts"] * Ta * Sa)
    elif kind == "teos10":
        rp = rho_teos10(T, S, z) - rho0
    else:
        raise ValueError(f"unknown eos '{kind}' (expected linear|seos|teos10)")
    return -g * rp / rho0


# ------------------------------------------------- column operators (torch)
def thomas_torch(sub, diag, sup, rhs):
    """libs/core/vertical.py:17-44, one level at a time (autograd-friendly)."""
    n = diag.shape[0]
    cs = [sup[0] / diag[0]]
    ds = [rhs[0] / diag[0]]
    for k in range(1, n):
        den = diag[k] - sub[k] * cs[k - 1]
        cs.append(sup[k] / den)
        ds.append((rhs[k] - sub[k] * ds[k - 1]) / den)
    x = [None] * n
    x[n - 1] = ds[n - 1]
    for k in range(n - 2, -1, -1):
        x[k] = ds[k] - cs[k] * x[k + 1]
    return torch.stack(x, dim=0)


def thomas_adjoint_torch(sub, diag, sup, x, g):
    """lambda = A^-T g for the tridiagonal A (sub[0] and sup[-1] unused) and
    the band gradients of -lambda^T (dA) x: diag -lambda_k x_k, sub
    -lambda_k x_{k-1}, sup -lambda_k x_{k+1}."""
    z = torch.zeros_like(diag[:1])
    lam = thomas_torch(torch.cat([z, sup[:-1]], dim=0), diag,
                       torch.cat([sub[1:], z], dim=0), g)
    g_diag = -lam * x
    g_sub = torch.cat([z, -lam[1:] * x[:-1]], dim=0)
    g_sup = torch.cat([-lam[:-1] * x[1:], z], dim=0)
    return lam, g_sub, g_diag, g_sup


class _ThomasHip(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sub, diag, sup, rhs, lib, stream):
        sub, diag, sup, rhs = (t.contiguous() for t in (sub, diag, sup, rhs))
        n = diag.shape[0]
        ncol = diag.numel() // n
        x = torch.empty_like(rhs)
        scratch = torch.empty_like(rhs)
        _lib.check(lib.hoc_thomas_batched(_ptr(sub), _ptr(diag), _ptr(sup), _ptr(rhs), _ptr(x),
                                          _ptr(scratch), n, ncol, stream()))
        ctx.save_for_backward(sub, diag, sup, x)
        ctx.lib, ctx.stream = lib, stream
        return x

    @staticmethod
    def backward(ctx, g):
        sub, diag, sup, x = ctx.saved_tensors
        g = g.contiguous()
        n = diag.shape[0]
        ncol = diag.numel() // n
        lam, d_sub, d_diag, d_sup, scratch = (torch.empty_like(x) for _ in range(5))
        _lib.check(ctx.lib.hoc_thomas_adjoint(_ptr(sub), _ptr(diag), _ptr(sup), _ptr(x), _ptr(g),
                                              _ptr(lam), _ptr(d_sub), _ptr(d_diag), _ptr(d_sup),
                                              _ptr(scratch), n, ncol, ctx.stream()))
        return d_sub, d_diag, d_sup, lam, None, None


class _GasparHip(torch.autograd.Function):
    """l_up, l_dn of the integral mixing length (closure.py:65-103): HIP
    forward march, hand-written reverse march (hoc_gaspar_lengths_adjoint)."""

    @staticmethod
    def forward(ctx, e, n2, dzi, zs, zb, wet, e_min, z_0, lib, stream):
        ins = [t.contiguous() for t in (e, n2, dzi, zs, zb, wet)]
        nz1 = ins[0].shape[0]
        ncol = ins[0].numel() // nz1
        l_up, l_dn = torch.empty_like(ins[0]), torch.empty_like(ins[0])
        _lib.check(lib.hoc_gaspar_lengths(*(_ptr(t) for t in ins), e_min, z_0, _ptr(l_up),
                                          _ptr(l_dn), nz1, ncol, stream()))
        ctx.save_for_backward(*ins)
        ctx.args = (e_min, z_0, lib, stream)
        return l_up, l_dn

    @staticmethod
    def backward(ctx, g_up, g_dn):
        ins = ctx.saved_tensors
        e_min, z_0, lib, stream = ctx.args
        e = ins[0]
        g_up = torch.zeros_like(e) if g_up is None else g_up.contiguous()
        g_dn = torch.zeros_like(e) if g_dn is None else g_dn.contiguous()
        g_e, g_n2 = torch.empty_like(e), torch.empty_like(e)
        nz1 = e.shape[0]
        _lib.check(lib.hoc_gaspar_lengths_adjoint(*(_ptr(t) for t in ins), e_min, z_0,
                                                  _ptr(g_up), _ptr(g_dn), _ptr(g_e), _ptr(g_n2),
                                                  nz1, e.numel() // nz1, stream()))
        return g_e, g_n2, None, None, None, None, None, None, None, None


class _MomentumAdvectionHip(torch.autograd.Function):
    """(au, av) of the kappa = 1/3 momentum advection: the HIP operator
    forward, the gathered reverse kernel
    (hoc_stepper_momentum_advection_adjoint) backward."""

    @staticmethod
    def forward(ctx, u, v, w, owner):
        u, v, w = u.contiguous(), v.contiguous(), w.contiguous()
        au, av = torch.empty_like(u), torch.empty_like(v)
        _lib.check(owner.lib.hoc_stepper_momentum_advection(owner._h, _ptr(u), _ptr(v), _ptr(w),
                                                            _ptr(au), _ptr(av),
                                                            owner._stream()))
        ctx.save_for_backward(u, v, w)
        ctx.owner = owner
        return au, av

    @staticmethod
    def backward(ctx, g_au, g_av):
        u, v, w = ctx.saved_tensors
        owner = ctx.owner
        g_au = torch.zeros_like(u) if g_au is None else g_au.contiguous()
        g_av = torch.zeros_like(v) if g_av is None else g_av.contiguous()
        gu, gv, gw = torch.empty_like(u), torch.empty_like(v), torch.empty_like(w)
        _lib.check(owner.lib.hoc_stepper_momentum_advection_adjoint(
            owner._h, _ptr(u), _ptr(v), _ptr(w), _ptr(g_au), _ptr(g_av), _ptr(gu), _ptr(gv),
            _ptr(gw), owner._stream()))
        return gu, gv, gw, None


_EOS_KIND = {"linear": 0, "seos": 1, "teos10": 2}


class _EosHip(torch.autograd.Function):
    """b(T, S) at fixed z: the HIP EOS forward, its derivative by dual numbers
    through the same expression (hoc_eos_buoyancy_adjoint)."""

    @staticmethod
    def forward(ctx, T, S, z, phys, lib, stream):
        T, S, z = T.contiguous(), S.contiguous(), z.contiguous()
        b = torch.empty_like(T)
        args = (_EOS_KIND[phys.eos], phys.g, phys.rho0, phys.alpha_T, phys.beta_S, phys.T0,
                phys.S0)
        _lib.check(lib.hoc_eos_buoyancy(*args, _ptr(T), _ptr(S), _ptr(z), _ptr(b), T.numel(),
                                        stream()))
        ctx.save_for_backward(T, S, z)
        ctx.args = (args, lib, stream)
        return b

    @staticmethod
    def backward(ctx, g):
        T, S, z = ctx.saved_tensors
        args, lib, stream = ctx.args
        g = g.contiguous()
        gT, gS = torch.empty_like(T), torch.empty_like(S)
        _lib.check(lib.hoc_eos_buoyancy_adjoint(*args, _ptr(T), _ptr(S), _ptr(z), _ptr(g),
                                                _ptr(gT), _ptr(gS), T.numel(), stream()))
        return gT, gS, None, None, None, None


class _HelmholtzHip(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Ku, Kv, rhs, op):
        Ku, Kv, rhs = Ku.contiguous(), Kv.contiguous(), rhs.contiguous()
        x = op.solve(Ku, Kv, rhs)
        ctx.save_for_backward(Ku, Kv, x)
        ctx.op = op
        return x

    @staticmethod
    def backward(ctx, g):
        Ku, Kv, x = ctx.saved_tensors
        op = ctx.op
        lam = op.solve(Ku, Kv, g.contiguous())
        dKu, dKv = op.coefficient_gradients(x, lam)
        return dKu, dKv, lam, None


class _TracerAdvectionHip(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, u, v, w, owner):
        c, u, v, w = (t.contiguous() for t in (c, u, v, w))
        out = torch.empty_like(c)
        lib, h = owner.lib, owner._h
        _lib.check(lib.hoc_stepper_tracer_advection(h, _ptr(c), _ptr(u), _ptr(v), _ptr(w),
                                                    _ptr(out), owner._stream()))
        ctx.save_for_backward(c, u, v, w)
        ctx.owner = owner
        return out

    @staticmethod
    def backward(ctx, g):
        c, u, v, w = ctx.saved_tensors
        owner = ctx.owner
        g = g.contiguous()
        gc, gu, gv = torch.empty_like(c), torch.empty_like(u), torch.empty_like(v)
        gw = torch.empty_like(w)
        _lib.check(owner.lib.hoc_stepper_tracer_advection_adjoint(
            owner._h, _ptr(c), _ptr(u), _ptr(v), _ptr(w), _ptr(g), _ptr(gc), _ptr(gu), _ptr(gv),
            _ptr(gw), owner._stream()))
        return gc, gu, gv, gw, None


class _HelmholtzOperator:
    """A = I - coef D[K G .] on the masked grid (libs/core/solvers_var.py:22-60)
    and its solves: the stepper's PCG (kernels="hip") or a dense solve over
    the wet cells (kernels="torch")."""

    def __init__(self, owner, mask, coef, dx, dy):
        self.owner, self.mask, self.coef, self.dx, self.dy = owner, mask, coef, dx, dy
        self.iterations = 0

    def apply(self, x, Ku, Kv):
        x = x * self.mask
        fx = Ku * (east(x) - x)
        fy = Kv * (north(x) - x)
        lap = (fx - west(fx)) / self.dx ** 2 + (fy - south(fy)) / self.dy ** 2
        return (x - self.coef * lap) * self.mask + x * (1.0 - self.mask)

    def solve(self, Ku, Kv, rhs):
        st = self.owner
        x = torch.empty_like(rhs)
        it, cv = C.c_int32(0), C.c_int32(0)
        _lib.check(st.lib.hoc_stepper_helmholtz_solve(st._h, _ptr(Ku), _ptr(Kv), _ptr(rhs),
                                                      _ptr(x), C.byref(it), C.byref(cv),
                                                      st._stream()))
        self.iterations += it.value
        return x

    def coefficient_gradients(self, x, lam):
        """dL/dK = -lambda^T (dA/dK) x for the face coefficients."""
        xm, lm = x * self.mask, lam * self.mask
        dKu = self.coef * (east(xm) - xm) * (lm - east(lm)) / self.dx ** 2
        dKv = self.coef * (north(xm) - xm) * (lm - north(lm)) / self.dy ** 2
        return dKu, dKv

    def solve_dense(self, Ku, Kv, rhs):
        ny, nx = rhs.shape
        eye = torch.eye(ny * nx, dtype=F64, device=rhs.device).reshape(ny * nx, ny, nx)
        A = self.apply(eye, Ku, Kv).reshape(ny * nx, ny * nx).T
        wet = (self.mask.reshape(-1) > 0).nonzero().squeeze(1)
        Aw = A[wet][:, wet]
        b = (rhs * self.mask).reshape(-1)[wet]
        xw = torch.linalg.solve(Aw, b)
        return torch.zeros(ny * nx, dtype=F64, device=rhs.device).index_put(
            (wet,), xw).reshape(ny, nx)


# ------------------------------------------------------------ the step
@dataclass
class DiffState:
    """State carried from step to step, with the mixing coefficients the
    next step starts from (the stepper's nu / kappa, fact of the reference's
    Stepper3DV05 carrying them, model3d_v05.py:97-98,181)."""
    u: torch.Tensor
    v: torch.Tensor
    b: torch.Tensor
    eta: torch.Tensor
    T: torch.Tensor
    S: torch.Tensor
    e: torch.Tensor | None
    Km: torch.Tensor
    Kh: torch.Tensor

    FIELDS = ("u", "v", "b", "eta", "T", "S", "e", "Km", "Kh")

    def tensors(self):
        return [getattr(self, k) for k in self.FIELDS if getattr(self, k) is not None]


class DiffStep:
    """The theta step as a differentiable function of (state, kappa_b)."""

    def __init__(self, domain, physics, params, dt, device="cuda:0", kernels="hip"):
        if params.name != "theta":
            raise ValueError("the adjoint covers the theta scheme")
        if params.advection not in ("up3", "up3_tvd"):
            raise ValueError("the adjoint covers advection up3 and up3_tvd")
        if getattr(domain, "vcoord", "zlevel") != "zlevel" or "open" in (domain.bc_x, domain.bc_y):
            raise ValueError("the adjoint covers z-level grids with closed or periodic edges")
        if physics.tracers != "TS":
            raise ValueError("the adjoint covers prognostic T and S")
        if physics.closure not in ("tke", "none"):
            raise ValueError("closure must be tke or none")
        if kernels not in ("hip", "torch"):
            raise ValueError("kernels must be hip or torch")
        self.kernels = kernels
        self.domain, self.physics, self.params, self.dt = domain, physics, params, float(dt)
        self.device = torch.device(device)
        d, g = domain, domain.grid
        self.grid = g
        dev = dict(dtype=F64, device=self.device)
        t = lambda a: torch.as_tensor(np.array(a, dtype=np.float64, order="C"), **dev)
        dz3 = d.layer_thickness(np.zeros(g.shape))
        dzu, dzv = d.face_thickness(dz3)
        self.dz3, self.dz3u, self.dz3v = t(dz3), t(dzu), t(dzv)
        self.zc = t(d.depth_at_centres())
        self.mask, self.mask_u, self.mask_v = t(d.mask), t(d.mask_u), t(d.mask_v)
        self.mask3, self.mask3u, self.mask3v = t(d.mask3), t(d.mask3u), t(d.mask3v)
        below = torch.cat([self.mask3[1:], torch.zeros_like(self.mask3[:1])], dim=0)
        self.bottom = (self.mask3 > 0) & (below == 0)
        self.H = t(d.H)
        p = physics
        self.tau_u = t(np.broadcast_to(np.asarray(p.tau_x) / p.rho0, g.shape)) * self.mask3u[0]
        self.tau_v = t(np.broadcast_to(np.asarray(p.tau_y) / p.rho0, g.shape)) * self.mask3v[0]
        self.ustar2 = t(np.broadcast_to(np.hypot(p.tau_x, p.tau_y) / p.rho0, g.shape))
        coef = p.g * params.theta ** 2 * self.dt ** 2
        self.owner = None
        if kernels == "hip":
            from .stepper import HipStepper
            # the handle of the reverse kernels and of the adjoint Helmholtz
            # solves; with pcg_mg those use the red-black Gauss-Seidel
            # smoother (9 iterations to the same tolerance where damped Jacobi
            # takes 10-11 on the 256^2 x 30 basin: any convergent
            # preconditioner gives the adjoint system's solution to rtol)
            rparams = params
            if params.solver == "pcg_mg" and 1 <= getattr(params, "mg_sweeps", 2) <= 3:
                rparams = dataclasses.replace(params, mg_smoother="rbgs")
            self.owner = HipStepper(domain, physics, rparams, dt, device=device)
        self.helm = _HelmholtzOperator(self.owner, self.mask, coef, g.dx, g.dy)

    # ------------------------------------------------------ operators
    def thomas(self, sub, diag, sup, rhs):
        if self.kernels == "hip":
            st = self.owner
            return _ThomasHip.apply(sub, diag, sup, rhs, st.lib, st._stream)
        return thomas_torch(sub, diag, sup, rhs)

    def helmholtz(self, Ku, Kv, rhs):
        if self.kernels == "hip":
            return _HelmholtzHip.apply(Ku, Kv, rhs, self.helm)
        return self.helm.solve_dense(Ku, Kv, rhs)

    def tracer_advection(self, c, u, v, w):
        """libs/core/model3d_v05.py:482-512"""
        if self.kernels == "hip":
            return _TracerAdvectionHip.apply(c, u, v, w, self.owner)
        g = self.grid
        lim = self.params.advection == "up3_tvd"
        cx = kappa_face(c, u, -1, self.mask3, lim)
        cy = kappa_face(c, v, -2, self.mask3, lim)
        fx = self.dz3u * u * cx
        fy = self.dz3v * v * cy
        inv = safe_inv(self.dz3)
        hdiv = ((fx - west(fx)) / g.dx + (fy - south(fy)) / g.dy) * inv
        nz = c.shape[0]
        cf = torch.cat([c[:1], kappa_face_z(c, w, self.mask3, lim),
                        torch.zeros_like(c[:1])], dim=0)
        vdiv = (w[:nz] * cf[:nz] - w[1:] * cf[1:]) * inv
        return (hdiv + vdiv) * self.mask3

    @staticmethod
    def interface_thickness(dz3):
        """libs/core/vertical_var.py:21-32"""
        return torch.cat([dz3[:1], 0.5 * (dz3[:-1] + dz3[1:]), dz3[-1:]], dim=0)

    def diffusion_coeffs(self, nu, dz3, theta_v, drag):
        """libs/core/vertical_var.py:35-77"""
        dt = self.dt
        nz = nu.shape[0] - 1
        dzi = self.interface_thickness(dz3)
        inv_dz = safe_inv(dz3)
        fac = theta_v * dt * inv_dz
        a_top = fac * nu[:nz] * safe_inv(dzi[:nz])
        a_bot = fac * nu[1:] * safe_inv(dzi[1:])
        sub = torch.cat([torch.zeros_like(a_top[:1]), -a_top[1:]], dim=0)
        sup = -a_bot
        diag = torch.cat([1.0 + a_bot[:1], (1.0 + a_top + a_bot)[1:]], dim=0)
        diag = torch.where(self.bottom, 1.0 + a_top + theta_v * dt * drag * inv_dz, diag)
        sup = torch.where(self.bottom, 0.0, sup)
        dry = self.mask3 == 0
        diag = torch.where(dry, 1.0, diag)
        sub = torch.where(dry, 0.0, sub)
        sup = torch.where(dry, 0.0, sup)
        return sub, diag, sup

    def apply_diffusion(self, field, nu, dz3, surface_flux, drag):
        """libs/core/vertical_var.py:80-106 (mask3 branch)"""
        nz = field.shape[0]
        dzi = self.interface_thickness(dz3)
        inner = nu[1:nz] * (field[:-1] - field[1:]) * safe_inv(dzi[1:nz])
        inner = torch.where(self.bottom[:-1], 0.0, inner)
        flux = torch.cat([surface_flux[None].expand_as(field[:1]), inner,
                          torch.zeros_like(field[:1])], dim=0)
        dragt = drag * field * self.bottom
        inv = safe_inv(dz3)
        out = (flux[:nz] - flux[1:]) * inv
        return (out - dragt * inv) * (self.mask3 > 0)

    def w_diag(self, u, v):
        """libs/core/model3d_v05.py:232-235, vertical_var.py:109-121"""
        g = self.grid
        tdiv = div(self.dz3u * u, self.dz3v * v, g.dx, g.dy, self.mask3)
        w = -torch.flip(torch.cumsum(torch.flip(tdiv, [0]), dim=0), [0])
        return torch.cat([w, torch.zeros_like(w[:1])], dim=0)

    def pressure_gradients(self, b):
        """libs/core/model3d_v05.py:195-230"""
        g, p = self.grid, self.physics
        dz3 = self.dz3
        above = torch.cat([torch.zeros_like(b[:1]), torch.cumsum((dz3 * b)[:-1], dim=0)], dim=0)
        phi = 0.5 * dz3 * b + above
        if p.pgf == "remove":
            total = torch.sum(dz3, dim=0, keepdim=True)
            phi = phi - torch.sum(dz3 * phi, dim=0, keepdim=True) * safe_inv(total)
        if not p.pgf_correction:
            return (gradx(phi, g.dx, self.mask3u) * self.mask3u,
                    grady(phi, g.dy, self.mask3v) * self.mask3v)
        zc = self.zc
        out = []
        for dim, fmask, delta in ((-1, self.mask3u, g.dx), (-2, self.mask3v, g.dy)):
            phi_r = torch.roll(phi, -1, dims=dim)
            b_r = torch.roll(b, -1, dims=dim)
            z_r = torch.roll(zc, -1, dims=dim)
            z_face = 0.5 * (zc + z_r)
            left = phi + b * (z_face - zc)
            right = phi_r + b_r * (z_face - z_r)
            out.append(fmask * (right - left) / delta)
        return out[0], out[1]

    def buoyancy(self, T, S):
        p = self.physics
        if self.kernels == "hip":
            st = self.owner
            return _EosHip.apply(T, S, self.zc, p, st.lib, st._stream)
        return buoyancy(p.eos, T, S, self.zc, g=p.g, rho0=p.rho0, alpha_T=p.alpha_T,
                        beta_S=p.beta_S, T0=p.T0, S0=p.S0)

    # ------------------------------------------------------------ closure
    @staticmethod
    def _march(l, budget, n2_j, dz_j):
        """libs/core/closure.py:89-96"""
        cost = n2_j * dz_j * (l + 0.5 * dz_j)
        full = cost <= budget
        a = torch.where(n2_j > 0.0, n2_j, 1.0)
        s_part = ssqrt(nmax(l * l + 2.0 * budget / a, 0.0)) - l
        clipped = nmin(nmax(s_part, 0.0), dz_j.expand_as(s_part))
        step = torch.where(full, dz_j.expand_as(l), torch.where(n2_j > 0.0, clipped, dz_j))
        return l + torch.where(budget > 0.0, step, 0.0), torch.where(full, budget - cost, 0.0)

    def tke_step(self, e, u, v, b, Km_old, Kh_old, kb):
        """libs/core/closure.py:141-205 (integral or recursive lengths)."""
        p, dt = self.physics, self.dt
        dz3, mask3 = self.dz3, self.mask3
        nz = dz3.shape[0]
        zi = torch.cat([torch.zeros_like(dz3[:1]), torch.cumsum(dz3, dim=0)], dim=0)
        z_b = nmax(zi[-1][None] - zi, 0.0)
        z_s = zi
        wet = torch.cat([torch.zeros_like(zi[:1]), mask3[:-1] * mask3[1:],
                         torch.zeros_like(zi[:1])], dim=0)
        dzi = self.interface_thickness(dz3)
        inv_dzi = safe_inv(dzi)
        uc = 0.5 * (u + west(u))
        vc = 0.5 * (v + south(v))
        zero1 = torch.zeros_like(e[:1])
        du = torch.cat([zero1, (uc[:-1] - uc[1:]) * inv_dzi[1:nz], zero1], dim=0)
        dv = torch.cat([zero1, (vc[:-1] - vc[1:]) * inv_dzi[1:nz], zero1], dim=0)
        db = torch.cat([zero1, (b[:-1] - b[1:]) * inv_dzi[1:nz], zero1], dim=0)
        shear2 = (du ** 2 + dv ** 2) * wet
        n2 = db * wet
        nz1 = nz + 1
        e_min = float(p.e_min)
        if p.mxl == "recursive":
            n2c = nmax(n2, float(p.n2_min))
            l = ssqrt(2.0 * nmax(e, e_min) / n2c)
            l = nmin(l, nmin(z_s + p.z_0, z_b + p.z_0))
            rows = [l[0]]
            for k in range(1, nz1):
                rows.append(nmin(l[k], rows[k - 1] + dzi[k]))
            down = [None] * nz1
            down[nz1 - 1] = rows[nz1 - 1]
            for k in range(nz1 - 2, -1, -1):
                down[k] = nmin(rows[k], down[k + 1] + dzi[k + 1])
            l = torch.stack(down, dim=0) * wet
            l_up = l_dn = l
        elif self.kernels == "hip" and nz1 <= 128:
            st = self.owner
            l_up, l_dn = _GasparHip.apply(e, n2, dzi, z_s, z_b, wet, float(p.e_min),
                                          float(p.z_0), st.lib, st._stream)
        else:
            n2p = nmax(n2, 0.0)
            e_k = nmax(e, e_min)
            l_up = torch.zeros_like(e)
            l_dn = torch.zeros_like(e)
            b_up, b_dn = e_k, e_k
            for j in range(1, nz1):
                lu, bu = self._march(l_up[j:], b_up[j:], n2p[:nz1 - j], dzi[:nz1 - j])
                l_up = torch.cat([l_up[:j], lu], dim=0)
                b_up = torch.cat([b_up[:j], bu], dim=0)
                ld, bd = self._march(l_dn[:nz1 - j], b_dn[:nz1 - j], n2p[j:], dzi[j:])
                l_dn = torch.cat([ld, l_dn[nz1 - j:]], dim=0)
                b_dn = torch.cat([bd, b_dn[nz1 - j:]], dim=0)
            l_up = nmin(l_up, z_s + p.z_0) * wet
            l_dn = nmin(l_dn, z_b + p.z_0) * wet
        l_min = float(p.l_min)
        l_k = nmax(l_min, ssqrt(l_up * l_dn))
        l_eps = nmax(l_min, nmin(l_up, l_dn))
        nu_b = p.nu
        kappa_b = kb[None]
        prod = (Km_old - nu_b) * shear2 - (Kh_old - kappa_b) * n2
        diss = p.c_eps * ssqrt(nmax(e, e_min)) / l_eps
        K_lay = 0.5 * (Km_old[:-1] + Km_old[1:])
        inv_dz3 = safe_inv(dz3)
        a_up = torch.cat([zero1, dt * K_lay[:-1] * inv_dz3[:-1] * inv_dzi[1:nz], zero1], dim=0)
        a_dn = torch.cat([zero1, dt * K_lay[1:] * inv_dz3[1:] * inv_dzi[1:nz], zero1], dim=0)
        sub, sup = -a_up, -a_dn
        diag = 1.0 + a_up + a_dn + dt * diss
        rhs = e + dt * prod
        bnd = wet <= 0.0
        diag = torch.where(bnd, 1.0, diag)
        sub = torch.where(bnd, 0.0, sub)
        sup = torch.where(bnd, 0.0, sup)
        rhs = torch.where(bnd, p.e_min, rhs)
        top = torch.where(mask3[0] > 0.0, p.e_bb * self.ustar2, p.e_min)
        rhs = torch.cat([top[None], rhs[1:]], dim=0)
        e_new = self.thomas(sub, diag, sup, rhs)
        e_new = nmax(e_new, e_min)
        e_new = torch.where(wet > 0.0, e_new, p.e_min)
        e_new = torch.cat([rhs[:1], e_new[1:]], dim=0)
        K_m = p.c_k * l_k * ssqrt(e_new) * wet + nu_b
        K_h = (K_m - nu_b) / p.pr_t + kappa_b
        return e_new, K_m, K_h

    # --------------------------------------------------------------- step
    def step(self, s: DiffState, kb: torch.Tensor, tape: dict | None = None) -> DiffState:
        """libs/core/model3d_v05.py:337-415, theta branch. With a dict `tape`
        the intermediates the reverse step reads are stored in it (q, Ku, Kv,
        w of the old velocity, the masked predictor solutions and the free
        surface of every Picard pass, w of the new velocity)."""
        d, p, g, pr = self.domain, self.physics, self.grid, self.params
        dt, th, tc, thv = self.dt, pr.theta, pr.theta_cor, pr.theta_v
        m3, m3u, m3v = self.mask3, self.mask3u, self.mask3v
        if p.closure == "tke":
            e_new, Km, Kh = self.tke_step(s.e, s.u, s.v, s.b, s.Km, s.Kh, kb)
        else:
            e_new, Km, Kh = s.e, s.Km, s.Kh
        tri_m = self.diffusion_coeffs(Km, self.dz3, thv, p.bottom_drag)
        tri_b = self.diffusion_coeffs(Kh, self.dz3, thv, 0.0)
        q = self.thomas(*tri_m, torch.ones_like(s.u))
        qu = nmin(q, east(q))
        qv = nmin(q, north(q))
        Ku = torch.sum(self.dz3u * qu, dim=0) * self.mask_u
        Kv = torch.sum(self.dz3v * qv, dim=0) * self.mask_v
        if tape is not None:
            tape.update(q=q, Ku=Ku, Kv=Kv, gh_u=[], gh_v=[], eta=[])

        pgx, pgy = self.pressure_gradients(s.b)
        av_prev_v = avg_v_to_u(s.v, m3v, m3u)
        av_prev_u = avg_u_to_v(s.u, m3u, m3v)
        dzu_diff = self.apply_diffusion(s.u, Km, self.dz3u, self.tau_u, p.bottom_drag)
        dzv_diff = self.apply_diffusion(s.v, Km, self.dz3v, self.tau_v, p.bottom_drag)
        w_old = self.w_diag(s.u, s.v)
        if tape is not None:
            tape["w_old"] = w_old
        if self.kernels == "hip":
            au, av = _MomentumAdvectionHip.apply(s.u, s.v, w_old, self.owner)
        else:
            au, av = momentum_advection_up3(s.u, s.v, w_old, g.dx, g.dy, g.dz, m3u, m3v,
                                            pr.advection == "up3_tvd")
        ex_u = -(au * m3u)
        ex_v = -(av * m3v)
        if p.A_h > 0.0:
            ex_u = ex_u + p.A_h * laplacian(s.u, g.dx, g.dy, m3u, m3v, m3u)
            ex_v = ex_v + p.A_h * laplacian(s.v, g.dx, g.dy, m3u, m3v, m3v)
        zero_rest = torch.zeros_like(s.u[1:])
        stress_u = torch.cat([(thv * dt * self.tau_u * safe_inv(self.dz3u[0]))[None], zero_rest])
        stress_v = torch.cat([(thv * dt * self.tau_v * safe_inv(self.dz3v[0]))[None], zero_rest])
        stress_u = stress_u + dt * ex_u
        stress_v = stress_v + dt * ex_v
        gx_eta = gradx(s.eta, g.dx, self.mask_u)[None] * m3u
        gy_eta = grady(s.eta, g.dy, self.mask_v)[None] * m3v
        U_old = torch.sum(self.dz3u * s.u, dim=0)
        V_old = torch.sum(self.dz3v * s.v, dim=0)
        div_old = div(U_old, V_old, g.dx, g.dy, self.mask)
        u_it, v_it, eta_new = s.u, s.v, s.eta
        for _ in range(pr.n_picard):
            v_cor = (1.0 - tc) * av_prev_v + tc * avg_v_to_u(v_it, m3v, m3u)
            u_cor = (1.0 - tc) * av_prev_u + tc * avg_u_to_v(u_it, m3u, m3v)
            gu = (s.u + dt * (p.f * v_cor + pgx) - p.g * dt * (1.0 - th) * gx_eta
                  + (1.0 - thv) * dt * dzu_diff)
            gv = (s.v + dt * (-p.f * u_cor + pgy) - p.g * dt * (1.0 - th) * gy_eta
                  + (1.0 - thv) * dt * dzv_diff)
            gh_u = self.thomas(*tri_m, gu + stress_u) * m3
            gh_v = self.thomas(*tri_m, gv + stress_v) * m3
            div_g = div(torch.sum(self.dz3u * gh_u, dim=0), torch.sum(self.dz3v * gh_v, dim=0),
                        g.dx, g.dy, self.mask)
            rhs = s.eta - dt * ((1.0 - th) * div_old + th * div_g)
            eta_new = self.helmholtz(Ku, Kv, rhs)
            if tape is not None:
                tape["gh_u"].append(gh_u)
                tape["gh_v"].append(gh_v)
                tape["eta"].append(eta_new)
            u_it = gh_u - p.g * dt * th * q * gradx(eta_new, g.dx, self.mask_u)[None] * m3u
            v_it = gh_v - p.g * dt * th * q * grady(eta_new, g.dy, self.mask_v)[None] * m3v
            u_it, v_it = u_it * m3, v_it * m3
        w = self.w_diag(u_it, v_it)
        if tape is not None:
            tape["w"] = w
        # tracers (model3d_v05.py:466-533)
        flux_T = torch.full(g.shape, p.q_heat / (p.rho0 * p.cp), dtype=F64, device=self.device)
        flux_S = torch.full(g.shape, p.q_salt, dtype=F64, device=self.device)
        T_new = self._advance_tracer(s.T, u_it, v_it, w, flux_T, Kh, tri_b)
        S_new = self._advance_tracer(s.S, u_it, v_it, w, flux_S, Kh, tri_b)
        b_new = self.buoyancy(T_new, S_new) * m3
        return DiffState(u=u_it, v=v_it, b=b_new, eta=eta_new, T=T_new, S=S_new, e=e_new,
                         Km=Km, Kh=Kh)

    def _advance_tracer(self, c, u, v, w, surf_flux, Kh, tri_b):
        """libs/core/model3d_v05.py:466-480"""
        g, p, dt, thv = self.grid, self.physics, self.dt, self.params.theta_v
        tend = -self.tracer_advection(c, u, v, w)
        if p.K_h > 0.0:
            tend = tend + p.K_h * laplacian(c, g.dx, g.dy, self.mask3u, self.mask3v, self.mask3)
        rhs = (c + dt * tend
               + (1.0 - thv) * dt * self.apply_diffusion(c, Kh, self.dz3, surf_flux, 0.0))
        top = rhs[0] + thv * dt * surf_flux * safe_inv(self.dz3[0])
        rhs = torch.cat([top[None], rhs[1:]], dim=0)
        return self.thomas(*tri_b, rhs) * self.mask3

    # ------------------------------------------------ hand-written reverse
    # Transposed operators of the step's pieces. kernels="hip" runs the HIP
    # reverse kernels; kernels="torch" takes the vector-Jacobian products of
    # the torch formulas (the reference the HIP kernels are checked against).
    @staticmethod
    def _vjp(fn, ins, grads):
        with torch.enable_grad():
            xs = [t.detach().requires_grad_() for t in ins]
            outs = fn(*xs)
            outs = outs if isinstance(outs, tuple) else (outs,)
            pairs = [(o, g) for o, g in zip(outs, grads) if g is not None]
            got = torch.autograd.grad([o for o, _ in pairs], xs, [g for _, g in pairs],
                                      allow_unused=True)
        return [torch.zeros_like(x) if gi is None else gi for x, gi in zip(xs, got)]

    def thomas_T(self, tri, x, g):
        """(lambda, d_sub, d_diag, d_sup) of x = A^-1 rhs for the adjoint g of x."""
        sub, diag, sup = (t.contiguous() for t in tri)
        if self.kernels != "hip":
            return thomas_adjoint_torch(sub, diag, sup, x, g)
        st = self.owner
        x, g = x.contiguous(), g.contiguous()
        n = diag.shape[0]
        lam, d_sub, d_diag, d_sup, scratch = (torch.empty_like(x) for _ in range(5))
        _lib.check(st.lib.hoc_thomas_adjoint(_ptr(sub), _ptr(diag), _ptr(sup), _ptr(x), _ptr(g),
                                             _ptr(lam), _ptr(d_sub), _ptr(d_diag), _ptr(d_sup),
                                             _ptr(scratch), n, diag.numel() // n, st._stream()))
        return lam, d_sub, d_diag, d_sup

    def tracer_advection_T(self, c, u, v, w, g):
        if self.kernels != "hip":
            return self._vjp(self.tracer_advection, [c, u, v, w], [g])
        st = self.owner
        c, u, v, w, g = (t.contiguous() for t in (c, u, v, w, g))
        gc, gu, gv, gw = (torch.empty_like(t) for t in (c, u, v, w))
        _lib.check(st.lib.hoc_stepper_tracer_advection_adjoint(
            st._h, _ptr(c), _ptr(u), _ptr(v), _ptr(w), _ptr(g), _ptr(gc), _ptr(gu), _ptr(gv),
            _ptr(gw), st._stream()))
        return gc, gu, gv, gw

    def momentum_advection_T(self, u, v, w, g_au, g_av):
        if self.kernels != "hip":
            g, pr = self.grid, self.params
            fn = lambda a, b, c: momentum_advection_up3(a, b, c, g.dx, g.dy, g.dz, self.mask3u,
                                                        self.mask3v, pr.advection == "up3_tvd")
            return self._vjp(fn, [u, v, w], [g_au, g_av])
        st = self.owner
        u, v, w, g_au, g_av = (t.contiguous() for t in (u, v, w, g_au, g_av))
        gu, gv, gw = torch.empty_like(u), torch.empty_like(v), torch.empty_like(w)
        _lib.check(st.lib.hoc_stepper_momentum_advection_adjoint(
            st._h, _ptr(u), _ptr(v), _ptr(w), _ptr(g_au), _ptr(g_av), _ptr(gu), _ptr(gv),
            _ptr(gw), st._stream()))
        return gu, gv, gw

    def buoyancy_T(self, T, S, g):
        if self.kernels != "hip":
            return self._vjp(self.buoyancy, [T, S], [g])
        p, st = self.physics, self.owner
        T, S, g = T.contiguous(), S.contiguous(), g.contiguous()
        gT, gS = torch.empty_like(T), torch.empty_like(S)
        _lib.check(st.lib.hoc_eos_buoyancy_adjoint(
            _EOS_KIND[p.eos], p.g, p.rho0, p.alpha_T, p.beta_S, p.T0, p.S0, _ptr(T), _ptr(S),
            _ptr(self.zc), _ptr(g), _ptr(gT), _ptr(gS), T.numel(), st._stream()))
        return gT, gS

    def helmholtz_solve(self, Ku, Kv, rhs):
        if self.kernels == "hip":
            return self.helm.solve(Ku.contiguous(), Kv.contiguous(), rhs.contiguous())
        return self.helm.solve_dense(Ku, Kv, rhs)

    def diffusion_coeffs_T(self, g_sub, g_diag, g_sup, dz3, theta_v):
        """Adjoint of diffusion_coeffs with respect to nu (vertical_var.py:35-77):
        the rows' a_top and a_bot terms with the bottom and dry overrides."""
        nz = dz3.shape[0]
        dzi = self.interface_thickness(dz3)
        fac = theta_v * self.dt * safe_inv(dz3)
        wet = self.mask3 > 0.0
        bot = self.bottom
        g_sub = torch.where(wet, g_sub, 0.0)
        g_diag = torch.where(wet, g_diag, 0.0)
        g_sup = torch.where(wet & ~bot, g_sup, 0.0)
        ga_top = torch.cat([torch.where(bot[:1], g_diag[:1], 0.0), g_diag[1:] - g_sub[1:]], dim=0)
        ga_bot = torch.where(bot, 0.0, g_diag - g_sup)
        g_nu = torch.zeros((nz + 1,) + dz3.shape[1:], dtype=dz3.dtype, device=dz3.device)
        g_nu[:nz] += fac * safe_inv(dzi[:nz]) * ga_top
        g_nu[1:] += fac * safe_inv(dzi[1:]) * ga_bot
        return g_nu

    def apply_diffusion_T(self, g_out, field, nu, dz3, drag):
        """Adjoint of apply_diffusion (vertical_var.py:80-106) with respect to
        the field and nu (the surface flux is a constant)."""
        nz = field.shape[0]
        dzi = self.interface_thickness(dz3)
        inv = safe_inv(dz3)
        oi = g_out * (self.mask3 > 0) * inv
        g_inner = torch.where(self.bottom[:-1], 0.0, oi[1:] - oi[:-1])
        c = nu[1:nz] * safe_inv(dzi[1:nz])
        g_f = (-drag) * oi * self.bottom
        g_f = g_f + torch.cat([c * g_inner, torch.zeros_like(field[:1])], dim=0)
        g_f = g_f - torch.cat([torch.zeros_like(field[:1]), c * g_inner], dim=0)
        g_nu = torch.zeros_like(nu)
        g_nu[1:nz] = (field[:-1] - field[1:]) * safe_inv(dzi[1:nz]) * g_inner
        return g_f, g_nu

    def w_diag_T(self, g_w):
        """Adjoint of w_diag: w_k = -sum_{j >= k} div_j, then the transposed
        divergence of the face transports."""
        g = self.grid
        nz = g_w.shape[0] - 1
        g_div = -torch.cumsum(g_w[:nz], dim=0)
        gU, gV = div_T(g_div, g.dx, g.dy, self.mask3)
        return self.dz3u * gU, self.dz3v * gV

    def pressure_gradients_T(self, g_pgx, g_pgy):
        """Adjoint of pressure_gradients with respect to b."""
        g, p = self.grid, self.physics
        dz3 = self.dz3
        if not p.pgf_correction:
            g_phi = (gradx_T(g_pgx * self.mask3u, g.dx, self.mask3u)
                     + grady_T(g_pgy * self.mask3v, g.dy, self.mask3v))
            g_b = torch.zeros_like(g_pgx)
        else:
            zc = self.zc
            g_phi = torch.zeros_like(g_pgx)
            g_b = torch.zeros_like(g_pgx)
            for dim, fmask, delta, G in ((-1, self.mask3u, g.dx, g_pgx),
                                         (-2, self.mask3v, g.dy, g_pgy)):
                P = fmask * G / delta
                z_r = torch.roll(zc, -1, dims=dim)
                z_face = 0.5 * (zc + z_r)
                g_phi = g_phi + torch.roll(P, 1, dims=dim) - P
                g_b = g_b + torch.roll(P * (z_face - z_r), 1, dims=dim) - P * (z_face - zc)
        if p.pgf == "remove":
            total = torch.sum(dz3, dim=0, keepdim=True)
            g_phi = g_phi - dz3 * safe_inv(total) * torch.sum(g_phi, dim=0, keepdim=True)
        below = torch.flip(torch.cumsum(torch.flip(g_phi, [0]), dim=0), [0]) - g_phi
        return g_b + 0.5 * dz3 * g_phi + dz3 * below

    def face_coefficients_T(self, q, g_Ku, g_Kv):
        """Adjoint of K_u = sum dz3u min(q, q_E) mask_u (and K_v) w.r.t. q."""
        out = torch.zeros_like(q)
        for G, dz, fm, fwd, back in ((g_Ku, self.dz3u, self.mask_u, east, west),
                                     (g_Kv, self.dz3v, self.mask_v, north, south)):
            Q = dz * (G * fm)[None]
            keep = (q <= fwd(q)) | (q != q)
            out = out + torch.where(keep, Q, 0.0) + back(torch.where(keep, 0.0, Q))
        return out

    def _tracer_T(self, c, u, v, w, Kh, tri_b, x, g_x):
        """Adjoint of _advance_tracer: (g_c, g_u, g_v, g_w, g_Kh)."""
        g, p, dt, thv = self.grid, self.physics, self.dt, self.params.theta_v
        lam, d_sub, d_diag, d_sup = self.thomas_T(tri_b, x, g_x * self.mask3)
        g_Kh = self.diffusion_coeffs_T(d_sub, d_diag, d_sup, self.dz3, thv)
        g_c, g_u, g_v, g_w = self.tracer_advection_T(c, u, v, w, -dt * lam)
        g_c = g_c + lam
        if p.K_h > 0.0:
            g_c = g_c + dt * p.K_h * laplacian_T(lam, g.dx, g.dy, self.mask3u, self.mask3v,
                                                  self.mask3)
        d_f, d_nu = self.apply_diffusion_T((1.0 - thv) * dt * lam, c, Kh, self.dz3, 0.0)
        return g_c + d_f, g_u, g_v, g_w, g_Kh + d_nu

    def reverse(self, s: DiffState, kb: torch.Tensor, y: DiffState, tape: dict, adj: dict,
                fused: bool | None = None):
        """Hand-written reverse of step(s, kb) -> y: from the adjoints `adj`
        of y's fields (None for none) the adjoints of s's fields and of kb,
        with the forward's intermediates from `tape` (step(..., tape=) or
        hoc_stepper_step_tape). fused (the default with kernels="hip"): the
        step after its closure reversed by hoc_stepper_step_adjoint's
        kernels; otherwise by the transposed operators below. The closure
        (tke_step) is differentiated by autograd either way."""
        if fused is None:
            fused = self.kernels == "hip"
        if fused:
            return self._reverse_fused(s, kb, y, tape, adj)
        return self._reverse_ops(s, kb, y, tape, adj)

    def _closure_T(self, s: DiffState, kb, out: dict, g_e, g_Km, g_Kh, fused: bool = False):
        """The closure's share of the reverse step added into `out`; returns
        dJ/dkappa_b of the step. fused: hoc_stepper_closure_adjoint (TKE with
        integral lengths); otherwise autograd through tke_step."""
        p = self.physics
        if fused and p.closure == "tke" and p.mxl == "integral":
            st = self.owner
            ptr = lambda t: t.data_ptr()
            ins = [t.contiguous() for t in (s.e, s.u, s.v, s.b, s.Km, s.Kh, kb, g_Km, g_Kh)]
            ge = None if g_e is None else g_e.contiguous()
            for k in ("u", "v", "b"):
                out[k] = out[k].contiguous()
            g_e0, g_Km0, g_Kh0 = (torch.empty_like(s.e) for _ in range(3))
            g_kb = torch.zeros_like(kb)
            _lib.check(st.lib.hoc_stepper_closure_adjoint(
                st._h, *(ptr(t) for t in ins[:7]), None if ge is None else ptr(ge),
                ptr(ins[7]), ptr(ins[8]), ptr(g_e0), ptr(out["u"]), ptr(out["v"]),
                ptr(out["b"]), ptr(g_Km0), ptr(g_Kh0), ptr(g_kb), st._stream()))
            out["e"], out["Km"], out["Kh"] = g_e0, g_Km0, g_Kh0
            return g_kb
        if p.closure == "tke":
            ins = [s.e, s.u, s.v, s.b, s.Km, s.Kh, kb]
            got = self._vjp(self.tke_step, ins, [g_e, g_Km, g_Kh])
            for k, gk in zip(("e", "u", "v", "b", "Km", "Kh"), got[:6]):
                out[k] = gk if out.get(k) is None else out[k] + gk
            return got[6]
        out["Km"], out["Kh"], out["e"] = g_Km, g_Kh, g_e
        return torch.zeros_like(kb)

    def _reverse_fused(self, s: DiffState, kb, y: DiffState, tape: dict, adj: dict):
        st = self.owner
        g = self.grid
        opts = dict(dtype=F64, device=self.device)
        gx = {k: torch.empty(g.shape3d, **opts) for k in ("u", "v", "b", "T", "S")}
        gx["eta"] = torch.empty(g.shape, **opts)
        shape_i = (g.nz + 1,) + g.shape
        g_Km, g_Kh = torch.empty(shape_i, **opts), torch.empty(shape_i, **opts)
        ptr = lambda t: None if t is None else t.contiguous().data_ptr()
        hx, hy, hgy, hgx = _lib.HocState(), _lib.HocState(), _lib.HocState(), _lib.HocState()
        keep = []
        for name in ("u", "v", "T", "S"):
            a, b = getattr(s, name).contiguous(), getattr(y, name).contiguous()
            keep += [a, b]
            setattr(hx, name, a.data_ptr())
            setattr(hy, name, b.data_ptr())
        for name in ("u", "v", "b", "eta", "T", "S"):
            t = adj.get(name)
            if t is not None:
                t = t.contiguous()
                keep.append(t)
            setattr(hgy, name, ptr(t))
            setattr(hgx, name, gx[name].data_ptr())
        gy_Km = None if adj.get("Km") is None else adj["Km"].contiguous()
        gy_Kh = None if adj.get("Kh") is None else adj["Kh"].contiguous()
        t = _lib.HocStepTape()
        for k in ("q", "Ku", "Kv", "w_old", "w"):
            setattr(t, k, tape[k].data_ptr())
        for k in ("gh_u", "gh_v", "eta"):
            arr = getattr(t, k)
            for i, buf in enumerate(tape[k]):
                arr[i] = buf.data_ptr()
        Km, Kh = y.Km.contiguous(), y.Kh.contiguous()
        _lib.check(st.lib.hoc_stepper_step_adjoint(
            st._h, C.byref(hx), C.byref(hy), Km.data_ptr(), Kh.data_ptr(), C.byref(t),
            C.byref(hgy), ptr(gy_Km), ptr(gy_Kh), C.byref(hgx), g_Km.data_ptr(),
            g_Kh.data_ptr(), st._stream()))
        out = dict(gx)
        out.update(e=None, Km=None, Kh=None)
        g_kb = self._closure_T(s, kb, out, adj.get("e"), g_Km, g_Kh, fused=True)
        return out, g_kb

    def _reverse_ops(self, s: DiffState, kb: torch.Tensor, y: DiffState, tape: dict,
                     adj: dict):
        d, p, g, pr = self.domain, self.physics, self.grid, self.params
        dt, th, tc, thv = self.dt, pr.theta, pr.theta_cor, pr.theta_v
        m3, m3u, m3v = self.mask3, self.mask3u, self.mask3v
        dx, dy = g.dx, g.dy
        zero3 = torch.zeros_like(s.u)
        A = lambda k, like: adj[k] if adj.get(k) is not None else torch.zeros_like(like)
        Km, Kh, q, Ku, Kv = y.Km, y.Kh, tape["q"], tape["Ku"], tape["Kv"]
        g_Km = A("Km", Km).clone()
        g_Kh = A("Kh", Kh).clone()
        # the EOS of the new tracers
        g_T, g_S = A("T", y.T), A("S", y.S)
        if adj.get("b") is not None:
            dT, dS = self.buoyancy_T(y.T, y.S, adj["b"] * m3)
            g_T, g_S = g_T + dT, g_S + dS
        # tracers (model3d_v05.py:466-533)
        tri_b = self.diffusion_coeffs(Kh, self.dz3, thv, 0.0)
        w = tape["w"]
        gT_in, gu1, gv1, gw, dKh = self._tracer_T(s.T, y.u, y.v, w, Kh, tri_b, y.T, g_T)
        gS_in, gu2, gv2, gw2, dKh2 = self._tracer_T(s.S, y.u, y.v, w, Kh, tri_b, y.S, g_S)
        g_Kh = g_Kh + dKh + dKh2
        gwu, gwv = self.w_diag_T(gw + gw2)
        g_ui = A("u", y.u) + gu1 + gu2 + gwu
        g_vi = A("v", y.v) + gv1 + gv2 + gwv
        # Picard passes in reverse (model3d_v05.py:387-408)
        tri_m = self.diffusion_coeffs(Km, self.dz3, thv, p.bottom_drag)
        cH = p.g * dt * th
        g_u, g_v, g_b = zero3.clone(), zero3.clone(), zero3.clone()
        g_eta = torch.zeros_like(s.eta)
        g_eta_p = A("eta", y.eta)
        g_q = torch.zeros_like(q)
        g_Ku, g_Kv = torch.zeros_like(Ku), torch.zeros_like(Kv)
        g_div_old = torch.zeros_like(s.eta)
        band = [zero3.clone(), zero3.clone(), zero3.clone()]
        L_u, L_v = zero3.clone(), zero3.clone()
        cor_v, cor_u = zero3.clone(), zero3.clone()
        for it in range(pr.n_picard - 1, -1, -1):
            eta_it, ghu, ghv = tape["eta"][it], tape["gh_u"][it], tape["gh_v"][it]
            Au, Av = g_ui * m3, g_vi * m3
            gxe = gradx(eta_it, dx, self.mask_u)[None] * m3u
            gye = grady(eta_it, dy, self.mask_v)[None] * m3v
            g_q = g_q - cH * (gxe * Au + gye * Av)
            X = torch.sum(-cH * q * m3u * Au, dim=0)
            Y = torch.sum(-cH * q * m3v * Av, dim=0)
            ge = g_eta_p + gradx_T(X, dx, self.mask_u) + grady_T(Y, dy, self.mask_v)
            lam_eta = self.helmholtz_solve(Ku, Kv, ge)
            dKu, dKv = self.helm.coefficient_gradients(eta_it, lam_eta)
            g_Ku, g_Kv = g_Ku + dKu, g_Kv + dKv
            g_eta = g_eta + lam_eta
            g_div_old = g_div_old - dt * (1.0 - th) * lam_eta
            GU, GV = div_T(-dt * th * lam_eta, dx, dy, self.mask)
            lu, a1, a2, a3 = self.thomas_T(tri_m, ghu, (Au + self.dz3u * GU[None]) * m3)
            lv, b1, b2, b3 = self.thomas_T(tri_m, ghv, (Av + self.dz3v * GV[None]) * m3)
            band = [band[0] + a1 + b1, band[1] + a2 + b2, band[2] + a3 + b3]
            L_u, L_v = L_u + lu, L_v + lv
            gvc, guc = dt * p.f * lu, -dt * p.f * lv
            cor_v, cor_u = cor_v + gvc, cor_u + guc
            g_eta_p = torch.zeros_like(s.eta)
            if it > 0:   # the previous pass's velocity enters this pass's Coriolis terms
                g_ui = tc * avg_u_to_v_T(guc, m3u, m3v)
                g_vi = tc * avg_v_to_u_T(gvc, m3v, m3u)
            else:
                g_u = g_u + tc * avg_u_to_v_T(guc, m3u, m3v)
                g_v = g_v + tc * avg_v_to_u_T(gvc, m3v, m3u)
        # the step's explicit terms, all linear in the L's
        g_u = g_u + (1.0 - tc) * avg_u_to_v_T(cor_u, m3u, m3v) + L_u
        g_v = g_v + (1.0 - tc) * avg_v_to_u_T(cor_v, m3v, m3u) + L_v
        g_b = g_b + self.pressure_gradients_T(dt * L_u, dt * L_v)
        ge_x = torch.sum(m3u * (-p.g * dt * (1.0 - th)) * L_u, dim=0)
        ge_y = torch.sum(m3v * (-p.g * dt * (1.0 - th)) * L_v, dim=0)
        g_eta = g_eta + gradx_T(ge_x, dx, self.mask_u) + grady_T(ge_y, dy, self.mask_v)
        du, dKm_u = self.apply_diffusion_T((1.0 - thv) * dt * L_u, s.u, Km, self.dz3u,
                                           p.bottom_drag)
        dv, dKm_v = self.apply_diffusion_T((1.0 - thv) * dt * L_v, s.v, Km, self.dz3v,
                                           p.bottom_drag)
        g_u, g_v, g_Km = g_u + du, g_v + dv, g_Km + dKm_u + dKm_v
        gex_u, gex_v = dt * L_u, dt * L_v
        if p.A_h > 0.0:
            g_u = g_u + p.A_h * laplacian_T(gex_u, dx, dy, m3u, m3v, m3u)
            g_v = g_v + p.A_h * laplacian_T(gex_v, dx, dy, m3u, m3v, m3v)
        du, dv, dw = self.momentum_advection_T(s.u, s.v, tape["w_old"], -(m3u * gex_u),
                                               -(m3v * gex_v))
        dwu, dwv = self.w_diag_T(dw)
        g_u, g_v = g_u + du + dwu, g_v + dv + dwv
        GU, GV = div_T(g_div_old, dx, dy, self.mask)
        g_u = g_u + self.dz3u * GU[None]
        g_v = g_v + self.dz3v * GV[None]
        # q = A_m^-1 1 and the face coefficients K_u, K_v (model3d_v05.py:150-167)
        g_q = g_q + self.face_coefficients_T(q, g_Ku, g_Kv)
        _, c1, c2, c3 = self.thomas_T(tri_m, q, g_q)
        band = [band[0] + c1, band[1] + c2, band[2] + c3]
        g_Km = g_Km + self.diffusion_coeffs_T(*band, self.dz3, thv)
        # the closure (closure.py:141-205) by autograd
        out = {"u": g_u, "v": g_v, "b": g_b, "eta": g_eta, "T": gT_in, "S": gS_in,
               "e": None, "Km": None, "Kh": None}
        g_kb = self._closure_T(s, kb, out, adj.get("e"), g_Km, g_Kh)
        return out, g_kb

    # ------------------------------------------------------- initial state
    def initial_state(self, T0, S0, kb) -> DiffState:
        """State at rest built from (T0, S0): b from the EOS, e at e_min, the
        mixing coefficients at their backgrounds nu_b and kappa_b
        (model3d_v05.py:97-98,175)."""
        g, p = self.grid, self.physics
        z3 = torch.zeros(g.shape3d, dtype=F64, device=self.device)
        shape_i = (g.nz + 1,) + g.shape
        e = (torch.full(shape_i, p.e_min, dtype=F64, device=self.device)
             if p.closure == "tke" else None)
        return DiffState(u=z3, v=z3.clone(), b=self.buoyancy(T0, S0) * self.mask3,
                         eta=torch.zeros(g.shape, dtype=F64, device=self.device), T=T0, S=S0,
                         e=e, Km=torch.full(shape_i, float(p.nu), dtype=F64, device=self.device),
                         Kh=kb[None].expand(shape_i) + 0.0)


# -------------------------------------------------------------- objective
def objective(domain, u, v, eta, dz3u, dz3v, dz3, H, mask):
    """J = sum_wet(H eta^2) / sum_wet(H) + 0.5 sum(dz3u u^2 + dz3v v^2) / sum(dz3):
    the depth-weighted mean squared SSH plus half the volume-mean squared
    speed with face thickness (the kinetic-energy weighting of
    BasinSeiche._energy, libs/core/cases3d_v05.py:289-302)."""
    wH = H * mask
    ssh = torch.sum(wH * eta ** 2) / torch.sum(wH)
    ke = 0.5 * torch.sum(dz3u * u ** 2 + dz3v * v ** 2) / torch.sum(dz3)
    return ssh + ke


# ---------------------------------------------- checkpointed gradient
@dataclass
class GradientResult:
    J: float
    dJ_dT0: torch.Tensor
    dJ_dkappa: torch.Tensor
    forward_s: float      # the plain HIP forward of n steps (one timed run)
    gradient_s: float     # forward with checkpoints + reverse sweep
    n_steps: int
    every: int
    solve_budget: int = 0     # sync-free iterations per recompute solve (0: polled)
    adjoint_budget: int = 0   # sync-free iterations per adjoint solve (0: polled)
    redone_polled: bool = False


class Adjoint:
    """Gradient of J(state after n steps) with respect to the initial T and
    the background diffusivity field kappa_b [ny, nx], by a reverse sweep
    over checkpoints: the plain HIP step runs forward keeping the full state
    (u, v, b, eta, T, S, e, K_m, K_h) every `every` steps; the reverse sweep
    recomputes each window's states with the HIP step, then back-propagates
    one step at a time: reverse="hand" (default) recomputes the window with
    hoc_stepper_step_tape and runs DiffStep.reverse on each step's tape;
    reverse="autograd" differentiates the torch restatement of each step."""

    def __init__(self, domain, physics, params, dt, device="cuda:0", kernels="hip",
                 reverse="hand"):
        if reverse not in ("hand", "autograd"):
            raise ValueError("reverse must be hand or autograd")
        self.domain, self.physics, self.params, self.dt = domain, physics, params, float(dt)
        self.device = torch.device(device)
        self.kernels = kernels
        self.reverse = reverse
        # the window recomputes run sync-free with a budget from the forward
        # sweep's iteration counts, the adjoint solves with one from the
        # first reverse step (the polled results when every solve converges
        # within its budget; otherwise the gradient is redone polled)
        self.sync_free_recompute = True
        self.adjoint_budget = 0
        self._work_stream = None
        self._budget_override = 0   # tests: a recompute budget forced from the host
        self.ds = DiffStep(domain, physics, params, dt, device=device, kernels=kernels)

    def _hip(self, kb_np):
        from .stepper import HipStepper
        phys = dataclasses.replace(self.physics,
                                   base=dataclasses.replace(self.physics.base, kappa=kb_np))
        return HipStepper(self.domain, phys, self.params, self.dt, device=self.device)

    @staticmethod
    def _capture(st, s) -> DiffState:
        return DiffState(u=s.u.clone(), v=s.v.clone(), b=s.b.clone(), eta=s.eta.clone(),
                         T=s.T.clone(), S=s.S.clone(),
                         e=None if s.e is None else s.e.clone(),
                         Km=st.nu.clone(), Kh=st.kappa.clone())

    def _restore(self, st, c: DiffState):
        from .types import State3D
        st.nu.copy_(c.Km)
        st.kappa.copy_(c.Kh)
        _lib.check(st.lib.hoc_stepper_build_coefficients(st._h, st._stream()))
        return State3D(u=c.u.clone(), v=c.v.clone(), b=c.b.clone(), eta=c.eta.clone(), t=0.0,
                       T=c.T.clone(), S=c.S.clone(), e=None if c.e is None else c.e.clone())

    @staticmethod
    def _tape_step(st, fin: dict):
        """One HIP step from the device fields fin with its tape: (the state
        after it with the mixing coefficients the step used, the tape)."""
        fout = st._new_state_buffers()
        tape = st._new_tape()
        st._step_tape_into(fin, fout, tape)
        return (DiffState(u=fout["u"], v=fout["v"], b=fout["b"], eta=fout["eta"], T=fout["T"],
                          S=fout["S"], e=fout.get("e"), Km=st.nu.clone(), Kh=st.kappa.clone()),
                tape)

    def _recompute(self, st, c0: DiffState, start: int, stop: int, side, holder: list):
        """Generator: steps start..stop-1 recomputed with their tapes from the
        checkpoint c0 on the stream `side`, one step per resumption; the
        states and tapes go to holder[0]."""
        states, tapes = [c0], []
        holder.append((states, tapes))
        with torch.cuda.stream(side):
            fin = st._inputs(self._restore(st, c0))
        yield
        for _ in range(start, stop):
            with torch.cuda.stream(side):
                y, tape = self._tape_step(st, fin)
            states.append(y)
            tapes.append(tape)
            fin = {k: getattr(y, k) for k in ("u", "v", "b", "eta", "T", "S", "e")}
            yield

    def _reverse_sweep(self, st, ckpt: dict, kept, n_steps: int, every: int, kb, adj: dict,
                       sync_free: bool = False):
        """The windows reversed last to first (the last from the forward
        sweep's tapes `kept`); while a window is reversed on the current
        stream, the one before it is recomputed on a second stream (its own
        handle `st`, step by step between the reverse steps), so the two
        sequences of small kernels share the device. The reverse of a window
        waits for the event that ends its recompute; tensors made on the
        second stream are recorded on the first before use. `adj` becomes
        the adjoint of the state before step 0; returns dJ/dkappa_b. With
        sync_free the adjoint solves after the first reverse step (polled)
        run sync-free with one iteration more than that step's largest
        solve."""
        main = torch.cuda.current_stream(self.device)
        own = self.ds.owner
        calibrate, it_a = sync_free, own.solver_iterations
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(main)
        starts = sorted(ckpt, reverse=True)
        g_kb = torch.zeros_like(kb)
        window, ev = kept, None
        for idx, start in enumerate(starts):
            stop = min(start + every, n_steps)
            if ev is not None:
                main.wait_event(ev)
                for y in window[0][1:]:
                    for t in y.tensors():
                        t.record_stream(main)
                for tape in window[1]:
                    for k, t in tape.items():
                        for tt in (t if isinstance(t, list) else [t]):
                            tt.record_stream(main)
            states, tapes = window
            holder, rec = [], None
            if idx + 1 < len(starts):
                nxt = starts[idx + 1]
                rec = self._recompute(st, ckpt[nxt], nxt, min(nxt + every, n_steps), side,
                                      holder)
                next(rec)
            for i in range(stop - start - 1, -1, -1):
                out, gk = self.ds.reverse(states[i], kb, states[i + 1], tapes[i], adj)
                adj.clear()
                adj.update(out)
                g_kb += gk
                if calibrate:
                    calibrate = False
                    per = (own.solver_iterations - it_a) / max(1, self.params.n_picard)
                    self.adjoint_budget = math.ceil(per) + 1
                    _lib.check(own.lib.hoc_stepper_set_sync_free(own._h, self.adjoint_budget))
                states[i + 1] = tapes[i] = None
                if rec is not None:
                    next(rec, None)
            if rec is not None:
                for _ in rec:
                    pass
                window = holder[0]
                ev = torch.cuda.Event()
                ev.record(side)
        main.wait_stream(side)
        return g_kb

    def J_forward(self, T0, S0, kb, n_steps):
        """J after n steps of the plain HIP step (for finite differences)."""
        ds = self.ds
        st = self._hip(kb.detach().cpu().numpy())
        from .types import State3D
        s0 = ds.initial_state(T0, S0, kb)
        s = State3D(u=s0.u, v=s0.v, b=s0.b, eta=s0.eta, t=0.0, T=s0.T, S=s0.S, e=s0.e)
        s = st.integrate(s, n_steps)
        return float(objective(self.domain, s.u, s.v, s.eta, ds.dz3u, ds.dz3v, ds.dz3, ds.H,
                               ds.mask))

    def gradient(self, T0, S0, kb, n_steps=60, every=10) -> GradientResult:
        """The gradient on a stream of its own (the library replays its
        sync-free solves as HIP graphs, which the legacy default stream
        cannot capture); the current stream waits for nothing else."""
        cur = torch.cuda.current_stream(self.device)
        if cur.cuda_stream != 0:
            return self._gradient(T0, S0, kb, n_steps, every)
        if self._work_stream is None:
            self._work_stream = torch.cuda.Stream(device=self.device)
        work = self._work_stream
        work.wait_stream(cur)
        with torch.cuda.stream(work):
            r = self._gradient(T0, S0, kb, n_steps, every)
        cur.wait_stream(work)
        return r

    def _gradient(self, T0, S0, kb, n_steps=60, every=10) -> GradientResult:
        ds = self.ds
        T0 = T0.detach().clone()
        kb = kb.detach().clone()
        st = self._hip(kb.detach().cpu().numpy())
        from .types import State3D
        # timed plain forward (the adjoint:forward denominator)
        with torch.no_grad():
            s0 = ds.initial_state(T0, S0, kb)
        it_f = st.solver_iterations
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        s = State3D(u=s0.u, v=s0.v, b=s0.b, eta=s0.eta, t=0.0, T=s0.T, S=s0.S, e=s0.e)
        s = st.integrate(s, n_steps)
        torch.cuda.synchronize(self.device)
        forward_s = time.perf_counter() - t0
        n_solves = max(1, n_steps * max(1, self.params.n_picard))
        sync_free = (self.reverse == "hand" and self.sync_free_recompute
                     and self.params.solver in ("pcg_mg", "pcg_rbgs"))
        f0 = st.solver_failures
        if sync_free:  # the checkpointed sweep repeats the timed forward's solves
            fwd_budget = math.ceil((st.solver_iterations - it_f) / n_solves) + 2
            _lib.check(st.lib.hoc_stepper_set_sync_free(st._h, fwd_budget))

        st.reset_mixing()
        t0 = time.perf_counter()
        # forward sweep with checkpoints; with the hand-written reverse the
        # last window's steps keep their tapes (that window needs no recompute)
        ckpt = {}
        last = ((n_steps - 1) // every) * every
        kept = None
        s = State3D(u=s0.u, v=s0.v, b=s0.b, eta=s0.eta, t=0.0, T=s0.T, S=s0.S, e=s0.e)
        # the solves' iterations per window (read at the checkpoints only, so
        # the sweep is not synchronised after every step)
        per_solve, it0 = 0.0, st.solver_iterations
        for n in range(n_steps):
            if n % every == 0:
                if n > 0:
                    it1 = st.solver_iterations
                    per_solve = max(per_solve, (it1 - it0) / (every * max(1, self.params.n_picard)))
                    it0 = it1
                ckpt[n] = self._capture(st, s)
            if self.reverse == "hand" and n >= last:
                if kept is None:
                    kept = ([ckpt[last]], [])
                    fin = st._inputs(s)
                y, tape = self._tape_step(st, fin)
                kept[0].append(y)
                kept[1].append(tape)
                fin = {k: getattr(y, k) for k in ("u", "v", "b", "eta", "T", "S", "e")}
                s = State3D(u=y.u, v=y.v, b=y.b, eta=y.eta, t=0.0, T=y.T, S=y.S, e=y.e)
            else:
                s = st.step(s)
        n_last = n_steps - last
        per_solve = max(per_solve, (st.solver_iterations - it0)
                        / (n_last * max(1, self.params.n_picard)))
        final = self._capture(st, s)
        # adjoint of J
        fin = [final.u.requires_grad_(), final.v.requires_grad_(), final.eta.requires_grad_()]
        J = objective(self.domain, fin[0], fin[1], fin[2], ds.dz3u, ds.dz3v, ds.dz3, ds.H,
                      ds.mask)
        gu, gv, geta = torch.autograd.grad(J, fin)
        adj = {k: None for k in DiffState.FIELDS}
        adj.update(u=gu, v=gv, eta=geta)
        g_kb = torch.zeros_like(kb)
        # reverse sweep, window by window
        budget = 0
        if self.reverse == "hand":
            if sync_free:
                budget = math.ceil(per_solve) + 2
                if self._budget_override:
                    budget = self._budget_override
            own = ds.owner
            fa0 = own.solver_failures
            _lib.check(st.lib.hoc_stepper_set_sync_free(st._h, budget))
            try:
                g_kb += self._reverse_sweep(st, ckpt, kept, n_steps, every, kb, adj,
                                            sync_free=budget > 0)
            finally:
                _lib.check(st.lib.hoc_stepper_set_sync_free(st._h, 0))
                _lib.check(own.lib.hoc_stepper_set_sync_free(own._h, 0))
            if st.solver_failures > f0 or own.solver_failures > fa0:
                # a solve outran its budget: the gradient again with polled solves
                self.sync_free_recompute = False
                try:
                    r = self._gradient(T0, S0, kb, n_steps=n_steps, every=every)
                finally:
                    self.sync_free_recompute = True
                r.redone_polled = True
                return r
        for start in (sorted(ckpt, reverse=True) if self.reverse != "hand" else ()):
            stop = min(start + every, n_steps)
            states = [ckpt[start]]
            if stop - start > 1:
                s = self._restore(st, ckpt[start])
                for n in range(start, stop - 1):
                    s = st.step(s)
                    states.append(self._capture(st, s))
            for n in range(stop - 1, start - 1, -1):
                x = states[n - start]
                xin = DiffState(**{k: (None if getattr(x, k) is None
                                       else getattr(x, k).detach().requires_grad_())
                                   for k in DiffState.FIELDS})
                kbv = kb.detach().requires_grad_()
                y = ds.step(xin, kbv)
                outs, grads = [], []
                for k in DiffState.FIELDS:
                    yk, gk = getattr(y, k), adj[k]
                    if yk is not None and gk is not None and yk.requires_grad:
                        outs.append(yk)
                        grads.append(gk)
                ins = [t for t in xin.tensors()] + [kbv]
                got = torch.autograd.grad(outs, ins, grads, allow_unused=True)
                names = [k for k in DiffState.FIELDS if getattr(xin, k) is not None]
                adj = {k: None for k in DiffState.FIELDS}
                for k, gk in zip(names, got[:-1]):
                    adj[k] = gk
                if got[-1] is not None:
                    g_kb += got[-1]
                del y, outs, grads, got
            del states
        # the initial state's dependence on (T0, kb)
        T0v = T0.requires_grad_()
        kbv = kb.detach().requires_grad_()
        x0 = ds.initial_state(T0v, S0, kbv)
        outs, grads = [], []
        for k in DiffState.FIELDS:
            xk, gk = getattr(x0, k), adj[k]
            if xk is not None and gk is not None and xk.requires_grad:
                outs.append(xk)
                grads.append(gk)
        gT, gk0 = torch.autograd.grad(outs, [T0v, kbv], grads, allow_unused=True)
        g_kb += gk0 if gk0 is not None else 0.0
        torch.cuda.synchronize(self.device)
        gradient_s = time.perf_counter() - t0
        return GradientResult(J=float(J.detach()), dJ_dT0=gT.detach(), dJ_dkappa=g_kb.detach(),
                              forward_s=forward_s, gradient_s=gradient_s, n_steps=n_steps,
                              every=every, solve_budget=budget,
                              adjoint_budget=self.adjoint_budget if budget else 0)


def directional_check(adj: Adjoint, T0, S0, kb, result: GradientResult, n_dirs: int = 16,
                      h: float = 1e-6, seed: int = 20261204) -> dict:
    """Central differences of J along n_dirs seeded random directions d of the
    initial T (N(0, 1) per wet cell: J(T0 + h d) - J(T0 - h d) over 2h, each J
    from the plain fp64 HIP forward of result.n_steps steps) against the
    adjoint's directional derivative sum(dJ/dT0 * d). Reports per direction
    |fd - ad| / |ad| and |fd - ad| / (|dJ/dT0|_2 |d|_2) (the difference on the
    scale of the whole gradient, which does not blow up for directions nearly
    orthogonal to it), with their maxima and medians. The difference quotient
    itself carries the rounding of the two 60-step forwards divided by 2h."""
    rng = np.random.default_rng(seed)
    g = adj.domain.grid
    m3 = torch.as_tensor(adj.domain.mask3, device=adj.device)
    gnorm = float(torch.linalg.vector_norm(result.dJ_dT0))
    rel, rel_g, pairs = [], [], []
    for _ in range(n_dirs):
        d = torch.as_tensor(rng.standard_normal(g.shape3d), device=adj.device) * m3
        fd = (adj.J_forward(T0 + h * d, S0, kb, result.n_steps)
              - adj.J_forward(T0 - h * d, S0, kb, result.n_steps)) / (2.0 * h)
        ad = float(torch.sum(result.dJ_dT0 * d))
        rel.append(abs(fd - ad) / max(abs(ad), 1e-300))
        rel_g.append(abs(fd - ad) / max(gnorm * float(torch.linalg.vector_norm(d)), 1e-300))
        pairs.append((fd, ad))
    return {"n_directions": n_dirs, "h": h, "n_steps": result.n_steps,
            "rel_diff_max": float(np.max(rel)), "rel_diff_median": float(np.median(rel)),
            "rel_diff": [float(x) for x in rel],
            "grad_scaled_diff_max": float(np.max(rel_g)),
            "grad_scaled_diff_median": float(np.median(rel_g)), "fd_ad": pairs}
