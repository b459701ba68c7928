"""One-equation TKE closure of the oracle (libs/core/closure.py:25-205).

TKE lives on interfaces 0..nz; interface k separates layer k-1 (above) from
layer k (below). The march of `gaspar_lengths` is written per interface with
two exact early exits (a march past the column end or with an exhausted
budget changes nothing), which leaves every value bitwise identical to the
reference's full j-loop.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .column import safe_inv, thomas


@dataclass(frozen=True)
class TKEParams:
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


def interface_geometry(dz3, mask3):
    """Depth below the surface, distance to the wet bottom, the interior-wet
    flag and the spacing on interfaces (libs/core/closure.py:43-62)."""
    nz = dz3.shape[0]
    zi = np.concatenate([np.zeros((1,) + dz3.shape[1:]), np.cumsum(dz3, axis=0)], axis=0)
    z_b = np.maximum(zi[-1][None] - zi, 0.0)
    wet = np.zeros_like(zi)
    wet[1:nz] = mask3[:-1] * mask3[1:]
    dzi = np.empty_like(zi)
    dzi[0] = dz3[0]
    dzi[1:nz] = 0.5 * (dz3[:-1] + dz3[1:])
    dzi[nz] = dz3[-1]
    return zi, z_b, wet, dzi


def _march(l, budget, n2_j, dz_j):
    """One interface crossed by the Gaspar budget march
    (libs/core/closure.py:89-96)."""
    cost = n2_j * dz_j * (l + 0.5 * dz_j)
    full = cost <= budget
    a = np.where(n2_j > 0.0, n2_j, 1.0)
    s_part = np.sqrt(np.maximum(l * l + 2.0 * budget / a, 0.0)) - l
    step = np.where(full, dz_j, np.where(n2_j > 0.0, np.clip(s_part, 0.0, dz_j), dz_j))
    return l + np.where(budget > 0.0, step, 0.0), np.where(full, budget - cost, 0.0)


def gaspar_lengths(e, n2, dzi, z_s, z_b, wet, p: TKEParams):
    """Distances a parcel of energy e can rise / sink against N^2
    (libs/core/closure.py:65-103)."""
    nz1 = e.shape[0]
    n2p = np.maximum(n2, 0.0)
    e_k = np.maximum(e, p.e_min)
    l_up = np.zeros_like(e)
    l_dn = np.zeros_like(e)
    b_up = e_k.copy()
    b_dn = e_k.copy()
    for j in range(1, nz1):
        # upward: interface k crosses k-j (only k >= j reach a real interface)
        ks = slice(j, nz1)
        l_up[ks], b_up[ks] = _march(l_up[ks], b_up[ks], n2p[:nz1 - j], dzi[:nz1 - j])
        # downward: interface k crosses k+j (only k <= nz1-1-j)
        kd = slice(0, nz1 - j)
        l_dn[kd], b_dn[kd] = _march(l_dn[kd], b_dn[kd], n2p[j:], dzi[j:])
    l_up = np.minimum(l_up, z_s + p.z_0) * wet
    l_dn = np.minimum(l_dn, z_b + p.z_0) * wet
    return l_up, l_dn


def recursive_lengths(e, n2, dzi, z_s, z_b, wet, p: TKEParams):
    """Buoyancy length capped by the walls, then limited by one downward and
    one upward sweep (libs/core/closure.py:106-129)."""
    nz1 = e.shape[0]
    n2c = np.maximum(n2, p.n2_min)
    e_k = np.maximum(e, p.e_min)
    l = np.sqrt(2.0 * e_k / n2c)
    l = np.minimum(l, np.minimum(z_s + p.z_0, z_b + p.z_0))
    for k in range(1, nz1):
        l[k] = np.minimum(l[k], l[k - 1] + dzi[k])
    for k in range(nz1 - 2, -1, -1):
        l[k] = np.minimum(l[k], l[k + 1] + dzi[k + 1])
    l = l * wet
    return l, l


def mixing_lengths(e, n2, dzi, z_s, z_b, wet, p: TKEParams):
    if p.mxl == "recursive":
        return recursive_lengths(e, n2, dzi, z_s, z_b, wet, p)
    if p.mxl != "integral":
        raise ValueError(f"unknown mixing length form {p.mxl!r} (integral|recursive)")
    return gaspar_lengths(e, n2, dzi, z_s, z_b, wet, p)


def tke_step(e, u, v, b, dz3, mask3, ustar2, dt, p: TKEParams, nu_b, kappa_b,
             K_m_old, K_h_old):
    """One TKE step; returns (e_new, K_m, K_h) on interfaces
    (libs/core/closure.py:141-205). nu_b / kappa_b may be scalars or arrays
    broadcastable to [nz+1, ny, nx]."""
    nz = dz3.shape[0]
    z_s, z_b, wet, dzi = interface_geometry(dz3, mask3)
    inv_dzi = safe_inv(dzi)
    uc = 0.5 * (u + np.roll(u, 1, axis=-1))
    vc = 0.5 * (v + np.roll(v, 1, axis=-2))
    du = np.zeros_like(e)
    dv = np.zeros_like(e)
    db = np.zeros_like(e)
    du[1:nz] = (uc[:-1] - uc[1:]) * inv_dzi[1:nz]
    dv[1:nz] = (vc[:-1] - vc[1:]) * inv_dzi[1:nz]
    db[1:nz] = (b[:-1] - b[1:]) * inv_dzi[1:nz]
    shear2 = (du ** 2 + dv ** 2) * wet
    n2 = db * wet

    l_up, l_dn = mixing_lengths(e, n2, dzi, z_s, z_b, wet, p)
    l_k = np.maximum(p.l_min, np.sqrt(l_up * l_dn))
    l_eps = np.maximum(p.l_min, np.minimum(l_up, l_dn))

    prod = (K_m_old - nu_b) * shear2 - (K_h_old - kappa_b) * n2
    diss = p.c_eps * np.sqrt(np.maximum(e, p.e_min)) / l_eps
    K_lay = 0.5 * (K_m_old[:-1] + K_m_old[1:])
    inv_dz3 = safe_inv(dz3)
    a_up = np.zeros_like(e)
    a_dn = np.zeros_like(e)
    a_up[1:nz] = dt * K_lay[:-1] * inv_dz3[:-1] * inv_dzi[1:nz]
    a_dn[1:nz] = dt * K_lay[1:] * inv_dz3[1:] * inv_dzi[1:nz]
    sub = -a_up
    sup = -a_dn
    diag = 1.0 + a_up + a_dn + dt * diss
    rhs = e + dt * prod
    bnd = wet <= 0.0
    diag = np.where(bnd, 1.0, diag)
    sub = np.where(bnd, 0.0, sub)
    sup = np.where(bnd, 0.0, sup)
    rhs = np.where(bnd, p.e_min, rhs)
    rhs[0] = np.where(mask3[0] > 0.0, p.e_bb * ustar2, p.e_min)
    e_new = thomas(sub, diag, sup, rhs)
    e_new = np.maximum(e_new, p.e_min)
    e_new = np.where(wet > 0.0, e_new, p.e_min)
    e_new[0] = rhs[0]
    K_m = p.c_k * l_k * np.sqrt(e_new) * wet + nu_b
    K_h = (K_m - nu_b) / p.pr_t + kappa_b
    return e_new, K_m, K_h
