"""Horizontal C-grid operators and advection of the oracle.

Periodic index layout: every shift wraps (np.roll); walls are face masks.
`east(a)[.., j, i] = a[.., j, i+1]`, `west(a) = a[.., j, i-1]`, and likewise
north/south along j.
"""

from __future__ import annotations

import numpy as np

KAPPA = 1.0 / 3.0


def east(a, n=1):
    return np.roll(a, -n, axis=-1)


def west(a, n=1):
    return np.roll(a, n, axis=-1)


def north(a, n=1):
    return np.roll(a, -n, axis=-2)


def south(a, n=1):
    return np.roll(a, n, axis=-2)


# ------------------------------------------------- masked operators (S10.4)
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
    """sum(w f) / sum(w) over wet partners, summed left to right from 0
    (libs/core/operators_masked.py:35-42)."""
    num = 0
    den = 0
    for f, w in pairs:
        num = num + w * f
    for _, w in pairs:
        den = den + w
    return np.where(den > 0.0, num / np.where(den > 0.0, den, 1.0), 0.0)


def avg_v_to_u(v, mask_v, mask_u):
    """libs/core/operators_masked.py:45-52"""
    vs, ms = south(v), south(mask_v)
    out = _wet_mean([(v, mask_v), (vs, ms), (east(v), east(mask_v)), (east(vs), east(ms))])
    return out * mask_u


def avg_u_to_v(u, mask_u, mask_v):
    """libs/core/operators_masked.py:55-61"""
    uw, mw = west(u), west(mask_u)
    out = _wet_mean([(u, mask_u), (uw, mw), (north(u), north(mask_u)), (north(uw), north(mw))])
    return out * mask_v


def laplacian(a, dx, dy, mask_u, mask_v, mask):
    """D o G (libs/core/operators_masked.py:64-68)."""
    return div(gradx(a, dx, mask_u), grady(a, dy, mask_v), dx, dy, mask)


# ------------------------------------------------------ kappa = 1/3 faces
def _superbee(r):
    return np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)))


def kappa_face(c, vel, axis, wet, limited):
    """Face value between c[i] and c[i+1] for the face velocity vel; upwind1
    where the far upstream cell is dry (libs/core/advection_v06.py:21-47)."""
    cm = np.roll(c, 1, axis=axis)
    cp = np.roll(c, -1, axis=axis)
    cpp = np.roll(c, -2, axis=axis)
    if limited:
        dc = cp - c
        zero = dc == 0.0
        eps = np.where(zero, 1.0, dc)
        rp = np.where(zero, 0.0, (c - cm) / eps)
        rn = np.where(zero, 0.0, (cpp - cp) / eps)
        fp = c + 0.5 * _superbee(rp) * dc
        fn = cp - 0.5 * _superbee(rn) * dc
    else:
        fp = c + 0.25 * ((1.0 - KAPPA) * (c - cm) + (1.0 + KAPPA) * (cp - c))
        fn = cp - 0.25 * ((1.0 - KAPPA) * (cpp - cp) + (1.0 + KAPPA) * (cp - c))
    if wet is not None:
        fp = np.where(np.roll(wet, 1, axis=axis) > 0.0, fp, c)
        fn = np.where(np.roll(wet, -2, axis=axis) > 0.0, fn, cp)
    return np.where(vel > 0.0, fp, fn)


def kappa_face_z(c, w, wet3, limited):
    """Values on interior interfaces 1..nz-1, w positive upward
    (libs/core/advection_v06.py:50-74). Returns [nz-1, ...]."""
    above, below = c[:-1], c[1:]
    zc = np.zeros_like(c[:1])
    far_up = np.concatenate([zc, c[:-2]], axis=0)
    far_dn = np.concatenate([c[2:], zc], axis=0)
    zw = np.zeros_like(wet3[:1])
    has_up = np.concatenate([zw, wet3[:-2]], axis=0) > 0.0
    has_dn = np.concatenate([wet3[2:], zw], axis=0) > 0.0
    d = above - below
    if limited:
        zero = d == 0.0
        eps = np.where(zero, 1.0, d)
        r_up = np.where(zero, 0.0, (below - far_dn) / eps)
        r_dn = np.where(zero, 0.0, (far_up - above) / eps)
        f_up = below + 0.5 * _superbee(r_up) * d
        f_dn = above - 0.5 * _superbee(r_dn) * d
    else:
        f_up = below + 0.25 * ((1.0 - KAPPA) * (below - far_dn) + (1.0 + KAPPA) * (above - below))
        f_dn = above - 0.25 * ((1.0 - KAPPA) * (far_up - above) + (1.0 + KAPPA) * (above - below))
    f_up = np.where(has_dn, f_up, below)
    f_dn = np.where(has_up, f_dn, above)
    nz = c.shape[0]
    return np.where(w[1:nz] > 0.0, f_up, f_dn)


def momentum_advection_up3(u, v, w, dx, dy, dz, mask3u, mask3v, limited):
    """Flux-form momentum advection with kappa-interpolated advected velocity
    (libs/core/advection_v06.py:77-108). dz is the NOMINAL layer thickness."""
    nz = u.shape[0]
    ub = 0.5 * (west(u) + u)                                   # centre i
    u_c = west(kappa_face(u, east(ub), -1, mask3u, limited))
    ax = (east(ub * u_c) - ub * u_c) / dx
    vb = 0.5 * (v + east(v))                                   # corner
    fy = vb * kappa_face(u, vb, -2, mask3u, limited)
    ay = (fy - south(fy)) / dy
    wb = 0.5 * (w + east(w))
    fz = np.zeros_like(wb)
    if nz > 1:
        fz[1:nz] = wb[1:nz] * kappa_face_z(u, wb, mask3u, limited)
    az = (fz[:nz] - fz[1:]) / dz

    vbc = 0.5 * (south(v) + v)
    v_c = south(kappa_face(v, north(vbc), -2, mask3v, limited))
    by = (north(vbc * v_c) - vbc * v_c) / dy
    ubr = 0.5 * (u + north(u))
    fx = ubr * kappa_face(v, ubr, -1, mask3v, limited)
    bx = (fx - west(fx)) / dx
    wbv = 0.5 * (w + north(w))
    gz = np.zeros_like(wbv)
    if nz > 1:
        gz[1:nz] = wbv[1:nz] * kappa_face_z(v, wbv, mask3v, limited)
    bz = (gz[:nz] - gz[1:]) / dz
    return ax + ay + az, bx + by + bz


def momentum_advection_centred(u, v, w, dx, dy, dz):
    """Second-order centred flux form (libs/core/advection.py:62-95)."""
    nz = u.shape[0]
    ub = 0.5 * (west(u) + u)
    ax = (east(ub) ** 2 - ub ** 2) / dx
    vb = 0.5 * (v + east(v))
    fy = vb * (0.5 * (u + north(u)))
    ay = (fy - south(fy)) / dy
    wb = 0.5 * (w + east(w))
    fz = np.zeros_like(wb)
    if nz > 1:
        fz[1:nz] = wb[1:nz] * 0.5 * (u[:-1] + u[1:])
    az = (fz[:nz] - fz[1:]) / dz

    vbc = 0.5 * (south(v) + v)
    by = (north(vbc) ** 2 - vbc ** 2) / dy
    ubr = 0.5 * (u + north(u))
    fx = ubr * (0.5 * (v + east(v)))
    bx = (fx - west(fx)) / dx
    wbv = 0.5 * (w + north(w))
    gz = np.zeros_like(wbv)
    if nz > 1:
        gz[1:nz] = wbv[1:nz] * 0.5 * (v[:-1] + v[1:])
    bz = (gz[:nz] - gz[1:]) / dz
    return ax + ay + az, bx + by + bz
