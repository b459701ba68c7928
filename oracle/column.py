"""Vertical (per-column) operators of the oracle.

Arrays are [k, j, i] (layers) or [k, j, i] with nz+1 interfaces. Every
expression keeps the reference's operation order, so results are bitwise
those of the reference.
"""

from __future__ import annotations

import numpy as np


def safe_inv(x: np.ndarray) -> np.ndarray:
    """1/x where x > 0, else 0 (libs/core/vertical_var.py:16-18)."""
    return np.where(x > 0.0, 1.0 / np.where(x > 0.0, x, 1.0), 0.0)


def thomas(sub, diag, sup, rhs):
    """Batched tridiagonal solve along axis 0; sub[0] and sup[-1] unused
    (libs/core/vertical.py:17-44)."""
    n = diag.shape[0]
    cs = np.empty_like(sup)
    ds = np.empty_like(rhs)
    cs[0] = sup[0] / diag[0]
    ds[0] = rhs[0] / diag[0]
    for k in range(1, n):
        den = diag[k] - sub[k] * cs[k - 1]
        cs[k] = sup[k] / den
        ds[k] = (rhs[k] - sub[k] * ds[k - 1]) / den
    x = np.empty_like(rhs)
    x[n - 1] = ds[n - 1]
    for k in range(n - 2, -1, -1):
        x[k] = ds[k] - cs[k] * x[k + 1]
    return x


def interface_thickness(dz3):
    """Centre-to-centre spacing on interfaces; ends take the adjacent layer
    (libs/core/vertical_var.py:21-32)."""
    nz = dz3.shape[0]
    out = np.empty((nz + 1,) + dz3.shape[1:])
    out[0] = dz3[0]
    out[1:nz] = 0.5 * (dz3[:-1] + dz3[1:])
    out[nz] = dz3[nz - 1]
    return out


def _bottom_cells(mask3):
    """True in the deepest wet cell of each column (wet, dry below)."""
    below = np.concatenate([mask3[1:], np.zeros((1,) + mask3.shape[1:])], axis=0)
    return (mask3 > 0) & (below == 0)


def diffusion_coeffs(nu, dz3, dt, theta_v, bottom_drag, mask3):
    """(sub, diag, sup) of A = I - theta_v dt Dz with the surface, bottom and
    dry-cell rows of libs/core/vertical_var.py:35-77."""
    nz = nu.shape[0] - 1
    dzi = interface_thickness(dz3)
    inv_dz = safe_inv(dz3)
    fac = theta_v * dt * inv_dz
    a_top = fac * nu[:nz] * safe_inv(dzi[:nz])
    a_bot = fac * nu[1:] * safe_inv(dzi[1:])
    sub = -a_top.copy()
    sup = -a_bot.copy()
    diag = 1.0 + a_top + a_bot
    diag[0] = 1.0 + a_bot[0]
    sub[0] = 0.0
    bottom = _bottom_cells(mask3)
    diag = np.where(bottom, 1.0 + a_top + theta_v * dt * bottom_drag * inv_dz, diag)
    sup = np.where(bottom, 0.0, sup)
    dry = mask3 == 0
    diag = np.where(dry, 1.0, diag)
    sub = np.where(dry, 0.0, sub)
    sup = np.where(dry, 0.0, sup)
    return sub, diag, sup


def apply_diffusion(field, nu, dz3, surface_flux, bottom_drag, mask3):
    """Explicit Dz[field] with a surface flux, no flux through the bottom
    interface and drag on the bottom cell (libs/core/vertical_var.py:80-106,
    mask3 branch)."""
    nz = field.shape[0]
    dzi = interface_thickness(dz3)
    flux = np.zeros((nz + 1,) + field.shape[1:])
    flux[1:nz] = nu[1:nz] * (field[:-1] - field[1:]) * safe_inv(dzi[1:nz])
    flux[0] = surface_flux
    bottom = _bottom_cells(mask3)
    flux[1:nz] = np.where(bottom[:-1], 0.0, flux[1:nz])
    flux[nz] = 0.0
    drag = bottom_drag * field * bottom
    inv = safe_inv(dz3)
    out = (flux[:nz] - flux[1:]) * inv
    return (out - drag * inv) * (mask3 > 0)


def w_from_transport_divergence(tdiv):
    """w on interfaces, zero at the bottom, integrated upward
    (libs/core/vertical_var.py:109-121)."""
    nz = tdiv.shape[0]
    w = np.zeros((nz + 1,) + tdiv.shape[1:])
    w[:nz] = -np.cumsum(tdiv[::-1], axis=0)[::-1]
    return w


def w_at_centres(w):
    return 0.5 * (w[:-1] + w[1:])


def buoyancy_potential(b, dz3):
    """Phi_k = 1/2 dz_k b_k + sum_{k'<k} dz b (libs/core/vertical_var.py:132-136)."""
    above = np.concatenate([np.zeros((1,) + b.shape[1:]),
                            np.cumsum((dz3 * b)[:-1], axis=0)], axis=0)
    return 0.5 * dz3 * b + above


def remove_depth_mean(phi, dz3):
    """Thickness-weighted depth mean removed (libs/core/vertical_var.py:139-145)."""
    total = np.sum(dz3, axis=0, keepdims=True)
    mean = np.sum(dz3 * phi, axis=0, keepdims=True) * safe_inv(total)
    return phi - mean
