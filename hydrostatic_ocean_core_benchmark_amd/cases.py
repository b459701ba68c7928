"""Problem builders for the benchmark workloads.

lock_exchange(): the 128x4x20 lock exchange -- flat 20 m, Lx = 64 km,
    Ly = 4 km (closed in x, periodic in y), linear EOS rho0 = 1000,
    alpha_T = 2e-4, T0 = 5; T = 5 degC for x < 32 km and 35 degC beyond
    (LockExchange.initial, libs/core/cases3d_v05.py:245-258, with
    delta_T = 30), S = 35, u = v = eta = 0; up3_tvd, TKE (integral), theta 0.5
    with PCG-Jacobi rtol 1e-12, dt = 1 s. Knobs the workload leaves open take
    the reference's lock_exchange_v06 values (tools/tier2_sweep.sh:85):
    f0 = 0, nu = 1e-5, kappa = 1e-6, tau_x = 0.05, A_h = K_h = 10, pgf = keep.

ladder_lock_exchange(): the reference's published hardware-ladder workload,
    lock_exchange_v06 (tools/tier2_sweep.sh:85 over config/*.toml): nx x nx x
    nz over `rough` topography (r 0.1, seed 20260911), Lx = 64 km, Ly = 6.4 km,
    H = 20 m, periodic, z-level; TEOS-10, up3_tvd, TKE (integral), A_h = K_h =
    10, nu = 1e-5, kappa = 1e-6, tau_x = 0.05, f0 = 0, pgf keep; the v0.5
    LockExchange initial state with delta_T = 2 (libs/core/cases3d_v05.py:
    245-258); dt = cfl x the barotropic limit of the deepest column
    (tools/toml2nml.py:143-170).

basin(): a closed basin of nx x ny x nz cells at spacing dx with z-level
    partial-cell bathymetry from a seeded Gaussian random field (correlation
    length 200 km) clipped to 500-4000 m, >= 3 wet layers; initial
    T(z) = 2 + 18 exp(z / 800 m) (z <= 0 up) plus seeded N(0, sigma_T) noise,
    S = 35 + 0.5 (y/Ly - 1/2) (a weak linear gradient, 1 psu across the
    basin); zonal wind stress tau_x = 0.1 cos(pi y / Ly) N m-2; polyTEOS10
    EOS; up3_tvd; TKE (integral); theta 0.5; dt = 600 s at dx = 10 km.
    No reference counterpart (the reference has no GRF generator or
    spatially varying stress).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .domain import CGrid, Domain, build_bathymetry
from .types import Physics3D, PhysicsV05, SchemeParams


@dataclass
class Problem:
    name: str
    domain: Domain
    physics: PhysicsV05
    params: SchemeParams
    dt: float
    n_steps: int
    initial_T: np.ndarray
    initial_S: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        g = self.domain.grid
        return g.nx * g.ny * g.nz


def lock_exchange(nx: int = 128, ny: int = 4, nz: int = 20, n_steps: int = 200,
                  solver: str = "pcg_jacobi") -> Problem:
    grid = CGrid(nx=nx, ny=ny, Lx=6.4e4, Ly=4.0e3, nz=nz, depth=20.0)
    dom = Domain(grid=grid, H=np.full(grid.shape, 20.0), mask=np.ones(grid.shape),
                 bc_x="closed", bc_y="periodic")
    base = Physics3D(g=9.80616, f=0.0, H=20.0, N2=0.0, nu=1e-5, kappa=1e-6, rho0=1000.0,
                     tau_x=0.05, tau_y=0.0, tracers="TS", alpha_T=2e-4, beta_S=7.4e-4,
                     T0=5.0, S0=35.0, A_h=10.0, K_h=10.0)
    phys = PhysicsV05(base=base, eos="linear", pgf="keep", closure="tke", mxl="integral")
    params = SchemeParams(name="theta", theta=0.5, theta_cor=0.5, n_picard=2, theta_v=0.5,
                          advection="up3_tvd", solver=solver, rtol=1e-12, max_iter=2000)
    x, _ = grid.coords("eta")
    warm = (x > 0.5 * grid.Lx).astype(np.float64)
    T = (5.0 + 30.0 * warm)[None] * np.ones(grid.shape3d) * dom.mask3
    S = np.full(grid.shape3d, 35.0) * dom.mask3
    return Problem("lock_exchange", dom, phys, params, 1.0, n_steps, T, S,
                   {"grid": f"{nx}x{ny}x{nz}", "Lx": grid.Lx, "Ly": grid.Ly})


def ladder_lock_exchange(nx: int, nz: int = 30, scheme: str = "fb", solver: str = "multigrid",
                         cfl: float = 0.5, n_steps: int = 50) -> Problem:
    grid = CGrid(nx=nx, ny=nx, Lx=6.4e4, Ly=6.4e3, nz=nz, depth=20.0)
    H, mask = build_bathymetry(grid, "rough", 20.0, h_rel=0.9, length=0.0, slope=0.5,
                               spectrum_slope=1.5, spectrum_kmax=0, r_target=0.1, seed=20260911,
                               island=False)
    dom = Domain(grid=grid, H=H, mask=mask, vcoord="zlevel", bc_x="periodic", bc_y="periodic",
                 min_partial=0.1, face_rule="min")
    base = Physics3D(g=9.80616, f=0.0, H=20.0, N2=0.0, nu=1e-5, kappa=1e-6, rho0=1025.0,
                     tau_x=0.05, tau_y=0.0, bottom_drag=0.0, tracers="TS", alpha_T=2e-4,
                     beta_S=7.4e-4, T0=10.0, S0=35.0, cp=3990.0, A_h=10.0, K_h=10.0)
    phys = PhysicsV05(base=base, eos="teos10", pgf="keep", closure="tke", mxl="integral")
    params = SchemeParams(name=scheme, theta=0.5, theta_cor=0.5, n_picard=2, theta_v=0.5,
                          advection="up3_tvd", solver=solver, rtol=1e-12, max_iter=2000)
    c = np.sqrt(base.g * float(np.max(H[mask > 0])))
    dt = cfl * (1.0 / (c * np.sqrt(1.0 / grid.dx ** 2 + 1.0 / grid.dy ** 2)))
    x, _ = grid.coords("eta")
    warm = (x > 0.5 * grid.Lx).astype(np.float64)
    T = (base.T0 + 2.0 * warm)[None] * np.ones(grid.shape3d) * dom.mask3
    S = np.full(grid.shape3d, base.S0) * dom.mask3
    label = "fb" if scheme == "fb" else f"theta0.5 {solver}"
    return Problem("lock_exchange_v06", dom, phys, params, dt, n_steps, T, S,
                   {"grid": f"{nx}x{nx}x{nz}", "scheme": label, "cfl": cfl})


def basin(nx: int = 512, ny: int | None = None, nz: int = 40, dx: float = 1.0e4,
          dt: float = 600.0, n_steps: int = 100, seed: int = 20261202,
          sigma_T: float = 0.05, solver: str = "pcg_mg", eos: str = "teos10",
          rtol: float = 1e-12, max_iter: int = 5000, kappa_b=1e-5) -> Problem:
    ny = nx if ny is None else ny
    Lx, Ly = nx * dx, ny * dx
    grid = CGrid(nx=nx, ny=ny, Lx=Lx, Ly=Ly, nz=nz, depth=4000.0)
    H, mask = build_bathymetry(grid, "grf", 4000.0, corr_length=2.0e5, h_min=500.0,
                               h_max=4000.0, seed=seed)
    dom = Domain(grid=grid, H=H, mask=mask, bc_x="closed", bc_y="closed")
    dzr = float(dom.dzr[0])
    if float(np.min(H)) < 3.0 * dzr:
        raise ValueError("bathymetry leaves fewer than 3 wet layers somewhere")
    _, yc = grid.coords("eta")
    tau_x = 0.1 * np.cos(np.pi * yc / Ly)
    base = Physics3D(g=9.80616, f=1.0e-4, H=4000.0, N2=0.0, nu=1.0e-4, kappa=kappa_b,
                     rho0=1025.0, tau_x=tau_x, tau_y=0.0, tracers="TS", alpha_T=2e-4,
                     beta_S=7.4e-4, T0=10.0, S0=35.0, A_h=100.0, K_h=100.0)
    phys = PhysicsV05(base=base, eos=eos, pgf="keep", closure="tke", mxl="integral")
    params = SchemeParams(name="theta", theta=0.5, theta_cor=0.5, n_picard=2, theta_v=0.5,
                          advection="up3_tvd", solver=solver, rtol=rtol, max_iter=max_iter)
    depth = dom.depth_at_centres()
    rng = np.random.default_rng(seed + 1)
    T = (2.0 + 18.0 * np.exp(-depth / 800.0) + sigma_T * rng.standard_normal(grid.shape3d))
    S = 35.0 + 0.5 * (yc / Ly - 0.5)
    T = T * dom.mask3
    S = np.broadcast_to(S, grid.shape3d) * dom.mask3
    return Problem("basin", dom, phys, params, dt, n_steps, T, S,
                   {"grid": f"{nx}x{ny}x{nz}", "dx": dx, "seed": seed,
                    "wet_fraction": float(np.mean(dom.mask3)),
                    "H_min": float(H.min()), "H_max": float(H.max())})
