"""Write the reference-produced golden fixtures under tests/golden/*.npz.

Runs ONLY in the build container, where the reference is mounted read-only at
/root/reference. It imports the reference's NumPy core (`libs.core`, never
`main`) for the functions that no closed form or recorded value pins to
round-off: the composed step (1-3 steps on tiny grids, and 200 steps of the
128x4x20 lock exchange), `tke_step` on stratified partial-cell columns,
`momentum_advection_up3` next to walls, the pressure gradient with partial
cells, the masked Coriolis averages at coasts, the Helmholtz solvers'
trajectories on arbitrary coefficients, and `Domain` geometry. Only inputs
and outputs are stored; the reference itself is never needed to run the tests.

Usage:  python tests/golden/make_golden.py [--out tests/golden] [--model2d-only]
"""

from __future__ import annotations

import argparse
import sys
import zlib
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
sys.path.insert(0, str(REF))
sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

from libs.core.bathymetry import build_bathymetry  # noqa: E402
from libs.core.closure import TKEParams, tke_step  # noqa: E402
from libs.core.advection_v06 import momentum_advection_up3  # noqa: E402
from libs.core.domain import Domain  # noqa: E402
from libs.core.grid import CGrid  # noqa: E402
from libs.core.model3d import Physics3D, State3D  # noqa: E402
from libs.core.model3d_v05 import PhysicsV05, Stepper3DV05  # noqa: E402
from libs.core.operators_masked import avg_u_to_v_m, avg_v_to_u_m  # noqa: E402
from libs.core.schemes import SchemeParams  # noqa: E402
from libs.core.solvers_var import HelmholtzVar  # noqa: E402


def _domain(nx, ny, nz, Lx, Ly, H0, topo, bc_x, bc_y, island=False, seed=7, vcoord="zlevel"):
    grid = CGrid(nx=nx, ny=ny, Lx=Lx, Ly=Ly, nz=nz, depth=H0)
    H, mask = build_bathymetry(grid, topo, H0, r_target=0.25, seed=seed, island=island,
                               h_rel=0.6)
    return Domain(grid=grid, H=H, mask=mask, vcoord=vcoord, bc_x=bc_x, bc_y=bc_y)


def _geometry(d: Domain, prefix: str) -> dict:
    dzu, dzv = d.face_thickness(d.dz3_ref)
    return {f"{prefix}H": d.H, f"{prefix}mask": d.mask, f"{prefix}dz3": d.dz3_ref,
            f"{prefix}mask3": d.mask3, f"{prefix}mask3u": d.mask3u, f"{prefix}mask3v": d.mask3v,
            f"{prefix}mask_u": d.mask_u, f"{prefix}mask_v": d.mask_v,
            f"{prefix}dz3u": dzu, f"{prefix}dz3v": dzv, f"{prefix}zc": d.depth_at_centres()}


def _random_state(d: Domain, rng, tracers, e=True):
    g = d.grid
    nz, ny, nx = g.shape3d
    depth = d.depth_at_centres()
    u = 0.05 * rng.standard_normal(g.shape3d) * d.mask3u
    v = 0.05 * rng.standard_normal(g.shape3d) * d.mask3v
    eta = 0.02 * rng.standard_normal(g.shape) * d.mask
    T = (4.0 + 14.0 * np.exp(-depth / 300.0) + 0.05 * rng.standard_normal(g.shape3d)) * d.mask3
    S = (35.0 + 0.1 * rng.standard_normal(g.shape3d)) * d.mask3
    ee = (1e-6 + 1e-4 * rng.random((nz + 1, ny, nx))) if e else None
    b = 0.01 * rng.standard_normal(g.shape3d) * d.mask3
    if tracers != "TS":
        T = S = None
    return dict(u=u, v=v, eta=eta, T=T, S=S, e=ee, b=b)


from golden_cases import COMPOSED  # noqa: E402  (tests/golden_cases.py)


def composed_fixtures() -> dict:
    out = {}
    for name, (dk, pk, vk, sk, dt, steps) in COMPOSED.items():
        d = _domain(**dk)
        phys = PhysicsV05(base=Physics3D(**pk), **vk)
        params = SchemeParams(**sk)
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        st0 = _random_state(d, rng, pk["tracers"], e=vk.get("closure") == "tke")
        stepper = Stepper3DV05(d, phys, params, dt)
        if pk["tracers"] == "TS":
            st0["b"] = stepper.buoyancy_of(st0["T"], st0["S"]) * d.mask3
        s = State3D(u=st0["u"], v=st0["v"], b=st0["b"], eta=st0["eta"], t=0.0,
                    T=st0["T"], S=st0["S"], e=st0["e"])
        rec = _geometry(d, "geo_")
        for k, v in st0.items():
            if v is not None:
                rec[f"in_{k}"] = v
        for n in range(1, steps + 1):
            s = stepper.step(s)
            for k in ("u", "v", "b", "eta", "T", "S", "e"):
                val = getattr(s, k)
                if val is not None:
                    rec[f"s{n}_{k}"] = val
            rec[f"s{n}_nu"] = stepper.nu
            rec[f"s{n}_kappa"] = stepper.kappa
            rec[f"s{n}_iters"] = np.array(stepper.solver_iterations)
            rec[f"s{n}_substeps"] = np.array(stepper.barotropic_substeps)
        rec["counters"] = np.array([stepper.solver_iterations, stepper.solver_failures,
                                    stepper.solver_rebuilds, stepper.tridiagonal_solves,
                                    stepper.eos_cells])
        out[name] = rec
    return out


def lock_exchange_fixture() -> dict:
    """The 128x4x20 lock exchange: flat 20 m, Lx=64 km, Ly=4 km, closed in x
    and periodic in y, linear EOS rho0=1000 alpha_T=2e-4 T0=5, T = 5|35,
    S = 35, up3_tvd, TKE (integral), theta 0.5 PCG-Jacobi rtol 1e-12, dt=1 s."""
    grid = CGrid(nx=128, ny=4, Lx=6.4e4, Ly=4.0e3, nz=20, depth=20.0)
    H = np.full(grid.shape, 20.0)
    mask = np.ones(grid.shape)
    d = Domain(grid=grid, H=H, mask=mask, vcoord="zlevel", bc_x="closed", bc_y="periodic")
    base = Physics3D(g=9.80616, f=0.0, H=20.0, N2=0.0, nu=1e-5, kappa=1e-6, rho0=1000.0,
                     tau_x=0.05, tau_y=0.0, tracers="TS", alpha_T=2e-4, beta_S=7.4e-4,
                     T0=5.0, S0=35.0, A_h=10.0, K_h=10.0)
    phys = PhysicsV05(base=base, eos="linear", pgf="keep", closure="tke", mxl="integral")
    params = SchemeParams(name="theta", theta=0.5, theta_cor=0.5, n_picard=2, theta_v=0.5,
                          advection="up3_tvd", solver="pcg_jacobi", rtol=1e-12, max_iter=2000)
    st = Stepper3DV05(d, phys, params, 1.0)
    x, _ = grid.coords("eta")
    warm = (x > 0.5 * grid.Lx).astype(np.float64)
    T = (5.0 + 30.0 * warm)[None] * np.ones(grid.shape3d)
    S = np.full(grid.shape3d, 35.0)
    b = st.buoyancy_of(T, S) * d.mask3
    zero = np.zeros(grid.shape3d)
    s = State3D(u=zero, v=zero, b=b, eta=np.zeros(grid.shape), t=0.0,
                T=T * d.mask3, S=S * d.mask3)
    rec = {"in_T": s.T, "in_S": s.S, "in_b": s.b}
    iters = []
    for n in range(1, 201):
        before = st.solver_iterations
        s = st.step(s)
        iters.append(st.solver_iterations - before)
        if n in (1, 10, 200):
            for k in ("u", "v", "b", "eta", "T", "S", "e"):
                rec[f"s{n}_{k}"] = getattr(s, k)
            rec[f"s{n}_nu"] = st.nu
            rec[f"s{n}_kappa"] = st.kappa
    rec["iters_per_step"] = np.array(iters)
    rec["counters"] = np.array([st.solver_iterations, st.solver_failures, st.solver_rebuilds,
                                st.tridiagonal_solves, st.eos_cells])
    return rec


def array_forcing_fixture() -> dict:
    """The step with array-valued tau_x(y), tau_y(x) and kappa_b(x, y)
    (golden_cases.ARRAY_FORCING) -- the reference takes them by NumPy
    broadcasting -- plus tke_step with an array kappa_b, checked here to equal
    the reference's scalar-kappa_b runs assembled column by column (the
    closure is columnwise), so broadcasting means per-column values."""
    from golden_cases import ARRAY_FORCING, array_forcing_fields
    dk, pk, vk, sk, dt, steps = ARRAY_FORCING
    d = _domain(**dk)
    tau_x, tau_y, kappa_b = array_forcing_fields(d.grid)
    phys = PhysicsV05(base=Physics3D(**pk, tau_x=tau_x, tau_y=tau_y, kappa=kappa_b), **vk)
    params = SchemeParams(**sk)
    rng = np.random.default_rng(zlib.crc32(b"array_forcing"))
    st0 = _random_state(d, rng, "TS", e=True)
    stepper = Stepper3DV05(d, phys, params, dt)
    assert np.array_equal(stepper.kappa[0], kappa_b)
    st0["b"] = stepper.buoyancy_of(st0["T"], st0["S"]) * d.mask3
    s = State3D(u=st0["u"], v=st0["v"], b=st0["b"], eta=st0["eta"], t=0.0,
                T=st0["T"], S=st0["S"], e=st0["e"])
    rec = _geometry(d, "geo_")
    rec.update(tau_x=tau_x, tau_y=tau_y, kappa_b=kappa_b)
    for k, v in st0.items():
        rec[f"in_{k}"] = v
    for n in range(1, steps + 1):
        s = stepper.step(s)
        for k in ("u", "v", "b", "eta", "T", "S", "e"):
            rec[f"s{n}_{k}"] = getattr(s, k)
        rec[f"s{n}_nu"] = stepper.nu
        rec[f"s{n}_kappa"] = stepper.kappa
        rec[f"s{n}_iters"] = np.array(stepper.solver_iterations)
    # tke_step with kappa_b on three values in a pattern of columns
    g = d.grid
    nz1 = g.nz + 1
    depth = d.depth_at_centres()
    e = 1e-6 + 3e-4 * rng.random((nz1,) + g.shape)
    u = 0.2 * rng.standard_normal(g.shape3d) * d.mask3u
    v = 0.2 * rng.standard_normal(g.shape3d) * d.mask3v
    b = (2e-3 * np.exp(-depth / 400.0) + 2e-5 * rng.standard_normal(g.shape3d)) * d.mask3
    Km = 1e-4 + 1e-2 * rng.random((nz1,) + g.shape)
    Kh = 1e-5 + 1e-2 * rng.random((nz1,) + g.shape)
    ustar2 = 1e-4 * (1.0 + rng.random(g.shape))
    values = np.array([0.8e-5, 1.0e-5, 1.2e-5])
    pick = rng.integers(0, 3, g.shape)
    kb = values[pick]
    rec.update(tke_e=e, tke_u=u, tke_v=v, tke_b=b, tke_Km=Km, tke_Kh=Kh, tke_ustar2=ustar2,
               tke_kappa_b=kb)
    for mxl in ("integral", "recursive"):
        p = TKEParams(mxl=mxl)
        whole = tke_step(e, u, v, b, d.dz3_ref, d.mask3, ustar2, 300.0, p, 1e-4, kb, Km, Kh)
        assembled = [np.zeros_like(a) for a in whole]
        for q, val in enumerate(values):
            part = tke_step(e, u, v, b, d.dz3_ref, d.mask3, ustar2, 300.0, p, 1e-4,
                            float(val), Km, Kh)
            for a, w in zip(assembled, part):
                a[:, pick == q] = w[:, pick == q]
        for a, w in zip(assembled, whole):
            assert np.array_equal(a, w)
        rec[f"tke_{mxl}_e"], rec[f"tke_{mxl}_Km"], rec[f"tke_{mxl}_Kh"] = whole
    return rec


# The lock_exchange_v06 settings of the reference's ladder sweep
# (tools/tier2_sweep.sh:85), theta with PCG-Jacobi.
LOCK_V06 = {"bathymetry.kind": "rough", "physics3d_v04.tracers": "TS",
            "scheme.advection": "up3_tvd", "physics3d_v06.closure": "tke",
            "physics3d_v05.eos": "teos10", "physics3d_v04.A_h": 10.0,
            "physics3d_v04.K_h": 10.0, "physics.f0": 0.0, "physics3d.N2": 0.0,
            "physics3d.nu": 1e-5, "physics3d.kappa": 1e-6, "physics3d.tau_x": 0.05,
            "physics3d_v05.pgf": "keep", "grid.Lx": 6.4e4, "grid.Ly": 6.4e3,
            "physics.H": 20.0, "scheme.name": "theta", "solver.kind": "pcg_jacobi"}


def reference_bundle_fixture(out: Path) -> None:
    """A v0.5 bundle written by the reference's own writer
    (tools/toml2nml.py:143-273, write_v05_bundle) from its own config tree
    with the lock_exchange_v06 settings at 16x16x8, CFL 2, 4 steps, into
    out/ref_bundle/ (prefix ref_lock, paths relative to that folder), and
    the reference stepper's state after those steps on the writer's own
    problem objects (expected.npz)."""
    import os
    import shutil
    import tempfile
    sys.path.insert(0, str(REF / "tools"))
    from toml2nml import write_v05_bundle
    from libs.utils.config import load_config
    cfg = load_config(REF / "config").with_overrides(LOCK_V06)
    dest = out / "ref_bundle"
    dest.mkdir(parents=True, exist_ok=True)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)
        try:
            res = write_v05_bundle("ref_lock", cfg, nx=16, nz=8, case_name="lock_exchange",
                                   cfl=2.0, n_steps_override=4, n_repeat=1, n_warmup=0)
        finally:
            os.chdir(cwd)
        for name in ("ref_lock.nml", "ref_lock_domain.bin", "ref_lock_init.bin"):
            shutil.copy(Path(tmp) / name, dest / name)
    st = Stepper3DV05(res["domain"], res["physics"], res["params"], res["dt"])
    s = res["state0"]
    iters = []
    for _ in range(res["n_steps"]):
        before = st.solver_iterations
        s = st.step(s)
        iters.append(st.solver_iterations - before)
    np.savez_compressed(dest / "expected.npz", eta=s.eta, u=s.u, v=s.v, b=s.b, T=s.T, S=s.S,
                        iters_per_step=np.array(iters), dt=np.array(res["dt"]))


def operator_fixtures() -> dict:
    rng = np.random.default_rng(20261202)
    d = _domain(nx=24, ny=16, nz=10, Lx=2.4e5, Ly=1.6e5, H0=1000.0, topo="rough",
                bc_x="closed", bc_y="periodic", island=True, seed=11)
    g = d.grid
    rec = _geometry(d, "geo_")
    dzu, dzv = d.face_thickness(d.dz3_ref)
    # tke_step on stratified partial-cell columns, both mixing-length forms
    nz1 = g.nz + 1
    depth = d.depth_at_centres()
    e = 1e-6 + 3e-4 * rng.random((nz1,) + g.shape)
    u = 0.2 * rng.standard_normal(g.shape3d) * d.mask3u
    v = 0.2 * rng.standard_normal(g.shape3d) * d.mask3v
    b = (2e-3 * np.exp(-depth / 400.0) + 2e-5 * rng.standard_normal(g.shape3d)) * d.mask3
    Km = 1e-4 + 1e-2 * rng.random((nz1,) + g.shape)
    Kh = 1e-5 + 1e-2 * rng.random((nz1,) + g.shape)
    ustar2 = 1e-4 * (1.0 + rng.random(g.shape))
    rec.update(tke_e=e, tke_u=u, tke_v=v, tke_b=b, tke_Km=Km, tke_Kh=Kh, tke_ustar2=ustar2)
    for mxl in ("integral", "recursive"):
        p = TKEParams(mxl=mxl)
        en, km, kh = tke_step(e, u, v, b, d.dz3_ref, d.mask3, ustar2, 300.0, p, 1e-4, 1e-5,
                              Km, Kh)
        rec[f"tke_{mxl}_e"], rec[f"tke_{mxl}_Km"], rec[f"tke_{mxl}_Kh"] = en, km, kh
    # momentum advection next to walls and the island
    w = np.zeros((nz1,) + g.shape)
    w[1:g.nz] = 1e-3 * rng.standard_normal((g.nz - 1,) + g.shape)
    rec.update(adv_u=u, adv_v=v, adv_w=w)
    for lim in (False, True):
        au, av = momentum_advection_up3(u, v, w, g.dx, g.dy, g.dz, d.mask3u, d.mask3v, lim)
        rec[f"adv_{'tvd' if lim else 'up3'}_u"], rec[f"adv_{'tvd' if lim else 'up3'}_v"] = au, av
    # masked Coriolis averages
    rec["cor_v_to_u"] = avg_v_to_u_m(v, d.mask3v, d.mask3u)
    rec["cor_u_to_v"] = avg_u_to_v_m(u, d.mask3u, d.mask3v)
    # pressure gradient with partial cells, both pgf modes
    for pgf in ("keep", "remove"):
        phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0), pgf=pgf)
        st = Stepper3DV05(d, phys, SchemeParams(name="fb"), 10.0)
        px, py = st._pressure_gradients(b)
        rec[f"pg_{pgf}_x"], rec[f"pg_{pgf}_y"] = px, py
    rec["pg_b"] = b
    # Helmholtz solver trajectories on arbitrary coefficients (nx, ny even)
    Ku = (200.0 + 800.0 * rng.random(g.shape)) * d.mask_u
    Kv = (200.0 + 800.0 * rng.random(g.shape)) * d.mask_v
    rhs = rng.standard_normal(g.shape) * d.mask
    coef = 9.80616 * 0.25 * 600.0 ** 2
    rec.update(helm_Ku=Ku, helm_Kv=Kv, helm_rhs=rhs, helm_coef=np.array(coef))
    for kind in ("pcg_jacobi", "pcg_rbgs", "rbgs", "multigrid"):
        solver = HelmholtzVar(g, coef, Ku, Kv, d.mask, kind=kind, rtol=1e-11, max_iter=5000)
        x, rep = solver.solve(rhs)
        rec[f"helm_{kind}_x"] = x
        rec[f"helm_{kind}_iters"] = np.array(rep.iterations)
    # bathymetry bits: rough (spectral) and seamount at a non-square size
    grid2 = CGrid(nx=40, ny=24, Lx=4.0e5, Ly=2.4e5, nz=6, depth=800.0)
    for topo in ("rough", "seamount", "slope", "ridge"):
        H, m = build_bathymetry(grid2, topo, 800.0, seed=20260911)
        rec[f"bathy_{topo}"] = H
    return rec


def model2d_fixtures() -> dict:
    """The reference's 2D model (libs/core/schemes.py Stepper): 4 steps of the
    broadband inertia-gravity wave (8 x 8 modes) at 32 x 24, from its own
    case initial state, for fb and theta with each of the five solver kinds
    (fft, pcg_jacobi, pcg_rbgs, rbgs, multigrid); per step the state and the
    solver iteration count. The 32 x 24 grid coarsens to a 4 x 3 level, so
    the multigrid fixture covers a coarsest level of odd size."""
    from libs.core.cases import InertiaGravityWaveBroadband
    from libs.core.schemes import Physics, SchemeParams as SP2, Stepper, explicit_dt_max
    grid = CGrid(nx=32, ny=24, Lx=1.0e7, Ly=0.75e7)
    phys = Physics(g=9.80616, f=1.0e-4, H=1000.0)
    case = InertiaGravityWaveBroadband(grid, phys, 8, 1.0, 1.0)
    s0 = case.initial()
    rec = {"u0": s0.u, "v0": s0.v, "eta0": s0.eta,
           "grid": np.array([grid.nx, grid.ny, grid.Lx, grid.Ly]),
           "physics": np.array([phys.g, phys.f, phys.H])}
    dt_max = explicit_dt_max(grid, phys)
    for tag, name, solver, cfl in (("fb", "fb", "fft", 0.5), ("pcgj", "theta", "pcg_jacobi", 4.0),
                                   ("fft", "theta", "fft", 4.0),
                                   ("pcgrb", "theta", "pcg_rbgs", 4.0),
                                   ("rbgs", "theta", "rbgs", 1.5),
                                   ("mg", "theta", "multigrid", 4.0)):
        params = SP2(name=name, theta=0.5, theta_cor=0.5, n_picard=2, solver=solver,
                     rtol=1e-12, max_iter=2000)
        dt = cfl * dt_max
        st = Stepper(grid, phys, params, dt)
        s = s0
        rec[f"{tag}_dt"] = np.array(dt)
        for n in range(1, 5):
            before = st.solver_iterations
            s = st.step(s)
            rec[f"{tag}_s{n}_u"], rec[f"{tag}_s{n}_v"], rec[f"{tag}_s{n}_eta"] = s.u, s.v, s.eta
            rec[f"{tag}_s{n}_iters"] = np.array(st.solver_iterations - before)
    return rec


MODEL3D_V02_CONFIGS = {
    # name: (scheme overrides, physics overrides, tracers)
    "theta_pcgj_c2": (dict(name="theta", solver="pcg_jacobi", advection="centered2"),
                      dict(A_h=5.0e3, K_h=2.0e3, tau_x=0.1, tau_y=-0.05, bottom_drag=1.0e-3),
                      "buoyancy"),
    "theta_mg_up1_ts": (dict(name="theta", solver="multigrid", advection="upwind1", theta=0.6),
                        dict(q_heat=50.0, q_salt=1.0e-6, K_h=1.0e3), "TS"),
    "fb_up1": (dict(name="fb", advection="upwind1", n_picard=3), dict(tau_x=0.05), "buoyancy"),
    "split_mean_c2_ts": (dict(name="split_explicit", advection="centered2",
                              barotropic_average="mean"),
                         dict(A_h=2.0e3, q_heat=-20.0, bottom_drag=5.0e-4), "TS"),
    "split_none": (dict(name="split_explicit", advection="none", n_split=7), {}, "buoyancy"),
    "theta_pcgrb_none": (dict(name="theta", solver="pcg_rbgs", advection="none", theta_v=1.0),
                         dict(nu=0.0, kappa=0.0), "buoyancy"),
}


def model3d_v02_fixtures() -> dict:
    """The reference's v0.2 3D model (libs/core/model3d.py Stepper3D): 3
    steps at 16 x 12 x 5 from a seeded smooth state, per configuration of
    MODEL3D_V02_CONFIGS (each scheme, the advection kinds, buoyancy and T/S
    tracers, stress, drag and horizontal mixing); per step the state and the
    solver iteration / barotropic substep counters."""
    from libs.core.model3d import Physics3D, State3D, Stepper3D
    from libs.core.schemes import SchemeParams as SP
    nx, ny, nz, Lx, Ly, H = 16, 12, 5, 1.6e6, 1.2e6, 1000.0
    grid = CGrid(nx=nx, ny=ny, Lx=Lx, Ly=Ly, nz=nz, depth=H)
    rng = np.random.default_rng(2024)
    x, y = grid.coords("eta")
    zc = grid.z_centre()[:, None, None]
    def smooth(scale):
        a, b2, c = rng.uniform(0.5, 1.5, 3)
        return scale * (np.cos(2 * np.pi * (x / Lx * a + y / Ly)) [None] * np.cos(zc / H * c * 3.0)
                        + 0.3 * np.sin(2 * np.pi * (2 * x / Lx - y / Ly * b2))[None])
    u0, v0 = smooth(0.2), smooth(0.15)
    b0 = 1.0e-4 * (zc - H) + smooth(1.0e-3)
    T0 = 10.0 + smooth(2.0)
    S0 = 35.0 + smooth(0.3)
    eta0 = 0.2 * np.cos(2 * np.pi * x / Lx) * np.sin(2 * np.pi * y / Ly)
    rec = {"u0": u0, "v0": v0, "b0": b0, "T0": T0, "S0": S0, "eta0": eta0,
           "grid": np.array([nx, ny, nz, Lx, Ly, H])}
    base_phys = dict(g=9.80616, f=1.0e-4, H=H, N2=1.0e-4, nu=1.0e-2, kappa=5.0e-3)
    dt = 600.0
    for name, (sch, ph, tracers) in MODEL3D_V02_CONFIGS.items():
        phys = Physics3D(**{**base_phys, **ph, "tracers": tracers})
        params = SP(**{"theta": 0.5, "theta_cor": 0.5, "n_picard": 2, "theta_v": 0.5,
                       "rtol": 1e-12, "max_iter": 2000, "solver": "fft", **sch})
        st = Stepper3D(grid, phys, params, dt)
        if tracers == "TS":
            s = State3D(u=u0, v=v0, b=phys.buoyancy(T0, S0), eta=eta0, T=T0, S=S0)
        else:
            s = State3D(u=u0, v=v0, b=b0, eta=eta0)
        for n in range(1, 4):
            s = st.step(s)
            for k in ("u", "v", "b", "eta", "T", "S"):
                a = getattr(s, k)
                if a is not None:
                    rec[f"{name}_s{n}_{k}"] = a
            rec[f"{name}_s{n}_counters"] = np.array([st.solver_iterations,
                                                     st.barotropic_substeps, st.n_split])
    rec["dt"] = np.array(dt)
    return rec


def case_fixtures() -> dict:
    """Initial and exact states of the reference's verification cases
    (libs/core/cases.py, libs/core/cases3d.py) on small grids: the 2D
    inertia-gravity wave, its broadband form and the geostrophic balance at
    12 x 10, the five v0.2 3D cases at 8 x 6 x 4, each at t = 0 and at a
    later time."""
    from libs.core.cases import (GeostrophicBalance, InertiaGravityWave,
                                 InertiaGravityWaveBroadband)
    from libs.core.cases3d import (BaroclinicIGW, Barotropic3D, EkmanSpiral, TracerAdvection,
                                   VerticalDiffusion)
    from libs.core.model3d import Physics3D
    from libs.core.schemes import Physics
    rec = {}
    g2 = CGrid(nx=12, ny=10, Lx=1.2e6, Ly=1.0e6)
    p2 = Physics(g=9.80616, f=1.0e-4, H=1000.0)
    for name, case in (("igw", InertiaGravityWave(g2, p2, 2, 1, 1.0)),
                       ("igw_broadband", InertiaGravityWaveBroadband(g2, p2, 3, 1.0, 1.5)),
                       ("geo_balance", GeostrophicBalance(g2, p2, 2, 1, 0.5))):
        for tag, t in (("t0", 0.0), ("t1", 0.37 * case.period)):
            s = case.initial() if t == 0.0 else case.exact(t)
            for k in ("u", "v", "eta"):
                rec[f"{name}_{tag}_{k}"] = getattr(s, k)
        rec[f"{name}_period"] = np.array(case.period)
    g3 = CGrid(nx=8, ny=6, Lx=8.0e5, Ly=6.0e5, nz=4, depth=500.0)
    base = dict(g=9.80616, f=1.0e-4, H=500.0, N2=1.0e-4, nu=1.0e-2, kappa=1.0e-3)
    cases = (("barotropic3d", lambda: Barotropic3D(g3, Physics3D(**base), 2, 1, 1.0)),
             ("vdiffusion", lambda: VerticalDiffusion(g3, Physics3D(**{**base, "f": 0.0}), 1, 1.0)),
             ("baroclinic_igw", lambda: BaroclinicIGW(g3, Physics3D(**base), 2, 1, 2, 1.0)),
             ("tracer_advect", lambda: TracerAdvection(
                 g3, Physics3D(**{**base, "f": 0.0, "tracers": "TS"}), 2, 1, 1.0, 0.5, 0.25)),
             ("ekman", lambda: EkmanSpiral(g3, Physics3D(**{**base, "tau_x": 0.1, "tau_y": 0.05}),
                                           2.0)))
    for name, make in cases:
        case = make()
        for tag, t in (("t0", 0.0), ("t1", 0.37 * min(case.period, 1.0e6))):
            s = case.initial() if t == 0.0 else case.exact(t)
            for k in ("u", "v", "b", "eta", "T", "S"):
                a = getattr(s, k)
                if a is not None:
                    rec[f"{name}_{tag}_{k}"] = a
        rec[f"{name}_period"] = np.array(case.period)
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", type=Path, default=Path(__file__).resolve().parent)
    ap.add_argument("--composed-only", action="store_true",
                    help="rewrite only the composed-step fixtures")
    ap.add_argument("--model2d-only", action="store_true",
                    help="rewrite only the 2D-model fixture")
    ap.add_argument("--model3d-v02-only", action="store_true",
                    help="rewrite only the v0.2 3D-model fixture")
    ap.add_argument("--cases-only", action="store_true",
                    help="rewrite only the verification-case fixture")
    ap.add_argument("--ref-bundle-only", action="store_true",
                    help="rewrite only the bundle written by the reference's writer")
    ap.add_argument("--array-forcing-only", action="store_true",
                    help="rewrite only the array-forcing fixture")
    args = ap.parse_args()
    args.out.mkdir(parents=True, exist_ok=True)
    if args.ref_bundle_only:
        reference_bundle_fixture(args.out)
        print("ref_bundle/ written")
        return 0
    if args.array_forcing_only:
        np.savez_compressed(args.out / "step_theta_array_forcing.npz", **array_forcing_fixture())
        print("step_theta_array_forcing.npz written")
        return 0
    if args.cases_only:
        np.savez_compressed(args.out / "cases_2d_3d.npz", **case_fixtures())
        print("cases_2d_3d.npz written")
        return 0
    if args.model3d_v02_only:
        np.savez_compressed(args.out / "model3d_v02_16x12x5.npz", **model3d_v02_fixtures())
        print("model3d_v02_16x12x5.npz written")
        return 0
    if args.model2d_only:
        np.savez_compressed(args.out / "model2d_igw_broadband_32x24.npz", **model2d_fixtures())
        print("model2d_igw_broadband_32x24.npz written")
        return 0
    if not args.composed_only:
        np.savez_compressed(args.out / "operators.npz", **operator_fixtures())
        print("operators.npz written")
    for name, rec in composed_fixtures().items():
        np.savez_compressed(args.out / f"step_{name}.npz", **rec)
        print(f"step_{name}.npz written (iterations {int(rec['counters'][0])})")
    if args.composed_only:
        return 0
    rec = lock_exchange_fixture()
    np.savez_compressed(args.out / "lock_exchange_128x4x20.npz", **rec)
    print("lock_exchange_128x4x20.npz written, iterations", rec["iters_per_step"].sum())
    np.savez_compressed(args.out / "model2d_igw_broadband_32x24.npz", **model2d_fixtures())
    print("model2d_igw_broadband_32x24.npz written")
    np.savez_compressed(args.out / "model3d_v02_16x12x5.npz", **model3d_v02_fixtures())
    print("model3d_v02_16x12x5.npz written")
    np.savez_compressed(args.out / "cases_2d_3d.npz", **case_fixtures())
    print("cases_2d_3d.npz written")
    np.savez_compressed(args.out / "step_theta_array_forcing.npz", **array_forcing_fixture())
    print("step_theta_array_forcing.npz written")
    reference_bundle_fixture(args.out)
    print("ref_bundle/ written")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
