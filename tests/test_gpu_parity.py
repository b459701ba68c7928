"""HIP path against the reference's golden fixtures and the oracle, through
the C-ABI. Bitwise agreement is the expectation (the kernels keep NumPy's
operation order, no FMA contraction, and NumPy's summation order in the
solver's dot products); the gates below are the reference's own: relative L2
< 1e-12 for one step, < 1e-9 for a full run, solver iteration counts equal."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, rel_l2
from golden_cases import (COMPOSED, build_array_forcing_case, build_case,
                          lock_exchange_problem)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def hoc(gpu_available):
    from hydrostatic_ocean_core_benchmark_amd import stepper
    return stepper


def test_pairwise_dot_matches_numpy_sum(hoc):
    rng = np.random.default_rng(3)
    for n in (1, 5, 8, 37, 128, 129, 384, 512, 1000, 8192, 8193, 36864, 90000, 262144):
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-6, 6, size=n)
        b = rng.standard_normal(n)
        got = hoc.pairwise_dot(torch.as_tensor(a, device="cuda"), torch.as_tensor(b, device="cuda"))
        assert got == float(np.sum(a * b)), n


def test_thomas_matches_oracle(hoc):
    from oracle.column import thomas
    rng = np.random.default_rng(4)
    for n, ncol in ((1, 7), (2, 5), (20, 512), (61, 300)):
        sub = rng.standard_normal((n, ncol))
        sup = rng.standard_normal((n, ncol))
        diag = 4.0 + np.abs(sub) + np.abs(sup) + rng.random((n, ncol))
        rhs = rng.standard_normal((n, ncol))
        want = thomas(sub, diag, sup, rhs)
        got = hoc.thomas_batched(*(torch.as_tensor(a, device="cuda") for a in (sub, diag, sup, rhs)))
        assert np.array_equal(_np(got), want), (n, ncol)


@pytest.mark.parametrize("kind", ["linear", "seos", "teos10"])
def test_eos_matches_oracle(hoc, kind):
    from oracle.eos import buoyancy
    rng = np.random.default_rng(5)
    T = 30.0 * rng.random(4099) - 2.0
    S = 30.0 + 8.0 * rng.random(4099)
    z = 5000.0 * rng.random(4099)
    kw = dict(g=9.80616, rho0=1025.0, alpha_T=2e-4, beta_S=7.4e-4, T0=10.0, S0=35.0)
    want = buoyancy(kind, T, S, z, **kw)
    got = hoc.eos_buoyancy(kind, *(torch.as_tensor(a, device="cuda") for a in (T, S, z)), **kw)
    assert np.array_equal(_np(got), want)


def test_teos10_check_value(hoc):
    """rho(SA=30, CT=10, z=1000) = 1027.45140 (Roquet et al. 2015), and the
    reference's computed value 1027.4513640377502
    (expr/E07_v05_physics/data/v05_verification_20260923.json, V5-4)."""
    t = lambda x: torch.tensor([x], dtype=torch.float64, device="cuda")  # noqa: E731
    # b = -g (rho - rho0)/rho0; with g = -1, rho0 = 1: b = rho - 1
    b = hoc.eos_buoyancy("teos10", t(10.0), t(30.0), t(1000.0), g=-1.0, rho0=1.0, alpha_T=0.0,
                         beta_S=0.0, T0=0.0, S0=0.0)
    rho = float(b.item()) + 1.0
    assert abs(rho - 1027.45140) < 1e-4
    assert abs(rho - 1027.4513640377502) < 1e-9


def _stepper_for(hoc, name, golden, solver=None):
    dom, phys, params, dt, steps, s0 = build_case(name, golden)
    if solver is not None:
        params.solver = solver
    return hoc.HipStepper(dom, phys, params, dt), dom, phys, params, dt, steps, s0


@pytest.mark.parametrize("mxl", ["integral", "recursive"])
def test_tke_step_matches_reference(hoc, mxl):
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    f = load_golden("operators.npz")
    grid = CGrid(nx=24, ny=16, Lx=2.4e5, Ly=1.6e5, nz=10, depth=1000.0)
    dom = Domain(grid=grid, H=f["geo_H"], mask=f["geo_mask"], bc_x="closed", bc_y="periodic")
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0, nu=1e-4, kappa=1e-5),
                      closure="tke", mxl=mxl)
    st = hoc.HipStepper(dom, phys, SchemeParams(name="fb"), 300.0, ustar2=f["tke_ustar2"])
    st.nu.copy_(torch.as_tensor(f["tke_Km"], device="cuda"))
    st.kappa.copy_(torch.as_tensor(f["tke_Kh"], device="cuda"))
    e = st.tke_step(f["tke_e"], f["tke_u"], f["tke_v"], f["tke_b"])
    assert np.array_equal(_np(e), f[f"tke_{mxl}_e"])
    assert np.array_equal(_np(st.nu), f[f"tke_{mxl}_Km"])
    assert np.array_equal(_np(st.kappa), f[f"tke_{mxl}_Kh"])


@pytest.mark.parametrize("mxl", ["integral", "recursive"])
def test_tke_step_array_kappa_b_matches_reference(hoc, mxl):
    """The closure with kappa_b on three values over the columns, against
    the reference's own output."""
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    f = load_golden("step_theta_array_forcing.npz")
    grid = CGrid(nx=24, ny=16, Lx=2.4e5, Ly=1.6e5, nz=8, depth=1000.0)
    dom = Domain(grid=grid, H=f["geo_H"], mask=f["geo_mask"], bc_x="closed", bc_y="closed")
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0, nu=1e-4,
                                     kappa=f["tke_kappa_b"]), closure="tke", mxl=mxl)
    st = hoc.HipStepper(dom, phys, SchemeParams(name="fb"), 300.0, ustar2=f["tke_ustar2"])
    st.nu.copy_(torch.as_tensor(f["tke_Km"], device="cuda"))
    st.kappa.copy_(torch.as_tensor(f["tke_Kh"], device="cuda"))
    e = st.tke_step(f["tke_e"], f["tke_u"], f["tke_v"], f["tke_b"])
    assert np.array_equal(_np(e), f[f"tke_{mxl}_e"])
    assert np.array_equal(_np(st.nu), f[f"tke_{mxl}_Km"])
    assert np.array_equal(_np(st.kappa), f[f"tke_{mxl}_Kh"])


def test_array_forcing_step_matches_reference(hoc):
    """Array tau_x(y), tau_y(x) and kappa_b(x, y) (the basin's forcing)
    through 3 steps: the reference's own states, K_m and K_h bit for bit,
    iteration counts step by step."""
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    f = load_golden("step_theta_array_forcing.npz")
    dom, phys, params, dt, steps, s0 = build_array_forcing_case(f)
    st = hoc.HipStepper(dom, phys, params, dt)
    assert np.array_equal(_np(st.kappa), np.broadcast_to(f["kappa_b"], st.kappa.shape))
    s = State3D(**s0, t=0.0)
    for n in range(1, steps + 1):
        s = st.step(s)
        for k in FIELDS:
            got = _np(getattr(s, k))
            assert np.array_equal(got, f[f"s{n}_{k}"]), (n, k, rel_l2(got, f[f"s{n}_{k}"]))
        assert np.array_equal(_np(st.nu), f[f"s{n}_nu"])
        assert np.array_equal(_np(st.kappa), f[f"s{n}_kappa"])
        assert st.solver_iterations == int(f[f"s{n}_iters"]), n


@pytest.mark.parametrize("limited", [False, True])
def test_momentum_advection_walls_matches_reference(hoc, limited):
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    f = load_golden("operators.npz")
    grid = CGrid(nx=24, ny=16, Lx=2.4e5, Ly=1.6e5, nz=10, depth=1000.0)
    dom = Domain(grid=grid, H=f["geo_H"], mask=f["geo_mask"], bc_x="closed", bc_y="periodic")
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0))
    st = hoc.HipStepper(dom, phys, SchemeParams(name="fb", advection="up3_tvd" if limited
                                                else "up3"), 10.0)
    au, av = st.momentum_advection(f["adv_u"], f["adv_v"], f["adv_w"])
    tag = "tvd" if limited else "up3"
    assert np.array_equal(_np(au), f[f"adv_{tag}_u"])
    assert np.array_equal(_np(av), f[f"adv_{tag}_v"])


@pytest.mark.parametrize("pgf", ["keep", "remove"])
def test_pressure_gradient_matches_reference(hoc, pgf):
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    f = load_golden("operators.npz")
    grid = CGrid(nx=24, ny=16, Lx=2.4e5, Ly=1.6e5, nz=10, depth=1000.0)
    dom = Domain(grid=grid, H=f["geo_H"], mask=f["geo_mask"], bc_x="closed", bc_y="periodic")
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0), pgf=pgf)
    st = hoc.HipStepper(dom, phys, SchemeParams(name="fb"), 10.0)
    px, py = st.pressure_gradients(f["pg_b"])
    assert np.array_equal(_np(px), f[f"pg_{pgf}_x"])
    assert np.array_equal(_np(py), f[f"pg_{pgf}_y"])


@pytest.mark.parametrize("scheme", ["centered2", "upwind1", "up3", "up3_tvd"])
def test_tracer_advection_matches_oracle(hoc, scheme):
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    from oracle.stepper import OracleStepper
    f = load_golden("operators.npz")
    grid = CGrid(nx=24, ny=16, Lx=2.4e5, Ly=1.6e5, nz=10, depth=1000.0)
    dom = Domain(grid=grid, H=f["geo_H"], mask=f["geo_mask"], bc_x="closed", bc_y="periodic")
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0))
    params = SchemeParams(name="fb", advection=scheme)
    rng = np.random.default_rng(6)
    c = (10.0 + rng.standard_normal(grid.shape3d)) * dom.mask3
    want = OracleStepper(dom, phys, params, 10.0).tracer_advection(c, f["adv_u"], f["adv_v"],
                                                                   f["adv_w"])
    got = hoc.HipStepper(dom, phys, params, 10.0).tracer_advection(c, f["adv_u"], f["adv_v"],
                                                                    f["adv_w"])
    assert np.array_equal(_np(got), want)


@pytest.mark.parametrize("kind", ["pcg_jacobi", "pcg_rbgs", "rbgs", "multigrid"])
def test_helmholtz_matches_reference(hoc, kind):
    """Every reference solver kind on the fixture's 24 x 16 island problem:
    the reference's solution bit for bit and its iteration count."""
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    f = load_golden("operators.npz")
    grid = CGrid(nx=24, ny=16, Lx=2.4e5, Ly=1.6e5, nz=10, depth=1000.0)
    dom = Domain(grid=grid, H=f["geo_H"], mask=f["geo_mask"], bc_x="closed", bc_y="periodic")
    # coef = g theta^2 dt^2 with theta = 0.5, dt = 600 (the fixture's coef)
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=1000.0))
    params = SchemeParams(name="theta", solver=kind, rtol=1e-11, max_iter=5000)
    st = hoc.HipStepper(dom, phys, params, 600.0)
    assert st.cfg.g * 0.25 * 600.0 ** 2 == float(f["helm_coef"])
    x, iters, conv = st.helmholtz_solve(f["helm_Ku"], f["helm_Kv"], f["helm_rhs"])
    assert conv and iters == int(f[f"helm_{kind}_iters"])
    assert np.array_equal(_np(x), f[f"helm_{kind}_x"])


def _random_helmholtz(shape, bc, seed=7):
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    ny, nx = shape
    rng = np.random.default_rng(seed)
    grid = CGrid(nx=nx, ny=ny, Lx=nx * 1.0e4, Ly=ny * 1.0e4, nz=4, depth=4000.0)
    mask = (rng.random(grid.shape) > 0.1).astype(np.float64)
    dom = Domain(grid=grid, H=np.full(grid.shape, 4000.0), mask=mask, bc_x=bc,
                 bc_y=bc)
    Ku = (1000.0 + 3000.0 * rng.random(grid.shape)) * dom.mask_u
    Kv = (1000.0 + 3000.0 * rng.random(grid.shape)) * dom.mask_v
    rhs = rng.standard_normal(grid.shape) * mask
    return grid, dom, mask, Ku, Kv, rhs


@pytest.mark.parametrize("kind,shape,bc", [
    ("pcg_rbgs", (300, 500), "closed"), ("pcg_rbgs", (25, 33), "periodic"),
    ("rbgs", (40, 24), "closed"), ("rbgs", (15, 21), "periodic"),
    ("multigrid", (300, 500), "closed"), ("multigrid", (100, 60), "periodic"),
    ("pcg_mg", (512, 512), "closed"), ("pcg_mg", (300, 500), "periodic"),
    ("pcg_mg", (16, 24), "periodic"), ("pcg_mg", (13, 24), "closed")])
def test_helmholtz_kinds_match_oracle(hoc, kind, shape, bc):
    """Each solver kind against the oracle, bitwise, iteration counts equal:
    odd sides (no coarsening, or coarse levels with odd sides and periodic
    wrap, where red-black neighbours of one colour meet), several chunks with
    and without a trailing partial chunk."""
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    from oracle.helmholtz import HelmholtzVar
    grid, dom, mask, Ku, Kv, rhs = _random_helmholtz(shape, bc)
    ny, nx = shape
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=4000.0))
    rtol = 1e-8 if kind == "rbgs" else 1e-12
    params = SchemeParams(name="theta", solver=kind, rtol=rtol, max_iter=5000)
    st = hoc.HipStepper(dom, phys, params, 600.0)
    ref = HelmholtzVar(nx, ny, grid.dx, grid.dy, st.cfg.g * 0.25 * 600.0 ** 2, Ku, Kv, mask,
                       kind=kind, rtol=rtol, max_iter=5000)
    want, rep = ref.solve(rhs)
    for call in range(2):  # the second call starts from the last call's batch size
        x, iters, conv = st.helmholtz_solve(Ku, Kv, rhs)
        assert conv == rep.converged and iters == rep.iterations, (call, iters, rep.iterations)
        assert np.array_equal(_np(x), want), (call, rel_l2(_np(x), want))


FIELDS = ("u", "v", "b", "eta", "T", "S", "e")


@pytest.mark.parametrize("shape,bc,levels,sweeps,omega", [
    ((512, 512), "closed", 5, 1, 0.8), ((512, 512), "closed", 5, 3, 0.7),
    ((256, 384), "periodic", 5, 4, 0.8), ((300, 500), "closed", 2, 2, 0.8),
    ((200, 136), "periodic", 6, 2, 0.9), ((64, 64), "closed", 1, 2, 0.8),
    ((512, 512), "periodic", 7, 2, 0.8)])
def test_pcg_mg_variants_match_oracle(hoc, shape, bc, levels, sweeps, omega):
    """pcg_mg over its knobs: the fused V-cycle legs (1-3 sweeps per side),
    the per-sweep path (4 sweeps), a coarsest level solved in one workgroup
    and one too large for it (150 x 250), odd coarse sides with periodic
    wrap (200 x 136 -> 25 x 17), a single level, and 7 levels (512 -> 8: a
    coarse cycle of three levels, coarse faces six levels below the fine
    ones, formed in one launch)."""
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    from oracle.helmholtz import HelmholtzVar
    grid, dom, mask, Ku, Kv, rhs = _random_helmholtz(shape, bc, seed=11)
    ny, nx = shape
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=4000.0))
    params = SchemeParams(name="theta", solver="pcg_mg", rtol=1e-12, max_iter=5000,
                          mg_levels=levels, mg_sweeps=sweeps, mg_omega=omega)
    st = hoc.HipStepper(dom, phys, params, 600.0)
    ref = HelmholtzVar(nx, ny, grid.dx, grid.dy, st.cfg.g * 0.25 * 600.0 ** 2, Ku, Kv, mask,
                       kind="pcg_mg", rtol=1e-12, max_iter=5000, mg_levels=levels,
                       mg_sweeps=sweeps, mg_omega=omega)
    want, rep = ref.solve(rhs)
    x, iters, conv = st.helmholtz_solve(Ku, Kv, rhs)
    assert conv == rep.converged and iters == rep.iterations, (iters, rep.iterations)
    assert np.array_equal(_np(x), want), rel_l2(_np(x), want)


@pytest.mark.parametrize("shape,bc,levels,sweeps,tile", [
    ((512, 512), "closed", 5, 2, 0), ((512, 512), "closed", 5, 2, 1),
    ((512, 512), "closed", 5, 2, 2), ((512, 512), "closed", 5, 2, 3),
    ((512, 512), "closed", 5, 1, 2), ((512, 512), "periodic", 5, 3, 2),
    ((300, 500), "closed", 2, 2, 0), ((200, 136), "periodic", 6, 2, 2),
    ((64, 64), "closed", 1, 2, 0), ((96, 80), "periodic", 5, 2, 1)])
def test_pcg_mg_rbgs_smoother_matches_oracle(hoc, shape, bc, levels, sweeps, tile):
    """pcg_mg with the red-black Gauss-Seidel smoother: the fused legs at
    1-3 sweeps per side and every tile shape (partial tiles at 96 x 80),
    the one-workgroup coarse cycle (with odd periodic sides at 200 x 136 ->
    17 x 25, and a 64 x 64 single level holding three fields in LDS), and a
    coarse level too large for one workgroup (150 x 250, separate sweeps)."""
    from hydrostatic_ocean_core_benchmark_amd import _lib
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    from oracle.helmholtz import HelmholtzVar
    grid, dom, mask, Ku, Kv, rhs = _random_helmholtz(shape, bc, seed=13)
    ny, nx = shape
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=4000.0))
    params = SchemeParams(name="theta", solver="pcg_mg", rtol=1e-12, max_iter=5000,
                          mg_levels=levels, mg_sweeps=sweeps, mg_smoother="rbgs")
    st = hoc.HipStepper(dom, phys, params, 600.0)
    _lib.check(st.lib.hoc_stepper_set_mg_tile(st._h, tile))
    ref = HelmholtzVar(nx, ny, grid.dx, grid.dy, st.cfg.g * 0.25 * 600.0 ** 2, Ku, Kv, mask,
                       kind="pcg_mg", rtol=1e-12, max_iter=5000, mg_levels=levels,
                       mg_sweeps=sweeps, mg_smoother="rbgs")
    want, rep = ref.solve(rhs)
    x, iters, conv = st.helmholtz_solve(Ku, Kv, rhs)
    assert conv == rep.converged and iters == rep.iterations, (iters, rep.iterations)
    assert np.array_equal(_np(x), want), rel_l2(_np(x), want)


@pytest.mark.parametrize("name", list(COMPOSED))
def test_composed_step_matches_reference(hoc, name):
    """1-3 steps of the reference stepper (each fixture's own solver kind):
    every field bitwise and the solver iterations step by step."""
    f = load_golden(f"step_{name}.npz")
    dom, phys, params, dt, steps, s0 = build_case(name, f)
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    st = hoc.HipStepper(dom, phys, params, dt)
    s = State3D(**s0, t=0.0)
    for n in range(1, steps + 1):
        s = st.step(s)
        want = {k: f[f"s{n}_{k}"] for k in FIELDS if f"s{n}_{k}" in f}
        for k, w in want.items():
            got = _np(getattr(s, k))
            assert rel_l2(got, w) < 1e-12, (n, k, rel_l2(got, w))
            assert np.array_equal(got, w), (n, k, "not bitwise", rel_l2(got, w))
        assert st.solver_iterations == int(f[f"s{n}_iters"]), n
        if f"s{n}_substeps" in f:
            assert st.barotropic_substeps == int(f[f"s{n}_substeps"]), n


@pytest.mark.parametrize("name", list(COMPOSED))
def test_composed_integrate_matches_reference(hoc, name):
    """The same fixtures through integrate (hoc_stepper_step_n: from the
    second step on, w of the input velocity is the one the previous step
    formed, on a static geometry): the final state bitwise and the total
    solver iterations."""
    f = load_golden(f"step_{name}.npz")
    dom, phys, params, dt, steps, s0 = build_case(name, f)
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    st = hoc.HipStepper(dom, phys, params, dt)
    s = st.integrate(State3D(**s0, t=0.0), steps)
    for k in FIELDS:
        if f"s{steps}_{k}" in f:
            got, w = _np(getattr(s, k)), f[f"s{steps}_{k}"]
            assert np.array_equal(got, w), (k, "not bitwise", rel_l2(got, w))
    assert st.solver_iterations == int(f[f"s{steps}_iters"])
    t = 0.0
    for _ in range(steps):
        t = t + dt
    assert s.t == t


@pytest.mark.parametrize("name", list(COMPOSED))
def test_composed_step_both_geometry_forms(hoc, name):
    """The kernels read a static z-level geometry in column form (wet-level
    counts per column) and a moving one from the stored arrays. Every z-level
    case must get the column form, every z* / sigma case the arrays; and a
    z-level case forced back to the stored arrays must reproduce the same
    reference fixture bit for bit, so both forms are pinned."""
    f = load_golden(f"step_{name}.npz")
    dom, phys, params, dt, steps, s0 = build_case(name, f)
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    st = hoc.HipStepper(dom, phys, params, dt)
    moving = dom.vcoord != "zlevel"
    assert st.geometry_form() == (0 if moving else 1), name
    if moving:
        with pytest.raises(NotImplementedError):
            st.geometry_form(1)
        return
    assert st.geometry_form(0) == 0
    s = State3D(**s0, t=0.0)
    for n in range(1, steps + 1):
        s = st.step(s)
        for k in FIELDS:
            key = f"s{n}_{k}"
            if key in f:
                assert np.array_equal(_np(getattr(s, k)), f[key]), (n, k)
        assert st.solver_iterations == int(f[f"s{n}_iters"]), n


@pytest.mark.parametrize("name", list(COMPOSED))
def test_composed_step_pcg_mg_matches_oracle(hoc, name):
    """The same cases with the pcg_mg solver (no reference counterpart)
    against the oracle, bitwise."""
    f = load_golden(f"step_{name}.npz")
    dom, phys, params, dt, steps, s0 = build_case(name, f)
    if params.name != "theta":
        pytest.skip("forward-backward scheme has no Helmholtz solve")
    params.solver = "pcg_mg"
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    from oracle.stepper import OracleStepper, OState
    ost = OracleStepper(dom, phys, params, dt)
    os_ = OState(**s0, t=0.0)
    st = hoc.HipStepper(dom, phys, params, dt)
    s = State3D(**s0, t=0.0)
    for n in range(1, steps + 1):
        s = st.step(s)
        os_ = ost.step(os_)
        for k in FIELDS:
            w = getattr(os_, k)
            if w is None:
                continue
            got = _np(getattr(s, k))
            assert np.array_equal(got, w), (n, k, "not bitwise", rel_l2(got, w))
        assert st.solver_iterations == ost.solver_iterations, n


def test_lock_exchange_200_steps_matches_reference(hoc):
    """The 128x4x20 lock exchange for 200 steps against the reference
    (full-run gate 1e-9; iteration counts step by step)."""
    f = load_golden("lock_exchange_128x4x20.npz")
    dom, phys, params, dt = lock_exchange_problem()
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    st = hoc.HipStepper(dom, phys, params, dt)
    g = dom.grid
    s = State3D(u=np.zeros(g.shape3d), v=np.zeros(g.shape3d), b=f["in_b"],
                eta=np.zeros(g.shape), T=f["in_T"], S=f["in_S"])
    iters = []
    for n in range(1, 201):
        before = st.solver_iterations
        s = st.step(s)
        iters.append(st.solver_iterations - before)
        if n in (1, 10, 200):
            tol = 1e-12 if n == 1 else 1e-9
            for k in FIELDS:
                got = _np(getattr(s, k))
                assert rel_l2(got, f[f"s{n}_{k}"]) < tol, (n, k)
    assert iters == [int(x) for x in f["iters_per_step"]]
    assert np.array_equal(_np(s.T), f["s200_T"])


@pytest.mark.parametrize("shape", [(300, 500), (512, 512), (16, 24)])
def test_helmholtz_persistent_equals_multikernel_and_oracle(hoc, shape):
    """The persistent (one cooperative launch) and the per-phase PCG give the
    oracle's iterates bit for bit, with and without a trailing partial
    8192-element chunk (300 x 500 = 18 full chunks + 2544)."""
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    from oracle.helmholtz import HelmholtzVar
    ny, nx = shape
    grid, dom, mask, Ku, Kv, rhs = _random_helmholtz(shape, "closed")
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=4000.0))
    params = SchemeParams(name="theta", solver="pcg_jacobi", rtol=1e-12, max_iter=5000)
    st = hoc.HipStepper(dom, phys, params, 600.0)
    ref = HelmholtzVar(nx, ny, grid.dx, grid.dy, st.cfg.g * 0.25 * 600.0 ** 2, Ku, Kv, mask,
                       kind="pcg_jacobi", rtol=1e-12, max_iter=5000)
    want, rep = ref.solve(rhs)
    results = {}
    for path in ("persistent", "multikernel"):
        st.solver_path = path
        x, iters, conv = st.helmholtz_solve(Ku, Kv, rhs)
        results[path] = (st.solver_path, _np(x), iters, conv)
    assert results["multikernel"][0] == "multikernel"
    for path, (used, x, iters, conv) in results.items():
        assert conv and iters == rep.iterations, (path, used, iters, rep.iterations)
        assert np.array_equal(x, want), (path, used, rel_l2(x, want))


@pytest.mark.parametrize("kind", ["pcg_jacobi", "pcg_mg"])
def test_helmholtz_large_grid_matches_oracle(hoc, kind):
    """A 1450 x 1450 grid (256 full reduction chunks and a partial one,
    just past the size where the PCG phase scalars are formed once per grid
    by k_pcg_totals instead of in every workgroup): the solution and the
    iteration count stay the oracle's bit for bit."""
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    from oracle.helmholtz import HelmholtzVar
    shape = (1450, 1450)
    grid, dom, mask, Ku, Kv, rhs = _random_helmholtz(shape, "closed", seed=13)
    ny, nx = shape
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=4000.0))
    params = SchemeParams(name="theta", solver=kind, rtol=1e-10, max_iter=3000)
    st = hoc.HipStepper(dom, phys, params, 600.0)
    ref = HelmholtzVar(nx, ny, grid.dx, grid.dy, st.cfg.g * 0.25 * 600.0 ** 2, Ku, Kv, mask,
                       kind=kind, rtol=1e-10, max_iter=3000)
    want, rep = ref.solve(rhs)
    x, iters, conv = st.helmholtz_solve(Ku, Kv, rhs)
    print(f"large {kind}: {iters} iterations (oracle {rep.iterations})", flush=True)
    assert conv == rep.converged and iters == rep.iterations
    assert np.array_equal(_np(x), want)


def test_step_rejects_an_output_that_aliases_the_input(hoc):
    """The step reads the old state after writing parts of the new one, so
    stepping a state in place is refused before any kernel runs."""
    from hydrostatic_ocean_core_benchmark_amd.domain import CGrid, Domain
    from hydrostatic_ocean_core_benchmark_amd.types import Physics3D, PhysicsV05, SchemeParams
    grid = CGrid(nx=16, ny=8, Lx=1.6e5, Ly=8.0e4, nz=4, depth=100.0)
    dom = Domain(grid=grid, H=np.full(grid.shape, 100.0), mask=np.ones(grid.shape))
    phys = PhysicsV05(base=Physics3D(g=9.80616, f=1e-4, H=100.0))
    st = hoc.HipStepper(dom, phys, SchemeParams(name="theta", solver="pcg_jacobi"), 60.0)
    bufs = st._new_state_buffers()
    for t in bufs.values():
        t.zero_()
    with pytest.raises(ValueError, match="must not alias"):
        st._step_into(bufs, bufs)
    other = st._new_state_buffers()
    crossed = dict(other, u=bufs["v"])          # out.u is in.v
    with pytest.raises(ValueError, match="must not alias"):
        st._step_into(bufs, crossed)
    twice = dict(other, v=other["u"])           # out.u and out.v one array
    with pytest.raises(ValueError, match="must be distinct"):
        st._step_into(bufs, twice)
