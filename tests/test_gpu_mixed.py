"""The mixed-precision steps (fp32 pcg_mg V-cycle preconditioner, fp32
tracer advective fluxes, fp32 anomaly-form EOS; fp64 momentum fluxes, state
updates, column solves and outer PCG -- and the variant with the momentum
fluxes in fp32 too) against the fp64 oracle and the plain fp64 GPU step. There is
no reference counterpart: the bar is an error budget per field, written in
each test, and the outer PCG's convergence (the fp64 residual still reaches
rtol, in the same number of iterations)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# Per-field relative L2 budgets of the mixed steps after the steps of each
# test, measured against fp64. SSH (eta) and T carry the stated target of
# 1e-6 relative; u, v, S and b have no stated target and keep budgets from
# measurement (the 1024 x 1024 x 60 basin after 50 steps gives eta 1.6e-6,
# u and v 1.1e-5, T 3.6e-8, S 4.6e-9, b 1.5e-8 for both forms: the momentum
# fluxes' precision moves none of these in the first three digits).
BUDGET = {"eta": 1e-6, "u": 1e-4, "v": 1e-4, "T": 1e-6, "S": 1e-7, "b": 1e-6}
MIXED = ("mixed", "mixed_fp32_momentum")


def _np(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


@pytest.fixture(scope="module")
def hoc(gpu_available):
    from hydrostatic_ocean_core_benchmark_amd import stepper
    return stepper


def _initial(prob, st):
    from hydrostatic_ocean_core_benchmark_amd.types import State3D
    g = prob.domain.grid
    b = _np(st.buoyancy_of(prob.initial_T, prob.initial_S)) * prob.domain.mask3
    return State3D(u=np.zeros(g.shape3d), v=np.zeros(g.shape3d), b=b, eta=np.zeros(g.shape),
                   T=prob.initial_T, S=prob.initial_S)


@pytest.mark.parametrize("precision", MIXED)
def test_mixed_step_against_the_oracle(hoc, precision):
    """64 x 64 x 40 basin (TEOS-10, up3_tvd, TKE, pcg_mg), 4 steps: every
    field of the mixed step within its budget of the fp64 oracle, and the
    same outer PCG iteration count step by step."""
    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    from oracle.stepper import OracleStepper, OState
    prob = basin(nx=64, nz=40)
    st = hoc.HipStepper(prob.domain, prob.physics, prob.params, prob.dt, precision=precision)
    ost = OracleStepper(prob.domain, prob.physics, prob.params, prob.dt)
    s = _initial(prob, st)
    os_ = OState(u=s.u, v=s.v, b=ost.buoyancy_of(prob.initial_T, prob.initial_S)
                 * prob.domain.mask3, eta=s.eta, T=s.T, S=s.S, t=0.0)
    for n in range(1, 5):
        it0, oit0 = st.solver_iterations, ost.solver_iterations
        s = st.step(s)
        os_ = ost.step(os_)
        assert st.solver_iterations - it0 == ost.solver_iterations - oit0, n
        assert st.solver_failures == 0
    for k, tol in BUDGET.items():
        err = rel_l2(_np(getattr(s, k)), getattr(os_, k))
        print(f"{precision} {k}: {err:.3e}", flush=True)
        assert err < tol, (k, err, tol)
        assert err > 0.0, (k, "mixed step equals fp64 bit for bit: fp32 path not taken")


def test_mixed_step_against_fp64_step_and_repeatable(hoc):
    """128 x 128 x 40 basin, 10 steps: the mixed step within budget of the
    plain fp64 GPU step (which the parity tests hold bitwise to the oracle),
    equal total PCG iterations, and two mixed runs bitwise identical."""
    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    prob = basin(nx=128, nz=40)
    out = {}
    for prec in ("fp64", "mixed", "mixed", "mixed_fp32_momentum", "mixed_fp32_momentum"):
        st = hoc.HipStepper(prob.domain, prob.physics, prob.params, prob.dt, precision=prec)
        s = st.integrate(_initial(prob, st), 10)
        out.setdefault(prec, []).append(({k: _np(getattr(s, k)).copy() for k in BUDGET},
                                         st.solver_iterations))
    ref, it64 = out["fp64"][0]
    for prec in MIXED:
        (mix, itm), (mix2, _) = out[prec]
        assert itm == it64, prec
        for k, tol in BUDGET.items():
            assert rel_l2(mix[k], ref[k]) < tol, (prec, k, rel_l2(mix[k], ref[k]))
            assert np.array_equal(mix[k], mix2[k]), (prec, k)


def test_mixed_rejects_other_solvers(hoc):
    from hydrostatic_ocean_core_benchmark_amd.cases import basin
    prob = basin(nx=64, nz=40, solver="pcg_jacobi")
    with pytest.raises(ValueError, match="pcg_mg only"):
        hoc.HipStepper(prob.domain, prob.physics, prob.params, prob.dt, precision="mixed")
