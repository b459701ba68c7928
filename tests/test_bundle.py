"""The v0.5 bundle boundary (namelist + domain/init binaries -> final state +
metrics JSON): parsing, the writer/reader round trip, the compiled backends'
validation messages (CPU), and a bundle run on the GPU against the
reference's own fixture."""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import load_golden
from golden_cases import build_case

from hydrostatic_ocean_core_benchmark_amd import bundle as B
from hydrostatic_ocean_core_benchmark_amd.types import State3D

ROOT = Path(__file__).resolve().parents[1]


def _write(tmp_path, name="theta_pcgj_tke_tvd_teos", steps=None, **over):
    f = load_golden(f"step_{name}.npz")
    dom, phys, params, dt, n, s0 = build_case(name, f)
    for k, v in over.items():
        setattr(params, k, v)
    s = State3D(u=s0["u"], v=s0["v"], b=s0["b"], eta=s0["eta"], T=s0["T"], S=s0["S"])
    nml = B.write_bundle(tmp_path / "case", dom, phys, params, s, dt, steps or n,
                         case_name=name)
    return nml, (dom, phys, params, dt, steps or n, s0, f)


def test_fortran_literals():
    assert B.fortran_double("1.00000000000000000d-12") == 1e-12
    assert B.fortran_double("3.5D0") == 3.5
    assert B.fortran_double("2.0e3") == 2000.0
    for x in (0.1, 1.0 / 3.0, 9.80616, -5e-5, 5.12e6, 1e-300):
        assert B.fortran_double(B.format_double(x)) == x
    assert B.unquote("'pcg_mg'") == "pcg_mg" and B.unquote('"x"') == "x"


def test_namelist_reader(tmp_path):
    p = tmp_path / "a.nml"
    p.write_text("! header\n&grid_nml\n  nx = 12 ! trailing\n  lx = 1.5d3\n/\n"
                 "&scheme_nml\n  solver_kind = 'rbgs'\n  nx = 13\n/\n")
    kv = B.read_namelist(p)
    assert kv == {"nx": "13", "lx": "1.5d3", "solver_kind": "'rbgs'"}


def test_bundle_round_trip(tmp_path):
    nml, (dom, phys, params, dt, steps, s0, f) = _write(tmp_path)
    b = B.parse_bundle(nml)
    assert b.dt == dt and b.n_steps == steps and b.case_name == "theta_pcgj_tke_tvd_teos"
    assert np.array_equal(b.domain.H, dom.H) and np.array_equal(b.domain.mask, dom.mask)
    assert np.array_equal(b.domain.dz3_ref, dom.dz3_ref)
    for k in ("g", "f", "H", "nu", "kappa", "rho0", "tau_x", "tau_y", "alpha_T", "A_h",
              "K_h", "eos", "pgf", "closure", "mxl", "c_k", "e_min", "tracers"):
        assert getattr(b.physics, k) == getattr(phys, k), k
    for k in ("name", "theta", "theta_cor", "theta_v", "n_picard", "advection", "solver",
              "rtol", "max_iter"):
        assert getattr(b.params, k) == getattr(params, k), k
    g = dom.grid
    init = B.read_state_file(b.init_file, g.nx, g.ny, g.nz, b.ts)
    for k in ("eta", "u", "v", "b", "T", "S"):
        assert np.array_equal(init[k], s0[k]), k


@pytest.mark.parametrize("edit,message", [
    (("scheme_name = 'theta'", "scheme_name = 'leapfrog'"), "unknown scheme leapfrog"),
    (("theta = 5.0", "theta = 1.5"), "theta must lie in (0,1]"),
    (("n_steps = 3", "n_steps = 0"), "dt, n_steps and n_picard must come from the driver"),
    (("mxl = 'integral'", "mxl = 'other'"), "mxl must be integral|recursive"),
    (("vcoord = 'zlevel'", "vcoord = 'sigma'"), "vcoord=zlevel only"),
    (("bc_x = 'closed'", "bc_x = 'open'"), "periodic|closed only"),
])
def test_bundle_validation(tmp_path, edit, message):
    nml, _ = _write(tmp_path)
    text = nml.read_text()
    old, new = edit
    if old == "theta = 5.0":
        text = text.replace("  theta = 5.00000000000000000d-01", "  theta = 1.5")
    else:
        assert old in text, old
        text = text.replace(old, new)
    nml.write_text(text)
    with pytest.raises(B.BundleError, match=message.replace("(", r"\(").replace(")", r"\)")
                       .replace("|", r"\|")):
        B.parse_bundle(nml)


def test_bundle_cli_reports_fatal(tmp_path):
    """An invalid bundle: `FATAL:` on stderr, exit status 1, before any GPU use."""
    nml, _ = _write(tmp_path)
    nml.write_text(nml.read_text().replace("mxl = 'integral'", "mxl = 'x'"))
    r = subprocess.run([sys.executable, "-m", "hydrostatic_ocean_core_benchmark_amd.bundle",
                        str(nml)], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and r.stderr.startswith("FATAL: mxl must be integral|recursive")


def test_missing_files(tmp_path):
    nml, _ = _write(tmp_path)
    Path(str(tmp_path / "case") + "_domain.bin").unlink()
    with pytest.raises(B.BundleError, match="cannot read|expected"):
        B.parse_bundle(nml)


@pytest.mark.gpu
def test_bundle_run_matches_reference(tmp_path, gpu_available):
    """The 128x4x20 lock exchange as a bundle, 10 steps, 1 warm-up + 2 timed
    repeats from the same initial state (TKE and mixing reset each run): the
    state file holds the reference's state after 10 steps, bit for bit, and
    the metrics carry the compiled backends' keys and per-run counts."""
    from golden_cases import lock_exchange_problem
    f = load_golden("lock_exchange_128x4x20.npz")
    dom, phys, params, dt = lock_exchange_problem()
    g = dom.grid
    zero3 = np.zeros(g.shape3d)
    s0 = State3D(u=zero3, v=zero3, b=f["in_b"], eta=np.zeros(g.shape), T=f["in_T"],
                 S=f["in_S"])
    nml = B.write_bundle(tmp_path / "lock", dom, phys, params, s0, dt, 10,
                         case_name="lock_exchange", n_repeat=2, n_warmup=1)
    m = B.run_bundle(nml, log=lambda *a: None)
    got = B.read_state_file(str(tmp_path / "lock") + "_state3d5.bin", g.nx, g.ny, g.nz, True)
    for k in ("eta", "u", "v", "b", "T", "S"):
        assert np.array_equal(got[k], f[f"s10_{k}"]), k
    keys = ("backend gpu nx ny nz cells n_steps scheme solver eos advection tracers theta dt "
            "solver_iterations solver_failures tridiagonal_solves solver_rebuilds n_split "
            "barotropic_substeps device_bytes bytes_per_cell diverged wall_s wall_mad_s "
            "n_repeat omp_num_threads").split()
    on_disk = json.loads(Path(str(tmp_path / "lock") + "_metrics.json").read_text())
    assert list(on_disk) == keys and on_disk == m
    assert m["n_repeat"] == 2 and not m["diverged"] and m["cells"] == g.nx * g.ny * g.nz
    assert m["solver_iterations"] == int(np.sum(f["iters_per_step"][:10]))


def test_launcher_runs_the_bundle_entry_point(tmp_path):
    """cfd_exp3d5_hip (the backend binary for the reference's gate scripts)
    reaches the bundle runner: a missing namelist is the runner's FATAL, exit 1,
    before any device use."""
    import os
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parents[1] / "hydrostatic_ocean_core_benchmark_amd" / "cfd_exp3d5_hip"
    assert os.access(exe, os.X_OK)
    proc = subprocess.run([str(exe), str(tmp_path / "missing.nml")], capture_output=True,
                          text=True, timeout=120)
    assert proc.returncode == 1
    assert "FATAL" in proc.stderr
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 1 and "usage" in proc.stderr


REF_BUNDLE = ROOT / "tests" / "golden" / "ref_bundle"


def _ref_bundle_copy(tmp_path, monkeypatch):
    """The bundle the reference's writer produced (its paths are relative to
    the folder it was written in), copied into a scratch working folder."""
    import shutil
    for p in REF_BUNDLE.iterdir():
        shutil.copy(p, tmp_path / p.name)
    monkeypatch.chdir(tmp_path)
    return tmp_path / "ref_lock.nml"


def test_reference_written_bundle_parses_and_rewrites_identically(tmp_path, monkeypatch):
    """tools/toml2nml.py's own output (lock_exchange_v06 at 16x16x8): parsed
    here, and written back by this package's writer, the domain and initial
    state files are byte-identical to the reference's and every namelist key
    the reference writes holds the same value."""
    nml = _ref_bundle_copy(tmp_path, monkeypatch)
    b = B.parse_bundle(nml)
    exp = np.load(REF_BUNDLE / "expected.npz")
    g = b.domain.grid
    assert (g.nx, g.ny, g.nz, b.n_steps) == (16, 16, 8, 4) and b.dt == float(exp["dt"])
    assert b.params.name == "theta" and b.params.solver == "pcg_jacobi"
    assert b.physics.closure == "tke" and b.physics.eos == "teos10" and b.ts
    init = B.read_state_file(b.init_file, g.nx, g.ny, g.nz, True)
    s0 = State3D(u=init["u"], v=init["v"], b=init["b"], eta=init["eta"], T=init["T"],
                 S=init["S"])
    out = B.write_bundle(tmp_path / "mine", b.domain, b.physics, b.params, s0, b.dt,
                         b.n_steps, case_name=b.case_name, out_prefix="ref_lock")
    for suffix in ("_domain.bin", "_init.bin"):
        assert (tmp_path / f"mine{suffix}").read_bytes() == \
            (tmp_path / f"ref_lock{suffix}").read_bytes(), suffix
    ref_kv, my_kv = B.read_namelist(nml), B.read_namelist(out)
    skip = {"domain_file", "init_file", "n_repeat", "n_warmup", "omp_min_points",
            "tridiag_kernel", "pcg_sync", "pcg_check_every"}
    for key, raw in ref_kv.items():
        if key in skip:
            continue
        assert key in my_kv, key
        try:
            same = B.fortran_double(raw) == B.fortran_double(my_kv[key])
        except ValueError:
            same = B.unquote(raw) == B.unquote(my_kv[key])
        assert same, (key, raw, my_kv[key])


@pytest.mark.gpu
def test_reference_written_bundle_runs_to_the_reference_state(tmp_path, monkeypatch,
                                                              gpu_available):
    """Run the bundle the reference's writer produced: the final state is
    the reference stepper's after its 4 steps on the same problem, bit for
    bit, with the same solver iterations."""
    nml = _ref_bundle_copy(tmp_path, monkeypatch)
    m = B.run_bundle(nml, log=lambda *a: None)
    exp = np.load(REF_BUNDLE / "expected.npz")
    got = B.read_state_file(tmp_path / "ref_lock_state3d5.bin", 16, 16, 8, True)
    for k in ("eta", "u", "v", "b", "T", "S"):
        assert np.array_equal(got[k], exp[k]), k
    assert m["solver_iterations"] == int(np.sum(exp["iters_per_step"]))
    assert m["solver_failures"] == 0 and not m["diverged"]


def test_oracle_runs_the_reference_bundle_to_the_reference_state(tmp_path, monkeypatch):
    """The problem parsed from the reference-written bundle, stepped by the
    oracle, gives the reference's state (pins the parser's reading of every
    namelist key that reaches the step)."""
    from oracle.stepper import OracleStepper, OState
    nml = _ref_bundle_copy(tmp_path, monkeypatch)
    b = B.parse_bundle(nml)
    g = b.domain.grid
    init = B.read_state_file(b.init_file, g.nx, g.ny, g.nz, True)
    ost = OracleStepper(b.domain, b.physics, b.params, b.dt)
    s = ost.integrate(OState(u=init["u"], v=init["v"], b=init["b"], eta=init["eta"],
                             T=init["T"], S=init["S"]), b.n_steps)
    exp = np.load(REF_BUNDLE / "expected.npz")
    for k in ("eta", "u", "v", "b", "T", "S"):
        assert np.array_equal(getattr(s, k), exp[k]), k
    assert ost.solver_iterations == int(np.sum(exp["iters_per_step"]))
