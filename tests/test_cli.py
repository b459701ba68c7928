"""The v0.5 route of the CLI (main.py run|verify|bench): the [run] backend and
precision keys (config/settings.toml; reference config/settings.toml:23-26)
select what runs or stop with an error, and `main.py run` on the lock
exchange writes the reference's own state (GPU)."""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import load_golden

from hydrostatic_ocean_core_benchmark_amd import cli
from hydrostatic_ocean_core_benchmark_amd.config import apply_overrides, load_config

ROOT = Path(__file__).resolve().parents[1]
CONFIG = ROOT / "config"


def _config(*sets):
    return apply_overrides(load_config(str(CONFIG)), list(sets))


@pytest.mark.parametrize("value,want", [("float64", "fp64"), ("mixed", "mixed"),
                                        ("mixed_fp32_momentum", "mixed_fp32_momentum")])
def test_run_precision_selects_the_step(value, want):
    cfg = _config(f"run.precision={value}")
    assert cli.run_settings(cfg) == ("hip", want)
    prob = cli.build_problem(cfg, "basin", 32, None, 2)
    assert prob.meta["precision"] == want


@pytest.mark.parametrize("key,value,message", [
    ("run.precision", "float32", "not offered"),
    ("run.backend", "cuda", "not this package's backend"),
    ("run.backend", "ref", "not this package's backend")])
def test_unsupported_run_settings_exit(key, value, message):
    with pytest.raises(SystemExit, match=message):
        cli.build_problem(_config(f"{key}={value}"), "lock_exchange", None, None, 1)


def test_two_dimensional_models_refuse_mixed_precision():
    with pytest.raises(SystemExit, match="float64 only"):
        cli.main(["--config", str(CONFIG), "--set", "run.precision=mixed", "run", "--case",
                  "igw"])


def test_main_entry_reports_the_setting_error():
    res = subprocess.run([sys.executable, str(ROOT / "main.py"), "--config", str(CONFIG),
                          "--set", "run.backend=openacc", "run", "--case", "lock_exchange"],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode != 0 and "run.backend = 'openacc'" in res.stderr


@pytest.mark.gpu
def test_main_run_lock_exchange_matches_reference(tmp_path, gpu_available):
    """main.py run --case lock_exchange --steps 10 (the 128x4x20 benchmark
    lock exchange): the state file it writes is the reference's state after
    10 steps, bit for bit, and the metrics name the backend and precision."""
    from hydrostatic_ocean_core_benchmark_amd.bundle import read_state_file
    rc = cli.main(["--config", str(CONFIG), "--set", f"paths.output_dir={tmp_path}",
                   "run", "--case", "lock_exchange", "--steps", "10"])
    assert rc == 0
    (run,) = list(tmp_path.glob("run_lock_exchange_theta_*"))
    got = read_state_file(run / "state3d5.bin", 128, 4, 20, True)
    f = load_golden("lock_exchange_128x4x20.npz")
    for k in ("eta", "u", "v", "b", "T", "S"):
        assert np.array_equal(got[k], f[f"s10_{k}"]), k
    m = json.loads((run / "metrics.json").read_text())
    assert m["backend"] == "hip" and m["precision"] == "fp64" and m["n_steps"] == 10
    iters = [int(x) for x in f["iters_per_step"][:10]]
    assert m["solver_iterations"] == sum(iters)
