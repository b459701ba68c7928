"""The C-ABI library (no GPU needed): it loads, it exports every function
include/hoc.h declares, and the ctypes mirrors of the ABI structs have the
C layout (sizes and field offsets checked against a gcc-compiled probe)."""

from __future__ import annotations

import ctypes
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "hoc.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(hoc_\w+)\s*\(", text, re.M)))


def test_header_declares_the_step_api():
    names = declared_functions()
    for required in ("hoc_stepper_create", "hoc_stepper_step", "hoc_stepper_destroy",
                     "hoc_stepper_counters", "hoc_last_error", "hoc_thomas_batched",
                     "hoc_eos_buoyancy", "hoc_pairwise_dot"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from hydrostatic_ocean_core_benchmark_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libhoc.so not built")
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _lib.PROTOTYPES, f"{name} has no ctypes prototype"
    assert isinstance(lib.hoc_last_error(), bytes)


def test_ctypes_structs_match_c_layout(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    from hydrostatic_ocean_core_benchmark_amd import _lib
    structs = {"hoc_config": _lib.HocConfig, "hoc_geometry": _lib.HocGeometry,
               "hoc_state": _lib.HocState, "hoc_counters": _lib.HocCounters,
               "hoc_sw2d_config": _lib.HocSw2dConfig,
               "hoc_helm2d_config": _lib.HocHelm2dConfig,
               "hoc_m3_config": _lib.HocM3Config}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
             'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cname, key, val = line.split()
        got[(cname, key)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


def test_stepper_refuses_to_run_without_a_device():
    """No silent CPU fallback: on a machine without a HIP device the stepper
    raises instead of computing anything."""
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from hydrostatic_ocean_core_benchmark_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libhoc.so not built")
    import numpy as np
    from golden_cases import lock_exchange_problem
    from hydrostatic_ocean_core_benchmark_amd.stepper import HipStepper
    dom, phys, params, dt = lock_exchange_problem()
    with pytest.raises(RuntimeError):
        HipStepper(dom, phys, params, dt)
    del np


def test_stepper_precision_arguments_are_checked_before_any_device_use():
    """precision: fp64 | mixed, and mixed only with the pcg_mg theta solve;
    both refused with ValueError before the device is touched."""
    from golden_cases import lock_exchange_problem
    from hydrostatic_ocean_core_benchmark_amd.stepper import HipStepper
    dom, phys, params, dt = lock_exchange_problem()
    with pytest.raises(ValueError, match="unknown precision"):
        HipStepper(dom, phys, params, dt, precision="fp16")
    assert params.solver != "pcg_mg"
    for prec in ("mixed", "mixed_fp32_momentum"):
        with pytest.raises(ValueError, match="pcg_mg only"):
            HipStepper(dom, phys, params, dt, precision=prec)
    params.solver, params.mg_smoother = "pcg_mg", "rbgs"
    with pytest.raises(ValueError, match="rbgs pcg_mg smoother"):
        HipStepper(dom, phys, params, dt, precision="mixed")
    params.mg_sweeps = 4
    with pytest.raises(ValueError, match="rbgs pcg_mg smoother"):
        HipStepper(dom, phys, params, dt)


@pytest.mark.parametrize("module,cls", [("stepper", "HipStepper"), ("model2d", "Helmholtz2D"),
                                        ("model2d", "SWStepper"), ("model3d_v02", "Stepper3D")])
def test_finaliser_survives_module_teardown(monkeypatch, module, cls):
    """At interpreter exit, module teardown sets a module's globals to None
    before its last objects are collected; a finaliser that looked up
    `sys` there raised AttributeError. The finalisers bind
    sys.is_finalizing when the class is created."""
    import importlib
    mod = importlib.import_module(f"hydrostatic_ocean_core_benchmark_amd.{module}")
    obj = object.__new__(getattr(mod, cls))   # no library handle
    monkeypatch.setattr(mod, "sys", None)
    obj.__del__()
