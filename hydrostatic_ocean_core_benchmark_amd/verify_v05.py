"""The v0.5 verification suite V5-1 .. V5-6 on the GPU (tools/verify_v05.py).

Every case either has an exact answer or a diagnostic that is zero for the
exact equations, so the measured number is the error itself:

  V5-1 reduce_v04        the v0.5 step against the v0.2-v0.4 step it
                         generalises, on a flat periodic z-level domain
                         (tools/verify_v05.py:50-93); theta legs gated < 1e-12,
                         split-explicit legs reported (they differ by design in
                         how the Coriolis term leaves the barotropic forcing)
  V5-2 operator_order    truncation order of D[K G eta] over a smooth
                         non-uniform depth for the three face rules (:113-176)
  V5-3 seamount_rest     resting stratification over flat / seamount / rough
                         topography, three vertical coordinates, with and
                         without the pressure-gradient correction (:179-220)
  V5-4 eos_check         polyTEOS10-bsq against its published check value (:96-110)
  V5-5 flather_radiate   reflection from open boundaries against closed walls (:250-269)
  V5-6 zstar_constancy   a uniform tracer under a moving free surface (:223-247)

Both steppers, the operator and the EOS run as HIP kernels; the set-up and
the diagnostics run on the host.

    python -m hydrostatic_ocean_core_benchmark_amd.verify_v05 [--nx 48] [--nz 20]
        [--steps 100] [--only v5_3,v5_5] [--json out.json] [--config DIR]
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

from . import _lib
from .cases3d_v05 import build_domain3d, simulate3d_v05
from .config import load_config
from .domain import CGrid, Domain, build_bathymetry
from .types import Physics3D, PhysicsV05, SchemeParams, physics_v05_from_config

TOL_REDUCE = 1e-12      # V5-1
TOL_EOS = 1e-4          # V5-4, against the published check value
TEOS10_CHECK = 1027.45140   # rho(CT=10, SA=30, z=1000 m), the polynomial's published value


def _rel(a: np.ndarray, b: np.ndarray, scale: float) -> float:
    """Relative L2 of a against b, normalised by the state scale where b is
    (near) identically zero."""
    den = float(np.linalg.norm(b))
    if den < 1e-11 * scale:
        den = scale
    return float(np.linalg.norm(a - b) / den)


def _host(x) -> np.ndarray:
    return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def v5_1_reduce(cfg, device: str = "cuda:0") -> dict:
    from .cases3d_v02 import build_case3d
    from .model3d_v02 import Stepper3D
    from .stepper import HipStepper
    nx, nz, dt, n_steps = 32, 12, 600.0, 20
    grid = CGrid(nx=nx, ny=nx, Lx=4.0e5, Ly=4.0e5, nz=nz, depth=1000.0)
    H, mask = build_bathymetry(grid, "flat", 1000.0)
    domain = Domain(grid=grid, H=H, mask=mask, vcoord="zlevel")
    out = {}
    for scheme in ("theta", "split_explicit"):
        for case_name, phys in (
                ("barotropic3d", Physics3D(g=9.80616, f=1e-4, H=1000.0)),
                ("baroclinic_igw", Physics3D(g=9.80616, f=1e-4, H=1000.0, N2=1e-4,
                                             nu=1e-3, kappa=1e-3))):
            params = SchemeParams(name=scheme, theta=0.5, theta_cor=0.5, n_picard=2,
                                  solver="pcg_jacobi", rtol=1e-14, max_iter=5000)
            case, _ = build_case3d(case_name, grid, phys, cfg)
            s0 = case.initial()
            old = Stepper3D(grid, phys, params, dt, device=device).integrate(s0, n_steps)
            new = HipStepper(domain, PhysicsV05(base=phys), params, dt,
                             device=device).integrate(s0, n_steps)
            scale = max(float(np.linalg.norm(_host(old.u))), 1e-30)
            out[f"{scheme}/{case_name}"] = max(
                _rel(_host(getattr(new, f)), _host(getattr(old, f)), scale)
                for f in ("u", "v", "b", "eta"))
    identity = {k: v for k, v in out.items() if not k.startswith("split_explicit/")}
    known = {k: v for k, v in out.items() if k.startswith("split_explicit/")}
    worst = max(identity.values())
    return {"name": "V5-1 reduce_v04", "tol": TOL_REDUCE, "worst": worst, "detail": out,
            "known_difference_split_explicit": known, "pass": worst < TOL_REDUCE}


def divgrad(Ku: np.ndarray, Kv: np.ndarray, eta: np.ndarray, dx: float, dy: float,
            device: str = "cuda:0") -> np.ndarray:
    """D[K G eta] on the device over a doubly periodic, fully wet grid."""
    import torch
    lib = _lib.load()
    dev = dict(dtype=torch.float64, device=device)
    ku, kv, x = (torch.as_tensor(np.ascontiguousarray(a), **dev) for a in (Ku, Kv, eta))
    mask = torch.ones(eta.shape, dtype=torch.uint8, device=device)
    out = torch.empty_like(x)
    ny, nx = eta.shape
    stream = C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(lib.hoc_helm_divgrad(nx, ny, dx, dy, _lib.ptr(ku), _lib.ptr(kv), _lib.ptr(mask),
                                    _lib.ptr(x), _lib.ptr(out), stream))
    return out.cpu().numpy()


def v5_2_operator_order(device: str = "cuda:0",
                        sizes: tuple[int, ...] = (64, 128, 256, 512, 1024)) -> dict:
    """H = H0 (1 + a sin(kx) sin(ky)), eta = cos(m k x) cos(n k y), both
    periodic, so div(H grad eta) is known in closed form. The second-order
    face rules (mean, harmonic) are gated at an observed order > 1.9; the
    production "min" rule is first order by construction and is reported."""
    Lx = 4.0e5
    H0, a, m, n = 1000.0, 0.3, 2.0, 1.0
    k = 2.0 * np.pi / Lx

    def errors(rule: str) -> list[float]:
        out = []
        for nx in sizes:
            grid = CGrid(nx=nx, ny=nx, Lx=Lx, Ly=Lx, nz=1, depth=H0)
            x, y = grid.coords("eta")
            H = H0 * (1.0 + a * np.sin(k * x) * np.sin(k * y))
            eta = np.cos(m * k * x) * np.cos(n * k * y)
            He, Hn = np.roll(H, -1, axis=-1), np.roll(H, -1, axis=-2)
            if rule == "min":
                Ku, Kv = np.minimum(H, He), np.minimum(H, Hn)
            elif rule == "mean":
                Ku, Kv = 0.5 * (H + He), 0.5 * (H + Hn)
            else:
                Ku, Kv = 2.0 * H * He / (H + He), 2.0 * H * Hn / (H + Hn)
            dHdx = H0 * a * k * np.cos(k * x) * np.sin(k * y)
            dHdy = H0 * a * k * np.sin(k * x) * np.cos(k * y)
            dedx = -m * k * np.sin(m * k * x) * np.cos(n * k * y)
            dedy = -n * k * np.cos(m * k * x) * np.sin(n * k * y)
            lap = -((m * k) ** 2 + (n * k) ** 2) * eta
            exact = dHdx * dedx + dHdy * dedy + H * lap
            num = divgrad(Ku, Kv, eta, grid.dx, grid.dy, device)
            out.append(float(np.sqrt(np.mean((num - exact) ** 2))
                             / np.sqrt(np.mean(exact ** 2))))
        return out

    rows, orders = [], {}
    for rule in ("min", "mean", "harmonic"):
        e = errors(rule)
        orders[rule] = float(np.log2(e[-2] / e[-1]))
        rows += [{"face_rule": rule, "nx": nx, "l2_rel": v} for nx, v in zip(sizes, e)]
    worst = min(orders["mean"], orders["harmonic"])
    return {"name": "V5-2 operator_order", "tol": 1.9, "worst": worst, "detail": rows,
            "orders": orders, "pass": worst > 1.9}


def v5_3_seamount(cfg, nx: int, nz: int, steps: int, device: str = "cuda:0") -> dict:
    """Flat rows must stay at machine zero; over topography the pressure-
    gradient correction must cut max|u| at least tenfold per coordinate."""
    base = cfg.with_overrides({"solver.kind": "pcg_jacobi", "physics3d.N2": 0.0,
                               "physics3d_v05.pgf": "keep"})
    rows = []
    for topo in ("flat", "seamount", "rough"):
        for vcoord in ("zlevel", "zstar", "sigma"):
            for corr in (True, False):
                c = base.with_overrides({"physics3d_v05.pgf_correction": corr})
                d = build_domain3d(c, nx, nz, vcoord=vcoord, bathymetry=topo)
                r = simulate3d_v05(c, "seamount_rest", d, physics_v05_from_config(c), 2.0,
                                   n_steps_override=steps, device=device)
                rows.append({"topo": topo, "vcoord": vcoord, "pgf_correction": corr,
                             "rx0": r["rx0"], "max_u_cm_s": r["metric"],
                             "rms_u_cm_s": r["extra"]["rms_speed_cm_s"]})
    flat = [r["max_u_cm_s"] for r in rows if r["topo"] == "flat"]
    worst_flat = max(flat)
    gain = {}
    for vc in sorted({r["vcoord"] for r in rows if r["topo"] != "flat"}):
        sel = [r for r in rows if r["topo"] != "flat" and r["vcoord"] == vc]
        on = max(r["max_u_cm_s"] for r in sel if r["pgf_correction"])
        off = max(r["max_u_cm_s"] for r in sel if not r["pgf_correction"])
        gain[vc] = off / on if on > 0 else float("inf")
    ok = worst_flat < 1e-8 and all(g >= 10.0 for g in gain.values())
    return {"name": "V5-3 seamount_rest", "tol": 1e-8, "worst": worst_flat, "detail": rows,
            "pgf_correction_gain": gain, "pass": bool(ok)}


def v5_4_eos(device: str = "cuda:0") -> dict:
    """The device EOS against the polynomial's published check value, and
    the linear and SEOS anomalies at the reference point (both zero)."""
    import torch

    from .stepper import eos_buoyancy
    g, rho0 = 9.80616, 1025.0
    dev = dict(dtype=torch.float64, device=device)

    def rho_prime(kind: str, T: float, S: float, z: float) -> float:
        b = eos_buoyancy(kind, torch.tensor([T], **dev), torch.tensor([S], **dev),
                         torch.tensor([z], **dev), g=g, rho0=rho0, alpha_T=2.0e-4,
                         beta_S=7.4e-4, T0=10.0, S0=35.0)
        return -float(b.item()) * rho0 / g

    check = rho0 + rho_prime("teos10", 10.0, 30.0, 1000.0)
    err = abs(check - TEOS10_CHECK) / TEOS10_CHECK
    return {"name": "V5-4 eos_check", "tol": TOL_EOS, "worst": err,
            "detail": {"teos10_check": check, "published": TEOS10_CHECK,
                       "linear_anomaly_at_ref": rho_prime("linear", 10.0, 35.0, 0.0),
                       "seos_anomaly_at_ref": rho_prime("seos", 10.0, 35.0, 0.0)},
            "pass": err < TOL_EOS}


def v5_5_flather(cfg, nx: int, nz: int, device: str = "cuda:0") -> dict:
    """Open boundaries must pass the wave (reflected fraction < 0.1 and below
    half of the closed-wall control)."""
    base = cfg.with_overrides({"solver.kind": "pcg_jacobi", "physics3d.N2": 0.0,
                               "physics3d.nu": 0.0, "physics3d.kappa": 0.0,
                               "physics.f0": 0.0, "scheme.name": "split_explicit"})
    rows = []
    for bc in ("closed", "open"):
        d = build_domain3d(base, nx, nz, vcoord="zlevel", bathymetry="flat", bc=bc)
        r = simulate3d_v05(base, "flather_radiate", d, physics_v05_from_config(base), 0.5,
                           device=device)
        rows.append({"bc": bc, "reflected_fraction": r["metric"], "n_steps": r["n_steps"]})
    frac = {r["bc"]: r["reflected_fraction"] for r in rows}
    return {"name": "V5-5 flather_radiate", "tol": 0.1, "worst": frac["open"], "detail": rows,
            "pass": frac["open"] < 0.1 and frac["open"] < 0.5 * frac["closed"]}


def v5_6_constancy(cfg, nx: int, nz: int, steps: int, device: str = "cuda:0") -> dict:
    """A uniform tracer over a sloping bottom under a standing surface wave
    must stay uniform to < 1e-9 on every coordinate."""
    base = cfg.with_overrides({"solver.kind": "pcg_jacobi", "physics3d.N2": 0.0,
                               "physics3d.nu": 0.0, "physics3d.kappa": 0.0,
                               "physics3d_v04.tracers": "TS",
                               "scheme.advection": "centered2"})
    rows = []
    for vcoord in ("zlevel", "zstar", "sigma"):
        d = build_domain3d(base, nx, nz, vcoord=vcoord, bathymetry="slope", slope=0.5)
        r = simulate3d_v05(base, "zstar_constancy", d, physics_v05_from_config(base), 2.0,
                           n_steps_override=steps, device=device)
        rows.append({"vcoord": vcoord, "spread": r["metric"], "T_mean": r["extra"]["T_mean"],
                     "tracer_integral_drift": r["extra"]["tracer_integral_drift"],
                     "tracer_integral_drift_live": r["extra"]["tracer_integral_drift_live"]})
    worst = max(r["spread"] for r in rows)
    return {"name": "V5-6 zstar_constancy", "tol": 1e-9, "worst": worst, "detail": rows,
            "pass": worst < 1e-9}


def suite_config(config_dir: str | Path | None = None):
    """The suite's configuration: the config tree on a 400 km square domain."""
    return load_config(config_dir).with_overrides({"grid.Lx": 4.0e5, "grid.Ly": 4.0e5})


def run(checks: list[str] | None = None, *, nx: int = 48, nz: int = 20, steps: int = 100,
        config_dir: str | Path | None = None, device: str = "cuda:0",
        echo: bool = False) -> list[dict]:
    cfg = suite_config(config_dir)
    table = {
        "v5_1": lambda: v5_1_reduce(cfg, device),
        "v5_2": lambda: v5_2_operator_order(device),
        "v5_3": lambda: v5_3_seamount(cfg, nx, nz, steps, device),
        "v5_4": lambda: v5_4_eos(device),
        "v5_5": lambda: v5_5_flather(cfg, nx, nz, device),
        "v5_6": lambda: v5_6_constancy(cfg, nx, nz, steps, device),
    }
    wanted = list(table) if not checks else checks
    unknown = [c for c in wanted if c not in table]
    if unknown:
        raise ValueError(f"unknown check(s): {', '.join(unknown)} (expected {', '.join(table)})")
    results = []
    for key in wanted:
        r = table[key]()
        results.append(r)
        if echo:
            rel = ">" if r["name"].startswith("V5-2") else "<"
            print(f"{'PASS' if r['pass'] else 'FAIL'}  {r['name']:24s} "
                  f"worst={r['worst']:.4e} tol {rel} {r['tol']:.1e}", flush=True)
    return results


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description="v0.5 verification suite on the GPU")
    ap.add_argument("--nx", type=int, default=48)
    ap.add_argument("--nz", type=int, default=20)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--only", default="", help="comma-separated check ids (v5_1 .. v5_6)")
    ap.add_argument("--json", type=Path, default=None)
    ap.add_argument("--config", default=None, help="config directory (default: ./config)")
    args = ap.parse_args(argv)
    wanted = [c for c in args.only.split(",") if c] or None
    try:
        results = run(wanted, nx=args.nx, nz=args.nz, steps=args.steps,
                      config_dir=args.config, echo=True)
    except ValueError as exc:
        print(f"FATAL: {exc}")
        return 1
    if args.json:
        args.json.parent.mkdir(parents=True, exist_ok=True)
        args.json.write_text(json.dumps(results, indent=2, default=float))
        print(f"json: {args.json}")
    return 0 if all(r["pass"] for r in results) else 1


if __name__ == "__main__":
    sys.exit(main())
