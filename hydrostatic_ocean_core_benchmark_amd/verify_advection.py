"""Monotonicity, transport and order of the third-order (kappa = 1/3) tracer
reconstruction with the time integrator taken out of the question
(tools/verify_advection.py): constant-velocity periodic advection of a row
with SSP-RK3 in time, whose stages are convex combinations of forward-Euler
steps, so any overshoot that remains is the spatial reconstruction's.

The tendency is the device's tracer advection (the production kernel) on a
fully wet, doubly periodic nx x 4 x 2 grid with dx = 1 / nx, dz = 1,
u = +-1 and v = w = 0, every row and layer holding the same profile: there
its flux divergence is exactly -(u f_e - u f_w) / dx of the row. The RK3
stages run as fp64 tensor arithmetic on the device, in the reference's
operation order.

    python -m hydrostatic_ocean_core_benchmark_amd.verify_advection [--nx 128]
        [--courant 0.4] [--json out.json]

Reported (the keys of the reference's record): transport (half a crossing of
a sine must negate it), rk3 (a top-hat over 1, 2, 4, 8 crossings: the limited
scheme makes no new extremum, the unlimited one's overshoot saturates),
order (on a sine at nx = 32 .. 256), mirror (reflecting space and velocity
commutes with the scheme), quarter (a quarter crossing against the exact
solution, both directions) and saturation.
"""

from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

import numpy as np

NY, NZ = 4, 2


class RowAdvection:
    """-(flux divergence) of a periodic row on the device, for one scheme."""

    def __init__(self, nx: int, scheme: str, device: str = "cuda:0"):
        import torch

        from .domain import CGrid, Domain
        from .stepper import HipStepper
        from .types import Physics3D, PhysicsV05, SchemeParams
        grid = CGrid(nx=nx, ny=NY, Lx=1.0, Ly=NY / nx, nz=NZ, depth=float(NZ))
        dom = Domain(grid=grid, H=np.full(grid.shape, float(NZ)), mask=np.ones(grid.shape))
        phys = PhysicsV05(base=Physics3D(g=9.80616, f=0.0, H=float(NZ)))
        params = SchemeParams(name="fb", advection=scheme)
        self.st = HipStepper(dom, phys, params, 1.0, device=device)
        self.nx, self.torch = nx, torch
        dev = dict(dtype=torch.float64, device=self.st.device)
        self.v = torch.zeros((NZ, NY, nx), **dev)
        self.w = torch.zeros((NZ + 1, NY, nx), **dev)
        self.u = {s: torch.full((NZ, NY, nx), s, **dev) for s in (1.0, -1.0)}
        # a device-resident divisor: torch divides by a host scalar through
        # its reciprocal, which is not the correctly rounded quotient
        self.three = torch.tensor(3.0, **dev)

    def tend(self, c, sign: float):
        return -self.st.tracer_advection(c, self.u[sign], self.v, self.w)

    def rk3(self, c0: np.ndarray, courant: float, crossings: float,
            sign: float = 1.0) -> tuple[np.ndarray, int]:
        """SSP-RK3 over `crossings` domain lengths at speed 1 (dt = courant dx,
        rounded to land on the distance exactly)."""
        torch = self.torch
        dx = 1.0 / self.nx
        dt = courant * dx
        n_steps = max(1, int(round(crossings / dt)))
        dt = crossings / n_steps
        c = torch.as_tensor(np.broadcast_to(c0, (NZ, NY, self.nx)).copy(), dtype=torch.float64,
                            device=self.st.device)
        for _ in range(n_steps):
            c1 = c + dt * self.tend(c, sign)
            c2 = 0.75 * c + 0.25 * (c1 + dt * self.tend(c1, sign))
            c = (c + 2.0 * (c2 + dt * self.tend(c2, sign))) / self.three
        return c[0, 0].cpu().numpy(), n_steps


def run(nx: int = 128, courant: float = 0.4, device: str = "cuda:0") -> dict:
    x = (np.arange(nx) + 0.5) / nx
    top = np.where((x > 0.3) & (x < 0.6), 1.0, 0.0)
    rows = {name: RowAdvection(nx, name, device) for name in ("up3", "up3_tvd")}
    res: dict = {"nx": nx, "courant": courant}

    sine = np.sin(2.0 * np.pi * x)
    half, n = rows["up3_tvd"].rk3(sine, courant, 0.5)
    res["transport"] = {"steps": n, "amplitude": float(np.max(np.abs(half))),
                        "rel_error_vs_exact": float(np.max(np.abs(half + sine)))
                        / float(np.max(np.abs(sine)))}
    ok_transport = (res["transport"]["rel_error_vs_exact"] < 0.05
                    and res["transport"]["amplitude"] > 0.9)

    res["rk3"] = {}
    for name, row in rows.items():
        out = []
        for cr in (1.0, 2.0, 4.0, 8.0):
            c, n = row.rk3(top, courant, cr)
            out.append({"crossings": cr, "steps": n, "max": float(c.max()),
                        "min": float(c.min())})
        res["rk3"][name] = out

    res["order"] = {}
    for name in ("up3", "up3_tvd"):
        errs, sizes = [], (32, 64, 128, 256)
        for m in sizes:
            xx = (np.arange(m) + 0.5) / m
            s0 = np.sin(2.0 * np.pi * xx)
            c, _ = RowAdvection(m, name, device).rk3(s0, courant, 1.0)
            errs.append(float(np.sqrt(np.mean((c - s0) ** 2))))
        res["order"][name] = {"nx": list(sizes), "l2": errs,
                              "orders": [float(np.log2(errs[i] / errs[i + 1]))
                                         for i in range(len(errs) - 1)]}
    ok_order = res["order"]["up3"]["orders"][-1] > 2.6

    plus, _ = rows["up3"].rk3(top, courant, 0.25, +1.0)
    minus, _ = rows["up3"].rk3(top[::-1].copy(), courant, 0.25, -1.0)
    res["mirror"] = {"max_abs_difference": float(np.max(np.abs(plus - minus[::-1])))}
    ok_mirror = res["mirror"]["max_abs_difference"] < 1e-12

    nq = 256
    xq = (np.arange(nq) + 0.5) / nq
    s0 = np.sin(2.0 * np.pi * xq)
    quarter = RowAdvection(nq, "up3", device)
    res["quarter"] = {}
    ok_quarter = True
    for sgn, tag in ((1.0, "u>0"), (-1.0, "u<0")):
        got, _ = quarter.rk3(s0, courant, 0.25, sgn)
        exact = np.sin(2.0 * np.pi * (xq - sgn * 0.25))
        norm = np.sqrt(np.mean(exact ** 2))
        rel = float(np.sqrt(np.mean((got - exact) ** 2)) / norm)
        frozen = float(np.sqrt(np.mean((s0 - exact) ** 2)) / norm)
        res["quarter"][tag] = {"rel_l2": rel, "rel_l2_if_frozen": frozen}
        ok_quarter = ok_quarter and rel < 0.02 and rel < 0.1 * frozen

    over = [r["max"] - 1.0 for r in res["rk3"]["up3"]]
    under = [-r["min"] for r in res["rk3"]["up3"]]
    tvd_mono = all(r["max"] <= 1.0 + 1e-12 and r["min"] >= -1e-12
                   for r in res["rk3"]["up3_tvd"])
    present = min(over) > 1e-3 and min(under) > 1e-3
    no_growth = over[-1] <= 1.05 * max(over) and under[-1] <= 1.05 * max(under)
    res["saturation"] = {"over": over, "under": under, "present": bool(present),
                         "no_growth": bool(no_growth)}
    res["passed"] = bool(ok_transport and ok_order and ok_mirror and ok_quarter and tvd_mono
                         and present and no_growth)
    return res


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description="tracer reconstruction under SSP-RK3, on the GPU")
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--courant", type=float, default=0.4)
    ap.add_argument("--json", type=Path, default=None)
    args = ap.parse_args(argv)
    res = run(args.nx, args.courant)
    t = res["transport"]
    print(f"transport: steps={t['steps']} amplitude={t['amplitude']:.4f} "
          f"rel. error vs exact half crossing {t['rel_error_vs_exact']:.3e}")
    for name, out in res["rk3"].items():
        for r in out:
            print(f"rk3 {name:8s} crossings={r['crossings']:4.1f} steps={r['steps']:5d} "
                  f"max={r['max']:+.4f} min={r['min']:+.4f}")
    for name, o in res["order"].items():
        print(f"order {name:8s} L2 {['%.2e' % e for e in o['l2']]} "
              f"orders {['%.2f' % v for v in o['orders']]}")
    print(f"mirror |difference| {res['mirror']['max_abs_difference']:.2e}")
    for tag, q in res["quarter"].items():
        print(f"quarter {tag} rel L2 {q['rel_l2']:.3e} (frozen {q['rel_l2_if_frozen']:.3e})")
    print(f"ADVECTION {'PASS' if res['passed'] else 'FAIL'}", flush=True)
    if args.json:
        args.json.parent.mkdir(parents=True, exist_ok=True)
        args.json.write_text(json.dumps(res, indent=1))
        print(f"json: {args.json}")
    return 0 if res["passed"] else 1


if __name__ == "__main__":
    sys.exit(main())
