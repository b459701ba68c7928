"""The v0.5 bundle: the reference's process boundary for compiled backends.

A bundle is a Fortran namelist `<prefix>.nml` (groups grid_nml, physics_nml,
scheme_nml, case_nml, run_nml, v05_nml; tools/toml2nml.py:143-273), the
domain file (H then mask, fp64 [ny, nx]) and the initial-state file (eta,
u, v, b and, with tracers = 'TS', T and S; fp64 C order). A backend reads
them, integrates n_steps (n_warmup untimed repeats, then n_repeat timed
repeats from the same initial state), and writes `<prefix>_state3d5.bin` in
the order of the initial-state file plus `<prefix>_metrics.json`
(libs/cuda/src/cfd_exp3d5_cuda.cu:1104-1237). Invalid input ends the run
with a `FATAL:` line on stderr and exit status 1.

    python -m hydrostatic_ocean_core_benchmark_amd.bundle <prefix>.nml

`write_bundle` writes a bundle from a Problem (the writer side of the same
format), so a case built here can be run by any backend that reads it.
"""

from __future__ import annotations

import json
import sys
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .domain import CGrid, Domain
from .types import Physics3D, PhysicsV05, SchemeParams, State3D

BACKEND = "hip3d5"


class BundleError(Exception):
    """Invalid bundle: reported as `FATAL: <message>` with exit status 1."""


# ------------------------------------------------------------ namelist text
def read_namelist(path: str | Path) -> dict[str, str]:
    """Flat key -> raw value map of a namelist: '!' starts a comment, group
    lines ('&name') and terminators ('/') are skipped, later keys win
    (libs/cuda/src/config.hpp:75-93)."""
    try:
        text = Path(path).read_text(encoding="utf-8")
    except OSError as exc:
        raise BundleError(f"cannot open namelist {path}") from exc
    kv: dict[str, str] = {}
    for raw in text.splitlines():
        line = raw.split("!", 1)[0].strip()
        if not line or line[0] in "&/":
            continue
        key, eq, value = line.partition("=")
        if eq:
            kv[key.strip()] = value.strip()
    return kv


def fortran_double(text: str) -> float:
    """A Fortran real literal ('1.0d-12', '3.5D0', '2.0e3') as a float."""
    return float(text.replace("d", "e").replace("D", "e"))


def unquote(text: str) -> str:
    if len(text) >= 2 and text[0] in "'\"":
        return text[1:-1]
    return text


def format_double(value: float) -> str:
    """Full-precision Fortran double literal (round-trips bit for bit)."""
    return f"{float(value):.17e}".replace("e", "d")


def _flag(text: str) -> bool:
    return "t" in text.lower()


# ------------------------------------------------------------- the config
@dataclass
class Bundle:
    nml: Path
    domain: Domain
    physics: PhysicsV05
    params: SchemeParams
    case_name: str
    dt: float
    t_final: float
    n_steps: int
    n_repeat: int
    n_warmup: int
    out_prefix: str
    domain_file: str
    init_file: str

    @property
    def ts(self) -> bool:
        return self.physics.tracers == "TS"


def parse_bundle(path: str | Path) -> Bundle:
    """Namelist -> domain, physics, scheme and run settings, with the
    validation of the compiled backends (config.hpp:150-165,
    cfd_exp3d5_cuda.cu:89-115)."""
    kv = read_namelist(path)

    def num(key, default):
        return fortran_double(kv[key]) if key in kv else float(default)

    def inum(key, default):
        return int(num(key, default))

    def s(key, default):
        return unquote(kv[key]) if key in kv else default

    scheme = s("scheme_name", "theta")
    theta = num("theta", 0.5)
    if scheme not in ("fb", "theta", "split_explicit"):
        raise BundleError(f"unknown scheme {scheme} (expected fb|theta|split_explicit)")
    if scheme == "theta" and not (0.0 < theta <= 1.0):
        raise BundleError(f"theta must lie in (0,1]; got {theta:g}")
    n_picard, n_steps, dt = inum("n_picard", 2), inum("n_steps", 0), num("dt", 0.0)
    if n_picard < 1 or n_steps < 1 or dt <= 0.0:
        raise BundleError("dt, n_steps and n_picard must come from the driver")
    mxl = s("mxl", "integral")
    if mxl not in ("integral", "recursive"):
        raise BundleError("mxl must be integral|recursive")
    vcoord = s("vcoord", "zlevel")
    if vcoord != "zlevel":
        raise BundleError("HIP v0.5 backend: vcoord=zlevel only")
    bc_x, bc_y = s("bc_x", "periodic"), s("bc_y", "periodic")
    if "open" in (bc_x, bc_y):
        raise BundleError("HIP v0.5 backend: periodic|closed only")

    nx, ny, nz = inum("nx", 100), inum("ny", 100), inum("nz", 1)
    grid = CGrid(nx=nx, ny=ny, Lx=num("lx", 1.0e7), Ly=num("ly", 1.0e7), nz=nz,
                 depth=num("h0", 1000.0))
    domain_file, init_file = s("domain_file", ""), s("init_file", "")
    n2 = nx * ny
    H, mask = _read_exact(domain_file, 2 * n2, "domain").reshape(2, ny, nx)
    try:
        domain = Domain(grid=grid, H=H.copy(), mask=mask.copy(), bc_x=bc_x, bc_y=bc_y,
                        vcoord=vcoord, face_rule=s("face_rule", "min"),
                        min_partial=num("min_partial", 0.1))
    except ValueError as exc:
        raise BundleError(str(exc)) from exc

    base = Physics3D(g=num("g", 9.80616), f=num("f0", 1e-4), H=num("h0", 1000.0),
                     N2=num("n2", 0.0), nu=num("nu", 0.0), kappa=num("kappa", 0.0),
                     rho0=num("rho0", 1025.0), tau_x=num("tau_x", 0.0),
                     tau_y=num("tau_y", 0.0), bottom_drag=num("bottom_drag", 0.0),
                     tracers=s("tracers", "buoyancy"), alpha_T=num("alpha_t", 2e-4),
                     beta_S=num("beta_s", 7.4e-4), T0=num("t0", 10.0), S0=num("s0", 35.0),
                     cp=num("cp", 3990.0), A_h=num("a_h", 0.0), K_h=num("k_h", 0.0),
                     q_heat=num("q_heat", 0.0), q_salt=num("q_salt", 0.0))
    physics = PhysicsV05(base=base, eos=s("eos", "linear"), pgf=s("pgf", "remove"),
                         pgf_correction=_flag(s("pgf_correction", ".true.")),
                         closure=s("closure", "none"), c_k=num("c_k", 0.1),
                         c_eps=num("c_eps", 0.7), pr_t=num("pr_t", 1.0), z_0=num("z_0", 0.1),
                         e_min=num("e_min", 1e-6), l_min=num("l_min", 0.01),
                         e_bb=num("e_bb", 3.75), n2_min=num("n2_min", 1e-8), mxl=mxl)
    params = SchemeParams(name=scheme, theta=theta, theta_cor=num("theta_cor", 0.5),
                          n_picard=n_picard, theta_v=num("theta_v", 0.5),
                          advection=s("advection", "none"), n_split=inum("n_split", 0),
                          solver=s("solver_kind", "pcg_jacobi"), rtol=num("rtol", 1e-12),
                          max_iter=inum("max_iter", 2000), mg_levels=inum("mg_levels", 5),
                          mg_sweeps=inum("mg_sweeps", 2), mg_omega=num("mg_omega", 0.8),
                          mg_smoother=s("mg_smoother", "jacobi"))
    try:
        params.validate()
    except ValueError as exc:
        raise BundleError(str(exc)) from exc
    return Bundle(nml=Path(path), domain=domain, physics=physics, params=params,
                  case_name=s("case_name", "custom"), dt=dt, t_final=num("t_final", 0.0),
                  n_steps=n_steps, n_repeat=max(1, inum("n_repeat", 5)),
                  n_warmup=max(0, inum("n_warmup", 1)),
                  out_prefix=s("out_prefix", "output/hip_run"), domain_file=domain_file,
                  init_file=init_file)


def _read_exact(path: str, count: int, what: str) -> np.ndarray:
    try:
        raw = np.fromfile(path, dtype="<f8")
    except OSError as exc:
        raise BundleError(f"cannot read {path}") from exc
    if raw.size < count:
        raise BundleError(f"{what} file {path}: expected {count} doubles, got {raw.size}")
    return raw[:count]


def read_state_file(path: str | Path, nx: int, ny: int, nz: int, ts: bool) -> dict:
    """eta [ny, nx], then u, v, b (and T, S) [nz, ny, nx], fp64 C order
    (the initial-state and final-state files share this layout)."""
    names = ["u", "v", "b"] + (["T", "S"] if ts else [])
    n2, n3 = nx * ny, nx * ny * nz
    raw = np.fromfile(path, dtype="<f8")
    if raw.size != n2 + len(names) * n3:
        raise BundleError(f"{path}: expected {n2 + len(names) * n3} doubles, got {raw.size}")
    out = {"eta": raw[:n2].reshape(ny, nx)}
    off = n2
    for name in names:
        out[name] = raw[off:off + n3].reshape(nz, ny, nx)
        off += n3
    return out


def write_state_file(path: str | Path, fields: dict, ts: bool) -> None:
    names = ["eta", "u", "v", "b"] + (["T", "S"] if ts else [])
    with open(path, "wb") as fh:
        for name in names:
            fh.write(np.ascontiguousarray(fields[name], dtype="<f8").tobytes())


# ----------------------------------------------------------------- writer
def write_bundle(prefix: str | Path, domain: Domain, physics: PhysicsV05, params: SchemeParams,
                 state0: State3D, dt: float, n_steps: int, *, case_name: str = "custom",
                 n_repeat: int = 1, n_warmup: int = 0, out_prefix: str | None = None) -> Path:
    """Write `<prefix>.nml`, `<prefix>_domain.bin`, `<prefix>_init.bin` for
    a uniform-forcing problem (the namelist carries scalar tau and kappa)."""
    prefix = str(prefix)
    Path(prefix).parent.mkdir(parents=True, exist_ok=True)
    g, p = domain.grid, physics
    for name in ("tau_x", "tau_y", "kappa"):
        if np.ndim(getattr(p, name)) != 0:
            raise ValueError(f"the namelist holds a scalar {name}; got an array")
    domain_file, init_file = f"{prefix}_domain.bin", f"{prefix}_init.bin"
    with open(domain_file, "wb") as fh:
        fh.write(np.ascontiguousarray(domain.H, dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(domain.mask, dtype="<f8").tobytes())
    ts = p.tracers == "TS"
    write_state_file(init_file, {"eta": state0.eta, "u": state0.u, "v": state0.v,
                                 "b": state0.b, "T": state0.T, "S": state0.S}, ts)
    fd = format_double
    lines = [
        "! v0.5 bundle written by hydrostatic_ocean_core_benchmark_amd.bundle",
        "&grid_nml", f"  nx = {g.nx}", f"  ny = {g.ny}", f"  nz = {g.nz}",
        f"  lx = {fd(g.Lx)}", f"  ly = {fd(g.Ly)}", "/",
        "&physics_nml", f"  g  = {fd(p.g)}", f"  f0 = {fd(p.f)}", f"  h0 = {fd(p.H)}",
        f"  n2 = {fd(p.N2)}", f"  nu = {fd(p.nu)}", f"  kappa = {fd(p.kappa)}",
        f"  rho0 = {fd(p.rho0)}", f"  tau_x = {fd(p.tau_x)}", f"  tau_y = {fd(p.tau_y)}",
        f"  bottom_drag = {fd(p.bottom_drag)}", "/",
        "&scheme_nml", f"  scheme_name = '{params.name}'", f"  theta = {fd(params.theta)}",
        f"  theta_cor = {fd(params.theta_cor)}", f"  theta_v = {fd(params.theta_v)}",
        f"  n_picard = {params.n_picard}", f"  n_split = {params.n_split}",
        f"  solver_kind = '{params.solver}'", f"  rtol = {fd(params.rtol)}",
        f"  max_iter = {params.max_iter}", f"  mg_levels = {params.mg_levels}",
        f"  mg_sweeps = {params.mg_sweeps}", f"  mg_omega = {fd(params.mg_omega)}",
        f"  mg_smoother = '{params.mg_smoother}'", "/",
        "&case_nml", f"  case_name = '{case_name}'", "/",
        "&run_nml", f"  dt = {fd(dt)}", f"  t_final = {fd(dt * n_steps)}",
        f"  n_steps = {n_steps}", f"  n_repeat = {n_repeat}", f"  n_warmup = {n_warmup}",
        f"  out_prefix = '{out_prefix or prefix}'", "/",
        "&v05_nml", f"  vcoord = '{domain.vcoord}'", f"  bc_x = '{domain.bc_x}'",
        f"  bc_y = '{domain.bc_y}'", f"  face_rule = '{domain.face_rule}'",
        f"  min_partial = {fd(domain.min_partial)}", f"  eos = '{p.eos}'",
        f"  pgf = '{p.pgf}'", f"  pgf_correction = {'.true.' if p.pgf_correction else '.false.'}",
        f"  advection = '{params.advection}'", f"  tracers = '{p.tracers}'",
        f"  alpha_t = {fd(p.alpha_T)}", f"  beta_s = {fd(p.beta_S)}", f"  t0 = {fd(p.T0)}",
        f"  s0 = {fd(p.S0)}", f"  cp = {fd(p.cp)}", f"  a_h = {fd(p.A_h)}",
        f"  k_h = {fd(p.K_h)}", f"  q_heat = {fd(p.q_heat)}", f"  q_salt = {fd(p.q_salt)}",
        f"  closure = '{p.closure}'", f"  c_k = {fd(p.c_k)}", f"  c_eps = {fd(p.c_eps)}",
        f"  pr_t = {fd(p.pr_t)}", f"  z_0 = {fd(p.z_0)}", f"  e_min = {fd(p.e_min)}",
        f"  l_min = {fd(p.l_min)}", f"  e_bb = {fd(p.e_bb)}", f"  n2_min = {fd(p.n2_min)}",
        f"  mxl = '{p.mxl}'", f"  domain_file = '{domain_file}'",
        f"  init_file = '{init_file}'", "/", ""]
    nml = Path(f"{prefix}.nml")
    nml.write_text("\n".join(lines), encoding="utf-8")
    return nml


# ----------------------------------------------------------------- runner
def _median(x: list[float]) -> float:
    y = sorted(x)
    n = len(y)
    return y[n // 2] if n % 2 else 0.5 * (y[n // 2 - 1] + y[n // 2])


def run_bundle(path: str | Path, device: str = "cuda:0", log=print) -> dict:
    """Run a bundle on the GPU; returns the metrics written to
    `<out_prefix>_metrics.json`."""
    import torch

    from .stepper import HipStepper

    b = parse_bundle(path)
    g = b.domain.grid
    init = read_state_file(b.init_file, g.nx, g.ny, g.nz, b.ts)
    free0 = torch.cuda.mem_get_info(device)[0]
    st = HipStepper(b.domain, b.physics, b.params, b.dt, device)
    dev_bytes = float(free0 - torch.cuda.mem_get_info(device)[0])
    s0 = State3D(u=init["u"], v=init["v"], b=init["b"], eta=init["eta"], t=0.0,
                 T=init.get("T"), S=init.get("S"))
    fin0 = st._inputs(s0)
    bufs = [st._new_state_buffers(), st._new_state_buffers()]
    gpu = torch.cuda.get_device_name(st.device)
    log(f"{BACKEND}: case={b.case_name} nx={g.nx} nz={g.nz} scheme={b.params.name} "
        f"solver={b.params.solver} eos={b.physics.eos} dt={b.dt:12.5E} steps={b.n_steps} "
        f"gpu={gpu}")

    def diverged(eta) -> bool:
        m = float(torch.max(torch.abs(eta)))
        return not (m < 1.0e6)

    def one_run():
        st.reset_mixing()
        cur = fin0
        for n in range(b.n_steps):
            out = bufs[n % 2]
            st._step_into(cur, out)
            cur = out
            if n % 10 == 9 and diverged(cur["eta"]):
                return cur, True
        return cur, False

    run_diverged = False
    for _ in range(b.n_warmup):
        _, run_diverged = one_run()
        torch.cuda.synchronize(st.device)
        if run_diverged:
            break
    c0 = st.counters()
    samples: list[float] = []
    cur = fin0
    if not run_diverged:
        for _ in range(b.n_repeat):
            torch.cuda.synchronize(st.device)
            t0 = time.perf_counter()
            cur, run_diverged = one_run()
            torch.cuda.synchronize(st.device)
            samples.append(time.perf_counter() - t0)
            if run_diverged:
                log("  DIVERGED (|eta| > 1e6 or NaN) - run stopped early")
                break
    c1 = st.counters()
    n_done = max(1, len(samples))
    per_run = {k: (c1[k] - c0[k]) // n_done for k in c1}
    t_med = _median(samples) if samples else float("nan")
    t_mad = _median([abs(x - t_med) for x in samples]) if samples else float("nan")
    log(f"  wall median={t_med:12.5E} s  mad={t_mad:12.5E} s  "
        f"solver_iters/run={per_run['solver_iterations']}  "
        f"substeps/run={per_run['barotropic_substeps']}")

    host = {k: v.detach().cpu().numpy() for k, v in cur.items() if isinstance(v, torch.Tensor)}
    write_state_file(f"{b.out_prefix}_state3d5.bin", host, b.ts)

    def jnum(x):
        return float(x) if np.isfinite(x) else None

    cells = g.nx * g.ny * g.nz
    metrics = {
        "backend": BACKEND, "gpu": gpu, "nx": g.nx, "ny": g.ny, "nz": g.nz, "cells": cells,
        "n_steps": b.n_steps, "scheme": b.params.name, "solver": b.params.solver,
        "eos": b.physics.eos, "advection": b.params.advection, "tracers": b.physics.tracers,
        "theta": jnum(b.params.theta), "dt": jnum(b.dt),
        "solver_iterations": per_run["solver_iterations"],
        "solver_failures": c1["solver_failures"] - c0["solver_failures"],
        "tridiagonal_solves": per_run["tridiagonal_solves"],
        "solver_rebuilds": per_run["solver_rebuilds"], "n_split": st.n_split,
        "barotropic_substeps": per_run["barotropic_substeps"], "device_bytes": dev_bytes,
        "bytes_per_cell": jnum(dev_bytes / cells), "diverged": run_diverged,
        "wall_s": jnum(t_med), "wall_mad_s": jnum(t_mad), "n_repeat": len(samples),
        "omp_num_threads": 0,
    }
    Path(f"{b.out_prefix}_metrics.json").write_text(json.dumps(metrics, indent=2) + "\n",
                                                    encoding="utf-8")
    return metrics


def main(argv: list[str] | None = None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m hydrostatic_ocean_core_benchmark_amd.bundle <namelist>",
              file=sys.stderr)
        return 1
    try:
        run_bundle(argv[0])
    except (BundleError, ValueError, NotImplementedError) as exc:
        print(f"FATAL: {exc}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
