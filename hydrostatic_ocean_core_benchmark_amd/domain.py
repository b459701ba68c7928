"""Static problem geometry: the Arakawa C-grid, bathymetry generators and the
z-level domain (masks, partial bottom cells, face thickness, centre depth).

This is set-up code that runs once on the host; the step itself never
recomputes geometry for a z-level coordinate. Behaviour matches:
  CGrid                 libs/core/grid.py:17-83 (array layout [j, i])
  Domain                libs/core/domain.py:24-195
  build_bathymetry      libs/core/bathymetry.py:47-124
The arrays it produces are checked against fixtures written by the reference
implementation itself (tests/golden).

Additions with no reference counterpart:
  * ``bathymetry kind="grf"``: a seeded Gaussian random field with a Gaussian
    correlation function of length ``corr_length``, scaled to a mean and
    standard deviation and clipped to [h_min, h_max].
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

VCOORDS = ("zlevel", "zstar", "sigma")
BCS = ("periodic", "closed", "open")


@dataclass(frozen=True)
class CGrid:
    nx: int
    ny: int
    Lx: float
    Ly: float
    bc_x: str = "periodic"
    bc_y: str = "periodic"
    nz: int = 1
    depth: float = 0.0     # nominal resting depth; dz = depth / nz

    @property
    def dx(self) -> float:
        return self.Lx / self.nx

    @property
    def dy(self) -> float:
        return self.Ly / self.ny

    @property
    def dz(self) -> float:
        return self.depth / self.nz

    @property
    def shape(self) -> tuple[int, int]:
        return (self.ny, self.nx)

    @property
    def shape3d(self) -> tuple[int, int, int]:
        return (self.nz, self.ny, self.nx)

    @property
    def cell_area(self) -> float:
        return self.dx * self.dy

    def coords(self, variable: str) -> tuple[np.ndarray, np.ndarray]:
        """(x, y) arrays [ny, nx] at eta (centre), u (east face) or v (north face)."""
        offset = {"eta": (0.5, 0.5), "u": (1.0, 0.5), "v": (0.5, 1.0)}
        if variable not in offset:
            raise ValueError(f"unknown variable '{variable}' (expected eta|u|v)")
        ox, oy = offset[variable]
        x = (np.arange(self.nx, dtype=np.float64) + ox) * self.dx
        y = (np.arange(self.ny, dtype=np.float64) + oy) * self.dy
        return np.meshgrid(x, y, indexing="xy")

    def z_centre(self) -> np.ndarray:
        """Height above the flat bottom of the layer centres (libs/core/grid.py:63-66)."""
        k = np.arange(self.nz, dtype=np.float64)
        return (self.nz - k - 0.5) * self.dz

    def z_interface(self) -> np.ndarray:
        """Height above the flat bottom of the interfaces, k = 0 the surface
        (libs/core/grid.py:68-71)."""
        k = np.arange(self.nz + 1, dtype=np.float64)
        return (self.nz - k) * self.dz


# ----------------------------------------------------------------- bathymetry
def _powerlaw_field(nx: int, ny: int, slope: float, seed: int, kmax: int = 0) -> np.ndarray:
    """Zero-mean periodic field with amplitude spectrum |k|^-slope and uniform
    random phases (the reference 'rough' generator, bathymetry.py:47-82)."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    KX, KY = np.meshgrid(kx, ky, indexing="xy")
    kk = np.hypot(KX, KY)
    amp = np.zeros_like(kk)
    pos = kk > 0
    amp[pos] = kk[pos] ** (-slope)
    if kmax > 0:
        amp[kk > kmax] = 0.0
        m = int(kmax)
        lattice = rng.uniform(0.0, 2.0 * np.pi, size=(2 * m + 1, 2 * m + 1))
        ix = np.round(KX).astype(int)
        iy = np.round(KY).astype(int)
        sel = (np.abs(ix) <= m) & (np.abs(iy) <= m)
        phase = np.zeros_like(kk)
        phase[sel] = lattice[iy[sel] + m, ix[sel] + m]
    else:
        phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
    f = np.fft.ifft2(amp * np.exp(1j * phase)).real
    return f - f.mean()


def gaussian_random_field(nx: int, ny: int, dx: float, dy: float, corr_length: float,
                          seed: int) -> np.ndarray:
    """Zero-mean, unit-variance periodic Gaussian random field whose
    covariance is exp(-r^2 / corr_length^2): white noise filtered by the
    square root of that covariance's spectrum, exp(-k^2 corr_length^2 / 8)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ny, nx))
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    KX, KY = np.meshgrid(kx, ky, indexing="xy")
    filt = np.exp(-(KX ** 2 + KY ** 2) * corr_length ** 2 / 8.0)
    f = np.fft.ifft2(np.fft.fft2(noise) * filt).real
    f -= f.mean()
    sd = f.std()
    return f / sd if sd > 0 else f


def build_bathymetry(grid: CGrid, kind: str, H0: float, *, h_rel: float = 0.9,
                     length: float = 0.0, slope: float = 0.5, spectrum_slope: float = 1.5,
                     spectrum_kmax: int = 0, r_target: float = 0.1, seed: int = 20260911,
                     island: bool = False, corr_length: float = 2.0e5,
                     h_min: float = 500.0, h_max: float = 4000.0, h_std: float = 0.0,
                     ) -> tuple[np.ndarray, np.ndarray]:
    """(H[ny, nx] positive down, mask[ny, nx]) for the named topography."""
    x, y = grid.coords("eta")
    Lx, Ly = grid.Lx, grid.Ly
    L = length if length > 0.0 else 0.1 * Lx
    if kind == "flat":
        H = np.full(grid.shape, H0)
    elif kind == "slope":
        H = H0 * (1.0 - slope * x / Lx)
    elif kind == "seamount":
        r2 = (x - 0.5 * Lx) ** 2 + (y - 0.5 * Ly) ** 2
        H = H0 * (1.0 - h_rel * np.exp(-r2 / L ** 2))
    elif kind == "ridge":
        H = H0 * (1.0 - h_rel * np.exp(-((x - 0.5 * Lx) / L) ** 2))
    elif kind == "rough":
        f = _powerlaw_field(grid.nx, grid.ny, spectrum_slope, seed, spectrum_kmax)
        f = f / np.std(f)
        H = np.maximum(H0 * (1.0 + r_target * f), 0.1 * H0)
    elif kind == "grf":
        f = gaussian_random_field(grid.nx, grid.ny, grid.dx, grid.dy, corr_length, seed)
        mean = 0.5 * (h_min + h_max)
        sd = h_std if h_std > 0.0 else 0.25 * (h_max - h_min)
        H = np.clip(mean + sd * f, h_min, h_max)
    else:
        raise ValueError(f"unknown bathymetry '{kind}' (expected flat|slope|seamount|"
                         "ridge|rough|grf)")
    mask = np.ones(grid.shape, dtype=np.float64)
    if island:
        r = (np.abs(x - 0.5 * Lx) < 0.1 * Lx) & (np.abs(y - 0.5 * Ly) < 0.1 * Ly)
        mask[r] = 0.0
    return H, mask


def bathymetry_from_config(grid: CGrid, config) -> tuple[np.ndarray, np.ndarray]:
    g = lambda k, d: config.get(f"bathymetry.{k}", d)   # noqa: E731
    return build_bathymetry(
        grid, str(g("kind", "flat")), float(config.get("physics.H")),
        h_rel=float(g("h_rel", 0.9)), length=float(g("length", 0.0)),
        slope=float(g("slope", 0.5)), spectrum_slope=float(g("spectrum_slope", 1.5)),
        spectrum_kmax=int(g("spectrum_kmax", 0)), r_target=float(g("r_target", 0.1)),
        seed=int(g("seed", 20260911)), island=bool(g("island", False)),
        corr_length=float(g("corr_length", 2.0e5)), h_min=float(g("h_min", 500.0)),
        h_max=float(g("h_max", 4000.0)), h_std=float(g("h_std", 0.0)))


# --------------------------------------------------------------------- domain
def _shift(a: np.ndarray, by: int, axis: int) -> np.ndarray:
    """Periodic shift: _shift(a, -1, -1)[j, i] == a[j, i+1]."""
    return np.roll(a, by, axis=axis)


@dataclass
class Domain:
    grid: CGrid
    H: np.ndarray
    mask: np.ndarray
    vcoord: str = "zlevel"
    bc_x: str = "periodic"
    bc_y: str = "periodic"
    min_partial: float = 0.1
    face_rule: str = "min"

    mask_u: np.ndarray = field(init=False)
    mask_v: np.ndarray = field(init=False)
    Hu: np.ndarray = field(init=False)
    Hv: np.ndarray = field(init=False)
    dzr: np.ndarray = field(init=False)
    dz3_ref: np.ndarray = field(init=False)
    mask3: np.ndarray = field(init=False)
    mask3u: np.ndarray = field(init=False)
    mask3v: np.ndarray = field(init=False)

    def __post_init__(self) -> None:
        if self.vcoord not in VCOORDS:
            raise ValueError(f"vcoord must be one of {VCOORDS}")
        for name, bc in (("bc_x", self.bc_x), ("bc_y", self.bc_y)):
            if bc not in BCS:
                raise ValueError(f"{name} must be one of {BCS}")
        if np.any(self.H[self.mask > 0] <= 0.0):
            raise ValueError("bathymetry must be positive in every wet cell")
        g = self.grid
        mu = self.mask * _shift(self.mask, -1, -1)
        mv = self.mask * _shift(self.mask, -1, -2)
        if self.bc_x != "periodic":
            mu[:, -1] = 0.0
        if self.bc_y != "periodic":
            mv[-1, :] = 0.0
        self.mask_u, self.mask_v = mu, mv
        self.Hu = mu * self._face(self.H, -1)
        self.Hv = mv * self._face(self.H, -2)
        h_max = float(np.max(self.H[self.mask > 0]))
        self.dzr = np.full(g.nz, h_max / g.nz)
        self.dz3_ref, self.mask3 = self._resting_thickness()
        self.mask3u = self.mask3 * _shift(self.mask3, -1, -1) * self.mask_u[None]
        self.mask3v = self.mask3 * _shift(self.mask3, -1, -2) * self.mask_v[None]

    def _face(self, a: np.ndarray, axis: int) -> np.ndarray:
        b = _shift(a, -1, axis)
        if self.face_rule == "min":
            return np.minimum(a, b)
        if self.face_rule == "mean":
            return 0.5 * (a + b)
        if self.face_rule == "harmonic":
            s = a + b
            return np.where(s > 0.0, 2.0 * a * b / np.where(s > 0.0, s, 1.0), 0.0)
        raise ValueError(f"face_rule must be min|mean|harmonic, got '{self.face_rule}'")

    def _resting_thickness(self) -> tuple[np.ndarray, np.ndarray]:
        nz, ny, nx = self.grid.shape3d
        if self.vcoord == "sigma":
            frac = self.dzr / self.dzr.sum()
            dz3 = frac[:, None, None] * self.H[None]
            m3 = np.broadcast_to(self.mask, (nz, ny, nx)).astype(np.float64)
            return dz3 * m3, m3.copy()
        # z-level: fixed levels filled from the top; the deepest wet cell is
        # partial and is dropped when thinner than min_partial * dzr.
        top = np.concatenate([[0.0], np.cumsum(self.dzr)])[:-1]
        left = self.H[None] - top[:, None, None]
        dz3 = np.clip(left, 0.0, self.dzr[:, None, None])
        thin = (dz3 > 0.0) & (dz3 < self.min_partial * self.dzr[:, None, None])
        dz3 = np.where(thin, 0.0, dz3)
        m3 = (dz3 > 0.0).astype(np.float64) * self.mask[None]
        return dz3 * m3, m3

    def layer_thickness(self, eta: np.ndarray) -> np.ndarray:
        if self.vcoord == "zlevel":
            return self.dz3_ref
        scale = np.where(self.mask > 0, 1.0 + eta / np.where(self.H > 0, self.H, 1.0), 1.0)
        return self.dz3_ref * scale[None]

    def face_thickness(self, dz3: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        return self.mask3u * self._face(dz3, -1), self.mask3v * self._face(dz3, -2)

    def depth_at_centres(self, eta: np.ndarray | None = None) -> np.ndarray:
        dz3 = self.dz3_ref if eta is None else self.layer_thickness(eta)
        above = np.concatenate([np.zeros((1,) + dz3.shape[1:]),
                                np.cumsum(dz3[:-1], axis=0)], axis=0)
        return above + 0.5 * dz3

    @property
    def is_flat_periodic(self) -> bool:
        return (self.vcoord == "zlevel" and self.bc_x == "periodic"
                and self.bc_y == "periodic" and bool(np.all(self.mask == 1.0))
                and float(np.ptp(self.H)) == 0.0)

    @property
    def roughness(self) -> Roughness:
        """libs/core/domain.py:157-170: r_std and H_max/H_min over wet cells;
        rx0 = max |H_a - H_b| / (H_a + H_b) over wet faces only, so the jump
        across a closed boundary's wrap-around is not counted."""
        h = self.H[self.mask > 0]
        rx = 0.0
        for axis, face_mask in ((-1, self.mask_u), (-2, self.mask_v)):
            sel = face_mask > 0
            if np.any(sel):
                a, b = self.H, _shift(self.H, -1, axis)
                rx = max(rx, float(np.max(np.abs(a - b)[sel] / (a + b)[sel])))
        return Roughness(r_std=float(np.std(h) / np.mean(h)), rx0=rx,
                         h_ratio=float(np.max(h) / np.min(h)))

    def depth_error(self) -> float:
        """Largest |sum_k dz3_ref - H| over wet columns (libs/core/domain.py:
        172-178): zero for sigma, the partial-cell truncation for z-level."""
        wet = self.mask > 0
        if not np.any(wet):
            return 0.0
        return float(np.max(np.abs(self.dz3_ref.sum(axis=0) - self.H)[wet]))

    def summary(self) -> dict:
        """libs/core/domain.py:187-196"""
        wet = self.mask > 0
        return {"vcoord": self.vcoord, "bc_x": self.bc_x, "bc_y": self.bc_y,
                "face_rule": self.face_rule, "H_min": float(np.min(self.H[wet])),
                "H_max": float(np.max(self.H[wet])),
                "wet_fraction": float(np.mean(self.mask3)),
                "depth_error_max": self.depth_error(), **self.roughness.as_dict()}


@dataclass(frozen=True)
class Roughness:
    """Roughness measures of a bathymetry (libs/core/bathymetry.py:20-29)."""

    r_std: float       # std(H) / mean(H)
    rx0: float         # max |dH| / (H_a + H_b) over wet faces
    h_ratio: float     # H_max / H_min

    def as_dict(self) -> dict[str, float]:
        return {"r_std": self.r_std, "rx0": self.rx0, "h_ratio": self.h_ratio}


def build_domain(grid: CGrid, H: np.ndarray, mask: np.ndarray, config=None, **over) -> Domain:
    get = (lambda k, d: config.get(k, d)) if config is not None else (lambda k, d: d)
    kw = dict(vcoord=str(get("domain.vcoord", "zlevel")),
              bc_x=str(get("domain.bc_x", "periodic")),
              bc_y=str(get("domain.bc_y", "periodic")),
              min_partial=float(get("domain.min_partial", 0.1)),
              face_rule=str(get("domain.face_rule", "min")))
    kw.update(over)
    return Domain(grid=grid, H=H, mask=mask, **kw)
