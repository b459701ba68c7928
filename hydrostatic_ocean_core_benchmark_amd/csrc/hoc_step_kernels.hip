// Kernels of the ocean-core time step (everything except the 2D Helmholtz
// solve, which lives in hoc_helmholtz.hip).
//
// Two launch shapes:
//   column kernels: one lane per (j, i) column, marching k; lanes are
//     consecutive in i, so every level's load is one coalesced row segment.
//     Used where the work is a recurrence in k (tridiagonal sweeps, prefix
//     sums, the TKE mixing-length march) and the horizontal stencil at each
//     level is fused into the same pass.
//   cell kernels: one lane per (k, j, i) for the pure 3D stencils.
//
// Column recurrences keep their forward-sweep values (c*, d*) in global
// scratch rather than registers so that any nz runs through the same code.
#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

// ------------------------------------------------------------ geometry
// Interface thickness of dz3 (libs/core/vertical_var.py:21-32) and the
// surface stress per unit density on faces, (tau / rho0) * mask3[u|v][0]
// (libs/core/model3d_v05.py:355-356).
__global__ void k_geometry(Dims d, StepParams p, const double* __restrict__ dz3,
                           const uint8_t* __restrict__ m3u, const uint8_t* __restrict__ m3v,
                           const double* __restrict__ tau_x, const double* __restrict__ tau_y,
                           double* __restrict__ dzi, double* __restrict__ tau_u,
                           double* __restrict__ tau_v) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int nz = d.nz;
  const long long n2 = d.n2;
  dzi[c] = dz3[c];
  for (int k = 1; k < nz; ++k) dzi[k * n2 + c] = 0.5 * (dz3[(k - 1) * n2 + c] + dz3[k * n2 + c]);
  dzi[nz * n2 + c] = dz3[(nz - 1) * n2 + c];
  tau_u[c] = (tau_x[c] / p.rho0) * mk(m3u[c]);
  tau_v[c] = (tau_y[c] / p.rho0) * mk(m3v[c]);
}

// ------------------------------------------ moving vertical coordinates
// z* and sigma (libs/core/domain.py layer_thickness / face_thickness /
// depth_at_centres, rebuilt from eta every step, model3d_v05.py:128-139):
// dz3 = dz3_ref * (1 + eta / H) in wet columns, centre depths from the
// surface, interface spacing; then the face thickness by the face rule.
__global__ void k_moving_dz(Dims d, const uint8_t* __restrict__ m2, const double* __restrict__ H,
                            const double* __restrict__ eta, const double* __restrict__ dz3_ref,
                            double* __restrict__ dz3, double* __restrict__ zc,
                            double* __restrict__ dzi) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const int nz = d.nz;
  const double scale = m2[c] ? 1.0 + eta[c] / (H[c] > 0.0 ? H[c] : 1.0) : 1.0;
  double above = 0.0, prev = 0.0;
  for (int k = 0; k < nz; ++k) {
    const double dz = dz3_ref[k * n2 + c] * scale;
    dz3[k * n2 + c] = dz;
    if (k == 1) above = prev;
    else if (k > 1) above = above + prev;
    zc[k * n2 + c] = above + 0.5 * dz;
    dzi[k * n2 + c] = (k == 0) ? dz : 0.5 * (prev + dz);
    prev = dz;
  }
  dzi[nz * n2 + c] = prev;
}

__global__ void k_moving_faces(Dims d, StepParams p, const uint8_t* __restrict__ m3u,
                               const uint8_t* __restrict__ m3v, const double* __restrict__ dz3,
                               double* __restrict__ dz3u, double* __restrict__ dz3v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2, o = t - c;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cn = at2(d, wrap(j, 1, d.ny), i);
  dz3u[t] = mk(m3u[t]) * face_value(p.face_rule, dz3[t], dz3[o + ce]);
  dz3v[t] = mk(m3v[t]) * face_value(p.face_rule, dz3[t], dz3[o + cn]);
}

// ---------------------------------------------------------- TKE closure
// libs/core/closure.py:141-205 in three passes over [nz+1][ny][nx] interface
// fields:
//   k_closure_lengths (integral form) or k_closure_shear + k_mixing_recursive
//     (recursive form): N^2 and shear^2 (s_n2, s_sh) and the mixing lengths
//     l_k = max(l_min, sqrt(l_up l_dn)) (s_l), l_eps = max(l_min, min(l_up,
//     l_dn)) (s_le);
//   k_closure: per column, the implicit interface tridiagonal for e (c* in
//     s_c, d* in e_out), then back-substitution, floors, K_m and K_h.

// N^2 and shear^2 on interface k of column c (closure.py:43-62,150-156).
__device__ __forceinline__ void shear_n2(const Dims& d, const StepParams& p, const Geo& geo,
                                         const double* u, const double* v, const double* b,
                                         long long c, int k, double& n2k, double& sh) {
  n2k = 0.0;
  sh = 0.0;
  if (k < 1 || k >= d.nz) return;
  const ColumnGeo cg(geo, p, d, c);
  const long long n2 = d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cs = at2(d, wrap(j, -1, d.ny), i);
  const long long a = (long long)(k - 1) * n2, o = (long long)k * n2;
  const double uc_prev = 0.5 * (u[a + c] + u[a + cw]);
  const double vc_prev = 0.5 * (v[a + c] + v[a + cs]);
  const double uc = 0.5 * (u[o + c] + u[o + cw]);
  const double vc = 0.5 * (v[o + c] + v[o + cs]);
  const double wet = (cg.wet(k - 1) ? 1.0 : 0.0) * (cg.wet(k) ? 1.0 : 0.0);
  const double idz = safe_inv(cg.dzi(k));
  const double du = (uc_prev - uc) * idz;
  const double dv = (vc_prev - vc) * idz;
  const double db = (b[a + c] - b[o + c]) * idz;
  sh = (du * du + dv * dv) * wet;
  n2k = db * wet;
}

// Integral (Gaspar) mixing lengths, closure.py:65-103. A workgroup owns
// MX_COLS consecutive columns: N^2 (clamped at 0) and the interface
// thicknesses of all their interfaces are staged in LDS once, then
// MX_TPC lanes per column march the interfaces k, k + MX_TPC, ... from LDS.
// Each march is the reference's loop with two exact shortcuts: it stops
// once the budget is spent (l and the budget are then frozen), and it skips
// the interfaces above the column's first and below its last one with
// N^2 != 0 or dz != 0 (a march from a dry interface under the bottom starts
// at the bottom), since a step across them changes neither l nor the budget.
constexpr int MX_COLS = 64, MX_TPC = 8, MX_THREADS = MX_COLS * MX_TPC;

__global__ __launch_bounds__(MX_THREADS) void k_closure_lengths(
    Dims d, StepParams p, Geo geo, const double* __restrict__ e_in, const double* __restrict__ u,
    const double* __restrict__ v, const double* __restrict__ b, double* __restrict__ s_n2,
    double* __restrict__ s_sh, double* __restrict__ s_l, double* __restrict__ s_le) {
  extern __shared__ double mx_lds[];
  __shared__ int kbeg[MX_COLS], kend[MX_COLS];
  const int nz = d.nz, ni = nz + 1;
  const long long n2 = d.n2;
  double* n2p = mx_lds;                 // [ni][MX_COLS] max(N^2, 0)
  double* dzl = mx_lds + ni * MX_COLS;  // [ni][MX_COLS] interface thickness
  const long long c0 = (long long)blockIdx.x * MX_COLS;
  for (int t = threadIdx.x; t < ni * MX_COLS; t += MX_THREADS) {
    const int kk = t / MX_COLS, col = t % MX_COLS;
    const long long c = c0 + col;
    double n2k = 0.0, sh = 0.0, dz = 0.0;
    if (c < n2) {
      shear_n2(d, p, geo, u, v, b, c, kk, n2k, sh);
      dz = ColumnGeo(geo, p, d, c).dzi(kk);
      s_n2[(long long)kk * n2 + c] = n2k;
      s_sh[(long long)kk * n2 + c] = sh;
    }
    n2p[t] = nmax(n2k, 0.0);
    dzl[t] = dz;
  }
  __syncthreads();
  if (threadIdx.x < MX_COLS) {
    const int col = threadIdx.x;
    int lo = ni, hi = -1;
    for (int kk = 0; kk < ni; ++kk) {
      const bool live = n2p[kk * MX_COLS + col] != 0.0 || dzl[kk * MX_COLS + col] != 0.0;
      if (live) {
        lo = kk < lo ? kk : lo;
        hi = kk;
      }
    }
    kbeg[col] = lo;
    kend[col] = hi + 1;
  }
  __syncthreads();
  const int col = threadIdx.x % MX_COLS, k0 = threadIdx.x / MX_COLS;
  const long long c = c0 + col;
  if (c >= n2) return;
  const int lo = kbeg[col], hi = kend[col];
  const ColumnGeo cg(geo, p, d, c);
  // column depth and the depth of interface k: dz3[0] + ... + dz3[k-1], in
  // that order
  double h = 0.0, zi = 0.0;
  if (cg.col) {
    h = cg_above(geo.zs, cg.nw, cg.hb, nz);
  } else {
    for (int k = 0; k < nz; ++k) h = h + geo.dz3[(long long)k * n2 + c];
  }
  int kz = 0;
  for (int k = k0; k <= nz; k += MX_TPC) {
    if (cg.col) {
      zi = cg_above(geo.zs, cg.nw, cg.hb, k);
    } else {
      for (; kz < k; ++kz) zi = zi + geo.dz3[(long long)kz * n2 + c];
    }
    const long long o = (long long)k * n2 + c;
    const double wet =
        (k >= 1 && k < nz) ? (cg.wet(k - 1) ? 1.0 : 0.0) * (cg.wet(k) ? 1.0 : 0.0) : 0.0;
    const double budget0 = nmax(e_in[o], p.e_min);
    // upward: interface k crosses k-1, ..., 0
    double l = 0.0, bud = budget0;
    for (int kk = (k - 1 < hi - 1 ? k - 1 : hi - 1); kk >= lo && bud > 0.0; --kk) {
      const double n2j = n2p[kk * MX_COLS + col];
      const double dzj = dzl[kk * MX_COLS + col];
      const double cost = n2j * dzj * (l + 0.5 * dzj);
      double step;
      if (cost <= bud) {
        step = dzj;
        bud = bud - cost;
      } else {
        if (n2j > 0.0) {
          const double s = sqrt(nmax(l * l + 2.0 * bud / n2j, 0.0)) - l;
          step = nmin(nmax(s, 0.0), dzj);
        } else {
          step = dzj;
        }
        bud = 0.0;
      }
      l = l + step;
    }
    const double l_up = nmin(l, zi + p.z_0) * wet;
    // downward: interface k crosses k+1, ..., nz
    l = 0.0;
    bud = budget0;
    for (int kk = (k + 1 > lo ? k + 1 : lo); kk < hi && bud > 0.0; ++kk) {
      const double n2j = n2p[kk * MX_COLS + col];
      const double dzj = dzl[kk * MX_COLS + col];
      const double cost = n2j * dzj * (l + 0.5 * dzj);
      double step;
      if (cost <= bud) {
        step = dzj;
        bud = bud - cost;
      } else {
        if (n2j > 0.0) {
          const double s = sqrt(nmax(l * l + 2.0 * bud / n2j, 0.0)) - l;
          step = nmin(nmax(s, 0.0), dzj);
        } else {
          step = dzj;
        }
        bud = 0.0;
      }
      l = l + step;
    }
    const double zb = nmax(h - zi, 0.0);
    const double l_dn = nmin(l, zb + p.z_0) * wet;
    s_l[o] = nmax(p.l_min, sqrt(l_up * l_dn));
    s_le[o] = nmax(p.l_min, nmin(l_up, l_dn));
  }
}

size_t closure_lengths_lds(int nz) { return (size_t)2 * (nz + 1) * MX_COLS * sizeof(double); }

void launch_closure_lengths(const Dims& d, const StepParams& p, const Geo& geo,
                            const double* e_in, const double* u, const double* v,
                            const double* b, double* s_n2, double* s_sh, double* s_l,
                            double* s_le, hipStream_t s) {
  const unsigned nb = (unsigned)((d.n2 + MX_COLS - 1) / MX_COLS);
  k_closure_lengths<<<nb, MX_THREADS, closure_lengths_lds(d.nz), s>>>(d, p, geo, e_in, u, v, b,
                                                                      s_n2, s_sh, s_l, s_le);
}

int closure_lengths_prepare(int nz) {
  const size_t lds = closure_lengths_lds(nz);
  if (lds > 160 * 1024) return -1;
  return hipFuncSetAttribute((const void*)k_closure_lengths,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess
             ? 0
             : -2;
}

// N^2 and shear^2 alone, one lane per interface cell (the recursive form
// needs N^2 before its length recursion).
__global__ void k_closure_shear(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                                const double* __restrict__ v, const double* __restrict__ b,
                                double* __restrict__ s_n2, double* __restrict__ s_sh) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.ni) return;
  const int k = (int)(t / d.n2);
  const long long c = t % d.n2;
  double n2k, sh;
  shear_n2(d, p, geo, u, v, b, c, k, n2k, sh);
  s_n2[t] = n2k;
  s_sh[t] = sh;
}

// Recursive (NEMO nn_mxl=2) mixing length, closure.py:106-129; s_l is the
// work array of the two sweeps and ends as l_k, s_le as l_eps.
__global__ void k_mixing_recursive(Dims d, StepParams p, Geo geo, const double* __restrict__ e_in,
                                   const double* __restrict__ s_n2, double* __restrict__ s_l,
                                   double* __restrict__ s_le) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int nz = d.nz;
  const long long n2 = d.n2;
  const ColumnGeo cg(geo, p, d, c);
  double h = 0.0;
  for (int k = 0; k < nz; ++k) h = h + cg.dz(k);
  double zi = 0.0;
  for (int k = 0; k <= nz; ++k) {
    if (k > 0) zi = zi + cg.dz(k - 1);
    const double n2c = nmax(s_n2[k * n2 + c], p.n2_min);
    const double ek = nmax(e_in[k * n2 + c], p.e_min);
    double l = sqrt(2.0 * ek / n2c);
    const double zb = nmax(h - zi, 0.0);
    l = nmin(l, nmin(zi + p.z_0, zb + p.z_0));
    if (k > 0) l = nmin(l, s_l[(k - 1) * n2 + c] + cg.dzi(k));
    s_l[k * n2 + c] = l;
  }
  for (int k = nz - 1; k >= 0; --k) {
    const double l = nmin(s_l[k * n2 + c], s_l[(k + 1) * n2 + c] + cg.dzi(k + 1));
    s_l[k * n2 + c] = l;
  }
  for (int k = 0; k <= nz; ++k) {
    const double wet = (k >= 1 && k < nz)
                           ? (cg.wet(k - 1) ? 1.0 : 0.0) * (cg.wet(k) ? 1.0 : 0.0) : 0.0;
    const double l = s_l[k * n2 + c] * wet;  // l_up = l_dn = l
    s_l[k * n2 + c] = nmax(p.l_min, sqrt(l * l));
    s_le[k * n2 + c] = nmax(p.l_min, nmin(l, l));
  }
}

// The interface tridiagonal for e and the eddy coefficients, one lane per
// column, from s_n2, s_sh, s_l = l_k and s_le = l_eps.
__global__ void k_closure(Dims d, StepParams p, Geo geo, const double* __restrict__ e_in,
                          double* __restrict__ Km, double* __restrict__ Kh,
                          double* __restrict__ e_out, const double* __restrict__ s_n2,
                          const double* __restrict__ s_sh, const double* __restrict__ s_l,
                          const double* __restrict__ s_le, double* __restrict__ s_c) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int nz = d.nz;
  const long long n2 = d.n2;
  const ColumnGeo cg(geo, p, d, c);
  const double kb = geo.kappa_b[c];
  const double dt = p.dt;
  double c_prev = 0.0, d_prev = 0.0;
  double rhs0 = 0.0;
  // K_m at interfaces k-1, k, k+1 and the wet flag of the cell above
  // interface k, carried down the column
  double km_up = 0.0, km_0 = Km[c], km_dn = nz >= 1 ? Km[n2 + c] : 0.0;
  bool wet_above = false;
  for (int k = 0; k <= nz; ++k) {
    const bool wet_below = k < nz && cg.wet(k);
    const double wet = (k >= 1 && k < nz) ? (wet_above ? 1.0 : 0.0) * (wet_below ? 1.0 : 0.0)
                                          : 0.0;
    const double ek = e_in[k * n2 + c];
    const double n2k = s_n2[k * n2 + c];
    const double l_eps = s_le[k * n2 + c];
    const double kmo = km_0;
    const double kho = Kh[k * n2 + c];
    const double prod = (kmo - p.nu_b) * s_sh[k * n2 + c] - (kho - kb) * n2k;
    const double diss = p.c_eps * sqrt(nmax(ek, p.e_min)) / l_eps;
    double a_up = 0.0, a_dn = 0.0;
    if (k >= 1 && k < nz) {
      const double idzi = safe_inv(cg.dzi(k));
      const double klay_up = 0.5 * (km_up + kmo);
      const double klay_dn = 0.5 * (kmo + km_dn);
      a_up = dt * klay_up * safe_inv(cg.dz(k - 1)) * idzi;
      a_dn = dt * klay_dn * safe_inv(cg.dz(k)) * idzi;
    }
    double sub = -a_up, sup = -a_dn;
    double diag = 1.0 + a_up + a_dn + dt * diss;
    double rhs = ek + dt * prod;
    if (wet <= 0.0) {
      diag = 1.0;
      sub = 0.0;
      sup = 0.0;
      rhs = p.e_min;
    }
    if (k == 0) {
      rhs = wet_below ? p.e_bb * geo.ustar2[c] : p.e_min;
      rhs0 = rhs;
    }
    double cst, dst;
    if (k == 0) {
      cst = sup / diag;
      dst = rhs / diag;
    } else {
      const double den = diag - sub * c_prev;
      cst = sup / den;
      dst = (rhs - sub * d_prev) / den;
    }
    s_c[k * n2 + c] = cst;
    e_out[k * n2 + c] = dst;
    c_prev = cst;
    d_prev = dst;
    km_up = km_0;
    km_0 = km_dn;
    km_dn = k + 2 <= nz ? Km[(k + 2) * n2 + c] : 0.0;
    wet_above = wet_below;
  }

  // back-substitution, floors, eddy coefficients
  double x_next = 0.0;
  bool wet_below = false;  // cell k (below interface k)
  for (int k = nz; k >= 0; --k) {
    const bool wet_above = k >= 1 && cg.wet(k - 1);
    const double dst = e_out[k * n2 + c];
    const double x = (k == nz) ? dst : dst - s_c[k * n2 + c] * x_next;
    x_next = x;
    const double wet = (k >= 1 && k < nz) ? (wet_above ? 1.0 : 0.0) * (wet_below ? 1.0 : 0.0)
                                          : 0.0;
    double en = nmax(x, p.e_min);
    en = wet > 0.0 ? en : p.e_min;
    if (k == 0) en = rhs0;
    const double km = p.c_k * s_l[k * n2 + c] * sqrt(en) * wet + p.nu_b;
    const double kh = (km - p.nu_b) / p.pr_t + kb;
    e_out[k * n2 + c] = en;
    Km[k * n2 + c] = km;
    Kh[k * n2 + c] = kh;
    wet_below = wet_above;
  }
}

// ------------------------------------------------ vertical coefficients
// tri_m (viscosity Km, bottom drag), factored, then q = A_m^-1 1
// (libs/core/model3d_v05.py:141-153); tri_b is factored by the tracer
// kernels from the K_h they read anyway.
__global__ void k_coefficients(Dims d, StepParams p, Geo geo, const double* __restrict__ Km,
                               double* __restrict__ sub_m, double* __restrict__ den_m,
                               double* __restrict__ cst_m, double* __restrict__ q) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int nz = d.nz;
  const long long n2 = d.n2;
  // the factorization and the forward sweep of q in one pass over the
  // column, so the geometry is read once and tri_m is not read back
  const ColumnGeo cg(geo, p, d, c);
  double cm_prev = 0.0, d_prev = 0.0;
  double idzi_t = safe_inv(cg.dzi(0));
  double km_t = Km[c];
  bool wet = cg.wet(0);
  for (int k = 0; k < nz; ++k) {
    const long long o = k * n2 + c;
    const double inv_dz = safe_inv(cg.dz(k));
    const double idzi_b = safe_inv(cg.dzi(k + 1));
    const double km_b = Km[o + n2];
    const bool below = (k + 1 < nz) && cg.wet(k + 1);
    double sm, dm, cm;
    factor_row(p, k, inv_dz, km_t, km_b, idzi_t, idzi_b, wet, below, p.thv_dt_drag, cm_prev,
               sm, dm, cm);
    sub_m[o] = sm;
    den_m[o] = dm;
    cst_m[o] = cm;
    const double dst = (k == 0) ? 1.0 / dm : (1.0 - sm * d_prev) / dm;
    q[o] = dst;
    d_prev = dst;
    cm_prev = cm;
    idzi_t = idzi_b;
    km_t = km_b;
    wet = below;
  }
  double x_next = q[(nz - 1) * n2 + c];
  for (int k = nz - 2; k >= 0; --k) {
    const long long o = k * n2 + c;
    const double x = q[o] - cst_m[o] * x_next;
    q[o] = x;
    x_next = x;
  }
}

// Helmholtz face coefficients Ku = sum_k dz3u min(q, q_east) mask_u and Kv
// (libs/core/model3d_v05.py:154-157), summed in k order like NumPy's
// axis-0 reduction.
__global__ void k_kukv(Dims d, StepParams p, Geo geo, const double* __restrict__ q,
                       double* __restrict__ Ku, double* __restrict__ Kv) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i);
  const long long n2 = d.n2;
  const FaceGeo fu(geo, p, d, c, ce, 0), fv(geo, p, d, c, cn, 1);
  double su = 0.0, sv = 0.0;
  for (int k = 0; k < d.nz; ++k) {
    const double qc = q[k * n2 + c];
    const double tu = fu.dz(k) * nmin(qc, q[k * n2 + ce]);
    const double tv = fv.dz(k) * nmin(qc, q[k * n2 + cn]);
    su = (k == 0) ? tu : su + tu;
    sv = (k == 0) ? tv : sv + tv;
  }
  Ku[c] = su * mk(geo.mu2[c]);
  Kv[c] = sv * mk(geo.mv2[c]);
}

// --------------------------------------------------- pressure potential
// Phi_k = 1/2 dz_k b_k + sum_{k'<k} dz b (libs/core/vertical_var.py:132-136);
// optionally the thickness-weighted depth mean removed (:139-145).
__global__ void k_phi(Dims d, StepParams p, Geo geo, const double* __restrict__ b,
                      double* __restrict__ phi) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const ColumnGeo cg(geo, p, d, c);
  double above = 0.0;
  for (int k = 0; k < d.nz; ++k) {
    const double dz = cg.dz(k);
    const double bk = b[k * n2 + c];
    phi[k * n2 + c] = 0.5 * dz * bk + above;
    const double t = dz * bk;
    above = (k == 0) ? t : above + t;
  }
  if (p.pgf_remove) {
    double tot = 0.0, wsum = 0.0;
    for (int k = 0; k < d.nz; ++k) {
      const double dz = cg.dz(k);
      const double t = dz * phi[k * n2 + c];
      tot = (k == 0) ? dz : tot + dz;
      wsum = (k == 0) ? t : wsum + t;
    }
    const double mean = wsum * safe_inv(tot);
    for (int k = 0; k < d.nz; ++k) phi[k * n2 + c] = phi[k * n2 + c] - mean;
  }
}

// ------------------------------------------------- vertical velocity
// w_k = -sum_{k'>=k} div(dz3u u, dz3v v), w_nz = 0, summed from the bottom
// (libs/core/vertical_var.py:109-121 with model3d_v05.py:232-235).
__global__ void k_w_diag(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                         const double* __restrict__ v, double* __restrict__ w) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cs = at2(d, wrap(j, -1, d.ny), i);
  const long long n2 = d.n2;
  const int nz = d.nz;
  const ColumnGeo cg(geo, p, d, c);
  const FaceGeo fe(geo, p, d, c, at2(d, j, wrap(i, 1, d.nx)), 0), fw(geo, p, d, cw, c, 0);
  const FaceGeo fn(geo, p, d, c, at2(d, wrap(j, 1, d.ny), i), 1), fs(geo, p, d, cs, c, 1);
  double s = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long o = k * n2;
    const double U = fe.dz(k) * u[o + c];
    const double Uw = fw.dz(k) * u[o + cw];
    const double V = fn.dz(k) * v[o + c];
    const double Vs = fs.dz(k) * v[o + cs];
    const double dv = (cg.wet(k) ? 1.0 : 0.0) * ((U - Uw) / p.dx + (V - Vs) / p.dy);
    s = (k == nz - 1) ? dv : s + dv;
    w[o + c] = -s;
  }
  w[(long long)nz * n2 + c] = 0.0;
}

// ------------------------------------------------- the step's prologue
// k_phi, k_depth_sum and k_w_diag of one column in one pass over it (the
// same operations in the same order, so the same bits): the top-down march
// forms phi from b and the transports U0 = sum dz3u u, V0 = sum dz3v v
// (when U0 is set), then the bottom-up march forms w (when w is set) from u
// and the west u, v and the south v, the column's own velocities in cache.
__global__ void k_prologue(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                           const double* __restrict__ v, const double* __restrict__ b,
                           double* __restrict__ phi, double* __restrict__ U0,
                           double* __restrict__ V0, double* __restrict__ w) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const int nz = d.nz;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cn = at2(d, wrap(j, 1, d.ny), i);
  const ColumnGeo cg(geo, p, d, c);
  const FaceGeo fe(geo, p, d, c, ce, 0), fn(geo, p, d, c, cn, 1);
  double above = 0.0, su = 0.0, sv = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long t = k * n2 + c;
    const double dz = cg.dz(k);
    const double bk = b[t];
    phi[t] = 0.5 * dz * bk + above;
    const double tb = dz * bk;
    above = (k == 0) ? tb : above + tb;
    if (U0) {
      const double tu = fe.dz(k) * u[t];
      const double tv = fn.dz(k) * v[t];
      su = (k == 0) ? tu : su + tu;
      sv = (k == 0) ? tv : sv + tv;
    }
  }
  if (U0) {
    U0[c] = su;
    V0[c] = sv;
  }
  if (p.pgf_remove) {
    double tot = 0.0, wsum = 0.0;
    for (int k = 0; k < nz; ++k) {
      const double dz = cg.dz(k);
      const double t = dz * phi[k * n2 + c];
      tot = (k == 0) ? dz : tot + dz;
      wsum = (k == 0) ? t : wsum + t;
    }
    const double mean = wsum * safe_inv(tot);
    for (int k = 0; k < nz; ++k) phi[k * n2 + c] = phi[k * n2 + c] - mean;
  }
  if (!w) return;
  const long long cw = at2(d, j, wrap(i, -1, d.nx)), cs = at2(d, wrap(j, -1, d.ny), i);
  const FaceGeo fw(geo, p, d, cw, c, 0), fs(geo, p, d, cs, c, 1);
  double sw = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long o = k * n2;
    const double U = fe.dz(k) * u[o + c];
    const double Uw = fw.dz(k) * u[o + cw];
    const double V = fn.dz(k) * v[o + c];
    const double Vs = fs.dz(k) * v[o + cs];
    const double dv = (cg.wet(k) ? 1.0 : 0.0) * ((U - Uw) / p.dx + (V - Vs) / p.dy);
    sw = (k == nz - 1) ? dv : sw + dv;
    w[o + c] = -sw;
  }
  w[(long long)nz * n2 + c] = 0.0;
}

// ------------------------------------------------ explicit momentum terms
struct Nb {  // periodic neighbour offsets of a column (a plane has < 2^31 columns)
  int c, e, w, ee, ww, n, s, nn, ss, ne, nw, se, sw;
};
__device__ __forceinline__ Nb neighbours(const Dims& d, int j, int i) {
  Nb q;
  const int ie = wrap(i, 1, d.nx), iw = wrap(i, -1, d.nx);
  const int iee = wrap(i, 2, d.nx), iww = wrap(i, -2, d.nx);
  const int jn = wrap(j, 1, d.ny), js = wrap(j, -1, d.ny);
  const int jnn = wrap(j, 2, d.ny), jss = wrap(j, -2, d.ny);
  q.c = (int)at2(d, j, i);
  q.e = (int)at2(d, j, ie);
  q.w = (int)at2(d, j, iw);
  q.ee = (int)at2(d, j, iee);
  q.ww = (int)at2(d, j, iww);
  q.n = (int)at2(d, jn, i);
  q.s = (int)at2(d, js, i);
  q.nn = (int)at2(d, jnn, i);
  q.ss = (int)at2(d, jss, i);
  q.ne = (int)at2(d, jn, ie);
  q.nw = (int)at2(d, jn, iw);
  q.se = (int)at2(d, js, ie);
  q.sw = (int)at2(d, js, iw);
  return q;
}

// Coriolis partner of u: v averaged over the wet v faces around the u point
// (libs/core/operators_masked.py:45-52).
__device__ __forceinline__ double avg_v_to_u(const double* v, const uint8_t* mv, const uint8_t* mu,
                                             long long o, const Nb& q) {
  const double w0 = mk(mv[o + q.c]), w1 = mk(mv[o + q.s]), w2 = mk(mv[o + q.e]),
               w3 = mk(mv[o + q.se]);
  const double num = 0.0 + w0 * v[o + q.c] + w1 * v[o + q.s] + w2 * v[o + q.e] + w3 * v[o + q.se];
  const double den = 0.0 + w0 + w1 + w2 + w3;
  const double out = den > 0.0 ? num / den : 0.0;
  return out * mk(mu[o + q.c]);
}
// (libs/core/operators_masked.py:55-61)
__device__ __forceinline__ double avg_u_to_v(const double* u, const uint8_t* mu, const uint8_t* mv,
                                             long long o, const Nb& q) {
  const double w0 = mk(mu[o + q.c]), w1 = mk(mu[o + q.w]), w2 = mk(mu[o + q.n]),
               w3 = mk(mu[o + q.nw]);
  const double num = 0.0 + w0 * u[o + q.c] + w1 * u[o + q.w] + w2 * u[o + q.n] + w3 * u[o + q.nw];
  const double den = 0.0 + w0 + w1 + w2 + w3;
  const double out = den > 0.0 ? num / den : 0.0;
  return out * mk(mv[o + q.c]);
}

// Explicit vertical diffusion Dz[f] at layer k with surface flux sflux, no
// flux through a column's bottom interface and linear drag on the bottom
// cell (libs/core/vertical_var.py:80-106). dzf is the thickness of f's
// points (dz3u for u, dz3v for v, dz3 for tracers); cell mask3 decides the
// bottom.
__device__ __forceinline__ double explicit_diffusion(const Dims& d, const uint8_t* m3,
                                                     const double* f, const double* nu,
                                                     const double* dzf, double sflux,
                                                     double drag, long long c, int k) {
  const int nz = d.nz;
  const long long n2 = d.n2;
  auto dzi_f = [&](int kk) -> double {  // interface_thickness(dzf)[kk]
    if (kk == 0) return dzf[c];
    if (kk == nz) return dzf[(nz - 1) * n2 + c];
    return 0.5 * (dzf[(kk - 1) * n2 + c] + dzf[kk * n2 + c]);
  };
  auto is_bottom = [&](int kk) -> bool {
    return m3[kk * n2 + c] != 0 && !((kk + 1 < nz) && m3[(kk + 1) * n2 + c] != 0);
  };
  // flux through the top (interface k) and bottom (interface k+1) of cell k
  double ft;
  if (k == 0) {
    ft = sflux;
  } else {
    ft = nu[k * n2 + c] * (f[(k - 1) * n2 + c] - f[k * n2 + c]) * safe_inv(dzi_f(k));
    if (is_bottom(k - 1)) ft = 0.0;
  }
  double fb = 0.0;
  if (k + 1 < nz) {
    fb = nu[(k + 1) * n2 + c] * (f[k * n2 + c] - f[(k + 1) * n2 + c]) * safe_inv(dzi_f(k + 1));
    if (is_bottom(k)) fb = 0.0;
  }
  const double inv = safe_inv(dzf[k * n2 + c]);
  const double out = (ft - fb) * inv;
  const double dr = drag * f[k * n2 + c] * (is_bottom(k) ? 1.0 : 0.0);
  return (out - dr * inv) * (m3[k * n2 + c] != 0 ? 1.0 : 0.0);
}

// u-momentum advection at (k, j, i) (libs/core/advection_v06.py:77-99 for
// up3/up3_tvd, libs/core/advection.py:62-79 for centred2).
__device__ double advect_u(const Dims& d, const StepParams& p, const uint8_t* m3u,
                           const double* u, const double* v, const double* w, long long o,
                           int k, const Nb& q) {
  const long long n2 = d.n2;
  const int nz = d.nz;
  const bool lim = p.advection == ADV_UP3_TVD;
  const double* U = u + o;
  const double* V = v + o;
  double ax, ay, az;
  if (p.advection == ADV_CENTRED2) {
    const double ub = 0.5 * (U[q.w] + U[q.c]);
    const double ube = 0.5 * (U[q.c] + U[q.e]);
    ax = (ube * ube - ub * ub) / p.dx;
    const double vb = 0.5 * (V[q.c] + V[q.e]);
    const double vbs = 0.5 * (V[q.s] + V[q.se]);
    const double fy = vb * (0.5 * (U[q.c] + U[q.n]));
    const double fys = vbs * (0.5 * (U[q.s] + U[q.c]));
    ay = (fy - fys) / p.dy;
  } else {
    const uint8_t* M = m3u + o;
    // x flux at centres i (faces i-1|i) and i+1 (faces i|i+1)
    const double ub = 0.5 * (U[q.w] + U[q.c]);
    const double ube = 0.5 * (U[q.c] + U[q.e]);
    const double uc = kappa_face(U[q.ww], U[q.w], U[q.c], U[q.e], ub, M[q.ww] != 0,
                                 M[q.e] != 0, lim);
    const double uce = kappa_face(U[q.w], U[q.c], U[q.e], U[q.ee], ube, M[q.w] != 0,
                                  M[q.ee] != 0, lim);
    ax = (ube * uce - ub * uc) / p.dx;
    // y flux at corners j (between u[j], u[j+1]) and j-1
    const double vb = 0.5 * (V[q.c] + V[q.e]);
    const double vbs = 0.5 * (V[q.s] + V[q.se]);
    const double uy = kappa_face(U[q.s], U[q.c], U[q.n], U[q.nn], vb, M[q.s] != 0,
                                 M[q.nn] != 0, lim);
    const double uys = kappa_face(U[q.ss], U[q.s], U[q.c], U[q.n], vbs, M[q.ss] != 0,
                                  M[q.n] != 0, lim);
    ay = (vb * uy - vbs * uys) / p.dy;
  }
  // vertical flux on interfaces k (top) and k+1 (bottom)
  auto fz = [&](int kk) -> double {
    if (kk <= 0 || kk >= nz) return 0.0;
    const long long a = (long long)kk * n2;
    const double wb = 0.5 * (w[a + q.c] + w[a + q.e]);
    const double above = u[a - n2 + q.c], below = u[a + q.c];
    if (p.advection == ADV_CENTRED2) return wb * 0.5 * (above + below);
    const double far_up = kk >= 2 ? u[a - 2 * n2 + q.c] : 0.0;
    const double far_dn = kk <= nz - 2 ? u[a + n2 + q.c] : 0.0;
    const bool hu = kk >= 2 && m3u[a - 2 * n2 + q.c] != 0;
    const bool hd = kk <= nz - 2 && m3u[a + n2 + q.c] != 0;
    return wb * kappa_face_z(far_up, above, below, far_dn, wb, hu, hd, lim);
  };
  az = (fz(k) - fz(k + 1)) / p.dz_nom;
  return ax + ay + az;
}

// v-momentum advection (libs/core/advection_v06.py:100-108, advection.py:81-95).
__device__ double advect_v(const Dims& d, const StepParams& p, const uint8_t* m3v,
                           const double* u, const double* v, const double* w, long long o,
                           int k, const Nb& q) {
  const long long n2 = d.n2;
  const int nz = d.nz;
  const bool lim = p.advection == ADV_UP3_TVD;
  const double* U = u + o;
  const double* V = v + o;
  double bx, by, bz;
  if (p.advection == ADV_CENTRED2) {
    const double vb = 0.5 * (V[q.s] + V[q.c]);
    const double vbn = 0.5 * (V[q.c] + V[q.n]);
    by = (vbn * vbn - vb * vb) / p.dy;
    const double ub = 0.5 * (U[q.c] + U[q.n]);
    const double ubw = 0.5 * (U[q.w] + U[q.nw]);
    const double fx = ub * (0.5 * (V[q.c] + V[q.e]));
    const double fxw = ubw * (0.5 * (V[q.w] + V[q.c]));
    bx = (fx - fxw) / p.dx;
  } else {
    const uint8_t* M = m3v + o;
    const double vb = 0.5 * (V[q.s] + V[q.c]);
    const double vbn = 0.5 * (V[q.c] + V[q.n]);
    const double vc = kappa_face(V[q.ss], V[q.s], V[q.c], V[q.n], vb, M[q.ss] != 0,
                                 M[q.n] != 0, lim);
    const double vcn = kappa_face(V[q.s], V[q.c], V[q.n], V[q.nn], vbn, M[q.s] != 0,
                                  M[q.nn] != 0, lim);
    by = (vbn * vcn - vb * vc) / p.dy;
    const double ub = 0.5 * (U[q.c] + U[q.n]);
    const double ubw = 0.5 * (U[q.w] + U[q.nw]);
    const double vx = kappa_face(V[q.w], V[q.c], V[q.e], V[q.ee], ub, M[q.w] != 0,
                                 M[q.ee] != 0, lim);
    const double vxw = kappa_face(V[q.ww], V[q.w], V[q.c], V[q.e], ubw, M[q.ww] != 0,
                                  M[q.e] != 0, lim);
    bx = (ub * vx - ubw * vxw) / p.dx;
  }
  auto fz = [&](int kk) -> double {
    if (kk <= 0 || kk >= nz) return 0.0;
    const long long a = (long long)kk * n2;
    const double wb = 0.5 * (w[a + q.c] + w[a + q.n]);
    const double above = v[a - n2 + q.c], below = v[a + q.c];
    if (p.advection == ADV_CENTRED2) return wb * 0.5 * (above + below);
    const double far_up = kk >= 2 ? v[a - 2 * n2 + q.c] : 0.0;
    const double far_dn = kk <= nz - 2 ? v[a + n2 + q.c] : 0.0;
    const bool hu = kk >= 2 && m3v[a - 2 * n2 + q.c] != 0;
    const bool hd = kk <= nz - 2 && m3v[a + n2 + q.c] != 0;
    return wb * kappa_face_z(far_up, above, below, far_dn, wb, hu, hd, lim);
  };
  bz = (fz(k) - fz(k + 1)) / p.dz_nom;
  return bx + by + bz;
}

// Masked horizontal Laplacian D o G of a at (o + q.c) with gradient masks
// gmu (x faces), gmv (y faces) and divergence mask dm
// (libs/core/operators_masked.py:64-68).
__device__ __forceinline__ double laplacian(const StepParams& p, const double* a,
                                            const uint8_t* gmu, const uint8_t* gmv,
                                            const uint8_t* dm, long long o, const Nb& q) {
  const double* A = a + o;
  const double gx = mk(gmu[o + q.c]) * (A[q.e] - A[q.c]) / p.dx;
  const double gxw = mk(gmu[o + q.w]) * (A[q.c] - A[q.w]) / p.dx;
  const double gy = mk(gmv[o + q.c]) * (A[q.n] - A[q.c]) / p.dy;
  const double gys = mk(gmv[o + q.s]) * (A[q.c] - A[q.s]) / p.dy;
  return mk(dm[o + q.c]) * ((gx - gxw) / p.dx + (gy - gys) / p.dy);
}

// One lane per cell: the step's explicit momentum terms from the state at
// the start of the step (libs/core/model3d_v05.py:352-370):
//   pgx, pgy   baroclinic pressure gradient at the face depth (:195-230)
//   avu, avv   Coriolis partners avg_u_to_v(u), avg_v_to_u(v)
//   dzu, dzv   explicit vertical diffusion with surface stress and drag
//   stu, stv   theta_v dt tau / dz (k = 0) + dt * (-advection + A_h lap)
__global__ void k_momentum_explicit(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                                    const double* __restrict__ v, const double* __restrict__ b,
                                    const double* __restrict__ phi, const double* __restrict__ w,
                                    const double* __restrict__ Km, double* __restrict__ pgx,
                                    double* __restrict__ pgy, double* __restrict__ avu,
                                    double* __restrict__ avv, double* __restrict__ dzu,
                                    double* __restrict__ dzv, double* __restrict__ stu,
                                    double* __restrict__ stv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2;
  const int k = (int)(t / d.n2);
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const Nb q = neighbours(d, j, i);
  const long long o = (long long)k * d.n2;

  // pressure gradient
  double gx, gy;
  if (p.pgf_correction) {
    const double zc = geo.zc[o + c];
    {
      const double zr = geo.zc[o + q.e];
      const double zf = 0.5 * (zc + zr);
      const double left = phi[o + c] + b[o + c] * (zf - zc);
      const double right = phi[o + q.e] + b[o + q.e] * (zf - zr);
      gx = mk(geo.m3u[o + c]) * (right - left) / p.dx;
    }
    {
      const double zr = geo.zc[o + q.n];
      const double zf = 0.5 * (zc + zr);
      const double left = phi[o + c] + b[o + c] * (zf - zc);
      const double right = phi[o + q.n] + b[o + q.n] * (zf - zr);
      gy = mk(geo.m3v[o + c]) * (right - left) / p.dy;
    }
  } else {
    gx = mk(geo.m3u[o + c]) * (phi[o + q.e] - phi[o + c]) / p.dx * mk(geo.m3u[o + c]);
    gy = mk(geo.m3v[o + c]) * (phi[o + q.n] - phi[o + c]) / p.dy * mk(geo.m3v[o + c]);
  }
  pgx[t] = gx;
  pgy[t] = gy;

  {  // the Coriolis partners, for callers that ask for them
    if (avv) avv[t] = avg_v_to_u(v, geo.m3v, geo.m3u, o, q);
    if (avu) avu[t] = avg_u_to_v(u, geo.m3u, geo.m3v, o, q);
  }

  dzu[t] = explicit_diffusion(d, geo.m3, u, Km, geo.dz3u, k == 0 ? geo.tau_u[c] : 0.0,
                              p.bottom_drag, c, k);
  dzv[t] = explicit_diffusion(d, geo.m3, v, Km, geo.dz3v, k == 0 ? geo.tau_v[c] : 0.0,
                              p.bottom_drag, c, k);

  double exu = 0.0, exv = 0.0;
  if (p.advection != ADV_NONE) {
    double au, av;
    if (p.advection == ADV_UPWIND1) {
      // upwind1 momentum uses the centred2 form (libs/core/model3d_v05.py:251-254)
      StepParams pc = p;
      pc.advection = ADV_CENTRED2;
      au = advect_u(d, pc, geo.m3u, u, v, w, o, k, q);
      av = advect_v(d, pc, geo.m3v, u, v, w, o, k, q);
    } else {
      au = advect_u(d, p, geo.m3u, u, v, w, o, k, q);
      av = advect_v(d, p, geo.m3v, u, v, w, o, k, q);
    }
    exu = exu - au * mk(geo.m3u[o + c]);
    exv = exv - av * mk(geo.m3v[o + c]);
  }
  if (p.has_A_h) {
    exu = exu + p.A_h * laplacian(p, u, geo.m3u, geo.m3v, geo.m3u, o, q);
    exv = exv + p.A_h * laplacian(p, v, geo.m3u, geo.m3v, geo.m3v, o, q);
  }
  double s0u = 0.0, s0v = 0.0;
  if (k == 0) {
    s0u = p.thv_dt * geo.tau_u[c] * safe_inv(geo.dz3u[c]);
    s0v = p.thv_dt * geo.tau_v[c] * safe_inv(geo.dz3v[c]);
  }
  stu[t] = s0u + p.dt * exu;
  stv[t] = s0v + p.dt * exv;
}

// Depth-integrated transports U = sum_k dz3u u, V = sum_k dz3v v (k order).
__global__ void k_depth_sum(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                            const double* __restrict__ v, double* __restrict__ U,
                            double* __restrict__ V) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const FaceGeo fu(geo, p, d, c, at2(d, j, wrap(i, 1, d.nx)), 0);
  const FaceGeo fv(geo, p, d, c, at2(d, wrap(j, 1, d.ny), i), 1);
  double su = 0.0, sv = 0.0;
  for (int k = 0; k < d.nz; ++k) {
    const double tu = fu.dz(k) * u[k * n2 + c];
    const double tv = fv.dz(k) * v[k * n2 + c];
    su = (k == 0) ? tu : su + tu;
    sv = (k == 0) ? tv : sv + tv;
  }
  U[c] = su;
  V[c] = sv;
}

// div = mask ((U - U_w)/dx + (V - V_s)/dy) (libs/core/operators_masked.py:27-32).
__device__ __forceinline__ double div2(const Dims& d, const StepParams& p, const Geo& geo,
                                       const double* U, const double* V, long long c) {
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cs = at2(d, wrap(j, -1, d.ny), i);
  return mk(geo.m2[c]) * ((U[c] - U[cw]) / p.dx + (V[c] - V[cs]) / p.dy);
}

// Barotropic divergence with the Flather transports on open edges
// (model3d_v05.py:262-309): on an open edge column the wrapped neighbour
// flux of the masked divergence is swapped for the radiating transport
// u_ext -/+ sqrt(g max(H, 1)) (eta - eta_ext), x edges first, then the
// result is masked again.
__device__ __forceinline__ double div_bt(const Dims& d, const StepParams& p, const Geo& geo,
                                         const double* U, const double* V, const double* eta,
                                         long long c) {
  double dv = div2(d, p, geo, U, V, c);
  if (!p.open_x && !p.open_y) return dv;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  if (p.open_x) {
    const long long cw = at2(d, j, wrap(i, -1, d.nx));
    const double cg = sqrt(p.g * nmax(geo.H[c], 1.0));
    if (i == 0) {
      const double bw = p.u_ext - cg * (eta[c] - p.eta_ext);
      dv = dv + ((U[c] - bw) / p.dx - ((U[c] - U[cw]) / p.dx));
    }
    if (i == d.nx - 1) {
      const double be = p.u_ext + cg * (eta[c] - p.eta_ext);
      dv = dv + ((be - U[cw]) / p.dx - ((U[c] - U[cw]) / p.dx));
    }
  }
  if (p.open_y) {
    const long long cs = at2(d, wrap(j, -1, d.ny), i);
    const double cg = sqrt(p.g * nmax(geo.H[c], 1.0));
    if (j == 0) {
      const double bs = p.u_ext - cg * (eta[c] - p.eta_ext);
      dv = dv + ((V[c] - bs) / p.dy - ((V[c] - V[cs]) / p.dy));
    }
    if (j == d.ny - 1) {
      const double bn = p.u_ext + cg * (eta[c] - p.eta_ext);
      dv = dv + ((bn - V[cs]) / p.dy - ((V[c] - V[cs]) / p.dy));
    }
  }
  return dv * mk(geo.m2[c]);
}

__global__ void k_div2(Dims d, StepParams p, Geo geo, const double* __restrict__ U,
                       const double* __restrict__ V, const double* __restrict__ eta,
                       double* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  out[c] = div_bt(d, p, geo, U, V, eta, c);
}

// rhs = eta - dt ((1 - th) div_old + th div(U_g, V_g))  (model3d_v05.py:398-400)
__global__ void k_rhs_eta(Dims d, StepParams p, Geo geo, const double* __restrict__ eta,
                          const double* __restrict__ div_old, const double* __restrict__ Ug,
                          const double* __restrict__ Vg, double* __restrict__ rhs) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const double dg = div_bt(d, p, geo, Ug, Vg, eta, c);
  rhs[c] = eta[c] - p.dt * (p.one_m_th * div_old[c] + p.th * dg);
}

// fb: eta_new = eta - dt div(U, V)  (model3d_v05.py:410-413)
__global__ void k_eta_fb(Dims d, StepParams p, Geo geo, const double* __restrict__ eta,
                         const double* __restrict__ U, const double* __restrict__ V,
                         double* __restrict__ eta_new) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  eta_new[c] = eta[c] - p.dt * div_bt(d, p, geo, U, V, eta, c);
}

// ---------------------------------------------- split-explicit mode
// Barotropic forcing of the split scheme (model3d_v05.py:445-452): the
// Coriolis the barotropic step adds back, formed with the same 2D transport
// averages, removed from the depth-integrated baroclinic tendency.
__global__ void k_split_forcing(Dims d, StepParams p, Geo geo, const double* __restrict__ U0,
                                const double* __restrict__ V0, const double* __restrict__ Um,
                                const double* __restrict__ Vm, double* __restrict__ fx,
                                double* __restrict__ fy) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const Nb q = neighbours(d, j, i);
  const double cor_x = p.f * (p.one_m_tc * avg_v_to_u(V0, geo.mv2, geo.mu2, 0, q) +
                              p.theta_cor * avg_v_to_u(Vm, geo.mv2, geo.mu2, 0, q));
  const double cor_y = p.neg_f * (p.one_m_tc * avg_u_to_v(U0, geo.mu2, geo.mv2, 0, q) +
                                  p.theta_cor * avg_u_to_v(Um, geo.mu2, geo.mv2, 0, q));
  fx[c] = (Um[c] - U0[c]) / p.dt - cor_x;
  fy[c] = (Vm[c] - V0[c]) / p.dt - cor_y;
}

// One Picard pass of a barotropic substep (model3d_v05.py:321-327): from the
// substep's start U, V, e and the current iterate Ui, Vi.
__global__ void k_bt_picard(Dims d, StepParams p, Geo geo, const double* __restrict__ U,
                            const double* __restrict__ V, const double* __restrict__ e,
                            const double* __restrict__ Ui, const double* __restrict__ Vi,
                            const double* __restrict__ fx, const double* __restrict__ fy,
                            const double* __restrict__ Ku, const double* __restrict__ Kv,
                            double* __restrict__ Uo, double* __restrict__ Vo) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const Nb q = neighbours(d, j, i);
  const double mu = mk(geo.mu2[c]), mv = mk(geo.mv2[c]);
  const double gxe = mu * (e[q.e] - e[c]) / p.dx;
  const double gye = mv * (e[q.n] - e[c]) / p.dy;
  const double v_cor = p.one_m_tc * avg_v_to_u(V, geo.mv2, geo.mu2, 0, q) +
                       p.theta_cor * avg_v_to_u(Vi, geo.mv2, geo.mu2, 0, q);
  const double u_cor = p.one_m_tc * avg_u_to_v(U, geo.mu2, geo.mv2, 0, q) +
                       p.theta_cor * avg_u_to_v(Ui, geo.mu2, geo.mv2, 0, q);
  Uo[c] = (U[c] + p.ddt * (p.f * v_cor - p.g * Ku[c] * gxe + fx[c])) * mu;
  Vo[c] = (V[c] + p.ddt * (p.neg_f * u_cor - p.g * Kv[c] * gye + fy[c])) * mv;
}

// e_new = e - ddt div(U, V); the transports summed for the averaged mode
// (model3d_v05.py:329-331).
__global__ void k_bt_eta(Dims d, StepParams p, Geo geo, const double* __restrict__ e,
                         const double* __restrict__ U, const double* __restrict__ V,
                         double* __restrict__ e_new, double* __restrict__ acc_u,
                         double* __restrict__ acc_v) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  e_new[c] = e[c] - p.ddt * div_bt(d, p, geo, U, V, e, c);
  if (acc_u) {
    acc_u[c] = acc_u[c] + U[c];
    acc_v[c] = acc_v[c] + V[c];
  }
}

__global__ void k_bt_mean(Dims d, StepParams p, const double* __restrict__ acc_u,
                          const double* __restrict__ acc_v, double* __restrict__ U,
                          double* __restrict__ V) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  U[c] = acc_u[c] / (double)p.n_split;
  V[c] = acc_v[c] / (double)p.n_split;
}

// Depth-mean replacement (model3d_v05.py:454-462): u = (gh_u - Um/Hu + U/Hu)
// on the wet u faces, with Hu = sum_k dz3u.
__global__ void k_split_final(Dims d, Geo geo, const double* __restrict__ gh_u,
                              const double* __restrict__ gh_v, const double* __restrict__ Um,
                              const double* __restrict__ Vm, const double* __restrict__ U,
                              const double* __restrict__ V, const double* __restrict__ Hu,
                              const double* __restrict__ Hv, double* __restrict__ u,
                              double* __restrict__ v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2;
  const double ihu = safe_inv(Hu[c]), ihv = safe_inv(Hv[c]);
  u[t] = (gh_u[t] - Um[c] * ihu + U[c] * ihu) * mk(geo.m3u[t]);
  v[t] = (gh_v[t] - Vm[c] * ihv + V[c] * ihv) * mk(geo.m3v[t]);
}

// Hu = sum_k dz3u, Hv = sum_k dz3v, summed from the surface.
__global__ void k_face_depth(Dims d, Geo geo, double* __restrict__ Hu, double* __restrict__ Hv) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  double su = 0.0, sv = 0.0;
  for (int k = 0; k < d.nz; ++k) {
    su = (k == 0) ? geo.dz3u[c] : su + geo.dz3u[k * d.n2 + c];
    sv = (k == 0) ? geo.dz3v[c] : sv + geo.dz3v[k * d.n2 + c];
  }
  Hu[c] = su;
  Hv[c] = sv;
}

// ------------------------------------------- predictor and column solves
// One lane per column, one Picard pass (model3d_v05.py:387-396):
//   v_cor = (1-tc) avg_v_to_u(v) + tc avg_v_to_u(v_it),  u_cor likewise
// (both partners formed here, from the state and from the iterate)
//   gu = u + dt (f v_cor + pgx) - g dt (1-th) Gx(eta) + (1-thv) dt dzu
//   gh_u = A_m^-1 (gu + stu) * mask3, and the same for v,
// then the depth sums Ug = sum dz3u gh_u, Vg = sum dz3v gh_v for the
// Helmholtz right-hand side. gh_u / gh_v hold d* between the sweeps.
__global__ void k_predictor(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                            const double* __restrict__ v, const double* __restrict__ eta,
                            const double* __restrict__ u_it, const double* __restrict__ v_it,
                            const double* __restrict__ pgx, const double* __restrict__ pgy,
                            const double* __restrict__ dzu, const double* __restrict__ dzv,
                            const double* __restrict__ stu, const double* __restrict__ stv,
                            const double* __restrict__ sub, const double* __restrict__ den,
                            const double* __restrict__ cst, double* __restrict__ gh_u,
                            double* __restrict__ gh_v, double* __restrict__ Ug,
                            double* __restrict__ Vg) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const Nb q = neighbours(d, j, i);
  const long long n2 = d.n2;
  const int nz = d.nz;
  const bool split = p.scheme == SCHEME_SPLIT;
  const double gx2 = mk(geo.mu2[c]) * (eta[q.e] - eta[c]) / p.dx;
  const double gy2 = mk(geo.mv2[c]) * (eta[q.n] - eta[c]) / p.dy;
  double du_prev = 0.0, dv_prev = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long o = k * n2;
    const long long t = o + c;
    // Coriolis partners of the old state, formed here from u and v (their
    // neighbours are in cache) instead of stored and re-read; in the first
    // Picard pass the iterate is the old state itself
    const double av_v = avg_v_to_u(v, geo.m3v, geo.m3u, o, q);
    const double av_u = avg_u_to_v(u, geo.m3u, geo.m3v, o, q);
    const double av_v_it = v_it == v ? av_v : avg_v_to_u(v_it, geo.m3v, geo.m3u, o, q);
    const double av_u_it = u_it == u ? av_u : avg_u_to_v(u_it, geo.m3u, geo.m3v, o, q);
    const double vcor = p.one_m_tc * av_v + p.theta_cor * av_v_it;
    const double ucor = p.one_m_tc * av_u + p.theta_cor * av_u_it;
    const double gxe = gx2 * mk(geo.m3u[t]);
    const double gye = gy2 * mk(geo.m3v[t]);
    // split_explicit leaves the surface gradient to the barotropic substeps
    // (model3d_v05.py:431-432)
    const double gu = split ? u[t] + p.dt * (p.f * vcor + pgx[t]) + p.c_dzd * dzu[t]
                            : u[t] + p.dt * (p.f * vcor + pgx[t]) - p.c_gx * gxe + p.c_dzd * dzu[t];
    const double gv = split ? v[t] + p.dt * (p.neg_f * ucor + pgy[t]) + p.c_dzd * dzv[t]
                            : v[t] + p.dt * (p.neg_f * ucor + pgy[t]) - p.c_gx * gye +
                                  p.c_dzd * dzv[t];
    const double ru = gu + stu[t];
    const double rv = gv + stv[t];
    double du, dv;
    if (k == 0) {
      du = ru / den[t];
      dv = rv / den[t];
    } else {
      du = (ru - sub[t] * du_prev) / den[t];
      dv = (rv - sub[t] * dv_prev) / den[t];
    }
    gh_u[t] = du;
    gh_v[t] = dv;
    du_prev = du;
    dv_prev = dv;
  }
  double xu = 0.0, xv = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long t = k * n2 + c;
    if (k == nz - 1) {
      xu = gh_u[t];
      xv = gh_v[t];
    } else {
      xu = gh_u[t] - cst[t] * xu;
      xv = gh_v[t] - cst[t] * xv;
    }
    const double mm = mk(geo.m3[t]);
    gh_u[t] = xu * mm;
    gh_v[t] = xv * mm;
  }
  double su = 0.0, sv = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long t = k * n2 + c;
    const double tu = geo.dz3u[t] * gh_u[t];
    const double tv = geo.dz3v[t] * gh_v[t];
    su = (k == 0) ? tu : su + tu;
    sv = (k == 0) ? tv : sv + tv;
  }
  Ug[c] = su;
  Vg[c] = sv;
}

// Coriolis partners with the face masks given as flags: the same sums in
// the same order as avg_v_to_u / avg_u_to_v.
__device__ __forceinline__ double avg4(double a0, double a1, double a2, double a3, bool f0,
                                       bool f1, bool f2, bool f3, bool own) {
  const double w0 = f0 ? 1.0 : 0.0, w1 = f1 ? 1.0 : 0.0, w2 = f2 ? 1.0 : 0.0, w3 = f3 ? 1.0 : 0.0;
  const double num = 0.0 + w0 * a0 + w1 * a1 + w2 * a2 + w3 * a3;
  const double den = 0.0 + w0 + w1 + w2 + w3;
  const double out = den > 0.0 ? num / den : 0.0;
  return out * (own ? 1.0 : 0.0);
}

// k_predictor for a static z-level geometry in column form: the same
// operations on the same operands in the same order (the same bits), with
// the cell and face masks of the Coriolis averages and the face thicknesses
// of the depth sums formed from the wet-level counts of the column and its
// neighbours instead of read per cell. At most 128 VGPRs, so that a 512 x
// 512 plane of columns is resident at once (4 waves per SIMD): the column
// march is latency-bound.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_predictor_col(
    Dims d, StepParams p, Geo geo, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ eta, const double* __restrict__ u_it,
    const double* __restrict__ v_it, const double* __restrict__ pgx,
    const double* __restrict__ pgy, const double* __restrict__ dzu,
    const double* __restrict__ dzv, const double* __restrict__ stu,
    const double* __restrict__ stv, const double* __restrict__ sub,
    const double* __restrict__ den, const double* __restrict__ cst, double* __restrict__ gh_u,
    double* __restrict__ gh_v, double* __restrict__ Ug, double* __restrict__ Vg) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const Nb q = neighbours(d, j, i);
  const long long n2 = d.n2;
  const int nz = d.nz;
  const bool split = p.scheme == SCHEME_SPLIT;
  // a face mask at level k is k < its threshold: the 2D face mask and the
  // shallower of the face's two cells
  const int nwc = geo.nw[q.c], nwe = geo.nw[q.e], nwn = geo.nw[q.n], nwne = geo.nw[q.ne];
  auto thr = [](uint8_t open, int a, int b) { return open ? (a < b ? a : b) : 0; };
  const int tu_c = thr(geo.mu2[q.c], nwc, nwe), tu_w = thr(geo.mu2[q.w], geo.nw[q.w], nwc);
  const int tu_n = thr(geo.mu2[q.n], nwn, nwne), tu_nw = thr(geo.mu2[q.nw], geo.nw[q.nw], nwn);
  const int tv_c = thr(geo.mv2[q.c], nwc, nwn), tv_s = thr(geo.mv2[q.s], geo.nw[q.s], nwc);
  const int tv_e = thr(geo.mv2[q.e], nwe, nwne), tv_se = thr(geo.mv2[q.se], geo.nw[q.se], nwe);
  const double gx2 = mk(geo.mu2[c]) * (eta[q.e] - eta[c]) / p.dx;
  const double gy2 = mk(geo.mv2[c]) * (eta[q.n] - eta[c]) / p.dy;
  const bool first = u_it == u;
  double du_prev = 0.0, dv_prev = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long o = k * n2;
    const long long t = o + c;
    // Coriolis partners of the old state, and of the iterate in a second pass
    const bool fv0 = k < tv_c, fv1 = k < tv_s, fv2 = k < tv_e, fv3 = k < tv_se;
    const bool fu0 = k < tu_c, fu1 = k < tu_w, fu2 = k < tu_n, fu3 = k < tu_nw;
    const double av_v = avg4(v[t], v[o + q.s], v[o + q.e], v[o + q.se], fv0, fv1, fv2, fv3, fu0);
    const double av_u = avg4(u[t], u[o + q.w], u[o + q.n], u[o + q.nw], fu0, fu1, fu2, fu3, fv0);
    const double av_v_it = first ? av_v
                                 : avg4(v_it[t], v_it[o + q.s], v_it[o + q.e], v_it[o + q.se],
                                        fv0, fv1, fv2, fv3, fu0);
    const double av_u_it = first ? av_u
                                 : avg4(u_it[t], u_it[o + q.w], u_it[o + q.n], u_it[o + q.nw],
                                        fu0, fu1, fu2, fu3, fv0);
    const double vcor = p.one_m_tc * av_v + p.theta_cor * av_v_it;
    const double ucor = p.one_m_tc * av_u + p.theta_cor * av_u_it;
    const double gxe = gx2 * (fu0 ? 1.0 : 0.0);
    const double gye = gy2 * (fv0 ? 1.0 : 0.0);
    // split_explicit leaves the surface gradient to the barotropic substeps
    // (model3d_v05.py:431-432)
    const double gu = split ? u[t] + p.dt * (p.f * vcor + pgx[t]) + p.c_dzd * dzu[t]
                            : u[t] + p.dt * (p.f * vcor + pgx[t]) - p.c_gx * gxe + p.c_dzd * dzu[t];
    const double gv = split ? v[t] + p.dt * (p.neg_f * ucor + pgy[t]) + p.c_dzd * dzv[t]
                            : v[t] + p.dt * (p.neg_f * ucor + pgy[t]) - p.c_gx * gye +
                                  p.c_dzd * dzv[t];
    const double ru = gu + stu[t];
    const double rv = gv + stv[t];
    double du, dv;
    if (k == 0) {
      du = ru / den[t];
      dv = rv / den[t];
    } else {
      du = (ru - sub[t] * du_prev) / den[t];
      dv = (rv - sub[t] * dv_prev) / den[t];
    }
    gh_u[t] = du;
    gh_v[t] = dv;
    du_prev = du;
    dv_prev = dv;
  }
  double xu = 0.0, xv = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long t = k * n2 + c;
    if (k == nz - 1) {
      xu = gh_u[t];
      xv = gh_v[t];
    } else {
      xu = gh_u[t] - cst[t] * xu;
      xv = gh_v[t] - cst[t] * xv;
    }
    const double mm = k < nwc ? 1.0 : 0.0;
    gh_u[t] = xu * mm;
    gh_v[t] = xv * mm;
  }
  const FaceGeo fu(geo, p, d, c, q.e, 0), fv(geo, p, d, c, q.n, 1);
  double su = 0.0, sv = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long t = k * n2 + c;
    const double tu = fu.dz(k) * gh_u[t];
    const double tv = fv.dz(k) * gh_v[t];
    su = (k == 0) ? tu : su + tu;
    sv = (k == 0) ? tv : sv + tv;
  }
  Ug[c] = su;
  Vg[c] = sv;
}

// u_it = (gh_u - g dt th q Gx(eta_new) mask3u) mask3 (model3d_v05.py:404-406)
__global__ void k_backsub(Dims d, StepParams p, Geo geo, const double* __restrict__ gh_u,
                          const double* __restrict__ gh_v, const double* __restrict__ q,
                          const double* __restrict__ eta_n, double* __restrict__ u_it,
                          double* __restrict__ v_it) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i);
  const double gx = mk(geo.mu2[c]) * (eta_n[ce] - eta_n[c]) / p.dx;
  const double gy = mk(geo.mv2[c]) * (eta_n[cn] - eta_n[c]) / p.dy;
  bool m3, m3u, m3v;
  if (p.zcol) {
    const int k = (int)(t / d.n2);
    const int nwc = geo.nw[c];
    m3 = k < nwc;
    m3u = geo.mu2[c] && m3 && k < geo.nw[ce];
    m3v = geo.mv2[c] && m3 && k < geo.nw[cn];
  } else {
    m3 = geo.m3[t] != 0;
    m3u = geo.m3u[t] != 0;
    m3v = geo.m3v[t] != 0;
  }
  const double mm = m3 ? 1.0 : 0.0;
  u_it[t] = (gh_u[t] - p.c_bs * q[t] * gx * (m3u ? 1.0 : 0.0)) * mm;
  v_it[t] = (gh_v[t] - p.c_bs * q[t] * gy * (m3v ? 1.0 : 0.0)) * mm;
}

// The last Picard pass's back-substitution fused with w of the new
// velocity (model3d_v05.py:404-406 then 232-235), one lane per column
// marching up from the bottom: u and v of the column are formed and stored
// as k_backsub forms them; the west u and the south v that the transport
// divergence needs are formed again from the same operands (the same bits),
// so w is k_w_diag's of the stored fields without reading them back.
__global__ void k_backsub_w(Dims d, StepParams p, Geo geo, const double* __restrict__ gh_u,
                            const double* __restrict__ gh_v, const double* __restrict__ q,
                            const double* __restrict__ eta_n, double* __restrict__ u,
                            double* __restrict__ v, double* __restrict__ w) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i), cs = at2(d, wrap(j, -1, d.ny), i);
  const long long n2 = d.n2;
  const int nz = d.nz;
  // surface-gradient factors of the column and of its west (u) and south (v)
  // neighbours, as k_backsub forms them
  const double gx_c = mk(geo.mu2[c]) * (eta_n[ce] - eta_n[c]) / p.dx;
  const double gx_w = mk(geo.mu2[cw]) * (eta_n[c] - eta_n[cw]) / p.dx;
  const double gy_c = mk(geo.mv2[c]) * (eta_n[cn] - eta_n[c]) / p.dy;
  const double gy_s = mk(geo.mv2[cs]) * (eta_n[c] - eta_n[cs]) / p.dy;
  const ColumnGeo gc(geo, p, d, c), gw(geo, p, d, cw), gs(geo, p, d, cs);
  const FaceGeo fe(geo, p, d, c, ce, 0), fw(geo, p, d, cw, c, 0);
  const FaceGeo fn(geo, p, d, c, cn, 1), fs(geo, p, d, cs, c, 1);
  double sum = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long o = (long long)k * n2;
    const double mc = gc.wet(k) ? 1.0 : 0.0;
    const double uc = (gh_u[o + c] - p.c_bs * q[o + c] * gx_c * (fe.wet(k) ? 1.0 : 0.0)) * mc;
    const double vc = (gh_v[o + c] - p.c_bs * q[o + c] * gy_c * (fn.wet(k) ? 1.0 : 0.0)) * mc;
    const double uw = (gh_u[o + cw] - p.c_bs * q[o + cw] * gx_w * (fw.wet(k) ? 1.0 : 0.0)) *
                      (gw.wet(k) ? 1.0 : 0.0);
    const double vs = (gh_v[o + cs] - p.c_bs * q[o + cs] * gy_s * (fs.wet(k) ? 1.0 : 0.0)) *
                      (gs.wet(k) ? 1.0 : 0.0);
    u[o + c] = uc;
    v[o + c] = vc;
    const double U = fe.dz(k) * uc;
    const double Uw = fw.dz(k) * uw;
    const double V = fn.dz(k) * vc;
    const double Vs = fs.dz(k) * vs;
    const double dv = mc * ((U - Uw) / p.dx + (V - Vs) / p.dy);
    sum = (k == nz - 1) ? dv : sum + dv;
    w[o + c] = -sum;
  }
  w[(long long)nz * n2 + c] = 0.0;
}

// ------------------------------------------------------------- tracers
// Flux-form tracer advection divergence at (k, j, i) with face thickness
// (libs/core/model3d_v05.py:482-512).
__device__ double tracer_advection(const Dims& d, const StepParams& p, const Geo& geo,
                                   const double* cc, const double* u, const double* v,
                                   const double* w, long long c, int k, const Nb& q) {
  const long long n2 = d.n2;
  const int nz = d.nz;
  const long long o = (long long)k * n2;
  const double* C = cc + o;
  const uint8_t* M = geo.m3 + o;
  const bool up3 = p.advection == ADV_UP3 || p.advection == ADV_UP3_TVD;
  const bool lim = p.advection == ADV_UP3_TVD;
  auto face = [&](double cm, double c0, double c1, double c2, double vel, bool wm,
                  bool w2) -> double {
    if (up3) return kappa_face(cm, c0, c1, c2, vel, wm, w2, lim);
    if (p.advection == ADV_CENTRED2) return 0.5 * (c0 + c1);
    return vel > 0.0 ? c0 : c1;
  };
  const double ue = u[o + q.c], uw = u[o + q.w];
  const double vn = v[o + q.c], vs = v[o + q.s];
  const double cx = face(C[q.w], C[q.c], C[q.e], C[q.ee], ue, M[q.w] != 0, M[q.ee] != 0);
  const double cxw = face(C[q.ww], C[q.w], C[q.c], C[q.e], uw, M[q.ww] != 0, M[q.e] != 0);
  const double cy = face(C[q.s], C[q.c], C[q.n], C[q.nn], vn, M[q.s] != 0, M[q.nn] != 0);
  const double cys = face(C[q.ss], C[q.s], C[q.c], C[q.n], vs, M[q.ss] != 0, M[q.n] != 0);
  const double fx = geo.dz3u[o + q.c] * ue * cx;
  const double fxw = geo.dz3u[o + q.w] * uw * cxw;
  const double fy = geo.dz3v[o + q.c] * vn * cy;
  const double fys = geo.dz3v[o + q.s] * vs * cys;
  const double inv = safe_inv(geo.dz3[o + c]);
  const double hdiv = ((fx - fxw) / p.dx + (fy - fys) / p.dy) * inv;
  auto cf = [&](int kk) -> double {  // interface value, kk in 0..nz
    if (kk == 0) return cc[c];
    if (kk == nz) return 0.0;
    const long long a = (long long)kk * n2;
    const double above = cc[a - n2 + c], below = cc[a + c];
    if (!up3) return 0.5 * (above + below);
    const double far_up = kk >= 2 ? cc[a - 2 * n2 + c] : 0.0;
    const double far_dn = kk <= nz - 2 ? cc[a + n2 + c] : 0.0;
    const bool hu = kk >= 2 && geo.m3[a - 2 * n2 + c] != 0;
    const bool hd = kk <= nz - 2 && geo.m3[a + n2 + c] != 0;
    return kappa_face_z(far_up, above, below, far_dn, w[a + c], hu, hd, lim);
  };
  const double vdiv = (w[o + c] * cf(k) - w[o + n2 + c] * cf(k + 1)) * inv;
  return (hdiv + vdiv) * mk(M[q.c]);
}

// ---------------------------------------------- standalone operator kernels
// Batched tridiagonal solve, one lane per column (libs/core/vertical.py:17-44);
// x holds d* between the sweeps, cs holds c*.
__global__ void k_thomas(long long ncol, int n, const double* __restrict__ sub,
                         const double* __restrict__ diag, const double* __restrict__ sup,
                         const double* __restrict__ rhs, double* __restrict__ x,
                         double* __restrict__ cs) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double cp = 0.0, dp = 0.0;
  for (int k = 0; k < n; ++k) {
    const long long t = k * ncol + c;
    double cc, dd;
    if (k == 0) {
      cc = sup[t] / diag[t];
      dd = rhs[t] / diag[t];
    } else {
      const double den = diag[t] - sub[t] * cp;
      cc = sup[t] / den;
      dd = (rhs[t] - sub[t] * dp) / den;
    }
    cs[t] = cc;
    x[t] = dd;
    cp = cc;
    dp = dd;
  }
  double xn = x[(long long)(n - 1) * ncol + c];
  for (int k = n - 2; k >= 0; --k) {
    const long long t = k * ncol + c;
    xn = x[t] - cs[t] * xn;
    x[t] = xn;
  }
}

// Buoyancy from T, S at depth z for n points (libs/core/eos.py:130-133).
__global__ void k_eos(long long n, StepParams p, const double* __restrict__ T,
                      const double* __restrict__ S, const double* __restrict__ z,
                      double* __restrict__ b) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  b[t] = buoyancy(p, T[t], S[t], z[t]);
}

// Tracer advection divergence over the whole 3D field (the operator alone).
__global__ void k_tracer_advection(Dims d, StepParams p, Geo geo, const double* __restrict__ cc,
                                   const double* __restrict__ u, const double* __restrict__ v,
                                   const double* __restrict__ w, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2;
  const int k = (int)(t / d.n2);
  const Nb q = neighbours(d, (int)(c / d.nx), (int)(c % d.nx));
  out[t] = tracer_advection(d, p, geo, cc, u, v, w, c, k, q);
}

// Momentum advection (u, v) over the whole field (the operator alone).
__global__ void k_momentum_advection(Dims d, StepParams p, Geo geo, const double* __restrict__ u,
                                     const double* __restrict__ v, const double* __restrict__ w,
                                     double* __restrict__ au, double* __restrict__ av) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2;
  const int k = (int)(t / d.n2);
  const Nb q = neighbours(d, (int)(c / d.nx), (int)(c % d.nx));
  const long long o = (long long)k * d.n2;
  au[t] = advect_u(d, p, geo.m3u, u, v, w, o, k, q);
  av[t] = advect_v(d, p, geo.m3v, u, v, w, o, k, q);
}

}  // namespace hoc
