// Hand-written adjoints (reverse-mode derivatives) of two operators of the
// step, for the differentiable step and its checkpointed gradient:
//   k_thomas_adjoint            the batched tridiagonal solve x = A^-1 r:
//                               lambda = A^-T g (a Thomas sweep on the
//                               transposed bands), r_bar = lambda and the
//                               band gradients -lambda x^T restricted to the
//                               three bands;
//   k_tracer_advection_adjoint  the flux-form kappa = 1/3 (optionally
//                               Superbee-limited) tracer advection of
//                               libs/core/model3d_v05.py:482-512, gathered per
//                               cell: gradients with respect to the tracer,
//                               the face velocities u, v and the interface w.
// The upwind selections (sign of the face velocity, wet far cell) and the
// limiter branch are held at their forward values; at a tie the branch the
// forward takes (NumPy's maximum/minimum keep the first operand) is the one
// differentiated. The face velocity enters only through the selection, so
// its derivative is the flux's linear factor.
#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

__global__ void k_thomas_adjoint(long long ncol, int n, const double* __restrict__ sub,
                                 const double* __restrict__ diag, const double* __restrict__ sup,
                                 const double* __restrict__ x, const double* __restrict__ g,
                                 double* __restrict__ lam, double* __restrict__ cs,
                                 double* __restrict__ d_sub, double* __restrict__ d_diag,
                                 double* __restrict__ d_sup) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  // A^T: sub_T[k] = sup[k-1], sup_T[k] = sub[k+1], same diagonal
  double cp = 0.0, dp = 0.0;
  for (int k = 0; k < n; ++k) {
    const long long t = k * ncol + c;
    const double supT = (k + 1 < n) ? sub[t + ncol] : 0.0;
    double cc, dd;
    if (k == 0) {
      cc = supT / diag[t];
      dd = g[t] / diag[t];
    } else {
      const double subT = sup[t - ncol];
      const double den = diag[t] - subT * cp;
      cc = supT / den;
      dd = (g[t] - subT * dp) / den;
    }
    cs[t] = cc;
    lam[t] = dd;
    cp = cc;
    dp = dd;
  }
  double ln = lam[(long long)(n - 1) * ncol + c];
  for (int k = n - 2; k >= 0; --k) {
    const long long t = k * ncol + c;
    ln = lam[t] - cs[t] * ln;
    lam[t] = ln;
  }
  // band gradients: dL/dA_kj = -lambda_k x_j
  double x_prev = 0.0, x_k = x[c];
  for (int k = 0; k < n; ++k) {
    const long long t = k * ncol + c;
    const double x_next = (k + 1 < n) ? x[t + ncol] : 0.0;
    const double l = lam[t];
    d_sub[t] = k > 0 ? -l * x_prev : 0.0;
    d_diag[t] = -l * x_k;
    d_sup[t] = (k + 1 < n) ? -l * x_next : 0.0;
    x_prev = x_k;
    x_k = x_next;
  }
}

// Superbee phi(r) and dphi/dr on the branch the forward selects
// (superbee in hoc_device.h).
__device__ __forceinline__ void superbee_d(double r, double& phi, double& dphi) {
  const bool a_lin = 2.0 * r <= 1.0;           // nmin(2r, 1) keeps 2r
  const double a = a_lin ? 2.0 * r : 1.0;
  const bool b_lin = r <= 2.0;                 // nmin(r, 2) keeps r
  const double bb = b_lin ? r : 2.0;
  const bool take_a = a >= bb;                 // nmax(a, bb) keeps a
  const double m = take_a ? a : bb;
  const double dm = take_a ? (a_lin ? 2.0 : 0.0) : (b_lin ? 1.0 : 0.0);
  if (0.0 >= m) {                              // nmax(0, m) keeps 0
    phi = 0.0;
    dphi = 0.0;
  } else {
    phi = m;
    dphi = dm;
  }
}

// Partial derivatives of kappa_face(cm, c0, c1, c2, vel, wm, w2, limited)
// with respect to (cm, c0, c1, c2), and its value.
__device__ __forceinline__ double kappa_face_d(double cm, double c0, double c1, double c2,
                                               double vel, bool wm, bool w2, bool limited,
                                               double dp[4]) {
  dp[0] = dp[1] = dp[2] = dp[3] = 0.0;
  const bool up = vel > 0.0;
  if (!(up ? wm : w2)) {
    if (up) dp[1] = 1.0;
    else dp[2] = 1.0;
    return up ? c0 : c1;
  }
  if (!limited) {
    const double a = 0.25 * K1M, b = 0.25 * K1P;
    if (up) {
      dp[0] = -a;
      dp[1] = 1.0 + a - b;
      dp[2] = b;
    } else {
      dp[3] = -a;
      dp[2] = 1.0 + a - b;
      dp[1] = b;
    }
    return kappa_face(cm, c0, c1, c2, vel, wm, w2, false);
  }
  const double dc = c1 - c0;
  const double num = up ? (c0 - cm) : (c2 - c1);
  double r = 0.0, phi = 0.0, dphi = 0.0;
  if (dc != 0.0) r = num / dc;
  superbee_d(r, phi, dphi);
  // t = phi(r) dc / 2, r = num / dc:  dt = (dphi dnum + (phi - dphi r) ddc) / 2
  const double a = 0.5 * dphi, b = 0.5 * (phi - dphi * r);
  if (up) {  // f = c0 + t
    dp[0] = -a;
    dp[1] = 1.0 + a - b;
    dp[2] = b;
  } else {   // f = c1 - t
    dp[3] = -a;
    dp[2] = 1.0 + a - b;
    dp[1] = b;
  }
  return kappa_face(cm, c0, c1, c2, vel, wm, w2, true);
}

__device__ __forceinline__ int pwrap(int a, int n) {
  const int r = a % n;
  return r < 0 ? r + n : r;
}

// Gradients of L = sum(G * adv(c, u, v, w)) for one cell (k, j, i): gc at
// the cell, gu / gv at its east / north face, gw at its top interface (and
// the bottom interface nz, which carries no flux, for k = nz - 1).
__global__ void k_tracer_advection_adjoint(Dims d, StepParams p, Geo geo,
                                           const double* __restrict__ cc,
                                           const double* __restrict__ u,
                                           const double* __restrict__ v,
                                           const double* __restrict__ w,
                                           const double* __restrict__ G,
                                           double* __restrict__ gc, double* __restrict__ gu,
                                           double* __restrict__ gv, double* __restrict__ gw) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long n2 = d.n2;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int k = (int)(t / n2);
  const long long c = t % n2;
  const int j = (int)(c / nx), i = (int)(c % nx);
  const long long o = (long long)k * n2;
  const bool lim = p.advection == ADV_UP3_TVD;
  auto col = [&](int jj, int ii) -> long long { return (long long)pwrap(jj, ny) * nx + pwrap(ii, nx); };
  // Ghat = G mask / dz at a cell of this level
  auto ghat = [&](long long cell, long long off) -> double {
    const long long q = off + cell;
    return G[q] * mk(geo.m3[q]) * safe_inv(geo.dz3[q]);
  };
  const double* C = cc + o;
  const uint8_t* M = geo.m3 + o;
  double acc = 0.0;
  // ---- x-faces: face f (east face of column f) reads columns f-1..f+2
  for (int s = -2; s <= 1; ++s) {  // f = i + s; this cell plays role 1 - s
    const long long f = col(j, i + s);
    const long long fw = col(j, i + s - 1), fe = col(j, i + s + 1), fee = col(j, i + s + 2);
    const double vel = u[o + f];
    double dp[4];
    const double face = kappa_face_d(C[fw], C[f], C[fe], C[fee], vel, M[fw] != 0, M[fee] != 0,
                                     lim, dp);
    const double fbar = (ghat(f, o) - ghat(fe, o)) / p.dx;
    acc += fbar * geo.dz3u[o + f] * vel * dp[1 - s];
    if (s == 0) gu[t] = fbar * geo.dz3u[o + f] * face;
  }
  // ---- y-faces
  for (int s = -2; s <= 1; ++s) {
    const long long f = col(j + s, i);
    const long long fs = col(j + s - 1, i), fn = col(j + s + 1, i), fnn = col(j + s + 2, i);
    const double vel = v[o + f];
    double dp[4];
    const double face = kappa_face_d(C[fs], C[f], C[fn], C[fnn], vel, M[fs] != 0, M[fnn] != 0,
                                     lim, dp);
    const double fbar = (ghat(f, o) - ghat(fn, o)) / p.dy;
    acc += fbar * geo.dz3v[o + f] * vel * dp[1 - s];
    if (s == 0) gv[t] = fbar * geo.dz3v[o + f] * face;
  }
  // ---- vertical: interface kk (1..nz-1) reads layers kk-2..kk+1 as
  // kappa_face(cm = far_dn, c0 = below, c1 = above, c2 = far_up)
  auto zbar = [&](int kk) -> double {  // adjoint of the flux w cf at interface kk
    const double below = ghat(c, (long long)kk * n2);
    const double above = kk >= 1 ? ghat(c, (long long)(kk - 1) * n2) : 0.0;
    return below - above;
  };
  for (int kk = k - 1; kk <= k + 2; ++kk) {
    if (kk < 1 || kk > nz - 1) continue;
    const long long a = (long long)kk * n2 + c;
    const double far_up = kk >= 2 ? cc[a - 2 * n2] : 0.0;
    const double far_dn = kk <= nz - 2 ? cc[a + n2] : 0.0;
    const bool hu = kk >= 2 && geo.m3[a - 2 * n2] != 0;
    const bool hd = kk <= nz - 2 && geo.m3[a + n2] != 0;
    double dp[4];
    const double face = kappa_face_d(far_dn, cc[a], cc[a - n2], far_up, w[a], hd, hu, lim, dp);
    const double wb = zbar(kk) * w[a];
    // this layer is far_dn (kk = k - 1), below (kk = k), above (k + 1), far_up (k + 2)
    acc += wb * dp[kk == k - 1 ? 0 : kk == k ? 1 : kk == k + 1 ? 2 : 3];
    if (kk == k) gw[t] = zbar(kk) * face;
  }
  if (k == 0) {  // the surface interface carries w[0] c[0]
    const double gz0 = ghat(c, 0);
    acc += gz0 * w[c];
    gw[c] = gz0 * cc[c];
  }
  if (k == nz - 1) gw[(long long)nz * n2 + c] = 0.0;
  gc[t] = acc;
}

}  // namespace hoc

namespace hoc {

// ------------------------------------------------ Gaspar mixing lengths
// The budget march of libs/core/closure.py:65-103 for one interface and one
// direction, every crossing evaluated (no early exit), in the arithmetic of
// the march: cost = N2 dz (l + dz/2); a full crossing adds dz and spends
// cost; otherwise the parcel stops inside the layer (l grows by
// clip(sqrt(l^2 + 2 budget / N2) - l, 0, dz)) and the budget is spent.
struct MarchIn {
  double l, b;
};

__device__ __forceinline__ MarchIn march_step(double l, double b, double n2j, double dzj) {
  const double cost = n2j * dzj * (l + 0.5 * dzj);
  const bool full = cost <= b;
  const double a = n2j > 0.0 ? n2j : 1.0;
  const double A = nmax(l * l + 2.0 * b / a, 0.0);
  const double sq = A > 0.0 ? sqrt(A) : 0.0;
  const double s = sq - l;
  const double clipped = nmin(nmax(s, 0.0), dzj);
  const double step = full ? dzj : (n2j > 0.0 ? clipped : dzj);
  return MarchIn{l + (b > 0.0 ? step : 0.0), full ? b - cost : 0.0};
}

// Reverse of march_step: given the adjoints of (l', b') return those of
// (l, b) and add the one of n2j. Branches held at their forward values;
// at a tie of a maximum/minimum the first operand's branch (NumPy's choice)
// is differentiated.
__device__ __forceinline__ void march_step_adjoint(double l, double b, double n2j, double dzj,
                                                   double lb_out, double bb_out, double& lb,
                                                   double& bb, double& n2b) {
  const double cost = n2j * dzj * (l + 0.5 * dzj);
  const bool full = cost <= b;
  lb = lb_out;
  bb = 0.0;
  if (full) {  // b' = b - cost
    bb += bb_out;
    lb += -bb_out * n2j * dzj;
    n2b += -bb_out * dzj * (l + 0.5 * dzj);
  }
  if (b > 0.0 && !full && n2j > 0.0) {  // l' = l + clip(s, 0, dz)
    const double a = n2j;
    const double A0 = l * l + 2.0 * b / a;
    const bool a_pass = A0 >= 0.0 || A0 != A0;  // nmax(A0, 0) keeps A0
    const double A = a_pass ? A0 : 0.0;
    const double sq = A > 0.0 ? sqrt(A) : 0.0;
    const double s = sq - l;
    const bool s_pass = s >= 0.0 || s != s;     // nmax(s, 0) keeps s
    const double m = s_pass ? s : 0.0;
    const bool m_pass = m <= dzj || m != m;     // nmin(m, dz) keeps m
    if (s_pass && m_pass) {
      const double dA = (a_pass && A > 0.0) ? lb_out / (2.0 * sq) : 0.0;  // adjoint of A0
      lb += dA * 2.0 * l - lb_out;
      bb += dA * 2.0 / a;
      n2b += -dA * 2.0 * b / (a * a);
    }
  }
}

// l_up, l_dn on every interface of a column (layout [nz1][ncol]): the
// march above from e_k = max(e, e_min), then min(l, z + z_0) * wet.
__global__ void k_gaspar_lengths(long long ncol, int nz1, double e_min, double z_0,
                                 const double* __restrict__ e, const double* __restrict__ n2,
                                 const double* __restrict__ dzi, const double* __restrict__ zs,
                                 const double* __restrict__ zb, const double* __restrict__ wet,
                                 double* __restrict__ l_up, double* __restrict__ l_dn) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  for (int k = 0; k < nz1; ++k) {
    const long long t = (long long)k * ncol + c;
    const double ek = nmax(e[t], e_min);
    MarchIn u{0.0, ek}, dn{0.0, ek};
    for (int j = 1; j <= k; ++j) {
      const long long q = (long long)(k - j) * ncol + c;
      u = march_step(u.l, u.b, nmax(n2[q], 0.0), dzi[q]);
    }
    for (int j = 1; k + j < nz1; ++j) {
      const long long q = (long long)(k + j) * ncol + c;
      dn = march_step(dn.l, dn.b, nmax(n2[q], 0.0), dzi[q]);
    }
    l_up[t] = nmin(u.l, zs[t] + z_0) * wet[t];
    l_dn[t] = nmin(dn.l, zb[t] + z_0) * wet[t];
  }
}

constexpr int GASPAR_MAX_NZ1 = 128;

// Adjoint of k_gaspar_lengths: g_e and g_n2 (overwritten) from the adjoints
// of l_up and l_dn. Each march is re-run forward with its states kept in a
// per-thread array, then reversed.
__global__ void k_gaspar_lengths_adjoint(long long ncol, int nz1, double e_min, double z_0,
                                         const double* __restrict__ e,
                                         const double* __restrict__ n2,
                                         const double* __restrict__ dzi,
                                         const double* __restrict__ zs,
                                         const double* __restrict__ zb,
                                         const double* __restrict__ wet,
                                         const double* __restrict__ g_up,
                                         const double* __restrict__ g_dn,
                                         double* __restrict__ g_e, double* __restrict__ g_n2) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double n2b[GASPAR_MAX_NZ1];  // adjoint of max(n2, 0) per interface
  double ls[GASPAR_MAX_NZ1], bs[GASPAR_MAX_NZ1];
  for (int k = 0; k < nz1; ++k) n2b[k] = 0.0;
  for (int k = 0; k < nz1; ++k) {
    const long long t = (long long)k * ncol + c;
    const double ek = nmax(e[t], e_min);
    double bbar0 = 0.0;
    for (int dir = 0; dir < 2; ++dir) {
      const int nstep = dir == 0 ? k : nz1 - 1 - k;
      // forward, keeping the state before every crossing
      double l = 0.0, b = ek;
      for (int j = 1; j <= nstep; ++j) {
        const int kk = dir == 0 ? k - j : k + j;
        const long long q = (long long)kk * ncol + c;
        ls[j] = l;
        bs[j] = b;
        const MarchIn o = march_step(l, b, nmax(n2[q], 0.0), dzi[q]);
        l = o.l;
        b = o.b;
      }
      // l_out = min(l, z + z_0) * wet
      const double cap = (dir == 0 ? zs[t] : zb[t]) + z_0;
      const double gout = (dir == 0 ? g_up[t] : g_dn[t]) * wet[t];
      double lb = (l <= cap || l != l) ? gout : 0.0;
      double bb = 0.0;
      for (int j = nstep; j >= 1; --j) {
        const int kk = dir == 0 ? k - j : k + j;
        const long long q = (long long)kk * ncol + c;
        double lb_in, bb_in, nb = 0.0;
        march_step_adjoint(ls[j], bs[j], nmax(n2[q], 0.0), dzi[q], lb, bb, lb_in, bb_in, nb);
        lb = lb_in;
        bb = bb_in;
        n2b[kk] += nb;
      }
      bbar0 += bb;  // the march starts from budget e_k (l starts at the constant 0)
    }
    g_e[t] = (e[t] >= e_min || e[t] != e[t]) ? bbar0 : 0.0;
  }
  for (int k = 0; k < nz1; ++k) {
    const long long t = (long long)k * ncol + c;
    g_n2[t] = (n2[t] >= 0.0 || n2[t] != n2[t]) ? n2b[k] : 0.0;
  }
}

}  // namespace hoc

namespace hoc {

// ------------------------------------------------------------ EOS adjoint
// Forward-mode dual number (value, derivative) for the templated EOS
// (buoyancy_t, rho_teos10 in hoc_device.h): one evaluation per seeded input
// gives the exact derivative of the evaluated expression.
struct Dual {
  double v, d;
  __device__ Dual() : v(0.0), d(0.0) {}
  __device__ Dual(double x) : v(x), d(0.0) {}
  __device__ Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual sqrt(Dual a) {
  const double r = ::sqrt(a.v);
  return Dual(r, r > 0.0 ? a.d / (2.0 * r) : 0.0);
}

// dL/dT, dL/dS of L = sum(g_b * b(T, S, z)).
__global__ void k_eos_adjoint(long long n, StepParams p, const double* __restrict__ T,
                              const double* __restrict__ S, const double* __restrict__ z,
                              const double* __restrict__ g_b, double* __restrict__ g_T,
                              double* __restrict__ g_S) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const Dual zz(z[t]);
  const Dual dT = buoyancy_t<Dual>(p, Dual(T[t], 1.0), Dual(S[t]), zz);
  const Dual dS = buoyancy_t<Dual>(p, Dual(T[t]), Dual(S[t], 1.0), zz);
  g_T[t] = g_b[t] * dT.d;
  g_S[t] = g_b[t] * dS.d;
}

}  // namespace hoc

namespace hoc {

// ------------------------------------------------ momentum advection adjoint
// Gradients of L = sum(Gu * au + Gv * av) for the kappa = 1/3 momentum
// advection of libs/core/advection_v06.py:77-108 (the operator of
// hoc_stepper_momentum_advection, before masking), gathered at one point
// (k, j, i): gu, gv at its u and v points and gw at interface k. For the
// u equation the x fluxes live at cell centres c (Fx = ub uc, ub the mean of
// u at c-1 and c, uc the face value from u[c-2..c+1]), the y fluxes at
// corners (Fy = vb uy, vb the mean of v at i and i+1, uy from u[j-1..j+2])
// and the vertical fluxes at interfaces (Fz = wb uz, wb the mean of w at i
// and i+1); the v equation is the transpose of that layout. Every flux
// enters its two neighbours' divergences with opposite signs, which gives
// the flux adjoints (Fbar) the gather reads.
__global__ void k_momentum_advection_adjoint(Dims d, StepParams p, Geo geo,
                                             const double* __restrict__ u,
                                             const double* __restrict__ v,
                                             const double* __restrict__ w,
                                             const double* __restrict__ Gu,
                                             const double* __restrict__ Gv,
                                             double* __restrict__ gu, double* __restrict__ gv,
                                             double* __restrict__ gw) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long n2 = d.n2;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int k = (int)(t / n2);
  const long long c = t % n2;
  const int j = (int)(c / nx), i = (int)(c % nx);
  const long long o = (long long)k * n2;
  const bool lim = p.advection == ADV_UP3_TVD;
  auto at = [&](int jj, int ii) -> long long {
    return (long long)pwrap(jj, ny) * nx + pwrap(ii, nx);
  };
  auto U = [&](int jj, int ii) { return u[o + at(jj, ii)]; };
  auto V = [&](int jj, int ii) { return v[o + at(jj, ii)]; };
  auto MU = [&](int jj, int ii) { return geo.m3u[o + at(jj, ii)] != 0; };
  auto MV = [&](int jj, int ii) { return geo.m3v[o + at(jj, ii)] != 0; };
  auto GU = [&](int jj, int ii) { return Gu[o + at(jj, ii)]; };
  auto GV = [&](int jj, int ii) { return Gv[o + at(jj, ii)]; };
  double gU = 0.0, gV = 0.0, dp[4];
  // ---------------- u equation
  for (int s = -1; s <= 2; ++s) {  // x flux at centre cc = i + s
    const int cc = i + s;
    const double ub = 0.5 * (U(j, cc - 1) + U(j, cc));
    const double uc = kappa_face_d(U(j, cc - 2), U(j, cc - 1), U(j, cc), U(j, cc + 1), ub,
                                   MU(j, cc - 2), MU(j, cc + 1), lim, dp);
    const double fbar = (GU(j, cc - 1) - GU(j, cc)) / p.dx;
    gU += fbar * ub * dp[2 - s];
    if (s == 0 || s == 1) gU += fbar * 0.5 * uc;
  }
  for (int s = -2; s <= 1; ++s) {  // y flux at corner jc = j + s
    const int jc = j + s;
    const double vb = 0.5 * (V(jc, i) + V(jc, i + 1));
    kappa_face_d(U(jc - 1, i), U(jc, i), U(jc + 1, i), U(jc + 2, i), vb, MU(jc - 1, i),
                 MU(jc + 2, i), lim, dp);
    const double fbar = (GU(jc, i) - GU(jc + 1, i)) / p.dy;
    gU += fbar * vb * dp[1 - s];
  }
  for (int s = -1; s <= 0; ++s) {  // v enters the corners (j, i - 1) and (j, i)
    const int ii = i + s;
    const double vb = 0.5 * (V(j, ii) + V(j, ii + 1));
    const double uy = kappa_face_d(U(j - 1, ii), U(j, ii), U(j + 1, ii), U(j + 2, ii), vb,
                                   MU(j - 1, ii), MU(j + 2, ii), lim, dp);
    gV += (GU(j, ii) - GU(j + 1, ii)) / p.dy * uy * 0.5;
  }
  // ---------------- v equation
  for (int s = -1; s <= 2; ++s) {  // y flux at centre cc = j + s
    const int cc = j + s;
    const double vb = 0.5 * (V(cc - 1, i) + V(cc, i));
    const double vc = kappa_face_d(V(cc - 2, i), V(cc - 1, i), V(cc, i), V(cc + 1, i), vb,
                                   MV(cc - 2, i), MV(cc + 1, i), lim, dp);
    const double fbar = (GV(cc - 1, i) - GV(cc, i)) / p.dy;
    gV += fbar * vb * dp[2 - s];
    if (s == 0 || s == 1) gV += fbar * 0.5 * vc;
  }
  for (int s = -2; s <= 1; ++s) {  // x flux at corner ic = i + s
    const int ic = i + s;
    const double ub = 0.5 * (U(j, ic) + U(j + 1, ic));
    kappa_face_d(V(j, ic - 1), V(j, ic), V(j, ic + 1), V(j, ic + 2), ub, MV(j, ic - 1),
                 MV(j, ic + 2), lim, dp);
    const double fbar = (GV(j, ic) - GV(j, ic + 1)) / p.dx;
    gV += fbar * ub * dp[1 - s];
  }
  for (int s = -1; s <= 0; ++s) {  // u enters the corners (j - 1, i) and (j, i)
    const int jj = j + s;
    const double ub = 0.5 * (U(jj, i) + U(jj + 1, i));
    const double vx = kappa_face_d(V(jj, i - 1), V(jj, i), V(jj, i + 1), V(jj, i + 2), ub,
                                   MV(jj, i - 1), MV(jj, i + 2), lim, dp);
    gU += (GV(jj, i) - GV(jj, i + 1)) / p.dx * vx * 0.5;
  }
  // ---------------- vertical fluxes (interfaces 1..nz-1 carry them)
  // face value and partials at interface kk of column col for field f
  auto zface = [&](const double* f, const uint8_t* m3, int kk, long long col, double wb,
                   double* pd) -> double {
    const long long a = (long long)kk * n2 + col;
    const double far_up = kk >= 2 ? f[a - 2 * n2] : 0.0;
    const double far_dn = kk <= nz - 2 ? f[a + n2] : 0.0;
    const bool hu = kk >= 2 && m3[a - 2 * n2] != 0;
    const bool hd = kk <= nz - 2 && m3[a + n2] != 0;
    return kappa_face_d(far_dn, f[a], f[a - n2], far_up, wb, hd, hu, lim, pd);
  };
  const long long ce = at(j, i + 1), cn = at(j + 1, i);
  for (int kk = k - 1; kk <= k + 2; ++kk) {
    if (kk < 1 || kk > nz - 1) continue;
    const long long a = (long long)kk * n2;
    const int role = kk == k - 1 ? 0 : kk == k ? 1 : kk == k + 1 ? 2 : 3;
    const double wbu = 0.5 * (w[a + c] + w[a + ce]);
    zface(u, geo.m3u, kk, c, wbu, dp);
    gU += (Gu[a + c] - Gu[a - n2 + c]) / p.dz_nom * wbu * dp[role];
    const double wbv = 0.5 * (w[a + c] + w[a + cn]);
    zface(v, geo.m3v, kk, c, wbv, dp);
    gV += (Gv[a + c] - Gv[a - n2 + c]) / p.dz_nom * wbv * dp[role];
  }
  double gW = 0.0;
  if (k >= 1) {  // interface k: w enters wb of the u columns i-1, i and v rows j-1, j
    const long long a = (long long)k * n2;
    for (int s = -1; s <= 0; ++s) {
      const long long cu = at(j, i + s), cue = at(j, i + s + 1);
      const double wbu = 0.5 * (w[a + cu] + w[a + cue]);
      const double fu = zface(u, geo.m3u, k, cu, wbu, dp);
      gW += (Gu[a + cu] - Gu[a - n2 + cu]) / p.dz_nom * fu * 0.5;
      const long long cv = at(j + s, i), cvn = at(j + s + 1, i);
      const double wbv = 0.5 * (w[a + cv] + w[a + cvn]);
      const double fv = zface(v, geo.m3v, k, cv, wbv, dp);
      gW += (Gv[a + cv] - Gv[a - n2 + cv]) / p.dz_nom * fv * 0.5;
    }
  }
  gu[t] = gU;
  gv[t] = gV;
  gw[t] = gW;
  if (k == nz - 1) gw[(long long)nz * n2 + c] = 0.0;
}

}  // namespace hoc
