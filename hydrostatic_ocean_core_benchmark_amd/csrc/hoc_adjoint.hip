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
// G(q) returns the adjoint of the advection term at layer index q. Results:
// gc_out at the cell, gu_out / gv_out at its east / north face, gw_out at its
// top interface k.
template <class GA>
__device__ __forceinline__ void tracer_adv_adjoint_point(const Dims& d, const StepParams& p,
                                                         const Geo& geo,
                                                         const double* __restrict__ cc,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ w, GA G,
                                                         long long t, double& gc_out,
                                                         double& gu_out, double& gv_out,
                                                         double& gw_out) {
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
    return G(q) * mk(geo.m3[q]) * safe_inv(geo.dz3[q]);
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
    if (s == 0) gu_out = fbar * geo.dz3u[o + f] * face;
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
    if (s == 0) gv_out = fbar * geo.dz3v[o + f] * face;
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
    if (kk == k) gw_out = zbar(kk) * face;
  }
  if (k == 0) {  // the surface interface carries w[0] c[0]
    const double gz0 = ghat(c, 0);
    acc += gz0 * w[c];
    gw_out = gz0 * cc[c];
  }
  gc_out = acc;
}

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
  double a = 0.0, b = 0.0, e = 0.0, f = 0.0;
  tracer_adv_adjoint_point(d, p, geo, cc, u, v, w, [&](long long q) { return G[q]; }, t, a, b,
                           e, f);
  gc[t] = a;
  gu[t] = b;
  gv[t] = e;
  gw[t] = f;
  if (t / d.n2 == d.nz - 1) gw[(long long)d.nz * d.n2 + t % d.n2] = 0.0;
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

// Two derivatives at once (with respect to T and S): one evaluation of the
// EOS expression gives both partials.
struct Dual2 {
  double v, a, b;
  __device__ Dual2() : v(0.0), a(0.0), b(0.0) {}
  __device__ Dual2(double x) : v(x), a(0.0), b(0.0) {}
  __device__ Dual2(double x, double da, double db) : v(x), a(da), b(db) {}
};
__device__ __forceinline__ Dual2 operator+(Dual2 x, Dual2 y) {
  return Dual2(x.v + y.v, x.a + y.a, x.b + y.b);
}
__device__ __forceinline__ Dual2 operator-(Dual2 x, Dual2 y) {
  return Dual2(x.v - y.v, x.a - y.a, x.b - y.b);
}
__device__ __forceinline__ Dual2 operator-(Dual2 x) { return Dual2(-x.v, -x.a, -x.b); }
__device__ __forceinline__ Dual2 operator*(Dual2 x, Dual2 y) {
  return Dual2(x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b);
}
__device__ __forceinline__ Dual2 operator/(Dual2 x, Dual2 y) {
  const double q = x.v / y.v;
  return Dual2(q, (x.a - q * y.a) / y.v, (x.b - q * y.b) / y.v);
}
__device__ __forceinline__ Dual2 sqrt(Dual2 x) {
  const double r = ::sqrt(x.v);
  return Dual2(r, r > 0.0 ? x.a / (2.0 * r) : 0.0, r > 0.0 ? x.b / (2.0 * r) : 0.0);
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
// Gu(q), Gv(q) return the adjoints of the two advection terms at layer
// index q; results gU, gV at the point and gW at its top interface.
template <class GUA, class GVA>
__device__ __forceinline__ void momentum_adv_adjoint_point(const Dims& d, const StepParams& p,
                                                           const Geo& geo,
                                                           const double* __restrict__ u,
                                                           const double* __restrict__ v,
                                                           const double* __restrict__ w,
                                                           GUA Gu, GVA Gv, long long t,
                                                           double& gU_out, double& gV_out,
                                                           double& gW_out) {
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
  auto GU = [&](int jj, int ii) { return Gu(o + at(jj, ii)); };
  auto GV = [&](int jj, int ii) { return Gv(o + at(jj, ii)); };
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
    gU += (Gu(a + c) - Gu(a - n2 + c)) / p.dz_nom * wbu * dp[role];
    const double wbv = 0.5 * (w[a + c] + w[a + cn]);
    zface(v, geo.m3v, kk, c, wbv, dp);
    gV += (Gv(a + c) - Gv(a - n2 + c)) / p.dz_nom * wbv * dp[role];
  }
  double gW = 0.0;
  if (k >= 1) {  // interface k: w enters wb of the u columns i-1, i and v rows j-1, j
    const long long a = (long long)k * n2;
    for (int s = -1; s <= 0; ++s) {
      const long long cu = at(j, i + s), cue = at(j, i + s + 1);
      const double wbu = 0.5 * (w[a + cu] + w[a + cue]);
      const double fu = zface(u, geo.m3u, k, cu, wbu, dp);
      gW += (Gu(a + cu) - Gu(a - n2 + cu)) / p.dz_nom * fu * 0.5;
      const long long cv = at(j + s, i), cvn = at(j + s + 1, i);
      const double wbv = 0.5 * (w[a + cv] + w[a + cvn]);
      const double fv = zface(v, geo.m3v, k, cv, wbv, dp);
      gW += (Gv(a + cv) - Gv(a - n2 + cv)) / p.dz_nom * fv * 0.5;
    }
  }
  gU_out = gU;
  gV_out = gV;
  gW_out = gW;
}

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
  double a = 0.0, b = 0.0, c = 0.0;
  momentum_adv_adjoint_point(d, p, geo, u, v, w, [&](long long q) { return Gu[q]; },
                             [&](long long q) { return Gv[q]; }, t, a, b, c);
  gu[t] = a;
  gv[t] = b;
  gw[t] = c;
  if (t / d.n2 == d.nz - 1) gw[(long long)d.nz * d.n2 + t % d.n2] = 0.0;
}

}  // namespace hoc

namespace hoc {

// ================================================================ reverse step
// The reverse of the theta step after its closure (model3d_v05.py:337-415,
// 466-533) for hoc_stepper_step_adjoint: tracers, the new velocity's w, the
// Picard passes in reverse order with the adjoint Helmholtz solves, the
// step's explicit terms, q and the face coefficients. Column kernels own a
// column (i, j) and march k; stencil kernels own one layer point. The
// stored geometry arrays are read (m3, m3u, m3v, dz3, dz3u, dz3v, zc).

// Row k of diffusion_coeffs (vertical_var.py:35-77) of one column over
// thickness dz and interface coefficient nu, with the factors ft, fb of
// nu_k in a_top and of nu_{k+1} in a_bot; wet and bottom flags of the cell.
struct DiffRow {
  double sub, diag, sup, ft, fb;
  bool wet, bot;
};

__device__ __forceinline__ DiffRow diff_row(const Dims& d, const Geo& geo, long long c, int k,
                                            const double* __restrict__ nu,
                                            const double* __restrict__ dz, double thv_dt,
                                            double thv_dt_drag) {
  const long long n2 = d.n2;
  const int nz = d.nz;
  const long long t = (long long)k * n2 + c;
  DiffRow r;
  r.wet = geo.m3[t] != 0;
  r.bot = r.wet && (k == nz - 1 || geo.m3[t + n2] == 0);
  const double dzk = dz[t];
  const double inv = safe_inv(dzk);
  const double fac = thv_dt * inv;
  const double dzi_t = k == 0 ? dzk : 0.5 * (dz[t - n2] + dzk);
  const double dzi_b = k == nz - 1 ? dzk : 0.5 * (dzk + dz[t + n2]);
  r.ft = fac * safe_inv(dzi_t);
  r.fb = fac * safe_inv(dzi_b);
  const double a_top = r.ft * nu[t], a_bot = r.fb * nu[t + n2];
  if (!r.wet) {
    r.sub = 0.0;
    r.diag = 1.0;
    r.sup = 0.0;
  } else {
    r.sub = k == 0 ? 0.0 : -a_top;
    if (r.bot) {
      r.diag = 1.0 + a_top + thv_dt_drag * inv;
      r.sup = 0.0;
    } else {
      r.diag = k == 0 ? 1.0 + a_bot : 1.0 + a_top + a_bot;
      r.sup = -a_bot;
    }
  }
  return r;
}

// lambda_r = A^-T g_r for one column's matrix A (rows diff_row) and NR
// right-hand sides (g(k, out) fills out[0..NR-1]); lam[r] receives lambda_r,
// cs is sweep scratch (all stride n2). Then the band gradients of
// -lambda_r^T (dA) x_r (x_r the forward solutions), summed over r, are turned
// into the adjoint of nu and added to gnu (vertical_var.py:35-77 reversed:
// the bottom row's sup and the dry rows carry no nu).
template <int NR, class GF>
__device__ void rev_column_solve(const Dims& d, const Geo& geo, long long c,
                                 const double* __restrict__ nu, const double* __restrict__ dz,
                                 double thv_dt, double thv_dt_drag, GF g,
                                 const double* const* x, double* const* lam,
                                 double* __restrict__ cs, double* __restrict__ gnu) {
  const long long n2 = d.n2;
  const int nz = d.nz;
  DiffRow r = diff_row(d, geo, c, 0, nu, dz, thv_dt, thv_dt_drag);
  double cp = 0.0, sup_prev = 0.0, dp[NR];
  for (int q = 0; q < NR; ++q) dp[q] = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long t = (long long)k * n2 + c;
    DiffRow rn{};
    double supT = 0.0;  // A^T: sub_T[k] = sup[k-1], sup_T[k] = sub[k+1]
    if (k + 1 < nz) {
      rn = diff_row(d, geo, c, k + 1, nu, dz, thv_dt, thv_dt_drag);
      supT = rn.sub;
    }
    const double den = k == 0 ? r.diag : r.diag - sup_prev * cp;
    const double cc = supT / den;
    double gk[NR];
    g(k, gk);
    for (int q = 0; q < NR; ++q) {
      const double dd = k == 0 ? gk[q] / den : (gk[q] - sup_prev * dp[q]) / den;
      lam[q][t] = dd;
      dp[q] = dd;
    }
    cs[t] = cc;
    cp = cc;
    sup_prev = r.sup;
    r = rn;
  }
  double ln[NR];
  for (int q = 0; q < NR; ++q) ln[q] = lam[q][(long long)(nz - 1) * n2 + c];
  for (int k = nz - 2; k >= 0; --k) {
    const long long t = (long long)k * n2 + c;
    for (int q = 0; q < NR; ++q) {
      ln[q] = lam[q][t] - cs[t] * ln[q];
      lam[q][t] = ln[q];
    }
  }
  // interface k takes row k-1's a_bot term, then row k's a_top term; each
  // interface is read and written once (the a_bot term carried over)
  bool carry = false;
  double bot_term = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long t = (long long)k * n2 + c;
    const DiffRow rk = diff_row(d, geo, c, k, nu, dz, thv_dt, thv_dt_drag);
    if (!rk.wet) {
      if (carry) gnu[t] = gnu[t] + bot_term;
      carry = false;
      continue;
    }
    double gd = 0.0, gs = 0.0, gp = 0.0;
    for (int q = 0; q < NR; ++q) {
      const double l = lam[q][t];
      gd -= l * x[q][t];
      if (k > 0) gs -= l * x[q][t - n2];
      if (k + 1 < nz && !rk.bot) gp -= l * x[q][t + n2];
    }
    const double ga_top = k >= 1 ? gd - gs : (rk.bot ? gd : 0.0);
    const double ga_bot = rk.bot ? 0.0 : gd - gp;
    double v = gnu[t];
    if (carry) v = v + bot_term;
    gnu[t] = v + rk.ft * ga_top;
    carry = true;
    bot_term = rk.fb * ga_bot;
  }
  if (carry) {
    const long long t = (long long)nz * n2 + c;
    gnu[t] = gnu[t] + bot_term;
  }
}

// Adjoint of apply_diffusion (vertical_var.py:80-106, cell masks, the
// surface flux a constant) for one column: with go(k) the adjoint of the
// output, adds dL/dfield to gf and dL/dnu to gnu.
template <class OF>
__device__ void rev_apply_diffusion(const Dims& d, const Geo& geo, long long c,
                                    const double* __restrict__ f, const double* __restrict__ nu,
                                    const double* __restrict__ dz, double drag, OF go,
                                    double* __restrict__ gf, double* __restrict__ gnu) {
  const long long n2 = d.n2;
  const int nz = d.nz;
  double oi_prev = 0.0, pend = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long t = (long long)k * n2 + c;
    const bool wet = geo.m3[t] != 0;
    const bool bot = wet && (k == nz - 1 || geo.m3[t + n2] == 0);
    const double oi = wet ? go(k) * safe_inv(dz[t]) : 0.0;
    double add = bot ? (-drag) * oi : 0.0;
    if (k >= 1) {  // interface k between cells k-1 and k (no flux below a bottom cell)
      const bool bot_above = geo.m3[t - n2] != 0 && !wet;
      const double gi = bot_above ? 0.0 : oi - oi_prev;
      const double rdzi = safe_inv(0.5 * (dz[t - n2] + dz[t]));
      const double cf = nu[t] * rdzi;
      gf[t - n2] += pend + cf * gi;
      add -= cf * gi;
      gnu[t] += (f[t - n2] - f[t]) * rdzi * gi;
    }
    pend = add;
    oi_prev = oi;
  }
  gf[(long long)(nz - 1) * n2 + c] += pend;
}

__global__ void k_fill(long long n, double value, double* __restrict__ a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = value;
}

// Right-hand sides of the reversed tracer solve, one cell per thread: the
// adjoint of T_new, S_new plus that of b_new through the EOS (dual numbers
// of buoyancy_t give both partials in one evaluation), masked, into rT, rS.
__global__ void k_rev_tracer_rhs(Dims d, StepParams p, Geo geo, const double* __restrict__ T1,
                                 const double* __restrict__ S1,
                                 const double* __restrict__ gT1,
                                 const double* __restrict__ gS1,
                                 const double* __restrict__ gb1, double* __restrict__ rT,
                                 double* __restrict__ rS) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const double m = mk(geo.m3[t]);
  double a = gT1 ? gT1[t] : 0.0, b = gS1 ? gS1[t] : 0.0;
  if (gb1 && geo.m3[t]) {
    const Dual2 db = buoyancy_t<Dual2>(p, Dual2(T1[t], 1.0, 0.0), Dual2(S1[t], 0.0, 1.0),
                                       Dual2(geo.zc[t]));
    const double gb = gb1[t] * m;
    a += gb * db.a;
    b += gb * db.b;
  }
  rT[t] = a * m;
  rS[t] = b * m;
}

// Tracers (model3d_v05.py:466-533) reversed per column for T and S at once:
// the right-hand sides k_rev_tracer_rhs left in lamT, lamS solved with
// tri_b^T in place; lambda_T, lambda_S out; gT0 = lambda + (1 - theta_v) dt
// D^T lambda (the explicit diffusion's adjoint); the nu adjoints of tri_b
// and of that diffusion added to gKh.
__global__ void k_rev_tracers(Dims d, StepParams p, Geo geo, const double* __restrict__ T0,
                              const double* __restrict__ S0, const double* __restrict__ T1,
                              const double* __restrict__ S1, const double* __restrict__ Kh,
                              double* __restrict__ lamT,
                              double* __restrict__ lamS, double* __restrict__ cs,
                              double* __restrict__ gT0, double* __restrict__ gS0,
                              double* __restrict__ gKh) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  // right-hand sides from k_rev_tracer_rhs; the solve overwrites them level
  // by level after reading them
  auto g = [&](int k, double* out) {
    const long long t = (long long)k * n2 + c;
    out[0] = lamT[t];
    out[1] = lamS[t];
  };
  const double* xs[2] = {T1, S1};
  double* ls[2] = {lamT, lamS};
  rev_column_solve<2>(d, geo, c, Kh, geo.dz3, p.thv_dt, 0.0, g, xs, ls, cs, gKh);
  for (int k = 0; k < d.nz; ++k) {
    const long long t = (long long)k * n2 + c;
    gT0[t] = lamT[t];
    gS0[t] = lamS[t];
  }
  rev_apply_diffusion(d, geo, c, T0, Kh, geo.dz3, 0.0,
                      [&](int k) { return p.c_dzd * lamT[(long long)k * n2 + c]; }, gT0, gKh);
  rev_apply_diffusion(d, geo, c, S0, Kh, geo.dz3, 0.0,
                      [&](int k) { return p.c_dzd * lamS[(long long)k * n2 + c]; }, gS0, gKh);
}

// Adjoint at point t of laplacian(., mask_u, mask_v, mask) (operators_masked.py:
// 64-68) applied to G: gradx^T and grady^T of div^T.
__device__ __forceinline__ double lap_T_point(const Dims& d, const StepParams& p,
                                              const uint8_t* __restrict__ mu,
                                              const uint8_t* __restrict__ mv,
                                              const uint8_t* __restrict__ m,
                                              const double* __restrict__ G, long long t) {
  const long long n2 = d.n2;
  const long long o = (t / n2) * n2, c = t % n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long e = o + at2(d, j, wrap(i, 1, d.nx)), w = o + at2(d, j, wrap(i, -1, d.nx));
  const long long n = o + at2(d, wrap(j, 1, d.ny), i), s = o + at2(d, wrap(j, -1, d.ny), i);
  const double mg = mk(m[t]) * G[t];
  const double Uc = (mg - mk(m[e]) * G[e]) / p.dx, Uw = (mk(m[w]) * G[w] - mg) / p.dx;
  const double Vc = (mg - mk(m[n]) * G[n]) / p.dy, Vs = (mk(m[s]) * G[s] - mg) / p.dy;
  return (mk(mu[w]) * Uw - mk(mu[t]) * Uc) / p.dx + (mk(mv[s]) * Vs - mk(mv[t]) * Vc) / p.dy;
}

// Tracer advection reversed on a RTS_TX x RTS_TY tile of one level: the
// same sums as tracer_adv_adjoint_point, with every face evaluated once.
// The tile's tracer and mask (3-cell halo) and Ghat = G mask / dz (2-cell
// halo) are staged in LDS; each x- and y-face the tile's cells read has its
// four products (fbar dz3 vel) dp[m] and its velocity adjoint fbar dz3 face
// formed once into LDS; each cell then adds its faces' products in the
// order of the gathered form (x faces s = -2..1, y faces, the vertical
// interfaces, the surface), so the result is that form's bit for bit
// (every product is rounded on its own). The vertical part stays per cell.
constexpr int RTS_CW = RTS_TX + 6, RTS_CH = RTS_TY + 6;  // tracer and mask, halo 3
constexpr int RTS_GW = RTS_TX + 4, RTS_GH = RTS_TY + 4;  // Ghat, halo 2
constexpr int RTS_XF = RTS_TX + 3, RTS_YF = RTS_TY + 3;  // faces i0-2 .. i0+TX (j likewise)
struct RtsLds {
  double C[RTS_CH][RTS_CW];
  double gh[RTS_GH][RTS_GW];
  double X[RTS_TY][RTS_XF][5];
  double Y[RTS_YF][RTS_TX][5];
  uint8_t M[RTS_CH][RTS_CW];
};

template <class GA>
__device__ __forceinline__ void tracer_adv_adjoint_tile(const Dims& d, const StepParams& p,
                                                        const Geo& geo,
                                                        const double* __restrict__ cc,
                                                        const double* __restrict__ u,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ w, GA G,
                                                        RtsLds& L, double& gc_out,
                                                        double& gu_out, double& gv_out,
                                                        double& gw_out) {
  const long long n2 = d.n2;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * RTS_TX, j0 = blockIdx.y * RTS_TY, k = blockIdx.z;
  const long long o = (long long)k * n2;
  const bool lim = p.advection == ADV_UP3_TVD;
  auto col = [&](int jj, int ii) -> long long { return (long long)pwrap(jj, ny) * nx + pwrap(ii, nx); };
  auto ghat = [&](long long q) -> double { return G(q) * mk(geo.m3[q]) * safe_inv(geo.dz3[q]); };
  for (int q = tid; q < RTS_CH * RTS_CW; q += RTS_TX * RTS_TY) {
    const int qy = q / RTS_CW, qx = q % RTS_CW;
    const long long cell = o + col(j0 - 3 + qy, i0 - 3 + qx);
    L.C[qy][qx] = cc[cell];
    L.M[qy][qx] = geo.m3[cell];
  }
  for (int q = tid; q < RTS_GH * RTS_GW; q += RTS_TX * RTS_TY) {
    const int qy = q / RTS_GW, qx = q % RTS_GW;
    L.gh[qy][qx] = ghat(o + col(j0 - 2 + qy, i0 - 2 + qx));
  }
  __syncthreads();
  // x-faces: face fi is the east face of column i0 - 2 + fi in row j0 + ty
  for (int q = tid; q < RTS_TY * RTS_XF; q += RTS_TX * RTS_TY) {
    const int ty = q / RTS_XF, fi = q % RTS_XF, cy = ty + 3;
    const long long f = o + col(j0 + ty, i0 - 2 + fi);
    const double vel = u[f];
    double dp[4];
    const double face = kappa_face_d(L.C[cy][fi], L.C[cy][fi + 1], L.C[cy][fi + 2],
                                     L.C[cy][fi + 3], vel, L.M[cy][fi] != 0,
                                     L.M[cy][fi + 3] != 0, lim, dp);
    const double fbar = (L.gh[ty + 2][fi] - L.gh[ty + 2][fi + 1]) / p.dx;
    const double a = fbar * geo.dz3u[f] * vel;
#pragma unroll
    for (int m = 0; m < 4; ++m) L.X[ty][fi][m] = a * dp[m];
    L.X[ty][fi][4] = fbar * geo.dz3u[f] * face;
  }
  // y-faces: face row fj is the north face of row j0 - 2 + fj in column i0 + tx
  for (int q = tid; q < RTS_YF * RTS_TX; q += RTS_TX * RTS_TY) {
    const int fj = q / RTS_TX, tx = q % RTS_TX, cx = tx + 3;
    const long long f = o + col(j0 - 2 + fj, i0 + tx);
    const double vel = v[f];
    double dp[4];
    const double face = kappa_face_d(L.C[fj][cx], L.C[fj + 1][cx], L.C[fj + 2][cx],
                                     L.C[fj + 3][cx], vel, L.M[fj][cx] != 0,
                                     L.M[fj + 3][cx] != 0, lim, dp);
    const double fbar = (L.gh[fj][tx + 2] - L.gh[fj + 1][tx + 2]) / p.dy;
    const double a = fbar * geo.dz3v[f] * vel;
#pragma unroll
    for (int m = 0; m < 4; ++m) L.Y[fj][tx][m] = a * dp[m];
    L.Y[fj][tx][4] = fbar * geo.dz3v[f] * face;
  }
  __syncthreads();
  const int tx = tid % RTS_TX, ty = tid / RTS_TX;
  const int i = i0 + tx, j = j0 + ty;
  if (i >= nx || j >= ny) return;
  const long long c = (long long)j * nx + i;
  double acc = 0.0;
#pragma unroll
  for (int s = -2; s <= 1; ++s) acc += L.X[ty][tx + s + 2][1 - s];
  gu_out = L.X[ty][tx + 2][4];
#pragma unroll
  for (int s = -2; s <= 1; ++s) acc += L.Y[ty + s + 2][tx][1 - s];
  gv_out = L.Y[ty + 2][tx][4];
  const double gh_k = L.gh[ty + 2][tx + 2];
  auto ghat_at = [&](int kk) -> double {
    return kk == k ? gh_k : ghat((long long)kk * n2 + c);
  };
  auto zbar = [&](int kk) -> double {
    const double below = ghat_at(kk);
    const double above = kk >= 1 ? ghat_at(kk - 1) : 0.0;
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
    const double zb = zbar(kk);
    const double wb = zb * w[a];
    acc += wb * dp[kk == k - 1 ? 0 : kk == k ? 1 : kk == k + 1 ? 2 : 3];
    if (kk == k) gw_out = zb * face;
  }
  if (k == 0) {
    acc += gh_k * w[c];
    gw_out = gh_k * cc[c];
  }
  gc_out = acc;
}

// Tracer advection (hand-written reverse of adv) and K_h Laplacian reversed
// for T and S on one tile of one level (grid: tiles x tiles x nz, RTS_TX x
// RTS_TY threads): adds to gT0, gS0; the new velocity's adjoint gu1 = gu_y +
// the flux-velocity parts (gv1 likewise); gw1 the adjoint of the new w
// (both tracers).
__global__ __launch_bounds__(RTS_TX* RTS_TY) void k_rev_tracer_stencil(
    Dims d, StepParams p, Geo geo, const double* __restrict__ T0, const double* __restrict__ S0,
    const double* __restrict__ u1, const double* __restrict__ v1, const double* __restrict__ w1,
    const double* __restrict__ lamT, const double* __restrict__ lamS,
    const double* __restrict__ gu_y, const double* __restrict__ gv_y, double* __restrict__ gT0,
    double* __restrict__ gS0, double* __restrict__ gu1, double* __restrict__ gv1,
    double* __restrict__ gw1) {
  __shared__ RtsLds L;
  const double ndt = -p.dt;
  double cT = 0.0, uT = 0.0, vT = 0.0, wT = 0.0, cS = 0.0, uS = 0.0, vS = 0.0, wS = 0.0;
  tracer_adv_adjoint_tile(d, p, geo, T0, u1, v1, w1, [&](long long q) { return ndt * lamT[q]; },
                          L, cT, uT, vT, wT);
  __syncthreads();
  tracer_adv_adjoint_tile(d, p, geo, S0, u1, v1, w1, [&](long long q) { return ndt * lamS[q]; },
                          L, cS, uS, vS, wS);
  const int i = blockIdx.x * RTS_TX + (int)threadIdx.x % RTS_TX;
  const int j = blockIdx.y * RTS_TY + (int)threadIdx.x / RTS_TX;
  if (i >= d.nx || j >= d.ny) return;
  const long long t = (long long)blockIdx.z * d.n2 + (long long)j * d.nx + i;
  if (p.has_K_h) {
    const double dk = p.dt * p.K_h;
    cT += dk * lap_T_point(d, p, geo.m3u, geo.m3v, geo.m3, lamT, t);
    cS += dk * lap_T_point(d, p, geo.m3u, geo.m3v, geo.m3, lamS, t);
  }
  gT0[t] += cT;
  gS0[t] += cS;
  gu1[t] = (gu_y ? gu_y[t] : 0.0) + uT + uS;
  gv1[t] = (gv_y ? gv_y[t] : 0.0) + vT + vS;
  gw1[t] = wT + wS;
  if (t / d.n2 == d.nz - 1) gw1[(long long)d.nz * d.n2 + t % d.n2] = 0.0;
}

// Back-substitution u_it = (gh_u - g dt th q Gx(eta) mask3u) mask3 reversed
// for one column (model3d_v05.py:404-406); with gw (the last pass) the
// adjoint of w = w_diag(u_it, v_it) is added first (w_k = -sum_{j>=k} div_j,
// div of the face transports). gui, gvi are replaced by the adjoints of
// the masked predictor solutions' pre-mask expression (A = gui mask3); q's
// adjoint is added to gq; X, Y (columns) are the sums the free surface's
// adjoint takes through Gx^T, Gy^T.
__global__ void k_rev_backsub(Dims d, StepParams p, Geo geo, const double* __restrict__ gw,
                              const double* __restrict__ q, const double* __restrict__ eta,
                              int first, double* __restrict__ gui, double* __restrict__ gvi,
                              double* __restrict__ gq, double* __restrict__ X,
                              double* __restrict__ Y) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cn = at2(d, wrap(j, 1, d.ny), i);
  const double gx2 = mk(geo.mu2[c]) * (eta[ce] - eta[c]) / p.dx;
  const double gy2 = mk(geo.mv2[c]) * (eta[cn] - eta[c]) / p.dy;
  double pc = 0.0, pe = 0.0, pn = 0.0, xs = 0.0, ys = 0.0;
  for (int k = 0; k < d.nz; ++k) {
    const long long o = (long long)k * n2, t = o + c;
    double ui = gui[t], vi = gvi[t];
    if (gw) {
      pc -= gw[t];
      pe -= gw[o + ce];
      pn -= gw[o + cn];
      const double mc = mk(geo.m3[t]) * pc;
      ui += geo.dz3u[t] * ((mc - mk(geo.m3[o + ce]) * pe) / p.dx);
      vi += geo.dz3v[t] * ((mc - mk(geo.m3[o + cn]) * pn) / p.dy);
    }
    const double m = mk(geo.m3[t]);
    const double Au = ui * m, Av = vi * m;
    const double mu = mk(geo.m3u[t]), mv = mk(geo.m3v[t]);
    const double dq = -(p.c_bs * (gx2 * mu * Au + gy2 * mv * Av));
    gq[t] = first ? dq : gq[t] + dq;
    xs += -p.c_bs * q[t] * mu * Au;
    ys += -p.c_bs * q[t] * mv * Av;
    gui[t] = Au;
    gvi[t] = Av;
  }
  X[c] = xs;
  Y[c] = ys;
}

// The right-hand side of the adjoint Helmholtz solve of one pass:
// ge = g_eta (the output's, last pass only) + Gx^T X + Gy^T Y.
__global__ void k_rev_eta_rhs(Dims d, StepParams p, Geo geo, const double* __restrict__ geta,
                              const double* __restrict__ X, const double* __restrict__ Y,
                              double* __restrict__ ge) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long cw = at2(d, j, wrap(i, -1, d.nx)), cs = at2(d, wrap(j, -1, d.ny), i);
  const double mxc = mk(geo.mu2[c]) * X[c], myc = mk(geo.mv2[c]) * Y[c];
  ge[c] = (geta ? geta[c] : 0.0) + (mk(geo.mu2[cw]) * X[cw] - mxc) / p.dx +
          (mk(geo.mv2[cs]) * Y[cs] - myc) / p.dy;
}

// After the adjoint solve lam = A^-1 ge of one pass: the face coefficients'
// adjoints -lam^T (dA/dK) x (x the pass's free surface), the old free
// surface's and old divergence's shares of rhs = eta - dt((1 - th) div_old
// + th div_g), and the face values GU, GV of div^T(-dt th lam).
__global__ void k_rev_eta_post(Dims d, StepParams p, Geo geo, const double* __restrict__ x,
                               const double* __restrict__ lam, int first, double* __restrict__ gKu,
                               double* __restrict__ gKv, double* __restrict__ geta0,
                               double* __restrict__ gdiv, double* __restrict__ GU,
                               double* __restrict__ GV) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cn = at2(d, wrap(j, 1, d.ny), i);
  const double mc = mk(geo.m2[c]), me = mk(geo.m2[ce]), mn = mk(geo.m2[cn]);
  const double xc = x[c] * mc, xe = x[ce] * me, xn = x[cn] * mn;
  const double lc = lam[c] * mc, le = lam[ce] * me, ln = lam[cn] * mn;
  const double ku = p.helm_coef * (xe - xc) * (lc - le) / p.dx2;
  const double kv = p.helm_coef * (xn - xc) * (lc - ln) / p.dy2;
  const double dv = -(p.dt * p.one_m_th * lam[c]);
  gKu[c] = first ? ku : gKu[c] + ku;
  gKv[c] = first ? kv : gKv[c] + kv;
  geta0[c] = first ? lam[c] : geta0[c] + lam[c];
  gdiv[c] = first ? dv : gdiv[c] + dv;
  const double dc = -p.dt * p.th * lam[c];
  const double de = -p.dt * p.th * lam[ce], dn = -p.dt * p.th * lam[cn];
  GU[c] = (mc * dc - me * de) / p.dx;
  GV[c] = (mc * dc - mn * dn) / p.dy;
}

// The predictor's two solves gh = A_m^-1 (g + stress) reversed for one
// column (model3d_v05.py:394-399): the adjoint of gh (A + dz3[u|v] G[U|V]),
// masked, solved with tri_m^T into lu, lv; tri_m's nu adjoint added to gKm;
// Lu += lu, Lv += lv (the sums over passes).
__global__ void k_rev_predictor(Dims d, StepParams p, Geo geo, const double* __restrict__ Km,
                                const double* __restrict__ Au, const double* __restrict__ Av,
                                const double* __restrict__ GU, const double* __restrict__ GV,
                                const double* __restrict__ ghu, const double* __restrict__ ghv,
                                int first, double* __restrict__ lu, double* __restrict__ lv,
                                double* __restrict__ cs, double* __restrict__ Lu,
                                double* __restrict__ Lv, double* __restrict__ gKm) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const double gUc = GU[c], gVc = GV[c];
  auto g = [&](int k, double* out) {
    const long long t = (long long)k * n2 + c;
    const double m = mk(geo.m3[t]);
    out[0] = (Au[t] + geo.dz3u[t] * gUc) * m;
    out[1] = (Av[t] + geo.dz3v[t] * gVc) * m;
  };
  const double* xs[2] = {ghu, ghv};
  double* ls[2] = {lu, lv};
  rev_column_solve<2>(d, geo, c, Km, geo.dz3, p.thv_dt, p.thv_dt_drag, g, xs, ls, cs, gKm);
  for (int k = 0; k < d.nz; ++k) {
    const long long t = (long long)k * n2 + c;
    Lu[t] = first ? lu[t] : Lu[t] + lu[t];
    Lv[t] = first ? lv[t] : Lv[t] + lv[t];
  }
}

// Coriolis terms reversed at one point (operators_masked.py:45-61
// transposed): the u point c takes part in the v-point means of c, e, s,
// se; the v point in the u-point means of c, n, w, nw. With first == 0
// (a pass after the first) the previous iterate's adjoints are written:
// gu = tc Au2v^T(-dt f lv), gv = tc Av2u^T(dt f lu); with first != 0 the
// old state's shares of both passes are written (the first contributions
// to its velocity adjoints): gu = Au2v^T(-dt f (tc lv + (1 - tc) Lv)),
// gv = Av2u^T(dt f (tc lu + (1 - tc) Lu)).
__global__ void k_rev_coriolis(Dims d, StepParams p, Geo geo, const double* __restrict__ lu,
                               const double* __restrict__ lv, const double* __restrict__ Lu,
                               const double* __restrict__ Lv, int first,
                               double* __restrict__ gu, double* __restrict__ gv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long n2 = d.n2;
  const long long o = (t / n2) * n2, c = t % n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  auto at = [&](int dj, int di) { return o + at2(d, wrap(j, dj, d.ny), wrap(i, di, d.nx)); };
  const uint8_t* mu = geo.m3u;
  const uint8_t* mv = geo.m3v;
  auto Gv = [&](long long q) {  // the v-point adjoint of u_cor
    return first ? -p.dt * p.f * (p.theta_cor * lv[q] + p.one_m_tc * Lv[q]) : -p.dt * p.f * lv[q];
  };
  auto Gu = [&](long long q) {  // the u-point adjoint of v_cor
    return first ? p.dt * p.f * (p.theta_cor * lu[q] + p.one_m_tc * Lu[q]) : p.dt * p.f * lu[q];
  };
  // H'(q) at a v point q: G mask_v / (mu(q) + mu(w q) + mu(n q) + mu(nw q))
  auto Hv = [&](int dj, int di) {
    const long long q = at(dj, di);
    const double den = mk(mu[q]) + mk(mu[at(dj, di - 1)]) + mk(mu[at(dj + 1, di)]) +
                       mk(mu[at(dj + 1, di - 1)]);
    return den > 0.0 ? Gv(q) * mk(mv[q]) / den : 0.0;
  };
  // H(q) at a u point q: G mask_u / (mv(q) + mv(s q) + mv(e q) + mv(se q))
  auto Hu = [&](int dj, int di) {
    const long long q = at(dj, di);
    const double den = mk(mv[q]) + mk(mv[at(dj - 1, di)]) + mk(mv[at(dj, di + 1)]) +
                       mk(mv[at(dj - 1, di + 1)]);
    return den > 0.0 ? Gu(q) * mk(mu[q]) / den : 0.0;
  };
  const double su = mk(mu[t]) * (Hv(0, 0) + Hv(0, 1) + Hv(-1, 0) + Hv(-1, 1));
  const double sv = mk(mv[t]) * (Hu(0, 0) + Hu(1, 0) + Hu(0, -1) + Hu(1, -1));
  if (first) {  // the first writer of the old state's velocity adjoints
    gu[t] = su;
    gv[t] = sv;
  } else {
    gu[t] = p.theta_cor * su;
    gv[t] = p.theta_cor * sv;
  }
}

// The step's explicit terms reversed, from Lu, Lv (the sums of the
// predictor solves' adjoints over the passes, model3d_v05.py:349-386), one
// cell per thread: gu0 += Lu + dz3u div^T(gdiv) + A_h dt Lap^T(Lu), gv0
// likewise. k_rev_explicit then adds the column parts.
__global__ void k_rev_explicit_cells(Dims d, StepParams p, Geo geo, const double* __restrict__ Lu,
                                     const double* __restrict__ Lv,
                                     const double* __restrict__ gdiv, double* __restrict__ gu0,
                                     double* __restrict__ gv0) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long c = t % d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cn = at2(d, wrap(j, 1, d.ny), i);
  const double Dc = mk(geo.m2[c]) * gdiv[c];
  const double DUx = (Dc - mk(geo.m2[ce]) * gdiv[ce]) / p.dx;
  const double DVy = (Dc - mk(geo.m2[cn]) * gdiv[cn]) / p.dy;
  const double dtA = p.dt * p.A_h;
  double au = Lu[t] + geo.dz3u[t] * DUx;
  double av = Lv[t] + geo.dz3v[t] * DVy;
  if (p.has_A_h) {
    au += dtA * lap_T_point(d, p, geo.m3u, geo.m3v, geo.m3u, Lu, t);
    av += dtA * lap_T_point(d, p, geo.m3u, geo.m3v, geo.m3v, Lv, t);
  }
  gu0[t] += au;
  gv0[t] += av;
}

// The column parts of the explicit terms reversed, after
// k_rev_explicit_cells: the explicit diffusion's field adjoint added to
// gu0, gv0 (gKm takes its nu adjoint); geta0 += Gx^T of the
// surface-gradient term; gb0 = the pressure gradients' adjoint
// (model3d_v05.py:195-230).
__global__ void k_rev_explicit(Dims d, StepParams p, Geo geo, const double* __restrict__ u0,
                               const double* __restrict__ v0, const double* __restrict__ Km,
                               const double* __restrict__ Lu, const double* __restrict__ Lv,
                               double* __restrict__ gu0,
                               double* __restrict__ gv0, double* __restrict__ gb0,
                               double* __restrict__ geta0, double* __restrict__ gKm) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const int nz = d.nz;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i), cs = at2(d, wrap(j, -1, d.ny), i);
  double xc = 0.0, xw = 0.0, yc = 0.0, ys = 0.0;
  for (int k = 0; k < nz; ++k) {
    const long long o = (long long)k * n2, t = o + c;
    xc += mk(geo.m3u[t]) * (-p.c_gx) * Lu[t];
    xw += mk(geo.m3u[o + cw]) * (-p.c_gx) * Lu[o + cw];
    yc += mk(geo.m3v[t]) * (-p.c_gx) * Lv[t];
    ys += mk(geo.m3v[o + cs]) * (-p.c_gx) * Lv[o + cs];
  }
  geta0[c] += (mk(geo.mu2[cw]) * xw - mk(geo.mu2[c]) * xc) / p.dx +
              (mk(geo.mv2[cs]) * ys - mk(geo.mv2[c]) * yc) / p.dy;
  rev_apply_diffusion(d, geo, c, u0, Km, geo.dz3u, p.bottom_drag,
                      [&](int k) { return p.c_dzd * Lu[(long long)k * n2 + c]; }, gu0, gKm);
  rev_apply_diffusion(d, geo, c, v0, Km, geo.dz3v, p.bottom_drag,
                      [&](int k) { return p.c_dzd * Lv[(long long)k * n2 + c]; }, gv0, gKm);
  // pressure gradients: phi's adjoint and b's direct share per level, then
  // phi = dz3 b / 2 + sum_{j<k} dz3 b reversed from the bottom up
  auto level = [&](int k, double& gphi, double& gbd) {
    const long long o = (long long)k * n2, t = o + c;
    const double Pc = mk(geo.m3u[t]) * (p.dt * Lu[t]) / p.dx;
    const double Pw = mk(geo.m3u[o + cw]) * (p.dt * Lu[o + cw]) / p.dx;
    const double Qc = mk(geo.m3v[t]) * (p.dt * Lv[t]) / p.dy;
    const double Qs = mk(geo.m3v[o + cs]) * (p.dt * Lv[o + cs]) / p.dy;
    gphi = (Pw - Pc) + (Qs - Qc);
    gbd = 0.0;
    if (p.pgf_correction) {
      const double zc = geo.zc[t];
      const double zfx = 0.5 * (zc + geo.zc[o + ce]), zfw = 0.5 * (geo.zc[o + cw] + zc);
      const double zfy = 0.5 * (zc + geo.zc[o + cn]), zfs = 0.5 * (geo.zc[o + cs] + zc);
      gbd = Pw * (zfw - zc) - Pc * (zfx - zc) + Qs * (zfs - zc) - Qc * (zfy - zc);
    }
  };
  double rm = 0.0;
  if (p.pgf_remove) {
    double tot = 0.0, sg = 0.0;
    for (int k = 0; k < nz; ++k) {
      double gphi, gbd;
      level(k, gphi, gbd);
      tot += geo.dz3[(long long)k * n2 + c];
      sg += gphi;
    }
    rm = safe_inv(tot) * sg;
  }
  double below = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long t = (long long)k * n2 + c;
    double gphi, gbd;
    level(k, gphi, gbd);
    const double dz = geo.dz3[t];
    if (p.pgf_remove) gphi -= dz * rm;
    gb0[t] = gbd + 0.5 * dz * gphi + dz * below;
    below += gphi;
  }
}

// Momentum advection of the old state reversed on a RMA_TX x RMA_TY tile of
// one level with the adjoints -dt m3[u|v] L[u|v] of the advection terms:
// the sums of momentum_adv_adjoint_point with every horizontal face
// evaluated once. u, v, their masks (3-cell halo) and the two adjoints
// (2-cell halo) are staged in LDS. Phase 1 forms the u equation's faces (A:
// x fluxes at centres, B: y fluxes at corners) into LDS, phase 2 the v
// equation's (C: y fluxes at centres, D: x fluxes at corners) in the same
// space; each face holds its four products (fbar wb) dp[m] and the product
// the other velocity's gather takes. Each point adds them in the gathered
// form's order for each accumulator (gU: A, B, then D's corner share; gV:
// B's corner share, C, D), then the vertical part as before, so the result
// is that form's bit for bit (every product is rounded on its own).
constexpr int RMA_SW = RMA_TX + 6, RMA_SH = RMA_TY + 6;  // u, v, masks: halo 3
constexpr int RMA_GW = RMA_TX + 4, RMA_GH = RMA_TY + 4;  // adjoints: halo 2
struct RmaLds {
  double U[RMA_SH][RMA_SW], V[RMA_SH][RMA_SW];
  double GU[RMA_GH][RMA_GW], GV[RMA_GH][RMA_GW];
  union {
    struct {
      double A[RMA_TY][RMA_TX + 3][5];      // rows j0.., centres i0-1 .. i0+TX+1
      double B[RMA_TY + 3][RMA_TX + 1][5];  // corner rows j0-2 .., columns i0-1 ..
    } p1;
    struct {
      double C[RMA_TY + 3][RMA_TX][5];      // centre rows j0-1 .., columns i0 ..
      double D[RMA_TY + 1][RMA_TX + 3][5];  // rows j0-1 .., corner columns i0-2 ..
    } p2;
  } f;
  uint8_t MU[RMA_SH][RMA_SW], MV[RMA_SH][RMA_SW];
};

__global__ __launch_bounds__(RMA_TX* RMA_TY) void k_rev_momentum_adv(
    Dims d, StepParams p, Geo geo, const double* __restrict__ u0, const double* __restrict__ v0,
    const double* __restrict__ w0, const double* __restrict__ Lu, const double* __restrict__ Lv,
    double* __restrict__ gu0, double* __restrict__ gv0, double* __restrict__ gw0) {
  __shared__ RmaLds L;
  const long long n2 = d.n2;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * RMA_TX, j0 = blockIdx.y * RMA_TY, k = blockIdx.z;
  const long long o = (long long)k * n2;
  const bool lim = p.advection == ADV_UP3_TVD;
  auto col = [&](int jj, int ii) -> long long { return (long long)pwrap(jj, ny) * nx + pwrap(ii, nx); };
  auto Gu = [&](long long q) { return -(mk(geo.m3u[q]) * (p.dt * Lu[q])); };
  auto Gv = [&](long long q) { return -(mk(geo.m3v[q]) * (p.dt * Lv[q])); };
  constexpr int NT = RMA_TX * RMA_TY;
  for (int q = tid; q < RMA_SH * RMA_SW; q += NT) {
    const int qy = q / RMA_SW, qx = q % RMA_SW;
    const long long cell = o + col(j0 - 3 + qy, i0 - 3 + qx);
    L.U[qy][qx] = u0[cell];
    L.V[qy][qx] = v0[cell];
    L.MU[qy][qx] = geo.m3u[cell];
    L.MV[qy][qx] = geo.m3v[cell];
  }
  for (int q = tid; q < RMA_GH * RMA_GW; q += NT) {
    const int qy = q / RMA_GW, qx = q % RMA_GW;
    const long long cell = o + col(j0 - 2 + qy, i0 - 2 + qx);
    L.GU[qy][qx] = Gu(cell);
    L.GV[qy][qx] = Gv(cell);
  }
  __syncthreads();
  double dp[4];
  // ---- phase 1: the u equation's faces
  for (int q = tid; q < RMA_TY * (RMA_TX + 3); q += NT) {  // A: x flux at centre i0 - 1 + a
    const int ty = q / (RMA_TX + 3), a = q % (RMA_TX + 3), ry = ty + 3;
    const double ub = 0.5 * (L.U[ry][a + 1] + L.U[ry][a + 2]);
    const double uc = kappa_face_d(L.U[ry][a], L.U[ry][a + 1], L.U[ry][a + 2], L.U[ry][a + 3],
                                   ub, L.MU[ry][a] != 0, L.MU[ry][a + 3] != 0, lim, dp);
    const double fbar = (L.GU[ty + 2][a] - L.GU[ty + 2][a + 1]) / p.dx;
#pragma unroll
    for (int m = 0; m < 4; ++m) L.f.p1.A[ty][a][m] = fbar * ub * dp[m];
    L.f.p1.A[ty][a][4] = fbar * 0.5 * uc;
  }
  for (int q = tid; q < (RMA_TY + 3) * (RMA_TX + 1); q += NT) {  // B: y flux at corner row
    const int b = q / (RMA_TX + 1), e = q % (RMA_TX + 1), cx = e + 2;  // j0 - 2 + b, column i0 - 1 + e
    const double vb = 0.5 * (L.V[b + 1][cx] + L.V[b + 1][cx + 1]);
    const double uy = kappa_face_d(L.U[b][cx], L.U[b + 1][cx], L.U[b + 2][cx], L.U[b + 3][cx],
                                   vb, L.MU[b][cx] != 0, L.MU[b + 3][cx] != 0, lim, dp);
    const double fbar = (L.GU[b][e + 1] - L.GU[b + 1][e + 1]) / p.dy;
#pragma unroll
    for (int m = 0; m < 4; ++m) L.f.p1.B[b][e][m] = fbar * vb * dp[m];
    L.f.p1.B[b][e][4] = fbar * uy * 0.5;
  }
  __syncthreads();
  const int tx = tid % RMA_TX, ty = tid / RMA_TX;
  const int i = i0 + tx, j = j0 + ty;
  const bool inside = i < nx && j < ny;
  double gU = 0.0, gV = 0.0;
  if (inside) {
#pragma unroll
    for (int s = -1; s <= 2; ++s) {
      gU += L.f.p1.A[ty][tx + s + 1][2 - s];
      if (s == 0 || s == 1) gU += L.f.p1.A[ty][tx + s + 1][4];
    }
#pragma unroll
    for (int s = -2; s <= 1; ++s) gU += L.f.p1.B[ty + s + 2][tx + 1][1 - s];
#pragma unroll
    for (int s = -1; s <= 0; ++s) gV += L.f.p1.B[ty + 2][tx + 1 + s][4];
  }
  __syncthreads();
  // ---- phase 2: the v equation's faces
  for (int q = tid; q < (RMA_TY + 3) * RMA_TX; q += NT) {  // C: y flux at centre row j0 - 1 + cr
    const int cr = q / RMA_TX, cx = q % RMA_TX + 3;
    const double vb = 0.5 * (L.V[cr + 1][cx] + L.V[cr + 2][cx]);
    const double vc = kappa_face_d(L.V[cr][cx], L.V[cr + 1][cx], L.V[cr + 2][cx],
                                   L.V[cr + 3][cx], vb, L.MV[cr][cx] != 0,
                                   L.MV[cr + 3][cx] != 0, lim, dp);
    const double fbar = (L.GV[cr][cx - 1] - L.GV[cr + 1][cx - 1]) / p.dy;
#pragma unroll
    for (int m = 0; m < 4; ++m) L.f.p2.C[cr][cx - 3][m] = fbar * vb * dp[m];
    L.f.p2.C[cr][cx - 3][4] = fbar * 0.5 * vc;
  }
  for (int q = tid; q < (RMA_TY + 1) * (RMA_TX + 3); q += NT) {  // D: x flux at corner column
    const int r = q / (RMA_TX + 3), a = q % (RMA_TX + 3);  // row j0 - 1 + r, column i0 - 2 + a
    const double ub = 0.5 * (L.U[r + 2][a + 1] + L.U[r + 3][a + 1]);
    const double vx = kappa_face_d(L.V[r + 2][a], L.V[r + 2][a + 1], L.V[r + 2][a + 2],
                                   L.V[r + 2][a + 3], ub, L.MV[r + 2][a] != 0,
                                   L.MV[r + 2][a + 3] != 0, lim, dp);
    const double fbar = (L.GV[r + 1][a] - L.GV[r + 1][a + 1]) / p.dx;
#pragma unroll
    for (int m = 0; m < 4; ++m) L.f.p2.D[r][a][m] = fbar * ub * dp[m];
    L.f.p2.D[r][a][4] = fbar * vx * 0.5;
  }
  __syncthreads();
  if (!inside) return;
#pragma unroll
  for (int s = -1; s <= 2; ++s) {
    gV += L.f.p2.C[ty + s + 1][tx][2 - s];
    if (s == 0 || s == 1) gV += L.f.p2.C[ty + s + 1][tx][4];
  }
#pragma unroll
  for (int s = -2; s <= 1; ++s) gV += L.f.p2.D[ty + 1][tx + s + 2][1 - s];
#pragma unroll
  for (int s = -1; s <= 0; ++s) gU += L.f.p2.D[ty + 1 + s][tx + 2][4];
  // ---- vertical fluxes (interfaces 1..nz-1 carry them), per point
  const long long c = (long long)j * nx + i;
  auto zface = [&](const double* f, const uint8_t* m3, int kk, long long cl, double wb,
                   double* pd) -> double {
    const long long a = (long long)kk * n2 + cl;
    const double far_up = kk >= 2 ? f[a - 2 * n2] : 0.0;
    const double far_dn = kk <= nz - 2 ? f[a + n2] : 0.0;
    const bool hu = kk >= 2 && m3[a - 2 * n2] != 0;
    const bool hd = kk <= nz - 2 && m3[a + n2] != 0;
    return kappa_face_d(far_dn, f[a], f[a - n2], far_up, wb, hd, hu, lim, pd);
  };
  const long long ce = col(j, i + 1), cn = col(j + 1, i);
  for (int kk = k - 1; kk <= k + 2; ++kk) {
    if (kk < 1 || kk > nz - 1) continue;
    const long long a = (long long)kk * n2;
    const int role = kk == k - 1 ? 0 : kk == k ? 1 : kk == k + 1 ? 2 : 3;
    const double wbu = 0.5 * (w0[a + c] + w0[a + ce]);
    zface(u0, geo.m3u, kk, c, wbu, dp);
    gU += (Gu(a + c) - Gu(a - n2 + c)) / p.dz_nom * wbu * dp[role];
    const double wbv = 0.5 * (w0[a + c] + w0[a + cn]);
    zface(v0, geo.m3v, kk, c, wbv, dp);
    gV += (Gv(a + c) - Gv(a - n2 + c)) / p.dz_nom * wbv * dp[role];
  }
  double gW = 0.0;
  if (k >= 1) {  // interface k: w enters wb of the u columns i-1, i and v rows j-1, j
    const long long a = (long long)k * n2;
    for (int s = -1; s <= 0; ++s) {
      const long long cu = col(j, i + s), cue = col(j, i + s + 1);
      const double wbu = 0.5 * (w0[a + cu] + w0[a + cue]);
      const double fu = zface(u0, geo.m3u, k, cu, wbu, dp);
      gW += (Gu(a + cu) - Gu(a - n2 + cu)) / p.dz_nom * fu * 0.5;
      const long long cv = col(j + s, i), cvn = col(j + s + 1, i);
      const double wbv = 0.5 * (w0[a + cv] + w0[a + cvn]);
      const double fv = zface(v0, geo.m3v, k, cv, wbv, dp);
      gW += (Gv(a + cv) - Gv(a - n2 + cv)) / p.dz_nom * fv * 0.5;
    }
  }
  const long long t = o + c;
  gu0[t] += gU;
  gv0[t] += gV;
  gw0[t] = gW;
  if (k == nz - 1) gw0[(long long)nz * n2 + c] = 0.0;
}

// Per column: the old w's adjoint into gu0, gv0 (w_diag reversed), then
// q = A_m^-1 1 reversed: q's adjoint gains the face coefficients' shares
// (K_u = sum dz3u min(q, q_E) mask_u, model3d_v05.py:150-167, the minimum's
// kept operand differentiated), is solved with tri_m^T and tri_m's nu
// adjoint is added to gKm.
__global__ void k_rev_wold_q(Dims d, StepParams p, Geo geo, const double* __restrict__ gw0,
                             const double* __restrict__ Km, const double* __restrict__ q,
                             const double* __restrict__ gKu, const double* __restrict__ gKv,
                             double* __restrict__ gq, double* __restrict__ lam,
                             double* __restrict__ cs, double* __restrict__ gu0,
                             double* __restrict__ gv0, double* __restrict__ gKm) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long ce = at2(d, j, wrap(i, 1, d.nx)), cw = at2(d, j, wrap(i, -1, d.nx));
  const long long cn = at2(d, wrap(j, 1, d.ny), i), cs2 = at2(d, wrap(j, -1, d.ny), i);
  const double Gxc = gKu[c] * mk(geo.mu2[c]), Gxw = gKu[cw] * mk(geo.mu2[cw]);
  const double Gyc = gKv[c] * mk(geo.mv2[c]), Gys = gKv[cs2] * mk(geo.mv2[cs2]);
  double pc = 0.0, pe = 0.0, pn = 0.0;
  for (int k = 0; k < d.nz; ++k) {
    const long long o = (long long)k * n2, t = o + c;
    pc -= gw0[t];
    pe -= gw0[o + ce];
    pn -= gw0[o + cn];
    const double mc = mk(geo.m3[t]) * pc;
    gu0[t] += geo.dz3u[t] * ((mc - mk(geo.m3[o + ce]) * pe) / p.dx);
    gv0[t] += geo.dz3v[t] * ((mc - mk(geo.m3[o + cn]) * pn) / p.dy);
    const double qc = q[t];
    double a = gq[t];
    const bool kx = qc <= q[o + ce] || qc != qc;           // min(q, q_E) keeps q at c
    const bool kxw = q[o + cw] <= qc || q[o + cw] != q[o + cw];  // ... keeps q_W at w
    const bool ky = qc <= q[o + cn] || qc != qc;
    const bool kys = q[o + cs2] <= qc || q[o + cs2] != q[o + cs2];
    if (kx) a += geo.dz3u[t] * Gxc;
    if (!kxw) a += geo.dz3u[o + cw] * Gxw;
    if (ky) a += geo.dz3v[t] * Gyc;
    if (!kys) a += geo.dz3v[o + cs2] * Gys;
    gq[t] = a;
  }
  auto g = [&](int k, double* out) { out[0] = gq[(long long)k * n2 + c]; };
  const double* xs[1] = {q};
  double* ls[1] = {lam};
  rev_column_solve<1>(d, geo, c, Km, geo.dz3, p.thv_dt, p.thv_dt_drag, g, xs, ls, cs, gKm);
}

}  // namespace hoc

namespace hoc {

// =========================================================== closure reverse
// tke_step (closure.py:141-205, integral mixing lengths) reversed, in three
// launches over the columns' interfaces (layout [nz+1][ny][nx]):
//   k_rev_closure_col0  per column: N^2 on the interfaces (sn2), the depth
//       of each interface below the surface (szs, sequential sums) and the
//       end of the column's live interfaces (shi): crossings below the last
//       interface with N^2 > 0 or a thickness (the dry ones under the bottom)
//       change neither l nor the budget and carry a zero adjoint, so every
//       march and its reverse stop there;
//   k_rev_closure_rows  per interface: the centred-velocity differences
//       (sdu, sdv), the Gaspar lengths (slu, sld; the march stops once its
//       budget is spent) and the row of the implicit interface tridiagonal
//       (srow: sub, diag, sup, rhs);
//   k_rev_closure       per column: the solve (sx, before the floors), the
//       outputs' adjoints (e after its floors, K_m, K_h = (K_m - nu_b)/Pr +
//       kappa_b) back through the floors and K_m's factors, the transposed
//       solve and its band gradients, dissipation, production, the length
//       maxima/minima and the reverse march (march_step_adjoint) of every
//       interface, then N^2 back to b; the centred velocities' adjoints go
//       to guc, gvc (k_rev_closure_uv spreads them to u, v).
// Scratch: scs sweep coefficients, slam the adjoint right-hand side and
// lambda, sgl the adjoint of l_k, sn2b the adjoint of N^2.

struct ClosureCol {
  const Dims& d;
  const Geo& geo;
  long long c;
  __device__ long long T(int k) const { return (long long)k * d.n2 + c; }
  __device__ double dzi(int k) const {
    if (k == 0) return geo.dz3[c];
    if (k == d.nz) return geo.dz3[T(d.nz - 1)];
    return 0.5 * (geo.dz3[T(k - 1)] + geo.dz3[T(k)]);
  }
  __device__ double wet(int k) const {
    return (k >= 1 && k < d.nz) ? mk(geo.m3[T(k - 1)]) * mk(geo.m3[T(k)]) : 0.0;
  }
};

__device__ __forceinline__ double rc_ssqrt(double x) { return x > 0.0 ? ::sqrt(x) : 0.0; }
__device__ __forceinline__ double rc_dssqrt(double x) { return x > 0.0 ? 0.5 / ::sqrt(x) : 0.0; }

__global__ void k_rev_closure_col0(Dims d, Geo geo, const double* __restrict__ b,
                                   double* __restrict__ sn2, double* __restrict__ sn2b,
                                   double* __restrict__ szs, int* __restrict__ shi) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const ClosureCol C{d, geo, c};
  int hi = 0;
  double zs = 0.0;
  for (int k = 0; k <= d.nz; ++k) {
    if (k > 0) zs = zs + geo.dz3[C.T(k - 1)];
    szs[C.T(k)] = zs;
    double n2k = 0.0;
    const double dz = C.dzi(k);
    if (k >= 1 && k < d.nz) n2k = (b[C.T(k - 1)] - b[C.T(k)]) * safe_inv(dz) * C.wet(k);
    sn2[C.T(k)] = n2k;
    sn2b[C.T(k)] = 0.0;
    if (nmax(n2k, 0.0) != 0.0 || dz != 0.0) hi = k + 1;
  }
  shi[c] = hi;
}

__global__ void k_rev_closure_rows(Dims d, StepParams p, Geo geo, const double* __restrict__ e,
                                   const double* __restrict__ u, const double* __restrict__ v,
                                   const double* __restrict__ Km0,
                                   const double* __restrict__ Kh0,
                                   const double* __restrict__ kb,
                                   const double* __restrict__ sn2,
                                   const double* __restrict__ szs, const int* __restrict__ shi,
                                   double* __restrict__ sdu, double* __restrict__ sdv,
                                   double* __restrict__ slu, double* __restrict__ sld,
                                   double* __restrict__ srow) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.ni) return;
  const long long n2 = d.n2, ni = d.ni;
  const int nz = d.nz;
  const int k = (int)(t / n2);
  const long long c = t % n2;
  const ClosureCol C{d, geo, c};
  const double w = C.wet(k);
  double* sub = srow;
  double* diag = srow + ni;
  double* sup = srow + 2 * ni;
  double* rhs = srow + 3 * ni;
  if (!(w > 0.0)) {  // lengths carry the factor wet; boundary rows are fixed
    sdu[t] = 0.0;
    sdv[t] = 0.0;
    slu[t] = 0.0;
    sld[t] = 0.0;
    sub[t] = 0.0;
    diag[t] = 1.0;
    sup[t] = 0.0;
    rhs[t] = k == 0 ? (geo.m3[c] ? p.e_bb * geo.ustar2[c] : p.e_min) : p.e_min;
    return;
  }
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long cw = at2(d, j, wrap(i, -1, d.nx)), cs = at2(d, wrap(j, -1, d.ny), i);
  const long long a = (long long)(k - 1) * n2, o = (long long)k * n2;
  const double idzi = safe_inv(C.dzi(k));
  const double du = (0.5 * (u[a + c] + u[a + cw]) - 0.5 * (u[o + c] + u[o + cw])) * idzi;
  const double dv = (0.5 * (v[a + c] + v[a + cs]) - 0.5 * (v[o + c] + v[o + cs])) * idzi;
  sdu[t] = du;
  sdv[t] = dv;
  const int hi = shi[c];
  const double zs = szs[t], H = szs[(long long)nz * n2 + c];
  const double ek = e[t];
  const double e0 = nmax(ek, p.e_min);
  double l = 0.0, bud = e0;
  for (int kk = k - 1; kk >= 0 && bud > 0.0; --kk) {
    const MarchIn m = march_step(l, bud, nmax(sn2[C.T(kk)], 0.0), C.dzi(kk));
    l = m.l;
    bud = m.b;
  }
  const double lu = nmin(l, zs + p.z_0) * w;
  l = 0.0;
  bud = e0;
  for (int kk = k + 1; kk < hi && bud > 0.0; ++kk) {
    const MarchIn m = march_step(l, bud, nmax(sn2[C.T(kk)], 0.0), C.dzi(kk));
    l = m.l;
    bud = m.b;
  }
  const double ld = nmin(l, nmax(H - zs, 0.0) + p.z_0) * w;
  slu[t] = lu;
  sld[t] = ld;
  const double l_eps = nmax(p.l_min, nmin(lu, ld));
  const double diss = p.c_eps * rc_ssqrt(e0) / l_eps;
  const double sh2 = (du * du + dv * dv) * w;
  const double prod = (Km0[t] - p.nu_b) * sh2 - (Kh0[t] - kb[c]) * sn2[t];
  const double klu = 0.5 * (Km0[t - n2] + Km0[t]), kld = 0.5 * (Km0[t] + Km0[t + n2]);
  const double a_up = p.dt * klu * safe_inv(geo.dz3[t - n2]) * idzi;
  const double a_dn = p.dt * kld * safe_inv(geo.dz3[t]) * idzi;
  sub[t] = -a_up;
  sup[t] = -a_dn;
  diag[t] = 1.0 + a_up + a_dn + p.dt * diss;
  rhs[t] = ek + p.dt * prod;
}

__global__ void k_rev_closure(Dims d, StepParams p, Geo geo, const double* __restrict__ e,
                              const double* __restrict__ Km0, const double* __restrict__ Kh0,
                              const double* __restrict__ kb, const double* __restrict__ ge1,
                              const double* __restrict__ gKm1, const double* __restrict__ gKh1,
                              const double* __restrict__ sn2, const double* __restrict__ szs,
                              const int* __restrict__ shi, const double* __restrict__ sdu,
                              const double* __restrict__ sdv, const double* __restrict__ slu,
                              const double* __restrict__ sld, const double* __restrict__ srow,
                              double* __restrict__ sx, double* __restrict__ scs,
                              double* __restrict__ slam, double* __restrict__ sgl,
                              double* __restrict__ sn2b, double* __restrict__ ge0,
                              double* __restrict__ guc, double* __restrict__ gvc,
                              double* __restrict__ gb0, double* __restrict__ gKm0,
                              double* __restrict__ gKh0, double* __restrict__ gkb) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.n2) return;
  const long long n2 = d.n2, ni = d.ni;
  const int nz = d.nz;
  const ClosureCol C{d, geo, c};
  auto T = [&](int k) { return C.T(k); };
  const double* sub = srow;
  const double* diag = srow + ni;
  const double* sup = srow + 2 * ni;
  const double* rhs = srow + 3 * ni;
  const double kbc = kb[c];
  const int hi = shi[c];
  const double H = szs[T(nz)];
  // the forward solve of the interface tridiagonal
  {
    double cp = 0.0, dp = 0.0;
    for (int k = 0; k <= nz; ++k) {
      const long long t = T(k);
      const double den = k == 0 ? diag[t] : diag[t] - sub[t] * cp;
      cp = sup[t] / den;
      dp = k == 0 ? rhs[t] / den : (rhs[t] - sub[t] * dp) / den;
      scs[t] = cp;
      sx[t] = dp;
    }
    double xn = sx[T(nz)];
    for (int k = nz - 1; k >= 0; --k) {
      xn = sx[T(k)] - scs[T(k)] * xn;
      sx[T(k)] = xn;
    }
  }
  // the outputs' adjoints back to the solve's
  double gkb_acc = 0.0;
  for (int k = 0; k <= nz; ++k) {
    const long long t = T(k);
    const double gK = gKm1[t] + gKh1[t] / p.pr_t;
    gkb_acc += gKh1[t];
    const double w = C.wet(k);
    const double x = sx[t];
    const double e_new = k == 0 ? rhs[t] : (w > 0.0 ? nmax(x, p.e_min) : p.e_min);
    const double lu = slu[t], ld = sld[t];
    const double l_k = nmax(p.l_min, rc_ssqrt(lu * ld));
    sgl[t] = p.c_k * rc_ssqrt(e_new) * w * gK;
    const double ge_new = (ge1 ? ge1[t] : 0.0) + p.c_k * l_k * w * gK * rc_dssqrt(e_new);
    slam[t] = (k >= 1 && w > 0.0 && (x >= p.e_min || x != x)) ? ge_new : 0.0;
  }
  {  // lambda = A^-T g: sub_T[k] = sup[k-1], sup_T[k] = sub[k+1]
    double cp = 0.0, dp = 0.0;
    for (int k = 0; k <= nz; ++k) {
      const long long t = T(k);
      const double supT = k < nz ? sub[t + n2] : 0.0;
      const double den = k == 0 ? diag[t] : diag[t] - sup[t - n2] * cp;
      const double cc = supT / den;
      const double dd = k == 0 ? slam[t] / den : (slam[t] - sup[t - n2] * dp) / den;
      scs[t] = cc;
      slam[t] = dd;
      cp = cc;
      dp = dd;
    }
    double ln = slam[T(nz)];
    for (int k = nz - 1; k >= 0; --k) {
      ln = slam[T(k)] - scs[T(k)] * ln;
      slam[T(k)] = ln;
    }
  }
  // per interface: the row's band and right-hand-side adjoints, then the
  // lengths and their reverse marches. The adjoints of K_m (interfaces
  // k-1, k, k+1) and of the centred velocities (layers k-1, k) are carried
  // in registers and each address is written once, the terms added in the
  // same order as a running sum per address would add them.
  double km_prev = 0.0, km_cur = 0.0;  // K_m adjoint of interfaces k-1 and k so far
  double uc_prev = 0.0, vc_prev = 0.0;  // centred-velocity adjoints of layer k-1 so far
  for (int k = 0; k <= nz; ++k) {
    const long long t = T(k);
    const double w = C.wet(k);
    const double lu = slu[t], ld = sld[t];
    double gl_eps = 0.0;
    const double ek = e[t];
    double gek = 0.0;
    double km_next = 0.0, uc_cur = 0.0, vc_cur = 0.0, gkh = 0.0;
    if (w > 0.0) {
      const double lam = slam[t];
      const double gd = -lam * sx[t], gs = -lam * sx[t - n2], gp = -lam * sx[t + n2];
      const double ga_up = gd - gs, ga_dn = gd - gp, gdiss = p.dt * gd;
      const double idzi = safe_inv(C.dzi(k));
      const double glu = p.dt * safe_inv(geo.dz3[t - n2]) * idzi * ga_up;
      const double gld = p.dt * safe_inv(geo.dz3[t]) * idzi * ga_dn;
      km_prev += 0.5 * glu;
      km_cur += 0.5 * glu + 0.5 * gld;
      km_next += 0.5 * gld;
      gek += lam;
      const double gprod = p.dt * lam;
      const double n2k = sn2[t];
      const double du = sdu[t], dv = sdv[t];
      const double sh2 = (du * du + dv * dv) * w;
      km_cur += sh2 * gprod;
      const double gsh = (Km0[t] - p.nu_b) * gprod;
      gkh += -n2k * gprod;
      gkb_acc += n2k * gprod;
      sn2b[t] += -(Kh0[t] - kbc) * gprod;
      const double gdu = 2.0 * du * w * gsh * idzi, gdv = 2.0 * dv * w * gsh * idzi;
      uc_prev += gdu;
      uc_cur -= gdu;
      vc_prev += gdv;
      vc_cur -= gdv;
      // dissipation c_eps sqrt(max(e, e_min)) / l_eps
      const double E = nmax(ek, p.e_min);
      const double l_eps = nmax(p.l_min, nmin(lu, ld));
      const double diss = p.c_eps * rc_ssqrt(E) / l_eps;
      if (ek >= p.e_min || ek != ek) gek += gdiss * p.c_eps * rc_dssqrt(E) / l_eps;
      gl_eps = -gdiss * diss / l_eps;
    }
    // l_eps = max(l_min, min(l_up, l_dn)), l_k = max(l_min, sqrt(l_up l_dn))
    double glu = 0.0, gld = 0.0;
    {
      const double X = nmin(lu, ld);
      const double gX = (X <= p.l_min) ? 0.0 : gl_eps;
      if (lu <= ld || lu != lu) glu += gX;
      else gld += gX;
      const double P = lu * ld;
      const double gY = (rc_ssqrt(P) <= p.l_min) ? 0.0 : sgl[t];
      const double gP = gY * rc_dssqrt(P);
      glu += gP * ld;
      gld += gP * lu;
    }
    // reverse marches of interface k (closure.py:65-103)
    const double ek0 = nmax(ek, p.e_min);
    const double zs = szs[t];
    double gb0_acc = 0.0;
    for (int dir = 0; dir < 2; ++dir) {
      const double gout = (dir == 0 ? glu : gld) * w;
      if (gout == 0.0) continue;
      const double cap = (dir == 0 ? zs : nmax(H - zs, 0.0)) + p.z_0;
      // the march forward to its end, then each crossing recomputed from the
      // start for its reverse (the marches here stop within a few crossings)
      int nstep = 0;
      double l = 0.0, bud = ek0;
      for (int kk = dir == 0 ? k - 1 : k + 1; (dir == 0 ? kk >= 0 : kk < hi) && bud > 0.0;
           kk += dir == 0 ? -1 : 1) {
        const MarchIn o = march_step(l, bud, nmax(sn2[T(kk)], 0.0), C.dzi(kk));
        l = o.l;
        bud = o.b;
        ++nstep;
      }
      double lb = (l <= cap || l != l) ? gout : 0.0, bb = 0.0;
      for (int s = nstep; s >= 1; --s) {
        double ls = 0.0, bs = ek0;
        int kk = dir == 0 ? k - 1 : k + 1;
        for (int r = 1; r < s; ++r) {
          const MarchIn o = march_step(ls, bs, nmax(sn2[T(kk)], 0.0), C.dzi(kk));
          ls = o.l;
          bs = o.b;
          kk += dir == 0 ? -1 : 1;
        }
        const double n2r = sn2[T(kk)];
        double lb_in, bb_in, nb = 0.0;
        march_step_adjoint(ls, bs, nmax(n2r, 0.0), C.dzi(kk), lb, bb, lb_in, bb_in, nb);
        lb = lb_in;
        bb = bb_in;
        if (n2r >= 0.0 || n2r != n2r) sn2b[T(kk)] += nb;
      }
      gb0_acc += bb;
    }
    if (ek >= p.e_min || ek != ek) gek += gb0_acc;
    ge0[t] = 0.0 + gek;
    gKh0[t] = gkh;
    if (k >= 1) {  // interface k-1 and layer k-1 have all their terms
      gKm0[t - n2] = km_prev;
      guc[t - n2] = uc_prev;
      gvc[t - n2] = vc_prev;
    }
    km_prev = km_cur;
    km_cur = km_next;
    uc_prev = uc_cur;
    vc_prev = vc_cur;
  }
  gKm0[T(nz)] = km_prev;
  gkb[c] += gkb_acc;
  // N^2 = (b[k-1] - b[k]) / dzi wet back to b
  for (int k = 1; k < nz; ++k) {
    const double gdb = sn2b[T(k)] * C.wet(k) * safe_inv(C.dzi(k));
    gb0[T(k - 1)] += gdb;
    gb0[T(k)] -= gdb;
  }
}

// The centred velocities' adjoints (uc = (u + u_W) / 2, vc = (v + v_S) / 2)
// added to u and v.
__global__ void k_rev_closure_uv(Dims d, const double* __restrict__ guc,
                                 const double* __restrict__ gvc, double* __restrict__ gu,
                                 double* __restrict__ gv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.n3) return;
  const long long o = (t / d.n2) * d.n2, c = t % d.n2;
  const int i = (int)(c % d.nx), j = (int)(c / d.nx);
  const long long te = o + at2(d, j, wrap(i, 1, d.nx)), tn = o + at2(d, wrap(j, 1, d.ny), i);
  gu[t] += 0.5 * guc[t] + 0.5 * guc[te];
  gv[t] += 0.5 * gvc[t] + 0.5 * gvc[tn];
}

}  // namespace hoc
