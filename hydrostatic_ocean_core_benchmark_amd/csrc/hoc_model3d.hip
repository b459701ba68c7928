// The reference's v0.2 3D hydrostatic model (libs/core/model3d.py:93-391)
// on the device: flat bottom, nz uniform layers of thickness dz, doubly
// periodic C-grid; fields [nz][ny][nx] (w on nz + 1 interfaces, k = 0 the
// surface). Column work (potential, depth sums, w, the predictor with its
// tridiagonal solve, tracers) runs one thread per column, stencil work one
// thread per cell. The Thomas factors are column-independent (uniform
// viscosity and diffusivity) and come from the host. Every expression keeps
// the operand order of the NumPy expression it replaces and sums over layers
// accumulate from the surface down, as NumPy's axis-0 reductions do, so
// states are bitwise the reference's.
#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

namespace {

struct Col {
  int i, j, ie, iw, jn, js;
};

__device__ __forceinline__ Col m3_col(const M3& M, long long c) {
  Col q;
  q.i = (int)(c % M.nx);
  q.j = (int)(c / M.nx);
  q.ie = wrap(q.i, 1, M.nx);
  q.iw = wrap(q.i, -1, M.nx);
  q.jn = wrap(q.j, 1, M.ny);
  q.js = wrap(q.j, -1, M.ny);
  return q;
}

__device__ __forceinline__ long long at2(const M3& M, int j, int i) {
  return (long long)j * M.nx + i;
}

__device__ __forceinline__ long long at3(const M3& M, int k, int j, int i) {
  return (long long)k * M.n2 + (long long)j * M.nx + i;
}

// A_u[v] = 1/4 (v + v_s + v_e + v_se)   (operators.py:37-40)
__device__ __forceinline__ double m3_avg_v_to_u(const M3& M, const double* v, int k, const Col& c) {
  return 0.25 * (v[at3(M, k, c.j, c.i)] + v[at3(M, k, c.js, c.i)] + v[at3(M, k, c.j, c.ie)] +
                 v[at3(M, k, c.js, c.ie)]);
}

// A_v[u] = 1/4 (u + u_w + u_n + u_nw)   (operators.py:43-46)
__device__ __forceinline__ double m3_avg_u_to_v(const M3& M, const double* u, int k, const Col& c) {
  return 0.25 * (u[at3(M, k, c.j, c.i)] + u[at3(M, k, c.j, c.iw)] + u[at3(M, k, c.jn, c.i)] +
                 u[at3(M, k, c.jn, c.iw)]);
}

// laplacian_h (advection.py:97-106)
__device__ __forceinline__ double m3_lap_h(const M3& M, const double* a, int k, const Col& c) {
  const double ac = a[at3(M, k, c.j, c.i)];
  const double two_a = 2.0 * ac;
  return (a[at3(M, k, c.j, c.ie)] - two_a + a[at3(M, k, c.j, c.iw)]) / M.dx2 +
         (a[at3(M, k, c.jn, c.i)] - two_a + a[at3(M, k, c.js, c.i)]) / M.dy2;
}

// apply_diffusion(a, nu, dz, top_flux, bottom_drag) at layer k of one
// column (vertical.py:75-86): (flux[k] - flux[k+1]) / dz with interior
// fluxes nu (a[k-1] - a[k]) / dz, flux[0] = top_flux, flux[nz] = drag a[nz-1]
__device__ __forceinline__ double m3_vdiff(const M3& M, const double* a, long long col, int k,
                                           double nu, double top_flux, double drag) {
  const double ak = a[(long long)k * M.n2 + col];
  const double f_top = k == 0 ? top_flux : nu * (a[(long long)(k - 1) * M.n2 + col] - ak) / M.dz;
  const double f_bot = k == M.nz - 1 ? drag * ak : nu * (ak - a[(long long)(k + 1) * M.n2 + col]) / M.dz;
  return (f_top - f_bot) / M.dz;
}

// Momentum advection, centred flux form (advection.py:63-94), at layer k.
__device__ __forceinline__ double m3_adv_u(const M3& M, const double* u, const double* v,
                                           const double* w, int k, const Col& c) {
  const double uc = u[at3(M, k, c.j, c.i)];
  const double ubar_c = 0.5 * (u[at3(M, k, c.j, c.iw)] + uc);
  const double ubar_ce = 0.5 * (uc + u[at3(M, k, c.j, c.ie)]);
  const double adv_ux = (ubar_ce * ubar_ce - ubar_c * ubar_c) / M.dx;
  const double fy_n = (0.5 * (v[at3(M, k, c.j, c.i)] + v[at3(M, k, c.j, c.ie)])) *
                      (0.5 * (uc + u[at3(M, k, c.jn, c.i)]));
  const double fy_s = (0.5 * (v[at3(M, k, c.js, c.i)] + v[at3(M, k, c.js, c.ie)])) *
                      (0.5 * (u[at3(M, k, c.js, c.i)] + uc));
  const double adv_uy = (fy_n - fy_s) / M.dy;
  double fz_top = 0.0, fz_bot = 0.0;
  if (k > 0) {
    const double wb = 0.5 * (w[at3(M, k, c.j, c.i)] + w[at3(M, k, c.j, c.ie)]);
    fz_top = wb * 0.5 * (u[at3(M, k - 1, c.j, c.i)] + uc);
  }
  if (k < M.nz - 1) {
    const double wb = 0.5 * (w[at3(M, k + 1, c.j, c.i)] + w[at3(M, k + 1, c.j, c.ie)]);
    fz_bot = wb * 0.5 * (uc + u[at3(M, k + 1, c.j, c.i)]);
  }
  const double adv_uz = (fz_top - fz_bot) / M.dz;
  return adv_ux + adv_uy + adv_uz;
}

__device__ __forceinline__ double m3_adv_v(const M3& M, const double* u, const double* v,
                                           const double* w, int k, const Col& c) {
  const double vc = v[at3(M, k, c.j, c.i)];
  const double vbar_c = 0.5 * (v[at3(M, k, c.js, c.i)] + vc);
  const double vbar_cn = 0.5 * (vc + v[at3(M, k, c.jn, c.i)]);
  const double adv_vy = (vbar_cn * vbar_cn - vbar_c * vbar_c) / M.dy;
  const double fx_e = (0.5 * (u[at3(M, k, c.j, c.i)] + u[at3(M, k, c.jn, c.i)])) *
                      (0.5 * (vc + v[at3(M, k, c.j, c.ie)]));
  const double fx_w = (0.5 * (u[at3(M, k, c.j, c.iw)] + u[at3(M, k, c.jn, c.iw)])) *
                      (0.5 * (v[at3(M, k, c.j, c.iw)] + vc));
  const double adv_vx = (fx_e - fx_w) / M.dx;
  double fz_top = 0.0, fz_bot = 0.0;
  if (k > 0) {
    const double wb = 0.5 * (w[at3(M, k, c.j, c.i)] + w[at3(M, k, c.jn, c.i)]);
    fz_top = wb * 0.5 * (v[at3(M, k - 1, c.j, c.i)] + vc);
  }
  if (k < M.nz - 1) {
    const double wb = 0.5 * (w[at3(M, k + 1, c.j, c.i)] + w[at3(M, k + 1, c.jn, c.i)]);
    fz_bot = wb * 0.5 * (vc + v[at3(M, k + 1, c.j, c.i)]);
  }
  const double adv_vz = (fz_top - fz_bot) / M.dz;
  return adv_vx + adv_vy + adv_vz;
}

// Tracer value on a face (advection.py:17-36): centred mean, or upwind by
// the sign of the face velocity.
__device__ __forceinline__ double m3_face(int scheme, double vel, double here, double next) {
  if (scheme == 1) return 0.5 * (here + next);
  return vel > 0.0 ? here : next;
}

// DIV(u C) in flux form (advection.py:39-60) at layer k.
__device__ __forceinline__ double m3_adv_c(const M3& M, const double* cf, const double* u,
                                           const double* v, const double* w, int k, const Col& q) {
  const double cc = cf[at3(M, k, q.j, q.i)];
  const double ue = u[at3(M, k, q.j, q.i)], uw = u[at3(M, k, q.j, q.iw)];
  const double vn = v[at3(M, k, q.j, q.i)], vs = v[at3(M, k, q.js, q.i)];
  const double fx = ue * m3_face(M.adv, ue, cc, cf[at3(M, k, q.j, q.ie)]);
  const double fxw = uw * m3_face(M.adv, uw, cf[at3(M, k, q.j, q.iw)], cc);
  const double fy = vn * m3_face(M.adv, vn, cc, cf[at3(M, k, q.jn, q.i)]);
  const double fys = vs * m3_face(M.adv, vs, cf[at3(M, k, q.js, q.i)], cc);
  double fz_top = 0.0, fz_bot = 0.0;
  if (k > 0) {
    const double wk = w[at3(M, k, q.j, q.i)];
    const double above = cf[at3(M, k - 1, q.j, q.i)];
    fz_top = wk * (M.adv == 1 ? 0.5 * (above + cc) : (wk > 0.0 ? cc : above));
  }
  if (k < M.nz - 1) {
    const double wk = w[at3(M, k + 1, q.j, q.i)];
    const double below = cf[at3(M, k + 1, q.j, q.i)];
    fz_bot = wk * (M.adv == 1 ? 0.5 * (cc + below) : (wk > 0.0 ? below : cc));
  }
  return (fx - fxw) / M.dx + (fy - fys) / M.dy + (fz_top - fz_bot) / M.dz;
}

}  // namespace

// Phi' = Phi - <Phi>_z with Phi[k] = dz (0.5 b[k] + sum_{k'<k} b[k'])
// (vertical.py:99-111, model3d.py:158-161); one thread per column.
__global__ void k_m3_phi(M3 M, const double* __restrict__ b, double* __restrict__ phi) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  double above = 0.0, sum = 0.0;
  for (int k = 0; k < M.nz; ++k) {
    const double bk = b[(long long)k * M.n2 + col];
    const double p = M.dz * (0.5 * bk + above);
    phi[(long long)k * M.n2 + col] = p;
    sum = k == 0 ? p : sum + p;
    above = k == 0 ? bk : above + bk;
  }
  const double mean = sum / (double)M.nz;
  for (int k = 0; k < M.nz; ++k) {
    const long long t = (long long)k * M.n2 + col;
    phi[t] = phi[t] - mean;
  }
}

// U = dz sum_k u, V = dz sum_k v
__global__ void k_m3_depth(M3 M, const double* __restrict__ u, const double* __restrict__ v,
                           double* __restrict__ U, double* __restrict__ V) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  double su = u[col], sv = v[col];
  for (int k = 1; k < M.nz; ++k) {
    su = su + u[(long long)k * M.n2 + col];
    sv = sv + v[(long long)k * M.n2 + col];
  }
  U[col] = M.dz * su;
  V[col] = M.dz * sv;
}

// w[k] = -dz sum_{k'>=k} D[u, v][k'], w[nz] = 0 (vertical.py:89-96)
__global__ void k_m3_w(M3 M, const double* __restrict__ u, const double* __restrict__ v,
                       double* __restrict__ w) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  const Col c = m3_col(M, col);
  double acc = 0.0;
  for (int k = M.nz - 1; k >= 0; --k) {
    const double d = (u[at3(M, k, c.j, c.i)] - u[at3(M, k, c.j, c.iw)]) / M.dx +
                     (v[at3(M, k, c.j, c.i)] - v[at3(M, k, c.js, c.i)]) / M.dy;
    acc = k == M.nz - 1 ? d : acc + d;
    w[(long long)k * M.n2 + col] = M.neg_dz * acc;
  }
  w[(long long)M.nz * M.n2 + col] = 0.0;
}

// The explicit stress terms of the predictor (model3d.py:166-180, 335-343):
// st = [theta_v dt tau / dz at the surface] + dt (-advection + A_h lap_h)
__global__ void k_m3_stress(M3 M, const double* __restrict__ u, const double* __restrict__ v,
                            const double* __restrict__ w, double* __restrict__ st_u,
                            double* __restrict__ st_v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.n3) return;
  const int k = (int)(t / M.n2);
  const Col c = m3_col(M, t % M.n2);
  double ex_u = 0.0, ex_v = 0.0;
  if (M.adv != 0) {
    ex_u = ex_u - m3_adv_u(M, u, v, w, k, c);
    ex_v = ex_v - m3_adv_v(M, u, v, w, k, c);
  }
  if (M.A_h > 0.0) {
    ex_u = ex_u + M.A_h * m3_lap_h(M, u, k, c);
    ex_v = ex_v + M.A_h * m3_lap_h(M, v, k, c);
  }
  st_u[t] = (k == 0 ? M.s0u : 0.0) + M.dt * ex_u;
  st_v[t] = (k == 0 ? M.s0v : 0.0) + M.dt * ex_v;
}

// One Picard pass of the baroclinic predictor and its implicit vertical
// viscosity (model3d.py:352-364; split form :276-288), one thread per
// column: gu/gv per layer, the tridiagonal solve with the momentum factors,
// and the depth sums of the result (and, split, of f v_cor and f u_cor).
__global__ void k_m3_predict(M3 M, int split, const double* __restrict__ u0,
                             const double* __restrict__ v0, const double* __restrict__ phi,
                             const double* __restrict__ eta, const double* __restrict__ u_it,
                             const double* __restrict__ v_it, const double* __restrict__ st_u,
                             const double* __restrict__ st_v, double* __restrict__ gh_u,
                             double* __restrict__ gh_v, double* __restrict__ GU,
                             double* __restrict__ GV, double* __restrict__ cor_x,
                             double* __restrict__ cor_y) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  const Col c = m3_col(M, col);
  const double gx_eta = split ? 0.0 : (eta[at2(M, c.j, c.ie)] - eta[col]) / M.dx;
  const double gy_eta = split ? 0.0 : (eta[at2(M, c.jn, c.i)] - eta[col]) / M.dy;
  double du = 0.0, dv = 0.0, sfv = 0.0, sfu = 0.0;
  for (int k = 0; k < M.nz; ++k) {
    const long long t = (long long)k * M.n2 + col;
    const double v_cor = M.one_m_tc * m3_avg_v_to_u(M, v0, k, c) + M.tc * m3_avg_v_to_u(M, v_it, k, c);
    const double u_cor = M.one_m_tc * m3_avg_u_to_v(M, u0, k, c) + M.tc * m3_avg_u_to_v(M, u_it, k, c);
    const double pgx = (phi[at3(M, k, c.j, c.ie)] - phi[t]) / M.dx;
    const double pgy = (phi[at3(M, k, c.jn, c.i)] - phi[t]) / M.dy;
    const double dz_u = m3_vdiff(M, u0, col, k, M.nu, M.tau_u, M.bd);
    const double dz_v = m3_vdiff(M, v0, col, k, M.nu, M.tau_v, M.bd);
    double gu, gv;
    if (split) {
      gu = u0[t] + M.dt * (M.f * v_cor + pgx) + M.c_diff * dz_u;
      gv = v0[t] + M.dt * (M.neg_f * u_cor + pgy) + M.c_diff * dz_v;
      const double fv = M.f * v_cor, fu = M.f * u_cor;
      sfv = k == 0 ? fv : sfv + fv;
      sfu = k == 0 ? fu : sfu + fu;
    } else {
      gu = u0[t] + M.dt * (M.f * v_cor + pgx) - M.c_eta * gx_eta + M.c_diff * dz_u;
      gv = v0[t] + M.dt * (M.neg_f * u_cor + pgy) - M.c_eta * gy_eta + M.c_diff * dz_v;
    }
    const double ru = gu + st_u[t], rv = gv + st_v[t];
    // forward elimination (vertical.py:31-36)
    du = k == 0 ? ru / M.den_m[0] : (ru - M.sub_m[k] * du) / M.den_m[k];
    dv = k == 0 ? rv / M.den_m[0] : (rv - M.sub_m[k] * dv) / M.den_m[k];
    gh_u[t] = du;
    gh_v[t] = dv;
  }
  double xu = du, xv = dv;
  for (int k = M.nz - 2; k >= 0; --k) {
    const long long t = (long long)k * M.n2 + col;
    xu = gh_u[t] - M.cst_m[k] * xu;
    xv = gh_v[t] - M.cst_m[k] * xv;
    gh_u[t] = xu;
    gh_v[t] = xv;
  }
  double su = gh_u[col], sv = gh_v[col];
  for (int k = 1; k < M.nz; ++k) {
    su = su + gh_u[(long long)k * M.n2 + col];
    sv = sv + gh_v[(long long)k * M.n2 + col];
  }
  GU[col] = M.dz * su;
  GV[col] = M.dz * sv;
  if (split) {
    cor_x[col] = M.dz * sfv;
    cor_y[col] = M.neg_dz * sfu;
  }
}

// rhs = eta - dt ((1 - th) D[U0, V0] + th D[GU, GV])   (model3d.py:367-369)
__global__ void k_m3_rhs(M3 M, const double* __restrict__ eta, const double* __restrict__ U0,
                         const double* __restrict__ V0, const double* __restrict__ GU,
                         const double* __restrict__ GV, double* __restrict__ rhs) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  const Col c = m3_col(M, col);
  const double d_old = (U0[col] - U0[at2(M, c.j, c.iw)]) / M.dx + (V0[col] - V0[at2(M, c.js, c.i)]) / M.dy;
  const double d_g = (GU[col] - GU[at2(M, c.j, c.iw)]) / M.dx + (GV[col] - GV[at2(M, c.js, c.i)]) / M.dy;
  rhs[col] = eta[col] - M.dt * (M.one_m_th * d_old + M.th * d_g);
}

// u = gh_u - g dt th q Gx eta_new, v likewise   (model3d.py:374-375)
__global__ void k_m3_backsub(M3 M, const double* __restrict__ gh_u, const double* __restrict__ gh_v,
                             const double* __restrict__ eta, double* __restrict__ u,
                             double* __restrict__ v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.n3) return;
  const int k = (int)(t / M.n2);
  const long long col = t % M.n2;
  const Col c = m3_col(M, col);
  const double cq = M.c_bs * M.q[k];
  u[t] = gh_u[t] - cq * ((eta[at2(M, c.j, c.ie)] - eta[col]) / M.dx);
  v[t] = gh_v[t] - cq * ((eta[at2(M, c.jn, c.i)] - eta[col]) / M.dy);
}

// eta' = eta - dt D[U, V]   (model3d.py:379-381; barotropic substep :235)
__global__ void k_m3_eta_div(M3 M, double dt, const double* __restrict__ eta,
                             const double* __restrict__ U, const double* __restrict__ V,
                             double* __restrict__ out) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  const Col c = m3_col(M, col);
  out[col] = eta[col] - dt * ((U[col] - U[at2(M, c.j, c.iw)]) / M.dx + (V[col] - V[at2(M, c.js, c.i)]) / M.dy);
}

// fx = (GU - U0) / dt - cor_x, fy likewise   (model3d.py:292-293)
__global__ void k_m3_forcing(M3 M, const double* __restrict__ GU, const double* __restrict__ GV,
                             const double* __restrict__ U0, const double* __restrict__ V0,
                             const double* __restrict__ cor_x, const double* __restrict__ cor_y,
                             double* __restrict__ fx, double* __restrict__ fy) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  fx[col] = (GU[col] - U0[col]) / M.dt - cor_x[col];
  fy[col] = (GV[col] - V0[col]) / M.dt - cor_y[col];
}

// One Picard pass of a barotropic substep (model3d.py:224-232) on the 2D
// transports: ui = u_i + ddt (f v_cor - g H Gx e + fx), vi likewise.
__global__ void k_m3_bt_momentum(M3 M, const double* __restrict__ u_i, const double* __restrict__ v_i,
                                 const double* __restrict__ e, const double* __restrict__ ui,
                                 const double* __restrict__ vi, const double* __restrict__ fx,
                                 const double* __restrict__ fy, double* __restrict__ ui_out,
                                 double* __restrict__ vi_out) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  const Col c = m3_col(M, col);
  const double v_cor = M.one_m_tc * m3_avg_v_to_u(M, v_i, 0, c) + M.tc * m3_avg_v_to_u(M, vi, 0, c);
  const double u_cor = M.one_m_tc * m3_avg_u_to_v(M, u_i, 0, c) + M.tc * m3_avg_u_to_v(M, ui, 0, c);
  const double gxe = (e[at2(M, c.j, c.ie)] - e[col]) / M.dx;
  const double gye = (e[at2(M, c.jn, c.i)] - e[col]) / M.dy;
  ui_out[col] = u_i[col] + M.ddt * (M.f * v_cor - M.gh * gxe + fx[col]);
  vi_out[col] = v_i[col] + M.ddt * (M.neg_f * u_cor - M.gh * gye + fy[col]);
}

// acc += u_i, acc_v += v_i   (model3d.py:236-237)
__global__ void k_m3_bt_acc(M3 M, const double* __restrict__ u_i, const double* __restrict__ v_i,
                            double* __restrict__ acc_u, double* __restrict__ acc_v) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  acc_u[col] = acc_u[col] + u_i[col];
  acc_v[col] = acc_v[col] + v_i[col];
}

// u = gh_u - <gh_u>_z + U / H with U the last substep or the substep mean
// (model3d.py:241-246, 297-300); one thread per column.
__global__ void k_m3_split_final(M3 M, int mean_n, const double* __restrict__ gh_u,
                                 const double* __restrict__ gh_v, const double* __restrict__ Ub,
                                 const double* __restrict__ Vb, double* __restrict__ u,
                                 double* __restrict__ v) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  double su = gh_u[col], sv = gh_v[col];
  for (int k = 1; k < M.nz; ++k) {
    su = su + gh_u[(long long)k * M.n2 + col];
    sv = sv + gh_v[(long long)k * M.n2 + col];
  }
  const double mu = su / (double)M.nz, mv = sv / (double)M.nz;
  const double Uc = mean_n > 0 ? Ub[col] / (double)mean_n : Ub[col];
  const double Vc = mean_n > 0 ? Vb[col] / (double)mean_n : Vb[col];
  const double bu = Uc / M.H, bv = Vc / M.H;
  for (int k = 0; k < M.nz; ++k) {
    const long long t = (long long)k * M.n2 + col;
    u[t] = gh_u[t] - mu + bu;
    v[t] = gh_v[t] - mv + bv;
  }
}

// One tracer step (model3d.py:182-198): explicit advection, horizontal
// diffusion and source, the theta_v-split vertical diffusion with the
// surface flux, then the tridiagonal solve with the tracer factors; one
// thread per column. which: 0 buoyancy (source -N2 w at centres, no surface
// flux), 1 temperature, 2 salinity.
__global__ void k_m3_tracer(M3 M, int which, const double* __restrict__ cf,
                            const double* __restrict__ u, const double* __restrict__ v,
                            const double* __restrict__ w, double* __restrict__ out) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= M.n2) return;
  const Col q = m3_col(M, col);
  const double flux = which == 1 ? M.flux_T : (which == 2 ? M.flux_S : 0.0);
  const double s0 = which == 1 ? M.s0T : (which == 2 ? M.s0S : M.s0b);
  double d = 0.0;
  for (int k = 0; k < M.nz; ++k) {
    const long long t = (long long)k * M.n2 + col;
    double tend = which == 0
                      ? M.neg_N2 * (0.5 * (w[t] + w[(long long)(k + 1) * M.n2 + col]))
                      : 0.0;
    if (M.adv != 0) tend = tend - m3_adv_c(M, cf, u, v, w, k, q);
    if (M.K_h > 0.0) tend = tend + M.K_h * m3_lap_h(M, cf, k, q);
    double rhs = cf[t] + M.dt * tend + M.c_diff * m3_vdiff(M, cf, col, k, M.kappa, flux, 0.0);
    if (k == 0) rhs = rhs + s0;
    d = k == 0 ? rhs / M.den_b[0] : (rhs - M.sub_b[k] * d) / M.den_b[k];
    out[t] = d;
  }
  double x = d;
  for (int k = M.nz - 2; k >= 0; --k) {
    const long long t = (long long)k * M.n2 + col;
    x = out[t] - M.cst_b[k] * x;
    out[t] = x;
  }
}

// b = g (alpha_T (T - T0) - beta_S (S - S0))   (model3d.py:84-88)
__global__ void k_m3_eos(M3 M, const double* __restrict__ T, const double* __restrict__ S,
                         double* __restrict__ b) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.n3) return;
  b[t] = M.g * (M.alpha_T * (T[t] - M.T0) - M.beta_S * (S[t] - M.S0));
}

}  // namespace hoc
