// The 2D rotating shallow-water step of the reference's original (v0.1)
// model, libs/core/schemes.py:97-177, on a doubly periodic C-grid with
// constant depth H: forward-backward and theta (Casulli) schemes with a
// Picard loop on the Coriolis term (the Helmholtz solves are in
// hoc_helm2d.hip). One thread per cell; every expression keeps the
// reference's operand order (host constants are formed in Python's
// left-to-right order), so results are bitwise the NumPy model's.
#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

__device__ __forceinline__ void sw_ij(const Sw2d& S, long long t, int& i, int& j) {
  i = (int)(t % S.nx);
  j = (int)(t / S.nx);
}

// A_u[v] = 1/4 (v[j,i] + v[j-1,i] + v[j,i+1] + v[j-1,i+1]) (operators.py:37-40)
__device__ __forceinline__ double sw_avg_v_to_u(const Sw2d& S, const double* v, int j, int i) {
  const int js = wrap(j, -1, S.ny), ie = wrap(i, 1, S.nx);
  return 0.25 * (v[(long long)j * S.nx + i] + v[(long long)js * S.nx + i] +
                 v[(long long)j * S.nx + ie] + v[(long long)js * S.nx + ie]);
}

// A_v[u] = 1/4 (u[j,i] + u[j,i-1] + u[j+1,i] + u[j+1,i-1]) (operators.py:43-46)
__device__ __forceinline__ double sw_avg_u_to_v(const Sw2d& S, const double* u, int j, int i) {
  const int iw = wrap(i, -1, S.nx), jn = wrap(j, 1, S.ny);
  return 0.25 * (u[(long long)j * S.nx + i] + u[(long long)j * S.nx + iw] +
                 u[(long long)jn * S.nx + i] + u[(long long)jn * S.nx + iw]);
}

__device__ __forceinline__ double sw_gradx(const Sw2d& S, const double* e, int j, int i) {
  return (e[(long long)j * S.nx + wrap(i, 1, S.nx)] - e[(long long)j * S.nx + i]) / S.dx;
}

__device__ __forceinline__ double sw_grady(const Sw2d& S, const double* e, int j, int i) {
  return (e[(long long)wrap(j, 1, S.ny) * S.nx + i] - e[(long long)j * S.nx + i]) / S.dy;
}

// D[u, v] at eta points (operators.py:32-34)
__device__ __forceinline__ double sw_div(const Sw2d& S, const double* u, const double* v, int j,
                                         int i) {
  const long long c = (long long)j * S.nx + i;
  return (u[c] - u[(long long)j * S.nx + wrap(i, -1, S.nx)]) / S.dx +
         (v[c] - v[(long long)wrap(j, -1, S.ny) * S.nx + i]) / S.dy;
}

// One Picard update of the momentum (schemes.py:127-135 fb, 149-152 theta):
// fb:    u' = u0 + dt (f v_cor - g Gx eta0),   v' = v0 + dt (-f u_cor - g Gy eta0)
// theta: gu = u0 + dt f v_cor - g dt (1 - th) Gx eta0, gv likewise with -dt f u_cor
__global__ void k_sw_momentum(Sw2d S, int theta, const double* __restrict__ u0,
                              const double* __restrict__ v0, const double* __restrict__ eta0,
                              const double* __restrict__ u_it, const double* __restrict__ v_it,
                              double* __restrict__ out_u, double* __restrict__ out_v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.n) return;
  int i, j;
  sw_ij(S, t, i, j);
  const double v_cor = S.one_m_tc * sw_avg_v_to_u(S, v0, j, i) + S.tc * sw_avg_v_to_u(S, v_it, j, i);
  const double u_cor = S.one_m_tc * sw_avg_u_to_v(S, u0, j, i) + S.tc * sw_avg_u_to_v(S, u_it, j, i);
  const double gx = sw_gradx(S, eta0, j, i), gy = sw_grady(S, eta0, j, i);
  if (theta) {
    out_u[t] = u0[t] + S.dt_f * v_cor - S.g_dt_1mth * gx;
    out_v[t] = v0[t] - S.dt_f * u_cor - S.g_dt_1mth * gy;
  } else {
    out_u[t] = u0[t] + S.dt * (S.f * v_cor - S.g * gx);
    out_v[t] = v0[t] + S.dt * (S.neg_f * u_cor - S.g * gy);
  }
}

// rhs = eta0 - dt H ((1 - th) D[u0, v0] + th D[gu, gv])   (schemes.py:154-155)
__global__ void k_sw_rhs(Sw2d S, const double* __restrict__ eta0, const double* __restrict__ u0,
                         const double* __restrict__ v0, const double* __restrict__ gu,
                         const double* __restrict__ gv, double* __restrict__ rhs) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.n) return;
  int i, j;
  sw_ij(S, t, i, j);
  rhs[t] = eta0[t] - S.dt_H * (S.one_m_th * sw_div(S, u0, v0, j, i) + S.th * sw_div(S, gu, gv, j, i));
}

// u = gu - g dt th Gx eta, v = gv - g dt th Gy eta   (schemes.py:161-162)
__global__ void k_sw_backsub(Sw2d S, const double* __restrict__ gu, const double* __restrict__ gv,
                             const double* __restrict__ eta, double* __restrict__ out_u,
                             double* __restrict__ out_v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.n) return;
  int i, j;
  sw_ij(S, t, i, j);
  out_u[t] = gu[t] - S.g_dt_th * sw_gradx(S, eta, j, i);
  out_v[t] = gv[t] - S.g_dt_th * sw_grady(S, eta, j, i);
}

// eta' = eta0 - dt H D[u, v]   (schemes.py:137)
__global__ void k_sw_fb_eta(Sw2d S, const double* __restrict__ eta0, const double* __restrict__ u,
                            const double* __restrict__ v, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.n) return;
  int i, j;
  sw_ij(S, t, i, j);
  out[t] = eta0[t] - S.dt_H * sw_div(S, u, v, j, i);
}

}  // namespace hoc
