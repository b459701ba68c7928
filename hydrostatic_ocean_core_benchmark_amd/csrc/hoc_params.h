// Parameter blocks passed by value to the kernels of the ocean-core step.
#pragma once

#include <stdint.h>

namespace hoc {

enum Scheme { SCHEME_THETA = 0, SCHEME_FB = 1, SCHEME_SPLIT = 2 };
enum Advection { ADV_NONE = 0, ADV_CENTRED2 = 1, ADV_UPWIND1 = 2, ADV_UP3 = 3, ADV_UP3_TVD = 4 };
enum Eos { EOS_LINEAR = 0, EOS_SEOS = 1, EOS_TEOS10 = 2 };
enum Solver { SOLVER_PCG_JACOBI = 0, SOLVER_PCG_RBGS = 1, SOLVER_RBGS = 2,
              SOLVER_MULTIGRID = 3, SOLVER_PCG_MG = 4 };

struct Dims {
  int nx, ny, nz;
  long long n2, n3, ni;  // columns, layer cells, interface points
};

// Scalars of one step. The derived coefficients are formed on the host in
// the reference's evaluation order (see hoc_abi.hip) so that every kernel
// multiplies by the same rounded constant the reference does.
struct StepParams {
  int scheme, advection, tracers_ts, eos;
  int pgf_remove, pgf_correction, closure_tke, mxl_recursive;
  int n_picard, solver, max_iter, has_A_h, has_K_h;
  int n_split, bt_mean;  // split_explicit: barotropic substeps, averaged transports
  int open_x, open_y;    // Flather open edges
  int vcoord, face_rule; // vertical coordinate (0 z-level, 1 z*, 2 sigma), face rule
  int zcol;              // 1: the static z-level geometry is read in column form (Geo nw/hb/zs)
  int eos_mixed;         // 1: the tracer kernels evaluate the EOS in the fp32 anomaly form
  double dzr;            // uniform reference layer thickness of the column form
  double dx, dy, dz_nom, dt;
  double theta, theta_cor, theta_v;
  double g, f, rho0, nu_b, bottom_drag, N2;
  double alpha_T, beta_S, T0, S0, A_h, K_h;
  double c_k, c_eps, pr_t, z_0, e_min, n2_min, l_min, e_bb;
  double rtol;
  // derived, host-computed with Python/NumPy evaluation order
  double thv_dt;        // theta_v * dt
  double thv_dt_drag;   // theta_v * dt * bottom_drag
  double one_m_tc;      // 1 - theta_cor
  double c_gx;          // g * dt * (1 - th)
  double c_dzd;         // (1 - theta_v) * dt
  double c_bs;          // g * dt * th
  double one_m_th;      // 1 - th
  double th;            // theta (0 for fb)
  double helm_coef;     // g * theta**2 * dt**2
  double helm_cx, helm_cy, dx2, dy2;
  double flux_T, flux_S;
  double neg_alpha_T, neg_g, neg_f, neg_N2;
  double ddt;           // dt / n_split (split_explicit)
  double eta_ext, u_ext; // Flather external state
  // mixed-precision EOS (anomaly form about (T0, S0)): eos_ref[k] are the
  // polyTEOS10 z-power-k coefficients at the reference point (fp64), eos_g
  // the device table [4][7][7] of the T/S anomaly polynomial (fp32)
  double eos_ref[4];
  double eos_lin[4][2];   // the anomaly polynomial's linear terms (a, c) per z power, fp64
  double ss_ref;          // sqrt((S0 + 32) * 0.875 / 35.16504)
  const float* eos_g;
};

// Static geometry and forcing on the device (set once per stepper).
struct Geo {
  const uint8_t *m2, *mu2, *mv2;        // mask, mask_u, mask_v      [ny,nx]
  const uint8_t *m3, *m3u, *m3v;        // mask3, mask3u, mask3v  [nz,ny,nx]
  const double *dz3, *dz3u, *dz3v, *zc; // thickness, face thickness, depth
  const double *dzi;                    // centre spacing on interfaces [nz+1]
  const double *tau_u, *tau_v;          // (tau/rho0) * mask3[u|v][0]   [ny,nx]
  const double *ustar2, *kappa_b;       // surface u*^2, background kappa [ny,nx]
  const double *H;                      // depth [ny,nx] (open edges only, else null)
  // z-level column form (when StepParams::zcol): wet levels per column, the
  // thickness of each column's deepest wet cell, and zs[k] = the sequential
  // sum of k reference thicknesses, k = 0..nz
  const uint8_t *nw;                    // [ny,nx]
  const double *hb;                     // [ny,nx]
  const double *zs;                     // [nz+1]
};

}  // namespace hoc
