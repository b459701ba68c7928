// Kernel declarations and the small device structs shared between the
// kernels and the host orchestration in hoc_abi.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hoc_params.h"

namespace hoc {

constexpr long long PW_CHUNK = 8192;  // NumPy's reduction buffer length
constexpr int PW_MAX_LEAVES = 128;

// Pairwise-summation plan of one chunk length: the leaf blocks (<= 128
// elements each, in order) and the internal nodes of the halving tree,
// level by level from the leaves up. Values are indexed leaves 0..L-1 then
// nodes L..L+N-1; the root is the last node.
struct PwPlan {
  int nleaves, nnodes, nlevels, pad;
  short leaf_off[PW_MAX_LEAVES];
  short leaf_len[PW_MAX_LEAVES];
  short node_l[PW_MAX_LEAVES];
  short node_r[PW_MAX_LEAVES];
  int level_end[16];
};

struct HelmOp {
  const double *Ku, *Kv;
  const uint8_t* mask;
  double coef, cx, cy, dx2, dy2;
};

struct PcgScalars {
  double rz, alpha, beta, tol, norm_b, res;
  double rz2[2];  // rz by iteration parity (per-phase path)
  int it, iters, done, converged;
  double tot[3];  // sums of part_a, part_b, part_c when computed once (k_pcg_totals)
};

// hoc_step_kernels.hip
__global__ void k_geometry(Dims d, StepParams p, const double* dz3, const uint8_t* m3u,
                           const uint8_t* m3v, const double* tau_x, const double* tau_y,
                           double* dzi, double* tau_u, double* tau_v);
__global__ void k_closure(Dims d, StepParams p, Geo geo, const double* e_in, double* Km,
                          double* Kh, double* e_out, const double* s_n2, const double* s_sh,
                          const double* s_l, const double* s_le, double* s_c);
__global__ void k_closure_shear(Dims d, StepParams p, Geo geo, const double* u, const double* v,
                                const double* b, double* s_n2, double* s_sh);
__global__ void k_mixing_recursive(Dims d, StepParams p, Geo geo, const double* e_in,
                                   const double* s_n2, double* s_l, double* s_le);
void launch_closure_lengths(const Dims& d, const StepParams& p, const Geo& geo,
                            const double* e_in, const double* u, const double* v,
                            const double* b, double* s_n2, double* s_sh, double* s_l,
                            double* s_le, hipStream_t s);
// Sets the LDS limit of the integral-length kernel for nz levels; < 0 if
// its staging does not fit one workgroup's LDS.
int closure_lengths_prepare(int nz);
__global__ void k_coefficients(Dims d, StepParams p, Geo geo, const double* Km,
                               double* sub_m, double* den_m, double* cst_m, double* q);
__global__ void k_kukv(Dims d, StepParams p, Geo geo, const double* q, double* Ku, double* Kv);
__global__ void k_phi(Dims d, StepParams p, Geo geo, const double* b, double* phi);
// k_phi, k_depth_sum (U0 set) and k_w_diag (w set) in one column pass
__global__ void k_prologue(Dims d, StepParams p, Geo geo, const double* u, const double* v,
                           const double* b, double* phi, double* U0, double* V0, double* w);
__global__ void k_w_diag(Dims d, StepParams p, Geo geo, const double* u, const double* v,
                         double* w);
__global__ void k_momentum_explicit(Dims d, StepParams p, Geo geo, const double* u,
                                    const double* v, const double* b, const double* phi,
                                    const double* w, const double* Km, double* pgx,
                                    double* pgy, double* avu, double* avv, double* dzu,
                                    double* dzv, double* stu, double* stv);
__global__ void k_depth_sum(Dims d, StepParams p, Geo geo, const double* u, const double* v, double* U,
                            double* V);
__global__ void k_div2(Dims d, StepParams p, Geo geo, const double* U, const double* V,
                       const double* eta, double* out);
__global__ void k_rhs_eta(Dims d, StepParams p, Geo geo, const double* eta,
                          const double* div_old, const double* Ug, const double* Vg,
                          double* rhs);
__global__ void k_eta_fb(Dims d, StepParams p, Geo geo, const double* eta, const double* U,
                         const double* V, double* eta_new);
__global__ void k_moving_dz(Dims d, const uint8_t* m2, const double* H, const double* eta,
                            const double* dz3_ref, double* dz3, double* zc, double* dzi);
__global__ void k_moving_faces(Dims d, StepParams p, const uint8_t* m3u, const uint8_t* m3v,
                               const double* dz3, double* dz3u, double* dz3v);
__global__ void k_split_forcing(Dims d, StepParams p, Geo geo, const double* U0, const double* V0,
                                const double* Um, const double* Vm, double* fx, double* fy);
__global__ void k_bt_picard(Dims d, StepParams p, Geo geo, const double* U, const double* V,
                            const double* e, const double* Ui, const double* Vi,
                            const double* fx, const double* fy, const double* Ku,
                            const double* Kv, double* Uo, double* Vo);
__global__ void k_bt_eta(Dims d, StepParams p, Geo geo, const double* e, const double* U,
                         const double* V, double* e_new, double* acc_u, double* acc_v);
__global__ void k_bt_mean(Dims d, StepParams p, const double* acc_u, const double* acc_v,
                          double* U, double* V);
__global__ void k_split_final(Dims d, Geo geo, const double* gh_u, const double* gh_v,
                              const double* Um, const double* Vm, const double* U,
                              const double* V, const double* Hu, const double* Hv, double* u,
                              double* v);
__global__ void k_face_depth(Dims d, Geo geo, double* Hu, double* Hv);
__global__ void k_predictor(Dims d, StepParams p, Geo geo, const double* u, const double* v,
                            const double* eta, const double* u_it, const double* v_it,
                            const double* pgx, const double* pgy, const double* dzu,
                            const double* dzv, const double* stu, const double* stv,
                            const double* sub,
                            const double* den, const double* cst, double* gh_u, double* gh_v,
                            double* Ug, double* Vg);
__global__ void k_predictor_col(Dims d, StepParams p, Geo geo, const double* u, const double* v,
                                const double* eta, const double* u_it, const double* v_it,
                                const double* pgx, const double* pgy, const double* dzu,
                                const double* dzv, const double* stu, const double* stv,
                                const double* sub, const double* den, const double* cst,
                                double* gh_u, double* gh_v, double* Ug, double* Vg);
__global__ void k_backsub_w(Dims d, StepParams p, Geo geo, const double* gh_u,
                            const double* gh_v, const double* q, const double* eta_n, double* u,
                            double* v, double* w);
__global__ void k_backsub(Dims d, StepParams p, Geo geo, const double* gh_u,
                          const double* gh_v, const double* q, const double* eta_n,
                          double* u_it, double* v_it);
__global__ void k_thomas(long long ncol, int n, const double* sub, const double* diag,
                         const double* sup, const double* rhs, double* x, double* cs);
__global__ void k_eos(long long n, StepParams p, const double* T, const double* S,
                      const double* z, double* b);
__global__ void k_tracer_advection(Dims d, StepParams p, Geo geo, const double* cc,
                                   const double* u, const double* v, const double* w,
                                   double* out);
__global__ void k_momentum_advection(Dims d, StepParams p, Geo geo, const double* u,
                                     const double* v, const double* w, double* au,
                                     double* av);

// hoc_tiled.hip: the LDS-tiled tracer and explicit-momentum passes; mixed
// selects the fp32 advective fluxes (and EOS), else the fp64 plain step
void launch_tracers_tiled(bool mixed, hipStream_t s, Dims d, StepParams p, Geo geo,
                          const double* T, const double* S, const double* b, const double* u,
                          const double* v, const double* w, const double* Kh, double* cst,
                          double* T_out, double* S_out, double* b_out);
void launch_momentum_explicit_tiled(bool mixed, hipStream_t s, Dims d, StepParams p, Geo geo,
                                    const double* u, const double* v, const double* b,
                                    const double* phi, const double* w, const double* Km,
                                    double* pgx, double* pgy, double* avu, double* avv,
                                    double* dzu, double* dzv, double* stu, double* stv);

// hoc_adjoint.hip
__global__ void k_thomas_adjoint(long long ncol, int n, const double* sub, const double* diag,
                                 const double* sup, const double* x, const double* g,
                                 double* lam, double* cs, double* d_sub, double* d_diag,
                                 double* d_sup);
__global__ void k_tracer_advection_adjoint(Dims d, StepParams p, Geo geo, const double* cc,
                                           const double* u, const double* v, const double* w,
                                           const double* G, double* gc, double* gu, double* gv,
                                           double* gw);

__global__ void k_gaspar_lengths(long long ncol, int nz1, double e_min, double z_0,
                                 const double* e, const double* n2, const double* dzi,
                                 const double* zs, const double* zb, const double* wet,
                                 double* l_up, double* l_dn);
__global__ void k_gaspar_lengths_adjoint(long long ncol, int nz1, double e_min, double z_0,
                                         const double* e, const double* n2, const double* dzi,
                                         const double* zs, const double* zb, const double* wet,
                                         const double* g_up, const double* g_dn, double* g_e,
                                         double* g_n2);

__global__ void k_momentum_advection_adjoint(Dims d, StepParams p, Geo geo, const double* u,
                                             const double* v, const double* w, const double* Gu,
                                             const double* Gv, double* gu, double* gv,
                                             double* gw);
__global__ void k_eos_adjoint(long long n, StepParams p, const double* T, const double* S,
                              const double* z, const double* g_b, double* g_T, double* g_S);

// the reverse step (hoc_stepper_step_adjoint)
__global__ void k_fill(long long n, double value, double* a);
__global__ void k_rev_tracer_rhs(Dims d, StepParams p, Geo geo, const double* T1,
                                 const double* S1, const double* gT1, const double* gS1,
                                 const double* gb1, double* rT, double* rS);
__global__ void k_rev_tracers(Dims d, StepParams p, Geo geo, const double* T0, const double* S0,
                              const double* T1, const double* S1, const double* Kh,
                              double* lamT, double* lamS, double* cs, double* gT0, double* gS0,
                              double* gKh);
// k_rev_tracer_stencil and k_rev_momentum_adv: one workgroup per tile of one level
constexpr int RTS_TX = 32, RTS_TY = 8;
constexpr int RMA_TX = 32, RMA_TY = 8;
__global__ void k_rev_tracer_stencil(Dims d, StepParams p, Geo geo, const double* T0,
                                     const double* S0, const double* u1, const double* v1,
                                     const double* w1, const double* lamT, const double* lamS,
                                     const double* gu_y, const double* gv_y, double* gT0,
                                     double* gS0, double* gu1, double* gv1, double* gw1);
__global__ void k_rev_backsub(Dims d, StepParams p, Geo geo, const double* gw, const double* q,
                              const double* eta, int first, double* gui, double* gvi, double* gq,
                              double* X, double* Y);
__global__ void k_rev_eta_rhs(Dims d, StepParams p, Geo geo, const double* geta, const double* X,
                              const double* Y, double* ge);
__global__ void k_rev_eta_post(Dims d, StepParams p, Geo geo, const double* x, const double* lam,
                               int first, double* gKu, double* gKv, double* geta0, double* gdiv,
                               double* GU, double* GV);
__global__ void k_rev_predictor(Dims d, StepParams p, Geo geo, const double* Km, const double* Au,
                                const double* Av, const double* GU, const double* GV,
                                const double* ghu, const double* ghv, int first, double* lu,
                                double* lv, double* cs, double* Lu, double* Lv, double* gKm);
__global__ void k_rev_coriolis(Dims d, StepParams p, Geo geo, const double* lu, const double* lv,
                               const double* Lu, const double* Lv, int first, double* gu,
                               double* gv);
__global__ void k_rev_explicit_cells(Dims d, StepParams p, Geo geo, const double* Lu,
                                     const double* Lv, const double* gdiv, double* gu0,
                                     double* gv0);
__global__ void k_rev_explicit(Dims d, StepParams p, Geo geo, const double* u0, const double* v0,
                               const double* Km, const double* Lu, const double* Lv,
                               double* gu0, double* gv0, double* gb0, double* geta0,
                               double* gKm);
__global__ void k_rev_momentum_adv(Dims d, StepParams p, Geo geo, const double* u0,
                                   const double* v0, const double* w0, const double* Lu,
                                   const double* Lv, double* gu0, double* gv0, double* gw0);
__global__ void k_rev_closure_col0(Dims d, Geo geo, const double* b, double* sn2, double* sn2b,
                                   double* szs, int* shi);
__global__ void k_rev_closure_rows(Dims d, StepParams p, Geo geo, const double* e,
                                   const double* u, const double* v, const double* Km0,
                                   const double* Kh0, const double* kb, const double* sn2,
                                   const double* szs, const int* shi, double* sdu, double* sdv,
                                   double* slu, double* sld, double* srow);
__global__ void k_rev_closure(Dims d, StepParams p, Geo geo, const double* e, const double* Km0,
                              const double* Kh0, const double* kb, const double* ge1,
                              const double* gKm1, const double* gKh1, const double* sn2,
                              const double* szs, const int* shi, const double* sdu,
                              const double* sdv, const double* slu, const double* sld,
                              const double* srow, double* sx, double* scs, double* slam,
                              double* sgl, double* sn2b, double* ge0, double* guc, double* gvc,
                              double* gb0, double* gKm0, double* gKh0, double* gkb);
__global__ void k_rev_closure_uv(Dims d, const double* guc, const double* gvc, double* gu,
                                 double* gv);
__global__ void k_rev_wold_q(Dims d, StepParams p, Geo geo, const double* gw0, const double* Km,
                             const double* q, const double* gKu, const double* gKv, double* gq,
                             double* lam, double* cs, double* gu0, double* gv0, double* gKm);

// hoc_sw2d.hip: the 2D shallow-water model (libs/core/schemes.py:97-177).
// Scalars formed on the host in Python's left-to-right order.
struct Sw2d {
  int nx, ny;
  long long n;
  double dx, dy, g, f, neg_f, H, dt, th, tc;
  double one_m_tc, one_m_th;
  double dt_f;       // dt * f
  double g_dt_1mth;  // g * dt * (1 - th)
  double dt_H;       // dt * H
  double g_dt_th;    // g * dt * th
};
__global__ void k_sw_momentum(Sw2d S, int theta, const double* u0, const double* v0,
                              const double* eta0, const double* u_it, const double* v_it,
                              double* out_u, double* out_v);
__global__ void k_sw_rhs(Sw2d S, const double* eta0, const double* u0, const double* v0,
                         const double* gu, const double* gv, double* rhs);
__global__ void k_sw_backsub(Sw2d S, const double* gu, const double* gv, const double* eta,
                             double* out_u, double* out_v);
__global__ void k_sw_fb_eta(Sw2d S, const double* eta0, const double* u, const double* v,
                            double* out);

// hoc_helm2d.hip: one level of the constant-coefficient periodic Helmholtz
struct H2Level {
  int nx, ny;
  long long n;
  double dx, dy;
  double dx2, dy2;     // dx**2, dy**2
  double cx, cy;       // coef / dx**2, coef / dy**2
  double diag;         // 1 + 2 cx + 2 cy
  double coef;
};
constexpr int H2_COARSE_THREADS = 1024;
constexpr int H2_COARSE_PER = 16;
constexpr long long H2_COARSE_MAX = (long long)H2_COARSE_THREADS * H2_COARSE_PER;
__global__ void k_h2_apply_dg(H2Level L, const double* p, double* ap);
__global__ void k_h2_resid(H2Level L, const double* x, const double* b, double* r);
__global__ void k_h2_gs_colour(H2Level L, int colour, double* x, const double* b);
__global__ void k_h2_gs_colour_copy(H2Level L, int colour, const double* in, const double* b,
                                    double* out);
__global__ void k_h2_gs_block(H2Level L, int n_sweeps, unsigned pattern, double* x,
                              const double* b);
__global__ void k_h2_restrict(H2Level F, H2Level Cl, const double* x, const double* b,
                              double* bc, double* xc);
__global__ void k_h2_prolong(H2Level F, H2Level Cl, double* x, const double* ec);
__global__ void k_h2_xr(long long n, double alpha, const double* p, const double* ap, double* x,
                        double* r);
__global__ void k_h2_jacobi(long long n, double inv_diag, const double* r, double* z);
__global__ void k_h2_p(long long n, double beta, const double* z, double* p);
__global__ void k_h2_fill(long long n, double value, double* x);
__global__ void k_h2_copy(long long n, const double* in, double* out);

// hoc_model3d.hip: the v0.2 3D model (flat bottom, uniform layers, periodic)
struct M3 {
  int nx, ny, nz;
  long long n2, n3;
  double dx, dy, dz, dx2, dy2, neg_dz;
  double dt, f, neg_f, g, H;
  double one_m_tc, tc, one_m_th, th;
  double c_eta;          // g * dt * (1 - th), th = 0 for fb
  double c_diff;         // (1 - theta_v) * dt
  double c_bs;           // g * dt * th
  double nu, kappa, bd;  // viscosity, diffusivity, bottom drag
  double tau_u, tau_v;   // tau_x / rho0, tau_y / rho0
  double s0u, s0v;       // theta_v * dt * tau / dz
  double A_h, K_h;
  int adv;               // 0 none, 1 centered2, 2 upwind1
  double neg_N2;
  double flux_T, flux_S; // q_heat / (rho0 * cp), q_salt
  double s0T, s0S, s0b;  // theta_v * dt * flux / dz
  double alpha_T, beta_S, T0, S0;
  double ddt, gh;        // dt / n_split, g * H
  const double *sub_m, *den_m, *cst_m, *sub_b, *den_b, *cst_b, *q;
};
__global__ void k_m3_phi(M3 M, const double* b, double* phi);
__global__ void k_m3_depth(M3 M, const double* u, const double* v, double* U, double* V);
__global__ void k_m3_w(M3 M, const double* u, const double* v, double* w);
__global__ void k_m3_stress(M3 M, const double* u, const double* v, const double* w, double* st_u,
                            double* st_v);
__global__ void k_m3_predict(M3 M, int split, const double* u0, const double* v0,
                             const double* phi, const double* eta, const double* u_it,
                             const double* v_it, const double* st_u, const double* st_v,
                             double* gh_u, double* gh_v, double* GU, double* GV, double* cor_x,
                             double* cor_y);
__global__ void k_m3_rhs(M3 M, const double* eta, const double* U0, const double* V0,
                         const double* GU, const double* GV, double* rhs);
__global__ void k_m3_backsub(M3 M, const double* gh_u, const double* gh_v, const double* eta,
                             double* u, double* v);
__global__ void k_m3_eta_div(M3 M, double dt, const double* eta, const double* U, const double* V,
                             double* out);
__global__ void k_m3_forcing(M3 M, const double* GU, const double* GV, const double* U0,
                             const double* V0, const double* cor_x, const double* cor_y,
                             double* fx, double* fy);
__global__ void k_m3_bt_momentum(M3 M, const double* u_i, const double* v_i, const double* e,
                                 const double* ui, const double* vi, const double* fx,
                                 const double* fy, double* ui_out, double* vi_out);
__global__ void k_m3_bt_acc(M3 M, const double* u_i, const double* v_i, double* acc_u,
                            double* acc_v);
__global__ void k_m3_split_final(M3 M, int mean_n, const double* gh_u, const double* gh_v,
                                 const double* Ub, const double* Vb, double* u, double* v);
__global__ void k_m3_tracer(M3 M, int which, const double* c, const double* u, const double* v,
                            const double* w, double* out);
__global__ void k_m3_eos(M3 M, const double* T, const double* S, double* b);

// hoc_helmholtz.hip
__global__ void k_helm_diag(Dims d, HelmOp op, double* inv_diag);
__global__ void k_helm_divgrad(Dims d, HelmOp op, const double* x, double* out);
__global__ void k_pw_dot(const double* a, const double* b, const PwPlan* plans,
                         long long nfull, double* part);
__global__ void k_seq_sum(const double* part, long long nchunks, double* out);

}  // namespace hoc

namespace hoc {

// PCG arguments. Full 8192-element chunks are cut into `sub` aligned
// subtrees of the pairwise tree, one workgroup each; the trailing partial
// chunk, if any, adds one workgroup per leaf of its plan. Partial sums land
// in part_*[part]; the consumer of a sum rebuilds each chunk's top levels
// and the in-order chunk sum.
struct PcgArgs {
  Dims d;
  HelmOp op;
  const double* rhs;
  const double* inv_diag;
  double *x, *r, *z, *pb0, *pb1, *Ap;
  double *part_a, *part_b, *part_c;
  const PwPlan* plans;  // [0] full chunk, [1] trailing chunk
  long long nfull;
  int sub, tail, nparts, nchunks;
  double rtol;
  int max_iter;
  unsigned* bar;        // grid barrier block, GRID_BAR_WORDS words
  int* err;
  PcgScalars* sc;
  long long* counters;
  int central;          // 1: k_pcg_totals forms the phase scalars once per grid
};

constexpr int PCG_WG = 256;         // threads per workgroup of a PCG phase
constexpr int GRID_BAR_LINE = 64;   // words per counter line of the grid barrier
constexpr int GRID_BAR_WORDS = 10 * GRID_BAR_LINE;  // 8 group counters, leaders, generation
constexpr int PCG_STAGE = 512;      // elements staged per part (8192 / sub)
constexpr int PCG_MAX_SUB = 16;     // subtrees per full chunk
constexpr int PCG_MAX_CHUNKS = 1024;

// One level of the multigrid hierarchy (level 0 is the model grid).
// One multigrid level: face coefficients, diagonal and constants in the
// arithmetic type R (double for the solvers, float for the mixed-precision
// V-cycle preconditioner); the mask is shared.
template <class R>
struct MgLevelT {
  int nx, ny;
  long long n;
  const R *Ku, *Kv;
  const uint8_t* mask;
  const R* inv_diag;
  R coef, cx, cy, dx2, dy2;
};
using MgLevel = MgLevelT<double>;
using MgLevelF = MgLevelT<float>;

__global__ void k_pcg_persistent(PcgArgs a);
__global__ void k_pcg_totals(PcgArgs a, int which);
__global__ void k_pcg_rz(PcgArgs a, int into_b);
__global__ void k_mg_coarsen(MgLevel f, MgLevel c, double* cKu, double* cKv, uint8_t* cm);
__global__ void k_mg_diag(MgLevel L, double* inv_diag);
// Every coarse level's faces and 1/diag from level 0 in one launch (blockIdx.y
// = level - l_first); the coarse masks, which do not change between solves,
// must already be in lev[l].mask. With fKu set, fp32 copies of each level's
// faces and 1/diag too (level 0 included when l_first = 0).
constexpr int MG_BUILD_LEVELS = 12;
struct MgBuildArgs {
  MgLevel lev[MG_BUILD_LEVELS];
  double *Ku[MG_BUILD_LEVELS], *Kv[MG_BUILD_LEVELS], *inv[MG_BUILD_LEVELS];
  float *fKu[MG_BUILD_LEVELS], *fKv[MG_BUILD_LEVELS], *finv[MG_BUILD_LEVELS];
  int nlev, l_first;
};
__global__ void k_mg_levels(MgBuildArgs A);
__global__ void k_mg_zero(long long n, double* x, const PcgScalars* sc);
constexpr int RB_TW = 32, RB_TH = 8;
__global__ void k_mg_rbgs2(MgLevel L, int c0, const double* b, const double* xin, double* xout,
                           const PcgScalars* sc);
__global__ void k_mg_rbgs(MgLevel L, int colour, int last, const double* b, const double* xin,
                          double* xout, const PcgScalars* sc);
__global__ void k_mg_jacobi(MgLevel L, double omega, const double* x, const double* b, double* xo,
                            const PcgScalars* sc);
__global__ void k_mg_residual_restrict(MgLevel f, MgLevel c, const double* x, const double* b,
                                       double* bc, double* xc_zero, const PcgScalars* sc);
__global__ void k_mg_prolong(MgLevel f, const double* xc, int cnx, double* x,
                             const PcgScalars* sc);
__global__ void k_mg_res_parts(PcgArgs a, MgLevel L, const double* x);
__global__ void k_mg_check(PcgArgs a, int it);
__global__ void k_mg_start(PcgArgs a);
constexpr int MG_COARSE_MAX = 1024;      // cells of a level solved in one workgroup
constexpr int MG_COARSE_LAST_MAX = 4096; // ... also when it is the hierarchy's last level
constexpr int MG_COARSE_THREADS = 1024;
constexpr int MG_FUSED_MAX_SWEEPS = 3;   // fused V-cycle legs for 1..3 sweeps per side
// The coarse end of a V-cycle solved in one workgroup (levels of at most
// MG_COARSE_MAX cells, all vectors in LDS). Measured at 512 x 512: taking
// the 64 x 64 level into the workgroup as well (4096 cells) costs 28 us per
// cycle against 19 us for its two tiled launches plus the 32 x 32 cycle.
// A last level of up to MG_COARSE_LAST_MAX cells (the 64 x 64 end of a
// 5-level hierarchy on 1024 x 1024) is one workgroup's 2S sweeps instead of
// 2S + 1 launches.
constexpr int MG_CYCLE_LEVELS = 8;
constexpr size_t MG_CYCLE_LDS = 160 * 1024;
template <class R>
struct MgCycleT {
  MgLevelT<R> lev[MG_CYCLE_LEVELS];
  int boff[MG_CYCLE_LEVELS], xoff[MG_CYCLE_LEVELS];  // LDS offsets (elements) of b, x_S
  int nl;
  int rb;  // 1: red-black Gauss-Seidel sweeps (a third work vector in LDS), 0: damped Jacobi
  // 1: the levels' Ku, Kv, 1/diag (at coff, elements) and masks (at moff, bytes) are
  // copied into LDS too, when they fit beside the vectors
  int stage;
  int coff[MG_CYCLE_LEVELS], moff[MG_CYCLE_LEVELS];
};
using MgCycle = MgCycleT<double>;
using MgCycleF = MgCycleT<float>;
template <class R>
size_t mg_cycle_layout(MgCycleT<R>& C);  // fills boff/xoff, returns the LDS bytes
int mg_cycle_prepare(size_t lds);        // < 0 if the fp64 cycle does not fit one workgroup
int mg_cycle_prepare_f32(size_t lds);    // the same for the fp32 cycles
void launch_mg_coarse_cycle(const MgCycle& C, size_t lds, double omega, int S, const double* b0,
                            double* out, const PcgScalars* sc, hipStream_t s);
constexpr int MG_TILE_SHAPES = 4;  // fused-leg tile shapes (32x16, 64x8, 16x16, 64x16)
// The pcg_mg V-cycle legs with red-black Gauss-Seidel smoothing (S = 1..3).
void launch_mg_down_rb(int S, int shape, hipStream_t s, const MgLevel& L, const MgLevel& C,
                       const double* b, double* xs_out, double* bc, const PcgScalars* sc);
void launch_mg_up_rb(int S, int shape, hipStream_t s, const MgLevel& L, int cnx,
                     const double* xs_in, const double* ec, const double* b, double* out,
                     const PcgScalars* sc);
void launch_mg_down(int S, int shape, hipStream_t s, const MgLevel& L, const MgLevel& C,
                    double omega, const double* b, double* xs_out, double* bc,
                    const PcgScalars* sc);
void launch_mg_up(int S, int shape, hipStream_t s, const MgLevel& L, int cnx, double omega,
                  const double* xs_in, const double* ec, const double* b, double* out,
                  const PcgScalars* sc);
// fp32 V-cycle of the mixed-precision preconditioner: level 0 reads the fp64
// residual and writes the fp64 result, every other vector is fp32
void launch_mg_coarse_cycle_f32(const MgCycleF& C, size_t lds, float omega, int S,
                                const double* b0_d, const float* b0_f, double* out_d,
                                float* out_f, const PcgScalars* sc, hipStream_t s);
void launch_mg_down_f32(int S, dim3 grid, hipStream_t s, const MgLevelF& L, const MgLevelF& C,
                        float omega, const double* b_d, const float* b_f, float* xs_out,
                        float* bc, const PcgScalars* sc);
void launch_mg_up_f32(int S, dim3 grid, hipStream_t s, const MgLevelF& L, int cnx, float omega,
                      const float* xs_in, const float* ec, const double* b_d, const float* b_f,
                      double* out_d, float* out_f, const PcgScalars* sc);
__global__ void k_mg_to_f32(long long n, const double* Ku, const double* Kv, const double* inv,
                            float* fKu, float* fKv, float* finv);
constexpr int MG_TW = 32, MG_TH = 16, MG_THREADS = 256;  // fused-leg tile and workgroup
template <class R>
dim3 mg_tile_grid(const MgLevelT<R>& L) {
  return dim3((unsigned)((L.nx + MG_TW - 1) / MG_TW), (unsigned)((L.ny + MG_TH - 1) / MG_TH));
}
__global__ void k_pcg_init(PcgArgs a);
__global__ void k_pcg_init_z(PcgArgs a);
__global__ void k_pcg_ap(PcgArgs a, int it);
__global__ void k_pcg_update(PcgArgs a, int it, int jacobi);
__global__ void k_pcg_budget_check(PcgArgs a, int n_it);

}  // namespace hoc
