// Host side of libhoc.so: the C-ABI declared in include/hoc.h, the stepper
// handle (geometry, workspaces, counters) and the launch sequence of one step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hoc.h"
#include "hoc_kernels.h"

using namespace hoc;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(HOC_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define LAUNCH_CHECK(name)                                                                   \
  do {                                                                                       \
    hipError_t _e = hipGetLastError();                                                       \
    if (_e != hipSuccess)                                                                    \
      return fail(HOC_ERR_DEVICE, std::string("launch ") + name + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr int TPB = 256;
inline unsigned blocks(long long n) { return (unsigned)((n + TPB - 1) / TPB); }

// ---------------------------------------------------- pairwise-sum plans
struct TreeBuild {
  std::vector<short> leaf_off, leaf_len;
  struct Node {
    int l, r, level;  // child ids: >= 0 leaf index, < 0 -(node index + 1)
  };
  std::vector<Node> nodes;
  // returns (id, level): id >= 0 leaf, < 0 node
  std::pair<int, int> build(int off, int n) {
    if (n <= 128) {
      leaf_off.push_back((short)off);
      leaf_len.push_back((short)n);
      return {(int)leaf_off.size() - 1, 0};
    }
    int n2 = n / 2;
    n2 -= n2 % 8;
    auto a = build(off, n2);
    auto b = build(off + n2, n - n2);
    const int lv = std::max(a.second, b.second) + 1;
    nodes.push_back({a.first, b.first, lv});
    return {-(int)nodes.size(), lv};
  }
};

PwPlan make_plan(int n) {
  PwPlan p;
  std::memset(&p, 0, sizeof(p));
  if (n <= 0) return p;
  TreeBuild t;
  t.build(0, n);
  const int L = (int)t.leaf_off.size();
  const int N = (int)t.nodes.size();
  p.nleaves = L;
  p.nnodes = N;
  for (int i = 0; i < L; ++i) {
    p.leaf_off[i] = t.leaf_off[i];
    p.leaf_len[i] = t.leaf_len[i];
  }
  // order nodes by level (stable); the root has the highest level and is last
  std::vector<int> order(N);
  for (int i = 0; i < N; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return t.nodes[a].level < t.nodes[b].level; });
  std::vector<int> rank(N);
  for (int i = 0; i < N; ++i) rank[order[i]] = i;
  auto value_index = [&](int id) { return id >= 0 ? id : L + rank[-id - 1]; };
  int maxlv = 0;
  for (int i = 0; i < N; ++i) {
    const auto& nd = t.nodes[order[i]];
    p.node_l[i] = (short)value_index(nd.l);
    p.node_r[i] = (short)value_index(nd.r);
    maxlv = std::max(maxlv, nd.level);
  }
  p.nlevels = maxlv;
  for (int lv = 1; lv <= maxlv; ++lv) {
    int cnt = 0;
    for (int i = 0; i < N; ++i)
      if (t.nodes[i].level <= lv) ++cnt;
    p.level_end[lv - 1] = cnt;
  }
  return p;
}

struct Reducer {
  long long n = 0, nfull = 0, nchunks = 0;
  PwPlan* plans = nullptr;  // device [2]
  int init(long long n_) {
    n = n_;
    nfull = n / PW_CHUNK;
    const long long tail = n % PW_CHUNK;
    nchunks = nfull + (tail ? 1 : 0);
    PwPlan h[2];
    h[0] = make_plan((int)PW_CHUNK);
    h[1] = make_plan((int)(tail ? tail : 0));
    HIP_TRY(hipMalloc(&plans, sizeof(h)));
    HIP_TRY(hipMemcpy(plans, h, sizeof(h), hipMemcpyHostToDevice));
    return HOC_OK;
  }
  void release() {
    if (plans) (void)hipFree(plans);
    plans = nullptr;
  }
};

}  // namespace

// ------------------------------------------------------------ the handle
struct hoc_stepper {
  hoc_config cfg{};
  Dims d{};
  StepParams p{};
  Geo geo{};
  std::vector<void*> allocs;
  // geometry owned on device
  uint8_t *m2 = nullptr, *mu2 = nullptr, *mv2 = nullptr, *m3 = nullptr, *m3u = nullptr,
          *m3v = nullptr;
  double *dz3 = nullptr, *dz3u = nullptr, *dz3v = nullptr, *zc = nullptr, *dzi = nullptr;
  double *tau_x = nullptr, *tau_y = nullptr, *tau_u = nullptr, *tau_v = nullptr;
  double *ustar2 = nullptr, *kappa_b = nullptr, *Hd = nullptr;
  double* dz3_ref = nullptr;  // resting thickness (moving coordinates)
  uint8_t* cg_nw = nullptr;   // z-level column form: wet levels per column
  double *cg_hb = nullptr, *cg_zs = nullptr;  // deepest wet thickness; zs[0..nz]
  // mixing (caller-owned)
  double *Km = nullptr, *Kh = nullptr;
  // coefficients
  double *sub_m, *den_m, *cst_m, *q, *Ku, *Kv, *inv_diag;
  double* cst_b;  // c* of tri_b, parked by the tracer kernel between its sweeps
  // step work arrays
  double *phi, *w, *pgx, *pgy, *avu, *avv, *dzu, *dzv, *stu, *stv;
  double *U0, *V0, *div0, *Ug, *Vg, *rhs;
  // split_explicit: face depths, barotropic forcing, substep transports and levels
  double *Hu = nullptr, *Hv = nullptr, *bfx = nullptr, *bfy = nullptr;
  double *bU[3] = {}, *bV[3] = {}, *be[2] = {}, *acc_u = nullptr, *acc_v = nullptr;
  long long barotropic_substeps = 0;
  double *gh_u[2], *gh_v[2];
  double *s_n2, *s_sh, *s_l, *s_le, *s_c;
  // reverse-step workspace (the first hoc_stepper_step_adjoint allocates it)
  bool rv_ready = false;
  double *rv_lamT, *rv_lamS, *rv_cs, *rv_gui, *rv_gvi, *rv_Lu, *rv_Lv, *rv_lu, *rv_lv, *rv_gq;
  double *rv_gw, *rv_gw0;
  double *rv_X, *rv_Y, *rv_ge, *rv_leta, *rv_GU, *rv_GV, *rv_gdiv, *rv_inv, *rv_gKu, *rv_gKv;
  double *rv_Ku, *rv_Kv;  // the tape's K_u, K_v at fixed addresses (graph-replayed solves)
  // closure-reverse workspace (the first hoc_stepper_closure_adjoint allocates it)
  bool cr_ready = false;
  double *cr_n2, *cr_lu, *cr_ld, *cr_x, *cr_cs, *cr_lam, *cr_gl, *cr_n2b, *cr_guc, *cr_gvc;
  double *cr_zs, *cr_du, *cr_dv, *cr_row;
  int* cr_hi;
  // Helmholtz
  double *r, *z, *pp, *pp2, *Ap, *part_a, *part_b, *part_c;
  unsigned* bar = nullptr;      // persistent-solver grid barrier [GRID_BAR_WORDS]
  int* bar_err = nullptr;
  int pcg_sub = 1;        // subtrees per full 8192-element chunk
  int pcg_parts_all = 0;  // workgroups of one PCG phase
  int pcg_parts = 0;      // == pcg_parts_all when the cooperative launch is usable
  PcgScalars* sc = nullptr;
  long long* dcount = nullptr;  // [0] solver iterations, [1] failures
  int* host_flag = nullptr;     // pinned: done, iters
  Reducer red;
  int last_iters = 8;
  int sync_free_budget = 0;
  // sync-free solves replayed as one HIP graph (captured on first use, again
  // when any operand pointer, the budget or the leg tile shape changes)
  hipGraph_t sf_graph = nullptr;
  hipGraphExec_t sf_exec = nullptr;
  const void* sf_key[5] = {};
  int sf_key_budget = 0, sf_key_tile = -1;
  double* sf_x = nullptr;  // fixed solution buffer of the step's graph-replayed solves
  int mg_tile_shape = 0;     // fused V-cycle leg tile shape (0..MG_TILE_SHAPES-1)
                             // default by size: 16x16 up to 512x512 columns
                             // (5.50 vs 5.66 ms/step at 512x512x40), else 32x16
                             // (26.11 vs 26.33 ms/step at 1024x1024x60)  // > 0: PCG solves launch this many iterations, no host polling
  // multigrid hierarchy (level 0 = the model grid)
  static constexpr int MAXLEV = 12;
  int nlev = 1;
  int mg_cycle_l0 = -1;  // first level of the one-workgroup coarse cycle, -1: none
  bool mg_masks_built = false;  // the coarse levels' wet masks (static) are formed
  long long coeff_gen = 0;       // rebuilds of Ku, Kv, 1/diag (build_coefficients)
  long long mg_gen = -1;         // coeff_gen the multigrid levels were built from
  MgLevel lev[MAXLEV];
  double *mg_Ku[MAXLEV] = {}, *mg_Kv[MAXLEV] = {}, *mg_inv[MAXLEV] = {};
  uint8_t* mg_mask[MAXLEV] = {};
  double *mg_x[MAXLEV] = {}, *mg_t[MAXLEV] = {}, *mg_b[MAXLEV] = {};
  // the fp32 V-cycle of the mixed-precision preconditioner (same hierarchy)
  MgLevelF levf[MAXLEV];
  float *mgf_Ku[MAXLEV] = {}, *mgf_Kv[MAXLEV] = {}, *mgf_inv[MAXLEV] = {};
  float *mgf_x[MAXLEV] = {}, *mgf_t[MAXLEV] = {}, *mgf_b[MAXLEV] = {};
  bool force_multikernel = true;   // per-phase launches unless path 0 is asked for
  // parts of a mixed-precision step run in fp32 (HOC_FP32_* bits); 0 for fp64
  int fp32_parts = 0;
  bool coeffs_built = false;
  // host counters
  long long solver_rebuilds = 0, tridiagonal_solves = 0, eos_cells = 0;
  // per-kernel event timing (one kernel name at a time)
  std::string prof_name;
  std::vector<hipEvent_t> prof_ev;  // start/stop pairs
  int prof_used = 0;
  long long prof_launches = 0;
  double prof_ms = 0.0;

  template <class T>
  int alloc(T** ptr, long long count) {
    void* v = nullptr;
    HIP_TRY(hipMalloc(&v, std::max<long long>(count, 1) * (long long)sizeof(T)));
    allocs.push_back(v);
    *ptr = (T*)v;
    return HOC_OK;
  }
  ~hoc_stepper() {
    if (sf_exec) (void)hipGraphExecDestroy(sf_exec);
    if (sf_graph) (void)hipGraphDestroy(sf_graph);
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    for (void* v : allocs) (void)hipFree(v);
    if (host_flag) (void)hipHostFree(host_flag);
    red.release();
  }
};

namespace {

// A handle's device buffers are freed only after the device has finished the
// work queued on them, and only while the runtime is healthy: at process exit
// the HIP runtime may already be torn down (and after a device error every
// call fails), and then the handle is left to the process's end rather than
// calling into the runtime again.
bool release_is_safe() { return hipDeviceSynchronize() == hipSuccess; }

int derive_params(const hoc_config& c, StepParams& p) {
  std::memset(&p, 0, sizeof(p));
  p.scheme = c.scheme;
  p.advection = c.advection;
  p.tracers_ts = c.tracers_ts;
  p.eos = c.eos;
  p.pgf_remove = c.pgf_remove;
  p.pgf_correction = c.pgf_correction;
  p.closure_tke = c.closure_tke;
  p.mxl_recursive = c.mxl_recursive;
  p.n_picard = c.n_picard;
  p.solver = c.solver;
  p.max_iter = c.max_iter;
  p.has_A_h = c.A_h > 0.0;
  p.has_K_h = c.K_h > 0.0;
  p.dx = c.dx;
  p.dy = c.dy;
  p.dz_nom = c.dz_nominal;
  p.dt = c.dt;
  p.theta = c.theta;
  p.theta_cor = c.theta_cor;
  p.theta_v = c.theta_v;
  p.g = c.g;
  p.f = c.f;
  p.rho0 = c.rho0;
  p.nu_b = c.nu_b;
  p.bottom_drag = c.bottom_drag;
  p.N2 = c.N2;
  p.alpha_T = c.alpha_T;
  p.beta_S = c.beta_S;
  p.T0 = c.T0;
  p.S0 = c.S0;
  p.A_h = c.A_h;
  p.K_h = c.K_h;
  p.c_k = c.c_k;
  p.c_eps = c.c_eps;
  p.pr_t = c.pr_t;
  p.z_0 = c.z_0;
  p.e_min = c.e_min;
  p.n2_min = c.n2_min;
  p.l_min = c.l_min;
  p.e_bb = c.e_bb;
  p.rtol = c.rtol;
  // Derived constants in the reference's (Python) evaluation order; `x**2`
  // in Python is libm pow(x, 2.0), used identically here.
  const double th = (c.scheme == 0) ? c.theta : 0.0;
  p.th = th;
  p.thv_dt = c.theta_v * c.dt;
  p.thv_dt_drag = c.theta_v * c.dt * c.bottom_drag;
  p.one_m_tc = 1.0 - c.theta_cor;
  p.c_gx = c.g * c.dt * (1.0 - th);
  p.c_dzd = (1.0 - c.theta_v) * c.dt;
  p.c_bs = c.g * c.dt * th;
  p.one_m_th = 1.0 - th;
  p.helm_coef = c.g * std::pow(c.theta, 2.0) * std::pow(c.dt, 2.0);
  p.dx2 = std::pow(c.dx, 2.0);
  p.dy2 = std::pow(c.dy, 2.0);
  p.helm_cx = p.helm_coef * (1.0 / p.dx2);
  p.helm_cy = p.helm_coef * (1.0 / p.dy2);
  p.flux_T = c.q_heat / (c.rho0 * c.cp);
  p.flux_S = c.q_salt;
  p.neg_alpha_T = -c.alpha_T;
  p.neg_g = -c.g;
  p.neg_f = -c.f;
  p.neg_N2 = -c.N2;
  p.n_split = c.n_split;
  p.bt_mean = c.barotropic_mean != 0;
  p.ddt = c.n_split > 0 ? c.dt / c.n_split : 0.0;
  p.vcoord = c.vcoord;
  p.face_rule = c.face_rule;
  p.open_x = c.open_x != 0;
  p.open_y = c.open_y != 0;
  p.eta_ext = c.eta_ext;
  p.u_ext = c.u_ext;
  return HOC_OK;
}

// polyTEOS10-bsq coefficients E_ijk of ss^i tt^j zh^k (libs/core/eos.py:68-100;
// Roquet et al. 2015), for the host-side re-expansion of the mixed EOS.
struct TeosTerm {
  int i, j, k;
  double e;
};
constexpr TeosTerm TEOS_TERMS[] = {
    {0, 0, 0, 8.0189615746e+02}, {1, 0, 0, 8.6672408165e+02}, {2, 0, 0, -1.7864682637e+03},
    {3, 0, 0, 2.0375295546e+03}, {4, 0, 0, -1.2849161071e+03}, {5, 0, 0, 4.3227585684e+02},
    {6, 0, 0, -6.0579916612e+01}, {0, 1, 0, 2.6010145068e+01}, {1, 1, 0, -6.5281885265e+01},
    {2, 1, 0, 8.1770425108e+01}, {3, 1, 0, -5.6888046321e+01}, {4, 1, 0, 1.7681814114e+01},
    {5, 1, 0, -1.9193502195e+00}, {0, 2, 0, -3.7074170417e+01}, {1, 2, 0, 6.1548258127e+01},
    {2, 2, 0, -6.0362551501e+01}, {3, 2, 0, 2.9130021253e+01}, {4, 2, 0, -5.4723692739e+00},
    {0, 3, 0, 2.1661789529e+01}, {1, 3, 0, -3.3449108469e+01}, {2, 3, 0, 1.9717078466e+01},
    {3, 3, 0, -3.1742946532e+00}, {0, 4, 0, -8.3627885467e+00}, {1, 4, 0, 1.1311538584e+01},
    {2, 4, 0, -5.3563304045e+00}, {0, 5, 0, 5.4048723791e-01}, {1, 5, 0, 4.5111434961e-01},
    {0, 6, 0, -1.9098268277e-01}, {0, 0, 1, 1.9681925209e+01}, {1, 0, 1, -4.2549998214e+01},
    {2, 0, 1, 5.0774768218e+01}, {3, 0, 1, -3.0938076334e+01}, {4, 0, 1, 6.6051753097e+00},
    {0, 1, 1, -1.3336301113e+01}, {1, 1, 1, -4.4870114575e+00}, {2, 1, 1, 5.0042598061e+00},
    {3, 1, 1, -6.5399043664e-01}, {0, 2, 1, 6.7080479603e+00}, {1, 2, 1, 3.5063081279e+00},
    {2, 2, 1, -1.8795372996e+00}, {0, 3, 1, -2.4649669534e+00}, {1, 3, 1, -5.5077101279e-01},
    {0, 4, 1, 5.5927935970e-01}, {0, 0, 2, 2.0660924175e+00}, {1, 0, 2, -4.9527603989e+00},
    {2, 0, 2, 2.5019633803e+00}, {0, 1, 2, 2.0564311499e+00}, {1, 1, 2, -2.1311365518e-01},
    {0, 2, 2, -1.2419983026e+00}, {0, 0, 3, -2.3342758797e-02}, {1, 0, 3, -1.8507636718e-02},
    {0, 1, 3, 3.7969820455e-01}};

// The mixed EOS's anomaly polynomial: each z-power k of polyTEOS10 as a
// polynomial in a = tt - tt_r, c = ss - ss_r (binomial re-expansion about
// tt_r = T0 / 40, ss_r = sqrt((S0 + 32) 0.875 / 35.16504)); G[k][p][q] holds
// the coefficient of a^p c^q, the constant terms go to ref[k] (fp64).
void teos_anomaly_table(double T0, double S0, double ref[4], double lin[4][2],
                        float G[4][7][7], double* ss_ref) {
  const double tt_r = T0 * (1.0 / 40.0);
  const double ss_r = std::sqrt((S0 + 32.0) * (0.875 / 35.16504));
  *ss_ref = ss_r;
  double acc[4][7][7] = {};
  auto binom = [](int n, int r) {
    double b = 1.0;
    for (int t = 1; t <= r; ++t) b = b * (n - r + t) / t;
    return b;
  };
  for (const TeosTerm& t : TEOS_TERMS)
    for (int pp = 0; pp <= t.j; ++pp)
      for (int q = 0; q <= t.i; ++q)
        acc[t.k][pp][q] += t.e * binom(t.j, pp) * std::pow(tt_r, t.j - pp) * binom(t.i, q) *
                           std::pow(ss_r, t.i - q);
  // the constant and linear terms stay fp64 (ref, lin); the fp32 table holds
  // the terms of degree >= 2, a fraction of the anomaly, so that its fp32
  // rounding is a fraction of the anomaly's rounding
  for (int k = 0; k < 4; ++k) {
    ref[k] = acc[k][0][0];
    lin[k][0] = acc[k][1][0];
    lin[k][1] = acc[k][0][1];
    for (int pp = 0; pp < 7; ++pp)
      for (int q = 0; q < 7; ++q) G[k][pp][q] = (pp + q < 2) ? 0.0f : (float)acc[k][pp][q];
  }
}

int validate(const hoc_config& c) {
  if (c.nx < 2 || c.ny < 2 || c.nz < 1)
    return fail(HOC_ERR_INVALID, "grid must have nx >= 2, ny >= 2, nz >= 1");
  if (!(c.dx > 0.0) || !(c.dy > 0.0) || !(c.dt > 0.0))
    return fail(HOC_ERR_INVALID, "dx, dy and dt must be positive");
  if (c.scheme < 0 || c.scheme > 2)
    return fail(HOC_ERR_INVALID, "scheme must be theta (0), fb (1) or split_explicit (2)");
  if (c.scheme == 2 && c.n_split < 1)
    return fail(HOC_ERR_INVALID, "split_explicit needs n_split >= 1");
  if (c.vcoord < 0 || c.vcoord > 2)
    return fail(HOC_ERR_INVALID, "vcoord must be zlevel (0), zstar (1) or sigma (2)");
  if (c.face_rule < 0 || c.face_rule > 2)
    return fail(HOC_ERR_INVALID, "face_rule must be min (0), mean (1) or harmonic (2)");
  if (c.scheme == 0 && !(c.theta > 0.0 && c.theta <= 1.0))
    return fail(HOC_ERR_INVALID, "theta must lie in (0, 1]");
  if (c.n_picard < 1) return fail(HOC_ERR_INVALID, "n_picard must be >= 1");
  if (c.advection < 0 || c.advection > 4) return fail(HOC_ERR_INVALID, "unknown advection");
  if (c.eos < 0 || c.eos > 2)
    return fail(HOC_ERR_INVALID, "unknown eos (expected linear|seos|teos10)");
  if (c.scheme == 0 && (c.solver < SOLVER_PCG_JACOBI || c.solver > SOLVER_PCG_MG))
    return fail(HOC_ERR_INVALID, "unknown variable-coefficient solver "
                                 "(expected pcg_jacobi|pcg_rbgs|rbgs|multigrid|pcg_mg)");
  if (c.solver == SOLVER_PCG_MG && (c.mg_sweeps < 1 || c.mg_levels < 1 || !(c.mg_omega > 0.0)))
    return fail(HOC_ERR_INVALID, "pcg_mg needs mg_levels >= 1, mg_sweeps >= 1, mg_omega > 0");
  if (c.max_iter < 1) return fail(HOC_ERR_INVALID, "max_iter must be >= 1");
  if (c.precision < 0 || c.precision > 2)
    return fail(HOC_ERR_INVALID, "precision must be fp64 (0), mixed (1) or mixed with fp32 "
                                 "momentum fluxes (2)");
  if (c.precision >= 1 && c.scheme == 0 && c.solver != SOLVER_PCG_MG)
    return fail(HOC_ERR_INVALID, "mixed precision runs the theta scheme with pcg_mg only");
  if (c.precision >= 1 && c.mg_sweeps > MG_FUSED_MAX_SWEEPS)
    return fail(HOC_ERR_INVALID, "mixed precision needs mg_sweeps <= 3");
  if (c.mg_smoother < 0 || c.mg_smoother > 1)
    return fail(HOC_ERR_INVALID, "mg_smoother must be jacobi (0) or rbgs (1)");
  if (c.mg_smoother == 1 && c.solver == SOLVER_PCG_MG &&
      (c.precision != 0 || c.mg_sweeps > MG_FUSED_MAX_SWEEPS))
    return fail(HOC_ERR_INVALID, "the rbgs pcg_mg smoother runs in fp64 with mg_sweeps <= 3");
  return HOC_OK;
}

template <class T>
int upload(T* dst, const T* src, long long n, const char* name) {
  if (!src) return fail(HOC_ERR_INVALID, std::string("geometry array missing: ") + name);
  HIP_TRY(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return HOC_OK;
}

#define TRY(expr)             \
  do {                        \
    int _rc = (expr);         \
    if (_rc != HOC_OK) return _rc; \
  } while (0)

}  // namespace

// Fixed device scratch of the no-allocation entry point: one copy per GPU,
// held for the life of the process. In host code `scratch` names a host-side
// stand-in, never the device address: host code uses only the per-device
// pointer noalloc_scratch() looks up once per device.
static __device__ __align__(256) double scratch[HOC_NOALLOC_SCRATCH_DOUBLES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up on the first
// call made with that device current and kept for later calls.
int noalloc_scratch(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(HOC_ERR_UNSUPPORTED, "no-alloc scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(HOC_ERR_DEVICE, "no-alloc scratch: null symbol address");
    g_noalloc_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_noalloc_ptr[dev];
  return HOC_OK;
}

// Column form of a static z-level geometry (see cg_dz in hoc_device.h): per
// column the number of wet levels nw and the thickness hb of the deepest
// wet cell, plus the uniform reference thickness dzr and the sequential
// sums zs[k] of k reference thicknesses. Returns false when any stored
// thickness, face thickness, mask or centre depth differs in a single bit
// from what the column form gives (the kernels then read the full arrays).
bool bits_equal(double a, double b) {
  uint64_t x, y;
  std::memcpy(&x, &a, 8);
  std::memcpy(&y, &b, 8);
  return x == y;
}

double face_rule_host(int rule, double a, double b) {
  if (rule == 0) return (a <= b || a != a) ? a : b;
  if (rule == 1) return 0.5 * (a + b);
  const double s = a + b;
  return s > 0.0 ? 2.0 * a * b / s : 0.0;
}

bool column_form(const Dims& d, int face_rule, const hoc_geometry* g, std::vector<uint8_t>& nw,
                 std::vector<double>& hb, std::vector<double>& zs, double& dzr) {
  const int nz = d.nz;
  const long long n2 = d.n2;
  if (nz > 255 || !g->mask3 || !g->dz3 || !g->dz3u || !g->dz3v || !g->z_centre || !g->mask3u ||
      !g->mask3v || !g->mask_u || !g->mask_v)
    return false;
  nw.assign(n2, 0);
  hb.assign(n2, 0.0);
  dzr = 0.0;
  for (long long c = 0; c < n2; ++c) {
    int n = 0;
    while (n < nz && g->mask3[(long long)n * n2 + c]) ++n;
    for (int k = n; k < nz; ++k)
      if (g->mask3[(long long)k * n2 + c]) return false;  // not a wet prefix
    nw[c] = (uint8_t)n;
    hb[c] = n > 0 ? g->dz3[(long long)(n - 1) * n2 + c] : 0.0;
    if (n >= 2) {
      const double top = g->dz3[c];
      if (dzr == 0.0) dzr = top;
      else if (!bits_equal(dzr, top)) return false;
    }
  }
  if (dzr == 0.0) {  // no column with a full level: any value reproduces the arrays
    for (long long c = 0; c < n2; ++c) dzr = std::max(dzr, hb[c]);
  }
  zs.assign(nz + 1, 0.0);
  for (int k = 1; k <= nz; ++k) zs[k] = zs[k - 1] + dzr;
  auto dz = [&](long long c, int k) {
    const int n = nw[c];
    return k < n - 1 ? dzr : (k == n - 1 ? hb[c] : 0.0);
  };
  for (int j = 0; j < d.ny; ++j)
    for (int i = 0; i < d.nx; ++i) {
      const long long c = (long long)j * d.nx + i;
      const long long ce = (long long)j * d.nx + (i + 1 == d.nx ? 0 : i + 1);
      const long long cn = (long long)(j + 1 == d.ny ? 0 : j + 1) * d.nx + i;
      const int n = nw[c];
      for (int k = 0; k < nz; ++k) {
        const long long t = (long long)k * n2 + c;
        if (!bits_equal(g->dz3[t], dz(c, k))) return false;
        const double zc = k < n ? zs[k] + 0.5 * dz(c, k) : (n > 0 ? zs[n - 1] + hb[c] : 0.0);
        if (!bits_equal(g->z_centre[t], zc)) return false;
        const bool wu = g->mask_u[c] && k < n && k < nw[ce];
        const bool wv = g->mask_v[c] && k < n && k < nw[cn];
        if ((g->mask3u[t] != 0) != wu || (g->mask3v[t] != 0) != wv) return false;
        const double fu = wu ? face_rule_host(face_rule, dz(c, k), dz(ce, k)) : 0.0;
        const double fv = wv ? face_rule_host(face_rule, dz(c, k), dz(cn, k)) : 0.0;
        if (!bits_equal(g->dz3u[t], fu) || !bits_equal(g->dz3v[t], fv)) return false;
      }
    }
  return true;
}

// Fold the recorded event pairs into the running total.
int prof_drain(hoc_stepper* h) {
  for (int i = 0; i < h->prof_used; ++i) {
    HIP_TRY(hipEventSynchronize(h->prof_ev[2 * i + 1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->prof_ev[2 * i], h->prof_ev[2 * i + 1]));
    h->prof_ms += ms;
  }
  h->prof_used = 0;
  return HOC_OK;
}

// Event pair around one launch of `name` when that kernel is being timed.
struct ProfScope {
  hoc_stepper* h;
  hipStream_t s;
  bool on;
  ProfScope(hoc_stepper* h_, const char* name, hipStream_t s_) : h(h_), s(s_) {
    on = !h->prof_name.empty() && h->prof_name == name;
    if (!on) return;
    if (2 * (h->prof_used + 1) > (int)h->prof_ev.size()) (void)prof_drain(h);
    (void)hipEventRecord(h->prof_ev[2 * h->prof_used], s);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(h->prof_ev[2 * h->prof_used + 1], s);
    h->prof_used += 1;
    h->prof_launches += 1;
  }
};

int build_coefficients(hoc_stepper* h, hipStream_t s) {
  if (!h->Km || !h->Kh) return fail(HOC_ERR_INVALID, "bind K_m / K_h before building");
  const Dims d = h->d;
  { ProfScope _ps(h, "k_coefficients", s);
  k_coefficients<<<blocks(d.n2), TPB, 0, s>>>(d, h->p, h->geo, h->Km, h->sub_m, h->den_m,
                                              h->cst_m, h->q); }
  LAUNCH_CHECK("k_coefficients");
  { ProfScope _ps(h, "k_kukv", s);
  k_kukv<<<blocks(d.n2), TPB, 0, s>>>(d, h->p, h->geo, h->q, h->Ku, h->Kv); }
  LAUNCH_CHECK("k_kukv");
  h->tridiagonal_solves += 1;
  h->coeff_gen += 1;  // Ku, Kv and 1/diag are new: the multigrid levels follow them
  if (h->cfg.scheme == 0) {
    HelmOp op{h->Ku, h->Kv, h->m2, h->p.helm_coef, h->p.helm_cx, h->p.helm_cy, h->p.dx2,
              h->p.dy2};
    k_helm_diag<<<blocks(d.n2), TPB, 0, s>>>(d, op, h->inv_diag);
    LAUNCH_CHECK("k_helm_diag");
    if (h->coeffs_built) h->solver_rebuilds += 1;
  }
  h->coeffs_built = true;
  return HOC_OK;
}

PcgArgs pcg_args(hoc_stepper* h, const HelmOp& op, const double* inv_diag, const double* rhs,
                 double* x) {
  PcgArgs a;
  std::memset(&a, 0, sizeof(a));
  a.d = h->d;
  a.op = op;
  a.rhs = rhs;
  a.inv_diag = inv_diag;
  a.x = x;
  a.r = h->r;
  a.z = h->z;
  a.pb0 = h->pp;
  a.pb1 = h->pp2;
  a.Ap = h->Ap;
  a.part_a = h->part_a;
  a.part_b = h->part_b;
  a.part_c = h->part_c;
  a.plans = h->red.plans;
  a.nfull = h->red.nfull;
  a.sub = h->pcg_sub;
  a.tail = (int)(h->red.nchunks - h->red.nfull);
  a.nparts = h->pcg_parts;
  a.nchunks = (int)h->red.nchunks;
  a.rtol = h->p.rtol;
  a.max_iter = h->p.max_iter;
  a.bar = h->bar;
  a.err = h->bar_err;
  a.sc = h->sc;
  a.counters = h->dcount;
  // beyond 256 chunks (2M cells) the phases' scalars are formed once per
  // grid (k_pcg_totals) instead of in every workgroup
  a.central = a.nchunks > 256 ? 1 : 0;
  return a;
}

// Poll of a per-phase solve between batches: it, iters, done, converged
// (the end of PcgScalars) into host_flag[0..3] in one copy; synchronises.
// Returns done.
int poll_done(hoc_stepper* h, hipStream_t s, int* done) {
  HIP_TRY(hipMemcpyAsync(h->host_flag, &h->sc->it, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *done = h->host_flag[2];
  return HOC_OK;
}

// Read back (iterations, converged) after a solve; synchronises, except
// after a poll that saw the solve done (polled: host_flag holds them).
int pcg_finish(hoc_stepper* h, hipStream_t s, bool persistent, int* iters_out, int* conv_out,
               bool polled = false) {
  if (!polled || persistent) {
    HIP_TRY(hipMemcpyAsync(h->host_flag, &h->sc->it, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(h->host_flag + 4, h->bar_err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  if (persistent && h->host_flag[4]) {
    (void)hipMemset(h->bar_err, 0, sizeof(int));
    (void)hipMemset(h->bar, 0, GRID_BAR_WORDS * sizeof(unsigned));
    return fail(HOC_ERR_DEVICE, "persistent Helmholtz solve: grid barrier timed out");
  }
  // PcgScalars ends with it, iters, done, converged
  if (iters_out) *iters_out = h->host_flag[1];
  if (conv_out) *conv_out = h->host_flag[3];
  if (h->host_flag[1] > 0) h->last_iters = h->host_flag[1];
  return HOC_OK;
}

// Level operators for the face coefficients Ku, Kv (level 0 given, the
// others agglomerated) and their diagonals.
int mg_build(hoc_stepper* h, const double* Ku, const double* Kv, const double* inv0,
             hipStream_t s) {
  MgLevel& L0 = h->lev[0];
  L0.Ku = Ku;
  L0.Kv = Kv;
  L0.mask = h->m2;
  L0.inv_diag = inv0;
  for (int l = 1; l < h->nlev; ++l) {
    MgLevel& L = h->lev[l];
    L.Ku = h->mg_Ku[l];
    L.Kv = h->mg_Kv[l];
    L.mask = h->mg_mask[l];
    L.inv_diag = h->mg_inv[l];
    // the coarse masks follow from the static wet mask: formed on the first
    // build only (with faces that k_mg_levels then forms again, equal bits)
    if (!h->mg_masks_built)
      k_mg_coarsen<<<blocks(L.n), TPB, 0, s>>>(h->lev[l - 1], L, h->mg_Ku[l], h->mg_Kv[l],
                                               h->mg_mask[l]);
  }
  h->mg_masks_built = true;
  const bool f32 = h->cfg.precision >= 1;  // the fp32 levels exist for every mixed handle
  MgBuildArgs A{};
  if (h->nlev > MG_BUILD_LEVELS) return fail(HOC_ERR_INVALID, "too many multigrid levels");
  A.nlev = h->nlev;
  A.l_first = f32 ? 0 : 1;
  long long nmax = 0;
  for (int l = 0; l < h->nlev; ++l) {
    A.lev[l] = h->lev[l];
    A.Ku[l] = h->mg_Ku[l];
    A.Kv[l] = h->mg_Kv[l];
    A.inv[l] = h->mg_inv[l];
    if (f32) {
      MgLevelF& F = h->levf[l];
      F.Ku = A.fKu[l] = h->mgf_Ku[l];
      F.Kv = A.fKv[l] = h->mgf_Kv[l];
      F.inv_diag = A.finv[l] = h->mgf_inv[l];
      F.mask = h->lev[l].mask;
    }
    if (l >= A.l_first) nmax = std::max(nmax, h->lev[l].n);
  }
  if (A.nlev > A.l_first)
    k_mg_levels<<<dim3((unsigned)blocks(nmax), (unsigned)(A.nlev - A.l_first)), TPB, 0, s>>>(A);
  LAUNCH_CHECK("multigrid levels");
  return HOC_OK;
}

// One red-black sweep on x (forward: red then black; reverse: black then
// red), through the level's spare buffer t; the result is back in x.
void rbgs_sweep(hoc_stepper* h, int l, const double* b, double* x, bool reverse, hipStream_t s) {
  const MgLevel& L = h->lev[l];
  double* t = h->mg_t[l];
  const int c0 = reverse ? 1 : 0;
  k_mg_rbgs<<<blocks(L.n), TPB, 0, s>>>(L, c0, 0, b, x, t, h->sc);
  k_mg_rbgs<<<blocks(L.n), TPB, 0, s>>>(L, 1 - c0, 1, b, t, x, h->sc);
}

// Two sweeps in a row on x (directions r1 then r2), one launch each through
// the level's spare buffer t and back; the result is back in x and equals
// two calls of rbgs_sweep bit for bit.
void rbgs_sweep_pair(hoc_stepper* h, int l, const double* b, double* x, bool r1, bool r2,
                     hipStream_t s) {
  const MgLevel& L = h->lev[l];
  double* t = h->mg_t[l];
  const dim3 grid((unsigned)((L.nx + RB_TW - 1) / RB_TW), (unsigned)((L.ny + RB_TH - 1) / RB_TH));
  k_mg_rbgs2<<<grid, RB_TW * RB_TH, 0, s>>>(L, r1 ? 1 : 0, b, x, t, h->sc);
  k_mg_rbgs2<<<grid, RB_TW * RB_TH, 0, s>>>(L, r2 ? 1 : 0, b, t, x, h->sc);
}

// The reference multigrid V-cycle on x (solvers_var.py:139-151): n_pre
// forward RBGS sweeps, coarse correction, n_post reverse sweeps.
void vcycle_rbgs(hoc_stepper* h, int l, double* x, const double* b, hipStream_t s) {
  rbgs_sweep_pair(h, l, b, x, false, false, s);
  if (l + 1 < h->nlev) {
    const MgLevel& C = h->lev[l + 1];
    k_mg_residual_restrict<<<blocks(C.n), TPB, 0, s>>>(h->lev[l], C, x, b, h->mg_b[l + 1],
                                                       h->mg_x[l + 1], h->sc);
    vcycle_rbgs(h, l + 1, h->mg_x[l + 1], h->mg_b[l + 1], s);
    k_mg_prolong<<<blocks(h->lev[l].n), TPB, 0, s>>>(h->lev[l], h->mg_x[l + 1], C.nx, x, h->sc);
  }
  rbgs_sweep_pair(h, l, b, x, true, true, s);
}

// pcg_mg preconditioner: one damped-Jacobi V-cycle from a zero guess
// (the oracle's _mg_precond). Returns the buffer holding the result.
double* vcycle_jacobi(hoc_stepper* h, int l, const double* b, double* x, double* t, hipStream_t s) {
  const MgLevel& L = h->lev[l];
  const double om = h->cfg.mg_omega;
  k_mg_zero<<<blocks(L.n), TPB, 0, s>>>(L.n, x, h->sc);
  for (int k = 0; k < h->cfg.mg_sweeps; ++k) {
    k_mg_jacobi<<<blocks(L.n), TPB, 0, s>>>(L, om, x, b, t, h->sc);
    std::swap(x, t);
  }
  if (l + 1 < h->nlev) {
    const MgLevel& C = h->lev[l + 1];
    k_mg_residual_restrict<<<blocks(C.n), TPB, 0, s>>>(L, C, x, b, h->mg_b[l + 1], nullptr,
                                                       h->sc);
    double* xc = vcycle_jacobi(h, l + 1, h->mg_b[l + 1], h->mg_x[l + 1], h->mg_t[l + 1], s);
    k_mg_prolong<<<blocks(L.n), TPB, 0, s>>>(L, xc, C.nx, x, h->sc);
  }
  for (int k = 0; k < h->cfg.mg_sweeps; ++k) {
    k_mg_jacobi<<<blocks(L.n), TPB, 0, s>>>(L, om, x, b, t, h->sc);
    std::swap(x, t);
  }
  return x;
}

// z = M^-1 r for pcg_rbgs (symmetric RBGS sweep pair from zero,
// solvers_var.py:154-160) and pcg_mg.
int precondition(hoc_stepper* h, const double* r, double* z, hipStream_t s) {
  if (h->cfg.solver == SOLVER_PCG_RBGS) {
    k_mg_zero<<<blocks(h->d.n2), TPB, 0, s>>>(h->d.n2, z, h->sc);
    rbgs_sweep_pair(h, 0, r, z, false, true, s);
  } else if (h->fp32_parts & HOC_FP32_VCYCLE) {
    // the fp32 V-cycle: as the fused fp64 path below, every level's vectors
    // and coefficients in fp32; level 0 reads the fp64 residual r and writes
    // the fp64 z (the outer PCG stays fp64)
    const int S = h->cfg.mg_sweeps, l0 = h->mg_cycle_l0;
    const float om = (float)h->cfg.mg_omega;
    if (l0 < 0) return fail(HOC_ERR_DEVICE, "mixed pcg_mg: no one-workgroup coarse cycle");
    for (int l = 0; l < l0; ++l)
      launch_mg_down_f32(S, mg_tile_grid(h->levf[l]), s, h->levf[l], h->levf[l + 1], om,
                         l == 0 ? r : nullptr, l == 0 ? nullptr : h->mgf_b[l], h->mgf_x[l],
                         h->mgf_b[l + 1], h->sc);
    {
      MgCycleF C{};
      C.nl = h->nlev - l0;
      for (int l = 0; l < C.nl; ++l) C.lev[l] = h->levf[l0 + l];
      const size_t lds = mg_cycle_layout(C);
      launch_mg_coarse_cycle_f32(C, lds, om, S, l0 == 0 ? r : nullptr,
                                 l0 == 0 ? nullptr : h->mgf_b[l0], l0 == 0 ? z : nullptr,
                                 l0 == 0 ? nullptr : h->mgf_x[l0], h->sc, s);
    }
    const float* ec = h->mgf_x[l0];
    for (int l = l0 - 1; l >= 0; --l) {
      launch_mg_up_f32(S, mg_tile_grid(h->levf[l]), s, h->levf[l], h->levf[l + 1].nx, om,
                       h->mgf_x[l], ec, l == 0 ? r : nullptr, l == 0 ? nullptr : h->mgf_b[l],
                       l == 0 ? z : nullptr, l == 0 ? nullptr : h->mgf_t[l], h->sc);
      ec = h->mgf_t[l];
    }
  } else if (h->cfg.mg_sweeps <= MG_FUSED_MAX_SWEEPS) {
    // fused legs: one launch down and one up per level above the first level
    // of at most MG_COARSE_MAX cells; from there to the last level the cycle
    // runs in one workgroup in LDS
    const int S = h->cfg.mg_sweeps, last = h->nlev - 1;
    const double om = h->cfg.mg_omega;
    const bool rb = h->cfg.mg_smoother == 1;
    auto level_b = [&](int l) -> const double* { return l == 0 ? r : h->mg_b[l]; };
    const int l0 = h->mg_cycle_l0 >= 0 ? h->mg_cycle_l0 : last;
    for (int l = 0; l < l0; ++l) {
      if (rb)
        launch_mg_down_rb(S, h->mg_tile_shape, s, h->lev[l], h->lev[l + 1], level_b(l),
                          h->mg_x[l], h->mg_b[l + 1], h->sc);
      else
        launch_mg_down(S, h->mg_tile_shape, s, h->lev[l], h->lev[l + 1], om, level_b(l),
                       h->mg_x[l], h->mg_b[l + 1], h->sc);
    }
    double* ec = l0 == 0 ? z : h->mg_x[l0];
    if (h->mg_cycle_l0 >= 0) {
      MgCycle C{};
      C.nl = last - l0 + 1;
      C.rb = rb ? 1 : 0;
      for (int l = 0; l < C.nl; ++l) C.lev[l] = h->lev[l0 + l];
      const size_t lds = mg_cycle_layout(C);
      launch_mg_coarse_cycle(C, lds, om, S, level_b(l0), ec, h->sc, s);
    } else if (rb) {
      // the last level alone: S forward and S reverse sweeps from zero
      k_mg_zero<<<blocks(h->lev[last].n), TPB, 0, s>>>(h->lev[last].n, ec, h->sc);
      for (int k = 0; k < 2 * S; ++k) rbgs_sweep(h, last, level_b(last), ec, k >= S, s);
    } else {
      double* out = vcycle_jacobi(h, last, level_b(last), ec, h->mg_t[last], s);
      if (out != ec) return fail(HOC_ERR_DEVICE, "pcg_mg: coarse result left its buffer");
    }
    for (int l = l0 - 1; l >= 0; --l) {
      double* out = l == 0 ? z : h->mg_t[l];
      if (rb)
        launch_mg_up_rb(S, h->mg_tile_shape, s, h->lev[l], h->lev[l + 1].nx, h->mg_x[l], ec,
                        level_b(l), out, h->sc);
      else
        launch_mg_up(S, h->mg_tile_shape, s, h->lev[l], h->lev[l + 1].nx, om, h->mg_x[l], ec,
                     level_b(l), out, h->sc);
      ec = out;
    }
  } else {
    // 2 * mg_sweeps buffer swaps per level: the result is back in z
    double* out = vcycle_jacobi(h, 0, r, z, h->mg_t[0], s);
    if (out != z) return fail(HOC_ERR_DEVICE, "pcg_mg: V-cycle result left the z buffer");
  }
  LAUNCH_CHECK("preconditioner");
  return HOC_OK;
}

// The sync-free PCG of pcg_mg / pcg_rbgs: a fixed launch sequence with no
// host read-back (capturable in a HIP graph). Every phase returns at once
// after the device `done` flag is set, and a solve that has not converged
// within the budget counts as a failure.
int pcg_sync_free(hoc_stepper* h, const PcgArgs& a, unsigned np, hipStream_t s, int* iters_out,
                  int* conv_out) {
  const int max_iter = h->p.max_iter;
  k_pcg_init<<<np, PCG_WG, 0, s>>>(a);   // r = rhs * mask, x = 0, ||rhs||^2 parts
  k_mg_start<<<1, PCG_WG, 0, s>>>(a);
  TRY(precondition(h, h->r, h->z, s));
  k_pcg_rz<<<np, PCG_WG, 0, s>>>(a, 1);
  const int n_it = std::min(h->sync_free_budget, max_iter + 1);
  for (int it = 1; it <= n_it; ++it) {
    if (a.central) k_pcg_totals<<<1, PCG_WG, 0, s>>>(a, it == 1 ? 1 : 2);
    k_pcg_ap<<<np, PCG_WG, 0, s>>>(a, it);
    if (it <= max_iter) {
      if (a.central) k_pcg_totals<<<1, PCG_WG, 0, s>>>(a, 0);
      k_pcg_update<<<np, PCG_WG, 0, s>>>(a, it, 0);
      TRY(precondition(h, h->r, h->z, s));
      k_pcg_rz<<<np, PCG_WG, 0, s>>>(a, 0);
    }
  }
  k_pcg_budget_check<<<1, 64, 0, s>>>(a, n_it);
  LAUNCH_CHECK("pcg iteration (sync-free)");
  if (iters_out) *iters_out = -1;
  if (conv_out) *conv_out = -1;
  return HOC_OK;
}

// Whether a sync-free solve on stream s may be replayed from the handle's
// graph: not on the legacy default stream, not while the caller captures
// the stream itself (its own graph then holds the launches), not while a
// kernel is being profiled by name.
int sf_graph_usable(hoc_stepper* h, hipStream_t s, bool* ok) {
  *ok = false;
  if (s == nullptr || !h->prof_name.empty()) return HOC_OK;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(s, &st));
  *ok = st == hipStreamCaptureStatusNone;
  return HOC_OK;
}

// PCG-Jacobi on A = I - coef D[K G .]; x is the output (device). Path 0: one
// cooperative launch; path 1: one launch per phase, the host polling the
// device `done` flag between batches of iterations.
int helmholtz_solve(hoc_stepper* h, const double* Ku, const double* Kv, const double* inv_diag,
                    const double* rhs, double* x, hipStream_t s, int* iters_out,
                    int* conv_out) {
  HelmOp op{Ku, Kv, h->m2, h->p.helm_coef, h->p.helm_cx, h->p.helm_cy, h->p.dx2, h->p.dy2};
  PcgArgs a = pcg_args(h, op, inv_diag, rhs, x);
  const unsigned np = (unsigned)h->pcg_parts_all;
  if (np == 0) return fail(HOC_ERR_INVALID, "empty grid");
  const int kind = h->cfg.solver;
  if (kind != SOLVER_PCG_JACOBI) {
    // the levels of the step's own coefficients are built once per rebuild
    // of them (both Picard passes solve with the same Ku, Kv); any other
    // operands are built for this solve alone
    const bool own = Ku == h->Ku && Kv == h->Kv && inv_diag == h->inv_diag;
    if (!own || h->mg_gen != h->coeff_gen) {
      TRY(mg_build(h, Ku, Kv, inv_diag, s));
      h->mg_gen = own ? h->coeff_gen : -1;
    }
    const int max_iter = h->p.max_iter;
    if (h->sync_free_budget > 0 && kind != SOLVER_RBGS && kind != SOLVER_MULTIGRID) {
      bool replay = false;
      TRY(sf_graph_usable(h, s, &replay));
      if (!replay) return pcg_sync_free(h, a, np, s, iters_out, conv_out);
      const void* key[5] = {Ku, Kv, inv_diag, rhs, x};
      bool same = h->sf_exec != nullptr && h->sf_key_budget == h->sync_free_budget &&
                  h->sf_key_tile == h->mg_tile_shape;
      for (int q = 0; q < 5 && same; ++q) same = h->sf_key[q] == key[q];
      if (!same) {
        if (h->sf_exec) (void)hipGraphExecDestroy(h->sf_exec);
        if (h->sf_graph) (void)hipGraphDestroy(h->sf_graph);
        h->sf_exec = nullptr;
        h->sf_graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        const int rc = pcg_sync_free(h, a, np, s, iters_out, conv_out);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != HOC_OK || e != hipSuccess || !g) {
          if (g) (void)hipGraphDestroy(g);
          return rc != HOC_OK ? rc : fail(HOC_ERR_DEVICE, "capture of the sync-free solve failed");
        }
        const hipError_t ie = hipGraphInstantiate(&h->sf_exec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
          (void)hipGraphDestroy(g);
          h->sf_exec = nullptr;
          return fail(HOC_ERR_DEVICE, "instantiation of the sync-free solve failed");
        }
        h->sf_graph = g;
        for (int q = 0; q < 5; ++q) h->sf_key[q] = key[q];
        h->sf_key_budget = h->sync_free_budget;
        h->sf_key_tile = h->mg_tile_shape;
      }
      HIP_TRY(hipGraphLaunch(h->sf_exec, s));
      if (iters_out) *iters_out = -1;
      if (conv_out) *conv_out = -1;
      return HOC_OK;
    }
    k_pcg_init<<<np, PCG_WG, 0, s>>>(a);   // r = rhs * mask, x = 0, ||rhs||^2 parts
    k_mg_start<<<1, PCG_WG, 0, s>>>(a);
    LAUNCH_CHECK("solver init");
    if (kind == SOLVER_RBGS || kind == SOLVER_MULTIGRID) {
      // stationary iteration on x with the masked rhs (a.r); ||rhs - A x||
      PcgArgs ar = a;
      ar.rhs = h->r;
      int it = 1;
      int batch = std::max(1, std::min(h->last_iters, max_iter));
      while (true) {
        for (int b = 0; b < batch && it <= max_iter; ++b, ++it) {
          if (kind == SOLVER_MULTIGRID) vcycle_rbgs(h, 0, x, h->r, s);
          else rbgs_sweep(h, 0, h->r, x, false, s);
          k_mg_res_parts<<<np, PCG_WG, 0, s>>>(ar, h->lev[0], x);
          k_mg_check<<<1, PCG_WG, 0, s>>>(ar, it);
        }
        LAUNCH_CHECK("stationary iteration");
        int done = 0;
        TRY(poll_done(h, s, &done));
        if (done) break;
        if (it > max_iter) return fail(HOC_ERR_DEVICE, "solver did not report completion");
        batch = std::min(2 * batch, 64);  // slow stationary convergence: fewer polls
      }
      return pcg_finish(h, s, false, iters_out, conv_out, true);
    }
    // PCG with the RBGS or multigrid preconditioner
    TRY(precondition(h, h->r, h->z, s));
    k_pcg_rz<<<np, PCG_WG, 0, s>>>(a, 1);
    int it = 1;
    int batch = std::max(1, std::min(h->last_iters + 1, max_iter + 1));
    while (true) {
      for (int b = 0; b < batch && it <= max_iter + 1; ++b, ++it) {
        if (a.central) k_pcg_totals<<<1, PCG_WG, 0, s>>>(a, it == 1 ? 1 : 2);
        { ProfScope _ps(h, "k_pcg_ap", s); k_pcg_ap<<<np, PCG_WG, 0, s>>>(a, it); }
        if (it <= max_iter) {
          if (a.central) k_pcg_totals<<<1, PCG_WG, 0, s>>>(a, 0);
          { ProfScope _ps(h, "k_pcg_update", s); k_pcg_update<<<np, PCG_WG, 0, s>>>(a, it, 0); }
          TRY(precondition(h, h->r, h->z, s));
          k_pcg_rz<<<np, PCG_WG, 0, s>>>(a, 0);
        }
      }
      LAUNCH_CHECK("pcg iteration");
      int done = 0;
      TRY(poll_done(h, s, &done));
      if (done) break;
      if (it > max_iter + 1) return fail(HOC_ERR_DEVICE, "Helmholtz solve did not report completion");
      batch = 4;
    }
    return pcg_finish(h, s, false, iters_out, conv_out, true);
  }
  if (h->pcg_parts > 0 && !h->force_multikernel) {
    a.sc = h->sc;
    void* args[] = {&a};
    { ProfScope _ps(h, "k_pcg_persistent", s);
      HIP_TRY(hipLaunchCooperativeKernel((const void*)k_pcg_persistent, dim3(h->pcg_parts),
                                         dim3(PCG_WG), args, 0, s)); }
    return pcg_finish(h, s, true, iters_out, conv_out);
  }
  k_pcg_init<<<np, PCG_WG, 0, s>>>(a);
  k_pcg_init_z<<<np, PCG_WG, 0, s>>>(a);
  LAUNCH_CHECK("pcg init");
  const int max_iter = h->p.max_iter;
  // Iteration it's A launch also finishes iteration it-1, so max_iter
  // iterations need max_iter + 1 A launches; the device flag stops early.
  int it = 1;
  int batch = std::max(1, std::min(h->last_iters + 1, max_iter + 1));
  while (true) {
    for (int b = 0; b < batch && it <= max_iter + 1; ++b, ++it) {
      if (a.central) k_pcg_totals<<<1, PCG_WG, 0, s>>>(a, it == 1 ? 1 : 2);
      { ProfScope _ps(h, "k_pcg_ap", s); k_pcg_ap<<<np, PCG_WG, 0, s>>>(a, it); }
      if (it <= max_iter) {
        if (a.central) k_pcg_totals<<<1, PCG_WG, 0, s>>>(a, 0);
        ProfScope _ps(h, "k_pcg_update", s);
        k_pcg_update<<<np, PCG_WG, 0, s>>>(a, it, 1);
      }
    }
    LAUNCH_CHECK("pcg iteration");
    int done = 0;
    TRY(poll_done(h, s, &done));
    if (done) break;
    if (it > max_iter + 1)
      return fail(HOC_ERR_DEVICE, "Helmholtz solve did not report completion");
    batch = 4;
  }
  return pcg_finish(h, s, false, iters_out, conv_out, true);
}

}  // namespace

// ------------------------------------------------------- the 2D model
struct hoc_sw2d {
  hoc_sw2d_config cfg{};
  Sw2d S{};
};

// ------------------------------------- the constant-coefficient Helmholtz
struct hoc_helm2d {
  hoc_helm2d_config cfg{};
  std::vector<H2Level> lv;
  std::vector<double*> lx, lb;   // per level from 1: correction and restricted residual
  Reducer red;
  double inv_diag = 0.0;
  double *r = nullptr, *z = nullptr, *p = nullptr, *ap = nullptr, *tmp = nullptr;
  double *part = nullptr, *dot = nullptr;
  double* host = nullptr;        // pinned read-back of one dot product
  std::vector<void*> allocs;
  int alloc(double** ptr, long long count) {
    void* v = nullptr;
    HIP_TRY(hipMalloc(&v, std::max<long long>(count, 1) * (long long)sizeof(double)));
    allocs.push_back(v);
    *ptr = (double*)v;
    return HOC_OK;
  }
  ~hoc_helm2d() {
    for (void* v : allocs) (void)hipFree(v);
    if (host) (void)hipHostFree(host);
    red.release();
  }
};

// ------------------------------------------------------- the v0.2 3D model
struct hoc_m3 {
  hoc_m3_config cfg{};
  M3 M{};
  std::vector<double> q;        // host copy of the effective-depth profile
  double* coeffs = nullptr;     // device: sub_m, den_m, cst_m, sub_b, den_b, cst_b, q
  ~hoc_m3() {
    if (coeffs) (void)hipFree(coeffs);
  }
};

namespace {

H2Level h2_level(int nx, int ny, double dx, double dy, double coef) {
  H2Level L{};
  L.nx = nx;
  L.ny = ny;
  L.n = (long long)nx * ny;
  L.dx = dx;
  L.dy = dy;
  L.dx2 = std::pow(dx, 2.0);
  L.dy2 = std::pow(dy, 2.0);
  L.cx = coef / L.dx2;
  L.cy = coef / L.dy2;
  L.diag = 1.0 + 2.0 * L.cx + 2.0 * L.cy;
  L.coef = coef;
  return L;
}

// sum(a * b) over the finest level in NumPy's pairwise order, read back.
int h2_dot(hoc_helm2d* h, const double* a, const double* b, hipStream_t s, double* out) {
  if (h->red.nchunks > 0)
    k_pw_dot<<<(unsigned)h->red.nchunks, 512, 0, s>>>(a, b, h->red.plans, h->red.nfull, h->part);
  k_seq_sum<<<1, 1, 0, s>>>(h->part, h->red.nchunks, h->dot);
  LAUNCH_CHECK("helm2d dot");
  HIP_TRY(hipMemcpyAsync(h->host, h->dot, sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *out = h->host[0];
  return HOC_OK;
}

// n_sweeps red-black sweeps of level l on x (bit k of pattern: sweep k runs
// black first). Small levels run whole in one workgroup; larger even levels
// update each colour in place; larger odd levels go through tmp.
int h2_sweeps(hoc_helm2d* h, int l, double* x, const double* b, int n_sweeps, unsigned pattern,
              hipStream_t s) {
  const H2Level& L = h->lv[l];
  if (L.n <= H2_COARSE_MAX) {
    for (int s0 = 0; s0 < n_sweeps; s0 += 32) {
      const int m = std::min(32, n_sweeps - s0);
      k_h2_gs_block<<<1, H2_COARSE_THREADS, 0, s>>>(L, m, s0 < 32 ? pattern >> s0 : 0u, x, b);
    }
    LAUNCH_CHECK("k_h2_gs_block");
    return HOC_OK;
  }
  const bool even = (L.nx % 2 == 0) && (L.ny % 2 == 0);
  for (int k = 0; k < n_sweeps; ++k) {
    const int first = (k < 32 && ((pattern >> k) & 1u)) ? 1 : 0;
    for (int c = 0; c < 2; ++c) {
      const int colour = c == 0 ? first : 1 - first;
      if (even) {
        k_h2_gs_colour<<<blocks(L.n / 2), TPB, 0, s>>>(L, colour, x, b);
      } else {
        k_h2_gs_colour_copy<<<blocks(L.n), TPB, 0, s>>>(L, colour, x, b, h->tmp);
        k_h2_copy<<<blocks(L.n), TPB, 0, s>>>(L.n, h->tmp, x);
      }
    }
  }
  LAUNCH_CHECK("k_h2_gs_colour");
  return HOC_OK;
}

// HelmholtzMultigrid._vcycle (solvers.py:226-245): 2 pre-sweeps forward then
// backward, restriction, the coarser cycle, prolongation, 2 post-sweeps
// backward then forward; 40 forward sweeps on the coarsest level.
int h2_vcycle(hoc_helm2d* h, int l, double* x, const double* b, hipStream_t s) {
  const int last = (int)h->lv.size() - 1;
  if (l == last) return h2_sweeps(h, l, x, b, 40, 0u, s);
  TRY(h2_sweeps(h, l, x, b, 2, 0x2u, s));
  const H2Level& F = h->lv[l];
  const H2Level& Cl = h->lv[l + 1];
  k_h2_restrict<<<blocks(Cl.n), TPB, 0, s>>>(F, Cl, x, b, h->lb[l + 1], h->lx[l + 1]);
  LAUNCH_CHECK("k_h2_restrict");
  TRY(h2_vcycle(h, l + 1, h->lx[l + 1], h->lb[l + 1], s));
  k_h2_prolong<<<blocks(F.n), TPB, 0, s>>>(F, Cl, x, h->lx[l + 1]);
  LAUNCH_CHECK("k_h2_prolong");
  return h2_sweeps(h, l, x, b, 2, 0x1u, s);
}

// z = M^-1 r (HelmholtzPCG._precondition, solvers.py:75-86)
int h2_precondition(hoc_helm2d* h, const double* r, double* z, hipStream_t s) {
  const H2Level& L = h->lv[0];
  if (h->cfg.kind == 0) {
    k_h2_jacobi<<<blocks(L.n), TPB, 0, s>>>(L.n, h->inv_diag, r, z);
    LAUNCH_CHECK("k_h2_jacobi");
    return HOC_OK;
  }
  k_h2_fill<<<blocks(L.n), TPB, 0, s>>>(L.n, 0.0, z);
  LAUNCH_CHECK("k_h2_fill");
  return h2_sweeps(h, 0, z, r, 2, 0x2u, s);
}

}  // namespace

extern "C" {

int hoc_sw2d_create(const hoc_sw2d_config* cfg, hoc_sw2d** out) {
  if (!cfg || !out) return fail(HOC_ERR_INVALID, "null argument");
  *out = nullptr;
  if (cfg->nx < 2 || cfg->ny < 2) return fail(HOC_ERR_INVALID, "grid must have nx, ny >= 2");
  if (!(cfg->dx > 0.0) || !(cfg->dy > 0.0) || !(cfg->dt > 0.0))
    return fail(HOC_ERR_INVALID, "dx, dy and dt must be positive");
  if (cfg->scheme != 0 && cfg->scheme != 1)
    return fail(HOC_ERR_INVALID, "scheme must be theta (0) or fb (1)");
  if (cfg->scheme == 0 && !(cfg->theta > 0.0 && cfg->theta <= 1.0))
    return fail(HOC_ERR_INVALID, "theta must lie in (0, 1]");
  hoc_sw2d* h = new hoc_sw2d();
  h->cfg = *cfg;
  Sw2d& S = h->S;
  S.nx = cfg->nx;
  S.ny = cfg->ny;
  S.n = (long long)cfg->nx * cfg->ny;
  S.dx = cfg->dx;
  S.dy = cfg->dy;
  S.g = cfg->g;
  S.f = cfg->f;
  S.neg_f = -cfg->f;
  S.H = cfg->H;
  S.dt = cfg->dt;
  // schemes.py evaluates these products left to right on Python floats
  const double th = cfg->theta, tc = cfg->theta_cor;
  S.th = th;
  S.tc = tc;
  S.one_m_tc = 1.0 - tc;
  S.one_m_th = 1.0 - th;
  S.dt_f = cfg->dt * cfg->f;
  S.g_dt_1mth = cfg->g * cfg->dt * (1.0 - th);
  S.dt_H = cfg->dt * cfg->H;
  S.g_dt_th = cfg->g * cfg->dt * th;
  *out = h;
  return HOC_OK;
}

int hoc_sw2d_destroy(hoc_sw2d* h) {
  if (h && release_is_safe()) delete h;
  return HOC_OK;
}

int hoc_sw2d_momentum(hoc_sw2d* h, const double* u0, const double* v0, const double* eta0,
                      const double* u_it, const double* v_it, double* out_u, double* out_v,
                      void* stream) {
  if (!h || !u0 || !v0 || !eta0 || !u_it || !v_it || !out_u || !out_v)
    return fail(HOC_ERR_INVALID, "null argument");
  k_sw_momentum<<<blocks(h->S.n), TPB, 0, (hipStream_t)stream>>>(
      h->S, h->cfg.scheme == 0 ? 1 : 0, u0, v0, eta0, u_it, v_it, out_u, out_v);
  LAUNCH_CHECK("k_sw_momentum");
  return HOC_OK;
}

int hoc_sw2d_rhs(hoc_sw2d* h, const double* eta0, const double* u0, const double* v0,
                 const double* gu, const double* gv, double* rhs, void* stream) {
  if (!h || !eta0 || !u0 || !v0 || !gu || !gv || !rhs) return fail(HOC_ERR_INVALID, "null argument");
  k_sw_rhs<<<blocks(h->S.n), TPB, 0, (hipStream_t)stream>>>(h->S, eta0, u0, v0, gu, gv, rhs);
  LAUNCH_CHECK("k_sw_rhs");
  return HOC_OK;
}

int hoc_sw2d_backsub(hoc_sw2d* h, const double* gu, const double* gv, const double* eta,
                     double* out_u, double* out_v, void* stream) {
  if (!h || !gu || !gv || !eta || !out_u || !out_v) return fail(HOC_ERR_INVALID, "null argument");
  k_sw_backsub<<<blocks(h->S.n), TPB, 0, (hipStream_t)stream>>>(h->S, gu, gv, eta, out_u, out_v);
  LAUNCH_CHECK("k_sw_backsub");
  return HOC_OK;
}

int hoc_sw2d_fb_eta(hoc_sw2d* h, const double* eta0, const double* u, const double* v,
                    double* out, void* stream) {
  if (!h || !eta0 || !u || !v || !out) return fail(HOC_ERR_INVALID, "null argument");
  k_sw_fb_eta<<<blocks(h->S.n), TPB, 0, (hipStream_t)stream>>>(h->S, eta0, u, v, out);
  LAUNCH_CHECK("k_sw_fb_eta");
  return HOC_OK;
}

int hoc_helm2d_create(const hoc_helm2d_config* cfg, hoc_helm2d** out) {
  if (!cfg || !out) return fail(HOC_ERR_INVALID, "null argument");
  *out = nullptr;
  if (cfg->nx < 2 || cfg->ny < 2) return fail(HOC_ERR_INVALID, "grid must have nx, ny >= 2");
  if (!(cfg->dx > 0.0) || !(cfg->dy > 0.0))
    return fail(HOC_ERR_INVALID, "dx and dy must be positive");
  if (!(cfg->coef >= 0.0)) return fail(HOC_ERR_INVALID, "coef must be >= 0");
  if (cfg->kind < 0 || cfg->kind > 3)
    return fail(HOC_ERR_INVALID, "kind must be 0 pcg_jacobi, 1 pcg_rbgs, 2 rbgs or 3 multigrid");
  if (cfg->kind > 0 && (cfg->nx % 2 || cfg->ny % 2))
    return fail(HOC_ERR_INVALID, "the red-black solvers need even nx and ny for a consistent "
                                 "colouring under periodic wrap");
  if (cfg->max_iter < 1) return fail(HOC_ERR_INVALID, "max_iter must be >= 1");
  hoc_helm2d* h = new hoc_helm2d();
  h->cfg = *cfg;
  int nx = cfg->nx, ny = cfg->ny;
  double dx = cfg->dx, dy = cfg->dy;
  // HelmholtzMultigrid.__init__ (solvers.py:213-220): halve while both sizes
  // are even and the smaller is above 4
  while (true) {
    h->lv.push_back(h2_level(nx, ny, dx, dy, cfg->coef));
    if (cfg->kind != 3 || nx % 2 || ny % 2 || std::min(nx, ny) <= 4) break;
    nx /= 2;
    ny /= 2;
    dx = dx * 2;
    dy = dy * 2;
  }
  h->inv_diag = 1.0 / (1.0 + cfg->coef * (2.0 / h->lv[0].dx2 + 2.0 / h->lv[0].dy2));
  const long long n = h->lv[0].n;
  int rc = h->red.init(n);
  rc = rc ? rc : h->alloc(&h->r, n);
  rc = rc ? rc : h->alloc(&h->z, n);
  rc = rc ? rc : h->alloc(&h->p, n);
  rc = rc ? rc : h->alloc(&h->ap, n);
  rc = rc ? rc : h->alloc(&h->tmp, n);
  rc = rc ? rc : h->alloc(&h->part, std::max<long long>(h->red.nchunks, 1));
  rc = rc ? rc : h->alloc(&h->dot, 1);
  h->lx.assign(h->lv.size(), nullptr);
  h->lb.assign(h->lv.size(), nullptr);
  for (size_t l = 1; l < h->lv.size() && rc == HOC_OK; ++l) {
    rc = h->alloc(&h->lx[l], h->lv[l].n);
    rc = rc ? rc : h->alloc(&h->lb[l], h->lv[l].n);
  }
  if (rc == HOC_OK && hipHostMalloc((void**)&h->host, sizeof(double), hipHostMallocDefault) != hipSuccess)
    rc = fail(HOC_ERR_DEVICE, "hipHostMalloc failed");
  if (rc != HOC_OK) {
    delete h;
    return rc;
  }
  *out = h;
  return HOC_OK;
}

int hoc_helm2d_destroy(hoc_helm2d* h) {
  if (h && release_is_safe()) delete h;
  return HOC_OK;
}

int hoc_helm2d_levels(hoc_helm2d* h, int32_t* n_levels) {
  if (!h || !n_levels) return fail(HOC_ERR_INVALID, "null argument");
  *n_levels = (int32_t)h->lv.size();
  return HOC_OK;
}

int hoc_helm2d_solve(hoc_helm2d* h, const double* rhs, double* x, int32_t* iterations,
                     double* residual_out, int32_t* converged, void* stream) {
  if (!h || !rhs || !x) return fail(HOC_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const H2Level& L = h->lv[0];
  const long long n = L.n;
  const int kind = h->cfg.kind, max_iter = h->cfg.max_iter;
  const double rtol = h->cfg.rtol;
  k_h2_fill<<<blocks(n), TPB, 0, s>>>(n, 0.0, x);
  LAUNCH_CHECK("k_h2_fill");
  double nb2 = 0.0;
  TRY(h2_dot(h, rhs, rhs, s, &nb2));
  double norm_b = std::sqrt(nb2);
  int it_out = max_iter, conv = 0;
  double residual = 0.0;
  if (kind <= 1) {
    // HelmholtzPCG.solve (solvers.py:91-117); r = rhs - A 0 = rhs exactly
    if (norm_b == 0.0) {
      it_out = 0;
      conv = 1;
    } else {
      k_h2_copy<<<blocks(n), TPB, 0, s>>>(n, rhs, h->r);
      TRY(h2_precondition(h, h->r, h->z, s));
      k_h2_copy<<<blocks(n), TPB, 0, s>>>(n, h->z, h->p);
      double rz = 0.0;
      TRY(h2_dot(h, h->r, h->z, s, &rz));
      for (int it = 1; it <= max_iter; ++it) {
        k_h2_apply_dg<<<blocks(n), TPB, 0, s>>>(L, h->p, h->ap);
        double pap = 0.0;
        TRY(h2_dot(h, h->p, h->ap, s, &pap));
        const double alpha = rz / pap;
        k_h2_xr<<<blocks(n), TPB, 0, s>>>(n, alpha, h->p, h->ap, x, h->r);
        double rr = 0.0;
        TRY(h2_dot(h, h->r, h->r, s, &rr));
        residual = std::sqrt(rr) / norm_b;
        if (residual < rtol) {
          it_out = it;
          conv = 1;
          break;
        }
        TRY(h2_precondition(h, h->r, h->z, s));
        double rz_new = 0.0;
        TRY(h2_dot(h, h->r, h->z, s, &rz_new));
        k_h2_p<<<blocks(n), TPB, 0, s>>>(n, rz_new / rz, h->z, h->p);
        rz = rz_new;
      }
      LAUNCH_CHECK("helm2d pcg");
    }
  } else {
    // HelmholtzRBGS.solve / HelmholtzMultigrid.solve (solvers.py:165-184,
    // 247-258): norm_b = |rhs| or 1
    if (norm_b == 0.0) norm_b = 1.0;
    for (int it = 1; it <= max_iter; ++it) {
      if (kind == 2) {
        TRY(h2_sweeps(h, 0, x, rhs, 1, 0u, s));
        if (!(it % 10 == 0 || it == max_iter)) continue;
      } else {
        TRY(h2_vcycle(h, 0, x, rhs, s));
      }
      k_h2_resid<<<blocks(n), TPB, 0, s>>>(L, x, rhs, h->r);
      LAUNCH_CHECK("k_h2_resid");
      double rr = 0.0;
      TRY(h2_dot(h, h->r, h->r, s, &rr));
      residual = std::sqrt(rr) / norm_b;
      if (residual < rtol) {
        it_out = it;
        conv = 1;
        break;
      }
    }
  }
  if (iterations) *iterations = it_out;
  if (residual_out) *residual_out = residual;
  if (converged) *converged = conv;
  return HOC_OK;
}

int hoc_m3_create(const hoc_m3_config* cfg, hoc_m3** out) {
  if (!cfg || !out) return fail(HOC_ERR_INVALID, "null argument");
  *out = nullptr;
  const hoc_m3_config& c = *cfg;
  if (c.nx < 2 || c.ny < 2 || c.nz < 1) return fail(HOC_ERR_INVALID, "grid must have nx, ny >= 2 and nz >= 1");
  if (!(c.dx > 0.0) || !(c.dy > 0.0) || !(c.dz > 0.0) || !(c.dt > 0.0) || !(c.H > 0.0))
    return fail(HOC_ERR_INVALID, "dx, dy, dz, dt and H must be positive");
  if (c.scheme < 0 || c.scheme > 2) return fail(HOC_ERR_INVALID, "scheme must be theta (0), fb (1) or split_explicit (2)");
  if (c.advection < 0 || c.advection > 2)
    return fail(HOC_ERR_INVALID, "advection must be none (0), centered2 (1) or upwind1 (2)");
  if (c.tracers != 0 && c.tracers != 1) return fail(HOC_ERR_INVALID, "tracers must be buoyancy (0) or TS (1)");
  if (c.scheme == 2 && c.n_split < 1) return fail(HOC_ERR_INVALID, "split_explicit needs n_split >= 1");
  hoc_m3* h = new hoc_m3();
  h->cfg = c;
  M3& M = h->M;
  M.nx = c.nx;
  M.ny = c.ny;
  M.nz = c.nz;
  M.n2 = (long long)c.nx * c.ny;
  M.n3 = M.n2 * c.nz;
  M.dx = c.dx;
  M.dy = c.dy;
  M.dz = c.dz;
  M.dx2 = std::pow(c.dx, 2.0);
  M.dy2 = std::pow(c.dy, 2.0);
  M.neg_dz = -c.dz;
  M.dt = c.dt;
  M.f = c.f;
  M.neg_f = -c.f;
  M.g = c.g;
  M.H = c.H;
  // model3d.py evaluates these scalar products left to right on Python floats
  const double th = c.scheme == 0 ? c.theta : 0.0, tc = c.theta_cor, thv = c.theta_v;
  M.th = th;
  M.one_m_th = 1.0 - th;
  M.tc = tc;
  M.one_m_tc = 1.0 - tc;
  M.c_eta = c.g * c.dt * (1.0 - th);
  M.c_diff = (1.0 - thv) * c.dt;
  M.c_bs = c.g * c.dt * th;
  M.nu = c.nu;
  M.kappa = c.kappa;
  M.bd = c.bottom_drag;
  M.tau_u = c.tau_x / c.rho0;
  M.tau_v = c.tau_y / c.rho0;
  M.s0u = thv * c.dt * M.tau_u / c.dz;
  M.s0v = thv * c.dt * M.tau_v / c.dz;
  M.A_h = c.A_h;
  M.K_h = c.K_h;
  M.adv = c.advection;
  M.neg_N2 = -c.N2;
  M.flux_T = c.q_heat / (c.rho0 * c.cp);
  M.flux_S = c.q_salt;
  M.s0T = thv * c.dt * M.flux_T / c.dz;
  M.s0S = thv * c.dt * M.flux_S / c.dz;
  M.s0b = thv * c.dt * 0.0 / c.dz;
  M.alpha_T = c.alpha_T;
  M.beta_S = c.beta_S;
  M.T0 = c.T0;
  M.S0 = c.S0;
  M.ddt = c.scheme == 2 ? c.dt / c.n_split : 0.0;
  M.gh = c.g * c.H;
  // diffusion_coeffs + the column-independent part of thomas_batched
  // (vertical.py:17-72): sub, the pivots (den[0] = diag[0]) and c*.
  const int nz = c.nz;
  auto factor = [&](double visc, double drag, double* sub, double* den, double* cst) {
    const double fac = thv * c.dt / std::pow(c.dz, 2.0);
    std::vector<double> sup(nz), diag(nz);
    for (int k = 0; k < nz; ++k) {
      const double a_top = fac * visc, a_bot = fac * visc;
      sub[k] = -a_top;
      sup[k] = -a_bot;
      diag[k] = 1.0 + a_top + a_bot;
    }
    diag[0] = 1.0 + fac * visc;
    sub[0] = 0.0;
    diag[nz - 1] = 1.0 + fac * visc + thv * c.dt * drag / c.dz;
    sup[nz - 1] = 0.0;
    cst[0] = sup[0] / diag[0];
    den[0] = diag[0];
    for (int k = 1; k < nz; ++k) {
      den[k] = diag[k] - sub[k] * cst[k - 1];
      cst[k] = sup[k] / den[k];
    }
  };
  std::vector<double> hc(7 * (size_t)nz);
  double *sub_m = &hc[0], *den_m = &hc[nz], *cst_m = &hc[2 * nz];
  double *sub_b = &hc[3 * nz], *den_b = &hc[4 * nz], *cst_b = &hc[5 * nz], *qv = &hc[6 * nz];
  factor(c.nu, c.bottom_drag, sub_m, den_m, cst_m);
  factor(c.kappa, 0.0, sub_b, den_b, cst_b);
  // q = A^-1 1 with the momentum factors (model3d.py:125-126)
  std::vector<double> d(nz);
  for (int k = 0; k < nz; ++k) d[k] = k == 0 ? 1.0 / den_m[0] : (1.0 - sub_m[k] * d[k - 1]) / den_m[k];
  qv[nz - 1] = d[nz - 1];
  for (int k = nz - 2; k >= 0; --k) qv[k] = d[k] - cst_m[k] * qv[k + 1];
  h->q.assign(qv, qv + nz);
  if (hipMalloc((void**)&h->coeffs, hc.size() * sizeof(double)) != hipSuccess ||
      hipMemcpy(h->coeffs, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    delete h;
    return fail(HOC_ERR_DEVICE, "device allocation of the column factors failed");
  }
  M.sub_m = h->coeffs;
  M.den_m = h->coeffs + nz;
  M.cst_m = h->coeffs + 2 * nz;
  M.sub_b = h->coeffs + 3 * nz;
  M.den_b = h->coeffs + 4 * nz;
  M.cst_b = h->coeffs + 5 * nz;
  M.q = h->coeffs + 6 * nz;
  *out = h;
  return HOC_OK;
}

int hoc_m3_destroy(hoc_m3* h) {
  if (h && release_is_safe()) delete h;
  return HOC_OK;
}

int hoc_m3_q(hoc_m3* h, double* q) {
  if (!h || !q) return fail(HOC_ERR_INVALID, "null argument");
  std::copy(h->q.begin(), h->q.end(), q);
  return HOC_OK;
}

int hoc_m3_prelude(hoc_m3* h, const double* u, const double* v, const double* b, double* phi,
                   double* U0, double* V0, double* w, double* st_u, double* st_v, void* stream) {
  if (!h || !u || !v || !b || !phi || !U0 || !V0 || !w || !st_u || !st_v)
    return fail(HOC_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const M3& M = h->M;
  k_m3_phi<<<blocks(M.n2), TPB, 0, s>>>(M, b, phi);
  k_m3_depth<<<blocks(M.n2), TPB, 0, s>>>(M, u, v, U0, V0);
  k_m3_w<<<blocks(M.n2), TPB, 0, s>>>(M, u, v, w);
  k_m3_stress<<<blocks(M.n3), TPB, 0, s>>>(M, u, v, w, st_u, st_v);
  LAUNCH_CHECK("m3 prelude");
  return HOC_OK;
}

int hoc_m3_predict(hoc_m3* h, const double* u0, const double* v0, const double* phi,
                   const double* eta, const double* u_it, const double* v_it, const double* st_u,
                   const double* st_v, double* gh_u, double* gh_v, double* GU, double* GV,
                   double* cor_x, double* cor_y, void* stream) {
  if (!h || !u0 || !v0 || !phi || !eta || !u_it || !v_it || !st_u || !st_v || !gh_u || !gh_v ||
      !GU || !GV)
    return fail(HOC_ERR_INVALID, "null argument");
  const int split = h->cfg.scheme == 2 ? 1 : 0;
  if (split && (!cor_x || !cor_y)) return fail(HOC_ERR_INVALID, "split_explicit needs cor_x and cor_y");
  k_m3_predict<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(
      h->M, split, u0, v0, phi, eta, u_it, v_it, st_u, st_v, gh_u, gh_v, GU, GV, cor_x, cor_y);
  LAUNCH_CHECK("k_m3_predict");
  return HOC_OK;
}

int hoc_m3_rhs(hoc_m3* h, const double* eta, const double* U0, const double* V0,
               const double* GU, const double* GV, double* rhs, void* stream) {
  if (!h || !eta || !U0 || !V0 || !GU || !GV || !rhs) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_rhs<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, eta, U0, V0, GU, GV, rhs);
  LAUNCH_CHECK("k_m3_rhs");
  return HOC_OK;
}

int hoc_m3_backsub(hoc_m3* h, const double* gh_u, const double* gh_v, const double* eta,
                   double* u, double* v, void* stream) {
  if (!h || !gh_u || !gh_v || !eta || !u || !v) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_backsub<<<blocks(h->M.n3), TPB, 0, (hipStream_t)stream>>>(h->M, gh_u, gh_v, eta, u, v);
  LAUNCH_CHECK("k_m3_backsub");
  return HOC_OK;
}

int hoc_m3_eta_div(hoc_m3* h, double dt_sub, const double* eta, const double* U,
                   const double* V, double* out, void* stream) {
  if (!h || !eta || !U || !V || !out) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_eta_div<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, dt_sub, eta, U, V, out);
  LAUNCH_CHECK("k_m3_eta_div");
  return HOC_OK;
}

int hoc_m3_forcing(hoc_m3* h, const double* GU, const double* GV, const double* U0,
                   const double* V0, const double* cor_x, const double* cor_y, double* fx,
                   double* fy, void* stream) {
  if (!h || !GU || !GV || !U0 || !V0 || !cor_x || !cor_y || !fx || !fy)
    return fail(HOC_ERR_INVALID, "null argument");
  k_m3_forcing<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, GU, GV, U0, V0, cor_x,
                                                                 cor_y, fx, fy);
  LAUNCH_CHECK("k_m3_forcing");
  return HOC_OK;
}

int hoc_m3_bt_momentum(hoc_m3* h, const double* u_i, const double* v_i, const double* e,
                       const double* ui, const double* vi, const double* fx, const double* fy,
                       double* ui_out, double* vi_out, void* stream) {
  if (!h || !u_i || !v_i || !e || !ui || !vi || !fx || !fy || !ui_out || !vi_out)
    return fail(HOC_ERR_INVALID, "null argument");
  if (ui_out == ui || vi_out == vi || ui_out == u_i || vi_out == v_i)
    return fail(HOC_ERR_INVALID, "barotropic momentum outputs must not alias its inputs");
  k_m3_bt_momentum<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, u_i, v_i, e, ui, vi,
                                                                     fx, fy, ui_out, vi_out);
  LAUNCH_CHECK("k_m3_bt_momentum");
  return HOC_OK;
}

int hoc_m3_bt_acc(hoc_m3* h, const double* u_i, const double* v_i, double* acc_u, double* acc_v,
                  void* stream) {
  if (!h || !u_i || !v_i || !acc_u || !acc_v) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_bt_acc<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, u_i, v_i, acc_u, acc_v);
  LAUNCH_CHECK("k_m3_bt_acc");
  return HOC_OK;
}

int hoc_m3_split_final(hoc_m3* h, int32_t mean_n, const double* gh_u, const double* gh_v,
                       const double* Ub, const double* Vb, double* u, double* v, void* stream) {
  if (!h || !gh_u || !gh_v || !Ub || !Vb || !u || !v) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_split_final<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, mean_n, gh_u, gh_v, Ub,
                                                                     Vb, u, v);
  LAUNCH_CHECK("k_m3_split_final");
  return HOC_OK;
}

int hoc_m3_w(hoc_m3* h, const double* u, const double* v, double* w, void* stream) {
  if (!h || !u || !v || !w) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_w<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, u, v, w);
  LAUNCH_CHECK("k_m3_w");
  return HOC_OK;
}

int hoc_m3_tracer(hoc_m3* h, int32_t which, const double* c, const double* u, const double* v,
                  const double* w, double* out, void* stream) {
  if (!h || !c || !u || !v || !w || !out) return fail(HOC_ERR_INVALID, "null argument");
  if (which < 0 || which > 2) return fail(HOC_ERR_INVALID, "which must be 0 (b), 1 (T) or 2 (S)");
  if (out == c) return fail(HOC_ERR_INVALID, "tracer output must not alias its input");
  k_m3_tracer<<<blocks(h->M.n2), TPB, 0, (hipStream_t)stream>>>(h->M, which, c, u, v, w, out);
  LAUNCH_CHECK("k_m3_tracer");
  return HOC_OK;
}

int hoc_m3_eos(hoc_m3* h, const double* T, const double* S, double* b, void* stream) {
  if (!h || !T || !S || !b) return fail(HOC_ERR_INVALID, "null argument");
  k_m3_eos<<<blocks(h->M.n3), TPB, 0, (hipStream_t)stream>>>(h->M, T, S, b);
  LAUNCH_CHECK("k_m3_eos");
  return HOC_OK;
}

const char* hoc_last_error(void) { return g_error.c_str(); }

int hoc_stepper_create(const hoc_config* cfg, const hoc_geometry* g, hoc_stepper** out) {
  if (!cfg || !g || !out) return fail(HOC_ERR_INVALID, "null argument");
  TRY(validate(*cfg));
  hoc_stepper* h = new (std::nothrow) hoc_stepper();
  if (!h) return fail(HOC_ERR_DEVICE, "out of host memory");
  h->cfg = *cfg;
  Dims& d = h->d;
  d.nx = cfg->nx;
  d.ny = cfg->ny;
  d.nz = cfg->nz;
  d.n2 = (long long)cfg->nx * cfg->ny;
  d.n3 = d.n2 * cfg->nz;
  d.ni = d.n2 * (cfg->nz + 1);
  derive_params(*cfg, h->p);
  h->mg_tile_shape = d.n2 <= 512LL * 512LL ? 2 : 0;
  h->fp32_parts = cfg->precision == 0 ? 0
                  : cfg->precision == 1 ? HOC_FP32_MIXED
                                        : HOC_FP32_MIXED | HOC_FP32_MOMENTUM_FLUX;
  h->p.eos_mixed = (h->fp32_parts & HOC_FP32_EOS) != 0;
  auto cleanup = [&](int rc) {
    delete h;
    return rc;
  };
#define A(ptr, n)                        \
  do {                                   \
    int _rc = h->alloc(&(ptr), (n));     \
    if (_rc != HOC_OK) return cleanup(_rc); \
  } while (0)
  if (cfg->precision >= 1) {
    float G[4][7][7];
    teos_anomaly_table(cfg->T0, cfg->S0, h->p.eos_ref, h->p.eos_lin, G, &h->p.ss_ref);
    float* dg = nullptr;
    A(dg, 4 * 7 * 7);
    if (upload(dg, &G[0][0][0], 4 * 7 * 7, "eos table") != HOC_OK) return cleanup(HOC_ERR_DEVICE);
    h->p.eos_g = dg;
  }
  A(h->m2, d.n2);
  A(h->mu2, d.n2);
  A(h->mv2, d.n2);
  A(h->m3, d.n3);
  A(h->m3u, d.n3);
  A(h->m3v, d.n3);
  A(h->dz3, d.n3);
  A(h->dz3u, d.n3);
  A(h->dz3v, d.n3);
  A(h->zc, d.n3);
  A(h->dzi, d.ni);
  A(h->tau_x, d.n2);
  A(h->tau_y, d.n2);
  A(h->tau_u, d.n2);
  A(h->tau_v, d.n2);
  A(h->ustar2, d.n2);
  A(h->kappa_b, d.n2);
  A(h->sub_m, d.n3);
  A(h->den_m, d.n3);
  A(h->cst_m, d.n3);
  A(h->cst_b, d.n3);
  A(h->q, d.n3);
  A(h->Ku, d.n2);
  A(h->Kv, d.n2);
  A(h->inv_diag, d.n2);
  A(h->phi, d.n3);
  A(h->w, d.ni);
  A(h->pgx, d.n3);
  A(h->pgy, d.n3);
  A(h->avu, d.n3);
  A(h->avv, d.n3);
  A(h->dzu, d.n3);
  A(h->dzv, d.n3);
  A(h->stu, d.n3);
  A(h->stv, d.n3);
  A(h->U0, d.n2);
  A(h->V0, d.n2);
  if (cfg->scheme == 2) {
    A(h->Hu, d.n2);
    A(h->Hv, d.n2);
    A(h->bfx, d.n2);
    A(h->bfy, d.n2);
    for (int q = 0; q < 3; ++q) {
      A(h->bU[q], d.n2);
      A(h->bV[q], d.n2);
    }
    A(h->be[0], d.n2);
    A(h->be[1], d.n2);
    A(h->acc_u, d.n2);
    A(h->acc_v, d.n2);
  }
  A(h->div0, d.n2);
  A(h->Ug, d.n2);
  A(h->Vg, d.n2);
  A(h->rhs, d.n2);
  A(h->gh_u[0], d.n3);
  A(h->gh_v[0], d.n3);
  A(h->gh_u[1], d.n3);
  A(h->gh_v[1], d.n3);
  A(h->s_n2, d.ni);
  A(h->s_sh, d.ni);
  A(h->s_l, d.ni);
  A(h->s_le, d.ni);
  A(h->s_c, d.ni);
  A(h->r, d.n2);
  A(h->z, d.n2);
  A(h->pp, d.n2);
  A(h->pp2, d.n2);
  A(h->Ap, d.n2);
  A(h->bar, GRID_BAR_WORDS);
  A(h->bar_err, 1);
  A(h->sc, 1);
  A(h->dcount, 4);
#undef A
  {
    int rc = h->red.init(d.n2);
    if (rc != HOC_OK) return cleanup(rc);
    double* pa = nullptr;
    double* pb = nullptr;
    rc = h->alloc(&pa, h->red.nchunks);
    if (rc == HOC_OK) rc = h->alloc(&pb, h->red.nchunks);
    if (rc != HOC_OK) return cleanup(rc);
    h->part_a = pa;
    h->part_b = pb;
    // persistent solver: split each full chunk into `sub` subtrees so that
    // about 256-512 single-wave workgroups share the work
    const long long nfull = h->red.nfull;
    // parts of 8192 / PCG_MAX_SUB = 512 elements: one element per thread
    // in two passes at most, ~8 waves per CU at 512 x 512
    const int sub = PCG_MAX_SUB;
    PwPlan tail_plan = make_plan((int)(d.n2 % PW_CHUNK));
    const long long parts = nfull * sub + (h->red.nchunks > nfull ? tail_plan.nleaves : 0);
    h->pcg_sub = sub;
    h->pcg_parts_all = (int)parts;
    int dev = 0, ncu = 0, coop = 0, per_cu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_pcg_persistent,
                                                     PCG_WG, 0) == hipSuccess &&
        coop && parts <= (long long)ncu * per_cu) {
      h->pcg_parts = (int)parts;
    }
    if (cfg->closure_tke && !cfg->mxl_recursive) {
      const int lrc = closure_lengths_prepare(d.nz);
      if (lrc == -1)
        return cleanup(fail(HOC_ERR_UNSUPPORTED, "nz too large for the integral mixing-length "
                                                 "kernel (at most 159 levels)"));
      if (lrc != 0) return cleanup(fail(HOC_ERR_DEVICE, "hipFuncSetAttribute failed"));
    }
    if (h->red.nchunks > PCG_MAX_CHUNKS)
      return cleanup(fail(HOC_ERR_UNSUPPORTED, "grid too large for the Helmholtz solver "
                                               "(more than 1024 x 8192 columns)"));
    double* pc = nullptr;
    rc = h->alloc(&pc, std::max<long long>(parts, h->red.nchunks));
    if (rc != HOC_OK) return cleanup(rc);
    h->part_c = pc;
    if (parts > h->red.nchunks) {  // part_a / part_b must hold one value per part
      double *pa2 = nullptr, *pb2 = nullptr;
      rc = h->alloc(&pa2, parts);
      if (rc == HOC_OK) rc = h->alloc(&pb2, parts);
      if (rc != HOC_OK) return cleanup(rc);
      h->part_a = pa2;
      h->part_b = pb2;
    }
    if (hipHostMalloc((void**)&h->host_flag, 8 * sizeof(int), hipHostMallocDefault) != hipSuccess)
      return cleanup(fail(HOC_ERR_DEVICE, "hipHostMalloc failed"));
  }
  {
    // multigrid levels: the reference's rule for `multigrid` (halve while
    // both sides are even and stay >= 8), mg_levels for `pcg_mg`
    int nx = d.nx, ny = d.ny, nl = 1;
    if (cfg->solver == SOLVER_MULTIGRID) {
      while (nl < hoc_stepper::MAXLEV && nx % 2 == 0 && ny % 2 == 0 && nx / 2 >= 8 && ny / 2 >= 8) {
        nx /= 2;
        ny /= 2;
        ++nl;
      }
    } else if (cfg->solver == SOLVER_PCG_MG) {
      while (nl < std::min(cfg->mg_levels, hoc_stepper::MAXLEV) && nx % 2 == 0 && ny % 2 == 0 &&
             nx >= 4 && ny >= 4) {
        nx /= 2;
        ny /= 2;
        ++nl;
      }
    }
    h->nlev = nl;
    int lx = d.nx, ly = d.ny;
    double ldx = cfg->dx, ldy = cfg->dy;
    for (int l = 0; l < nl; ++l) {
      MgLevel& L = h->lev[l];
      L.nx = lx;
      L.ny = ly;
      L.n = (long long)lx * ly;
      L.coef = h->p.helm_coef;
      L.dx2 = std::pow(ldx, 2.0);
      L.dy2 = std::pow(ldy, 2.0);
      L.cx = L.coef * (1.0 / L.dx2);
      L.cy = L.coef * (1.0 / L.dy2);
      int arc = HOC_OK;
      if (l > 0) {
        arc = arc ? arc : h->alloc(&h->mg_Ku[l], L.n);
        arc = arc ? arc : h->alloc(&h->mg_Kv[l], L.n);
        arc = arc ? arc : h->alloc(&h->mg_inv[l], L.n);
        arc = arc ? arc : h->alloc(&h->mg_mask[l], L.n);
      }
      arc = arc ? arc : h->alloc(&h->mg_x[l], L.n);
      arc = arc ? arc : h->alloc(&h->mg_t[l], L.n);
      arc = arc ? arc : h->alloc(&h->mg_b[l], L.n);
      if (cfg->precision >= 1) {
        MgLevelF& F = h->levf[l];
        F.nx = lx;
        F.ny = ly;
        F.n = L.n;
        F.coef = (float)L.coef;
        F.dx2 = (float)L.dx2;
        F.dy2 = (float)L.dy2;
        F.cx = (float)L.cx;
        F.cy = (float)L.cy;
        arc = arc ? arc : h->alloc(&h->mgf_Ku[l], L.n);
        arc = arc ? arc : h->alloc(&h->mgf_Kv[l], L.n);
        arc = arc ? arc : h->alloc(&h->mgf_inv[l], L.n);
        arc = arc ? arc : h->alloc(&h->mgf_x[l], L.n);
        arc = arc ? arc : h->alloc(&h->mgf_t[l], L.n);
        arc = arc ? arc : h->alloc(&h->mgf_b[l], L.n);
      }
      if (arc != HOC_OK) return cleanup(arc);
      lx /= 2;
      ly /= 2;
      ldx = ldx * 2.0;
      ldy = ldy * 2.0;
    }
  }
  if (cfg->solver == SOLVER_PCG_MG && cfg->mg_sweeps <= MG_FUSED_MAX_SWEEPS) {
    int l0 = -1;
    for (int l = 0; l < h->nlev; ++l)
      if (h->lev[l].n <= MG_COARSE_MAX) {
        l0 = l;
        break;
      }
    if (l0 < 0 && h->lev[h->nlev - 1].n <= MG_COARSE_LAST_MAX) l0 = h->nlev - 1;
    if (l0 >= 0 && h->nlev - l0 <= MG_CYCLE_LEVELS) {
      MgCycle C{};
      C.nl = h->nlev - l0;
      C.rb = cfg->mg_smoother == 1 ? 1 : 0;
      for (int l = 0; l < C.nl; ++l) C.lev[l] = h->lev[l0 + l];
      if (mg_cycle_prepare(mg_cycle_layout(C)) == 0) h->mg_cycle_l0 = l0;
      if (cfg->precision >= 1) {
        MgCycleF CF{};
        CF.nl = C.nl;
        for (int l = 0; l < CF.nl; ++l) CF.lev[l] = h->levf[l0 + l];
        if (mg_cycle_prepare_f32(mg_cycle_layout(CF)) != 0) h->mg_cycle_l0 = -1;
      }
    }
    if (cfg->precision >= 1 && h->mg_cycle_l0 < 0)
      return cleanup(fail(HOC_ERR_INVALID, "mixed precision: the coarse end of the multigrid "
                                           "hierarchy does not fit one workgroup"));
  }
  int rc = HOC_OK;
  rc = rc ? rc : upload(h->m2, g->mask, d.n2, "mask");
  rc = rc ? rc : upload(h->mu2, g->mask_u, d.n2, "mask_u");
  rc = rc ? rc : upload(h->mv2, g->mask_v, d.n2, "mask_v");
  rc = rc ? rc : upload(h->m3, g->mask3, d.n3, "mask3");
  rc = rc ? rc : upload(h->m3u, g->mask3u, d.n3, "mask3u");
  rc = rc ? rc : upload(h->m3v, g->mask3v, d.n3, "mask3v");
  rc = rc ? rc : upload(h->dz3, g->dz3, d.n3, "dz3");
  rc = rc ? rc : upload(h->dz3u, g->dz3u, d.n3, "dz3u");
  rc = rc ? rc : upload(h->dz3v, g->dz3v, d.n3, "dz3v");
  rc = rc ? rc : upload(h->zc, g->z_centre, d.n3, "z_centre");
  rc = rc ? rc : upload(h->tau_x, g->tau_x, d.n2, "tau_x");
  rc = rc ? rc : upload(h->tau_y, g->tau_y, d.n2, "tau_y");
  rc = rc ? rc : upload(h->ustar2, g->ustar2, d.n2, "ustar2");
  rc = rc ? rc : upload(h->kappa_b, g->kappa_b, d.n2, "kappa_b");
  if (rc == HOC_OK && (cfg->open_x || cfg->open_y || cfg->vcoord != 0)) {
    if (!g->H)
      return cleanup(fail(HOC_ERR_INVALID,
                          "open boundaries and moving coordinates need geometry H"));
    rc = h->alloc(&h->Hd, d.n2);
    rc = rc ? rc : upload(h->Hd, g->H, d.n2, "H");
  }
  if (rc == HOC_OK && cfg->vcoord != 0) {
    rc = h->alloc(&h->dz3_ref, d.n3);
    rc = rc ? rc : upload(h->dz3_ref, g->dz3, d.n3, "dz3");
  }
  if (rc != HOC_OK) return cleanup(rc);
  if (hipMemset(h->bar, 0, GRID_BAR_WORDS * sizeof(unsigned)) != hipSuccess ||
      hipMemset(h->bar_err, 0, sizeof(int)) != hipSuccess)
    return cleanup(fail(HOC_ERR_DEVICE, "hipMemset failed"));
  if (hipMemset(h->dcount, 0, 4 * sizeof(long long)) != hipSuccess)
    return cleanup(fail(HOC_ERR_DEVICE, "hipMemset failed"));
  k_geometry<<<blocks(d.n2), TPB>>>(d, h->p, h->dz3, h->m3u, h->m3v, h->tau_x, h->tau_y, h->dzi,
                                    h->tau_u, h->tau_v);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return cleanup(fail(HOC_ERR_DEVICE, "geometry kernel failed"));
  Geo& geo = h->geo;
  geo.m2 = h->m2;
  geo.mu2 = h->mu2;
  geo.mv2 = h->mv2;
  geo.m3 = h->m3;
  geo.m3u = h->m3u;
  geo.m3v = h->m3v;
  geo.dz3 = h->dz3;
  geo.dz3u = h->dz3u;
  geo.dz3v = h->dz3v;
  geo.zc = h->zc;
  geo.dzi = h->dzi;
  geo.tau_u = h->tau_u;
  geo.tau_v = h->tau_v;
  geo.ustar2 = h->ustar2;
  geo.kappa_b = h->kappa_b;
  geo.H = h->Hd;
  if (cfg->vcoord == 0) {
    // static z-level geometry: read in column form wherever it reproduces the
    // stored arrays bit for bit
    std::vector<uint8_t> nw;
    std::vector<double> hb, zs;
    double dzr = 0.0;
    if (column_form(d, cfg->face_rule, g, nw, hb, zs, dzr)) {
      int crc = h->alloc(&h->cg_nw, d.n2);
      crc = crc ? crc : h->alloc(&h->cg_hb, d.n2);
      crc = crc ? crc : h->alloc(&h->cg_zs, d.nz + 1);
      crc = crc ? crc : upload(h->cg_nw, nw.data(), d.n2, "nw");
      crc = crc ? crc : upload(h->cg_hb, hb.data(), d.n2, "hb");
      crc = crc ? crc : upload(h->cg_zs, zs.data(), d.nz + 1, "zs");
      if (crc != HOC_OK) return cleanup(crc);
      geo.nw = h->cg_nw;
      geo.hb = h->cg_hb;
      geo.zs = h->cg_zs;
      h->p.zcol = 1;
      h->p.dzr = dzr;
    }
  }
  if (cfg->scheme == 2) {
    k_face_depth<<<blocks(d.n2), TPB>>>(d, geo, h->Hu, h->Hv);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
      return cleanup(fail(HOC_ERR_DEVICE, "face-depth kernel failed"));
  }
  *out = h;
  return HOC_OK;
}

int hoc_stepper_destroy(hoc_stepper* h) {
  if (h && release_is_safe()) delete h;
  return HOC_OK;
}

int hoc_stepper_bind_mixing(hoc_stepper* h, double* K_m, double* K_h) {
  if (!h || !K_m || !K_h) return fail(HOC_ERR_INVALID, "null argument");
  h->Km = K_m;
  h->Kh = K_h;
  return HOC_OK;
}

int hoc_stepper_build_coefficients(hoc_stepper* h, void* stream) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  return build_coefficients(h, (hipStream_t)stream);
}

int hoc_stepper_tke_step(hoc_stepper* h, const double* e_in, const double* u, const double* v,
                         const double* b, double* e_out, void* stream) {
  if (!h || !e_in || !u || !v || !b || !e_out) return fail(HOC_ERR_INVALID, "null argument");
  if (!h->Km || !h->Kh) return fail(HOC_ERR_INVALID, "bind K_m / K_h first");
  hipStream_t s = (hipStream_t)stream;
  const Dims d = h->d;
  if (h->p.mxl_recursive) {
    k_closure_shear<<<blocks(d.ni), TPB, 0, s>>>(d, h->p, h->geo, u, v, b, h->s_n2, h->s_sh);
    k_mixing_recursive<<<blocks(d.n2), TPB, 0, s>>>(d, h->p, h->geo, e_in, h->s_n2, h->s_l,
                                                     h->s_le);
  } else {
    ProfScope _ps(h, "k_closure_lengths", s);
    launch_closure_lengths(d, h->p, h->geo, e_in, u, v, b, h->s_n2, h->s_sh, h->s_l, h->s_le, s);
  }
  { ProfScope _ps(h, "k_closure", s);
  k_closure<<<blocks(d.n2), TPB, 0, s>>>(d, h->p, h->geo, e_in, h->Km, h->Kh, e_out, h->s_n2,
                                         h->s_sh, h->s_l, h->s_le, h->s_c); }
  LAUNCH_CHECK("k_closure");
  return HOC_OK;
}

// The barotropic part of split_explicit (model3d_v05.py:312-335,441-462)
// after the baroclinic Picard passes left gh in (gh_u, gh_v) and their depth
// sums in (Ug, Vg): forcing, n_split substeps, depth-mean replacement.
int split_barotropic(hoc_stepper* h, const double* eta, const double* gh_u, const double* gh_v,
                     hoc_state* out, hipStream_t s) {
  const Dims d = h->d;
  const StepParams& p = h->p;
  const Geo& geo = h->geo;
  const unsigned nb = blocks(d.n2);
  k_split_forcing<<<nb, TPB, 0, s>>>(d, p, geo, h->U0, h->V0, h->Ug, h->Vg, h->bfx, h->bfy);
  if (p.bt_mean) {
    HIP_TRY(hipMemsetAsync(h->acc_u, 0, d.n2 * sizeof(double), s));
    HIP_TRY(hipMemsetAsync(h->acc_v, 0, d.n2 * sizeof(double), s));
  }
  // transports rotate over three buffers: the substep's start and two
  // Picard iterates; the levels over two, the last substep writing out->eta
  const double* U = h->U0;
  const double* V = h->V0;
  const double* e = eta;
  int cur = -1;  // buffer index holding U, V (-1: U0, V0)
  for (int n = 0; n < p.n_split; ++n) {
    const double* Ui = U;
    const double* Vi = V;
    int slot = -1;
    for (int m = 0; m < p.n_picard; ++m) {
      int q = 0;
      while (q == cur || q == slot) ++q;
      k_bt_picard<<<nb, TPB, 0, s>>>(d, p, geo, U, V, e, Ui, Vi, h->bfx, h->bfy, h->Ku, h->Kv,
                                     h->bU[q], h->bV[q]);
      Ui = h->bU[q];
      Vi = h->bV[q];
      slot = q;
    }
    U = Ui;
    V = Vi;
    cur = slot;
    double* e_new = (n == p.n_split - 1) ? out->eta : h->be[n & 1];
    k_bt_eta<<<nb, TPB, 0, s>>>(d, p, geo, e, U, V, e_new, p.bt_mean ? h->acc_u : nullptr,
                                h->acc_v);
    e = e_new;
  }
  LAUNCH_CHECK("barotropic substeps");
  h->barotropic_substeps += p.n_split;
  if (p.bt_mean) {
    int q = 0;
    while (q == cur) ++q;
    k_bt_mean<<<nb, TPB, 0, s>>>(d, p, h->acc_u, h->acc_v, h->bU[q], h->bV[q]);
    U = h->bU[q];
    V = h->bV[q];
  }
  k_split_final<<<blocks(d.n3), TPB, 0, s>>>(d, geo, gh_u, gh_v, h->Ug, h->Vg, U, V, h->Hu,
                                             h->Hv, out->u, out->v);
  LAUNCH_CHECK("split depth-mean replacement");
  return HOC_OK;
}

// One step; w_ready: h->w already holds w of in->u, in->v (the previous step's
// output on a static geometry), so the step does not form it again.
static int step_impl(hoc_stepper* h, const hoc_state* in, hoc_state* out, void* stream,
                     bool w_ready, const hoc_step_tape* tape = nullptr) {
  if (!h || !in || !out) return fail(HOC_ERR_INVALID, "null argument");
  if (tape) {
    if (h->cfg.scheme != 0 || h->cfg.vcoord != 0)
      return fail(HOC_ERR_UNSUPPORTED, "the step tape covers the theta scheme on z-level grids");
    if (h->cfg.n_picard > HOC_TAPE_MAX_PICARD)
      return fail(HOC_ERR_UNSUPPORTED, "the step tape holds at most HOC_TAPE_MAX_PICARD passes");
    if (!tape->q || !tape->Ku || !tape->Kv || !tape->w_old || !tape->w)
      return fail(HOC_ERR_INVALID, "tape arrays q, Ku, Kv, w_old, w are required");
    for (int it = 0; it < h->cfg.n_picard; ++it)
      if (!tape->gh_u[it] || !tape->gh_v[it] || !tape->eta[it])
        return fail(HOC_ERR_INVALID, "tape arrays gh_u, gh_v, eta are required for every pass");
  }
  if (!in->u || !in->v || !in->b || !in->eta || !out->u || !out->v || !out->b || !out->eta)
    return fail(HOC_ERR_INVALID, "state arrays u, v, b, eta are required");
  const bool ts = h->cfg.tracers_ts != 0;
  if (ts && (!in->T || !in->S || !out->T || !out->S))
    return fail(HOC_ERR_INVALID, "tracers_ts needs T and S");
  const bool tke = h->cfg.closure_tke != 0;
  if (tke && (!in->e || !out->e)) return fail(HOC_ERR_INVALID, "closure_tke needs e");
  // the step reads the old state after it has written parts of the new one:
  // no output array may be any input array, and the outputs are distinct
  {
    const double* ins[7] = {in->u, in->v, in->b, in->eta, ts ? in->T : nullptr,
                            ts ? in->S : nullptr, tke ? in->e : nullptr};
    const double* outs[7] = {out->u, out->v, out->b, out->eta, ts ? out->T : nullptr,
                             ts ? out->S : nullptr, tke ? out->e : nullptr};
    for (int a = 0; a < 7; ++a) {
      if (!outs[a]) continue;
      for (int q = 0; q < 7; ++q)
        if (outs[a] == ins[q])
          return fail(HOC_ERR_INVALID, "the output state must not alias the input state");
      for (int q = a + 1; q < 7; ++q)
        if (outs[a] == outs[q])
          return fail(HOC_ERR_INVALID, "the output state's arrays must be distinct");
    }
  }
  if (!h->coeffs_built) TRY(build_coefficients(h, (hipStream_t)stream));
  hipStream_t s = (hipStream_t)stream;
  const Dims d = h->d;
  const StepParams& p = h->p;
  const Geo& geo = h->geo;
  const bool theta = h->cfg.scheme == 0;

  if (h->cfg.vcoord != 0) {
    // z* / sigma: thickness, faces, centre depths and coefficients follow
    // the free surface (model3d_v05.py:346-347,128-139); the rebuilt solver
    // is a new one, not counted as a rebuild
    k_moving_dz<<<blocks(d.n2), TPB, 0, s>>>(d, h->m2, h->Hd, in->eta, h->dz3_ref, h->dz3,
                                             h->zc, h->dzi);
    k_moving_faces<<<blocks(d.n3), TPB, 0, s>>>(d, p, h->m3u, h->m3v, h->dz3, h->dz3u,
                                                h->dz3v);
    if (h->cfg.scheme == 2) k_face_depth<<<blocks(d.n2), TPB, 0, s>>>(d, geo, h->Hu, h->Hv);
    LAUNCH_CHECK("moving geometry");
    h->coeffs_built = false;
    TRY(build_coefficients(h, s));
  }
  if (tke) {
    TRY(hoc_stepper_tke_step(h, in->e, in->u, in->v, in->b, out->e, stream));
    h->tridiagonal_solves += 1;
    TRY(build_coefficients(h, s));
  } else if (in->e && out->e && in->e != out->e) {
    HIP_TRY(hipMemcpyAsync(out->e, in->e, d.ni * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  const bool split = h->cfg.scheme == 2;
  // phi, the transports of the state (theta, split_explicit) and, unless the
  // previous step formed it, w of the state: one column pass
  { ProfScope _ps(h, "k_prologue", s);
  k_prologue<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, in->u, in->v, in->b, h->phi,
                                          (theta || split) ? h->U0 : nullptr, h->V0,
                                          w_ready ? nullptr : h->w); }
  if (tape)
    HIP_TRY(hipMemcpyAsync(tape->w_old, h->w, d.ni * sizeof(double), hipMemcpyDeviceToDevice, s));
  { ProfScope _ps(h, "k_momentum_explicit", s);
  launch_momentum_explicit_tiled((h->fp32_parts & HOC_FP32_MOMENTUM_FLUX) != 0, s, d, p, geo,
                                 in->u, in->v, in->b, h->phi,
                                 h->w, h->Km, h->pgx, h->pgy, nullptr, nullptr, h->dzu, h->dzv,
                                 h->stu, h->stv); }
  LAUNCH_CHECK("explicit momentum");
  if (theta) k_div2<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, h->U0, h->V0, in->eta, h->div0);
  const double* u_it = in->u;
  const double* v_it = in->v;
  for (int it = 0; it < h->cfg.n_picard; ++it) {
    double* gu = tape ? tape->gh_u[it] : h->gh_u[it & 1];
    double* gv = tape ? tape->gh_v[it] : h->gh_v[it & 1];
    { ProfScope _ps(h, "k_predictor", s);
    if (p.zcol)
      k_predictor_col<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, in->u, in->v, in->eta, u_it, v_it,
                                                   h->pgx, h->pgy, h->dzu, h->dzv, h->stu,
                                                   h->stv, h->sub_m, h->den_m, h->cst_m, gu, gv,
                                                   h->Ug, h->Vg);
    else
      k_predictor<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, in->u, in->v, in->eta, u_it, v_it,
                                               h->pgx, h->pgy, h->dzu, h->dzv, h->stu, h->stv,
                                               h->sub_m, h->den_m, h->cst_m, gu, gv, h->Ug,
                                               h->Vg); }
    LAUNCH_CHECK("k_predictor");
    h->tridiagonal_solves += 2;
    if (theta) {
      k_rhs_eta<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, in->eta, h->div0, h->Ug, h->Vg, h->rhs);
      LAUNCH_CHECK("k_rhs_eta");
      int iters = 0, conv = 0;
      // a graph-replayed sync-free solve writes a fixed buffer (the graph's
      // operands stay the same from step to step), copied to the state
      double* xs = out->eta;
      bool replay = false;
      if (h->sync_free_budget > 0) TRY(sf_graph_usable(h, s, &replay));
      if (replay) {
        if (!h->sf_x) TRY(h->alloc(&h->sf_x, d.n2));
        xs = h->sf_x;
      }
      {  // the solve's span on the device timeline, host polls included
        ProfScope _ps(h, "helmholtz_solve", s);
        TRY(helmholtz_solve(h, h->Ku, h->Kv, h->inv_diag, h->rhs, xs, s, &iters, &conv));
      }
      if (xs != out->eta)
        HIP_TRY(hipMemcpyAsync(out->eta, xs, d.n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
      if (tape)
        HIP_TRY(hipMemcpyAsync(tape->eta[it], out->eta, d.n2 * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
      if (it + 1 == h->cfg.n_picard) {
        // the last pass: back-substitution and w of the new velocity in one
        // column pass (the tracers read w next)
        ProfScope _ps(h, "k_backsub_w", s);
        k_backsub_w<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, gu, gv, h->q, out->eta, out->u,
                                                 out->v, h->w);
      } else {
        ProfScope _ps(h, "k_backsub", s);
        k_backsub<<<blocks(d.n3), TPB, 0, s>>>(d, p, geo, gu, gv, h->q, out->eta, out->u,
                                               out->v);
      }
      LAUNCH_CHECK("k_backsub");
      u_it = out->u;
      v_it = out->v;
    } else {
      u_it = gu;
      v_it = gv;
    }
  }
  if (split) {
    TRY(split_barotropic(h, in->eta, u_it, v_it, out, s));
  } else if (!theta) {
    HIP_TRY(hipMemcpyAsync(out->u, u_it, d.n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(out->v, v_it, d.n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    k_depth_sum<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, out->u, out->v, h->U0, h->V0);
    k_eta_fb<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, in->eta, h->U0, h->V0, out->eta);
    LAUNCH_CHECK("k_eta_fb");
  }
  if (!theta) {  // the theta scheme's last back-substitution formed w already
    ProfScope _ps(h, "k_w_diag", s);
    k_w_diag<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, out->u, out->v, h->w);
  }
  { ProfScope _ps(h, "k_tracers", s);
  launch_tracers_tiled((h->fp32_parts & HOC_FP32_TRACER_FLUX) != 0, s, d, p, geo, in->T, in->S,
                       in->b, out->u, out->v,
                       h->w, h->Kh, h->cst_b, out->T, out->S, out->b); }
  LAUNCH_CHECK("k_tracers");
  h->tridiagonal_solves += ts ? 2 : 1;
  if (ts) h->eos_cells += d.n3;
  if (tape) {
    HIP_TRY(hipMemcpyAsync(tape->q, h->q, d.n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(tape->Ku, h->Ku, d.n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(tape->Kv, h->Kv, d.n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(tape->w, h->w, d.ni * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  return HOC_OK;
}

int hoc_stepper_step(hoc_stepper* h, const hoc_state* in, hoc_state* out, void* stream) {
  return step_impl(h, in, out, stream, false);
}

int hoc_stepper_step_tape(hoc_stepper* h, const hoc_state* in, hoc_state* out,
                          const hoc_step_tape* tape, void* stream) {
  if (!tape) return fail(HOC_ERR_INVALID, "null tape");
  return step_impl(h, in, out, stream, false, tape);
}

int hoc_stepper_step_adjoint(hoc_stepper* h, const hoc_state* x, const hoc_state* y,
                             const double* Km, const double* Kh, const hoc_step_tape* tape,
                             const hoc_state* gy, const double* gy_Km, const double* gy_Kh,
                             hoc_state* gx, double* g_Km, double* g_Kh, void* stream) {
  if (!h || !x || !y || !Km || !Kh || !tape || !gy || !gx || !g_Km || !g_Kh)
    return fail(HOC_ERR_INVALID, "null argument");
  const hoc_config& cfg = h->cfg;
  const StepParams& p = h->p;
  if (cfg.scheme != 0 || cfg.vcoord != 0)
    return fail(HOC_ERR_UNSUPPORTED, "the reverse step covers the theta scheme on z-level grids");
  if (p.open_x || p.open_y)
    return fail(HOC_ERR_UNSUPPORTED, "the reverse step covers closed or periodic edges");
  if (p.advection != ADV_UP3 && p.advection != ADV_UP3_TVD)
    return fail(HOC_ERR_UNSUPPORTED, "the reverse step covers advection up3 and up3_tvd");
  if (!cfg.tracers_ts) return fail(HOC_ERR_UNSUPPORTED, "the reverse step needs tracers T and S");
  if (h->fp32_parts) return fail(HOC_ERR_UNSUPPORTED, "the reverse step is the fp64 step's");
  if (h->d.nx < 4 || h->d.ny < 4)
    return fail(HOC_ERR_UNSUPPORTED, "the reverse step needs nx, ny >= 4");
  if (cfg.n_picard < 1 || cfg.n_picard > HOC_TAPE_MAX_PICARD)
    return fail(HOC_ERR_UNSUPPORTED, "the reverse step covers 1..HOC_TAPE_MAX_PICARD passes");
  if (!x->u || !x->v || !x->T || !x->S || !y->u || !y->v || !y->T || !y->S)
    return fail(HOC_ERR_INVALID, "x needs u, v, T, S and y needs u, v, T, S");
  if (!gx->u || !gx->v || !gx->b || !gx->eta || !gx->T || !gx->S)
    return fail(HOC_ERR_INVALID, "gx needs u, v, b, eta, T, S");
  if (!tape->q || !tape->Ku || !tape->Kv || !tape->w_old || !tape->w)
    return fail(HOC_ERR_INVALID, "tape arrays q, Ku, Kv, w_old, w are required");
  for (int it = 0; it < cfg.n_picard; ++it)
    if (!tape->gh_u[it] || !tape->gh_v[it] || !tape->eta[it])
      return fail(HOC_ERR_INVALID, "tape arrays gh_u, gh_v, eta are required for every pass");
  const Dims d = h->d;
  const Geo& geo = h->geo;
  hipStream_t s = (hipStream_t)stream;
  if (!h->rv_ready) {
    for (double** b : {&h->rv_lamT, &h->rv_lamS, &h->rv_cs, &h->rv_gui, &h->rv_gvi, &h->rv_Lu,
                       &h->rv_Lv, &h->rv_lu, &h->rv_lv, &h->rv_gq})
      TRY(h->alloc(b, d.n3));
    TRY(h->alloc(&h->rv_gw, d.ni));
    TRY(h->alloc(&h->rv_gw0, d.ni));
    for (double** b : {&h->rv_X, &h->rv_Y, &h->rv_ge, &h->rv_leta, &h->rv_GU, &h->rv_GV,
                       &h->rv_gdiv, &h->rv_inv, &h->rv_gKu, &h->rv_gKv, &h->rv_Ku, &h->rv_Kv})
      TRY(h->alloc(b, d.n2));
    h->rv_ready = true;
  }
  auto init = [&](double* dst, const double* src, long long n) -> int {
    if (src)
      HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    else
      k_fill<<<blocks(n), TPB, 0, s>>>(n, 0.0, dst);
    return HOC_OK;
  };
  TRY(init(g_Km, gy_Km, d.ni));
  TRY(init(g_Kh, gy_Kh, d.ni));
  // the other accumulators (gq, Lu, Lv, gx->u, gx->v, gx->eta, gdiv, gKu,
  // gKv) are written by their first writer in the last Picard pass's reverse
  // tracers and the new velocity's w (model3d_v05.py:466-533)
  k_rev_tracer_rhs<<<blocks(d.n3), TPB, 0, s>>>(d, p, geo, y->T, y->S, gy->T, gy->S, gy->b,
                                                h->rv_lamT, h->rv_lamS);
  k_rev_tracers<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, x->T, x->S, y->T, y->S, Kh, h->rv_lamT,
                                             h->rv_lamS, h->rv_cs, gx->T, gx->S, g_Kh);
  k_rev_tracer_stencil<<<dim3((d.nx + RTS_TX - 1) / RTS_TX, (d.ny + RTS_TY - 1) / RTS_TY, d.nz),
                         RTS_TX * RTS_TY, 0, s>>>(d, p, geo, x->T, x->S, y->u, y->v, tape->w,
                                                    h->rv_lamT, h->rv_lamS, gy->u, gy->v, gx->T,
                                                    gx->S, h->rv_gui, h->rv_gvi, h->rv_gw);
  LAUNCH_CHECK("reverse tracers");
  // Picard passes in reverse (model3d_v05.py:387-408)
  HIP_TRY(hipMemcpyAsync(h->rv_Ku, tape->Ku, d.n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(h->rv_Kv, tape->Kv, d.n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
  HelmOp op{h->rv_Ku, h->rv_Kv, h->m2, p.helm_coef, p.helm_cx, p.helm_cy, p.dx2, p.dy2};
  k_helm_diag<<<blocks(d.n2), TPB, 0, s>>>(d, op, h->rv_inv);
  const int last = cfg.n_picard - 1;
  for (int it = last; it >= 0; --it) {
    const int first = it == last ? 1 : 0;
    k_rev_backsub<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, it == last ? h->rv_gw : nullptr,
                                               tape->q, tape->eta[it], first, h->rv_gui,
                                               h->rv_gvi, h->rv_gq, h->rv_X, h->rv_Y);
    k_rev_eta_rhs<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, it == last ? gy->eta : nullptr,
                                               h->rv_X, h->rv_Y, h->rv_ge);
    LAUNCH_CHECK("reverse back-substitution");
    int iters = 0, conv = 0;
    TRY(helmholtz_solve(h, h->rv_Ku, h->rv_Kv, h->rv_inv, h->rv_ge, h->rv_leta, s, &iters, &conv));
    k_rev_eta_post<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, tape->eta[it], h->rv_leta, first,
                                                h->rv_gKu, h->rv_gKv, gx->eta, h->rv_gdiv,
                                                h->rv_GU, h->rv_GV);
    k_rev_predictor<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, Km, h->rv_gui, h->rv_gvi, h->rv_GU,
                                                 h->rv_GV, tape->gh_u[it], tape->gh_v[it], first,
                                                 h->rv_lu, h->rv_lv, h->rv_cs, h->rv_Lu,
                                                 h->rv_Lv, g_Km);
    k_rev_coriolis<<<blocks(d.n3), TPB, 0, s>>>(d, p, geo, h->rv_lu, h->rv_lv, h->rv_Lu,
                                                h->rv_Lv, it == 0 ? 1 : 0,
                                                it == 0 ? gx->u : h->rv_gui,
                                                it == 0 ? gx->v : h->rv_gvi);
    LAUNCH_CHECK("reverse predictor");
  }
  // the explicit terms of the old state, its w and q (model3d_v05.py:150-167,349-386)
  k_rev_explicit_cells<<<blocks(d.n3), TPB, 0, s>>>(d, p, geo, h->rv_Lu, h->rv_Lv, h->rv_gdiv,
                                                    gx->u, gx->v);
  k_rev_explicit<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, x->u, x->v, Km, h->rv_Lu, h->rv_Lv,
                                              gx->u, gx->v, gx->b, gx->eta, g_Km);
  k_rev_momentum_adv<<<dim3((d.nx + RMA_TX - 1) / RMA_TX, (d.ny + RMA_TY - 1) / RMA_TY, d.nz),
                       RMA_TX * RMA_TY, 0, s>>>(d, p, geo, x->u, x->v, tape->w_old, h->rv_Lu,
                                                h->rv_Lv, gx->u, gx->v, h->rv_gw0);
  k_rev_wold_q<<<blocks(d.n2), TPB, 0, s>>>(d, p, geo, h->rv_gw0, Km, tape->q, h->rv_gKu,
                                            h->rv_gKv, h->rv_gq, h->rv_lu, h->rv_cs, gx->u, gx->v,
                                            g_Km);
  LAUNCH_CHECK("reverse explicit terms");
  return HOC_OK;
}

int hoc_stepper_closure_adjoint(hoc_stepper* h, const double* e, const double* u,
                                const double* v, const double* b, const double* Km,
                                const double* Kh, const double* kappa_b, const double* g_e_new,
                                const double* g_Km_new, const double* g_Kh_new, double* g_e,
                                double* g_u, double* g_v, double* g_b, double* g_Km,
                                double* g_Kh, double* g_kappa_b, void* stream) {
  if (!h || !e || !u || !v || !b || !Km || !Kh || !kappa_b || !g_Km_new || !g_Kh_new || !g_e ||
      !g_u || !g_v || !g_b || !g_Km || !g_Kh || !g_kappa_b)
    return fail(HOC_ERR_INVALID, "null argument");
  if (!h->cfg.closure_tke || h->p.mxl_recursive)
    return fail(HOC_ERR_UNSUPPORTED, "the closure reverse covers the TKE closure with integral lengths");
  if (h->cfg.vcoord != 0) return fail(HOC_ERR_UNSUPPORTED, "the closure reverse covers z-level grids");
  const Dims d = h->d;
  hipStream_t s = (hipStream_t)stream;
  if (!h->cr_ready) {
    for (double** q : {&h->cr_n2, &h->cr_lu, &h->cr_ld, &h->cr_x, &h->cr_cs, &h->cr_lam,
                       &h->cr_gl, &h->cr_n2b, &h->cr_zs, &h->cr_du, &h->cr_dv})
      TRY(h->alloc(q, d.ni));
    TRY(h->alloc(&h->cr_row, 4 * d.ni));
    TRY(h->alloc(&h->cr_hi, d.n2));
    TRY(h->alloc(&h->cr_guc, d.n3));
    TRY(h->alloc(&h->cr_gvc, d.n3));
    h->cr_ready = true;
  }
  k_rev_closure_col0<<<blocks(d.n2), TPB, 0, s>>>(d, h->geo, b, h->cr_n2, h->cr_n2b, h->cr_zs,
                                                  h->cr_hi);
  k_rev_closure_rows<<<blocks(d.ni), TPB, 0, s>>>(d, h->p, h->geo, e, u, v, Km, Kh, kappa_b,
                                                  h->cr_n2, h->cr_zs, h->cr_hi, h->cr_du,
                                                  h->cr_dv, h->cr_lu, h->cr_ld, h->cr_row);
  k_rev_closure<<<blocks(d.n2), TPB, 0, s>>>(d, h->p, h->geo, e, Km, Kh, kappa_b, g_e_new,
                                             g_Km_new, g_Kh_new, h->cr_n2, h->cr_zs, h->cr_hi,
                                             h->cr_du, h->cr_dv, h->cr_lu, h->cr_ld, h->cr_row,
                                             h->cr_x, h->cr_cs, h->cr_lam, h->cr_gl, h->cr_n2b,
                                             g_e, h->cr_guc, h->cr_gvc, g_b, g_Km, g_Kh,
                                             g_kappa_b);
  k_rev_closure_uv<<<blocks(d.n3), TPB, 0, s>>>(d, h->cr_guc, h->cr_gvc, g_u, g_v);
  LAUNCH_CHECK("k_rev_closure");
  return HOC_OK;
}

int hoc_stepper_z_centre(hoc_stepper* h, double* out, void* stream) {
  if (!h || !out) return fail(HOC_ERR_INVALID, "null argument");
  HIP_TRY(hipMemcpyAsync(out, h->zc, h->d.n3 * sizeof(double), hipMemcpyDeviceToDevice,
                         (hipStream_t)stream));
  return HOC_OK;
}

int hoc_stepper_counters(hoc_stepper* h, hoc_counters* out) {
  if (!h || !out) return fail(HOC_ERR_INVALID, "null argument");
  long long dc[4];
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(dc, h->dcount, sizeof(dc), hipMemcpyDeviceToHost));
  out->solver_iterations = dc[0];
  out->solver_failures = dc[1];
  out->solver_rebuilds = h->solver_rebuilds;
  out->tridiagonal_solves = h->tridiagonal_solves;
  out->barotropic_substeps = h->barotropic_substeps;
  out->eos_cells = h->eos_cells;
  return HOC_OK;
}

int hoc_stepper_set_sync_free(hoc_stepper* h, int32_t budget) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  if (budget < 0) return fail(HOC_ERR_INVALID, "budget must be >= 0");
  if (budget > 0 && (h->cfg.scheme != 0 || (h->cfg.solver != SOLVER_PCG_MG &&
                                            h->cfg.solver != SOLVER_PCG_RBGS)))
    return fail(HOC_ERR_UNSUPPORTED, "sync-free stepping needs the theta scheme with pcg_mg "
                                     "or pcg_rbgs");
  h->sync_free_budget = budget;
  return HOC_OK;
}

int hoc_stepper_set_mg_tile(hoc_stepper* h, int32_t shape) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  if (shape < 0 || shape >= MG_TILE_SHAPES)
    return fail(HOC_ERR_INVALID, "mg tile shape must be 0 (32x16), 1 (64x8), 2 (16x16) or "
                                 "3 (64x16)");
  h->mg_tile_shape = shape;
  return HOC_OK;
}

int hoc_stepper_geometry_form(hoc_stepper* h, int32_t column, int32_t* active) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  if (column == 1 && !h->cg_nw)
    return fail(HOC_ERR_UNSUPPORTED, "no column form: moving vertical coordinate, or the stored "
                                     "geometry is not a z-level column form");
  if (column == 0 || column == 1) h->p.zcol = column;
  else if (column >= 0) return fail(HOC_ERR_INVALID, "column must be -1 (query), 0 or 1");
  if (active) *active = h->p.zcol;
  return HOC_OK;
}

int hoc_stepper_set_fp32_parts(hoc_stepper* h, int32_t parts) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  if (h->cfg.precision == 0)
    return fail(HOC_ERR_INVALID, "fp32 parts need a handle created with precision 1 or 2");
  if (parts < 0 || parts > HOC_FP32_ALL) return fail(HOC_ERR_INVALID, "unknown fp32 part bits");
  h->fp32_parts = parts;
  h->p.eos_mixed = (parts & HOC_FP32_EOS) != 0;
  return HOC_OK;
}

int hoc_stepper_set_solver_path(hoc_stepper* h, int32_t path) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  if (path != 0 && path != 1) return fail(HOC_ERR_INVALID, "path must be 0 or 1");
  h->force_multikernel = path == 1;
  return HOC_OK;
}

int32_t hoc_stepper_solver_path(hoc_stepper* h) {
  if (!h) return -1;
  return (h->pcg_parts > 0 && !h->force_multikernel) ? 0 : 1;
}

int hoc_stepper_profile_kernel(hoc_stepper* h, const char* name) {
  if (!h) return fail(HOC_ERR_INVALID, "null handle");
  TRY(prof_drain(h));
  h->prof_name = name ? name : "";
  h->prof_launches = 0;
  h->prof_ms = 0.0;
  if (!h->prof_name.empty() && h->prof_ev.empty()) {
    h->prof_ev.resize(2 * 512);
    for (auto& e : h->prof_ev) HIP_TRY(hipEventCreate(&e));
  }
  return HOC_OK;
}

int hoc_stepper_profile_read(hoc_stepper* h, int64_t* launches, double* total_ms) {
  if (!h || !launches || !total_ms) return fail(HOC_ERR_INVALID, "null argument");
  TRY(prof_drain(h));
  *launches = h->prof_launches;
  *total_ms = h->prof_ms;
  h->prof_launches = 0;
  h->prof_ms = 0.0;
  return HOC_OK;
}

int hoc_stepper_momentum_advection(hoc_stepper* h, const double* u, const double* v,
                                   const double* w, double* au, double* av, void* stream) {
  if (!h || !u || !v || !w || !au || !av) return fail(HOC_ERR_INVALID, "null argument");
  const Dims d = h->d;
  StepParams p = h->p;
  if (p.advection == ADV_UPWIND1) p.advection = ADV_CENTRED2;
  if (p.advection == ADV_NONE) return fail(HOC_ERR_INVALID, "advection is 'none'");
  k_momentum_advection<<<blocks(d.n3), TPB, 0, (hipStream_t)stream>>>(d, p, h->geo, u, v, w, au,
                                                                      av);
  LAUNCH_CHECK("k_momentum_advection");
  return HOC_OK;
}

int hoc_stepper_tracer_advection(hoc_stepper* h, const double* c, const double* u,
                                 const double* v, const double* w, double* out, void* stream) {
  if (!h || !c || !u || !v || !w || !out) return fail(HOC_ERR_INVALID, "null argument");
  const Dims d = h->d;
  k_tracer_advection<<<blocks(d.n3), TPB, 0, (hipStream_t)stream>>>(d, h->p, h->geo, c, u, v, w,
                                                                    out);
  LAUNCH_CHECK("k_tracer_advection");
  return HOC_OK;
}

int hoc_stepper_pressure_gradients(hoc_stepper* h, const double* b, double* pgx, double* pgy,
                                   void* stream) {
  if (!h || !b || !pgx || !pgy) return fail(HOC_ERR_INVALID, "null argument");
  if (!h->Km) return fail(HOC_ERR_INVALID, "bind K_m / K_h first");
  hipStream_t s = (hipStream_t)stream;
  const Dims d = h->d;
  // The explicit-momentum kernel computes the gradient along with the other
  // start-of-step terms; u = v = 0 and w = 0 make those terms harmless here.
  HIP_TRY(hipMemsetAsync(h->gh_u[0], 0, d.n3 * sizeof(double), s));
  HIP_TRY(hipMemsetAsync(h->s_n2, 0, d.ni * sizeof(double), s));
  k_phi<<<blocks(d.n2), TPB, 0, s>>>(d, h->p, h->geo, b, h->phi);
  k_momentum_explicit<<<blocks(d.n3), TPB, 0, s>>>(d, h->p, h->geo, h->gh_u[0], h->gh_u[0], b,
                                                   h->phi, h->s_n2, h->Km, pgx, pgy, h->avu,
                                                   h->avv, h->dzu, h->dzv, h->stu, h->stv);
  LAUNCH_CHECK("pressure gradients");
  return HOC_OK;
}

int hoc_stepper_w_diag(hoc_stepper* h, const double* u, const double* v, double* w,
                       void* stream) {
  if (!h || !u || !v || !w) return fail(HOC_ERR_INVALID, "null argument");
  const Dims d = h->d;
  k_w_diag<<<blocks(d.n2), TPB, 0, (hipStream_t)stream>>>(d, h->p, h->geo, u, v, w);
  LAUNCH_CHECK("k_w_diag");
  return HOC_OK;
}

int hoc_stepper_helmholtz_solve(hoc_stepper* h, const double* Ku, const double* Kv,
                                const double* rhs, double* x, int32_t* iterations,
                                int32_t* converged, void* stream) {
  if (!h || !Ku || !Kv || !rhs || !x) return fail(HOC_ERR_INVALID, "null argument");
  if (h->cfg.scheme != 0) return fail(HOC_ERR_INVALID, "the fb scheme has no Helmholtz solve");
  hipStream_t s = (hipStream_t)stream;
  const Dims d = h->d;
  HelmOp op{Ku, Kv, h->m2, h->p.helm_coef, h->p.helm_cx, h->p.helm_cy, h->p.dx2, h->p.dy2};
  double* inv = h->gh_u[1];  // n3-sized work array, unused by the solve itself
  k_helm_diag<<<blocks(d.n2), TPB, 0, s>>>(d, op, inv);
  LAUNCH_CHECK("k_helm_diag");
  int it = 0, cv = 0;
  TRY(helmholtz_solve(h, Ku, Kv, inv, rhs, x, s, &it, &cv));
  if (iterations) *iterations = it;
  if (converged) *converged = cv;
  return HOC_OK;
}

int hoc_thomas_batched(const double* sub, const double* diag, const double* sup,
                       const double* rhs, double* x, double* scratch, int64_t n, int64_t ncol,
                       void* stream) {
  if (!sub || !diag || !sup || !rhs || !x || !scratch) return fail(HOC_ERR_INVALID, "null argument");
  if (n < 1 || ncol < 1) return fail(HOC_ERR_INVALID, "n and ncol must be positive");
  k_thomas<<<blocks(ncol), TPB, 0, (hipStream_t)stream>>>(ncol, (int)n, sub, diag, sup, rhs, x,
                                                          scratch);
  LAUNCH_CHECK("k_thomas");
  return HOC_OK;
}

int hoc_thomas_batched_noalloc(const double* sub, const double* diag, const double* sup,
                               const double* rhs, double* x, int64_t n, int64_t ncol,
                               void* stream) {
  if (!sub || !diag || !sup || !rhs || !x) return fail(HOC_ERR_INVALID, "null argument");
  if (n < 1 || ncol < 1) return fail(HOC_ERR_INVALID, "n and ncol must be positive");
  if (n > HOC_NOALLOC_SCRATCH_DOUBLES / ncol)
    return fail(HOC_ERR_INVALID, "hoc_thomas_batched_noalloc: n * ncol = " +
                                     std::to_string((long long)n * (long long)ncol) +
                                     " exceeds the fixed scratch of " +
                                     std::to_string((long long)HOC_NOALLOC_SCRATCH_DOUBLES) +
                                     " doubles; use hoc_thomas_batched with a workspace");
  double* cs = nullptr;
  TRY(noalloc_scratch(&cs));
  // k_thomas writes every c* it later reads, so the scratch needs no clearing
  k_thomas<<<blocks(ncol), TPB, 0, (hipStream_t)stream>>>(ncol, (int)n, sub, diag, sup, rhs, x,
                                                          cs);
  LAUNCH_CHECK("k_thomas (no-alloc)");
  return HOC_OK;
}

int hoc_thomas_adjoint(const double* sub, const double* diag, const double* sup, const double* x,
                       const double* g, double* lambda, double* d_sub, double* d_diag,
                       double* d_sup, double* scratch, int64_t n, int64_t ncol, void* stream) {
  if (!sub || !diag || !sup || !x || !g || !lambda || !d_sub || !d_diag || !d_sup || !scratch)
    return fail(HOC_ERR_INVALID, "null argument");
  if (n < 1 || ncol < 1) return fail(HOC_ERR_INVALID, "n and ncol must be positive");
  k_thomas_adjoint<<<blocks(ncol), TPB, 0, (hipStream_t)stream>>>(
      ncol, (int)n, sub, diag, sup, x, g, lambda, scratch, d_sub, d_diag, d_sup);
  LAUNCH_CHECK("k_thomas_adjoint");
  return HOC_OK;
}

int hoc_stepper_tracer_advection_adjoint(hoc_stepper* h, const double* c, const double* u,
                                         const double* v, const double* w, const double* g_adv,
                                         double* g_c, double* g_u, double* g_v, double* g_w,
                                         void* stream) {
  if (!h || !c || !u || !v || !w || !g_adv || !g_c || !g_u || !g_v || !g_w)
    return fail(HOC_ERR_INVALID, "null argument");
  if (h->p.advection != ADV_UP3 && h->p.advection != ADV_UP3_TVD)
    return fail(HOC_ERR_UNSUPPORTED, "the advection adjoint covers up3 and up3_tvd");
  if (h->d.nx < 4 || h->d.ny < 4)
    return fail(HOC_ERR_UNSUPPORTED, "the advection adjoint needs nx, ny >= 4");
  const Dims d = h->d;
  k_tracer_advection_adjoint<<<blocks(d.n3), TPB, 0, (hipStream_t)stream>>>(
      d, h->p, h->geo, c, u, v, w, g_adv, g_c, g_u, g_v, g_w);
  LAUNCH_CHECK("k_tracer_advection_adjoint");
  return HOC_OK;
}

int hoc_gaspar_lengths(const double* e, const double* n2, const double* dzi, const double* zs,
                       const double* zb, const double* wet, double e_min, double z_0,
                       double* l_up, double* l_dn, int64_t nz1, int64_t ncol, void* stream) {
  if (!e || !n2 || !dzi || !zs || !zb || !wet || !l_up || !l_dn)
    return fail(HOC_ERR_INVALID, "null argument");
  if (nz1 < 1 || ncol < 1) return fail(HOC_ERR_INVALID, "nz1 and ncol must be positive");
  k_gaspar_lengths<<<blocks(ncol), TPB, 0, (hipStream_t)stream>>>(ncol, (int)nz1, e_min, z_0, e,
                                                                  n2, dzi, zs, zb, wet, l_up, l_dn);
  LAUNCH_CHECK("k_gaspar_lengths");
  return HOC_OK;
}

int hoc_gaspar_lengths_adjoint(const double* e, const double* n2, const double* dzi,
                               const double* zs, const double* zb, const double* wet,
                               double e_min, double z_0, const double* g_up, const double* g_dn,
                               double* g_e, double* g_n2, int64_t nz1, int64_t ncol,
                               void* stream) {
  if (!e || !n2 || !dzi || !zs || !zb || !wet || !g_up || !g_dn || !g_e || !g_n2)
    return fail(HOC_ERR_INVALID, "null argument");
  if (nz1 < 1 || ncol < 1) return fail(HOC_ERR_INVALID, "nz1 and ncol must be positive");
  if (nz1 > 128) return fail(HOC_ERR_UNSUPPORTED, "the mixing-length adjoint needs nz + 1 <= 128");
  k_gaspar_lengths_adjoint<<<blocks(ncol), TPB, 0, (hipStream_t)stream>>>(
      ncol, (int)nz1, e_min, z_0, e, n2, dzi, zs, zb, wet, g_up, g_dn, g_e, g_n2);
  LAUNCH_CHECK("k_gaspar_lengths_adjoint");
  return HOC_OK;
}

int hoc_stepper_momentum_advection_adjoint(hoc_stepper* h, const double* u, const double* v,
                                           const double* w, const double* g_au,
                                           const double* g_av, double* g_u, double* g_v,
                                           double* g_w, void* stream) {
  if (!h || !u || !v || !w || !g_au || !g_av || !g_u || !g_v || !g_w)
    return fail(HOC_ERR_INVALID, "null argument");
  if (h->p.advection != ADV_UP3 && h->p.advection != ADV_UP3_TVD)
    return fail(HOC_ERR_UNSUPPORTED, "the advection adjoint covers up3 and up3_tvd");
  if (h->d.nx < 4 || h->d.ny < 4)
    return fail(HOC_ERR_UNSUPPORTED, "the advection adjoint needs nx, ny >= 4");
  const Dims d = h->d;
  k_momentum_advection_adjoint<<<blocks(d.n3), TPB, 0, (hipStream_t)stream>>>(
      d, h->p, h->geo, u, v, w, g_au, g_av, g_u, g_v, g_w);
  LAUNCH_CHECK("k_momentum_advection_adjoint");
  return HOC_OK;
}

int hoc_stepper_step_n(hoc_stepper* h, int32_t n, hoc_state* a, hoc_state* b, void* stream) {
  if (!h || !a || !b) return fail(HOC_ERR_INVALID, "null argument");
  if (n < 0) return fail(HOC_ERR_INVALID, "n must be >= 0");
  // every step leaves h->w = w of its output velocity (the theta scheme's
  // last back-substitution or the closing k_w_diag), which is the next
  // step's input: on a static geometry the next step takes it as it is
  const bool reuse_w = h->cfg.vcoord == 0;
  for (int32_t k = 0; k < n; ++k) {
    const bool w_ready = reuse_w && k > 0;
    const int rc = (k % 2 == 0) ? step_impl(h, a, b, stream, w_ready)
                                : step_impl(h, b, a, stream, w_ready);
    if (rc != HOC_OK) return rc;
  }
  return HOC_OK;
}

int hoc_eos_buoyancy(int32_t eos, double g, double rho0, double alpha_T, double beta_S,
                     double T0, double S0, const double* T, const double* S, const double* z,
                     double* b, int64_t n, void* stream) {
  if (!T || !S || !z || !b) return fail(HOC_ERR_INVALID, "null argument");
  if (eos < 0 || eos > 2) return fail(HOC_ERR_INVALID, "unknown eos (expected linear|seos|teos10)");
  if (n <= 0) return HOC_OK;
  StepParams p;
  std::memset(&p, 0, sizeof(p));
  p.eos = eos;
  p.g = g;
  p.rho0 = rho0;
  p.alpha_T = alpha_T;
  p.beta_S = beta_S;
  p.T0 = T0;
  p.S0 = S0;
  p.neg_alpha_T = -alpha_T;
  p.neg_g = -g;
  k_eos<<<blocks(n), TPB, 0, (hipStream_t)stream>>>(n, p, T, S, z, b);
  LAUNCH_CHECK("k_eos");
  return HOC_OK;
}

int hoc_eos_buoyancy_adjoint(int32_t eos, double g, double rho0, double alpha_T, double beta_S,
                             double T0, double S0, const double* T, const double* S,
                             const double* z, const double* g_b, double* g_T, double* g_S,
                             int64_t n, void* stream) {
  if (!T || !S || !z || !g_b || !g_T || !g_S) return fail(HOC_ERR_INVALID, "null argument");
  if (eos < 0 || eos > 2) return fail(HOC_ERR_INVALID, "unknown eos (expected linear|seos|teos10)");
  if (n <= 0) return HOC_OK;
  StepParams p;
  std::memset(&p, 0, sizeof(p));
  p.eos = eos;
  p.g = g;
  p.rho0 = rho0;
  p.alpha_T = alpha_T;
  p.beta_S = beta_S;
  p.T0 = T0;
  p.S0 = S0;
  p.neg_alpha_T = -alpha_T;
  p.neg_g = -g;
  k_eos_adjoint<<<blocks(n), TPB, 0, (hipStream_t)stream>>>(n, p, T, S, z, g_b, g_T, g_S);
  LAUNCH_CHECK("k_eos_adjoint");
  return HOC_OK;
}

int hoc_helm_divgrad(int32_t nx, int32_t ny, double dx, double dy, const double* Ku,
                     const double* Kv, const uint8_t* mask, const double* x, double* out,
                     void* stream) {
  if (!Ku || !Kv || !mask || !x || !out) return fail(HOC_ERR_INVALID, "null argument");
  if (nx < 1 || ny < 1) return fail(HOC_ERR_INVALID, "nx and ny must be positive");
  if (!(dx > 0.0) || !(dy > 0.0)) return fail(HOC_ERR_INVALID, "dx and dy must be positive");
  Dims d{};
  d.nx = nx;
  d.ny = ny;
  d.nz = 1;
  d.n2 = (long long)nx * ny;
  d.n3 = d.n2;
  d.ni = d.n2;
  const HelmOp op{Ku, Kv, mask, 0.0, 0.0, 0.0, dx * dx, dy * dy};
  k_helm_divgrad<<<blocks(d.n2), TPB, 0, (hipStream_t)stream>>>(d, op, x, out);
  LAUNCH_CHECK("k_helm_divgrad");
  return HOC_OK;
}

int hoc_pairwise_dot(const double* a, const double* b, int64_t n, double* out, double* scratch,
                     void* stream) {
  if (!a || !b || !out || !scratch) return fail(HOC_ERR_INVALID, "null argument");
  if (n < 0) return fail(HOC_ERR_INVALID, "negative length");
  hipStream_t s = (hipStream_t)stream;
  Reducer R;
  TRY(R.init(n));
  if (R.nchunks > 0) {
    k_pw_dot<<<(unsigned)R.nchunks, 512, 0, s>>>(a, b, R.plans, R.nfull, scratch);
  }
  k_seq_sum<<<1, 1, 0, s>>>(scratch, R.nchunks, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  R.release();
  if (e != hipSuccess) return fail(HOC_ERR_DEVICE, hipGetErrorString(e));
  return HOC_OK;
}

}  // extern "C"
