/* C-ABI of the MI355X ocean-core step (libhoc.so).
 *
 * The library replaces the hot path of the reference's NumPy core behind its
 * step API; every entry point names the reference interface it stands in for.
 * Conventions:
 *   - Arrays are fp64 (masks uint8 0/1), C order: layers [nz][ny][nx],
 *     interfaces [nz+1][ny][nx], columns [ny][nx]; i (x) is fastest.
 *   - `hoc_geometry` arrays are HOST pointers, copied once at create.
 *   - State, mixing and operator arrays are DEVICE pointers owned by the
 *     caller (torch tensors in the Python binding).
 *   - `stream` is a hipStream_t (NULL = the default stream). Calls are
 *     asynchronous on it except where noted ("synchronises").
 *   - Every entry point returns HOC_OK or an error code; hoc_last_error()
 *     gives the message of the last failure on the calling thread. No C++
 *     exception crosses the ABI.
 *   - One handle per stream; a handle is not thread-safe.
 */
#ifndef HOC_H
#define HOC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HOC_OK 0
#define HOC_ERR_INVALID 1     /* invalid argument or configuration (ValueError) */
#define HOC_ERR_UNSUPPORTED 2 /* a configuration this backend does not run     */
#define HOC_ERR_DEVICE 3      /* HIP runtime failure (RuntimeError)            */

/* Scheme, physics and solver settings. Field meanings are those of
 * SchemeParams (libs/core/schemes.py:54-67), Physics3D (libs/core/model3d.py:57-81)
 * and PhysicsV05 (libs/core/model3d_v05.py:40-72). */
typedef struct hoc_config {
  int32_t nx, ny, nz;
  double dx, dy;
  double dz_nominal;       /* grid.dz = physics.H / nz, used by momentum vertical advection */
  double dt;
  int32_t scheme;          /* 0 theta, 1 fb, 2 split_explicit */
  double theta, theta_cor, theta_v;
  int32_t n_picard;
  int32_t advection;       /* 0 none, 1 centered2, 2 upwind1, 3 up3, 4 up3_tvd */
  int32_t solver;          /* 0 pcg_jacobi, 1 pcg_rbgs, 2 rbgs, 3 multigrid, 4 pcg_mg */
  double rtol;
  int32_t max_iter;
  int32_t mg_levels, mg_sweeps;  /* pcg_mg: levels and smoothing sweeps per side */
  double mg_omega;               /* pcg_mg: Jacobi damping */
  int32_t n_split;               /* split_explicit: barotropic substeps per step (>= 1) */
  int32_t barotropic_mean;       /* split_explicit: 1 = mean of the substep transports */
  int32_t open_x, open_y;        /* Flather open boundaries on the x / y domain edges */
  int32_t vcoord;                /* 0 z-level, 1 z* (zstar), 2 sigma: moving unless 0 */
  int32_t face_rule;             /* face thickness: 0 min, 1 mean, 2 harmonic */
  double eta_ext, u_ext;         /* Flather external level and transport */
  double g, f, rho0, nu_b, bottom_drag, N2;
  int32_t tracers_ts;      /* 1: T, S prognostic and b diagnosed; 0: b prognostic */
  double alpha_T, beta_S, T0, S0, cp, A_h, K_h, q_heat, q_salt;
  int32_t eos;             /* 0 linear, 1 seos, 2 teos10 */
  int32_t pgf_remove, pgf_correction;
  int32_t closure_tke, mxl_recursive;
  double c_k, c_eps, pr_t, kappa_vk, z_0, e_min, n2_min, l_min, e_bb;
  int32_t precision;       /* 0 fp64 (the plain step); 1 mixed: the pcg_mg V-cycle
                              preconditioner, the tracer advective fluxes and the EOS in
                              fp32, momentum fluxes, state updates and the outer PCG in
                              fp64; 2 as 1 with the momentum advective fluxes in fp32 */
  int32_t mg_smoother;     /* pcg_mg smoother: 0 damped Jacobi (mg_omega), 1 red-black
                              Gauss-Seidel (red then black down, black then red up; fp64,
                              mg_sweeps <= 3) */
} hoc_config;

/* Static z-level geometry and surface forcing (host arrays). Built like
 * Domain (libs/core/domain.py:24-195); the forcing fields are the array form
 * of the reference's scalar tau_x, tau_y (model3d_v05.py:176,355-356) and
 * kappa (model3d_v05.py:98, closure.py:143). */
typedef struct hoc_geometry {
  const uint8_t *mask, *mask_u, *mask_v;       /* [ny][nx] */
  const uint8_t *mask3, *mask3u, *mask3v;      /* [nz][ny][nx] */
  const double *dz3, *dz3u, *dz3v, *z_centre;  /* [nz][ny][nx] */
  const double *tau_x, *tau_y;                 /* [ny][nx] wind stress, N m-2 */
  const double *ustar2;                        /* [ny][nx] hypot(tau_x, tau_y) / rho0 */
  const double *kappa_b;                       /* [ny][nx] background diffusivity */
  const double *H;                             /* [ny][nx] depth; read for open edges and
                                                  moving vertical coordinates */
} hoc_geometry;

/* Prognostic state (device). T and S may be NULL when tracers_ts == 0;
 * e may be NULL when closure_tke == 0. Mirrors State3D (libs/core/model3d.py:34-47). */
typedef struct hoc_state {
  double *u, *v, *b, *eta, *T, *S, *e;
} hoc_state;

/* Stepper counters (libs/core/model3d_v05.py:89-94). */
typedef struct hoc_counters {
  int64_t solver_iterations, solver_failures, solver_rebuilds, tridiagonal_solves;
  int64_t barotropic_substeps, eos_cells;
} hoc_counters;

typedef struct hoc_stepper hoc_stepper;

/* Stepper3DV05.__init__ (libs/core/model3d_v05.py:81-125): validates, uploads
 * the geometry and allocates the workspaces. */
int hoc_stepper_create(const hoc_config* cfg, const hoc_geometry* geo, hoc_stepper** out);
/* Waits for the device, then frees the handle. If the runtime reports an
 * error (torn down at process exit, or a device fault) nothing is freed and
 * no further runtime call is made. Every *_destroy below behaves the same. */
int hoc_stepper_destroy(hoc_stepper* h);

/* The stepper's vertical viscosity / diffusivity on interfaces (the
 * reference's self.nu / self.kappa, model3d_v05.py:97-98,179), as
 * caller-owned device buffers [nz+1][ny][nx] that the TKE closure updates in
 * place every step. Must be bound before the first build or step. */
int hoc_stepper_bind_mixing(hoc_stepper* h, double* K_m, double* K_h);

/* Stepper3DV05._build_coefficients (model3d_v05.py:141-167): tri_m, tri_b,
 * q = A_m^-1 1, Ku, Kv and the Helmholtz diagonal from the bound K_m, K_h. */
int hoc_stepper_build_coefficients(hoc_stepper* h, void* stream);

/* Stepper3DV05.step (model3d_v05.py:337-415): reads `in`, writes `out`
 * (distinct buffers: an out field that aliases its in field is refused with
 * HOC_ERR_INVALID before anything runs). Synchronises on `stream` once per
 * Helmholtz solve to poll convergence. */
int hoc_stepper_step(hoc_stepper* h, const hoc_state* in, hoc_state* out, void* stream);

/* The intermediates of one step that its reverse (the adjoint sweep) reads,
 * caller-owned device arrays: q = A_m^-1 1 and the face coefficients K_u,
 * K_v of the step's mixing (model3d_v05.py:150-167), w of the input and of
 * the output velocity (:232-235), and per Picard pass the masked predictor
 * solutions gh_u, gh_v and the free surface it solved for (:387-408). */
#define HOC_TAPE_MAX_PICARD 4
typedef struct hoc_step_tape {
  double* q;                          /* [nz][ny][nx] */
  double* Ku;                         /* [ny][nx] */
  double* Kv;                         /* [ny][nx] */
  double* w_old;                      /* [nz+1][ny][nx] */
  double* w;                          /* [nz+1][ny][nx] */
  double* gh_u[HOC_TAPE_MAX_PICARD];  /* [nz][ny][nx], passes 0..n_picard-1 */
  double* gh_v[HOC_TAPE_MAX_PICARD];
  double* eta[HOC_TAPE_MAX_PICARD];   /* [ny][nx] */
} hoc_step_tape;

/* hoc_stepper_step that also writes the tape (no reference counterpart; the
 * forward half of the checkpointed adjoint). The same results as
 * hoc_stepper_step, bit for bit. Theta scheme, z-level. */
int hoc_stepper_step_tape(hoc_stepper* h, const hoc_state* in, hoc_state* out,
                          const hoc_step_tape* tape, void* stream);

/* The reverse of one step after its closure (no reference counterpart: the
 * reference has no adjoint; the step reversed is Stepper3DV05.step,
 * model3d_v05.py:337-415, theta branch). x is the step's input state (u, v,
 * T, S read), y its output (u, v, T, S read), Km, Kh ([nz+1][ny][nx]) the
 * mixing coefficients the step used (the closure's output) and tape the
 * step's tape (hoc_stepper_step_tape). gy holds the adjoints of y's fields
 * (u, v, b, eta, T, S; a NULL field counts as zero) and gy_Km, gy_Kh those
 * of Km, Kh as state of the next step (NULL: zero). Writes gx's u, v, b,
 * eta, T, S: the adjoints of x's fields through everything of the step
 * except its closure, and g_Km, g_Kh: the total adjoints of the step's Km,
 * Kh (gy's plus the step's uses), which the closure's reverse takes. The
 * Helmholtz solves of the reverse use the handle's solver on the tape's
 * K_u, K_v (the masked operator is symmetric). Theta scheme, z-level,
 * closed or periodic edges, up3 / up3_tvd, tracers T and S, fp64. */
int hoc_stepper_step_adjoint(hoc_stepper* h, const hoc_state* x, const hoc_state* y,
                             const double* Km, const double* Kh, const hoc_step_tape* tape,
                             const hoc_state* gy, const double* gy_Km, const double* gy_Kh,
                             hoc_state* gx, double* g_Km, double* g_Kh, void* stream);

/* The reverse of the step's closure, tke_step with integral mixing lengths
 * (no reference counterpart; closure.py:141-205 reversed). Inputs: the
 * step's input e ([nz+1][ny][nx]), u, v, b ([nz][ny][nx]), the mixing
 * coefficients Km, Kh it started from and the background kappa_b ([ny][nx]);
 * the adjoints of its outputs e_new (g_e_new, NULL: zero), K_m and K_h
 * (g_Km_new, g_Kh_new; hoc_stepper_step_adjoint's g_Km, g_Kh). Writes g_e,
 * g_Km, g_Kh (the adjoints of e, Km, Kh) and adds to g_u, g_v, g_b and
 * g_kappa_b. Branches of maxima, minima and the march held at their forward
 * values (a tie differentiates the first operand). */
int hoc_stepper_closure_adjoint(hoc_stepper* h, const double* e, const double* u,
                                const double* v, const double* b, const double* Km,
                                const double* Kh, const double* kappa_b, const double* g_e_new,
                                const double* g_Km_new, const double* g_Kh_new, double* g_e,
                                double* g_u, double* g_v, double* g_b, double* g_Km,
                                double* g_Kh, double* g_kappa_b, void* stream);

/* The current depth of the cell centres [nz][ny][nx] (device, written on
 * `stream`): fixed for z-level, rebuilt from eta every step for z* and sigma
 * (the reference's self.z_centre, model3d_v05.py:128-139). */
int hoc_stepper_z_centre(hoc_stepper* h, double* out, void* stream);

/* Counters accumulated since create (synchronises). */
int hoc_stepper_counters(hoc_stepper* h, hoc_counters* out);

/* Helmholtz solver launch strategy: 1 = two launches per PCG iteration (the
 * default), 0 = one persistent cooperative launch per solve with grid
 * barriers (when the device can co-schedule it). Both give bitwise the same
 * iterates. No reference counterpart. */
int hoc_stepper_set_solver_path(hoc_stepper* h, int32_t path);
/* The path the next solve will take (0 or 1; -1 for a null handle). */
int32_t hoc_stepper_solver_path(hoc_stepper* h);

/* Time every launch of one step kernel (e.g. "k_tracers") with HIP events
 * recorded on the launch stream; name NULL or "" stops timing. No reference
 * counterpart: measurement support for the roofline figures. */
int hoc_stepper_profile_kernel(hoc_stepper* h, const char* name);

/* Launches and total milliseconds of the timed kernel since the last read
 * (synchronises). */
int hoc_stepper_profile_read(hoc_stepper* h, int64_t* launches, double* total_ms);

/* --- single operators, for parity tests and for callers of the reference's
 *     module-level functions; all use the handle's geometry and settings --- */

/* closure.tke_step (libs/core/closure.py:141-205) with K_m, K_h the bound
 * buffers (old values in, new values out). e_out is [nz+1][ny][nx]. */
int hoc_stepper_tke_step(hoc_stepper* h, const double* e_in, const double* u, const double* v,
                         const double* b, double* e_out, void* stream);

/* advection_v06.momentum_advection_up3 / advection.momentum_advection
 * (libs/core/advection_v06.py:77-108, libs/core/advection.py:62-95) per the
 * configured scheme. w is [nz+1][ny][nx]. */
int hoc_stepper_momentum_advection(hoc_stepper* h, const double* u, const double* v,
                                   const double* w, double* adv_u, double* adv_v, void* stream);

/* Stepper3DV05._tracer_advection (model3d_v05.py:482-512). */
int hoc_stepper_tracer_advection(hoc_stepper* h, const double* c, const double* u,
                                 const double* v, const double* w, double* out, void* stream);

/* Stepper3DV05._pressure_gradients (model3d_v05.py:195-230). */
int hoc_stepper_pressure_gradients(hoc_stepper* h, const double* b, double* pgx, double* pgy,
                                   void* stream);

/* Stepper3DV05._w_diag (model3d_v05.py:232-235). */
int hoc_stepper_w_diag(hoc_stepper* h, const double* u, const double* v, double* w,
                       void* stream);

/* HelmholtzVar(..., kind=configured solver).solve (libs/core/solvers_var.py:162-197)
 * on face coefficients Ku, Kv [ny][nx] (device). Writes x; iterations and
 * converged are host outputs (synchronises). */
int hoc_stepper_helmholtz_solve(hoc_stepper* h, const double* Ku, const double* Kv,
                                const double* rhs, double* x, int32_t* iterations,
                                int32_t* converged, void* stream);

/* vertical.thomas_batched (libs/core/vertical.py:17-44): n rows per column,
 * ncol columns, arrays [n][ncol]; scratch is n*ncol doubles (device). */
int hoc_thomas_batched(const double* sub, const double* diag, const double* sup,
                       const double* rhs, double* x, double* scratch, int64_t n, int64_t ncol,
                       void* stream);

/* hoc_thomas_batched with no workspace argument and no device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). The c* sweep goes to a fixed device array of
 * HOC_NOALLOC_SCRATCH_DOUBLES doubles (16 MiB), one per GPU, held until the
 * process ends; n * ncol above that size is rejected (HOC_ERR_INVALID).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other threads) race. Same results as
 * hoc_thomas_batched, bit for bit. */
#define HOC_NOALLOC_SCRATCH_DOUBLES (1 << 21)
int hoc_thomas_batched_noalloc(const double* sub, const double* diag, const double* sup,
                               const double* rhs, double* x, int64_t n, int64_t ncol,
                               void* stream);

/* Adjoint of hoc_thomas_batched for x = A^-1 rhs (x from the forward solve):
 * lambda = A^-T g, rhs_bar = lambda, and the band gradients
 * d_sub[k] = -lambda[k] x[k-1], d_diag[k] = -lambda[k] x[k],
 * d_sup[k] = -lambda[k] x[k+1] (zero in the unused sub[0], sup[n-1]).
 * No reference counterpart (the reference has no adjoint); it is the
 * transposed solve of vertical.thomas_batched (libs/core/vertical.py:17-44).
 * All arrays [n][ncol] on the device; scratch n*ncol doubles. */
int hoc_thomas_adjoint(const double* sub, const double* diag, const double* sup, const double* x,
                       const double* g, double* lambda, double* d_sub, double* d_diag,
                       double* d_sup, double* scratch, int64_t n, int64_t ncol, void* stream);

/* Adjoint of hoc_stepper_tracer_advection (model3d_v05.py:482-512; no
 * reference counterpart): for L = sum(g_adv * adv(c, u, v, w)) writes dL/dc,
 * dL/du, dL/dv ([nz][ny][nx]) and dL/dw ([nz+1][ny][nx]), the upwind and
 * limiter branches held at their forward values. Needs advection up3 or
 * up3_tvd. */
int hoc_stepper_tracer_advection_adjoint(hoc_stepper* h, const double* c, const double* u,
                                         const double* v, const double* w, const double* g_adv,
                                         double* g_c, double* g_u, double* g_v, double* g_w,
                                         void* stream);

/* Gaspar integral mixing lengths of closure.gaspar_lengths
 * (libs/core/closure.py:65-103) on every interface, arrays [nz1][ncol] on
 * the device: from e, the interface N^2 (n2), the interface spacing dzi, the
 * depth below the surface zs and above the bottom zb and the interior-wet
 * flag wet (closure.py:43-62), write l_up and l_dn. */
int hoc_gaspar_lengths(const double* e, const double* n2, const double* dzi, const double* zs,
                       const double* zb, const double* wet, double e_min, double z_0,
                       double* l_up, double* l_dn, int64_t nz1, int64_t ncol, void* stream);

/* Adjoint of hoc_gaspar_lengths (no reference counterpart): g_e and g_n2 from
 * the adjoints g_up, g_dn of l_up, l_dn, the march's branches held at their
 * forward values. nz1 <= 128. */
int hoc_gaspar_lengths_adjoint(const double* e, const double* n2, const double* dzi,
                               const double* zs, const double* zb, const double* wet,
                               double e_min, double z_0, const double* g_up, const double* g_dn,
                               double* g_e, double* g_n2, int64_t nz1, int64_t ncol,
                               void* stream);

/* Sync-free stepping (no reference counterpart; for HIP-graph capture of the
 * step, e.g. an ensemble's): budget > 0 makes every Helmholtz solve launch
 * exactly `budget` PCG iterations with no host read-back; phases after
 * convergence return at once, and a solve not converged within the budget
 * counts as a solver failure (read with hoc_stepper_counters). 0 restores
 * the polled solve. Theta scheme with pcg_mg or pcg_rbgs. */
int hoc_stepper_set_sync_free(hoc_stepper* h, int32_t budget);

/* Tile shape of the fused multigrid V-cycle legs of pcg_mg (no reference
 * counterpart; a launch choice for the autotuner, the results are the same
 * bits for every shape): 0 32x16 (default), 1 64x8, 2 16x16, 3 64x16 cells. */
int hoc_stepper_set_mg_tile(hoc_stepper* h, int32_t shape);

/* The parts of a mixed-precision step that run in fp32 (no reference
 * counterpart; for attributing the mixed step's error and time to its
 * parts). A handle created with precision 1 runs HOC_FP32_MIXED, with 2
 * HOC_FP32_MIXED | HOC_FP32_MOMENTUM_FLUX; any subset may be selected on a
 * handle created with precision 1 or 2 (0 gives the fp64 step's results). */
#define HOC_FP32_VCYCLE 1         /* the pcg_mg V-cycle preconditioner */
#define HOC_FP32_TRACER_FLUX 2    /* tracer advective fluxes and K_h gradients */
#define HOC_FP32_EOS 4            /* the EOS (anomaly form) */
#define HOC_FP32_MOMENTUM_FLUX 8  /* momentum advective fluxes and A_h gradients */
#define HOC_FP32_MIXED (HOC_FP32_VCYCLE | HOC_FP32_TRACER_FLUX | HOC_FP32_EOS)
#define HOC_FP32_ALL (HOC_FP32_MIXED | HOC_FP32_MOMENTUM_FLUX)
int hoc_stepper_set_fp32_parts(hoc_stepper* h, int32_t parts);

/* Geometry form of the kernels (no reference counterpart): for a static
 * z-level grid the create call derives each column's wet-level count and
 * deepest wet thickness and, when they reproduce every stored thickness,
 * face thickness, mask and centre depth bit for bit, the kernels form those
 * from the columns instead of reading them per cell. column = 1 selects that
 * form (HOC_ERR_UNSUPPORTED when it is not available), 0 the stored arrays
 * (same results), -1 only queries; *active (may be NULL) receives the form
 * in use. */
int hoc_stepper_geometry_form(hoc_stepper* h, int32_t column, int32_t* active);

/* Adjoint of hoc_eos_buoyancy (no reference counterpart): g_T = g_b db/dT,
 * g_S = g_b db/dS at n points, the derivatives of the evaluated EOS
 * expression (forward-mode dual numbers through the same formula). */
int hoc_eos_buoyancy_adjoint(int32_t eos, double g, double rho0, double alpha_T, double beta_S,
                             double T0, double S0, const double* T, const double* S,
                             const double* z, const double* g_b, double* g_T, double* g_S,
                             int64_t n, void* stream);

/* Adjoint of hoc_stepper_momentum_advection (advection_v06.py:77-108; no
 * reference counterpart): from the adjoints g_au, g_av of the two advection
 * terms, dL/du, dL/dv ([nz][ny][nx]) and dL/dw ([nz+1][ny][nx]), the upwind
 * and limiter branches held at their forward values. up3 / up3_tvd. */
int hoc_stepper_momentum_advection_adjoint(hoc_stepper* h, const double* u, const double* v,
                                           const double* w, const double* g_au,
                                           const double* g_av, double* g_u, double* g_v,
                                           double* g_w, void* stream);

/* ------------------------------------------------------------------------
 * The 2D rotating shallow-water model of the reference's original driver
 * (libs/core/schemes.py:97-177; State u, v, eta [ny][nx] on a doubly
 * periodic C-grid, Physics g, f, H). The Python SWStepper drives the Picard
 * loop with these operators, as Stepper._step_fb/_step_theta do. */
typedef struct hoc_sw2d_config {
  int32_t nx, ny;
  double dx, dy, g, f, H, dt;
  int32_t scheme;              /* 0 theta, 1 fb */
  double theta, theta_cor;
} hoc_sw2d_config;

typedef struct hoc_sw2d hoc_sw2d;

int hoc_sw2d_create(const hoc_sw2d_config* cfg, hoc_sw2d** out);
int hoc_sw2d_destroy(hoc_sw2d* h);
/* One Picard update (schemes.py:127-135, 149-152): fb writes u', v'; theta
 * writes the predictors gu, gv. u0, v0, eta0: the step's start; u_it, v_it:
 * the current iterate. */
int hoc_sw2d_momentum(hoc_sw2d* h, const double* u0, const double* v0, const double* eta0,
                      const double* u_it, const double* v_it, double* out_u, double* out_v,
                      void* stream);
/* rhs = eta0 - dt H ((1 - theta) D[u0, v0] + theta D[gu, gv]) (schemes.py:154-155). */
int hoc_sw2d_rhs(hoc_sw2d* h, const double* eta0, const double* u0, const double* v0,
                 const double* gu, const double* gv, double* rhs, void* stream);
/* u = gu - g dt theta Gx eta, v likewise (schemes.py:161-162). */
int hoc_sw2d_backsub(hoc_sw2d* h, const double* gu, const double* gv, const double* eta,
                     double* out_u, double* out_v, void* stream);
/* eta' = eta0 - dt H D[u, v] (schemes.py:137). */
int hoc_sw2d_fb_eta(hoc_sw2d* h, const double* eta0, const double* u, const double* v,
                    double* out, void* stream);

/* ------------------------------------------------------------------------
 * The constant-coefficient free-surface Helmholtz solve of the original
 * models, (I - coef L) x = rhs on a doubly periodic nx x ny grid
 * (libs/core/solvers.py:55-281, built by build_solver :261-281): kind
 * 0 pcg_jacobi, 1 pcg_rbgs (HelmholtzPCG, precond "jacobi"/"rbgs"),
 * 2 rbgs (HelmholtzRBGS), 3 multigrid (HelmholtzMultigrid). The spectral
 * kind is the device FFT on the Python side. coef is g H theta^2 dt^2 as the
 * caller's stepper forms it. Kinds 1-3 need even nx and ny (the reference's
 * ValueError). */
typedef struct hoc_helm2d_config {
  int32_t nx, ny;
  double dx, dy, coef;
  int32_t kind;
  double rtol;
  int32_t max_iter;
} hoc_helm2d_config;

typedef struct hoc_helm2d hoc_helm2d;

int hoc_helm2d_create(const hoc_helm2d_config* cfg, hoc_helm2d** out);
int hoc_helm2d_destroy(hoc_helm2d* h);
/* solve(rhs) -> (x, SolveReport): writes x, the iteration (or sweep) count,
 * the final relative residual and whether it met rtol; synchronises (the
 * reference's convergence tests are host decisions). */
int hoc_helm2d_solve(hoc_helm2d* h, const double* rhs, double* x, int32_t* iterations,
                     double* residual, int32_t* converged, void* stream);
/* Number of multigrid levels (1 for the other kinds). */
int hoc_helm2d_levels(hoc_helm2d* h, int32_t* n_levels);

/* ------------------------------------------------------------------------
 * The v0.2 3D model (libs/core/model3d.py:93-391, Stepper3D): flat bottom,
 * nz uniform layers, doubly periodic; fields [nz][ny][nx], w [nz+1][ny][nx],
 * 2D fields [ny][nx]. The Python Stepper3D drives the step with these
 * operators (the Picard loop, the Helmholtz solve through hoc_helm2d or the
 * device FFT, the barotropic substeps), as Stepper3D.step/_step_split do.
 * Every pointer is a device buffer; every call is asynchronous on stream. */
typedef struct hoc_m3_config {
  int32_t nx, ny, nz;
  double dx, dy, dz, dt;
  double g, f, H, N2, nu, kappa, rho0, tau_x, tau_y, bottom_drag;
  int32_t tracers;             /* 0 buoyancy, 1 T and S */
  double alpha_T, beta_S, T0, S0, cp, A_h, K_h, q_heat, q_salt;
  int32_t scheme;              /* 0 theta, 1 fb, 2 split_explicit */
  double theta, theta_cor, theta_v;
  int32_t advection;           /* 0 none, 1 centered2, 2 upwind1 */
  int32_t n_split;             /* barotropic substeps (split_explicit) */
} hoc_m3_config;

typedef struct hoc_m3 hoc_m3;

/* Stepper3D.__init__ (model3d.py:96-151): the tridiagonal factors of
 * A = I - theta_v dt Dz for momentum (viscosity, bottom drag) and tracers
 * (diffusivity), and the effective-depth profile q = A^-1 1. */
int hoc_m3_create(const hoc_m3_config* cfg, hoc_m3** out);
int hoc_m3_destroy(hoc_m3* h);
/* q[nz] on the host (every column has the same profile). */
int hoc_m3_q(hoc_m3* h, double* q);
/* Once per step (model3d.py:158-161, 166-180, 329-343): phi' (the depth-mean
 * free buoyancy potential), U0 = dz sum u, V0 = dz sum v, w from D[u, v] and
 * the explicit stress st_u, st_v (surface stress, advection, A_h). */
int hoc_m3_prelude(hoc_m3* h, const double* u, const double* v, const double* b, double* phi,
                   double* U0, double* V0, double* w, double* st_u, double* st_v, void* stream);
/* One Picard pass of the predictor and its vertical solve
 * (model3d.py:352-364; split form :276-288): gh_u, gh_v, GU = dz sum gh_u,
 * GV, and for split_explicit cor_x = dz sum f v_cor, cor_y = -dz sum f u_cor
 * (cor_x, cor_y may be null otherwise). */
int hoc_m3_predict(hoc_m3* h, const double* u0, const double* v0, const double* phi,
                   const double* eta, const double* u_it, const double* v_it, const double* st_u,
                   const double* st_v, double* gh_u, double* gh_v, double* GU, double* GV,
                   double* cor_x, double* cor_y, void* stream);
/* rhs = eta - dt ((1 - theta) D[U0, V0] + theta D[GU, GV]) (model3d.py:367-369). */
int hoc_m3_rhs(hoc_m3* h, const double* eta, const double* U0, const double* V0,
               const double* GU, const double* GV, double* rhs, void* stream);
/* u = gh_u - g dt theta q Gx eta, v likewise (model3d.py:374-375). */
int hoc_m3_backsub(hoc_m3* h, const double* gh_u, const double* gh_v, const double* eta,
                   double* u, double* v, void* stream);
/* out = eta - dt_sub D[U, V] (fb surface, model3d.py:379-381; barotropic
 * substep, :235). */
int hoc_m3_eta_div(hoc_m3* h, double dt_sub, const double* eta, const double* U,
                   const double* V, double* out, void* stream);
/* Barotropic forcing fx = (GU - U0) / dt - cor_x, fy likewise (model3d.py:292-293). */
int hoc_m3_forcing(hoc_m3* h, const double* GU, const double* GV, const double* U0,
                   const double* V0, const double* cor_x, const double* cor_y, double* fx,
                   double* fy, void* stream);
/* One Picard pass of a barotropic substep (model3d.py:224-232). */
int hoc_m3_bt_momentum(hoc_m3* h, const double* u_i, const double* v_i, const double* e,
                       const double* ui, const double* vi, const double* fx, const double* fy,
                       double* ui_out, double* vi_out, void* stream);
/* acc_u += u_i, acc_v += v_i (model3d.py:236-237). */
int hoc_m3_bt_acc(hoc_m3* h, const double* u_i, const double* v_i, double* acc_u, double* acc_v,
                  void* stream);
/* u = gh_u - <gh_u>_z + Ub / (mean_n > 0 ? mean_n : 1) / H (model3d.py:241-246, 297-300). */
int hoc_m3_split_final(hoc_m3* h, int32_t mean_n, const double* gh_u, const double* gh_v,
                       const double* Ub, const double* Vb, double* u, double* v, void* stream);
/* w at interfaces from D[u, v] (model3d.py:384). */
int hoc_m3_w(hoc_m3* h, const double* u, const double* v, double* w, void* stream);
/* One tracer step (model3d.py:182-198): which 0 buoyancy (source -N2 w),
 * 1 temperature, 2 salinity (surface fluxes q_heat / (rho0 cp), q_salt). */
int hoc_m3_tracer(hoc_m3* h, int32_t which, const double* c, const double* u, const double* v,
                  const double* w, double* out, void* stream);
/* b = g (alpha_T (T - T0) - beta_S (S - S0)) (model3d.py:84-88). */
int hoc_m3_eos(hoc_m3* h, const double* T, const double* S, double* b, void* stream);

/* Stepper3DV05.integrate (libs/core/model3d_v05.py:535-538) on the device: n
 * steps from *a, ping-ponging between the caller's state sets a and b; the
 * final state is in a when n is even, in b when n is odd. */
int hoc_stepper_step_n(hoc_stepper* h, int32_t n, hoc_state* a, hoc_state* b, void* stream);

/* eos.buoyancy (libs/core/eos.py:130-133) for n points; eos 0/1/2 = linear,
 * seos, teos10; g, rho0, alpha_T, beta_S, T0, S0 as in Physics3D. */
int hoc_eos_buoyancy(int32_t eos, double g, double rho0, double alpha_T, double beta_S,
                     double T0, double S0, const double* T, const double* S, const double* z,
                     double* b, int64_t n, void* stream);

/* out = mask * D[K G x] on a doubly periodic ny x nx grid: the masked
 * five-point variable-coefficient operator of the barotropic Helmholtz
 * solve, with face coefficients Ku (east faces) and Kv (north faces). It is
 * operators_masked.div_m(Ku * gradx_u_m(x), Kv * grady_v_m(x)) of
 * libs/core/operators_masked.py:18-36 with unit face masks, the operator whose
 * truncation order tools/verify_v05.py:113-176 measures; x is masked by
 * `mask` before differencing, as in the Helmholtz solve. */
int hoc_helm_divgrad(int32_t nx, int32_t ny, double dx, double dy, const double* Ku,
                     const double* Kv, const uint8_t* mask, const double* x, double* out,
                     void* stream);

/* numpy.sum(a * b) for float64, bit for bit (sequential over 8192-element
 * chunks, pairwise inside each). out is one device double; scratch holds
 * ceil(n / 8192) doubles. */
int hoc_pairwise_dot(const double* a, const double* b, int64_t n, double* out, double* scratch,
                     void* stream);

const char* hoc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HOC_H */
