// Multigrid hierarchy and smoothers of the Helmholtz solvers
// (libs/core/solvers_var.py:63-160, plus the pcg_mg preconditioner):
//   coarse operators by face agglomeration, red-black Gauss-Seidel and
//   damped Jacobi sweeps, residual + full-weighting restriction, masked
//   piecewise-constant prolongation. Elementwise only -- every kernel is one
//   pass over a level -- and in the reference's operand order, so the
//   iterates are bitwise those of the oracle.
#include "hoc_device.h"
#include "hoc_kernels.h"
#include "hoc_reduce.h"

namespace hoc {

template <class R>
__device__ __forceinline__ void lvl_ij(const MgLevelT<R>& L, long long c, int& i, int& j) {
  i = (int)(c % L.nx);
  j = (int)(c / L.nx);
}

template <class R>
__device__ __forceinline__ R mkr(uint8_t m) {
  return m ? R(1) : R(0);
}

// A x at c from the values at c and its four neighbours, x masked first
// (solvers_var.py:47-53); ce/cw/cn/cs are the neighbours' indices. In the
// level's type R throughout (for R = double the reference's arithmetic).
// The same from the operands themselves: the faces' Ku[c], Ku[cw], Kv[c],
// Kv[cs] and the five cells' masks (as 0/1 values of R).
template <class R>
__device__ __forceinline__ R apply_from(R coef, R dx2, R dy2, R ku_c, R ku_w, R kv_c, R kv_s,
                                        R mc, R me, R mw, R mn, R ms, R x0, R x_e, R x_w, R x_n,
                                        R x_s) {
  const R xc = x0 * mc;
  const R xe = x_e * me;
  const R xw = x_w * mw;
  const R xn = x_n * mn;
  const R xs = x_s * ms;
  const R fx = ku_c * (xe - xc);
  const R fxw = ku_w * (xc - xw);
  const R fy = kv_c * (xn - xc);
  const R fys = kv_s * (xc - xs);
  const R lap = (fx - fxw) / dx2 + (fy - fys) / dy2;
  return (xc - coef * lap) * mc + xc * (R(1) - mc);
}

template <class R>
__device__ __forceinline__ R apply_vals(const MgLevelT<R>& L, long long c, long long ce,
                                        long long cw, long long cn, long long cs, R x0, R x_e,
                                        R x_w, R x_n, R x_s) {
  return apply_from<R>(L.coef, L.dx2, L.dy2, L.Ku[c], L.Ku[cw], L.Kv[c], L.Kv[cs],
                       mkr<R>(L.mask[c]), mkr<R>(L.mask[ce]), mkr<R>(L.mask[cw]),
                       mkr<R>(L.mask[cn]), mkr<R>(L.mask[cs]), x0, x_e, x_w, x_n, x_s);
}

// Damped Jacobi at c: x + omega (inv_diag (b - A x)) mask.
template <class R>
__device__ __forceinline__ R jacobi_vals(const MgLevelT<R>& L, R omega, long long c, long long ce,
                                         long long cw, long long cn, long long cs, R x0, R x_e,
                                         R x_w, R x_n, R x_s, R b) {
  const R ax = apply_vals<R>(L, c, ce, cw, cn, cs, x0, x_e, x_w, x_n, x_s);
  return x0 + omega * (L.inv_diag[c] * (b - ax)) * mkr<R>(L.mask[c]);
}

// Red-black Gauss-Seidel update at c: inv_diag (b + neighbour_sum(x)), the
// neighbours' values unmasked (solvers_var.py:55-69), in the operand order
// of mg_neighbour_sum.
template <class R>
__device__ __forceinline__ R rbgs_from(R inv, R b, R cx, R cy, R ku_c, R ku_w, R kv_c, R kv_s,
                                       R x_e, R x_w, R x_n, R x_s) {
  return inv * (b + (cx * (ku_c * x_e + ku_w * x_w) + cy * (kv_c * x_n + kv_s * x_s)));
}

template <class R>
__device__ __forceinline__ R rbgs_vals(const MgLevelT<R>& L, long long c, long long cw,
                                       long long cs, R b, R x_e, R x_w, R x_n, R x_s) {
  return rbgs_from<R>(L.inv_diag[c], b, L.cx, L.cy, L.Ku[c], L.Ku[cw], L.Kv[c], L.Kv[cs], x_e,
                      x_w, x_n, x_s);
}

struct Nbr {
  long long c, ce, cw, cn, cs;
};

template <class R>
__device__ __forceinline__ Nbr nbr_of(const MgLevelT<R>& L, int i, int j) {
  const long long row = (long long)j * L.nx;
  return Nbr{row + i, row + wrap(i, 1, L.nx), row + wrap(i, -1, L.nx),
             (long long)wrap(j, 1, L.ny) * L.nx + i, (long long)wrap(j, -1, L.ny) * L.nx + i};
}

__device__ __forceinline__ double mg_apply(const MgLevel& L, const double* x, long long c) {
  int i, j;
  lvl_ij(L, c, i, j);
  const Nbr n = nbr_of(L, i, j);
  return apply_vals(L, c, n.ce, n.cw, n.cn, n.cs, x[c], x[n.ce], x[n.cw], x[n.cn], x[n.cs]);
}

// -(A - diag) x, unmasked x (solvers_var.py:55-60).
__device__ __forceinline__ double mg_neighbour_sum(const MgLevel& L, const double* x, long long c) {
  int i, j;
  lvl_ij(L, c, i, j);
  const long long ce = (long long)j * L.nx + wrap(i, 1, L.nx);
  const long long cw = (long long)j * L.nx + wrap(i, -1, L.nx);
  const long long cn = (long long)wrap(j, 1, L.ny) * L.nx + i;
  const long long cs = (long long)wrap(j, -1, L.ny) * L.nx + i;
  return L.cx * (L.Ku[c] * x[ce] + L.Ku[cw] * x[cw]) +
         L.cy * (L.Kv[c] * x[cn] + L.Kv[cs] * x[cs]);
}

__device__ __forceinline__ int mg_pmod(int a, int n) {
  const int r = a % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ bool finished(const PcgScalars* sc) { return sc && sc->done; }

// Coarse faces: mean of the two fine faces on the coarse cell's east
// (north) edge; coarse cell wet if any child is (solvers_var.py:118-127).
__global__ void k_mg_coarsen(MgLevel f, MgLevel c, double* __restrict__ cKu,
                             double* __restrict__ cKv, uint8_t* __restrict__ cm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const int I = (int)(t % c.nx), J = (int)(t / c.nx);
  const long long f00 = (long long)(2 * J) * f.nx + 2 * I;
  const long long f01 = f00 + 1, f10 = f00 + f.nx, f11 = f10 + 1;
  cKu[t] = 0.5 * (f.Ku[f01] + f.Ku[f11]);
  cKv[t] = 0.5 * (f.Kv[f10] + f.Kv[f11]);
  uint8_t m = f.mask[f00];
  m = f.mask[f01] > m ? f.mask[f01] : m;
  m = f.mask[f10] > m ? f.mask[f10] : m;
  m = f.mask[f11] > m ? f.mask[f11] : m;
  cm[t] = m;
}

__global__ void k_mg_diag(MgLevel L, double* __restrict__ inv_diag) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  int i, j;
  lvl_ij(L, c, i, j);
  const long long cw = (long long)j * L.nx + wrap(i, -1, L.nx);
  const long long cs = (long long)wrap(j, -1, L.ny) * L.nx + i;
  double dg = 1.0 + L.cx * (L.Ku[c] + L.Ku[cw]) + L.cy * (L.Kv[c] + L.Kv[cs]);
  dg = L.mask[c] ? dg : 1.0;
  inv_diag[c] = 1.0 / dg;
}

// The face coefficient of a level `depth` levels below the fine one: the
// fine faces v[0], v[stride], ..., v[(2^depth - 1) stride] along the coarse
// cell's edge, averaged pairwise as k_mg_coarsen's recursion does, level by
// level (0.5 (left + right) at each node of the balanced tree).
__device__ __forceinline__ double mean_tree(const double* __restrict__ v, long long stride,
                                            int depth) {
  double st[MG_BUILD_LEVELS];
  const int len = 1 << depth;
  for (int q = 0; q < len; ++q) {
    double x = v[(long long)q * stride];
    int lv = 0;
    for (int qq = q; qq & 1; qq >>= 1) x = 0.5 * (st[lv++] + x);
    st[lv] = x;
  }
  return st[depth];
}

// The same tree for a depth known at compile time: every load issued at
// once, then the combine levels.
template <int D>
__device__ __forceinline__ double mean_tree_d(const double* __restrict__ v, long long stride) {
  constexpr int N = 1 << D;
  double x[N];
#pragma unroll
  for (int q = 0; q < N; ++q) x[q] = v[(long long)q * stride];
#pragma unroll
  for (int w = N; w > 1; w >>= 1)
#pragma unroll
    for (int q = 0; q < w / 2; ++q) x[q] = 0.5 * (x[2 * q] + x[2 * q + 1]);
  return x[0];
}

__device__ __forceinline__ double mean_tree_any(const double* __restrict__ v, long long stride,
                                                int depth) {
  switch (depth) {
    case 1: return mean_tree_d<1>(v, stride);
    case 2: return mean_tree_d<2>(v, stride);
    case 3: return mean_tree_d<3>(v, stride);
    case 4: return mean_tree_d<4>(v, stride);
    case 5: return mean_tree_d<5>(v, stride);
    default: return mean_tree(v, stride, depth);
  }
}

// Level l's Ku at (J, I) is the tree over the level-0 Ku of column
// (I + 1) 2^l - 1, rows J 2^l .. (J + 1) 2^l - 1; its Kv the tree over the
// level-0 Kv of row (J + 1) 2^l - 1, columns I 2^l .. (I + 1) 2^l - 1. Every
// level halves even sides, so level 0 is exactly 2^l times as wide.
__global__ void k_mg_levels(MgBuildArgs A) {
  const int l = (int)blockIdx.y + A.l_first;
  if (l >= A.nlev) return;
  const MgLevel& L = A.lev[l];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n) return;
  if (l == 0) {
    A.fKu[0][t] = (float)L.Ku[t];
    A.fKv[0][t] = (float)L.Kv[t];
    A.finv[0][t] = (float)L.inv_diag[t];
    return;
  }
  const MgLevel& F = A.lev[0];
  const int I = (int)(t % L.nx), J = (int)(t / L.nx);
  const long long span = 1ll << l;
  auto ku_at = [&](int jj, int ii) {
    return mean_tree_any(F.Ku + jj * span * F.nx + (ii + 1) * span - 1, F.nx, l);
  };
  auto kv_at = [&](int jj, int ii) {
    return mean_tree_any(F.Kv + ((jj + 1) * span - 1) * F.nx + ii * span, 1, l);
  };
  const double ku = ku_at(J, I), kv = kv_at(J, I);
  const double kuw = ku_at(J, wrap(I, -1, L.nx)), kvs = kv_at(wrap(J, -1, L.ny), I);
  double dg = 1.0 + L.cx * (ku + kuw) + L.cy * (kv + kvs);
  dg = L.mask[t] ? dg : 1.0;
  const double inv = 1.0 / dg;
  A.Ku[l][t] = ku;
  A.Kv[l][t] = kv;
  A.inv[l][t] = inv;
  if (A.fKu[l]) {
    A.fKu[l][t] = (float)ku;
    A.fKv[l][t] = (float)kv;
    A.finv[l][t] = (float)inv;
  }
}

__global__ void k_mg_zero(long long n, double* __restrict__ x, const PcgScalars* sc) {
  if (finished(sc)) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) x[c] = 0.0;
}

// One colour of a red-black Gauss-Seidel sweep, out of place: the colour's
// cells take inv_diag (b + neighbour_sum(xin)), the others keep xin. The
// second colour of a sweep (last = 1) masks the whole field, as the
// reference does after both colours (solvers_var.py:63-69). Out of place
// because a periodic level with an odd side has same-colour neighbours
// across the wrap, whose old values the reference reads.
__global__ void k_mg_rbgs(MgLevel L, int colour, int last, const double* __restrict__ b,
                          const double* __restrict__ xin, double* __restrict__ xout,
                          const PcgScalars* sc) {
  if (finished(sc)) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  int i, j;
  lvl_ij(L, c, i, j);
  const double v = (((i + j) & 1) == colour) ? L.inv_diag[c] * (b[c] + mg_neighbour_sum(L, xin, c))
                                             : xin[c];
  xout[c] = last ? v * mk(L.mask[c]) : v;
}

// One whole red-black sweep (first colour c0, then the other) in one launch,
// xout = sweep(xin) masked, exactly as the two launches of k_mg_rbgs compute
// it: a RB_TW x RB_TH tile stages the old field with a two-cell halo in LDS,
// computes the first colour on the tile and a one-cell halo from it (the
// other colour keeps its old value there), then the second colour on the
// tile from those values.
// Every first-colour update reads the old field only and every second-colour
// update reads first-colour values and old second-colour values, so odd
// periodic sides (where same-coloured cells meet) keep the whole-array
// semantics. Halo positions wrap; on sides shorter than the tile they alias
// cells of the tile and hold the same values.
__global__ __launch_bounds__(RB_TW * RB_TH) void k_mg_rbgs2(MgLevel L, int c0,
                                                            const double* __restrict__ b,
                                                            const double* __restrict__ xin,
                                                            double* __restrict__ xout,
                                                            const PcgScalars* sc) {
  if (finished(sc)) return;
  constexpr int SW = RB_TW + 2, SH = RB_TH + 2;   // first-colour region: tile + 1
  constexpr int XW = RB_TW + 4, XH = RB_TH + 4;   // old field: tile + 2
  __shared__ double xs[XH][XW];
  __shared__ double v1[SH][SW];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * RB_TW, j0 = blockIdx.y * RB_TH;
  for (int q = tid; q < XW * XH; q += RB_TW * RB_TH) {
    const int ly = q / XW, lx = q % XW;
    xs[ly][lx] = xin[(long long)mg_pmod(j0 + ly - 2, L.ny) * L.nx + mg_pmod(i0 + lx - 2, L.nx)];
  }
  __syncthreads();
  for (int q = tid; q < SW * SH; q += RB_TW * RB_TH) {
    const int ly = q / SW, lx = q % SW;
    const int i = mg_pmod(i0 + lx - 1, L.nx), j = mg_pmod(j0 + ly - 1, L.ny);
    const long long g = (long long)j * L.nx + i;
    if (((i + j) & 1) == c0) {
      const long long cw = (long long)j * L.nx + wrap(i, -1, L.nx);
      const long long cs = (long long)wrap(j, -1, L.ny) * L.nx + i;
      const double nb = L.cx * (L.Ku[g] * xs[ly + 1][lx + 2] + L.Ku[cw] * xs[ly + 1][lx]) +
                        L.cy * (L.Kv[g] * xs[ly + 2][lx + 1] + L.Kv[cs] * xs[ly][lx + 1]);
      v1[ly][lx] = L.inv_diag[g] * (b[g] + nb);
    } else {
      v1[ly][lx] = xs[ly + 1][lx + 1];
    }
  }
  __syncthreads();
  const int tx = tid % RB_TW, ty = tid / RB_TW;
  const int i = i0 + tx, j = j0 + ty;
  if (i >= L.nx || j >= L.ny) return;
  const long long c = (long long)j * L.nx + i;
  double v;
  if (((i + j) & 1) != c0) {
    const long long cw = (long long)j * L.nx + wrap(i, -1, L.nx);
    const long long cs = (long long)wrap(j, -1, L.ny) * L.nx + i;
    const double nb = L.cx * (L.Ku[c] * v1[ty + 1][tx + 2] + L.Ku[cw] * v1[ty + 1][tx]) +
                      L.cy * (L.Kv[c] * v1[ty + 2][tx + 1] + L.Kv[cs] * v1[ty][tx + 1]);
    v = L.inv_diag[c] * (b[c] + nb);
  } else {
    v = v1[ty + 1][tx + 1];
  }
  xout[c] = v * mk(L.mask[c]);
}

// Damped Jacobi: xo = x + omega (inv_diag (b - A x)) mask.
__global__ void k_mg_jacobi(MgLevel L, double omega, const double* __restrict__ x,
                            const double* __restrict__ b, double* __restrict__ xo,
                            const PcgScalars* sc) {
  if (finished(sc)) return;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= L.n) return;
  xo[c] = x[c] + omega * (L.inv_diag[c] * (b[c] - mg_apply(L, x, c))) * mk(L.mask[c]);
}

// bc = restrict((b - A x) mask), 1/4 of the 2 x 2 children in the order
// [0::2,0::2] + [0::2,1::2] + [1::2,0::2] + [1::2,1::2].
__global__ void k_mg_residual_restrict(MgLevel f, MgLevel c, const double* __restrict__ x,
                                       const double* __restrict__ b, double* __restrict__ bc,
                                       double* __restrict__ xc_zero, const PcgScalars* sc) {
  if (finished(sc)) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const int I = (int)(t % c.nx), J = (int)(t / c.nx);
  const long long f00 = (long long)(2 * J) * f.nx + 2 * I;
  const long long ch[4] = {f00, f00 + 1, f00 + f.nx, f00 + f.nx + 1};
  double r[4];
  for (int q = 0; q < 4; ++q) r[q] = (b[ch[q]] - mg_apply(f, x, ch[q])) * mk(f.mask[ch[q]]);
  bc[t] = 0.25 * (r[0] + r[1] + r[2] + r[3]);
  if (xc_zero) xc_zero[t] = 0.0;  // the coarse correction's zero start, when asked
}

// x += prolong(xc) * mask (piecewise constant).
__global__ void k_mg_prolong(MgLevel f, const double* __restrict__ xc, int cnx,
                             double* __restrict__ x, const PcgScalars* sc) {
  if (finished(sc)) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.n) return;
  int i, j;
  lvl_ij(f, t, i, j);
  x[t] = x[t] + xc[(long long)(j / 2) * cnx + i / 2] * mk(f.mask[t]);
}

// Partial sums of (rhs - A x)^2 per part (residual norm of the stationary
// solvers, solvers_var.py:176).
__global__ __launch_bounds__(PCG_WG) void k_mg_res_parts(PcgArgs a, MgLevel L,
                                                         const double* __restrict__ x) {
  if (a.sc->done) return;
  __shared__ double lds[2 * PCG_STAGE];
  __shared__ double leafs[128];
  const int part = blockIdx.x;
  const Two s = part_sum<1>(a, part, [&](long long e) -> Two {
    const double r = a.rhs[e] - mg_apply(L, x, e);
    return Two{r * r, 0.0};
  }, lds, leafs);
  if (threadIdx.x == 0) a.part_b[part] = s.a;
}

// Stationary-solver bookkeeping after iteration `it`: ||rhs - A x|| <= tol.
__global__ __launch_bounds__(PCG_WG) void k_mg_check(PcgArgs a, int it) {
  if (a.sc->done) return;
  __shared__ double cs[PCG_MAX_CHUNKS];
  const double res = sqrt(total_from_parts(a, a.part_b, cs));
  if (threadIdx.x != 0) return;
  PcgScalars* sc = a.sc;
  sc->res = res;
  if (res <= sc->tol) {
    sc->done = 1;
    sc->converged = 1;
    sc->iters = it;
    a.counters[0] += it;
  } else if (it >= a.max_iter) {
    sc->done = 1;
    sc->converged = 0;
    sc->iters = a.max_iter;
    a.counters[0] += a.max_iter;
    a.counters[1] += 1;
  }
}

// ||rhs|| and the zero-rhs shortcut for the stationary solvers; x = 0.
__global__ __launch_bounds__(PCG_WG) void k_mg_start(PcgArgs a) {
  __shared__ double cs[PCG_MAX_CHUNKS];
  const double nb = sqrt(total_from_parts(a, a.part_a, cs));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc->norm_b = nb;
    a.sc->tol = a.rtol * nb;
    a.sc->iters = 0;
    a.sc->converged = 1;
    a.sc->done = (nb == 0.0) ? 1 : 0;
    a.sc->res = 0.0;
  }
}

// ------------------------------------------------- fused pcg_mg V-cycle legs
// The pcg_mg preconditioner with S damped-Jacobi sweeps per side, as LDS
// tiles of MG_TW x MG_TH fine cells. Each sweep is evaluated on the tile
// plus the halo the later steps still need (radius 1 per remaining sweep,
// neighbours wrapping like the operator), so one launch replaces the
// zero / S sweeps / residual-restrict launches of the way down and the
// prolong / S sweeps launches of the way up, with the same arithmetic per
// cell (apply_vals, jacobi_vals). R is the arithmetic and storage type of
// the level's vectors; RB the type of the right-hand side read (double on
// level 0 of the fp32 cycle, where it is the fp64 PCG residual), RO the
// type of the way-up result (double on level 0 of the fp32 cycle).

// Way down on one level: x_S = S sweeps from zero (written out for the way
// up) and bc = restrict((b - A x_S) mask).
template <int S, class R, class RB, int TW = MG_TW, int TH = MG_TH>
__global__ __launch_bounds__(MG_THREADS) void k_mg_down(MgLevelT<R> L, MgLevelT<R> C, R omega,
                                                        const RB* __restrict__ b,
                                                        R* __restrict__ xs_out,
                                                        R* __restrict__ bc,
                                                        const PcgScalars* sc) {
  if (finished(sc)) return;
  constexpr int RW = TW + 2 * S, RH = TH + 2 * S;
  __shared__ R buf[2][RH * RW];
  const int i0 = blockIdx.x * TW, j0 = blockIdx.y * TH;
  for (int k = 1; k <= S; ++k) {
    const int h = S - k + 1;  // halo of x_k
    const int w = TW + 2 * h, hh = TH + 2 * h;
    R* dst = buf[(k - 1) & 1];
    const R* src = buf[k & 1];
    for (int t = threadIdx.x; t < w * hh; t += MG_THREADS) {
      const int li = S - h + t % w, lj = S - h + t / w;
      const Nbr n = nbr_of(L, mg_pmod(i0 - S + li, L.nx), mg_pmod(j0 - S + lj, L.ny));
      const int q = lj * RW + li;
      const R bb = (R)b[n.c];
      dst[q] = (k == 1)
                   ? jacobi_vals<R>(L, omega, n.c, n.ce, n.cw, n.cn, n.cs, R(0), R(0), R(0), R(0),
                                    R(0), bb)
                   : jacobi_vals<R>(L, omega, n.c, n.ce, n.cw, n.cn, n.cs, src[q], src[q + 1],
                                    src[q - 1], src[q + RW], src[q - RW], bb);
    }
    __syncthreads();
  }
  const R* xs = buf[(S - 1) & 1];
  for (int t = threadIdx.x; t < TW * TH; t += MG_THREADS) {
    const int gi = i0 + t % TW, gj = j0 + t / TW;
    if (gi < L.nx && gj < L.ny)
      xs_out[(long long)gj * L.nx + gi] = xs[(S + t / TW) * RW + S + t % TW];
  }
  constexpr int CW = TW / 2, CH = TH / 2;
  for (int t = threadIdx.x; t < CW * CH; t += MG_THREADS) {
    const int I = i0 / 2 + t % CW, J = j0 / 2 + t / CW;
    if (I >= C.nx || J >= C.ny) continue;
    R r[4];
    for (int qd = 0; qd < 4; ++qd) {  // [0::2,0::2], [0::2,1::2], [1::2,0::2], [1::2,1::2]
      const int a = qd & 1, bq = qd >> 1;
      const int gi = 2 * I + a, gj = 2 * J + bq;
      const Nbr n = nbr_of(L, gi, gj);
      const int q = (S + gj - j0) * RW + S + gi - i0;
      const R ax = apply_vals<R>(L, n.c, n.ce, n.cw, n.cn, n.cs, xs[q], xs[q + 1], xs[q - 1],
                                 xs[q + RW], xs[q - RW]);
      r[qd] = ((R)b[n.c] - ax) * mkr<R>(L.mask[n.c]);
    }
    bc[(long long)J * C.nx + I] = R(0.25) * (r[0] + r[1] + r[2] + r[3]);
  }
}

// Way up on one level: x = x_S + prolong(ec) mask, then S sweeps; out = x.
template <int S, class R, class RB, class RO, int TW = MG_TW, int TH = MG_TH>
__global__ __launch_bounds__(MG_THREADS) void k_mg_up(MgLevelT<R> L, int cnx, R omega,
                                                      const R* __restrict__ xs_in,
                                                      const R* __restrict__ ec,
                                                      const RB* __restrict__ b,
                                                      RO* __restrict__ out,
                                                      const PcgScalars* sc) {
  if (finished(sc)) return;
  constexpr int RW = TW + 2 * S, RH = TH + 2 * S;
  __shared__ R buf[2][RH * RW];
  const int i0 = blockIdx.x * TW, j0 = blockIdx.y * TH;
  for (int t = threadIdx.x; t < RW * RH; t += MG_THREADS) {
    const int gi = mg_pmod(i0 - S + t % RW, L.nx), gj = mg_pmod(j0 - S + t / RW, L.ny);
    const long long c = (long long)gj * L.nx + gi;
    buf[0][t] = xs_in[c] + ec[(long long)(gj / 2) * cnx + gi / 2] * mkr<R>(L.mask[c]);
  }
  __syncthreads();
  for (int k = 1; k <= S; ++k) {
    const int h = S - k;
    const int w = TW + 2 * h, hh = TH + 2 * h;
    R* dst = buf[k & 1];
    const R* src = buf[(k - 1) & 1];
    for (int t = threadIdx.x; t < w * hh; t += MG_THREADS) {
      const int li = S - h + t % w, lj = S - h + t / w;
      const Nbr n = nbr_of(L, mg_pmod(i0 - S + li, L.nx), mg_pmod(j0 - S + lj, L.ny));
      const int q = lj * RW + li;
      dst[q] = jacobi_vals<R>(L, omega, n.c, n.ce, n.cw, n.cn, n.cs, src[q], src[q + 1],
                              src[q - 1], src[q + RW], src[q - RW], (R)b[n.c]);
    }
    __syncthreads();
  }
  const R* x = buf[S & 1];
  for (int t = threadIdx.x; t < TW * TH; t += MG_THREADS) {
    const int gi = i0 + t % TW, gj = j0 + t / TW;
    if (gi < L.nx && gj < L.ny)
      out[(long long)gj * L.nx + gi] = (RO)x[(S + t / TW) * RW + S + t % TW];
  }
}

// ---------------------------------------- fused legs, RBGS smoothing
// The pcg_mg V-cycle with S red-black Gauss-Seidel sweeps per side (red then
// black on the way down, black then red on the way up; the wet mask after
// each sweep, as rbgs_sweep): each half-sweep updates one colour from the
// values of the previous half-sweep, evaluated on the tile plus the halo the
// later half-sweeps still need (radius 1 per remaining half-sweep), with the
// two LDS buffers alternating per half-sweep. The level's coefficients and
// right-hand side over the tile's region are staged in LDS once, so the
// half-sweeps read no global memory. Each thread takes a pair of
// neighbouring cells along x, one of each colour: one update per thread and
// half-sweep, no lane idles on the other colour. These legs run on the
// levels above the coarse cycle, which are all coarsened and so have even
// sides: a cell's colour is then the parity of its unwrapped coordinates.

// A tile region's copy of the level: face coefficients, 1/diag, right-hand
// side, wet mask (0/1).
template <int RN>
struct RbStage {
  double ku[RN], kv[RN], inv[RN], b[RN], m[RN];
};

// Stage the region of halo H around the tile at (i0, j0), positions wrapping
// like the operator.
template <int H, int TW, int TH>
__device__ __forceinline__ void rb_stage(RbStage<(TW + 2 * H) * (TH + 2 * H)>& st,
                                         const MgLevel& L, const double* __restrict__ b, int i0,
                                         int j0) {
  constexpr int RW = TW + 2 * H, RN = RW * (TH + 2 * H);
  for (int t = threadIdx.x; t < RN; t += MG_THREADS) {
    const int gi = mg_pmod(i0 - H + t % RW, L.nx), gj = mg_pmod(j0 - H + t / RW, L.ny);
    const long long c = (long long)gj * L.nx + gi;
    st.ku[t] = L.Ku[c];
    st.kv[t] = L.Kv[c];
    st.inv[t] = L.inv_diag[c];
    st.b[t] = b[c];
    st.m[t] = mkr<double>(L.mask[c]);
  }
}

// One half-sweep of colour `colour` on the region of halo h, src -> dst; the
// wet mask applied after the sweep's second colour (mask_all).
template <int H, int TW, int TH>
__device__ __forceinline__ void rb_half_sweep(const RbStage<(TW + 2 * H) * (TH + 2 * H)>& st,
                                              double cx, double cy, int i0, int j0, int h,
                                              int colour, bool mask_all, const double* src,
                                              double* dst) {
  constexpr int RW = TW + 2 * H;
  const int w2 = (TW + 2 * h) / 2, hh = TH + 2 * h;
  for (int t = threadIdx.x; t < w2 * hh; t += MG_THREADS) {
    const int lj = H - h + t / w2;
    const int li0 = H - h + 2 * (t % w2);
    const int a = ((i0 + li0 + j0 + lj) & 1) == colour ? 0 : 1;  // the pair's cell of this colour
    const int q = lj * RW + li0 + a, qo = lj * RW + li0 + 1 - a;
    const double v = rbgs_from<double>(st.inv[q], st.b[q], cx, cy, st.ku[q], st.ku[q - 1],
                                       st.kv[q], st.kv[q - RW], src[q + 1], src[q - 1],
                                       src[q + RW], src[q - RW]);
    const double vo = src[qo];
    dst[q] = mask_all ? v * st.m[q] : v;
    dst[qo] = mask_all ? vo * st.m[qo] : vo;
  }
}

// Way down: x_S = S sweeps from zero (written out for the way up) and
// bc = restrict((b - A x_S) mask).
template <int S, int TW = MG_TW, int TH = MG_TH>
__global__ __launch_bounds__(MG_THREADS) void k_mg_down_rb(MgLevel L, MgLevel C,
                                                           const double* __restrict__ b,
                                                           double* __restrict__ xs_out,
                                                           double* __restrict__ bc,
                                                           const PcgScalars* sc) {
  if (finished(sc)) return;
  constexpr int H = 2 * S + 1;  // 2S half-sweeps, then the residual's stencil
  constexpr int RW = TW + 2 * H, RN = RW * (TH + 2 * H);
  __shared__ RbStage<RN> st;
  __shared__ double buf[2][RN];
  const int i0 = blockIdx.x * TW, j0 = blockIdx.y * TH;
  rb_stage<H, TW, TH>(st, L, b, i0, j0);
  for (int t = threadIdx.x; t < RN; t += MG_THREADS) buf[0][t] = 0.0;
  __syncthreads();
  for (int ph = 1; ph <= 2 * S; ++ph) {  // red first
    rb_half_sweep<H, TW, TH>(st, L.cx, L.cy, i0, j0, H - ph, (ph - 1) & 1, (ph & 1) == 0,
                             buf[(ph - 1) & 1], buf[ph & 1]);
    __syncthreads();
  }
  const double* xs = buf[0];  // 2S half-sweeps end in buf[0]
  for (int t = threadIdx.x; t < TW * TH; t += MG_THREADS) {
    const int gi = i0 + t % TW, gj = j0 + t / TW;
    if (gi < L.nx && gj < L.ny)
      xs_out[(long long)gj * L.nx + gi] = xs[(H + t / TW) * RW + H + t % TW];
  }
  constexpr int CW = TW / 2, CH = TH / 2;
  for (int t = threadIdx.x; t < CW * CH; t += MG_THREADS) {
    const int I = i0 / 2 + t % CW, J = j0 / 2 + t / CW;
    if (I >= C.nx || J >= C.ny) continue;
    double r[4];
    for (int qd = 0; qd < 4; ++qd) {  // [0::2,0::2], [0::2,1::2], [1::2,0::2], [1::2,1::2]
      const int gi = 2 * I + (qd & 1), gj = 2 * J + (qd >> 1);
      const int q = (H + gj - j0) * RW + H + gi - i0;
      const double ax = apply_from<double>(
          L.coef, L.dx2, L.dy2, st.ku[q], st.ku[q - 1], st.kv[q], st.kv[q - RW], st.m[q],
          st.m[q + 1], st.m[q - 1], st.m[q + RW], st.m[q - RW], xs[q], xs[q + 1], xs[q - 1],
          xs[q + RW], xs[q - RW]);
      r[qd] = (st.b[q] - ax) * st.m[q];
    }
    bc[(long long)J * C.nx + I] = 0.25 * (r[0] + r[1] + r[2] + r[3]);
  }
}

// Way up: x = x_S + prolong(ec) mask, then S reversed sweeps; out = x.
template <int S, int TW = MG_TW, int TH = MG_TH>
__global__ __launch_bounds__(MG_THREADS) void k_mg_up_rb(MgLevel L, int cnx,
                                                         const double* __restrict__ xs_in,
                                                         const double* __restrict__ ec,
                                                         const double* __restrict__ b,
                                                         double* __restrict__ out,
                                                         const PcgScalars* sc) {
  if (finished(sc)) return;
  constexpr int H = 2 * S;
  constexpr int RW = TW + 2 * H, RN = RW * (TH + 2 * H);
  __shared__ RbStage<RN> st;
  __shared__ double buf[2][RN];
  const int i0 = blockIdx.x * TW, j0 = blockIdx.y * TH;
  rb_stage<H, TW, TH>(st, L, b, i0, j0);
  for (int t = threadIdx.x; t < RN; t += MG_THREADS) {  // st.m[t] was written by this thread
    const int gi = mg_pmod(i0 - H + t % RW, L.nx), gj = mg_pmod(j0 - H + t / RW, L.ny);
    const long long c = (long long)gj * L.nx + gi;
    buf[0][t] = xs_in[c] + ec[(long long)(gj / 2) * cnx + gi / 2] * st.m[t];
  }
  __syncthreads();
  for (int ph = 1; ph <= 2 * S; ++ph) {  // black first
    rb_half_sweep<H, TW, TH>(st, L.cx, L.cy, i0, j0, H - ph, ph & 1, (ph & 1) == 0,
                             buf[(ph - 1) & 1], buf[ph & 1]);
    __syncthreads();
  }
  const double* x = buf[0];
  for (int t = threadIdx.x; t < TW * TH; t += MG_THREADS) {
    const int gi = i0 + t % TW, gj = j0 + t / TW;
    if (gi < L.nx && gj < L.ny) out[(long long)gj * L.nx + gi] = x[(H + t / TW) * RW + H + t % TW];
  }
}

// The coarse end of the pcg_mg V-cycle in one workgroup, every vector in
// LDS: levels C.lev[0..nl) (the first with at most MG_COARSE_MAX cells, or
// the last level alone), S sweeps from zero + residual restriction down, 2S
// sweeps from zero on the last level, prolongation + S sweeps up; the same
// arithmetic per cell as the tiled legs (jacobi_vals or rbgs_vals,
// apply_vals, from their operands). With C.rb the sweeps are red-black
// Gauss-Seidel, reversed (black first) on the way up and in the last level's
// second S. b0 is level 0's right-hand side (staged in LDS), out receives
// level 0's result.

// One level's operands inside the cycle: its coefficients from their LDS
// copies (STAGED) or from the level's arrays, never a mix, so that the
// compiler addresses the copies as LDS.
template <class R, bool STAGED>
struct CycLevel {
  int nx, ny, n;
  R coef, cx, cy, dx2, dy2;
  const R *ku, *kv, *inv;
  const uint8_t* m;
};

// One cell of a level with its neighbours and colour (the parity of i + j).
struct CycCell {
  int c, ce, cw, cn, cs, par;
};

template <class R, bool STAGED>
__device__ __forceinline__ CycCell cyc_cell(const CycLevel<R, STAGED>& L, int c) {
  const int i = c % L.nx, j = c / L.nx;
  return CycCell{c, j * L.nx + wrap(i, 1, L.nx), j * L.nx + wrap(i, -1, L.nx),
                 wrap(j, 1, L.ny) * L.nx + i, wrap(j, -1, L.ny) * L.nx + i, (i + j) & 1};
}

template <class R, bool STAGED>
__device__ __forceinline__ R cyc_apply(const CycLevel<R, STAGED>& L, int c, int ce, int cw,
                                       int cn, int cs, R x0, R x_e, R x_w, R x_n, R x_s) {
  return apply_from<R>(L.coef, L.dx2, L.dy2, L.ku[c], L.ku[cw], L.kv[c], L.kv[cs],
                       mkr<R>(L.m[c]), mkr<R>(L.m[ce]), mkr<R>(L.m[cw]), mkr<R>(L.m[cn]),
                       mkr<R>(L.m[cs]), x0, x_e, x_w, x_n, x_s);
}

template <class R, class RB, class RO, bool STAGED>
__device__ __forceinline__ void coarse_cycle(const MgCycleT<R>& C, R omega, int S,
                                             const RB* __restrict__ b0, RO* __restrict__ out,
                                             unsigned char* cyc_raw) {
  R* cyc = reinterpret_cast<R*>(cyc_raw);
  const int n0 = (int)C.lev[0].n;
  const int tid = threadIdx.x;
  auto lev = [&](int l) {
    const MgLevelT<R>& G = C.lev[l];
    CycLevel<R, STAGED> L;
    L.nx = G.nx;
    L.ny = G.ny;
    L.n = (int)G.n;
    L.coef = G.coef;
    L.cx = G.cx;
    L.cy = G.cy;
    L.dx2 = G.dx2;
    L.dy2 = G.dy2;
    if (STAGED) {
      L.ku = cyc + C.coff[l];
      L.kv = cyc + C.coff[l] + L.n;
      L.inv = cyc + C.coff[l] + 2 * L.n;
      L.m = cyc_raw + C.moff[l];
    } else {
      L.ku = G.Ku;
      L.kv = G.Kv;
      L.inv = G.inv_diag;
      L.m = G.mask;
    }
    return L;
  };
  R* wA = cyc;
  R* wB = cyc + n0;
  R* wC = cyc + 2 * n0;  // red-black sweeps only: the field between the two colours
  // one sweep on level L from src (zero when `zero`) into dst: damped
  // Jacobi, or red-black Gauss-Seidel (red first unless reverse) through wC.
  // K0 is the thread's first cell (tid), formed once per visit of the
  // level; the others (levels of more than MG_COARSE_THREADS cells) are
  // formed as they come.
  auto for_cells = [&](const CycLevel<R, STAGED>& L, const CycCell& K0, auto&& f) {
    if (tid < L.n) f(K0);
    for (int c = tid + MG_COARSE_THREADS; c < L.n; c += MG_COARSE_THREADS) f(cyc_cell(L, c));
  };
  auto sweep = [&](const CycLevel<R, STAGED>& L, const CycCell& K0, const R* b, const R* src,
                   bool zero, R* dst, bool reverse) {
    if (!C.rb) {
      for_cells(L, K0, [&](const CycCell& K) {
        const int c = K.c;
        const R x0 = zero ? R(0) : src[c], xe = zero ? R(0) : src[K.ce];
        const R xw = zero ? R(0) : src[K.cw], xn = zero ? R(0) : src[K.cn];
        const R xs = zero ? R(0) : src[K.cs];
        const R ax = cyc_apply(L, c, K.ce, K.cw, K.cn, K.cs, x0, xe, xw, xn, xs);
        dst[c] = x0 + omega * (L.inv[c] * (b[c] - ax)) * mkr<R>(L.m[c]);
      });
      __syncthreads();
      return;
    }
    const int first = reverse ? 1 : 0;
    for_cells(L, K0, [&](const CycCell& K) {
      const int c = K.c;
      R v = zero ? R(0) : src[c];
      if (K.par == first) {
        const R xe = zero ? R(0) : src[K.ce], xw = zero ? R(0) : src[K.cw];
        const R xn = zero ? R(0) : src[K.cn], xs = zero ? R(0) : src[K.cs];
        v = rbgs_from<R>(L.inv[c], b[c], L.cx, L.cy, L.ku[c], L.ku[K.cw], L.kv[c], L.kv[K.cs],
                         xe, xw, xn, xs);
      }
      wC[c] = v;
    });
    __syncthreads();
    for_cells(L, K0, [&](const CycCell& K) {
      const int c = K.c;
      R v = wC[c];
      if (K.par != first)
        v = rbgs_from<R>(L.inv[c], b[c], L.cx, L.cy, L.ku[c], L.ku[K.cw], L.kv[c], L.kv[K.cs],
                         wC[K.ce], wC[K.cw], wC[K.cn], wC[K.cs]);
      dst[c] = v * mkr<R>(L.m[c]);
    });
    __syncthreads();
  };
  // level 0's right-hand side, staged (the restriction of the levels below
  // overwrites nothing of it), and the levels' coefficients when staged
  {
    R* st = cyc + C.boff[0];
    for (int c = tid; c < n0; c += MG_COARSE_THREADS) st[c] = (R)b0[c];
    if (STAGED) {
      for (int l = 0; l < C.nl; ++l) {
        const MgLevelT<R>& G = C.lev[l];
        const int n = (int)G.n;
        R* k = cyc + C.coff[l];
        unsigned char* m = cyc_raw + C.moff[l];
        for (int c = tid; c < n; c += MG_COARSE_THREADS) {
          k[c] = G.Ku[c];
          k[n + c] = G.Kv[c];
          k[2 * n + c] = G.inv_diag[c];
          m[c] = G.mask[c];
        }
      }
    }
    __syncthreads();
  }
  auto level_b = [&](int l) -> const R* { return cyc + C.boff[l]; };
  // down: x_S into XS[l], restricted residual into B[l + 1]
  for (int l = 0; l + 1 < C.nl; ++l) {
    const CycLevel<R, STAGED> L = lev(l);
    const CycCell K0 = cyc_cell(L, tid);
    const R* b = level_b(l);
    R* xs = cyc + C.xoff[l];
    const R* src = wA;
    for (int k = 1; k <= S; ++k) {
      R* dst = (k == S) ? xs : ((k & 1) ? wA : wB);
      sweep(L, K0, b, src, k == 1, dst, false);
      src = dst;
    }
    const MgLevelT<R>& Cl = C.lev[l + 1];  // sizes only
    R* bc = cyc + C.boff[l + 1];
    const int nc = (int)Cl.n;
    for (int t = tid; t < nc; t += MG_COARSE_THREADS) {
      const int I = t % Cl.nx, J = t / Cl.nx;
      R r[4];
      for (int qd = 0; qd < 4; ++qd) {  // [0::2,0::2], [0::2,1::2], [1::2,0::2], [1::2,1::2]
        const int gi = 2 * I + (qd & 1), gj = 2 * J + (qd >> 1);
        const int c = gj * L.nx + gi;
        const int ce = gj * L.nx + wrap(gi, 1, L.nx), cw = gj * L.nx + wrap(gi, -1, L.nx);
        const int cn = wrap(gj, 1, L.ny) * L.nx + gi, cs = wrap(gj, -1, L.ny) * L.nx + gi;
        const R ax = cyc_apply(L, c, ce, cw, cn, cs, xs[c], xs[ce], xs[cw], xs[cn], xs[cs]);
        r[qd] = (b[c] - ax) * mkr<R>(L.m[c]);
      }
      bc[t] = R(0.25) * (r[0] + r[1] + r[2] + r[3]);
    }
    __syncthreads();
  }
  // the last level: 2S sweeps from zero (the second S reversed)
  R* e = wA;
  {
    const int l = C.nl - 1;
    const CycLevel<R, STAGED> L = lev(l);
    const CycCell K0 = cyc_cell(L, tid);
    const R* b = level_b(l);
    const R* src = wB;
    for (int k = 1; k <= 2 * S; ++k) {
      R* dst = (k & 1) ? wA : wB;
      sweep(L, K0, b, src, k == 1, dst, k > S);
      src = dst;
      e = dst;
    }
  }
  // up: x = XS[l] + prolong(e) mask, then S sweeps
  for (int l = C.nl - 2; l >= 0; --l) {
    const CycLevel<R, STAGED> L = lev(l);
    const CycCell K0 = cyc_cell(L, tid);
    const R* b = level_b(l);
    const R* xs = cyc + C.xoff[l];
    const int cnx = C.lev[l + 1].nx;
    R* xp = (e == wA) ? wB : wA;
    R* other = (e == wA) ? wA : wB;
    for (int c = tid; c < L.n; c += MG_COARSE_THREADS) {
      const int i = c % L.nx, j = c / L.nx;
      xp[c] = xs[c] + e[(j / 2) * cnx + i / 2] * mkr<R>(L.m[c]);
    }
    __syncthreads();
    const R* src = xp;
    for (int k = 1; k <= S; ++k) {
      R* dst = (k & 1) ? other : xp;
      sweep(L, K0, b, src, false, dst, true);
      src = dst;
      e = dst;
    }
  }
  for (int c = tid; c < n0; c += MG_COARSE_THREADS) out[c] = (RO)e[c];
}

template <class R, class RB, class RO>
__global__ __launch_bounds__(MG_COARSE_THREADS) void k_mg_coarse_cycle(
    MgCycleT<R> C, R omega, int S, const RB* __restrict__ b0, RO* __restrict__ out,
    const PcgScalars* sc) {
  if (finished(sc)) return;
  extern __shared__ __align__(16) unsigned char cyc_raw[];
  if (C.stage)
    coarse_cycle<R, RB, RO, true>(C, omega, S, b0, out, cyc_raw);
  else
    coarse_cycle<R, RB, RO, false>(C, omega, S, b0, out, cyc_raw);
}

template <class R>
size_t mg_cycle_layout(MgCycleT<R>& C) {
  const int n0 = (int)C.lev[0].n;
  int off = (C.rb ? 3 : 2) * n0;
  for (int l = 0; l < C.nl; ++l) {
    C.boff[l] = off;
    off += (int)C.lev[l].n;
    C.xoff[l] = off;
    off += (int)C.lev[l].n;
  }
  const size_t vectors = (size_t)off * sizeof(R);
  for (int l = 0; l < C.nl; ++l) {
    C.coff[l] = off;
    off += 3 * (int)C.lev[l].n;
  }
  size_t bytes = (size_t)off * sizeof(R);
  for (int l = 0; l < C.nl; ++l) {
    C.moff[l] = (int)bytes;
    bytes += (size_t)C.lev[l].n;
  }
  bytes = (bytes + 15) / 16 * 16;
  C.stage = bytes <= MG_CYCLE_LDS ? 1 : 0;
  return C.stage ? bytes : vectors;
}
template size_t mg_cycle_layout<double>(MgCycle&);
template size_t mg_cycle_layout<float>(MgCycleF&);

namespace {
int set_cycle_lds(const void* fn, size_t lds) {
  if (lds > MG_CYCLE_LDS) return -1;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
                 hipSuccess
             ? 0
             : -2;
}
}  // namespace

int mg_cycle_prepare(size_t lds) {
  return set_cycle_lds((const void*)k_mg_coarse_cycle<double, double, double>, lds);
}

int mg_cycle_prepare_f32(size_t lds) {
  int rc = set_cycle_lds((const void*)k_mg_coarse_cycle<float, double, double>, lds);
  return rc ? rc : set_cycle_lds((const void*)k_mg_coarse_cycle<float, float, float>, lds);
}

void launch_mg_coarse_cycle(const MgCycle& C, size_t lds, double omega, int S, const double* b0,
                            double* out, const PcgScalars* sc, hipStream_t s) {
  k_mg_coarse_cycle<double, double, double>
      <<<1, MG_COARSE_THREADS, lds, s>>>(C, omega, S, b0, out, sc);
}

void launch_mg_coarse_cycle_f32(const MgCycleF& C, size_t lds, float omega, int S,
                                const double* b0_d, const float* b0_f, double* out_d,
                                float* out_f, const PcgScalars* sc, hipStream_t s) {
  if (b0_d)
    k_mg_coarse_cycle<float, double, double>
        <<<1, MG_COARSE_THREADS, lds, s>>>(C, omega, S, b0_d, out_d, sc);
  else
    k_mg_coarse_cycle<float, float, float>
        <<<1, MG_COARSE_THREADS, lds, s>>>(C, omega, S, b0_f, out_f, sc);
}

// Fused-leg tile shapes (cells per workgroup tile; every shape gives the
// same values, they differ only in halo recomputation and occupancy):
// 0: 32 x 16, 1: 64 x 8, 2: 16 x 16, 3: 64 x 16. Shapes 1-3 are built for
// 2 sweeps per side (the default), other sweep counts use shape 0.
dim3 mg_tile_grid_shape(const MgLevel& L, int shape) {
  static const int tw[MG_TILE_SHAPES] = {32, 64, 16, 64}, th[MG_TILE_SHAPES] = {16, 8, 16, 16};
  const int q = (shape >= 0 && shape < MG_TILE_SHAPES) ? shape : 0;
  return dim3((unsigned)((L.nx + tw[q] - 1) / tw[q]), (unsigned)((L.ny + th[q] - 1) / th[q]));
}

void launch_mg_down(int S, int shape, hipStream_t s, const MgLevel& L, const MgLevel& C,
                    double omega, const double* b, double* xs_out, double* bc,
                    const PcgScalars* sc) {
  if (S != 2) shape = 0;
  const dim3 grid = mg_tile_grid_shape(L, shape);
  if (S == 1)
    k_mg_down<1, double, double><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b, xs_out, bc, sc);
  else if (S == 3)
    k_mg_down<3, double, double><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b, xs_out, bc, sc);
  else if (shape == 1)
    k_mg_down<2, double, double, 64, 8><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b, xs_out, bc, sc);
  else if (shape == 2)
    k_mg_down<2, double, double, 16, 16><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b, xs_out, bc,
                                                                    sc);
  else if (shape == 3)
    k_mg_down<2, double, double, 64, 16><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b, xs_out, bc,
                                                                    sc);
  else
    k_mg_down<2, double, double><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b, xs_out, bc, sc);
}

void launch_mg_up(int S, int shape, hipStream_t s, const MgLevel& L, int cnx, double omega,
                  const double* xs_in, const double* ec, const double* b, double* out,
                  const PcgScalars* sc) {
  if (S != 2) shape = 0;
  const dim3 grid = mg_tile_grid_shape(L, shape);
  if (S == 1)
    k_mg_up<1, double, double, double><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in, ec, b,
                                                                   out, sc);
  else if (S == 3)
    k_mg_up<3, double, double, double><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in, ec, b,
                                                                   out, sc);
  else if (shape == 1)
    k_mg_up<2, double, double, double, 64, 8><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in, ec,
                                                                          b, out, sc);
  else if (shape == 2)
    k_mg_up<2, double, double, double, 16, 16><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in,
                                                                           ec, b, out, sc);
  else if (shape == 3)
    k_mg_up<2, double, double, double, 64, 16><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in,
                                                                           ec, b, out, sc);
  else
    k_mg_up<2, double, double, double><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in, ec, b,
                                                                   out, sc);
}

void launch_mg_down_rb(int S, int shape, hipStream_t s, const MgLevel& L, const MgLevel& C,
                       const double* b, double* xs_out, double* bc, const PcgScalars* sc) {
  if (S != 2) shape = 0;
  const dim3 grid = mg_tile_grid_shape(L, shape);
  if (S == 1)
    k_mg_down_rb<1><<<grid, MG_THREADS, 0, s>>>(L, C, b, xs_out, bc, sc);
  else if (S == 3)
    k_mg_down_rb<3><<<grid, MG_THREADS, 0, s>>>(L, C, b, xs_out, bc, sc);
  else if (shape == 1)
    k_mg_down_rb<2, 64, 8><<<grid, MG_THREADS, 0, s>>>(L, C, b, xs_out, bc, sc);
  else if (shape == 2)
    k_mg_down_rb<2, 16, 16><<<grid, MG_THREADS, 0, s>>>(L, C, b, xs_out, bc, sc);
  else if (shape == 3)
    k_mg_down_rb<2, 64, 16><<<grid, MG_THREADS, 0, s>>>(L, C, b, xs_out, bc, sc);
  else
    k_mg_down_rb<2><<<grid, MG_THREADS, 0, s>>>(L, C, b, xs_out, bc, sc);
}

void launch_mg_up_rb(int S, int shape, hipStream_t s, const MgLevel& L, int cnx,
                     const double* xs_in, const double* ec, const double* b, double* out,
                     const PcgScalars* sc) {
  if (S != 2) shape = 0;
  const dim3 grid = mg_tile_grid_shape(L, shape);
  if (S == 1)
    k_mg_up_rb<1><<<grid, MG_THREADS, 0, s>>>(L, cnx, xs_in, ec, b, out, sc);
  else if (S == 3)
    k_mg_up_rb<3><<<grid, MG_THREADS, 0, s>>>(L, cnx, xs_in, ec, b, out, sc);
  else if (shape == 1)
    k_mg_up_rb<2, 64, 8><<<grid, MG_THREADS, 0, s>>>(L, cnx, xs_in, ec, b, out, sc);
  else if (shape == 2)
    k_mg_up_rb<2, 16, 16><<<grid, MG_THREADS, 0, s>>>(L, cnx, xs_in, ec, b, out, sc);
  else if (shape == 3)
    k_mg_up_rb<2, 64, 16><<<grid, MG_THREADS, 0, s>>>(L, cnx, xs_in, ec, b, out, sc);
  else
    k_mg_up_rb<2><<<grid, MG_THREADS, 0, s>>>(L, cnx, xs_in, ec, b, out, sc);
}

namespace {
template <int S>
void down_f32(dim3 grid, hipStream_t s, const MgLevelF& L, const MgLevelF& C, float omega,
              const double* b_d, const float* b_f, float* xs_out, float* bc,
              const PcgScalars* sc) {
  if (b_d)
    k_mg_down<S, float, double><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b_d, xs_out, bc, sc);
  else
    k_mg_down<S, float, float><<<grid, MG_THREADS, 0, s>>>(L, C, omega, b_f, xs_out, bc, sc);
}
template <int S>
void up_f32(dim3 grid, hipStream_t s, const MgLevelF& L, int cnx, float omega, const float* xs_in,
            const float* ec, const double* b_d, const float* b_f, double* out_d, float* out_f,
            const PcgScalars* sc) {
  if (b_d)
    k_mg_up<S, float, double, double><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in, ec, b_d,
                                                                  out_d, sc);
  else
    k_mg_up<S, float, float, float><<<grid, MG_THREADS, 0, s>>>(L, cnx, omega, xs_in, ec, b_f,
                                                                out_f, sc);
}
}  // namespace

void launch_mg_down_f32(int S, dim3 grid, hipStream_t s, const MgLevelF& L, const MgLevelF& C,
                        float omega, const double* b_d, const float* b_f, float* xs_out,
                        float* bc, const PcgScalars* sc) {
  if (S == 1) down_f32<1>(grid, s, L, C, omega, b_d, b_f, xs_out, bc, sc);
  else if (S == 2) down_f32<2>(grid, s, L, C, omega, b_d, b_f, xs_out, bc, sc);
  else down_f32<3>(grid, s, L, C, omega, b_d, b_f, xs_out, bc, sc);
}

void launch_mg_up_f32(int S, dim3 grid, hipStream_t s, const MgLevelF& L, int cnx, float omega,
                      const float* xs_in, const float* ec, const double* b_d, const float* b_f,
                      double* out_d, float* out_f, const PcgScalars* sc) {
  if (S == 1) up_f32<1>(grid, s, L, cnx, omega, xs_in, ec, b_d, b_f, out_d, out_f, sc);
  else if (S == 2) up_f32<2>(grid, s, L, cnx, omega, xs_in, ec, b_d, b_f, out_d, out_f, sc);
  else up_f32<3>(grid, s, L, cnx, omega, xs_in, ec, b_d, b_f, out_d, out_f, sc);
}

// fp32 copies of one level's face coefficients and inverse diagonal
// (round to nearest).
__global__ void k_mg_to_f32(long long n, const double* __restrict__ Ku,
                            const double* __restrict__ Kv, const double* __restrict__ inv,
                            float* __restrict__ fKu, float* __restrict__ fKv,
                            float* __restrict__ finv) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  fKu[c] = (float)Ku[c];
  fKv[c] = (float)Kv[c];
  finv[c] = (float)inv[c];
}

}  // namespace hoc
