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

__device__ __forceinline__ void lvl_ij(const MgLevel& L, long long c, int& i, int& j) {
  i = (int)(c % L.nx);
  j = (int)(c / L.nx);
}

// A x at c with x masked first (solvers_var.py:47-53).
__device__ __forceinline__ double mg_apply(const MgLevel& L, const double* x, long long c) {
  int i, j;
  lvl_ij(L, c, i, j);
  const long long ce = (long long)j * L.nx + wrap(i, 1, L.nx);
  const long long cw = (long long)j * L.nx + wrap(i, -1, L.nx);
  const long long cn = (long long)wrap(j, 1, L.ny) * L.nx + i;
  const long long cs = (long long)wrap(j, -1, L.ny) * L.nx + i;
  const double mc = mk(L.mask[c]);
  const double xc = x[c] * mc;
  const double xe = x[ce] * mk(L.mask[ce]);
  const double xw = x[cw] * mk(L.mask[cw]);
  const double xn = x[cn] * mk(L.mask[cn]);
  const double xs = x[cs] * mk(L.mask[cs]);
  const double fx = L.Ku[c] * (xe - xc);
  const double fxw = L.Ku[cw] * (xc - xw);
  const double fy = L.Kv[c] * (xn - xc);
  const double fys = L.Kv[cs] * (xc - xs);
  const double lap = (fx - fxw) / L.dx2 + (fy - fys) / L.dy2;
  return (xc - L.coef * lap) * mc + xc * (1.0 - mc);
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
                                       const PcgScalars* sc) {
  if (finished(sc)) return;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.n) return;
  const int I = (int)(t % c.nx), J = (int)(t / c.nx);
  const long long f00 = (long long)(2 * J) * f.nx + 2 * I;
  const long long ch[4] = {f00, f00 + 1, f00 + f.nx, f00 + f.nx + 1};
  double r[4];
  for (int q = 0; q < 4; ++q) r[q] = (b[ch[q]] - mg_apply(f, x, ch[q])) * mk(f.mask[ch[q]]);
  bc[t] = 0.25 * (r[0] + r[1] + r[2] + r[3]);
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

}  // namespace hoc
