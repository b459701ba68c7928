// The constant-coefficient free-surface Helmholtz solves of the reference's
// original models, (I - coef L) eta = rhs on a doubly periodic grid
// (libs/core/solvers.py:55-281): the operator in its two written forms
// (D o G for PCG, the explicit 5-point form for the Gauss-Seidel family),
// the Jacobi and red-black Gauss-Seidel preconditioners/smoothers, and the
// multigrid transfers (2x2 mean restriction, piecewise-constant
// prolongation). One thread per cell; expressions keep the reference's
// operand order, so solves are bitwise the NumPy solvers'.
#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

__device__ __forceinline__ long long h2_at(const H2Level& L, int j, int i) {
  return (long long)j * L.nx + i;
}

// A p = p - coef D[G p]   (solvers.py:83-84, operators.py:25-53)
__global__ void k_h2_apply_dg(H2Level L, const double* __restrict__ p, double* __restrict__ ap) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n) return;
  const int i = (int)(t % L.nx), j = (int)(t / L.nx);
  const int ie = wrap(i, 1, L.nx), iw = wrap(i, -1, L.nx);
  const int jn = wrap(j, 1, L.ny), js = wrap(j, -1, L.ny);
  const double pc = p[t];
  const double gx = (p[h2_at(L, j, ie)] - pc) / L.dx;
  const double gxw = (pc - p[h2_at(L, j, iw)]) / L.dx;
  const double gy = (p[h2_at(L, jn, i)] - pc) / L.dy;
  const double gys = (pc - p[h2_at(L, js, i)]) / L.dy;
  const double lap = (gx - gxw) / L.dx + (gy - gys) / L.dy;
  ap[t] = pc - L.coef * lap;
}

// A x with the 5-point Laplacian written out (solvers.py:139-143):
// lap = ((x_e - 2 x) + x_w) / dx^2 + ((x_n - 2 x) + x_s) / dy^2
__device__ __forceinline__ double h2_apply_5pt(const H2Level& L, const double* x, int j, int i) {
  const double xc = x[h2_at(L, j, i)];
  const double two_x = 2.0 * xc;
  const double lx = (x[h2_at(L, j, wrap(i, 1, L.nx))] - two_x + x[h2_at(L, j, wrap(i, -1, L.nx))]) / L.dx2;
  const double ly = (x[h2_at(L, wrap(j, 1, L.ny), i)] - two_x + x[h2_at(L, wrap(j, -1, L.ny), i)]) / L.dy2;
  return xc - L.coef * (lx + ly);
}

// r = b - A x (5-point form)
__global__ void k_h2_resid(H2Level L, const double* __restrict__ x, const double* __restrict__ b,
                           double* __restrict__ r) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n) return;
  const int i = (int)(t % L.nx), j = (int)(t / L.nx);
  r[t] = b[t] - h2_apply_5pt(L, x, j, i);
}

// Red-black Gauss-Seidel update of one cell (solvers.py:131-136):
// x = (b + (cx (x_e + x_w) + cy (x_n + x_s))) / diag
__device__ __forceinline__ double h2_gs(const H2Level& L, const double* x, const double* b, int j,
                                        int i) {
  const double nb = L.cx * (x[h2_at(L, j, wrap(i, 1, L.nx))] + x[h2_at(L, j, wrap(i, -1, L.nx))]) +
                    L.cy * (x[h2_at(L, wrap(j, 1, L.ny), i)] + x[h2_at(L, wrap(j, -1, L.ny), i)]);
  return (b[h2_at(L, j, i)] + nb) / L.diag;
}

// One colour of a sweep in place, one thread per cell of that colour. With nx
// and ny even every neighbour of a cell has the other colour, so reading x in
// place sees exactly the array as it stood before this colour.
__global__ void k_h2_gs_colour(H2Level L, int colour, double* __restrict__ x,
                               const double* __restrict__ b) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int half = L.nx / 2;
  if (t >= (long long)half * L.ny) return;
  const int j = (int)(t / half);
  const int i = 2 * (int)(t % half) + ((j + colour) & 1);
  x[h2_at(L, j, i)] = h2_gs(L, x, b, j, i);
}

// One colour into a second array (any nx, ny): out = colour ? update(in) : in.
__global__ void k_h2_gs_colour_copy(H2Level L, int colour, const double* __restrict__ in,
                                    const double* __restrict__ b, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n) return;
  const int i = (int)(t % L.nx), j = (int)(t / L.nx);
  out[t] = (((i + j) & 1) == colour) ? h2_gs(L, in, b, j, i) : in[t];
}

// Whole sweeps of one small level in one workgroup. Each colour's new values
// are computed from the array as it stood, held in registers across a
// barrier and then stored, so odd-sized levels (whose periodic wrap puts
// same-coloured cells side by side) follow the reference's whole-array
// update exactly. `pattern` bit k set: sweep k runs black then red.
__global__ __launch_bounds__(H2_COARSE_THREADS) void k_h2_gs_block(H2Level L, int n_sweeps,
                                                                   unsigned pattern,
                                                                   double* __restrict__ x,
                                                                   const double* __restrict__ b) {
  double nv[H2_COARSE_PER];
  for (int s = 0; s < n_sweeps; ++s) {
    const int first = ((pattern >> s) & 1u) ? 1 : 0;
    for (int c = 0; c < 2; ++c) {
      const int colour = c == 0 ? first : 1 - first;
#pragma unroll
      for (int k = 0; k < H2_COARSE_PER; ++k) {
        const long long t = threadIdx.x + (long long)k * H2_COARSE_THREADS;
        if (t < L.n) {
          const int i = (int)(t % L.nx), j = (int)(t / L.nx);
          if (((i + j) & 1) == colour) nv[k] = h2_gs(L, x, b, j, i);
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < H2_COARSE_PER; ++k) {
        const long long t = threadIdx.x + (long long)k * H2_COARSE_THREADS;
        if (t < L.n) {
          const int i = (int)(t % L.nx), j = (int)(t / L.nx);
          if (((i + j) & 1) == colour) x[t] = nv[k];
        }
      }
      __syncthreads();
    }
  }
}

// Restriction of the fine residual (solvers.py:236-239), one thread per
// coarse cell: the 2x2 block's residuals b - A x summed as NumPy's two-axis
// mean sums them, ((r00 + r01) + (r10 + r11)) / 4; the coarse correction
// starts at zero.
__global__ void k_h2_restrict(H2Level F, H2Level Cl, const double* __restrict__ x,
                              const double* __restrict__ b, double* __restrict__ bc,
                              double* __restrict__ xc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Cl.n) return;
  const int I = (int)(t % Cl.nx), J = (int)(t / Cl.nx);
  const int i0 = 2 * I, j0 = 2 * J;
  const double r00 = b[h2_at(F, j0, i0)] - h2_apply_5pt(F, x, j0, i0);
  const double r01 = b[h2_at(F, j0, i0 + 1)] - h2_apply_5pt(F, x, j0, i0 + 1);
  const double r10 = b[h2_at(F, j0 + 1, i0)] - h2_apply_5pt(F, x, j0 + 1, i0);
  const double r11 = b[h2_at(F, j0 + 1, i0 + 1)] - h2_apply_5pt(F, x, j0 + 1, i0 + 1);
  bc[t] = ((r00 + r01) + (r10 + r11)) / 4.0;
  xc[t] = 0.0;
}

// x += P e, P piecewise constant over 2x2 blocks (solvers.py:241-243)
__global__ void k_h2_prolong(H2Level F, H2Level Cl, double* __restrict__ x,
                             const double* __restrict__ ec) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  const int i = (int)(t % F.nx), j = (int)(t / F.nx);
  x[t] = x[t] + ec[h2_at(Cl, j / 2, i / 2)];
}

// x = x + alpha p, r = r - alpha ap
__global__ void k_h2_xr(long long n, double alpha, const double* __restrict__ p,
                        const double* __restrict__ ap, double* __restrict__ x,
                        double* __restrict__ r) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  x[t] = x[t] + alpha * p[t];
  r[t] = r[t] - alpha * ap[t];
}

// z = inv_diag r
__global__ void k_h2_jacobi(long long n, double inv_diag, const double* __restrict__ r,
                            double* __restrict__ z) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  z[t] = inv_diag * r[t];
}

// p = z + beta p
__global__ void k_h2_p(long long n, double beta, const double* __restrict__ z,
                       double* __restrict__ p) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  p[t] = z[t] + beta * p[t];
}

__global__ void k_h2_fill(long long n, double value, double* __restrict__ x) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  x[t] = value;
}

__global__ void k_h2_copy(long long n, const double* __restrict__ in, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = in[t];
}

}  // namespace hoc
