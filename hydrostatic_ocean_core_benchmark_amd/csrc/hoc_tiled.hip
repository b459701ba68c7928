// LDS-tiled column kernels ("2.5D"): a workgroup owns a TX x TY tile of
// columns and marches down the levels. At each level the fields that enter
// horizontal stencils are staged once in LDS with a 2-cell periodic halo,
// so every value is read from HBM about once instead of once per stencil
// point; the column's own vertical neighbours stay in a register window.
// Arithmetic and operand order are those of the untiled kernels (and of the
// reference's NumPy expressions), so the results are bitwise the same.
#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

constexpr int TX = 32, TY = 8, HALO = 2;
constexpr int SX = TX + 2 * HALO, SY = TY + 2 * HALO;
constexpr int TILE_THREADS = TX * TY;

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Tracer face value for the configured scheme (model3d_v05.py:486-494).
__device__ __forceinline__ double tracer_face(int adv, double cm, double c0, double c1, double c2,
                                              double vel, bool wm, bool w2) {
  if (adv == ADV_UP3 || adv == ADV_UP3_TVD)
    return kappa_face(cm, c0, c1, c2, vel, wm, w2, adv == ADV_UP3_TVD);
  if (adv == ADV_CENTRED2) return 0.5 * (c0 + c1);
  return vel > 0.0 ? c0 : c1;
}

// Vertical interface value at interface kk of a column whose window holds
// c[kk-2], c[kk-1], c[kk], c[kk+1] and the wet flags of c[kk-2], c[kk+1].
__device__ __forceinline__ double tracer_face_z(int adv, int kk, int nz, double far_up,
                                                double above, double below, double far_dn,
                                                bool mup, bool mdn, double w) {
  if (kk == nz) return 0.0;
  if (!(adv == ADV_UP3 || adv == ADV_UP3_TVD)) return 0.5 * (above + below);
  const bool hu = kk >= 2 && mup;
  const bool hd = kk <= nz - 2 && mdn;
  return kappa_face_z(kk >= 2 ? far_up : 0.0, above, below, kk <= nz - 2 ? far_dn : 0.0, w, hu,
                      hd, adv == ADV_UP3_TVD);
}

struct TileLds {
  double c[2][SY][SX];      // the tracers (T, S) or b, with halo
  uint8_t m[SY][SX];        // mask3 with halo
  double u[TY][TX + 1];     // u at i-1 .. i
  double v[TY + 1][TX];     // v at j-1 .. j
  double du[TY][TX + 1];    // dz3u
  double dv[TY + 1][TX];    // dz3v
  uint8_t mu[TY][TX + 1];   // mask3u
  uint8_t mv[TY + 1][TX];   // mask3v
};

// Both tracers of the tile's columns: advection + K_h Laplacian + explicit
// part of the vertical diffusion + surface flux, the forward sweep on tri_b
// (d* parked in the outputs), then the back-substitution and the EOS
// (libs/core/model3d_v05.py:466-533).
__global__ __launch_bounds__(TILE_THREADS) void k_tracers_tiled(
    Dims d, StepParams p, Geo geo, const double* __restrict__ T, const double* __restrict__ S,
    const double* __restrict__ b, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ Kh, const double* __restrict__ sub,
    const double* __restrict__ den, const double* __restrict__ cst, double* __restrict__ T_out,
    double* __restrict__ S_out, double* __restrict__ b_out) {
  __shared__ TileLds L;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long n2 = d.n2;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool active = i < nx && j < ny;
  const long long c = active ? (long long)j * nx + i : 0;
  const bool ts = p.tracers_ts != 0;
  const int ntr = ts ? 2 : 1;
  const double* A = ts ? T : b;
  double* Ao = ts ? T_out : b_out;
  const double flux[2] = {ts ? p.flux_T : 0.0, p.flux_S};
  const int adv = p.advection;

  // own-column windows: tracer at k-2..k+2, mask3 at k-2..k+2
  double win[2][5];
  bool mw[5];
  for (int q = 0; q < 5; ++q) {
    win[0][q] = 0.0;
    win[1][q] = 0.0;
    mw[q] = false;
  }
  if (active) {
    for (int q = 2; q < 5; ++q) {
      const int kk = q - 2;
      if (kk < nz) {
        win[0][q] = A[kk * n2 + c];
        if (ts) win[1][q] = S[kk * n2 + c];
        mw[q] = geo.m3[kk * n2 + c] != 0;
      }
    }
  }
  double w_k = active ? w[c] : 0.0;
  double kh_k = active ? Kh[c] : 0.0;
  double dzi_k = active ? geo.dzi[c] : 0.0;
  double cf_top[2] = {0.0, 0.0};  // interface value at the top of level k
  double dprev[2] = {0.0, 0.0};

  for (int k = 0; k < nz; ++k) {
    const long long o = (long long)k * n2;
    __syncthreads();
    // ---- stage level k
    for (int t = tid; t < SX * SY; t += TILE_THREADS) {
      const int lx = t % SX, ly = t / SX;
      const long long g = (long long)pmod(j0 + ly - HALO, ny) * nx + pmod(i0 + lx - HALO, nx);
      L.c[0][ly][lx] = A[o + g];
      if (ts) L.c[1][ly][lx] = S[o + g];
      L.m[ly][lx] = geo.m3[o + g];
    }
    for (int t = tid; t < TY * (TX + 1); t += TILE_THREADS) {
      const int lx = t % (TX + 1), ly = t / (TX + 1);
      const long long g = (long long)pmod(j0 + ly, ny) * nx + pmod(i0 + lx - 1, nx);
      L.u[ly][lx] = u[o + g];
      L.du[ly][lx] = geo.dz3u[o + g];
      L.mu[ly][lx] = geo.m3u[o + g];
    }
    for (int t = tid; t < (TY + 1) * TX; t += TILE_THREADS) {
      const int lx = t % TX, ly = t / TX;
      const long long g = (long long)pmod(j0 + ly - 1, ny) * nx + pmod(i0 + lx, nx);
      L.v[ly][lx] = v[o + g];
      L.dv[ly][lx] = geo.dz3v[o + g];
      L.mv[ly][lx] = geo.m3v[o + g];
    }
    __syncthreads();
    if (!active) continue;

    // ---- own-column values for level k
    const double w_k1 = w[o + n2 + c];
    const double kh_k1 = (k + 1 <= nz) ? Kh[o + n2 + c] : 0.0;
    const double dzi_k1 = geo.dzi[o + n2 + c];
    const double dz = geo.dz3[o + c];
    const double inv = safe_inv(dz);
    const bool m_k = mw[2];
    const bool bot_km1 = k >= 1 && mw[1] && !mw[2];          // cell k-1 is a bottom cell
    const bool bot_k = m_k && !((k + 1 < nz) && mw[3]);      // cell k is a bottom cell
    const int cx = tx + HALO, cy = ty + HALO;

    for (int r = 0; r < ntr; ++r) {
      const double(*C)[SX] = L.c[r];
      const double cc = C[cy][cx];
      double tend = 0.0;
      if (!ts) tend = p.neg_N2 * (0.5 * (w_k + w_k1));
      if (adv != ADV_NONE) {
        const double ue = L.u[ty][tx + 1], uw = L.u[ty][tx];
        const double vn = L.v[ty + 1][tx], vs = L.v[ty][tx];
        const double fcx = tracer_face(adv, C[cy][cx - 1], cc, C[cy][cx + 1], C[cy][cx + 2], ue,
                                       L.m[cy][cx - 1] != 0, L.m[cy][cx + 2] != 0);
        const double fcxw = tracer_face(adv, C[cy][cx - 2], C[cy][cx - 1], cc, C[cy][cx + 1], uw,
                                        L.m[cy][cx - 2] != 0, L.m[cy][cx + 1] != 0);
        const double fcy = tracer_face(adv, C[cy - 1][cx], cc, C[cy + 1][cx], C[cy + 2][cx], vn,
                                       L.m[cy - 1][cx] != 0, L.m[cy + 2][cx] != 0);
        const double fcys = tracer_face(adv, C[cy - 2][cx], C[cy - 1][cx], cc, C[cy + 1][cx], vs,
                                        L.m[cy - 2][cx] != 0, L.m[cy + 1][cx] != 0);
        const double fx = L.du[ty][tx + 1] * ue * fcx;
        const double fxw = L.du[ty][tx] * uw * fcxw;
        const double fy = L.dv[ty + 1][tx] * vn * fcy;
        const double fys = L.dv[ty][tx] * vs * fcys;
        const double hdiv = ((fx - fxw) / p.dx + (fy - fys) / p.dy) * inv;
        const double cft = (k == 0) ? win[r][2] : cf_top[r];
        const double cfb = tracer_face_z(adv, k + 1, nz, win[r][1], win[r][2], win[r][3],
                                         win[r][4], mw[1], mw[4], w_k1);
        cf_top[r] = cfb;
        const double vdiv = (w_k * cft - w_k1 * cfb) * inv;
        tend = tend - (hdiv + vdiv) * (m_k ? 1.0 : 0.0);
      }
      if (p.has_K_h) {
        const double gx = mk(L.mu[ty][tx + 1]) * (C[cy][cx + 1] - cc) / p.dx;
        const double gxw = mk(L.mu[ty][tx]) * (cc - C[cy][cx - 1]) / p.dx;
        const double gy = mk(L.mv[ty + 1][tx]) * (C[cy + 1][cx] - cc) / p.dy;
        const double gys = mk(L.mv[ty][tx]) * (cc - C[cy - 1][cx]) / p.dy;
        tend = tend + p.K_h * ((m_k ? 1.0 : 0.0) * ((gx - gxw) / p.dx + (gy - gys) / p.dy));
      }
      // explicit vertical diffusion with the surface flux (vertical_var.py:80-106)
      double ft;
      if (k == 0) {
        ft = flux[r];
      } else {
        ft = kh_k * (win[r][1] - win[r][2]) * safe_inv(dzi_k);
        if (bot_km1) ft = 0.0;
      }
      double fb = 0.0;
      if (k + 1 < nz) {
        fb = kh_k1 * (win[r][2] - win[r][3]) * safe_inv(dzi_k1);
        if (bot_k) fb = 0.0;
      }
      const double out = (ft - fb) * inv;
      const double dr = 0.0 * win[r][2] * (bot_k ? 1.0 : 0.0);
      const double dzd = (out - dr * inv) * (m_k ? 1.0 : 0.0);
      double rhs = cc + p.dt * tend + p.c_dzd * dzd;
      if (k == 0) rhs = rhs + p.thv_dt * flux[r] * safe_inv(dz);
      const long long t = o + c;
      const double dd = (k == 0) ? rhs / den[t] : (rhs - sub[t] * dprev[r]) / den[t];
      dprev[r] = dd;
      (r == 0 ? Ao : S_out)[t] = dd;
    }
    // ---- advance the windows
    for (int r = 0; r < 2; ++r)
      for (int q = 0; q < 4; ++q) win[r][q] = win[r][q + 1];
    for (int q = 0; q < 4; ++q) mw[q] = mw[q + 1];
    const int kn = k + 3;
    if (kn < nz) {
      win[0][4] = A[(long long)kn * n2 + c];
      if (ts) win[1][4] = S[(long long)kn * n2 + c];
      mw[4] = geo.m3[(long long)kn * n2 + c] != 0;
    } else {
      win[0][4] = 0.0;
      win[1][4] = 0.0;
      mw[4] = false;
    }
    w_k = w_k1;
    kh_k = kh_k1;
    dzi_k = dzi_k1;
  }
  if (!active) return;
  double xa = 0.0, xb = 0.0;
  for (int k = nz - 1; k >= 0; --k) {
    const long long t = (long long)k * n2 + c;
    const double mm = mk(geo.m3[t]);
    xa = (k == nz - 1) ? Ao[t] : Ao[t] - cst[t] * xa;
    Ao[t] = xa * mm;
    if (ts) {
      xb = (k == nz - 1) ? S_out[t] : S_out[t] - cst[t] * xb;
      S_out[t] = xb * mm;
      b_out[t] = buoyancy(p, xa * mm, xb * mm, geo.zc[t]) * mm;
    }
  }
}

}  // namespace hoc
