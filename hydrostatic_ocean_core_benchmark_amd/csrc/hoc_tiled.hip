// LDS-tiled column kernels ("2.5D"): a workgroup owns a TX x TY tile of
// columns and marches down the levels. At each level the fields that enter
// horizontal stencils are staged once in LDS with a 2-cell periodic halo,
// so every value is read from HBM about once instead of once per stencil
// point; the column's own vertical neighbours stay in a register window.
// Arithmetic and operand order are those of the untiled kernels (and of the
// reference's NumPy expressions), so the results are bitwise the same.
#include <type_traits>

#include "hoc_device.h"
#include "hoc_kernels.h"

namespace hoc {

// Tile shape: 32 x 8 (HOC_TILE_TX, HOC_TILE_TY override it at build time;
// every shape with TX * TY = 256 gives the same bits).
#ifndef HOC_TILE_TX
#define HOC_TILE_TX 32
#endif
#ifndef HOC_TILE_TY
#define HOC_TILE_TY 8
#endif
constexpr int TX = HOC_TILE_TX, TY = HOC_TILE_TY, HALO = 2;
static_assert(TX * TY == 256 && TX + TY <= 256, "tile of 256 threads");
constexpr int SX = TX + 2 * HALO, SY = TY + 2 * HALO;
constexpr int TILE_THREADS = TX * TY;

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// Tracer face value for the configured scheme (model3d_v05.py:486-494), in
// the flux type R (double: the reference's arithmetic; float: the mixed step).
template <class R>
__device__ __forceinline__ R tracer_face(int adv, R cm, R c0, R c1, R c2, R vel, bool wm,
                                         bool w2) {
  if (adv == ADV_UP3 || adv == ADV_UP3_TVD)
    return kappa_face<R>(cm, c0, c1, c2, vel, wm, w2, adv == ADV_UP3_TVD);
  if (adv == ADV_CENTRED2) return R(0.5) * (c0 + c1);
  return vel > R(0) ? c0 : c1;
}

// Vertical interface value at interface kk of a column whose window holds
// c[kk-2], c[kk-1], c[kk], c[kk+1] and the wet flags of c[kk-2], c[kk+1].
template <class R>
__device__ __forceinline__ R tracer_face_z(int adv, int kk, int nz, R far_up, R above, R below,
                                           R far_dn, bool mup, bool mdn, R w) {
  if (kk == nz) return R(0);
  if (!(adv == ADV_UP3 || adv == ADV_UP3_TVD)) return R(0.5) * (above + below);
  const bool hu = kk >= 2 && mup;
  const bool hd = kk <= nz - 2 && mdn;
  return kappa_face_z<R>(kk >= 2 ? far_up : R(0), above, below, kk <= nz - 2 ? far_dn : R(0), w,
                         hu, hd, adv == ADV_UP3_TVD);
}

// Staged tile of one level in the flux type R: what the face pass and the
// cell pass both read (double-buffered, so level k + 1 lands while level k
// is still read) ...
template <class R>
struct TileLds {
  R c[2][SY][SX];           // the tracers (T, S) or b, with halo
  uint8_t m[SY][SX];        // mask3 with halo
  uint8_t mu[TY][TX + 1];   // mask3u
  uint8_t mv[TY + 1][TX];   // mask3v
};

// ... and what only the face pass reads (one buffer: the barrier after the
// face pass ends its reads before level k + 1 is stored).
template <class R>
struct VelLds {
  R u[TY][TX + 1];          // u at i-1 .. i
  R v[TY + 1][TX];          // v at j-1 .. j
  R du[TY][TX + 1];         // dz3u
  R dv[TY + 1][TX];         // dz3v
};

// Advective face fluxes dz_f u_f c_f of one level, each computed once, on
// the x-faces (index q: between tile columns q-1 and q) and y-faces
// (between rows q-1 and q), for both tracers. The K_h face gradients are
// formed by the cell that uses them, from the staged tile (each interior
// gradient twice, from the same operands): with the single velocity buffer
// this holds a workgroup to 33 KB of LDS, so four fit on a CU.
template <class R>
struct FaceLds {
  R fx[2][TY][TX + 1];
  R fy[2][TY + 1][TX];
};

// K_h gradient mask_f (c_1 - c_0) / d across one face of the staged tile.
template <class R>
__device__ __forceinline__ R face_grad(uint8_t mf, R c0, R c1, double d) {
  return R(mk(mf)) * (c1 - c0) / R(d);
}

// Level-k staging of one tile held in registers between its global loads
// and its LDS stores, so the loads of level k + 1 are in flight while
// level k is computed (two LDS buffers, one barrier per level).
constexpr int ST_C = (SX * SY + TILE_THREADS - 1) / TILE_THREADS;        // tracer + mask slots
constexpr int ST_U = (TY * (TX + 1) + TILE_THREADS - 1) / TILE_THREADS;  // u-face slots
constexpr int ST_V = ((TY + 1) * TX + TILE_THREADS - 1) / TILE_THREADS;  // v-face slots

template <class R>
struct TileRegs {
  R c0[ST_C], c1[ST_C];
  uint8_t m[ST_C];
  R u[ST_U], du[ST_U];
  uint8_t mu[ST_U];
  R v[ST_V], dv[ST_V];
  uint8_t mv[ST_V];
};

// Plane offsets of this thread's staging slots (independent of k; -1: no slot).
struct TileOffs {
  int c[ST_C], u[ST_U], v[ST_V];
};

__device__ __forceinline__ TileOffs tile_offsets(int tid, int i0, int j0, int nx, int ny) {
  TileOffs F;
#pragma unroll
  for (int q = 0; q < ST_C; ++q) {
    const int t = tid + q * TILE_THREADS;
    F.c[q] = t < SX * SY ? pmod(j0 + t / SX - HALO, ny) * nx + pmod(i0 + t % SX - HALO, nx) : -1;
  }
#pragma unroll
  for (int q = 0; q < ST_U; ++q) {
    const int t = tid + q * TILE_THREADS;
    F.u[q] = t < TY * (TX + 1) ? pmod(j0 + t / (TX + 1), ny) * nx + pmod(i0 + t % (TX + 1) - 1, nx)
                               : -1;
  }
#pragma unroll
  for (int q = 0; q < ST_V; ++q) {
    const int t = tid + q * TILE_THREADS;
    F.v[q] = t < (TY + 1) * TX ? pmod(j0 + t / TX - 1, ny) * nx + pmod(i0 + t % TX, nx) : -1;
  }
  return F;
}

template <class RT>
__device__ __forceinline__ void tile_load(TileRegs<RT>& R, const TileOffs& F, long long o, bool ts,
                                          const double* A, const double* S, const double* u,
                                          const double* v, const Geo& geo) {
#pragma unroll
  for (int q = 0; q < ST_C; ++q) {
    if (F.c[q] >= 0) {
      const long long g = o + F.c[q];
      R.c0[q] = (RT)A[g];
      R.c1[q] = ts ? (RT)S[g] : RT(0);
      R.m[q] = geo.m3[g];
    }
  }
#pragma unroll
  for (int q = 0; q < ST_U; ++q) {
    if (F.u[q] >= 0) {
      const long long g = o + F.u[q];
      R.u[q] = (RT)u[g];
      R.du[q] = (RT)geo.dz3u[g];
      R.mu[q] = geo.m3u[g];
    }
  }
#pragma unroll
  for (int q = 0; q < ST_V; ++q) {
    if (F.v[q] >= 0) {
      const long long g = o + F.v[q];
      R.v[q] = (RT)v[g];
      R.dv[q] = (RT)geo.dz3v[g];
      R.mv[q] = geo.m3v[g];
    }
  }
}

template <class RT>
__device__ __forceinline__ void tile_store(TileLds<RT>& L, VelLds<RT>& V, const TileRegs<RT>& R,
                                           int tid) {
#pragma unroll
  for (int q = 0; q < ST_C; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < SX * SY) {
      const int lx = t % SX, ly = t / SX;
      L.c[0][ly][lx] = R.c0[q];
      L.c[1][ly][lx] = R.c1[q];
      L.m[ly][lx] = R.m[q];
    }
  }
#pragma unroll
  for (int q = 0; q < ST_U; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < TY * (TX + 1)) {
      const int lx = t % (TX + 1), ly = t / (TX + 1);
      V.u[ly][lx] = R.u[q];
      V.du[ly][lx] = R.du[q];
      L.mu[ly][lx] = R.mu[q];
    }
  }
#pragma unroll
  for (int q = 0; q < ST_V; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < (TY + 1) * TX) {
      const int lx = t % TX, ly = t / TX;
      V.v[ly][lx] = R.v[q];
      V.dv[ly][lx] = R.dv[q];
      L.mv[ly][lx] = R.mv[q];
    }
  }
}

// Both tracers of the tile's columns: advection + K_h Laplacian + explicit
// part of the vertical diffusion + surface flux, the forward sweep on tri_b
// (d* parked in the outputs), then the back-substitution and the EOS
// (libs/core/model3d_v05.py:466-533). R is the type of the horizontal and
// vertical advective fluxes, the K_h gradients and the EOS: double for the
// plain step (bitwise the reference), float for the mixed step; the column
// update, the vertical diffusion and the tridiagonal solve stay fp64.
// The fp32-flux instantiation runs at 4 waves per SIMD (128 VGPRs): measured
// at 1024 x 1024 x 60, 3.65 ms per step against 4.19 ms at the compiler's
// default of 3 waves.
template <class R>
__global__ __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1))) void k_tracers_tiled(
    Dims d, StepParams p, Geo geo, const double* __restrict__ T, const double* __restrict__ S,
    const double* __restrict__ b, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ Kh, double* __restrict__ cst,
    double* __restrict__ T_out, double* __restrict__ S_out, double* __restrict__ b_out) {
  __shared__ TileLds<R> LB[2];
  __shared__ VelLds<R> VL;
  __shared__ FaceLds<R> F;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long n2 = d.n2;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * TX, j0 = by * TY;
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
  R cf_top[2] = {R(0), R(0)};  // interface value at the top of level k
  double dprev[2] = {0.0, 0.0};
  double cb_prev = 0.0;  // c* of tri_b at the level above

  TileRegs<R> RG;
  const TileOffs TO = tile_offsets(tid, i0, j0, nx, ny);
  tile_load(RG, TO, 0, ts, A, S, u, v, geo);
  tile_store(LB[0], VL, RG, tid);
  __syncthreads();
  const bool has_adv = adv != ADV_NONE;
  for (int k = 0; k < nz; ++k) {
    const long long o = (long long)k * n2;
    const TileLds<R>& L = LB[k & 1];
    if (k + 1 < nz) tile_load(RG, TO, o + n2, ts, A, S, u, v, geo);
    // the own column's operands of level k, issued before the face pass so
    // their latency hides behind it
    double w_k1 = 0.0, kh_k1 = 0.0, dzi_k1 = 0.0, dz = 0.0;
    if (active) {
      w_k1 = w[o + n2 + c];
      kh_k1 = Kh[o + n2 + c];
      dzi_k1 = geo.dzi[o + n2 + c];
      dz = geo.dz3[o + c];
    }
    // ---- faces: the east and north face of every cell, plus the tile's
    // west column (8) and south row (32) of faces by threads 0..39
    {
      auto xface = [&](int fy_, int q) {
        const int cy = fy_ + HALO, c0x = q - 1 + HALO;
        const R vel = VL.u[fy_][q];
        const R ddu = VL.du[fy_][q];
        const bool wm = L.m[cy][c0x - 1] != 0, w2 = L.m[cy][c0x + 2] != 0;
        for (int r = 0; r < ntr; ++r) {
          const R(*C)[SX] = L.c[r];
          F.fx[r][fy_][q] = ddu * vel * tracer_face<R>(adv, C[cy][c0x - 1], C[cy][c0x],
                                                       C[cy][c0x + 1], C[cy][c0x + 2], vel, wm,
                                                       w2);
        }
      };
      auto yface = [&](int q, int fx_) {
        const int cx = fx_ + HALO, c0y = q - 1 + HALO;
        const R vel = VL.v[q][fx_];
        const R ddv = VL.dv[q][fx_];
        const bool wm = L.m[c0y - 1][cx] != 0, w2 = L.m[c0y + 2][cx] != 0;
        for (int r = 0; r < ntr; ++r) {
          const R(*C)[SX] = L.c[r];
          F.fy[r][q][fx_] = ddv * vel * tracer_face<R>(adv, C[c0y - 1][cx], C[c0y][cx],
                                                       C[c0y + 1][cx], C[c0y + 2][cx], vel, wm,
                                                       w2);
        }
      };
      if (has_adv) {
        xface(ty, tx + 1);
        yface(ty + 1, tx);
        if (tid < TY) xface(tid, 0);
        else if (tid < TY + TX) yface(0, tid - TY);
      }
    }
    __syncthreads();
    if (active) {
      // ---- own-column values for level k
      const double inv = safe_inv(dz);
      const bool m_k = mw[2];
      const bool bot_km1 = k >= 1 && mw[1] && !mw[2];          // cell k-1 is a bottom cell
      const bool bot_k = m_k && !((k + 1 < nz) && mw[3]);      // cell k is a bottom cell
      const int cx = tx + HALO, cy = ty + HALO;
      // row k of tri_b (K_h, no drag), factored here from the operands the
      // diffusion terms read anyway; c* parked in `cst` for the back-substitution
      double sub_k, den_k, cst_k;
      factor_row(p, k, inv, kh_k, kh_k1, safe_inv(dzi_k), safe_inv(dzi_k1), m_k,
                 (k + 1 < nz) && mw[3], p.thv_dt * 0.0, cb_prev, sub_k, den_k, cst_k);
      cst[o + c] = cst_k;
      cb_prev = cst_k;

      for (int r = 0; r < ntr; ++r) {
        const double cc = win[r][2];  // the tile's own cell, in fp64
        double tend = 0.0;
        if (!ts) tend = p.neg_N2 * (0.5 * (w_k + w_k1));
        if (has_adv) {
          const R hdiv = ((F.fx[r][ty][tx + 1] - F.fx[r][ty][tx]) / R(p.dx) +
                          (F.fy[r][ty + 1][tx] - F.fy[r][ty][tx]) / R(p.dy)) * R(inv);
          const R cft = (k == 0) ? R(win[r][2]) : cf_top[r];
          const R cfb = tracer_face_z<R>(adv, k + 1, nz, R(win[r][1]), R(win[r][2]),
                                         R(win[r][3]), R(win[r][4]), mw[1], mw[4], R(w_k1));
          cf_top[r] = cfb;
          const R vdiv = (R(w_k) * cft - R(w_k1) * cfb) * R(inv);
          tend = tend - (double)((hdiv + vdiv) * (m_k ? R(1) : R(0)));
        }
        if (p.has_K_h) {
          const R(*C)[SX] = L.c[r];
          const R ge = face_grad<R>(L.mu[ty][tx + 1], C[cy][cx], C[cy][cx + 1], p.dx);
          const R gw = face_grad<R>(L.mu[ty][tx], C[cy][cx - 1], C[cy][cx], p.dx);
          const R gn = face_grad<R>(L.mv[ty + 1][tx], C[cy][cx], C[cy + 1][cx], p.dy);
          const R gs = face_grad<R>(L.mv[ty][tx], C[cy - 1][cx], C[cy][cx], p.dy);
          tend = tend + p.K_h * (double)((m_k ? R(1) : R(0)) *
                                         ((ge - gw) / R(p.dx) + (gn - gs) / R(p.dy)));
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
        const double dd = (k == 0) ? rhs / den_k : (rhs - sub_k * dprev[r]) / den_k;
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
    if (k + 1 < nz) tile_store(LB[(k + 1) & 1], VL, RG, tid);
    __syncthreads();
  }
  if (!active) return;
  // back-substitution with the next level's operands loaded one level ahead
  double xa = 0.0, xb = 0.0;
  long long t = (long long)(nz - 1) * n2 + c;
  double a_n = Ao[t], s_n = ts ? S_out[t] : 0.0, c_n = cst[t], z_n = ts ? geo.zc[t] : 0.0;
  uint8_t m_n = geo.m3[t];
  for (int k = nz - 1; k >= 0; --k) {
    const double a_k = a_n, s_k = s_n, c_k = c_n, z_k = z_n;
    const double mm = mk(m_n);
    const long long tk = t;
    if (k > 0) {
      t -= n2;
      a_n = Ao[t];
      if (ts) s_n = S_out[t];
      c_n = cst[t];
      if (ts) z_n = geo.zc[t];
      m_n = geo.m3[t];
    }
    xa = (k == nz - 1) ? a_k : a_k - c_k * xa;
    Ao[tk] = xa * mm;
    if (ts) {
      xb = (k == nz - 1) ? s_k : s_k - c_k * xb;
      S_out[tk] = xb * mm;
      if (p.eos_mixed)
        b_out[tk] = buoyancy_mixed(p, xa * mm, xb * mm, z_k) * mm;
      else
        b_out[tk] = buoyancy(p, xa * mm, xb * mm, z_k) * mm;
    }
  }
}

}  // namespace hoc

namespace hoc {

// ------------------------------------------------ column-form geometry
// Column geometry of a tile and its 2-cell halo, staged once per launch:
// wet levels nw and deepest wet thickness hb per column, and the 2D face
// masks. Every thickness, face thickness and 3D mask of the tile at any
// level follows from these (cg_dz, cg_face in hoc_device.h).
struct ColTile {  // [SY + 1][SX + 1]: the face flags of the last halo row and column
  double hb[SY + 1][SX + 1];  // need the column beyond them
  uint8_t nw[SY + 1][SX + 1];
  uint8_t mu[SY + 1][SX + 1];  // mask_u: the east face of the column is open
  uint8_t mv[SY + 1][SX + 1];  // mask_v: the north face of the column is open
};

__device__ __forceinline__ void col_tile_load(ColTile& CG, const Geo& geo, int tid, int i0,
                                              int j0, int nx, int ny) {
  for (int t = tid; t < (SX + 1) * (SY + 1); t += TILE_THREADS) {
    const int lx = t % (SX + 1), ly = t / (SX + 1);
    const long long g = (long long)pmod(j0 + ly - HALO, ny) * nx + pmod(i0 + lx - HALO, nx);
    CG.hb[ly][lx] = geo.hb[g];
    CG.nw[ly][lx] = geo.nw[g];
    CG.mu[ly][lx] = geo.mu2[g];
    CG.mv[ly][lx] = geo.mv2[g];
  }
}

// Level-k wet flags of the east face of tile column (ly, lx) and of its
// north face (mask3u, mask3v).
__device__ __forceinline__ bool col_wet_u(const ColTile& CG, int ly, int lx, int k) {
  return CG.mu[ly][lx] && k < CG.nw[ly][lx] && k < CG.nw[ly][lx + 1];
}
__device__ __forceinline__ bool col_wet_v(const ColTile& CG, int ly, int lx, int k) {
  return CG.mv[ly][lx] && k < CG.nw[ly][lx] && k < CG.nw[ly + 1][lx];
}

// Per-level staging of the column-form tracer tile: the tracers with halo
// and the face velocities; no geometry.
template <class R>
struct TracerLdsC {
  R c[2][SY][SX];
};
template <class R>
struct VelLdsC {
  R u[TY][TX + 1];  // u at i-1 .. i
  R v[TY + 1][TX];  // v at j-1 .. j
};
template <class R>
struct TracerRegsC {
  R c0[ST_C], c1[ST_C];
  R u[ST_U];
  R v[ST_V];
};

template <class RT>
__device__ __forceinline__ void tracer_load_c(TracerRegsC<RT>& R, const TileOffs& F, long long o,
                                              bool ts, const double* A, const double* S,
                                              const double* u, const double* v) {
#pragma unroll
  for (int q = 0; q < ST_C; ++q) {
    if (F.c[q] >= 0) {
      const long long g = o + F.c[q];
      R.c0[q] = (RT)A[g];
      R.c1[q] = ts ? (RT)S[g] : RT(0);
    }
  }
#pragma unroll
  for (int q = 0; q < ST_U; ++q)
    if (F.u[q] >= 0) R.u[q] = (RT)u[o + F.u[q]];
#pragma unroll
  for (int q = 0; q < ST_V; ++q)
    if (F.v[q] >= 0) R.v[q] = (RT)v[o + F.v[q]];
}

template <class RT>
__device__ __forceinline__ void tracer_store_c(TracerLdsC<RT>& L, VelLdsC<RT>& V,
                                               const TracerRegsC<RT>& R, int tid) {
#pragma unroll
  for (int q = 0; q < ST_C; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < SX * SY) {
      const int lx = t % SX, ly = t / SX;
      L.c[0][ly][lx] = R.c0[q];
      L.c[1][ly][lx] = R.c1[q];
    }
  }
#pragma unroll
  for (int q = 0; q < ST_U; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < TY * (TX + 1)) V.u[t / (TX + 1)][t % (TX + 1)] = R.u[q];
  }
#pragma unroll
  for (int q = 0; q < ST_V; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < (TY + 1) * TX) V.v[t / TX][t % TX] = R.v[q];
  }
}

// k_tracers_tiled for a static z-level geometry in column form: the same
// operations on the same operands in the same order (so the same bits),
// with every thickness, face thickness, centre depth and 3D mask formed from
// the tile's column geometry instead of read per cell, and tri_b factored
// in the forward sweep. Per layer cell it reads T, S, u, v, w and K_h, and
// writes T, S, b, with d* (in the outputs) and c* (in `cst`) parked between
// the sweeps.
template <class R>
__global__ __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1))) void k_tracers_col(
    Dims d, StepParams p, Geo geo, const double* __restrict__ T, const double* __restrict__ S,
    const double* __restrict__ b, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ Kh, double* __restrict__ cst,
    double* __restrict__ T_out, double* __restrict__ S_out, double* __restrict__ b_out) {
  __shared__ TracerLdsC<R> LB[2];
  __shared__ VelLdsC<R> VL;
  __shared__ FaceLds<R> F;
  __shared__ ColTile CG;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long n2 = d.n2;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * TX, j0 = by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool active = i < nx && j < ny;
  const long long c = active ? (long long)j * nx + i : 0;
  const bool ts = p.tracers_ts != 0;
  const int ntr = ts ? 2 : 1;
  const double* A = ts ? T : b;
  double* Ao = ts ? T_out : b_out;
  const double flux[2] = {ts ? p.flux_T : 0.0, p.flux_S};
  const int adv = p.advection;
  const double dzr = p.dzr;
  const int rule = p.face_rule;
  const int cx = tx + HALO, cy = ty + HALO;

  TracerRegsC<R> RG;
  const TileOffs TO = tile_offsets(tid, i0, j0, nx, ny);
  tracer_load_c(RG, TO, 0, ts, A, S, u, v);
  col_tile_load(CG, geo, tid, i0, j0, nx, ny);
  tracer_store_c(LB[0], VL, RG, tid);
  __syncthreads();
  const int nwc = CG.nw[cy][cx];
  const double hbc = CG.hb[cy][cx];
  auto wet = [&](int kk) { return kk >= 0 && kk < nwc; };

  // own-column window: tracer at k-2..k+2
  double win[2][5];
  for (int q = 0; q < 5; ++q) win[0][q] = win[1][q] = 0.0;
  if (active) {
    for (int q = 2; q < 5; ++q) {
      const int kk = q - 2;
      if (kk < nz) {
        win[0][q] = A[kk * n2 + c];
        if (ts) win[1][q] = S[kk * n2 + c];
      }
    }
  }
  double w_k = active ? w[c] : 0.0;
  double kh_k = active ? Kh[c] : 0.0;
  double dzi_k = cg_dzi(dzr, nwc, hbc, 0, nz);
  R cf_top[2] = {R(0), R(0)};  // interface value at the top of level k
  double dprev[2] = {0.0, 0.0};
  double cb_prev = 0.0;  // c* of tri_b at the level above

  const bool has_adv = adv != ADV_NONE;
  for (int k = 0; k < nz; ++k) {
    const long long o = (long long)k * n2;
    const TracerLdsC<R>& L = LB[k & 1];
    if (k + 1 < nz) tracer_load_c(RG, TO, o + n2, ts, A, S, u, v);
    double w_k1 = 0.0, kh_k1 = 0.0;
    if (active) {
      w_k1 = w[o + n2 + c];
      kh_k1 = Kh[o + n2 + c];
    }
    const double dzi_k1 = cg_dzi(dzr, nwc, hbc, k + 1, nz);
    const double dz = cg_dz(dzr, nwc, hbc, k);
    // ---- faces: the east and north face of every cell, plus the tile's
    // west column (8) and south row (32) of faces by threads 0..39
    if (has_adv) {
      auto xface = [&](int fy_, int q) {
        const int yy = fy_ + HALO, c0x = q - 1 + HALO;
        const R vel = VL.u[fy_][q];
        const R ddu = R(cg_face(rule, dzr, CG.mu[yy][c0x] != 0, CG.nw[yy][c0x], CG.hb[yy][c0x],
                                CG.nw[yy][c0x + 1], CG.hb[yy][c0x + 1], k));
        const bool wm = k < CG.nw[yy][c0x - 1], w2 = k < CG.nw[yy][c0x + 2];
        for (int r = 0; r < ntr; ++r) {
          const R(*C)[SX] = L.c[r];
          F.fx[r][fy_][q] = ddu * vel * tracer_face<R>(adv, C[yy][c0x - 1], C[yy][c0x],
                                                       C[yy][c0x + 1], C[yy][c0x + 2], vel, wm,
                                                       w2);
        }
      };
      auto yface = [&](int q, int fx_) {
        const int xx = fx_ + HALO, c0y = q - 1 + HALO;
        const R vel = VL.v[q][fx_];
        const R ddv = R(cg_face(rule, dzr, CG.mv[c0y][xx] != 0, CG.nw[c0y][xx], CG.hb[c0y][xx],
                                CG.nw[c0y + 1][xx], CG.hb[c0y + 1][xx], k));
        const bool wm = k < CG.nw[c0y - 1][xx], w2 = k < CG.nw[c0y + 2][xx];
        for (int r = 0; r < ntr; ++r) {
          const R(*C)[SX] = L.c[r];
          F.fy[r][q][fx_] = ddv * vel * tracer_face<R>(adv, C[c0y - 1][xx], C[c0y][xx],
                                                       C[c0y + 1][xx], C[c0y + 2][xx], vel, wm,
                                                       w2);
        }
      };
      xface(ty, tx + 1);
      yface(ty + 1, tx);
      if (tid < TY) xface(tid, 0);
      else if (tid < TY + TX) yface(0, tid - TY);
    }
    __syncthreads();
    if (active) {
      const double inv = safe_inv(dz);
      const bool m_k = wet(k);
      const bool m_k1 = wet(k + 1);
      const bool bot_km1 = k >= 1 && wet(k - 1) && !m_k;  // cell k-1 is a bottom cell
      const bool bot_k = m_k && !((k + 1 < nz) && m_k1);   // cell k is a bottom cell
      double sub_k, den_k, cst_k;
      factor_row(p, k, inv, kh_k, kh_k1, safe_inv(dzi_k), safe_inv(dzi_k1), m_k,
                 (k + 1 < nz) && m_k1, p.thv_dt * 0.0, cb_prev, sub_k, den_k, cst_k);
      cst[o + c] = cst_k;
      cb_prev = cst_k;
      const uint8_t mue = col_wet_u(CG, cy, cx, k), muw = col_wet_u(CG, cy, cx - 1, k);
      const uint8_t mvn = col_wet_v(CG, cy, cx, k), mvs = col_wet_v(CG, cy - 1, cx, k);

      for (int r = 0; r < ntr; ++r) {
        const double cc = win[r][2];  // the tile's own cell, in fp64
        double tend = 0.0;
        if (!ts) tend = p.neg_N2 * (0.5 * (w_k + w_k1));
        if (has_adv) {
          const R hdiv = ((F.fx[r][ty][tx + 1] - F.fx[r][ty][tx]) / R(p.dx) +
                          (F.fy[r][ty + 1][tx] - F.fy[r][ty][tx]) / R(p.dy)) * R(inv);
          const R cft = (k == 0) ? R(win[r][2]) : cf_top[r];
          const R cfb = tracer_face_z<R>(adv, k + 1, nz, R(win[r][1]), R(win[r][2]),
                                         R(win[r][3]), R(win[r][4]), wet(k - 1), wet(k + 2),
                                         R(w_k1));
          cf_top[r] = cfb;
          const R vdiv = (R(w_k) * cft - R(w_k1) * cfb) * R(inv);
          tend = tend - (double)((hdiv + vdiv) * (m_k ? R(1) : R(0)));
        }
        if (p.has_K_h) {
          const R(*C)[SX] = L.c[r];
          const R ge = face_grad<R>(mue, C[cy][cx], C[cy][cx + 1], p.dx);
          const R gw = face_grad<R>(muw, C[cy][cx - 1], C[cy][cx], p.dx);
          const R gn = face_grad<R>(mvn, C[cy][cx], C[cy + 1][cx], p.dy);
          const R gs = face_grad<R>(mvs, C[cy - 1][cx], C[cy][cx], p.dy);
          tend = tend + p.K_h * (double)((m_k ? R(1) : R(0)) *
                                         ((ge - gw) / R(p.dx) + (gn - gs) / R(p.dy)));
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
        const double dd = (k == 0) ? rhs / den_k : (rhs - sub_k * dprev[r]) / den_k;
        dprev[r] = dd;
        (r == 0 ? Ao : S_out)[o + c] = dd;
      }
      // ---- advance the window
      for (int r = 0; r < 2; ++r)
        for (int q = 0; q < 4; ++q) win[r][q] = win[r][q + 1];
      const int kn = k + 3;
      if (kn < nz) {
        win[0][4] = A[(long long)kn * n2 + c];
        if (ts) win[1][4] = S[(long long)kn * n2 + c];
      } else {
        win[0][4] = 0.0;
        win[1][4] = 0.0;
      }
      w_k = w_k1;
      kh_k = kh_k1;
    }
    dzi_k = dzi_k1;
    if (k + 1 < nz) tracer_store_c(LB[(k + 1) & 1], VL, RG, tid);
    __syncthreads();
  }
  if (!active) return;
  // back-substitution with the next level's operands loaded one level ahead
  double xa = 0.0, xb = 0.0;
  long long t = (long long)(nz - 1) * n2 + c;
  double a_n = Ao[t], s_n = ts ? S_out[t] : 0.0, c_n = cst[t];
  for (int k = nz - 1; k >= 0; --k) {
    const double a_k = a_n, s_k = s_n, c_k = c_n;
    const double mm = wet(k) ? 1.0 : 0.0;
    const long long tk = t;
    if (k > 0) {
      t -= n2;
      a_n = Ao[t];
      if (ts) s_n = S_out[t];
      c_n = cst[t];
    }
    xa = (k == nz - 1) ? a_k : a_k - c_k * xa;
    Ao[tk] = xa * mm;
    if (ts) {
      xb = (k == nz - 1) ? s_k : s_k - c_k * xb;
      S_out[tk] = xb * mm;
      const double z_k = cg_zc(geo.zs, dzr, nwc, hbc, k);
      if (p.eos_mixed)
        b_out[tk] = buoyancy_mixed(p, xa * mm, xb * mm, z_k) * mm;
      else
        b_out[tk] = buoyancy(p, xa * mm, xb * mm, z_k) * mm;
    }
  }
}

}  // namespace hoc

namespace hoc {

// Staged tile of one level: the advected velocities and w in the flux type
// R, the pressure-gradient inputs in fp64.
template <class R>
struct MomLds {
  R u[SY][SX], v[SY][SX];               // velocities, 2-cell halo
  uint8_t mu[SY][SX], mv[SY][SX];       // mask3u, mask3v, 2-cell halo
};

// What only the cell pass reads (one buffer: level k + 1 is stored after the
// barrier that ends level k and before the face pass of level k + 1, whose
// closing barrier orders it before the cell pass).
template <class R>
struct MomPLds {
  double phi[TY + 1][TX + 1];           // potential, +1 east / north
  double b[TY + 1][TX + 1];
  double zc[TY + 1][TX + 1];
  R w[TY + 1][TX + 1];                  // w at interface k + 1
};

// Momentum face fluxes of one level, each computed once: advective fluxes
// (advecting velocity times the face value) and the A_h gradients, on the
// x-faces (index q: between tile columns q-1 and q) and y-faces (between
// rows q-1 and q) of the u and v points.
template <class R>
struct MomFaces {
  R fux[TY][TX + 1], fvx[TY][TX + 1], gux[TY][TX + 1], gvx[TY][TX + 1];
  R fuy[TY + 1][TX], fvy[TY + 1][TX], guy[TY + 1][TX], gvy[TY + 1][TX];
};

constexpr int MS_H = (SX * SY + TILE_THREADS - 1) / TILE_THREADS;              // halo slots
constexpr int MS_P = ((TY + 1) * (TX + 1) + TILE_THREADS - 1) / TILE_THREADS;  // +1 slots

template <class R>
struct MomRegs {
  R u[MS_H], v[MS_H];
  uint8_t mu[MS_H], mv[MS_H];
  double phi[MS_P], b[MS_P], zc[MS_P];
  R w[MS_P];
};

struct MomOffs {
  int h[MS_H], p[MS_P];
};

__device__ __forceinline__ MomOffs mom_offsets(int tid, int i0, int j0, int nx, int ny) {
  MomOffs F;
#pragma unroll
  for (int q = 0; q < MS_H; ++q) {
    const int t = tid + q * TILE_THREADS;
    F.h[q] = t < SX * SY ? pmod(j0 + t / SX - HALO, ny) * nx + pmod(i0 + t % SX - HALO, nx) : -1;
  }
#pragma unroll
  for (int q = 0; q < MS_P; ++q) {
    const int t = tid + q * TILE_THREADS;
    F.p[q] = t < (TY + 1) * (TX + 1)
                 ? pmod(j0 + t / (TX + 1), ny) * nx + pmod(i0 + t % (TX + 1), nx)
                 : -1;
  }
  return F;
}

template <class RT>
__device__ __forceinline__ void mom_load(MomRegs<RT>& R, const MomOffs& F, long long o, long long n2,
                                         const double* u, const double* v, const double* phi,
                                         const double* b, const double* w, const Geo& geo) {
#pragma unroll
  for (int q = 0; q < MS_H; ++q) {
    if (F.h[q] >= 0) {
      const long long g = o + F.h[q];
      R.u[q] = (RT)u[g];
      R.v[q] = (RT)v[g];
      R.mu[q] = geo.m3u[g];
      R.mv[q] = geo.m3v[g];
    }
  }
#pragma unroll
  for (int q = 0; q < MS_P; ++q) {
    if (F.p[q] >= 0) {
      const long long g = o + F.p[q];
      R.phi[q] = phi[g];
      R.b[q] = b[g];
      R.zc[q] = geo.zc[g];
      R.w[q] = (RT)w[g + n2];
    }
  }
}

template <class RT>
__device__ __forceinline__ void mom_store(MomLds<RT>& L, const MomRegs<RT>& R, int tid) {
#pragma unroll
  for (int q = 0; q < MS_H; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < SX * SY) {
      const int lx = t % SX, ly = t / SX;
      L.u[ly][lx] = R.u[q];
      L.v[ly][lx] = R.v[q];
      L.mu[ly][lx] = R.mu[q];
      L.mv[ly][lx] = R.mv[q];
    }
  }
}

template <class RT>
__device__ __forceinline__ void mom_store_p(MomPLds<RT>& P, const MomRegs<RT>& R, int tid) {
#pragma unroll
  for (int q = 0; q < MS_P; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < (TY + 1) * (TX + 1)) {
      const int lx = t % (TX + 1), ly = t / (TX + 1);
      P.phi[ly][lx] = R.phi[q];
      P.b[ly][lx] = R.b[q];
      P.zc[ly][lx] = R.zc[q];
      P.w[ly][lx] = R.w[q];
    }
  }
}

// Column-form staging: the velocities with halo and the pressure-gradient
// inputs are loaded; the level's face masks (with halo) and centre depths
// are formed from the tile's column geometry into the same LDS layout the
// field form loads them into.
template <class R>
struct MomRegsC {
  R u[MS_H], v[MS_H];
  double phi[MS_P], b[MS_P];
  R w[MS_P];
};

template <class RT>
__device__ __forceinline__ void mom_load_c(MomRegsC<RT>& R, const MomOffs& F, long long o,
                                           long long n2, const double* u, const double* v,
                                           const double* phi, const double* b, const double* w) {
#pragma unroll
  for (int q = 0; q < MS_H; ++q) {
    if (F.h[q] >= 0) {
      const long long g = o + F.h[q];
      R.u[q] = (RT)u[g];
      R.v[q] = (RT)v[g];
    }
  }
#pragma unroll
  for (int q = 0; q < MS_P; ++q) {
    if (F.p[q] >= 0) {
      const long long g = o + F.p[q];
      R.phi[q] = phi[g];
      R.b[q] = b[g];
      R.w[q] = (RT)w[g + n2];
    }
  }
}

template <class RT>
__device__ __forceinline__ void mom_store_c(MomLds<RT>& L, const MomRegsC<RT>& R,
                                            const ColTile& CG, int kk, int tid) {
#pragma unroll
  for (int q = 0; q < MS_H; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < SX * SY) {
      const int lx = t % SX, ly = t / SX;
      L.u[ly][lx] = R.u[q];
      L.v[ly][lx] = R.v[q];
      L.mu[ly][lx] = col_wet_u(CG, ly, lx, kk);
      L.mv[ly][lx] = col_wet_v(CG, ly, lx, kk);
    }
  }
}

template <class RT>
__device__ __forceinline__ void mom_store_p_c(MomPLds<RT>& P, const MomRegsC<RT>& R,
                                              const ColTile& CG, const double* zs, double dzr,
                                              int kk, int tid) {
#pragma unroll
  for (int q = 0; q < MS_P; ++q) {
    const int t = tid + q * TILE_THREADS;
    if (t < (TY + 1) * (TX + 1)) {
      const int lx = t % (TX + 1), ly = t / (TX + 1);
      P.phi[ly][lx] = R.phi[q];
      P.b[ly][lx] = R.b[q];
      P.zc[ly][lx] = cg_zc(zs, dzr, CG.nw[ly + HALO][lx + HALO], CG.hb[ly + HALO][lx + HALO], kk);
      P.w[ly][lx] = R.w[q];
    }
  }
}

struct NoTile {};

// Vertical momentum flux on interface kk of the column: wb times the
// interface value of the advected velocity (advection_v06.py:95-99,
// advection.py:75-79); f* hold the velocity at kk-2..kk+1, m* the wet flags
// of kk-2 and kk+1.
template <class R>
__device__ __forceinline__ R mom_flux_z(int adv, int kk, int nz, R wb, R far_up, R above, R below,
                                        R far_dn, bool mup, bool mdn) {
  if (kk <= 0 || kk >= nz) return R(0);
  if (adv == ADV_CENTRED2) return wb * R(0.5) * (above + below);
  const bool hu = kk >= 2 && mup;
  const bool hd = kk <= nz - 2 && mdn;
  return wb * kappa_face_z<R>(kk >= 2 ? far_up : R(0), above, below, kk <= nz - 2 ? far_dn : R(0),
                              wb, hu, hd, adv == ADV_UP3_TVD);
}

// The step's explicit momentum terms (libs/core/model3d_v05.py:352-370),
// tile by tile and level by level; same outputs as k_momentum_explicit.
// Staging is double-buffered (level k + 1 loads in flight while level k is
// computed) and every horizontal face flux is formed once. R is the type of
// the advective fluxes and the A_h gradients (double: bitwise the
// reference; float: the mixed step); the pressure gradient, the Coriolis
// partners, the vertical diffusion and the tendencies stay fp64. ZC: the
// static z-level geometry in column form (masks, face thicknesses and
// centre depths formed from the tile's column geometry, the same bits).
template <class R, bool ZC>
__global__ __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(ZC ? 3 : 1))) void k_momentum_explicit_tiled(
    Dims d, StepParams p, Geo geo, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ b, const double* __restrict__ phi, const double* __restrict__ w,
    const double* __restrict__ Km, double* __restrict__ pgx, double* __restrict__ pgy,
    double* __restrict__ avu, double* __restrict__ avv, double* __restrict__ dzu,
    double* __restrict__ dzv, double* __restrict__ stu, double* __restrict__ stv) {
  __shared__ MomLds<R> LB[2];
  __shared__ MomPLds<R> P;
  __shared__ MomFaces<R> F;
  __shared__ std::conditional_t<ZC, ColTile, NoTile> CG;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long n2 = d.n2;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * TX, j0 = by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool active = i < nx && j < ny;
  const long long c = active ? (long long)j * nx + i : 0;
  int adv = p.advection;
  if (adv == ADV_UPWIND1) adv = ADV_CENTRED2;  // momentum keeps the centred form
  const bool lim = adv == ADV_UP3_TVD;

  double km_k = active ? Km[c] : 0.0;
  R fzu_top = R(0), fzv_top = R(0);
  const int cx = tx + HALO, cy = ty + HALO;

  std::conditional_t<ZC, MomRegsC<R>, MomRegs<R>> RG;
  const MomOffs MO = mom_offsets(tid, i0, j0, nx, ny);
  if constexpr (ZC) {
    mom_load_c(RG, MO, 0, n2, u, v, phi, b, w);
    col_tile_load(CG, geo, tid, i0, j0, nx, ny);
    __syncthreads();
    mom_store_c(LB[0], RG, CG, 0, tid);
  } else {
    mom_load(RG, MO, 0, n2, u, v, phi, b, w, geo);
    mom_store(LB[0], RG, tid);
  }
  __syncthreads();
  // own-column windows at k-2..k+2: u, v, mask3u, mask3v; and k-1..k+1 of
  // dz3u, dz3v, mask3
  double uw5[5], vw5[5];
  bool muw[5], mvw[5];
  double duw[3], dvw[3];
  bool m3w[3];
  // own-column geometry at level kk (0 <= kk < nz)
  auto own_mu = [&](int kk) -> bool {
    if constexpr (ZC) return col_wet_u(CG, ty + HALO, tx + HALO, kk);
    else return geo.m3u[kk * n2 + c] != 0;
  };
  auto own_mv = [&](int kk) -> bool {
    if constexpr (ZC) return col_wet_v(CG, ty + HALO, tx + HALO, kk);
    else return geo.m3v[kk * n2 + c] != 0;
  };
  auto own_m3 = [&](int kk) -> bool {
    if constexpr (ZC) return kk < CG.nw[ty + HALO][tx + HALO];
    else return geo.m3[kk * n2 + c] != 0;
  };
  auto own_du = [&](int kk) -> double {
    if constexpr (ZC) {
      const int yy = ty + HALO, xx = tx + HALO;
      return cg_face(p.face_rule, p.dzr, CG.mu[yy][xx] != 0, CG.nw[yy][xx], CG.hb[yy][xx],
                     CG.nw[yy][xx + 1], CG.hb[yy][xx + 1], kk);
    } else {
      return geo.dz3u[kk * n2 + c];
    }
  };
  auto own_dv = [&](int kk) -> double {
    if constexpr (ZC) {
      const int yy = ty + HALO, xx = tx + HALO;
      return cg_face(p.face_rule, p.dzr, CG.mv[yy][xx] != 0, CG.nw[yy][xx], CG.hb[yy][xx],
                     CG.nw[yy + 1][xx], CG.hb[yy + 1][xx], kk);
    } else {
      return geo.dz3v[kk * n2 + c];
    }
  };
  for (int q = 0; q < 5; ++q) {
    uw5[q] = vw5[q] = 0.0;
    muw[q] = mvw[q] = false;
  }
  for (int q = 0; q < 3; ++q) {
    duw[q] = dvw[q] = 0.0;
    m3w[q] = false;
  }
  if (active) {
    for (int q = 2; q < 5; ++q) {
      const int kk = q - 2;
      if (kk < nz) {
        uw5[q] = u[kk * n2 + c];
        vw5[q] = v[kk * n2 + c];
        muw[q] = own_mu(kk);
        mvw[q] = own_mv(kk);
      }
    }
    for (int q = 1; q < 3; ++q) {
      const int kk = q - 1;
      if (kk < nz) {
        duw[q] = own_du(kk);
        dvw[q] = own_dv(kk);
        m3w[q] = own_m3(kk);
      }
    }
  }
  for (int k = 0; k < nz; ++k) {
    const long long o = (long long)k * n2;
    const MomLds<R>& L = LB[k & 1];
    // level k's cell-pass inputs, loaded during level k - 1 (before the
    // loads of level k + 1 reuse the registers)
    if constexpr (ZC) mom_store_p_c(P, RG, CG, geo.zs, p.dzr, k, tid);
    else mom_store_p(P, RG, tid);
    if (k + 1 < nz) {
      if constexpr (ZC) mom_load_c(RG, MO, o + n2, n2, u, v, phi, b, w);
      else mom_load(RG, MO, o + n2, n2, u, v, phi, b, w, geo);
    }
    // level-k wet flags of the east (u) and north (v) faces of tile column
    // (ly, lx), and the centre depth of tile point (ly, lx) of the +1 tile
    auto MUf = [&](int ly, int lx) -> bool { return L.mu[ly][lx] != 0; };
    auto MVf = [&](int ly, int lx) -> bool { return L.mv[ly][lx] != 0; };
    auto ZCf = [&](int ly, int lx) -> double { return P.zc[ly][lx]; };
    // the own column's viscosity below level k, issued before the face pass
    const double km_k1 = active ? Km[o + n2 + c] : 0.0;
    {
      const R(*U)[SX] = L.u;
      const R(*V)[SX] = L.v;
      // x-face q of row fy_ (between columns q-1 and q)
      auto xface = [&](int fy_, int q) {
        const int yy = fy_ + HALO, x0 = q - 1 + HALO;  // x0: the column west of the face
        if (adv != ADV_NONE) {
          const R ub = R(0.5) * (U[yy][x0] + U[yy][x0 + 1]);
          const R uu = R(0.5) * (U[yy][x0] + U[yy + 1][x0]);
          if (adv == ADV_CENTRED2) {
            F.fux[fy_][q] = ub * ub;
            F.fvx[fy_][q] = uu * (R(0.5) * (V[yy][x0] + V[yy][x0 + 1]));
          } else {
            F.fux[fy_][q] = ub * kappa_face<R>(U[yy][x0 - 1], U[yy][x0], U[yy][x0 + 1],
                                               U[yy][x0 + 2], ub, MUf(yy, x0 - 1),
                                               MUf(yy, x0 + 2), lim);
            F.fvx[fy_][q] = uu * kappa_face<R>(V[yy][x0 - 1], V[yy][x0], V[yy][x0 + 1],
                                               V[yy][x0 + 2], uu, MVf(yy, x0 - 1),
                                               MVf(yy, x0 + 2), lim);
          }
        }
        const R m = MUf(yy, x0) ? R(1) : R(0);
        F.gux[fy_][q] = m * (U[yy][x0 + 1] - U[yy][x0]) / R(p.dx);
        F.gvx[fy_][q] = m * (V[yy][x0 + 1] - V[yy][x0]) / R(p.dx);
      };
      // y-face q of column fx_ (between rows q-1 and q)
      auto yface = [&](int q, int fx_) {
        const int xx = fx_ + HALO, y0 = q - 1 + HALO;  // y0: the row south of the face
        if (adv != ADV_NONE) {
          const R vb = R(0.5) * (V[y0][xx] + V[y0][xx + 1]);
          const R vv = R(0.5) * (V[y0][xx] + V[y0 + 1][xx]);
          if (adv == ADV_CENTRED2) {
            F.fuy[q][fx_] = vb * (R(0.5) * (U[y0][xx] + U[y0 + 1][xx]));
            F.fvy[q][fx_] = vv * vv;
          } else {
            F.fuy[q][fx_] = vb * kappa_face<R>(U[y0 - 1][xx], U[y0][xx], U[y0 + 1][xx],
                                               U[y0 + 2][xx], vb, MUf(y0 - 1, xx),
                                               MUf(y0 + 2, xx), lim);
            F.fvy[q][fx_] = vv * kappa_face<R>(V[y0 - 1][xx], V[y0][xx], V[y0 + 1][xx],
                                               V[y0 + 2][xx], vv, MVf(y0 - 1, xx),
                                               MVf(y0 + 2, xx), lim);
          }
        }
        const R m = MVf(y0, xx) ? R(1) : R(0);
        F.guy[q][fx_] = m * (U[y0 + 1][xx] - U[y0][xx]) / R(p.dy);
        F.gvy[q][fx_] = m * (V[y0 + 1][xx] - V[y0][xx]) / R(p.dy);
      };
      xface(ty, tx + 1);
      yface(ty + 1, tx);
      if (tid < TY) xface(tid, 0);
      else if (tid < TY + TX) yface(0, tid - TY);
    }
    __syncthreads();
    if (active) {
      const long long t = o + c;
      const double mu_c = MUf(cy, cx) ? 1.0 : 0.0, mv_c = MVf(cy, cx) ? 1.0 : 0.0;
      // pressure gradient at the face depth (model3d_v05.py:195-230)
      double gx, gy;
      if (p.pgf_correction) {
        const double zc = ZCf(ty, tx);
        {
          const double zr = ZCf(ty, tx + 1);
          const double zf = 0.5 * (zc + zr);
          const double left = P.phi[ty][tx] + P.b[ty][tx] * (zf - zc);
          const double right = P.phi[ty][tx + 1] + P.b[ty][tx + 1] * (zf - zr);
          gx = mu_c * (right - left) / p.dx;
        }
        {
          const double zr = ZCf(ty + 1, tx);
          const double zf = 0.5 * (zc + zr);
          const double left = P.phi[ty][tx] + P.b[ty][tx] * (zf - zc);
          const double right = P.phi[ty + 1][tx] + P.b[ty + 1][tx] * (zf - zr);
          gy = mv_c * (right - left) / p.dy;
        }
      } else {
        gx = mu_c * (P.phi[ty][tx + 1] - P.phi[ty][tx]) / p.dx * mu_c;
        gy = mv_c * (P.phi[ty + 1][tx] - P.phi[ty][tx]) / p.dy * mv_c;
      }
      pgx[t] = gx;
      pgy[t] = gy;
      // Coriolis partners (operators_masked.py:45-61), fp64: from the tile
      // in the plain step, from the fp64 fields (neighbours in cache) when
      // the tile holds fp32 velocities
      double vc0, vs, ve, vse, uc0, uw, un, unw;
      if constexpr (sizeof(R) == sizeof(double)) {
        vc0 = L.v[cy][cx], vs = L.v[cy - 1][cx], ve = L.v[cy][cx + 1], vse = L.v[cy - 1][cx + 1];
        uc0 = L.u[cy][cx], uw = L.u[cy][cx - 1], un = L.u[cy + 1][cx], unw = L.u[cy + 1][cx - 1];
      } else {
        const long long rj = (long long)j * nx, rs = (long long)pmod(j - 1, ny) * nx,
                        rn = (long long)pmod(j + 1, ny) * nx;
        const int ie = pmod(i + 1, nx), iw = pmod(i - 1, nx);
        vc0 = vw5[2], vs = v[o + rs + i], ve = v[o + rj + ie], vse = v[o + rs + ie];
        uc0 = uw5[2], uw = u[o + rj + iw], un = u[o + rn + i], unw = u[o + rn + iw];
      }
      if (avv) {  // the Coriolis partners, for callers that ask for them
        const double w0 = mv_c, w1 = MVf(cy - 1, cx) ? 1.0 : 0.0,
                     w2 = MVf(cy, cx + 1) ? 1.0 : 0.0, w3 = MVf(cy - 1, cx + 1) ? 1.0 : 0.0;
        const double num = 0.0 + w0 * vc0 + w1 * vs + w2 * ve + w3 * vse;
        const double den = 0.0 + w0 + w1 + w2 + w3;
        avv[t] = (den > 0.0 ? num / den : 0.0) * mu_c;
      }
      if (avu) {
        const double w0 = mu_c, w1 = MUf(cy, cx - 1) ? 1.0 : 0.0,
                     w2 = MUf(cy + 1, cx) ? 1.0 : 0.0, w3 = MUf(cy + 1, cx - 1) ? 1.0 : 0.0;
        const double num = 0.0 + w0 * uc0 + w1 * uw + w2 * un + w3 * unw;
        const double den = 0.0 + w0 + w1 + w2 + w3;
        avu[t] = (den > 0.0 ? num / den : 0.0) * mv_c;
      }
      // explicit vertical diffusion with surface stress and bottom drag (for
      // callers that ask for it: the column-form predictor forms it itself)
      if (dzu)
        dzu[t] = explicit_diffusion_w(k, nz, uw5[1], uw5[2], uw5[3], duw[0], duw[1], duw[2],
                                      m3w[0], m3w[1], m3w[2], km_k, km_k1,
                                      k == 0 ? geo.tau_u[c] : 0.0, p.bottom_drag);
      if (dzv)
        dzv[t] = explicit_diffusion_w(k, nz, vw5[1], vw5[2], vw5[3], dvw[0], dvw[1], dvw[2],
                                      m3w[0], m3w[1], m3w[2], km_k, km_k1,
                                      k == 0 ? geo.tau_v[c] : 0.0, p.bottom_drag);
      // advection and horizontal viscosity
      double exu = 0.0, exv = 0.0;
      if (adv != ADV_NONE) {
        const R ax = (F.fux[ty][tx + 1] - F.fux[ty][tx]) / R(p.dx);
        const R ay = (F.fuy[ty + 1][tx] - F.fuy[ty][tx]) / R(p.dy);
        const R by_ = (F.fvy[ty + 1][tx] - F.fvy[ty][tx]) / R(p.dy);
        const R bx_ = (F.fvx[ty][tx + 1] - F.fvx[ty][tx]) / R(p.dx);
        // vertical fluxes: the top one is the bottom one of the level above
        const R wbu = R(0.5) * (P.w[ty][tx] + P.w[ty][tx + 1]);
        const R wbv = R(0.5) * (P.w[ty][tx] + P.w[ty + 1][tx]);
        const R fzu_bot = mom_flux_z<R>(adv, k + 1, nz, wbu, R(uw5[1]), R(uw5[2]), R(uw5[3]),
                                        R(uw5[4]), muw[1], muw[4]);
        const R fzv_bot = mom_flux_z<R>(adv, k + 1, nz, wbv, R(vw5[1]), R(vw5[2]), R(vw5[3]),
                                        R(vw5[4]), mvw[1], mvw[4]);
        const R az = (fzu_top - fzu_bot) / R(p.dz_nom);
        const R bz = (fzv_top - fzv_bot) / R(p.dz_nom);
        fzu_top = fzu_bot;
        fzv_top = fzv_bot;
        exu = exu - (double)((ax + ay + az) * R(mu_c));
        exv = exv - (double)((bx_ + by_ + bz) * R(mv_c));
      }
      if (p.has_A_h) {
        exu = exu + p.A_h * (double)(R(mu_c) * ((F.gux[ty][tx + 1] - F.gux[ty][tx]) / R(p.dx) +
                                                (F.guy[ty + 1][tx] - F.guy[ty][tx]) / R(p.dy)));
        exv = exv + p.A_h * (double)(R(mv_c) * ((F.gvx[ty][tx + 1] - F.gvx[ty][tx]) / R(p.dx) +
                                                (F.gvy[ty + 1][tx] - F.gvy[ty][tx]) / R(p.dy)));
      }
      double s0u = 0.0, s0v = 0.0;
      if (k == 0) {
        s0u = p.thv_dt * geo.tau_u[c] * safe_inv(duw[1]);
        s0v = p.thv_dt * geo.tau_v[c] * safe_inv(dvw[1]);
      }
      stu[t] = s0u + p.dt * exu;
      stv[t] = s0v + p.dt * exv;
      km_k = km_k1;
    }
    // advance the windows
    for (int q = 0; q < 4; ++q) {
      uw5[q] = uw5[q + 1];
      vw5[q] = vw5[q + 1];
      muw[q] = muw[q + 1];
      mvw[q] = mvw[q + 1];
    }
    for (int q = 0; q < 2; ++q) {
      duw[q] = duw[q + 1];
      dvw[q] = dvw[q + 1];
      m3w[q] = m3w[q + 1];
    }
    if (active) {
      const int k3 = k + 3, k2 = k + 2;
      uw5[4] = k3 < nz ? u[(long long)k3 * n2 + c] : 0.0;
      vw5[4] = k3 < nz ? v[(long long)k3 * n2 + c] : 0.0;
      muw[4] = k3 < nz && own_mu(k3);
      mvw[4] = k3 < nz && own_mv(k3);
      duw[2] = k2 < nz ? own_du(k2) : 0.0;
      dvw[2] = k2 < nz ? own_dv(k2) : 0.0;
      m3w[2] = k2 < nz && own_m3(k2);
    }
    if (k + 1 < nz) {
      if constexpr (ZC) mom_store_c(LB[(k + 1) & 1], RG, CG, k + 1, tid);
      else mom_store(LB[(k + 1) & 1], RG, tid);
    }
    __syncthreads();
  }
}

void launch_tracers_tiled(bool mixed, hipStream_t s, Dims d, StepParams p, Geo geo,
                          const double* T, const double* S, const double* b, const double* u,
                          const double* v, const double* w, const double* Kh, double* cst,
                          double* T_out, double* S_out, double* b_out) {
  const dim3 grid((d.nx + TX - 1) / TX, (d.ny + TY - 1) / TY);
  if (p.zcol) {
    if (mixed)
      k_tracers_col<float><<<grid, TILE_THREADS, 0, s>>>(d, p, geo, T, S, b, u, v, w, Kh, cst,
                                                         T_out, S_out, b_out);
    else
      k_tracers_col<double><<<grid, TILE_THREADS, 0, s>>>(d, p, geo, T, S, b, u, v, w, Kh, cst,
                                                          T_out, S_out, b_out);
  } else if (mixed) {
    k_tracers_tiled<float><<<grid, TILE_THREADS, 0, s>>>(d, p, geo, T, S, b, u, v, w, Kh, cst,
                                                         T_out, S_out, b_out);
  } else {
    k_tracers_tiled<double><<<grid, TILE_THREADS, 0, s>>>(d, p, geo, T, S, b, u, v, w, Kh, cst,
                                                          T_out, S_out, b_out);
  }
}

void launch_momentum_explicit_tiled(bool mixed, hipStream_t s, Dims d, StepParams p, Geo geo,
                                    const double* u, const double* v, const double* b,
                                    const double* phi, const double* w, const double* Km,
                                    double* pgx, double* pgy, double* avu, double* avv,
                                    double* dzu, double* dzv, double* stu, double* stv) {
  const dim3 grid((d.nx + TX - 1) / TX, (d.ny + TY - 1) / TY);
  if (p.zcol) {
    if (mixed)
      k_momentum_explicit_tiled<float, true><<<grid, TILE_THREADS, 0, s>>>(
          d, p, geo, u, v, b, phi, w, Km, pgx, pgy, avu, avv, dzu, dzv, stu, stv);
    else
      k_momentum_explicit_tiled<double, true><<<grid, TILE_THREADS, 0, s>>>(
          d, p, geo, u, v, b, phi, w, Km, pgx, pgy, avu, avv, dzu, dzv, stu, stv);
  } else if (mixed) {
    k_momentum_explicit_tiled<float, false><<<grid, TILE_THREADS, 0, s>>>(
        d, p, geo, u, v, b, phi, w, Km, pgx, pgy, avu, avv, dzu, dzv, stu, stv);
  } else {
    k_momentum_explicit_tiled<double, false><<<grid, TILE_THREADS, 0, s>>>(
        d, p, geo, u, v, b, phi, w, Km, pgx, pgy, avu, avv, dzu, dzv, stu, stv);
  }
}

}  // namespace hoc
