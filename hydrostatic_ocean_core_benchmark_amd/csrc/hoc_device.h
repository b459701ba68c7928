// Device-side helpers shared by every kernel of the ocean-core step.
//
// Layout: fp64, C order. A 3D layer field is [k][j][i] (nz x ny x nx), an
// interface field [k][j][i] with nz+1 levels, a 2D field [j][i]. i is the
// fastest index, so one lane per (j, i) column gives coalesced access at
// every level k. Masks are uint8 (0/1); multiplying by (double)m is bitwise
// the same as multiplying by the reference's 0.0/1.0 float masks.
//
// Every arithmetic expression keeps the operand order of the reference's
// NumPy code and the library is built with -ffp-contract=off, so no
// multiply-add is fused and results are bitwise those of the oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hoc_params.h"

namespace hoc {

__device__ __forceinline__ long long at3(const Dims& d, int k, int j, int i) {
  return ((long long)k * d.ny + j) * d.nx + i;
}
__device__ __forceinline__ long long at2(const Dims& d, int j, int i) {
  return (long long)j * d.nx + i;
}
// Periodic neighbour index, |s| <= n (every shift of the step is within 2).
__device__ __forceinline__ int wrap(int i, int s, int n) {
  int t = i + s;
  if (t < 0) t += n;
  if (t >= n) t -= n;
  return t;
}

// NumPy's maximum/minimum: NaN in either operand propagates.
__device__ __forceinline__ double nmax(double a, double b) {
  return (a >= b || a != a) ? a : b;
}
__device__ __forceinline__ double nmin(double a, double b) {
  return (a <= b || a != a) ? a : b;
}
__device__ __forceinline__ float nmax(float a, float b) {
  return (a >= b || a != a) ? a : b;
}
__device__ __forceinline__ float nmin(float a, float b) {
  return (a <= b || a != a) ? a : b;
}
// 1/x where x > 0, else 0 (the reference's `_safe`).
__device__ __forceinline__ double safe_inv(double x) {
  return x > 0.0 ? 1.0 / x : 0.0;
}
__device__ __forceinline__ double mk(uint8_t m) { return m ? 1.0 : 0.0; }

// Face thickness rule of the domain (libs/core/domain.py:98-110): 0 min,
// 1 mean, 2 harmonic.
__device__ __forceinline__ double face_value(int rule, double a, double b) {
  if (rule == 0) return nmin(a, b);
  if (rule == 1) return 0.5 * (a + b);
  const double s = a + b;
  return s > 0.0 ? 2.0 * a * b / s : 0.0;
}

// ------------------------------------- static z-level geometry by column
// A z-level column is fixed by its number of wet levels nw and the
// thickness hb of its deepest wet cell: levels fill from the surface with
// the uniform reference thickness dzr, the deepest wet level may be partial
// and every level below it is dry (libs/core/domain.py:114-131). The host
// derives (nw, hb) from the full arrays and checks that every thickness,
// face thickness, mask and centre depth below equals the stored value bit
// for bit before the kernels use this form; the kernels then read one byte
// and one double per column instead of five fp64 fields and three masks per
// cell. zs[k] is the sequential sum of k reference thicknesses (np.cumsum's
// order).
__device__ __forceinline__ double cg_dz(double dzr, int nw, double hb, int k) {
  return k < nw - 1 ? dzr : (k == nw - 1 ? hb : 0.0);
}
// Depth of interface k = 0..nz below the surface: the thicknesses above it
// summed from the top in order (exact zeros add nothing below the bottom).
__device__ __forceinline__ double cg_above(const double* zs, int nw, double hb, int k) {
  if (k < nw) return zs[k];
  return nw > 0 ? zs[nw - 1] + hb : 0.0;
}
// Depth of the centre of level k below the surface: the thickness above it
// summed from the top, plus half its own (libs/core/domain.py:141-145).
__device__ __forceinline__ double cg_zc(const double* zs, double dzr, int nw, double hb, int k) {
  if (k < nw) return zs[k] + 0.5 * cg_dz(dzr, nw, hb, k);
  return nw > 0 ? zs[nw - 1] + hb : 0.0;
}
// Centre spacing on interface k = 0..nz (libs/core/vertical_var.py:21-32).
__device__ __forceinline__ double cg_dzi(double dzr, int nw, double hb, int k, int nz) {
  if (k == 0) return cg_dz(dzr, nw, hb, 0);
  if (k == nz) return cg_dz(dzr, nw, hb, nz - 1);
  return 0.5 * (cg_dz(dzr, nw, hb, k - 1) + cg_dz(dzr, nw, hb, k));
}
// Thickness of the level-k face between a column (nw0, hb0) and its east or
// north neighbour (nw1, hb1): the face rule where the face is wet (both
// cells wet and the 2D face mask open), else 0 (mask3u * face, domain.py:138).
__device__ __forceinline__ double cg_face(int rule, double dzr, bool open, int nw0, double hb0,
                                          int nw1, double hb1, int k) {
  if (!(open && k < nw0 && k < nw1)) return 0.0;
  return face_value(rule, cg_dz(dzr, nw0, hb0, k), cg_dz(dzr, nw1, hb1, k));
}

// Vertical geometry of one column, from the stored arrays or, when the
// geometry is in column form (StepParams::zcol, a uniform branch), from the
// column's (nw, hb): the same bits either way.
struct ColumnGeo {
  const Geo& g;
  long long c, n2;
  int nz, nw;
  bool col;
  double hb, dzr;
  __device__ __forceinline__ ColumnGeo(const Geo& geo, const StepParams& p, const Dims& d,
                                       long long col_index)
      : g(geo), c(col_index), n2(d.n2), nz(d.nz), nw(0), col(p.zcol != 0), hb(0.0),
        dzr(p.dzr) {
    if (col) {
      nw = geo.nw[c];
      hb = geo.hb[c];
    }
  }
  __device__ __forceinline__ double dz(int k) const {
    return col ? cg_dz(dzr, nw, hb, k) : g.dz3[(long long)k * n2 + c];
  }
  __device__ __forceinline__ double dzi(int k) const {  // k = 0..nz
    return col ? cg_dzi(dzr, nw, hb, k, nz) : g.dzi[(long long)k * n2 + c];
  }
  __device__ __forceinline__ bool wet(int k) const {  // k = 0..nz-1
    return col ? k < nw : g.m3[(long long)k * n2 + c] != 0;
  }
  __device__ __forceinline__ double zc(int k) const {
    return col ? cg_zc(g.zs, dzr, nw, hb, k) : g.zc[(long long)k * n2 + c];
  }
};

// Face thickness and wet flag of the east (axis 0) or north (axis 1) face
// of one column at any level, from the stored arrays or the column form.
struct FaceGeo {
  const double* dzf;
  const uint8_t* mf;
  long long c, n2;
  bool col, open;
  int rule, nw0, nw1;
  double hb0, hb1, dzr;
  __device__ __forceinline__ FaceGeo(const Geo& geo, const StepParams& p, const Dims& d,
                                     long long c0, long long c1, int axis)
      : dzf(axis == 0 ? geo.dz3u : geo.dz3v), mf(axis == 0 ? geo.m3u : geo.m3v), c(c0),
        n2(d.n2), col(p.zcol != 0), open(false), rule(p.face_rule), nw0(0), nw1(0), hb0(0.0),
        hb1(0.0), dzr(p.dzr) {
    if (col) {
      open = (axis == 0 ? geo.mu2[c0] : geo.mv2[c0]) != 0;
      nw0 = geo.nw[c0];
      nw1 = geo.nw[c1];
      hb0 = geo.hb[c0];
      hb1 = geo.hb[c1];
    }
  }
  __device__ __forceinline__ double dz(int k) const {
    return col ? cg_face(rule, dzr, open, nw0, hb0, nw1, hb1, k) : dzf[(long long)k * n2 + c];
  }
  __device__ __forceinline__ bool wet(int k) const {
    return col ? (open && k < nw0 && k < nw1) : mf[(long long)k * n2 + c] != 0;
  }
};

// Tile of this workgroup in a 2D grid, remapped so that each of the 8 XCDs
// (workgroups are dispatched to them round-robin by linear id) owns one
// contiguous band of tile rows: neighbouring tiles then share their halo
// rows in the same L2. A bijection for any grid size.
__device__ __forceinline__ void xcd_tile(int& bx, int& by) {
  const int gx = (int)gridDim.x;
  const int nb = gx * (int)gridDim.y;
  const int id = (int)(blockIdx.y * gridDim.x + blockIdx.x);
  const int x = id % 8, q = nb / 8, r = nb % 8;
  const int L = x * q + (x < r ? x : r) + id / 8;
  bx = L % gx;
  by = L / gx;
}

// Superbee limiter phi(r) = max(0, max(min(2r, 1), min(r, 2))).
// F is the arithmetic type (double: the reference's; float: the mixed
// step's advective fluxes).
template <class F>
__device__ __forceinline__ F superbee(F r) {
  return nmax(F(0), nmax(nmin(F(2) * r, F(1)), nmin(r, F(2))));
}

constexpr double KAPPA = 1.0 / 3.0;
constexpr double K1M = 1.0 - KAPPA;
constexpr double K1P = 1.0 + KAPPA;

// kappa = 1/3 face value between c0 = c[i] and c1 = c[i+1] for the face
// velocity vel; cm = c[i-1], c2 = c[i+2]; wm / w2 tell whether the far
// upstream cell (i-1 for vel > 0, i+2 for vel < 0) is wet, else upwind1.
// (libs/core/advection_v06.py:21-47). The reference forms both upwind
// candidates and selects one by the sign of vel; only the selected one is
// formed here, with its operands chosen first (one division per face):
// c1 - t is c1 + (-t) in IEEE arithmetic, so the result is the same bits.
template <class F>
__device__ __forceinline__ F kappa_face(F cm, F c0, F c1, F c2, F vel, bool wm, bool w2,
                                        bool limited) {
  const bool up = vel > F(0);
  const F base = up ? c0 : c1;
  const F num = up ? (c0 - cm) : (c2 - c1);
  const F dc = c1 - c0;
  F t;
  if (limited) {
    const F r = dc == F(0) ? F(0) : num / dc;
    t = F(0.5) * superbee(r) * dc;
  } else {
    t = F(0.25) * (F(K1M) * num + F(K1P) * dc);
  }
  const F f = up ? base + t : base - t;
  return (up ? wm : w2) ? f : base;
}

// Interior interface value between above = c[k-1] and below = c[k], w > 0
// upward; far_up = c[k-2] (0 when k < 2), far_dn = c[k+1] (0 when k > nz-2);
// has_up / has_dn are the wet flags of those far cells.
// (libs/core/advection_v06.py:50-74); the selected candidate only, as above.
template <class F>
__device__ __forceinline__ F kappa_face_z(F far_up, F above, F below, F far_dn, F w,
                                          bool has_up, bool has_dn, bool limited) {
  const bool up = w > F(0);
  const F base = up ? below : above;
  const F num = up ? (below - far_dn) : (far_up - above);
  const F d = above - below;
  F t;
  if (limited) {
    const F r = d == F(0) ? F(0) : num / d;
    t = F(0.5) * superbee(r) * d;
  } else {
    t = F(0.25) * (F(K1M) * num + F(K1P) * d);
  }
  const F f = up ? base + t : base - t;
  return (up ? has_dn : has_up) ? f : base;
}

// Explicit vertical diffusion of a velocity at level k from its column
// values f0..f2 (levels k-1..k+1), thicknesses dz0..dz2 and cell masks
// m0..m2 (libs/core/vertical_var.py:80-106, mask3 branch).
__device__ __forceinline__ double explicit_diffusion_w(int k, int nz, double f0, double f1,
                                                       double f2, double dz0, double dz1,
                                                       double dz2, bool m0, bool m1, bool m2,
                                                       double nu_k, double nu_k1, double sflux,
                                                       double drag) {
  const bool bot_km1 = k >= 1 && m0 && !m1;
  const bool bot_k = m1 && !((k + 1 < nz) && m2);
  double ft;
  if (k == 0) {
    ft = sflux;
  } else {
    const double dzi = (k == nz) ? dz0 : 0.5 * (dz0 + dz1);
    ft = nu_k * (f0 - f1) * safe_inv(dzi);
    if (bot_km1) ft = 0.0;
  }
  double fb = 0.0;
  if (k + 1 < nz) {
    const double dzi = 0.5 * (dz1 + dz2);
    fb = nu_k1 * (f1 - f2) * safe_inv(dzi);
    if (bot_k) fb = 0.0;
  }
  const double inv = safe_inv(dz1);
  const double out = (ft - fb) * inv;
  const double dr = drag * f1 * (bot_k ? 1.0 : 0.0);
  return (out - dr * inv) * (m1 ? 1.0 : 0.0);
}

// Row k of one implicit vertical-diffusion matrix (I - theta_v dt D[nu]) of
// a column, factored for the Thomas sweep: sub, den = diag - sub c*_{k-1}
// and c* = sup / den (libs/core/vertical_var.py:35-77). inv_dz, idzi_t and
// idzi_b are 1/dz3 and 1/dzi above and below the cell.
__device__ __forceinline__ void factor_row(const StepParams& p, int k, double inv_dz,
                                           double nu_t, double nu_b, double idzi_t,
                                           double idzi_b, bool wet, bool below,
                                           double drag_term, double c_prev, double& sub_o,
                                           double& den_o, double& cst_o) {
  const double fac = p.thv_dt * inv_dz;
  const double a_top = fac * nu_t * idzi_t;
  const double a_bot = fac * nu_b * idzi_b;
  double sub = -a_top, sup = -a_bot;
  double diag = 1.0 + a_top + a_bot;
  if (k == 0) {
    diag = 1.0 + a_bot;
    sub = 0.0;
  }
  if (wet && !below) {
    diag = 1.0 + a_top + drag_term * inv_dz;
    sup = 0.0;
  }
  if (!wet) {
    diag = 1.0;
    sub = 0.0;
    sup = 0.0;
  }
  den_o = (k == 0) ? diag : diag - sub * c_prev;
  cst_o = sup / den_o;
  sub_o = sub;
}

// ------------------------------------------------------------------ EOS
// libs/core/eos.py:68-133; coefficients of Roquet et al. (2015).
template <class F>
__device__ __forceinline__ F rho_teos10(F CT, F SA, F z) {
  const F R00 = F(4.6494977072e+01), R01 = F(-5.2099962525e+00), R02 = F(2.2601900708e-01),
               R03 = F(6.4326772569e-02), R04 = F(1.5616995503e-02), R05 = F(-1.7243708991e-03);
  const F E000 = F(8.0189615746e+02), E100 = F(8.6672408165e+02), E200 = F(-1.7864682637e+03),
               E300 = F(2.0375295546e+03), E400 = F(-1.2849161071e+03), E500 = F(4.3227585684e+02),
               E600 = F(-6.0579916612e+01);
  const F E010 = F(2.6010145068e+01), E110 = F(-6.5281885265e+01), E210 = F(8.1770425108e+01),
               E310 = F(-5.6888046321e+01), E410 = F(1.7681814114e+01), E510 = F(-1.9193502195e+00);
  const F E020 = F(-3.7074170417e+01), E120 = F(6.1548258127e+01), E220 = F(-6.0362551501e+01),
               E320 = F(2.9130021253e+01), E420 = F(-5.4723692739e+00);
  const F E030 = F(2.1661789529e+01), E130 = F(-3.3449108469e+01), E230 = F(1.9717078466e+01),
               E330 = F(-3.1742946532e+00);
  const F E040 = F(-8.3627885467e+00), E140 = F(1.1311538584e+01), E240 = F(-5.3563304045e+00);
  const F E050 = F(5.4048723791e-01), E150 = F(4.5111434961e-01);
  const F E060 = F(-1.9098268277e-01);
  const F E001 = F(1.9681925209e+01), E101 = F(-4.2549998214e+01), E201 = F(5.0774768218e+01),
               E301 = F(-3.0938076334e+01), E401 = F(6.6051753097e+00);
  const F E011 = F(-1.3336301113e+01), E111 = F(-4.4870114575e+00), E211 = F(5.0042598061e+00),
               E311 = F(-6.5399043664e-01);
  const F E021 = F(6.7080479603e+00), E121 = F(3.5063081279e+00), E221 = F(-1.8795372996e+00);
  const F E031 = F(-2.4649669534e+00), E131 = F(-5.5077101279e-01);
  const F E041 = F(5.5927935970e-01);
  const F E002 = F(2.0660924175e+00), E102 = F(-4.9527603989e+00), E202 = F(2.5019633803e+00);
  const F E012 = F(2.0564311499e+00), E112 = F(-2.1311365518e-01);
  const F E022 = F(-1.2419983026e+00);
  const F E003 = F(-2.3342758797e-02), E103 = F(-1.8507636718e-02);
  const F E013 = F(3.7969820455e-01);
  const F RDELTA_S = F(32.0);
  const F R1_S0 = F(0.875) / 35.16504;
  const F R1_T0 = F(1.0) / 40.0;
  const F R1_Z0 = F(1.0e-4);

  const F zh = z * R1_Z0;
  const F ss = sqrt((SA + RDELTA_S) * R1_S0);
  const F tt = CT * R1_T0;
  const F r0 = (((((R05 * zh + R04) * zh + R03) * zh + R02) * zh + R01) * zh + R00) * zh;
  const F rz3 = E013 * tt + E103 * ss + E003;
  const F rz2 = ((E022 * tt + E112 * ss + E012) * tt + (E202 * ss + E102) * ss + E002);
  const F rz1 =
      ((((E041 * tt + E131 * ss + E031) * tt + (E221 * ss + E121) * ss + E021) * tt +
        ((E311 * ss + E211) * ss + E111) * ss + E011) * tt +
       (((E401 * ss + E301) * ss + E201) * ss + E101) * ss + E001);
  const F rz0 =
      ((((((E060 * tt + E150 * ss + E050) * tt + (E240 * ss + E140) * ss + E040) * tt +
          ((E330 * ss + E230) * ss + E130) * ss + E030) * tt +
         (((E420 * ss + E320) * ss + E220) * ss + E120) * ss + E020) * tt +
        ((((E510 * ss + E410) * ss + E310) * ss + E210) * ss + E110) * ss + E010) * tt +
       (((((E600 * ss + E500) * ss + E400) * ss + E300) * ss + E200) * ss + E100) * ss + E000);
  return ((rz3 * zh + rz2) * zh + rz1) * zh + rz0 + r0;
}

// b = -g rho'/rho0 for the configured EOS (libs/core/eos.py:103-133), in
// the arithmetic type F (double: the reference's; float: the mixed step).
template <class F>
__device__ __forceinline__ F buoyancy_t(const StepParams& p, F T, F S, F z) {
  F rp;
  if (p.eos == EOS_LINEAR) {
    rp = F(p.rho0) * (F(p.neg_alpha_T) * (T - F(p.T0)) + F(p.beta_S) * (S - F(p.S0)));
  } else if (p.eos == EOS_SEOS) {
    const F a0 = F(1.6550e-1), b0 = F(7.6554e-1), lam1 = F(5.9520e-2), lam2 = F(7.4914e-4),
            mu1 = F(1.4970e-4), mu2 = F(1.1090e-5), nu_ts = F(2.4341e-3);
    const F Ta = T - F(10), Sa = S - F(35);
    rp = (-a0 * (F(1) + F(0.5) * lam1 * Ta + mu1 * z) * Ta +
          b0 * (F(1) - F(0.5) * lam2 * Sa - mu2 * z) * Sa - nu_ts * Ta * Sa);
  } else {
    rp = rho_teos10<F>(T, S, z) - F(p.rho0);
  }
  return F(p.neg_g) * rp / F(p.rho0);
}
__device__ __forceinline__ double buoyancy(const StepParams& p, double T, double S, double z) {
  return buoyancy_t<double>(p, T, S, z);
}

// Buoyancy of the mixed-precision step. The T/S-dependent part of the EOS is
// evaluated in fp32 on anomalies formed in fp64, so that the large constant
// part of the density never passes through fp32 (rho ~ 1027 against
// rho' ~ 1 kg m-3): linear and SEOS forms on T - T0, S - S0 (T - 10,
// S - 35), polyTEOS10 as its polynomial re-expanded about (T0, S0) --
// rho = rho_ref(z) [fp64] + sum_k zh^k d_k(a, c) [fp32], a = (T - T0) / 40,
// c = ss - ss_ref = (S - S0) 0.875 / 35.16504 / (ss + ss_ref).
__device__ __forceinline__ float teos_anomaly_poly(const float* G, int k, int deg, float a,
                                                   float c) {
  float acc = 0.0f;
  for (int pp = deg; pp >= 0; --pp) {
    float inner = 0.0f;
    for (int q = deg - pp; q >= 0; --q) inner = inner * c + G[(k * 7 + pp) * 7 + q];
    acc = acc * a + inner;
  }
  return acc;
}

__device__ __forceinline__ double buoyancy_mixed(const StepParams& p, double T, double S,
                                                 double z) {
  double rp;
  if (p.eos == EOS_LINEAR) {
    const float dT = (float)(T - p.T0), dS = (float)(S - p.S0);
    rp = (double)((float)p.rho0 * ((float)p.neg_alpha_T * dT + (float)p.beta_S * dS));
  } else if (p.eos == EOS_SEOS) {
    const float a0 = 1.6550e-1f, b0 = 7.6554e-1f, lam1 = 5.9520e-2f, lam2 = 7.4914e-4f,
                mu1 = 1.4970e-4f, mu2 = 1.1090e-5f, nu_ts = 2.4341e-3f;
    const float Ta = (float)(T - 10.0), Sa = (float)(S - 35.0), zf = (float)z;
    rp = (double)(-a0 * (1.0f + 0.5f * lam1 * Ta + mu1 * zf) * Ta +
                  b0 * (1.0f - 0.5f * lam2 * Sa - mu2 * zf) * Sa - nu_ts * Ta * Sa);
  } else {
    const double R00 = 4.6494977072e+01, R01 = -5.2099962525e+00, R02 = 2.2601900708e-01,
                 R03 = 6.4326772569e-02, R04 = 1.5616995503e-02, R05 = -1.7243708991e-03;
    const double zh = z * 1.0e-4;
    const double r0 = (((((R05 * zh + R04) * zh + R03) * zh + R02) * zh + R01) * zh + R00) * zh;
    const double ref = ((p.eos_ref[3] * zh + p.eos_ref[2]) * zh + p.eos_ref[1]) * zh +
                       p.eos_ref[0] + r0;
    // anomalies in fp64; the linear terms in fp64, the terms of degree >= 2
    // (the fp32 table has zeros below degree 2) in fp32
    const double a64 = (T - p.T0) * (1.0 / 40.0);
    const double ss64 = sqrt((S + 32.0) * (0.875 / 35.16504));
    const double c64 = ((S - p.S0) * (0.875 / 35.16504)) / (ss64 + p.ss_ref);
    double lin = 0.0;
    for (int k = 3; k >= 0; --k)
      lin = lin * zh + (p.eos_lin[k][0] * a64 + p.eos_lin[k][1] * c64);
    const float a = (float)a64, c = (float)c64;
    const float zf = (float)zh;
    const float* G = p.eos_g;
    const float d = ((teos_anomaly_poly(G, 3, 1, a, c) * zf + teos_anomaly_poly(G, 2, 2, a, c)) *
                         zf + teos_anomaly_poly(G, 1, 4, a, c)) * zf +
                    teos_anomaly_poly(G, 0, 6, a, c);
    rp = (ref - p.rho0) + (lin + (double)d);
  }
  return p.neg_g * rp / p.rho0;
}

}  // namespace hoc
