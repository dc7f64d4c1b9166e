// Device-side plan constants and the FP64 visibility -> (tile, cell, plane)
// mapping shared by every kernel that needs it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace wg {

// Plain-old-data copy of the plan that kernels take by value.
struct DevPlan {
  double ufac, vfac;      // grid cells per wavelength
  double half_nu, half_nv;
  double w_lo, inv_dw, wofs;  // wofs = W/2 - 1
  double l_shift, m_shift, zc;
  double dw;
  double dx_rad, dy_rad;
  // per tap k < W, the ES kernel as a polynomial in g = 2 f - 1 of the
  // fractional coordinate f (EsPolyCoefficients); rows of taps >= W are 0
  float es_poly[kEsPolyTaps][kEsPolyCoefs];
  int W, T, ntu, ntv, P0, P, nu, nv;
  int tshift;         // log2(T) when T is a power of two, else -1
  int box_u0, box_v0, wrap_u, wrap_v;
  int BU, BV;
  // box storage: row pitch and plane stride in cells (BU and BU * BV for a
  // box of its own; larger when the box is a region of a facet atlas)
  size_t box_pitch, box_plane;
  int nx, ny;
  uint32_t nkeys;     // invalid key; valid keys are below
  uint32_t nbuckets;  // (tile, p0) buckets
  // key = (bucket * key_T + (dv >> key_shift)) * key_T + (du >> key_shift)
  int key_shift, key_T;
  int split;          // 1: split-precision matrix-core gridding (W <= 8)
};

// Sort key of a visibility in bucket `bucket` with first cell (du, dv).
__device__ __forceinline__ uint32_t SortKey(uint32_t bucket, int du, int dv, int key_shift,
                                            int key_T) {
  return (bucket * uint32_t(key_T) + uint32_t(dv >> key_shift)) * uint32_t(key_T) +
         uint32_t(du >> key_shift);
}

DevPlan ToDevPlan(const Plan& p);

// Location of one visibility on the gridding lattice.
struct VisLoc {
  int tile;      // tv * ntu + tu
  int du, dv;    // first kernel cell relative to the tile origin
  int p0;        // first w plane
  float fu, fv, fw;  // fractional part of the continuous grid coordinate
  bool flip;     // (u, v, w, V) -> (-u, -v, -w, V*) was applied
  double u, v, w;    // flipped coordinates in wavelengths
};

// FP64 mapping.  Contraction is off so that the result is bit-identical to
// the same IEEE-754 double operations on the host (numpy restatement in the
// tests).  The first kernel cell is floor(t) - W/2 + 1 and the fractional
// part t - floor(t) is kept separately, so that fp32 holds it to 6e-8.
__device__ __forceinline__ VisLoc LocateVis(double um, double vm, double wm,
                                            double scale, const DevPlan& dp) {
#pragma clang fp contract(off)
  VisLoc r;
  r.flip = wm < 0.0;
  if (r.flip) {
    um = -um;
    vm = -vm;
    wm = -wm;
  }
  const double u = um * scale;
  const double v = vm * scale;
  const double w = wm * scale;
  r.u = u;
  r.v = v;
  r.w = w;
  const double tu = u * dp.ufac + dp.half_nu;
  const double tv = v * dp.vfac + dp.half_nv;
  const double tw = (w - dp.w_lo) * dp.inv_dw + dp.wofs;
  const double flu = floor(tu), flv = floor(tv), flw = floor(tw);
  r.fu = float(tu - flu);
  r.fv = float(tv - flv);
  r.fw = float(tw - flw);
  const int hw = dp.W / 2 - 1;
  int a0u = int(flu) - hw;
  int a0v = int(flv) - hw;
  r.p0 = int(flw) - hw;
  int lu, lv;
  if (dp.wrap_u) {
    lu = a0u % dp.nu;
    if (lu < 0) lu += dp.nu;
  } else {
    lu = a0u - dp.box_u0;
  }
  if (dp.wrap_v) {
    lv = a0v % dp.nv;
    if (lv < 0) lv += dp.nv;
  } else {
    lv = a0v - dp.box_v0;
  }
  // lu, lv >= 0 (box or wrapped cells): a shift where T is a power of two
  const int tu_i = dp.tshift >= 0 ? lu >> dp.tshift : lu / dp.T;
  const int tv_i = dp.tshift >= 0 ? lv >> dp.tshift : lv / dp.T;
  r.du = lu - tu_i * dp.T;
  r.dv = lv - tv_i * dp.T;
  r.tile = tv_i * dp.ntu + tu_i;
  return r;
}

// One tap of the ES kernel from its polynomial c[0..D] in g = 2 f - 1
// (Horner, full-rate FMAs): phi(x) at x = (k - W/2 + 1 - f) * 2 / W for the
// tap k whose coefficients c are.  Accurate to ~4e-8 of phi's peak in FP32,
// against ~3e-7 for exp2 / sqrt of the closed form, whose sqrt(1 - x^2) - 1
// cancels.
template <int D>
__device__ __forceinline__ float EsPolyG(const float (&c)[D + 1], float g) {
  float p = c[D];
#pragma unroll
  for (int d = D - 1; d >= 0; --d) p = __builtin_fmaf(p, g, c[d]);
  return p;
}
template <int D>
__device__ __forceinline__ float EsPoly(const float (&c)[D + 1], float f) {
  return EsPolyG<D>(c, __builtin_fmaf(f, 2.0f, -1.0f));
}
// phi is even, so tap W - 1 - k at f equals tap k at 1 - f, i.e. the
// polynomial of tap k at -g: the taps k < W / 2 give all W.

// Even and odd parts of the tap polynomial c[0..D] (D odd) at y = g^2, so
// that p(g) = E + g O and p(-g) = E - g O: two taps from D + 1 FMAs
// (Horner in y for each part) and one FMA each.
template <int D>
__device__ __forceinline__ void EsPolyEvenOdd(const float (&c)[D + 1], float y, float& E,
                                              float& O) {
  static_assert(D % 2 == 1, "even / odd split of an odd-degree polynomial");
  E = c[D - 1];
  O = c[D];
#pragma unroll
  for (int d = D - 3; d >= 0; d -= 2) {
    E = __builtin_fmaf(E, y, c[d]);
    O = __builtin_fmaf(O, y, c[d + 1]);
  }
}

// The coefficients of tap k (any lane-varying k < kEsPolyTaps).
template <int D>
__device__ __forceinline__ void EsPolyLoad(const DevPlan& dp, int k, float (&c)[D + 1]) {
#pragma unroll
  for (int d = 0; d <= D; ++d) c[d] = dp.es_poly[k][d];
}

// Weights of the W taps for fractional coordinate f in [0, 1): tap k sits at
// x = (k - W/2 + 1 - f) * 2 / W.
template <int W>
__device__ __forceinline__ void EsTaps(const DevPlan& dp, float f, float* out) {
  constexpr int D = EsPolyDegree(W);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    float c[D + 1];
    EsPolyLoad<D>(dp, k, c);
    out[k] = EsPoly<D>(c, f);
  }
}

// max(a, b) that keeps a NaN operand (fmaxf returns the other one): the
// bounds of a work item holding a NaN or infinite visibility become NaN,
// and its subgrid is written as NaN, so the image shows the bad input
// instead of silently losing the item's visibilities.
__device__ __forceinline__ float MaxKeepNaN(float a, float b) {
  return (a > b || a != a) ? a : b;
}

// |V|^2 with one fixed rounding sequence (so every kernel that takes the
// max-modulus of the same visibilities gets the same bits).
__device__ __forceinline__ float SquaredModulus(float2 v) {
  return __builtin_fmaf(v.x, v.x, v.y * v.y);
}

// i / d for i, d < 2^31 without an integer division: the FP64 product with
// the rounded reciprocal is within one of the quotient, and one remainder
// test corrects it.
__device__ __forceinline__ uint32_t DivSmall(uint32_t i, uint32_t d, double inv_d) {
  uint32_t q = uint32_t(double(i) * inv_d);
  const int64_t r = int64_t(i) - int64_t(q) * int64_t(d);
  if (r < 0) --q;
  else if (r >= int64_t(d)) ++q;
  return q;
}

// Polarizations, with the values of the FITS STOKES axis (the values of
// aocommon::PolarizationEnum): Stokes I, Q, U, V, circular RR, LL, RL, LR
// and linear XX, YY, XY, YX.
enum PolCode : int {
  kPolI = 1, kPolQ = 2, kPolU = 3, kPolV = 4,
  kPolRR = -1, kPolLL = -2, kPolRL = -3, kPolLR = -4,
  kPolXX = -5, kPolYY = -6, kPolXY = -7, kPolYX = -8
};
inline bool ValidPolCode(int p) { return (p >= 1 && p <= 4) || (p >= -8 && p <= -1); }

// One visibility of polarization `pol` from the four linear correlations
// v = (XX, XY, YX, YY) (gridding/msgridderdata.h:45-49; aocommon's
// ConvertFromLinear, restated from the IAU definitions XX = I + Q,
// YY = I - Q, XY = U + iV, YX = U - iV and RR = I + V, LL = I - V,
// RL = Q + iU, LR = Q - iU).
__host__ __device__ __forceinline__ float2 FromLinear(const float2* v, int pol) {
  const float ix = 0.5f * (v[0].x + v[3].x), iy = 0.5f * (v[0].y + v[3].y);  // I
  const float qx = 0.5f * (v[0].x - v[3].x), qy = 0.5f * (v[0].y - v[3].y);  // Q
  const float ux = 0.5f * (v[1].x + v[2].x), uy = 0.5f * (v[1].y + v[2].y);  // U
  // V = -i (XY - YX) / 2
  const float vx = 0.5f * (v[1].y - v[2].y), vy = -0.5f * (v[1].x - v[2].x);
  switch (pol) {
    case kPolXX: return v[0];
    case kPolXY: return v[1];
    case kPolYX: return v[2];
    case kPolYY: return v[3];
    case kPolQ: return make_float2(qx, qy);
    case kPolU: return make_float2(ux, uy);
    case kPolV: return make_float2(vx, vy);
    case kPolRR: return make_float2(ix + vx, iy + vy);
    case kPolLL: return make_float2(ix - vx, iy - vy);
    case kPolRL: return make_float2(qx - uy, qy + ux);  // Q + iU
    case kPolLR: return make_float2(qx + uy, qy - ux);  // Q - iU
    default: return make_float2(ix, iy);                // Stokes I
  }
}

// Top bit of a sorted permutation entry: the visibility was flipped to w >= 0
// (the flat index itself is below 2^31).
constexpr uint32_t kPermFlip = 0x80000000u;

// Caller row of a renumbered row: the bits of the fourth word of its padded
// uvw4 entry (w, tag).
__device__ __forceinline__ uint32_t RowTag(double2 wt) {
  return uint32_t(__double_as_longlong(wt.y));
}
// The high half of that word: a facet batch's facet of the row (else 0).
__device__ __forceinline__ uint32_t RowHi(double2 wt) {
  return uint32_t(uint64_t(__double_as_longlong(wt.y)) >> 32);
}

// exp(sign 2 pi i (u l_shift + v m_shift + w zc)) for flipped coordinates.
__device__ __forceinline__ float2 ShiftPhasor(const DevPlan& dp, const VisLoc& L, float sign) {
  const double cyc = L.u * dp.l_shift + L.v * dp.m_shift + L.w * dp.zc;
  const double f = cyc - rint(cyc);
  float s, c;
  sincospif(2.0f * float(f), &s, &c);
  return make_float2(c, sign * s);
}

// Packed per-visibility word: du (8 bits) | dv (8 bits) | p0 (15 bits) | flip.
__device__ __forceinline__ uint32_t PackLoc(const VisLoc& l) {
  return uint32_t(l.du) | (uint32_t(l.dv) << 8) | (uint32_t(l.p0) << 16) |
         (l.flip ? 0x80000000u : 0u);
}

}  // namespace wg
