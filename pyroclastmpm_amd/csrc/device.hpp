// device.hpp -- device-side building blocks shared by the gfx950 kernels:
// B-spline weights per axis, symmetric 3x3 stress algebra and the four
// return maps. Header-only, included by the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pmpm.h"

namespace pmpm {

// ---------------------------------------------------------------------------
// Spline traits. SPAN = node extent per cell (tile halo), ENTRIES = nodes one
// particle touches per axis. NLO = first node offset relative to the cell.
// ---------------------------------------------------------------------------
template <int SF> struct Spline;
template <> struct Spline<PMPM_SF_LINEAR> { static constexpr int SPAN = 2, ENTRIES = 2, NLO = 0; };
template <> struct Spline<PMPM_SF_QUADRATIC> { static constexpr int SPAN = 4, ENTRIES = 3, NLO = -1; };
template <> struct Spline<PMPM_SF_CUBIC> { static constexpr int SPAN = 4, ENTRIES = 4, NLO = -1; };

// Boundary-modified cubic B-spline pieces (cubic.cuh:28-197): coefficients of
// c0 + c1 r + c2 r^2 + c3 r^3 for node type 0..4 (0 = untyped -> 0) and piece
// k covering r in [k-2, k-1).
__device__ __constant__ static const float kCubic[5][4][4] = {
    {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},
    {{4.f / 3.f, 2.f, 1.f, 1.f / 6.f}, {1.f, 1.f, 0.f, -1.f / 6.f}, {1.f, -1.f, 0.f, 1.f / 6.f}, {4.f / 3.f, -2.f, 1.f, -1.f / 6.f}},
    {{0, 0, 0, 0}, {2.f / 3.f, 0.f, -1.f, -1.f / 3.f}, {2.f / 3.f, 0.f, -1.f, 0.5f}, {4.f / 3.f, -2.f, 1.f, -1.f / 6.f}},
    {{4.f / 3.f, 2.f, 1.f, 1.f / 6.f}, {2.f / 3.f, 0.f, -1.f, -0.5f}, {2.f / 3.f, 0.f, -1.f, 0.5f}, {4.f / 3.f, -2.f, 1.f, -1.f / 6.f}},
    {{4.f / 3.f, 2.f, 1.f, 1.f / 6.f}, {2.f / 3.f, 0.f, -1.f, -0.5f}, {2.f / 3.f, 0.f, -1.f, 1.f / 3.f}, {0, 0, 0, 0}}};

__device__ inline int cubic_node_type(int id, int n, int built) {
  if (!built) return 0;
  if (id == 0 || id == n - 1) return 1;
  if (id == 1) return 2;
  return 3;  // the reference never assigns type 4 (nodes.cpp:160)
}

__device__ inline void cubic_eval(float r, int type, float inv_h, float &N, float &dN) {
  int k;
  if (r >= -2.f && r < -1.f) k = 0;
  else if (r >= -1.f && r < 0.f) k = 1;
  else if (r >= 0.f && r < 1.f) k = 2;
  else if (r >= 1.f && r < 2.f) k = 3;
  else { N = 0.f; dN = 0.f; return; }
  const float *c = kCubic[type][k];
  N = ((c[3] * r + c[2]) * r + c[1]) * r + c[0];
  dN = inv_h * ((3.f * c[3] * r + 2.f * c[2]) * r + c[1]);
}

__device__ inline void linear_eval(float r, float inv_h, float &N, float &dN) {
  const float a = fabsf(r);
  N = (a >= 1.f) ? 0.f : __fsub_rn(1.f, a);
  dN = (r > 1.f || r < -1.f || r == 0.f) ? 0.f : (r > 0.f ? -inv_h : inv_h);
}

__device__ inline void quad_eval(float r, float inv_h, float &N, float &dN) {
  const float a = fabsf(r);
  if (a < 0.5f) { N = 0.75f - r * r; dN = -2.f * r * inv_h; }
  else if (a < 1.5f) { const float b = 1.5f - a; N = 0.5f * b * b; dN = (r > 0.f ? -b : b) * inv_h; }
  else { N = 0.f; dN = 0.f; }
}

// Any-spline, any-offset evaluation (used by the materialised shape function
// and by the wrap-around slow path): r = relative coordinate, type = cubic
// node type.
template <int SF> __device__ inline void spline_eval(float r, int type, float inv_h, float &N, float &dN) {
  if (SF == PMPM_SF_LINEAR) linear_eval(r, inv_h, N, dN);
  else if (SF == PMPM_SF_QUADRATIC) quad_eval(r, inv_h, N, dN);
  else cubic_eval(r, type, inv_h, N, dN);
}

// Second derivative d2N/dx2 of the same pieces (for position adjoints): the
// linear spline's is 0, the quadratic's -2 and 1 (times inv_h^2), the cubic
// piece's 2 c2 + 6 c3 r.
template <int SF> __device__ inline float spline_d2(float r, int type, float inv_h) {
  const float h2 = inv_h * inv_h;
  if (SF == PMPM_SF_LINEAR) return 0.f;
  if (SF == PMPM_SF_QUADRATIC) {
    const float a = fabsf(r);
    return a < 0.5f ? -2.f * h2 : (a < 1.5f ? h2 : 0.f);
  }
  int k;
  if (r >= -2.f && r < -1.f) k = 0;
  else if (r >= -1.f && r < 0.f) k = 1;
  else if (r >= 0.f && r < 1.f) k = 2;
  else if (r >= 1.f && r < 2.f) k = 3;
  else return 0.f;
  const float *c = kCubic[type][k];
  return h2 * (2.f * c[2] + 6.f * c[3] * r);
}

// Coefficients c0 + c1 r + c2 r^2 + c3 r^3 of cubic piece 3 - S for node
// types 1-3 (the kCubic rows, chosen with selects instead of a table load).
// Node offset S (node bin - 1 + S) of a particle with frac = t - bin in
// [0, 1) lies on that piece: r = frac + 1 - S. Type 2 (node 1) never sits at
// S = 3 of such a particle (it would need bin = -2).
template <int S> __device__ inline void cubic_span_coeffs(int type, float c[4]) {
  const bool t1 = type == 1, t2 = type == 2;
  if (S == 0) {  // piece 3, identical for types 1-3
    c[0] = 4.f / 3.f; c[1] = -2.f; c[2] = 1.f; c[3] = -1.f / 6.f;
  } else if (S == 1) {  // piece 2
    c[0] = t1 ? 1.f : 2.f / 3.f; c[1] = t1 ? -1.f : 0.f; c[2] = t1 ? 0.f : -1.f; c[3] = t1 ? 1.f / 6.f : 0.5f;
  } else if (S == 2) {  // piece 1
    c[0] = t1 ? 1.f : 2.f / 3.f; c[1] = t1 ? 1.f : 0.f; c[2] = t1 ? 0.f : -1.f;
    c[3] = t1 ? -1.f / 6.f : (t2 ? -1.f / 3.f : -0.5f);
  } else {  // piece 0, zero for type 2
    c[0] = t2 ? 0.f : 4.f / 3.f; c[1] = t2 ? 0.f : 2.f; c[2] = t2 ? 0.f : 1.f; c[3] = t2 ? 0.f : 1.f / 6.f;
  }
}

// Weight, derivative and (if d2N) second derivative of node offset S from
// the span form. At frac = 1 - 2^-24 rounding puts r on the next knot: at
// S = 0 (r = 2) the weight is 0 as in cubic_eval; at S = 3 (r = -1) both
// pieces meet there (the spline is C2), so the values agree to rounding.
template <int S>
__device__ inline void cubic_span(float frac, int type, float inv_h, float &N, float &dN, float *d2N = nullptr) {
  float c[4];
  cubic_span_coeffs<S>(type, c);
  const float r = __fsub_rn(frac, (float)(S - 1));
  N = ((c[3] * r + c[2]) * r + c[1]) * r + c[0];
  dN = inv_h * ((3.f * c[3] * r + 2.f * c[2]) * r + c[1]);
  if (d2N) *d2N = inv_h * inv_h * (2.f * c[2] + 6.f * c[3] * r);
  if (S == 0 && !(r < 2.f)) {
    N = 0.f;
    dN = 0.f;
    if (d2N) *d2N = 0.f;
  }
}

// The four cubic window entries of one axis (nodes bin - 1 .. bin + 2). The
// span form serves frac in [0, 1), every particle at t >= 0; a particle left
// of the origin (t < 0, where trunc gives frac in (-1, 0)) takes the table
// form, whose pieces then shift by one.
__device__ inline void cubic_axis(float frac, int bin, int n_axis, int types_built, float inv_h, float *N, float *dN,
                                  float *d2N = nullptr) {
  if (!types_built) {  // untyped nodes: kCubic row 0
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      N[e] = 0.f;
      dN[e] = 0.f;
      if (d2N) d2N[e] = 0.f;
    }
    return;
  }
  int type[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) type[e] = cubic_node_type(bin - 1 + e, n_axis, 1);
  if (frac >= 0.f) {
    cubic_span<0>(frac, type[0], inv_h, N[0], dN[0], d2N ? d2N + 0 : nullptr);
    cubic_span<1>(frac, type[1], inv_h, N[1], dN[1], d2N ? d2N + 1 : nullptr);
    cubic_span<2>(frac, type[2], inv_h, N[2], dN[2], d2N ? d2N + 2 : nullptr);
    cubic_span<3>(frac, type[3], inv_h, N[3], dN[3], d2N ? d2N + 3 : nullptr);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float r = __fsub_rn(frac, (float)(e - 1));
      cubic_eval(r, type[e], inv_h, N[e], dN[e]);
      if (d2N) d2N[e] = spline_d2<PMPM_SF_CUBIC>(r, type[e], inv_h);
    }
  }
}

// t = (x - origin) * inv_h rounded after each operation, exactly as the
// reference evaluates it (no fused multiply-add: the cell index and the
// relative coordinates must come out bit-identical).
__device__ inline float cell_coord(float x, float origin, float inv_h) {
  // the product must not fuse with a later subtraction of the cell index
#pragma clang fp contract(off)
  return (x - origin) * inv_h;
}

// Per-axis weights of one particle. t = (x - origin) * inv_h, bin = trunc(t).
// first = offset (relative to bin) of entry 0; N[e], dN[e] for the ENTRIES
// nodes bin+first+e. The relative coordinate (t - bin) - o equals the
// reference's t - (bin + o) bit for bit (t - bin is exact).
template <int SF>
__device__ inline void axis_weights(float t, int bin, int n_axis, int types_built, float inv_h,
                                    int &first, float *N, float *dN) {
  const float frac = __fsub_rn(t, (float)bin);
  if (SF == PMPM_SF_LINEAR) {
    first = 0;
#pragma unroll
    for (int e = 0; e < 2; ++e) linear_eval(__fsub_rn(frac, (float)e), inv_h, N[e], dN[e]);
  } else if (SF == PMPM_SF_QUADRATIC) {
    first = (frac < 0.5f) ? -1 : 0;
#pragma unroll
    for (int e = 0; e < 3; ++e) quad_eval(__fsub_rn(frac, (float)(first + e)), inv_h, N[e], dN[e]);
  } else {
    first = -1;
    cubic_axis(frac, bin, n_axis, types_built, inv_h, N, dN);
  }
}

// ---------------------------------------------------------------------------
// Symmetric 3x3 (stress, elastic strain): xx yy zz xy xz yz.
// ---------------------------------------------------------------------------
struct Sym {
  float xx, yy, zz, xy, xz, yz;
};
__device__ inline float sym_at(const Sym &s, int i, int j) {
  if (i == j) return i == 0 ? s.xx : (i == 1 ? s.yy : s.zz);
  const int k = i + j;  // 1 -> xy, 2 -> xz, 3 -> yz
  return k == 1 ? s.xy : (k == 2 ? s.xz : s.yz);
}
__device__ inline Sym sym_zero() { return Sym{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }
__device__ inline Sym load_sym(const float *a, int n, int i) {
  return Sym{a[i], a[n + i], a[2 * n + i], a[3 * n + i], a[4 * n + i], a[5 * n + i]};
}
__device__ inline void store_sym(float *a, int n, int i, const Sym &s) {
  a[i] = s.xx; a[n + i] = s.yy; a[2 * n + i] = s.zz;
  a[3 * n + i] = s.xy; a[4 * n + i] = s.xz; a[5 * n + i] = s.yz;
}
__device__ inline float sym_frob2(const Sym &s) {
  return s.xx * s.xx + s.yy * s.yy + s.zz * s.zz + 2.f * (s.xy * s.xy + s.xz * s.xz + s.yz * s.yz);
}

// ---------------------------------------------------------------------------
// Return maps. L is the velocity gradient padded to 3x3 row-major (L[3*i+j]),
// entries outside the D x D block are 0. dt = time step.
// ---------------------------------------------------------------------------

// linearelastic_inline.h:40-107: sigma_DxD += lambda*deps + 2G*deps,
// deps = 0.5 (L + L^T) dt (the reference's lambda*deps*I is a matrix product).
template <int D>
__device__ inline void rm_linear_elastic(const pmpm_material &m, float dt, const float *L, Sym &s) {
  const float lam = m.p[0], twoG = 2.f * m.p[1];
  float de[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) de[i][j] = 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt;
  s.xx += lam * de[0][0] + twoG * de[0][0];
  if (D > 1) { s.yy += lam * de[1][1] + twoG * de[1][1]; s.xy += lam * de[0][1] + twoG * de[0][1]; }
  if (D > 2) {
    s.zz += lam * de[2][2] + twoG * de[2][2];
    s.xz += lam * de[0][2] + twoG * de[0][2];
    s.yz += lam * de[1][2] + twoG * de[1][2];
  }
}

// newtonfluid_inline.h:49-96: p = -K((rho/rho0)^gamma - 1) on the D x D block,
// tau = (2 mu / 3) dev(D); the stress is replaced, not incremented.
template <int D>
__device__ inline void rm_newton_fluid(const pmpm_material &m, const float *L, float mass, float vol,
                                       float vol0, Sym &s) {
  const float visc = m.p[0], K = m.p[1], gam = m.p[2];
  const float rho = mass / vol, rho0 = mass / vol0;
  const float pres = -K * (powf(rho / rho0, gam) - 1.f);
  float Dm[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Dm[i][j] = 0.5f * (L[3 * i + j] + L[3 * j + i]);
  const float tr = Dm[0][0] + (D > 1 ? Dm[1][1] : 0.f) + (D > 2 ? Dm[2][2] : 0.f);
  const float third = (float)((1.0 / 3.0) * (double)tr);
  const float fac = (float)(2.0 * (double)visc / 3.0);
  Sym o = sym_zero();
  o.xx = pres + (Dm[0][0] - third) * fac;
  if (D > 1) { o.yy = pres + (Dm[1][1] - third) * fac; o.xy = Dm[0][1] * fac; }
  if (D > 2) { o.zz = pres + (Dm[2][2] - third) * fac; o.xz = Dm[0][2] * fac; o.yz = Dm[1][2] * fac; }
  s = o;
}

// localrheo_inline.h:39-157: hypoelastic Jaumann trial, zero stress below the
// critical density or in tension, otherwise the rate-dependent radial return
// with the reference's hard-coded grain density 2450 and diameter 0.005.
__device__ inline double negative_root(double a, double b, double c) {
  if (b > 0) return (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
  return (2 * c) / (-b + sqrt(b * b - 4 * a * c));
}
__device__ inline void rm_local_rheo(const pmpm_material &m, float dt, const float *L, float mass,
                                     float vol, Sym &s) {
  const float G = m.p[0], lam = m.p[1], rho_c = m.p[3], mu_s = m.p[4], mu_2 = m.p[5], I0 = m.p[6];
  const float rho = mass / vol;
  float Dm[3][3], W[3][3], S[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Dm[i][j] = 0.5f * (L[3 * i + j] + L[3 * j + i]);
      W[i][j] = 0.5f * (L[3 * i + j] - L[3 * j + i]);
      S[i][j] = sym_at(s, i, j);
    }
  const float trD = Dm[0][0] + Dm[1][1] + Dm[2][2];
  float st[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float sw = 0.f, ws = 0.f;
#pragma unroll
      for (int k = 0; k < 3; ++k) { sw += S[i][k] * W[k][j]; ws += W[i][k] * S[k][j]; }
      st[i][j] = S[i][j] + dt * (2.f * G * Dm[i][j] + (i == j ? lam * trD : 0.f) + (-sw + ws));
    }
  const float ptr = (float)(-(1.0 / 3.0) * (double)(st[0][0] + st[1][1] + st[2][2]));
  float s0[3][3];
  float f2 = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      s0[i][j] = st[i][j] + (i == j ? ptr : 0.f);
      f2 += s0[i][j] * s0[i][j];
    }
  const float tau = (float)sqrt(0.5 * (double)f2);
  if (ptr < 0.f || rho < rho_c) { s = sym_zero(); return; }
  const float S0 = mu_s * ptr;
  float scale = 1.f;
  if (!(tau <= S0)) {
    const float S2 = mu_2 * ptr;
    const float al = G * I0 * dt * sqrtf(ptr / 2450.0f) / 0.005f;
    const float B = -(S2 + tau + al);
    const float H = S2 * tau + S0 * al;
    const float tn = (float)negative_root(1.0, (double)B, (double)H);
    scale = tn / tau;
  }
  s.xx = scale * s0[0][0] - ptr;
  s.yy = scale * s0[1][1] - ptr;
  s.zz = scale * s0[2][2] - ptr;
  s.xy = scale * s0[0][1];
  s.xz = scale * s0[0][2];
  s.yz = scale * s0[1][2];
}

// dp_rheo_inline.h:43-290 (+ dp_rheo.cpp:45-70): co-rotated elastic-strain
// trial, perfectly plastic Drucker-Prager cone (eta = M/sqrt3, tension cut Pt,
// dilatancy eta_bar) with the closed-form cone and apex returns.
template <int D>
__device__ inline void rm_drucker_prager(const pmpm_material &m, float dt, const float *L,
                                         const Sym &sref, Sym &s, Sym &eps) {
  const float K = m.p[0], G = m.p[1], M = m.p[2], Pt = m.p[3], etab = m.p[4];
  float de[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  float ep[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) ep[i][j] = (i < D && j < D) ? sym_at(eps, i, j) : 0.f;
  if (m.flags & PMPM_MATF_VELGRAD_IS_INCREMENT) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) de[i][j] = L[3 * i + j];
  } else {
    float W[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        de[i][j] = 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt;
        W[i][j] = 0.5f * (L[3 * i + j] - L[3 * j + i]) * dt;
      }
    float c[3][3];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        float we = 0.f, ew = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) { we += W[i][k] * ep[k][j]; ew += ep[i][k] * W[k][j]; }
        c[i][j] = -we + ew;
      }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) de[i][j] += c[i][j];
  }
  const float pref = (float)((1.0 / 3.0) * (double)(sref.xx + sref.yy + sref.zz));
  float sr[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) sr[i][j] = sym_at(sref, i, j) - (i == j ? pref : 0.f);
  float et[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) et[i][j] = (i < D && j < D) ? ep[i][j] + de[i][j] : 0.f;
  const float ev = et[0][0] + (D > 1 ? et[1][1] : 0.f) + (D > 2 ? et[2][2] : 0.f);
  const float ptr = pref + K * ev;
  const float ev3 = (1.f / 3.f) * ev;
  float str[3][3];
  float f2 = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float dev = et[i][j] - (i == j ? ev3 : 0.f);
      if (D < 3 && i == 2 && j == 2) dev = 0.f;
      str[i][j] = sr[i][j] + 2.f * G * dev;
      f2 += str[i][j] * str[i][j];
    }
  const float J2 = 0.5f * f2;
  const float eta = (float)(0.57735026918962576 * (double)M);
  const float sqJ2 = sqrtf(J2);
  const float Phi = sqJ2 + eta * (ptr - Pt);
  if (Phi <= 0.f) {
    s = Sym{str[0][0] + ptr, str[1][1] + ptr, str[2][2] + ptr, str[0][1], str[0][2], str[1][2]};
    if (m.flags & PMPM_MATF_UPDATE_HISTORY)
      eps = Sym{et[0][0], D > 1 ? et[1][1] : 0.f, D > 2 ? et[2][2] : 0.f, D > 1 ? et[0][1] : 0.f,
                D > 2 ? et[0][2] : 0.f, D > 2 ? et[1][2] : 0.f};
    return;
  }
  // cone (dp_rheo_inline.h:50-65): the return runs in double, as the
  // reference's return_mapping_cone takes its moduli, trial pressure, J2 and
  // plastic multiplier as double; the apex test compares the float sqrt(J2)
  // with the double G dgamma (dp_rheo_inline.h:237-238)
  const double Kd = (double)K, Gd = (double)G;
  const double dg = (double)Phi / (Kd * (double)eta * (double)etab + Gd);
  double pn = (double)ptr - Kd * (double)etab * dg;
  const float f = (float)(1.0 - (Gd * dg) / sqrt((double)J2));
  float sn[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) sn[i][j] = f * str[i][j];
  if ((double)sqJ2 - Gd * dg < 0.0) {  // apex (dp_rheo_inline.h:67-78)
    if ((m.flags & PMPM_MATF_APEX_GUARD) && etab == 0.f) {
      pn = (double)Pt;
    } else {
      const double dga = ((double)ptr - (double)Pt) / ((double)etab * Kd);
      pn = (double)ptr - Kd * (double)etab * dga;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) sn[i][j] = 0.f;
  }
  // update_plasticity (dp_rheo_inline.h:80-96): p_next stays double until the
  // stress is stored; the volumetric strain is formed in double
  const float pnf = (float)pn;
  s = Sym{sn[0][0] + pnf, sn[1][1] + pnf, sn[2][2] + pnf, sn[0][1], sn[0][2], sn[1][2]};
  if (m.flags & PMPM_MATF_UPDATE_HISTORY) {
    const float ish = (float)(2.0 * Gd);
    const float vol = (float)((pn - (double)pref) / (3.0 * Kd));
    float e[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) e[i][j] = (sn[i][j] - sr[i][j]) / ish + (i == j ? vol : 0.f);
    eps = Sym{e[0][0], D > 1 ? e[1][1] : 0.f, D > 2 ? e[2][2] : 0.f, D > 1 ? e[0][1] : 0.f,
              D > 2 ? e[0][2] : 0.f, D > 2 ? e[1][2] : 0.f};
  }
}

// ---------------------------------------------------------------------------
// Return maps of the materials the step runs in a pass after G2P (their
// history lives in pmpm_particles.hist): von Mises, Mohr-Coulomb and the
// modified Cam-Clay. eps is the D x D elastic strain (symmetric storage,
// components outside D x D zero), h0/h1 the particle's history scalars.
// ---------------------------------------------------------------------------
__host__ __device__ inline bool is_ext_material(int kind) {
  return kind == PMPM_MAT_VON_MISES || kind == PMPM_MAT_MOHR_COULOMB || kind == PMPM_MAT_MODIFIED_CAM_CLAY;
}
// Newton iterations of the plastic returns below stop here at the latest; the
// reference's loops have no cap (see each).
constexpr int kMaxReturnIters = 100;

// Strain increment of the step: L itself (servo-control form) or
// 0.5 (L + L^T) dt, both over the D x D block.
template <int D> __device__ inline void strain_increment(const pmpm_material &m, float dt, const float *L, float de[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const bool in = i < D && j < D;
      de[i][j] = !in ? 0.f
                     : ((m.flags & PMPM_MATF_VELGRAD_IS_INCREMENT) ? L[3 * i + j]
                                                                    : 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt);
    }
}
__device__ inline void sym_to_mat(const Sym &s, float m[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) m[i][j] = sym_at(s, i, j);
}
__device__ inline Sym mat_to_sym(const float m[3][3]) { return Sym{m[0][0], m[1][1], m[2][2], m[0][1], m[0][2], m[1][2]}; }
// (M M^T).trace() of a 3 x 3 matrix, row by row as a coefficient product sums it
__device__ inline float frob2_rows(const float m[3][3]) {
  float t = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float r = __fadd_rn(__fadd_rn(__fmul_rn(m[i][0], m[i][0]), __fmul_rn(m[i][1], m[i][1])),
                              __fmul_rn(m[i][2], m[i][2]));
    t = i == 0 ? r : __fadd_rn(t, r);
  }
  return t;
}

// vonmises_inline.h:41-140 (3D): elastic trial on the elastic strain,
// J2 yield with linear isotropic hardening sigma_y = yield + H acc, Newton on
// the plastic multiplier in double. h0 = accumulated plastic strain.
// (The reference's Newton loop has no cap; here it stops after
// kMaxReturnIters, which the linear hardening never reaches.)
__device__ inline void rm_von_mises(const pmpm_material &m, float dt, const float *L, Sym &s, Sym &eps, float &h0) {
  const float K = m.p[0], G = m.p[1], ys = m.p[2], H = m.p[3];
  float de[3][3], et[3][3];
  strain_increment<3>(m, dt, L, de);
  float ep[3][3];
  sym_to_mat(eps, ep);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) et[i][j] = __fadd_rn(ep[i][j], de[i][j]);
  const float ev = __fadd_rn(__fadd_rn(et[0][0], et[1][1]), et[2][2]);
  const float ptr = __fmul_rn(K, ev);
  const float ev3 = (float)((1 / 3.) * (double)ev);
  const float twoG = (float)(2. * (double)G);
  float st[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) st[i][j] = __fmul_rn(twoG, i == j ? __fsub_rn(et[i][j], ev3) : et[i][j]);
  const float acc_tr = h0;
  const float sy_tr = __fadd_rn(ys, __fmul_rn(H, acc_tr));
  const float q = (float)sqrt(1.5 * (double)frob2_rows(st));
  const float phi = __fsub_rn(q, sy_tr);
  const bool hist = (m.flags & PMPM_MATF_UPDATE_HISTORY) != 0;
  if (phi <= 0.f) {
    s = Sym{__fadd_rn(st[0][0], ptr), __fadd_rn(st[1][1], ptr), __fadd_rn(st[2][2], ptr), st[0][1], st[0][2], st[1][2]};
    if (hist) eps = mat_to_sym(et);
    return;
  }
  double dg = 0.0, acc = 0.0;
  const double Gd = G, Hd = H, qd = q;
  for (int it = 0; it < kMaxReturnIters; ++it) {
    acc = (double)acc_tr + dg;
    const double sy = (double)ys + Hd * acc;
    const double ph = qd - 3.0 * Gd * dg - sy;
    dg = dg - ph / (-3.0 * Gd - Hd);
    if (!(fabs(ph) > 1e-7)) break;
  }
  const float f = (float)(1.0 - (3.0 * Gd * dg) / qd);
  float sc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) sc[i][j] = __fmul_rn(f, st[i][j]);
  s = Sym{__fadd_rn(sc[0][0], ptr), __fadd_rn(sc[1][1], ptr), __fadd_rn(sc[2][2], ptr), sc[0][1], sc[0][2], sc[1][2]};
  if (hist) {
    const float g2 = (float)(2.0 * Gd);
    float e[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) e[i][j] = __fadd_rn(__fdiv_rn(sc[i][j], g2), i == j ? ev3 : 0.f);
    eps = mat_to_sym(e);
    h0 = (float)acc;
  }
}

// Eigen-decomposition of a symmetric 3 x 3 matrix by cyclic Jacobi rotations
// in double: w descending, column k of V the eigenvector of w[k].
__host__ __device__ inline void sym_eig3_desc(const double A[3][3], double w[3], double V[3][3]) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = A[i][j];
      V[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-32 * dia || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- A J (columns p, q)
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - sn * akq;
          a[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A (rows p, q)
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - sn * aqk;
          a[q][k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - sn * vkq;
          V[k][q] = sn * vkp + c * vkq;
        }
      }
  }
  int o[3] = {0, 1, 2};  // descending order
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2 - i; ++j)
      if (a[o[j]][o[j]] < a[o[j + 1]][o[j + 1]]) { const int t = o[j]; o[j] = o[j + 1]; o[j + 1] = t; }
  double Vs[3][3];
  for (int k = 0; k < 3; ++k) {
    w[k] = a[o[k]][o[k]];
    for (int i = 0; i < 3; ++i) Vs[i][k] = V[i][o[k]];
  }
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) V[i][k] = Vs[i][k];
}

// mohrcoulomb_inline.h:301-457 (update_mohrcoulomb; returns 34-289): elastic
// trial in principal stresses (sigma_1 >= sigma_2 >= sigma_3, tension
// positive), yield sigma_1 - sigma_3 + (sigma_1 + sigma_3) sin(phi)
// - 2 c cos(phi) with c = cohesion + H acc; return to the main plane, then
// the right edge, then the left edge (each accepted when the truncated
// principal stresses keep their order), else the apex. Newton loops in
// double, stopped after kMaxReturnIters at the latest (the reference's have
// no cap). The principal frame comes from a Jacobi decomposition in double
// (the reference's Eigen closed form in float agrees to its rounding).
// h0 = accumulated plastic strain. With dilatancy 0 the apex return divides
// by zero as the reference's does.
template <int D>
__device__ inline void rm_mohr_coulomb(const pmpm_material &m, float dt, const float *L, Sym &s, Sym &eps, float &h0) {
  const float K = m.p[0], G = m.p[1], coh = m.p[2], fric = m.p[3], dil = m.p[4], H = m.p[5];
  const bool hist = (m.flags & PMPM_MATF_UPDATE_HISTORY) != 0;
  float de[3][3], ep[3][3], et[3][3];
  strain_increment<D>(m, dt, L, de);
  sym_to_mat(eps, ep);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) et[i][j] = (i < D && j < D) ? __fadd_rn(ep[i][j], de[i][j]) : 0.f;
  float ev = et[0][0];
  if (D > 1) ev = __fadd_rn(ev, et[1][1]);
  if (D > 2) ev = __fadd_rn(ev, et[2][2]);
  const float ptr = __fmul_rn(K, ev);
  const float ev3 = (float)((1 / 3.) * (double)ev);
  const float twoG = (float)(2. * (double)G);
  float sig[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float sv = (i < D && j < D) ? __fmul_rn(twoG, i == j ? __fsub_rn(et[i][j], ev3) : et[i][j]) : 0.f;
      sig[i][j] = i == j ? __fadd_rn(sv, ptr) : sv;
    }
  const float acc_tr = h0;
  const float coh_tr = __fadd_rn(coh, __fmul_rn(H, acc_tr));
  double A[3][3], w[3], V[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = sig[i][j];
  sym_eig3_desc(A, w, V);
  const float ps[3] = {(float)w[0], (float)w[1], (float)w[2]};
  const float sfa = sinf(fric), cfa = cosf(fric);
  const float phi_tr = (float)((double)__fadd_rn(__fsub_rn(ps[0], ps[2]), __fmul_rn(__fadd_rn(ps[0], ps[2]), sfa)) -
                               2. * (double)coh_tr * (double)cfa);
  auto store_elastic_eps = [&]() {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) ep[i][j] = (i < D && j < D) ? et[i][j] : 0.f;
    eps = mat_to_sym(ep);
  };
  if (phi_tr <= 0.f) {
    s = mat_to_sym(sig);
    if (hist) store_elastic_eps();
    return;
  }
  const float sda = sinf(dil);
  const double Kd = K, Gd = G, Hd = H, sfad = sfa, cfad = cfa, sdad = sda;
  const float a = (float)(4. * Gd * (1. + (1. / 3.) * sfad * sdad) +
                          (double)__fmul_rn(__fmul_rn(__fmul_rn(4.f, K), sfa), sda));
  const double ad = a;
  // principal stresses back to the frame, the elastic strain of the result
  auto finish = [&](const float pc[3], double acc_next) {
    float out[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) t += V[i][k] * (double)pc[k] * V[j][k];
        out[i][j] = (float)t;
      }
    s = mat_to_sym(out);
    if (hist) {
      const float pn = (float)((double)__fadd_rn(__fadd_rn(out[0][0], out[1][1]), out[2][2]) / 3.0);
      const float i2g = (float)(1. / (2. * Gd)), i3k = (float)(1. / (3. * Kd));
      float e[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          e[i][j] = (i < D && j < D)
                        ? __fadd_rn(__fmul_rn(i == j ? __fsub_rn(out[i][j], pn) : out[i][j], i2g),
                                    i == j ? __fmul_rn(pn, i3k) : 0.f)
                        : 0.f;
      eps = mat_to_sym(e);
      h0 = (float)acc_next;
    }
  };
  auto ordered = [](const float pc[3]) {
    return truncf(pc[0]) >= truncf(pc[1]) && truncf(pc[1]) >= truncf(pc[2]);
  };
  // main plane (mohrcoulomb_inline.h:34-98)
  {
    double dg = 0.0, ph = phi_tr, acc_next = acc_tr;
    const double d = -ad - 4 * Hd * cfad * cfad;
    for (int it = 0; it < kMaxReturnIters; ++it) {
      dg = dg - ph / d;
      acc_next = (double)acc_tr + 2 * cfad * dg;
      const double ca = (double)coh + Hd * acc_next;
      ph = (double)ps[0] - (double)ps[2] + ((double)ps[0] + (double)ps[2]) * sfad - 2. * ca * cfad - ad * dg;
      if (!(fabs(ph) > 1.e-6)) break;
    }
    const float pc[3] = {(float)((double)ps[0] - dg * (2. * Gd * (1. + (1. / 3.) * sdad) + 2. * Kd * sdad)),
                         (float)((double)ps[1] + dg * ((4. / 3.) * Gd - 2. * Kd) * sdad),
                         (float)((double)ps[2] + dg * (2. * Gd * (1. - (1. / 3.) * sdad) - 2. * Kd * sdad))};
    if (ordered(pc)) { finish(pc, acc_next); return; }
  }
  // edges (mohrcoulomb_inline.h:100-229): right, then left
  for (int edge = 0; edge < 2; ++edge) {
    const bool right = edge == 0;
    const float s1 = __fadd_rn(__fsub_rn(ps[0], ps[2]), __fmul_rn(__fadd_rn(ps[0], ps[2]), sfa));
    const float s2 = right ? __fadd_rn(__fsub_rn(ps[0], ps[1]), __fmul_rn(__fadd_rn(ps[0], ps[1]), sfa))
                           : __fadd_rn(__fsub_rn(ps[1], ps[2]), __fmul_rn(__fadd_rn(ps[1], ps[2]), sfa));
    double ph0 = (double)s1 - 2. * (double)coh_tr * cfad, ph1 = (double)s2 - 2. * (double)coh_tr * cfad;
    const double b = right ? 2 * Gd * (1. + sfad + sdad - (1. / 3.) * sfad * sdad) + 4. * Kd * sfad * sdad
                           : 2 * Gd * (1. - sfad - sdad - (1 / 3.) * sfad * sdad) + 4. * Kd * sfad * sdad;
    double g0 = 0.0, g1 = 0.0, acc_next = acc_tr;
    for (int it = 0; it < kMaxReturnIters; ++it) {
      const double d00 = -ad - 4. * Hd * cfad * cfad;
      const double d01 = -b - (double)__fmul_rn(__fmul_rn(__fmul_rn(4.f, H), cfa), cfa);
      // dgamma -= d^-1 Phi, the 2 x 2 inverse formed first
      const double invdet = 1.0 / (d00 * d00 - d01 * d01);
      const double i00 = d00 * invdet, i01 = -d01 * invdet;
      const double n0 = g0 - (i00 * ph0 + i01 * ph1), n1 = g1 - (i01 * ph0 + i00 * ph1);
      g0 = n0;
      g1 = n1;
      acc_next = (double)acc_tr + 2 * cfad * (g0 + g1);
      const double ca = (double)coh + Hd * acc_next;
      ph0 = (double)s1 - 2. * ca * cfad - ad * g0 - b * g1;
      ph1 = (double)s2 - 2. * ca * cfad - ad * g1 - b * g0;
      if (!(fabs(ph0) + fabs(ph1) > 1.e-7)) break;
    }
    const double c1 = 2. * Gd * (1. + (1. / 3.) * sdad) + 2. * Kd * sdad, c2 = ((4. / 3.) * Gd - 2. * Kd) * sdad,
                 c3 = 2. * Gd * (1. - (1. / 3.) * sdad) - 2. * Kd * sdad;
    float pc[3];
    if (right) {
      pc[0] = (float)((double)ps[0] - c1 * (g0 + g1));
      pc[1] = (float)((double)ps[1] + c2 * g0 + c3 * g1);
      pc[2] = (float)((double)ps[2] + c3 * g0 + c2 * g1);
    } else {
      pc[0] = (float)((double)ps[0] - c1 * g0 + c2 * g1);
      pc[1] = (float)((double)ps[1] + c2 * g0 - c1 * g1);
      pc[2] = (float)((double)ps[2] + c3 * (g0 + g1));
    }
    if (ordered(pc)) { finish(pc, acc_next); return; }
  }
  // apex (mohrcoulomb_inline.h:231-270)
  double dv = 0.0, ph = phi_tr, pn = 0.0, acc_next = acc_tr;
  const double al = cfad / sdad, cotf = cfad / sfad;
  for (int it = 0; it < kMaxReturnIters; ++it) {
    acc_next = (double)acc_tr + al * dv;
    const double ca = (double)coh + Hd * acc_next;
    pn = (double)ptr - Kd * dv;
    ph = ca * cotf - pn;
    dv = dv - ph / ((Hd * cfad * cotf) / sdad + Kd);
    if (!(fabs(ph) > 1.e-6)) break;
  }
  const float pf = (float)pn;
  s = Sym{pf, pf, pf, 0.f, 0.f, 0.f};
  if (hist) {
    const float e = (float)(pn * (1. / (3. * Kd)));
    eps = Sym{e, D > 1 ? e : 0.f, D > 2 ? e : 0.f, 0.f, 0.f, 0.f};
    h0 = (float)acc_next;
  }
}

// modifiedcamclay_inline.h:269-413 (update_modifiedcamclay_nonlinear, the
// update modifiedcamclay.cpp:147-176 runs): co-rotated elastic-strain trial
// with pressure-dependent stiffness p = p_prev / (1 - v/kappa de_v), the
// ellipse (p_s - p - Pt)^2 + (q/M)^2 - p_s^2 with back stress
// p_s = (pc + Pt)/2, and a 2 x 2 Newton on (plastic multiplier, volumetric
// plastic strain) in double whose tolerance grows tenfold per iteration
// after the thousandth, as the reference's does. Compression positive in p.
// h0 = volumetric plastic strain, h1 = preconsolidation pressure.
__device__ inline float mcc_yield(float p, float Pt, float ps, float q, float M) {
  const float a = __fsub_rn(__fsub_rn(ps, p), Pt), b = __fdiv_rn(q, M);
  return __fsub_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(ps, ps));
}
template <int D>
__device__ inline void rm_modified_cam_clay(const pmpm_material &m, float dt, const float *L, float vol,
                                            const Sym &sref, Sym &s, Sym &eps, float &h0, float &h1) {
  const float K = m.p[0], G = m.p[1], M = m.p[2], lam = m.p[3], kap = m.p[4], Pt = m.p[5], Vs = m.p[6];
  const bool hist = (m.flags & PMPM_MATF_UPDATE_HISTORY) != 0;
  float de[3][3], ep[3][3];
  sym_to_mat(eps, ep);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) de[i][j] = 0.f;
  if (m.flags & PMPM_MATF_VELGRAD_IS_INCREMENT) {
    strain_increment<D>(m, dt, L, de);
  } else {
    float W[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const bool in = i < D && j < D;
        de[i][j] = in ? 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt : 0.f;
        W[i][j] = in ? 0.5f * (L[3 * i + j] - L[3 * j + i]) * dt : 0.f;
      }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        float we = 0.f, ew = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) { we += W[i][k] * ep[k][j]; ew += ep[i][k] * W[k][j]; }
        de[i][j] += -we + ew;
      }
  }
  const float tr_ref = __fadd_rn(__fadd_rn(sref.xx, sref.yy), sref.zz);
  const float pref = __fmul_rn(-(float)(1.0 / 3.0), tr_ref);
  float sr[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) sr[i][j] = i == j ? __fadd_rn(sym_at(sref, i, j), pref) : sym_at(sref, i, j);
  float et[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) et[i][j] = (i < D && j < D) ? __fadd_rn(ep[i][j], de[i][j]) : 0.f;
  float tre = et[0][0];
  if (D > 1) tre = __fadd_rn(tre, et[1][1]);
  if (D > 2) tre = __fadd_rn(tre, et[2][2]);
  const float evt = -tre;  // compression positive
  const float ev3 = (float)((1.0 / 3.0) * (double)evt);
  float dev[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) dev[i][j] = i == j ? __fadd_rn(et[i][j], ev3) : et[i][j];
  if (D < 3) dev[2][2] = 0.f;
  const float vspec = __fdiv_rn(vol, Vs);
  const float p_prev = (float)(-(double)__fadd_rn(__fadd_rn(s.xx, s.yy), s.zz) / 3.0);
  const float v_kap = __fdiv_rn(vspec, kap);
  float trd = de[0][0];
  if (D > 1) trd = __fadd_rn(trd, de[1][1]);
  if (D > 2) trd = __fadd_rn(trd, de[2][2]);
  const float dev_tr = -trd;
  const float ptr = (float)((double)p_prev / (1.0 - (double)__fmul_rn(v_kap, dev_tr)));
  const float twoG = (float)(2. * (double)G);
  float st[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) st[i][j] = __fadd_rn(sr[i][j], __fmul_rn(twoG, dev[i][j]));
  const float q_tr = (float)sqrt(1.5 * (double)frob2_rows(st));
  const float pc_prev = h1;
  const float ps_tr = (float)(0.5 * (double)__fadd_rn(pc_prev, Pt));
  const float f_tr = mcc_yield(ptr, Pt, ps_tr, q_tr, M);
  if (f_tr <= 0.f) {
    float o[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) o[i][j] = i == j ? __fsub_rn(st[i][j], ptr) : st[i][j];
    s = mat_to_sym(o);
    if (hist) eps = mat_to_sym(et);
    return;
  }
  const float epv_prev = h0;
  double x0 = 0.0, x1 = epv_prev;  // plastic multiplier, volumetric plastic strain
  double p_next = 0.0, q_next = 0.0, ps_next = 0.0, tol = 1e-2, sfac_s = 1.0;
  float devn = 0.f;
  const double vlt = (double)vspec / (double)__fsub_rn(lam, kap);
  const double Md = M, Gd = G, Kd = K;
  // M*M and 6*G are float products in the reference's expressions
  const double MM = (double)__fmul_rn(M, M), G6 = (double)__fmul_rn(6.f, G), G12 = (double)__fmul_rn(-12.f, G);
  for (int counter = 0; counter < 4096; ++counter) {
    const double mult = x0, epv = x1;
    const float dpv = (float)(epv - (double)epv_prev);
    devn = __fsub_rn(dev_tr, dpv);
    p_next = (double)p_prev / (1.0 - (double)__fmul_rn(v_kap, devn));
    q_next = (MM / (MM + G6 * mult)) * (double)q_tr;
    sfac_s = MM / (MM + 6.0 * Gd * mult);  // the deviator's factor (6.0 * G in double there)
    ps_next = 0.5 * ((double)pc_prev * (1.0 + vlt * (double)dpv) + (double)Pt);
    const double r0 = mcc_yield((float)p_next, Pt, (float)ps_next, (float)q_next, M);
    const double r1 = epv - (double)epv_prev + 2.0 * mult * (ps_next - p_next - (double)Pt);
    const double Hh = 0.5 * vlt * ((double)pc_prev / pow(1.0 - vlt * (double)dpv, 2));
    const double pover = ps_next - p_next - (double)Pt;
    const double d00 = (G12 / (MM + G6 * mult)) * (q_next / Md) * (q_next / Md);
    const double d01 = (2.0 * pover) * (Kd + Hh) - 2.0 * ps_next * Hh;
    const double d10 = 2.0 * pover;
    const double d11 = 1.0 + (2.0 * mult) * (Kd + Hh);
    // Solution -= d^-1 R with the 2 x 2 inverse formed first (cofactors over
    // the determinant)
    const double invdet = 1.0 / (d00 * d11 - d10 * d01);
    const double i00 = d11 * invdet, i10 = -d10 * invdet, i01 = -d01 * invdet, i11 = d00 * invdet;
    x0 = x0 - (i00 * r0 + i01 * r1);
    x1 = x1 - (i10 * r0 + i11 * r1);
    const double conv = sqrt(r0 * r0 + r1 * r1);
    if (counter > 1000) tol *= 10;
    if (!(fabs(conv) > tol)) break;
  }
  // s_next from the last iterate's scale (the reference keeps it from inside
  // the loop), sigma = s - p I
  const float sf = (float)sfac_s;
  float sn[3][3], o[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      sn[i][j] = __fmul_rn(sf, st[i][j]);
      o[i][j] = i == j ? (float)((double)sn[i][j] - p_next) : sn[i][j];
    }
  s = mat_to_sym(o);
  if (hist) {
    float tre0 = ep[0][0];
    if (D > 1) tre0 = __fadd_rn(tre0, ep[1][1]);
    if (D > 2) tre0 = __fadd_rn(tre0, ep[2][2]);
    const float vterm = (float)((1. / 3.) * (double)__fadd_rn(tre0, devn));
    const float g2 = (float)(2.0 * Gd);
    float e[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        e[i][j] = (i < D && j < D) ? __fsub_rn(__fdiv_rn(__fsub_rn(sn[i][j], sr[i][j]), g2), i == j ? vterm : 0.f)
                                   : 0.f;
    eps = mat_to_sym(e);
    h0 = (float)x1;
    h1 = (float)(2.0 * ps_next - (double)Pt);
  }
}

// The materials of the post-G2P pass, by the particle's colour.
template <int D>
__device__ inline void return_map_ext(const pmpm_params &prm, int color, const float *L, float vol, const Sym &sref,
                                      Sym &s, Sym &eps, float &h0, float &h1) {
  if (color >= prm.num_materials) return;
  const pmpm_material &m = prm.materials[color];
  switch (m.kind) {
    case PMPM_MAT_VON_MISES:
      if (D == 3) rm_von_mises(m, prm.dt, L, s, eps, h0);  // 3D only (vonmises.cpp:46-49)
      break;
    case PMPM_MAT_MOHR_COULOMB:
      rm_mohr_coulomb<D>(m, prm.dt, L, s, eps, h0);
      break;
    case PMPM_MAT_MODIFIED_CAM_CLAY:
      rm_modified_cam_clay<D>(m, prm.dt, L, vol, sref, s, eps, h0, h1);
      break;
    default:
      break;
  }
}

// Dispatch on the material of the particle's colour (solver.cpp:110-121: each
// material updates only particles whose colour equals its list index).
template <int D>
__device__ inline void return_map(const pmpm_params &prm, int color, uint8_t pflags, const float *L,
                                  float mass, float vol, float vol0, const Sym &sref, Sym &s, Sym &eps) {
  if (color >= prm.num_materials) return;
  const pmpm_material &m = prm.materials[color];
  const bool active = (pflags & PMPM_PF_ACTIVE) != 0;
  switch (m.kind) {
    case PMPM_MAT_LINEAR_ELASTIC:
      if (active) rm_linear_elastic<D>(m, prm.dt, L, s);
      break;
    case PMPM_MAT_NEWTON_FLUID:
      if (active) rm_newton_fluid<D>(m, L, mass, vol, vol0, s);
      break;
    case PMPM_MAT_LOCAL_RHEO:
      rm_local_rheo(m, prm.dt, L, mass, vol, s);
      break;
    case PMPM_MAT_DRUCKER_PRAGER:
      rm_drucker_prager<D>(m, prm.dt, L, sref, s, eps);
      break;
    default:
      break;
  }
}

// True when the step has a PLANAR_DOMAIN condition (a particle force).
__host__ __device__ inline bool has_planar(const pmpm_params &prm) {
  for (int k = 0; k < prm.num_bcs; ++k)
    if (prm.bcs[k].kind == PMPM_BC_PLANAR_DOMAIN) return true;
  return false;
}

// External force of one particle from the PLANAR_DOMAIN conditions
// (planardomain_inline.h:56-138, applied before P2G: usl.cpp:79-84): a
// particle of radius 0.5 V^(1/D) that overlaps a wall of the box
// [origin, end] by o > 0 gets (m / dt^2) o (n - mu (v - (n.v) n)), n the
// wall's inward normal and mu the wall's friction parameter. The scalar
// (m / dt^2) o is formed in double and rounded to float before it scales the
// vector, as the reference's double * float-vector product does. f is added
// to (the caller zeroes it).
template <int D>
__device__ inline void planar_force(const pmpm_params &prm, const pmpm_grid &g, const float *x, const float *v,
                                    float vol, float m, float *f) {
  for (int k = 0; k < prm.num_bcs; ++k) {
    const pmpm_bc &bc = prm.bcs[k];
    if (bc.kind != PMPM_BC_PLANAR_DOMAIN) continue;
    const float inv_dim = 1.f / (float)D;
    const float radius = 0.5f * powf(vol, inv_dim);
    const double s0 = (double)m / ((double)prm.dt * (double)prm.dt);
#pragma unroll
    for (int face = 0; face < 2; ++face) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        const float gap = face == 0 ? __fsub_rn(x[i], g.origin[i]) : __fsub_rn(bc.f[6 + i], x[i]);
        const float ov = __fsub_rn(radius, gap);
        if (!(ov > 0.f)) continue;
        const float mu = bc.f[3 * face + i];
        const float sc = (float)(s0 * (double)ov);
#pragma unroll
        for (int a = 0; a < D; ++a) {
          // n - mu (v - (n.v) n): the normal component of the bracket is 0
          const float t = a == i ? (face == 0 ? 1.f : -1.f) : __fsub_rn(0.f, __fmul_rn(mu, v[a]));
          f[a] = __fadd_rn(f[a], __fmul_rn(sc, t));
        }
      }
    }
  }
}

template <int D> __device__ inline float det_dd(const float *F) {
  if (D == 1) return F[0];
  if (D == 2) return F[0] * F[3] - F[1] * F[2];
  return F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
         F[2] * (F[3] * F[7] - F[4] * F[6]);
}

__device__ inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- the linear-index wrap-around (usl_inline.h:122-126: P2G bound-checks
// only the linear node index, so a node vector outside the grid on x, or on y
// in 3D, can land on an in-grid node).
//
// Cells whose window (node offsets -1 .. 2) can wrap into the grid: x in
// {0, nx-2, nx-1} (any y, z) and, in 3D, y in {0, ny-2, ny-1} (other x).
// Offsets outside the grid along the slowest axis always leave the index
// range, so no other cell wraps. wrap_cell maps w in [0, wrap_cell_count) to
// its cell (false for the duplicates of tiny grids), wrap_index inverts it.
template <int D> __host__ __device__ inline bool wrap_cell(const pmpm_grid &g, int w, int *c) {
  const int nx = g.num_cells[0], ny = D > 1 ? g.num_cells[1] : 1, nz = D > 2 ? g.num_cells[2] : 1;
  if (D < 2) return false;
  auto pick = [](int i, int len, int &v) {  // i-th of {0, len-2, len-1}, each cell once
    v = i == 0 ? 0 : len - 3 + i;
    return i == 0 || v > 0;
  };
  const int nxs = 3 * ny * (D > 2 ? nz : 1);
  if (w < nxs) {
    c[1] = (w / 3) % ny;
    c[2] = D > 2 ? w / (3 * ny) : 0;
    return pick(w % 3, nx, c[0]);
  }
  if (D < 3) return false;
  w -= nxs;
  const int mid = nx - 3;  // x in [1, nx-3]
  if (mid <= 0 || w >= 3 * mid * nz) return false;
  c[0] = 1 + (w / 3) % mid;
  c[2] = w / (3 * mid);
  return pick(w % 3, ny, c[1]);
}
template <int D> __host__ __device__ inline int wrap_cell_count(const pmpm_grid &g) {
  const int nx = g.num_cells[0], ny = D > 1 ? g.num_cells[1] : 1, nz = D > 2 ? g.num_cells[2] : 1;
  if (D < 2) return 0;
  return 3 * ny * (D > 2 ? nz : 1) + (D > 2 && nx > 3 ? 3 * (nx - 3) * nz : 0);
}
template <int D> __host__ __device__ inline int wrap_index(const pmpm_grid &g, const int *c) {
  if (D < 2) return -1;
  const int nx = g.num_cells[0], ny = g.num_cells[1], nz = D > 2 ? g.num_cells[2] : 1;
  auto slot = [](int v, int len) { return v == 0 ? 0 : (v >= len - 2 ? v - (len - 3) : -1); };
  const int i = slot(c[0], nx);
  if (i >= 0) return i + 3 * (c[1] + ny * (D > 2 ? c[2] : 0));
  if (D < 3) return -1;
  const int j = slot(c[1], ny);
  if (j < 0) return -1;
  return 3 * ny * nz + j + 3 * ((c[0] - 1) + (nx - 3) * c[2]);
}
// Nodes the wrap-around can reach: x in {0, 1, nx-1} (any y, z) and, in 3D,
// y in {0, 1, ny-1} (x in [2, nx-2]). wrap_target maps t in
// [0, wrap_target_count) to its node coordinates (false for the duplicates
// of tiny grids).
template <int D> __host__ __device__ inline bool wrap_target(const pmpm_grid &g, int t, int *ids) {
  const int nx = g.num_cells[0], ny = D > 1 ? g.num_cells[1] : 1, nz = D > 2 ? g.num_cells[2] : 1;
  if (D < 2) return false;
  auto pick = [](int i, int len, int &v) {  // i-th of {0, 1, len-1}, each node once
    v = i == 0 ? 0 : (i == 1 ? 1 : len - 1);
    return i == 0 || (i == 1 ? len > 1 : len - 1 > 1);
  };
  const int nxs = 3 * ny * (D > 2 ? nz : 1);
  if (t < nxs) {
    ids[1] = (t / 3) % ny;
    ids[2] = D > 2 ? t / (3 * ny) : 0;
    return pick(t % 3, nx, ids[0]);
  }
  if (D < 3) return false;
  t -= nxs;
  const int mid = nx - 3;  // x in [2, nx-2]
  if (mid <= 0 || t >= 3 * mid * nz) return false;
  ids[0] = 2 + (t / 3) % mid;
  ids[2] = t / (3 * mid);
  return pick(t % 3, ny, ids[1]);
}
template <int D> __host__ __device__ inline int wrap_target_count(const pmpm_grid &g) {
  return wrap_cell_count<D>(g);  // the same slab sizes
}

}  // namespace pmpm
