// adjoint_math.hpp -- scalar-generic return maps and their vector-Jacobian
// products, for host and device.
//
// The return maps are restated here over a scalar type T so that T = Dual
// (value + one tangent) gives exact directional derivatives of the same
// floating-point program; branch decisions read the value part only, so the
// active set (elastic / cone / apex) is the one the forward step took. A
// vector-Jacobian product is assembled from one tangent pass per input
// direction. Stresses and elastic strains are 6-vectors (xx, yy, zz, xy, xz,
// yz); an off-diagonal direction moves both matrix entries, which is the
// adjoint of the symmetric storage.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PMPM_HD __host__ __device__
#else
#define PMPM_HD
#endif

namespace pmpm {

struct Dual {
  float v, d;
};
PMPM_HD inline Dual mkd(float v, float d = 0.f) { return Dual{v, d}; }
PMPM_HD inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
PMPM_HD inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
PMPM_HD inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
PMPM_HD inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
PMPM_HD inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
PMPM_HD inline Dual operator+(Dual a, float b) { return {a.v + b, a.d}; }
PMPM_HD inline Dual operator+(float a, Dual b) { return {a + b.v, b.d}; }
PMPM_HD inline Dual operator-(Dual a, float b) { return {a.v - b, a.d}; }
PMPM_HD inline Dual operator-(float a, Dual b) { return {a - b.v, -b.d}; }
PMPM_HD inline Dual operator*(Dual a, float b) { return {a.v * b, a.d * b}; }
PMPM_HD inline Dual operator*(float a, Dual b) { return {a * b.v, a * b.d}; }
PMPM_HD inline Dual operator/(Dual a, float b) { return {a.v / b, a.d / b}; }
PMPM_HD inline Dual operator/(float a, Dual b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
PMPM_HD inline Dual gsqrt(Dual a) {
  const float s = sqrtf(a.v);
  return {s, s > 0.f ? a.d / (2.f * s) : 0.f};
}
PMPM_HD inline float gsqrt(float a) { return sqrtf(a); }
PMPM_HD inline double gsqrt(double a) { return sqrt(a); }
PMPM_HD inline float val(float a) { return a; }
PMPM_HD inline double val(double a) { return a; }
PMPM_HD inline float val(Dual a) { return a.v; }

// index of (i, j) in the 6-vector
PMPM_HD inline int sym_idx(int i, int j) {
  if (i == j) return i;
  const int k = i + j;  // 1 -> xy, 2 -> xz, 3 -> yz
  return k == 1 ? 3 : (k == 2 ? 4 : 5);
}

// Linear elastic (linearelastic_inline.h:40-107): s += (lam + 2G) deps on
// the D x D block, deps = 0.5 (L + L^T) dt.
template <class T, int D>
PMPM_HD void le_map(T lam, T G, float dt, const T *L, T *s6) {
  const T twoG = 2.f * G;
  for (int i = 0; i < D; ++i)
    for (int j = i; j < D; ++j) {
      const T de = 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt;
      T &s = s6[sym_idx(i, j)];
      s = s + (lam * de + twoG * de);
    }
}

// Drucker-Prager with elastic-strain history (dp_rheo_inline.h:43-290), the
// same operations as the forward kernel. p = {K, G, M, Pt}; etab and sref
// are constants. eps6 in, s6 / eps6_out out (eps6_out only when `history`).
template <class T, int D>
PMPM_HD void dp_map(const T *p, float etab, bool increment, bool history, bool apex_guard, float dt, const T *L,
                    const float *sref6, const T *eps6, T *s6, T *eps6_out) {
  const T K = p[0], G = p[1], M = p[2], Pt = p[3];
  T de[3][3], ep[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      de[i][j] = T{};
      ep[i][j] = (i < D && j < D) ? eps6[sym_idx(i, j)] : T{};
    }
  if (increment) {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) de[i][j] = L[3 * i + j];
  } else {
    T W[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        de[i][j] = 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt;
        W[i][j] = 0.5f * (L[3 * i + j] - L[3 * j + i]) * dt;
      }
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        T we = T{}, ew = T{};
        for (int k = 0; k < D; ++k) {
          we = we + W[i][k] * ep[k][j];
          ew = ew + ep[i][k] * W[k][j];
        }
        de[i][j] = de[i][j] + (-we + ew);
      }
  }
  const float pref = (float)((1.0 / 3.0) * (double)(sref6[0] + sref6[1] + sref6[2]));
  float sr[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) sr[i][j] = sref6[sym_idx(i, j)] - (i == j ? pref : 0.f);
  T et[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) et[i][j] = (i < D && j < D) ? ep[i][j] + de[i][j] : T{};
  T ev = et[0][0];
  if (D > 1) ev = ev + et[1][1];
  if (D > 2) ev = ev + et[2][2];
  const T ptr = pref + K * ev;
  const T ev3 = (1.f / 3.f) * ev;
  T str[3][3];
  T f2 = T{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T dev = et[i][j] - (i == j ? ev3 : T{});
      if (D < 3 && i == 2 && j == 2) dev = T{};
      str[i][j] = sr[i][j] + 2.f * G * dev;
      f2 = f2 + str[i][j] * str[i][j];
    }
  const T J2 = 0.5f * f2;
  const T eta = 0.57735026918962576f * M;
  const T sqJ2 = gsqrt(J2);
  const T Phi = sqJ2 + eta * (ptr - Pt);
  if (val(Phi) <= 0.f) {
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) s6[sym_idx(i, j)] = str[i][j] + (i == j ? ptr : T{});
    if (history)
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) eps6_out[sym_idx(i, j)] = (i < D && j < D) ? et[i][j] : T{};
    return;
  }
  const T dg = Phi / (K * eta * etab + G);
  T pn = ptr - K * etab * dg;
  const T f = 1.f - (G * dg) / sqJ2;
  T sn[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) sn[i][j] = f * str[i][j];
  if (val(sqJ2) - val(G * dg) < 0.f) {  // apex
    if (apex_guard && etab == 0.f) {
      pn = Pt;
    } else {
      const T dga = (ptr - Pt) / (etab * K);
      pn = ptr - K * etab * dga;
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) sn[i][j] = T{};
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) s6[sym_idx(i, j)] = sn[i][j] + (i == j ? pn : T{});
  if (history) {
    const T i2G = 1.f / (2.f * G);
    const T vol = (pn - pref) / (3.f * K);
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j)
        eps6_out[sym_idx(i, j)] =
            (i < D && j < D) ? (sn[i][j] - sr[i][j]) * i2G + (i == j ? vol : T{}) : T{};
  }
}

// Vector-Jacobian products. Inputs: the forward inputs and the adjoints of
// the outputs (sbar6, ebar6); results are added to the input adjoints.
//   dp: d/d eps6 (6), d/d L (9), d/d p (4: K, G, M, Pt)
template <int D>
PMPM_HD void dp_vjp(const float *p, float etab, bool increment, bool history, bool apex_guard, float dt,
                    const float *L, const float *sref6, const float *eps6, const float *sbar6, const float *ebar6,
                    float *eps_bar6, float *Lbar, float *pbar) {
  for (int k = 0; k < 19; ++k) {
    Dual pd[4], Ld[9], ed[6], so[6], eo[6];
    for (int i = 0; i < 4; ++i) pd[i] = mkd(p[i], k == 15 + i ? 1.f : 0.f);
    for (int i = 0; i < 9; ++i) Ld[i] = mkd(L[i], k == 6 + i ? 1.f : 0.f);
    for (int i = 0; i < 6; ++i) ed[i] = mkd(eps6[i], k == i ? 1.f : 0.f);
    for (int i = 0; i < 6; ++i) eo[i] = mkd(0.f);
    dp_map<Dual, D>(pd, etab, increment, history, apex_guard, dt, Ld, sref6, ed, so, eo);
    float acc = 0.f;
    for (int i = 0; i < 6; ++i) acc += so[i].d * sbar6[i] + (history ? eo[i].d * ebar6[i] : 0.f);
    if (k < 6) eps_bar6[k] += acc;
    else if (k < 15) Lbar[k - 6] += acc;
    else pbar[k - 15] += acc;
  }
}

// The same VJP as dp_vjp in one reverse sweep: dp_map's float program is
// run forward once, keeping its intermediates, and its adjoint is propagated
// backwards through the branch the forward took (elastic, cone return, apex).
// Costs a few forward maps instead of one dual-number pass per input.
template <int D>
PMPM_HD void dp_vjp_rev(const float *p, float etab, bool increment, bool history, bool apex_guard, float dt,
                        const float *L, const float *sref6, const float *eps6, const float *sbar6,
                        const float *ebar6, float *eps_bar6, float *Lbar, float *pbar) {
  const float K = p[0], G = p[1], M = p[2], Pt = p[3];
  // ---- forward (dp_map's operations, T = float)
  float ep[3][3], de[3][3], W[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ep[i][j] = (i < D && j < D) ? eps6[sym_idx(i, j)] : 0.f;
      de[i][j] = 0.f;
      W[i][j] = 0.f;
    }
  if (increment) {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) de[i][j] = L[3 * i + j];
  } else {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        de[i][j] = 0.5f * (L[3 * i + j] + L[3 * j + i]) * dt;
        W[i][j] = 0.5f * (L[3 * i + j] - L[3 * j + i]) * dt;
      }
    float rot[3][3];
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        float we = 0.f, ew = 0.f;
        for (int k = 0; k < D; ++k) {
          we = we + W[i][k] * ep[k][j];
          ew = ew + ep[i][k] * W[k][j];
        }
        rot[i][j] = -we + ew;
      }
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) de[i][j] = de[i][j] + rot[i][j];
  }
  const float pref = (float)((1.0 / 3.0) * (double)(sref6[0] + sref6[1] + sref6[2]));
  float sr[3][3], et[3][3], dev[3][3], str[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      sr[i][j] = sref6[sym_idx(i, j)] - (i == j ? pref : 0.f);
      et[i][j] = (i < D && j < D) ? ep[i][j] + de[i][j] : 0.f;
    }
  float ev = et[0][0];
  if (D > 1) ev = ev + et[1][1];
  if (D > 2) ev = ev + et[2][2];
  const float ptr = pref + K * ev;
  const float ev3 = (1.f / 3.f) * ev;
  float f2 = 0.f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      dev[i][j] = et[i][j] - (i == j ? ev3 : 0.f);
      if (D < 3 && i == 2 && j == 2) dev[i][j] = 0.f;
      str[i][j] = sr[i][j] + 2.f * G * dev[i][j];
      f2 = f2 + str[i][j] * str[i][j];
    }
  const float J2 = 0.5f * f2;
  const float eta = (float)(0.57735026918962576 * (double)M);
  const float sqJ2 = sqrtf(J2);
  const float Phi = sqJ2 + eta * (ptr - Pt);
  // ---- reverse
  float strb[3][3] = {}, etb[3][3] = {}, epb[3][3] = {}, deb[3][3] = {};
  float ptrb = 0.f, Kb = 0.f, Gb = 0.f, Mb = 0.f, Ptb = 0.f, sqJ2b = 0.f;
  if (Phi <= 0.f) {  // elastic: s = str + ptr I, eps_out = et
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        strb[i][j] += sbar6[sym_idx(i, j)];
        if (i == j) ptrb += sbar6[i];
        if (history && i < D && j < D) etb[i][j] += ebar6[sym_idx(i, j)];
      }
  } else {
    const float den = K * eta * etab + G;
    const float dg = Phi / den;
    float pn = ptr - K * etab * dg;
    const float f = 1.f - (G * dg) / sqJ2;
    // the branch the forward map took: its apex test runs in double
    // (rm_drucker_prager)
    const double dgd = (double)Phi / ((double)K * (double)eta * (double)etab + (double)G);
    const bool apex = (double)sqJ2 - (double)G * dgd < 0.0;
    float dga = 0.f;
    if (apex) {
      if (apex_guard && etab == 0.f) {
        pn = Pt;
      } else {
        dga = (ptr - Pt) / (etab * K);
        pn = ptr - K * etab * dga;
      }
    }
    // outputs: s = sn + pn I (upper triangle), eps_out = (sn - sr)/(2G) + vol I
    float snb[3][3] = {}, pnb = 0.f;
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        snb[i][j] += sbar6[sym_idx(i, j)];
        if (i == j) pnb += sbar6[i];
      }
    if (history) {
      const float i2G = 1.f / (2.f * G);
      float i2Gb = 0.f, volb = 0.f;
      for (int i = 0; i < D; ++i)
        for (int j = i; j < D; ++j) {
          const float e = ebar6[sym_idx(i, j)];
          const float sn = apex ? 0.f : f * str[i][j];
          snb[i][j] += e * i2G;
          i2Gb += e * (sn - sr[i][j]);
          if (i == j) volb += e;
        }
      Gb += i2Gb * (-2.f * i2G * i2G);
      pnb += volb / (3.f * K);
      Kb += -volb * (pn - pref) / (3.f * K * K);
    }
    float dgb = 0.f;
    if (!apex) {
      // sn = f str; f = 1 - G dg / sqJ2; pn = ptr - K etab dg
      float fb = 0.f;
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
          fb += snb[i][j] * str[i][j];
          strb[i][j] += f * snb[i][j];
        }
      Gb += -fb * dg / sqJ2;
      dgb += -fb * G / sqJ2;
      sqJ2b += fb * G * dg / (sqJ2 * sqJ2);
      ptrb += pnb;
      Kb += -pnb * etab * dg;
      dgb += -pnb * K * etab;
    } else if (apex_guard && etab == 0.f) {
      Ptb += pnb;
    } else {
      // pn = ptr - K etab dga, dga = (ptr - Pt) / (etab K)
      ptrb += pnb;
      Kb += -pnb * etab * dga;
      const float dgab = -pnb * K * etab;
      const float q = 1.f / (etab * K);
      ptrb += dgab * q;
      Ptb -= dgab * q;
      Kb += -dgab * (ptr - Pt) * q / K;
    }
    // dg = Phi / den, den = K eta etab + G, Phi = sqJ2 + eta (ptr - Pt)
    const float Phib = dgb / den;
    const float denb = -dgb * Phi / (den * den);
    Kb += denb * eta * etab;
    float etab_b = denb * K * etab;
    Gb += denb;
    sqJ2b += Phib;
    etab_b += Phib * (ptr - Pt);
    ptrb += Phib * eta;
    Ptb -= Phib * eta;
    Mb += 0.57735026918962576f * etab_b;
    // sqJ2 = sqrt(f2 / 2)
    if (sqJ2 > 0.f) {
      const float f2b = 0.5f * (sqJ2b / (2.f * sqJ2));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) strb[i][j] += 2.f * str[i][j] * f2b;
    }
  }
  // str = sr + 2G dev; dev = et - ev3 I (zz fixed at 0 below 3D)
  float ev3b = 0.f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (D < 3 && i == 2 && j == 2) continue;
      Gb += 2.f * dev[i][j] * strb[i][j];
      const float devb = 2.f * G * strb[i][j];
      if (i < D && j < D) etb[i][j] += devb;
      if (i == j) ev3b -= devb;
    }
  float evb = (1.f / 3.f) * ev3b;
  Kb += ptrb * ev;
  evb += ptrb * K;
  for (int i = 0; i < D; ++i) etb[i][i] += evb;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      epb[i][j] += etb[i][j];
      deb[i][j] += etb[i][j];
    }
  if (increment) {
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) Lbar[3 * i + j] += deb[i][j];
  } else {
    float Wb[3][3] = {};
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        const float r = deb[i][j];
        for (int k = 0; k < D; ++k) {
          Wb[i][k] -= r * ep[k][j];
          epb[k][j] -= r * W[i][k];
          epb[i][k] += r * W[k][j];
          Wb[k][j] += r * ep[i][k];
        }
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const float a = 0.5f * dt * deb[i][j], b = 0.5f * dt * Wb[i][j];
        Lbar[3 * i + j] += a + b;
        Lbar[3 * j + i] += a - b;
      }
  }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) eps_bar6[sym_idx(i, j)] += epb[i][j];
  pbar[0] += Kb;
  pbar[1] += Gb;
  pbar[2] += Mb;
  pbar[3] += Ptb;
}

//   le: d/d s6 (6), d/d L (9), d/d p (2: lam, G)
template <int D>
PMPM_HD void le_vjp(const float *p, float dt, const float *L, const float *s6, const float *sbar6, float *s_bar6,
                    float *Lbar, float *pbar) {
  for (int k = 0; k < 17; ++k) {
    Dual lam = mkd(p[0], k == 15 ? 1.f : 0.f), G = mkd(p[1], k == 16 ? 1.f : 0.f);
    Dual Ld[9], sd[6];
    for (int i = 0; i < 9; ++i) Ld[i] = mkd(L[i], k == 6 + i ? 1.f : 0.f);
    for (int i = 0; i < 6; ++i) sd[i] = mkd(s6[i], k == i ? 1.f : 0.f);
    le_map<Dual, D>(lam, G, dt, Ld, sd);
    float acc = 0.f;
    for (int i = 0; i < 6; ++i) acc += sd[i].d * sbar6[i];
    if (k < 6) s_bar6[k] += acc;
    else if (k < 15) Lbar[k - 6] += acc;
    else pbar[k - 15] += acc;
  }
}

}  // namespace pmpm
