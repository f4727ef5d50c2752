// CPU check of the return-map vector-Jacobian products in adjoint_math.hpp:
// the dual-number VJP (float) against central finite differences of the same
// map evaluated in double, on random states in each branch (elastic, cone,
// apex) of Drucker-Prager and for linear elasticity. Prints one line per
// case: "<name> <max relative error>".
#include <cmath>
#include <cstdio>
#include <random>

#include "adjoint_math.hpp"

using namespace pmpm;

template <int D>
static void dp_case(const char *name, std::mt19937 &rng, double strain_scale, double tension, bool increment = false,
                    bool history = true, bool apex_guard = true, float etab = 0.f) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  float p[4] = {833.3f, 384.6f, 1.2f, (float)tension};  // K, G, M, Pt
  float L[9] = {0}, eps[6] = {0}, sref[6] = {0};
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) L[3 * i + j] = (float)(u(rng) * 50.0);
  for (int i = 0; i < 6; ++i) eps[i] = (float)(u(rng) * strain_scale);
  for (int i = 0; i < 3; ++i) eps[i] -= (float)(0.5 * strain_scale);  // mostly compressive
  if (D < 3) eps[2] = eps[4] = eps[5] = 0.f;
  if (D < 2) eps[1] = eps[3] = 0.f;
  float sbar[6], ebar[6];
  for (int i = 0; i < 6; ++i) { sbar[i] = (float)u(rng); ebar[i] = (float)u(rng); }
  const float dt = 1e-3f;
  if (increment)
    for (int k = 0; k < 9; ++k) L[k] *= dt;  // a strain increment, not a rate
  float eps_bar[6] = {0}, Lbar[9] = {0}, pbar[4] = {0};
  dp_vjp<D>(p, etab, increment, history, apex_guard, dt, L, sref, eps, sbar, ebar, eps_bar, Lbar, pbar);
  // the reverse sweep, from pre-filled adjoints (it adds)
  float eps_bar_r[6], Lbar_r[9], pbar_r[4];
  for (int i = 0; i < 6; ++i) eps_bar_r[i] = 0.25f;
  for (int i = 0; i < 9; ++i) Lbar_r[i] = 0.25f;
  for (int i = 0; i < 4; ++i) pbar_r[i] = 0.25f;
  dp_vjp_rev<D>(p, etab, increment, history, apex_guard, dt, L, sref, eps, sbar, ebar, eps_bar_r, Lbar_r, pbar_r);
  // finite differences of the double map along each input
  auto f = [&](const double *pp, const double *LL, const double *ee) {
    double so[6], eo[6] = {0};
    dp_map<double, D>(pp, etab, increment, history, apex_guard, dt, LL, sref, ee, so, eo);
    double acc = 0;
    for (int i = 0; i < 6; ++i) acc += so[i] * sbar[i] + eo[i] * ebar[i];
    return acc;
  };
  double P[4], LL[9], E[6];
  for (int i = 0; i < 4; ++i) P[i] = p[i];
  for (int i = 0; i < 9; ++i) LL[i] = L[i];
  for (int i = 0; i < 6; ++i) E[i] = eps[i];
  double worst = 0, scale = 1e-30;
  double fd[19], ad[19], rv[19];
  for (int k = 0; k < 19; ++k) {
    double *x = k < 6 ? &E[k] : (k < 15 ? &LL[k - 6] : &P[k - 15]);
    const double x0 = *x, h = 1e-6 * (fabs(x0) > 1e-3 ? fabs(x0) : 1e-3);
    *x = x0 + h;
    const double fp = f(P, LL, E);
    *x = x0 - h;
    const double fm = f(P, LL, E);
    *x = x0;
    fd[k] = (fp - fm) / (2 * h);
    ad[k] = k < 6 ? eps_bar[k] : (k < 15 ? Lbar[k - 6] : pbar[k - 15]);
    rv[k] = (k < 6 ? eps_bar_r[k] : (k < 15 ? Lbar_r[k - 6] : pbar_r[k - 15])) - 0.25;
    scale = fmax(scale, fabs(fd[k]));
  }
  double worst_rev = 0;
  for (int k = 0; k < 19; ++k) {
    worst = fmax(worst, fabs(fd[k] - ad[k]) / scale);
    worst_rev = fmax(worst_rev, fabs(fd[k] - rv[k]) / scale);
  }
  // which branch the state took: from the returned stress, Phi < 0 inside
  // the cone, Phi = 0 on it, zero deviator at the apex
  double so[6], eo[6];
  dp_map<double, D>(P, etab, increment, history, apex_guard, dt, LL, sref, E, so, eo);
  const double pm = (so[0] + so[1] + so[2]) / 3.0;
  double j2 = 0;
  for (int i = 0; i < 3; ++i) j2 += 0.5 * (so[i] - pm) * (so[i] - pm);
  for (int i = 3; i < 6; ++i) j2 += so[i] * so[i];
  const double phi = sqrt(j2) + 0.57735026918962576 * P[2] * (pm - P[3]);
  const char *branch = j2 < 1e-20 ? "apex" : (fabs(phi) < 1e-6 * (1 + fabs(pm)) ? "cone" : "elastic");
  printf("%s %s dual %.3e reverse %.3e %.3e\n", name, branch, worst, worst_rev, fmax(worst, worst_rev));
}

template <int D>
static void le_case(const char *name, std::mt19937 &rng) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  float p[2] = {576.9f, 384.6f};
  float L[9] = {0}, s[6] = {0}, sbar[6];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) L[3 * i + j] = (float)(u(rng) * 5.0);
  for (int i = 0; i < 6; ++i) { s[i] = (float)(u(rng) * 10.0); sbar[i] = (float)u(rng); }
  const float dt = 1e-3f;
  float s_bar[6] = {0}, Lbar[9] = {0}, pbar[2] = {0};
  le_vjp<D>(p, dt, L, s, sbar, s_bar, Lbar, pbar);
  auto f = [&](const double *pp, const double *LL, const double *ss) {
    double so[6];
    for (int i = 0; i < 6; ++i) so[i] = ss[i];
    le_map<double, D>(pp[0], pp[1], dt, LL, so);
    double acc = 0;
    for (int i = 0; i < 6; ++i) acc += so[i] * sbar[i];
    return acc;
  };
  double P[2] = {p[0], p[1]}, LL[9], S[6];
  for (int i = 0; i < 9; ++i) LL[i] = L[i];
  for (int i = 0; i < 6; ++i) S[i] = s[i];
  double worst = 0, scale = 1e-30, fd[17], ad[17];
  for (int k = 0; k < 17; ++k) {
    double *x = k < 6 ? &S[k] : (k < 15 ? &LL[k - 6] : &P[k - 15]);
    const double x0 = *x, h = 1e-6 * (fabs(x0) > 1e-3 ? fabs(x0) : 1e-3);
    *x = x0 + h;
    const double fp = f(P, LL, S);
    *x = x0 - h;
    const double fm = f(P, LL, S);
    *x = x0;
    fd[k] = (fp - fm) / (2 * h);
    ad[k] = k < 6 ? s_bar[k] : (k < 15 ? Lbar[k - 6] : pbar[k - 15]);
    scale = fmax(scale, fabs(fd[k]));
  }
  for (int k = 0; k < 17; ++k) worst = fmax(worst, fabs(fd[k] - ad[k]) / scale);
  printf("%s %.3e\n", name, worst);
}

int main() {
  std::mt19937 rng(7);
  for (int r = 0; r < 5; ++r) {
    dp_case<3>("dp3_small", rng, 1e-5, 0.0);  // small strains
    dp_case<3>("dp3_large", rng, 5e-2, 0.0);  // large strains
    dp_case<2>("dp2_large", rng, 5e-2, 0.0);
    dp_case<3>("dp3_tension", rng, 5e-2, 5.0);  // with tension cut-off Pt > 0
    dp_case<3>("dp3_increment", rng, 5e-2, 0.0, true);  // L is the strain increment
    dp_case<3>("dp3_no_history", rng, 5e-2, 0.0, false, false);
    dp_case<3>("dp3_dilatant", rng, 5e-2, 2.0, false, true, false, 0.3f);  // etab > 0, no apex guard
    dp_case<2>("dp2_dilatant", rng, 5e-2, 2.0, false, true, false, 0.3f);
    le_case<3>("le3", rng);
    le_case<2>("le2", rng);
  }
  return 0;
}
