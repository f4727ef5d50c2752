// oracle.h -- CPU restatement of PyroclastMPM's per-step explicit MPM path.
//
// TEST INFRASTRUCTURE ONLY. This header is the parity oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may build or call
// it, and only as the checker. The product (pyroclastmpm_amd/, the HIP
// library) never links or calls it.
//
// It restates the reference's algorithm with the reference's loop order:
//   * spatial binning: hash -> std::stable_sort -> cell ranges
//     (src/spatialpartition/spatialpartition.cpp:77-136,
//      spatialpartition_inline.h:49-115)
//   * stored psi/dpsi per particle x window
//     (src/shapefunction/shapefunction.cpp:70-142, linear.cuh:28-52,
//      cubic.cuh:28-197)
//   * node-centric gather P2G over cell_start/cell_end
//     (src/solver/usl/usl_inline.h:82-180)
//   * per-particle G2P (usl_inline.h:245-332)
//   * material updates (linearelastic_inline.h:40-107,
//     newtonfluid_inline.h:49-96, localrheo_inline.h:39-157,
//     dp_rheo_inline.h:43-290, dp_rheo.cpp:45-118)
//   * node boundary conditions (gravity.cpp:65-88, gravity_inline.h:46-63,
//     nodedomain_inline.h:40-87, rigidbodylevelset.cpp:92-275,
//     rigidbodylevelset_inline.h:65-343)
//   * the solver loop (usl.cpp:61-115, solver.cpp:69-174)
// Additions with no reference implementation, written from their textbook
// definitions: MUSL double mapping (semantics of tests/unit/test_musl.h:37-156),
// the quadratic B-spline, and an analytic box signed-distance rigid body with a
// prescribed velocity.
//
// Templated on Real (float = the reference's FP32 build; double = the FP64
// oracle) and on the dimension D (1, 2, 3).
//
// Pinning (what checks each function against something independent of this
// file): tests/test_oracle_goldens.py (the reference's unit-test values) and
// tests/test_oracle_closed_forms.py (closed forms and invariants); each test
// names the function it pins.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

enum ShapeFn { SF_LINEAR = 0, SF_QUADRATIC = 1, SF_CUBIC = 2 };

inline ShapeFn shape_fn_from_string(const std::string &s) {
  if (s == "linear") return SF_LINEAR;
  if (s == "quadratic") return SF_QUADRATIC;
  if (s == "cubic") return SF_CUBIC;
  throw std::invalid_argument("unknown shape function: " + s);
}

// ---------------------------------------------------------------------------
// Small fixed-size algebra (column-major like Eigen: element (i,j) at i+j*N).
// ---------------------------------------------------------------------------
template <class R, int N> struct Vec {
  R v[N];
  R &operator[](int i) { return v[i]; }
  const R &operator[](int i) const { return v[i]; }
  static Vec zero() { Vec a; for (int i = 0; i < N; ++i) a.v[i] = R(0); return a; }
};
template <class R, int N> struct Mat {
  R m[N * N];
  R &operator()(int i, int j) { return m[i + j * N]; }
  const R &operator()(int i, int j) const { return m[i + j * N]; }
  static Mat zero() { Mat a; for (int i = 0; i < N * N; ++i) a.m[i] = R(0); return a; }
  static Mat eye() { Mat a = zero(); for (int i = 0; i < N; ++i) a(i, i) = R(1); return a; }
};
template <int N> struct VecI {
  int v[N];
  int &operator[](int i) { return v[i]; }
  const int &operator[](int i) const { return v[i]; }
};

template <class R, int N> Mat<R, N> matmul(const Mat<R, N> &a, const Mat<R, N> &b) {
  Mat<R, N> c;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      R s = a(i, 0) * b(0, j);
      for (int k = 1; k < N; ++k) s += a(i, k) * b(k, j);
      c(i, j) = s;
    }
  return c;
}
template <class R, int N> Mat<R, N> transpose(const Mat<R, N> &a) {
  Mat<R, N> t;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) t(i, j) = a(j, i);
  return t;
}
template <class R, int N> R trace(const Mat<R, N> &a) {
  R s = a(0, 0);
  for (int i = 1; i < N; ++i) s += a(i, i);
  return s;
}
template <class R> R det(const Mat<R, 1> &a) { return a(0, 0); }
template <class R> R det(const Mat<R, 2> &a) { return a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0); }
template <class R> R det(const Mat<R, 3> &a) {
  return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) -
         a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
         a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
}
// sum_ij a_ij^2  (== trace(a a^T))
template <class R, int N> R frob2(const Mat<R, N> &a) {
  R s = R(0);
  for (int i = 0; i < N * N; ++i) s += a.m[i] * a.m[i];
  return s;
}
template <class R, int D> Mat<R, 3> pad3(const Mat<R, D> &a) {
  Mat<R, 3> b = Mat<R, 3>::zero();
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) b(i, j) = a(i, j);
  return b;
}
template <class R, int D> Mat<R, D> block(const Mat<R, 3> &a) {
  Mat<R, D> b;
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) b(i, j) = a(i, j);
  return b;
}

// ---------------------------------------------------------------------------
// Grid (types_common.h:250-291): n_a = trunc((end_a-origin_a)/h)+1, in Real.
// ---------------------------------------------------------------------------
template <class R, int D> struct Grid {
  Vec<R, D> origin, end;
  R cell_size = R(1), inv_cell_size = R(1);
  VecI<D> num_cells;
  int num_cells_total = 1;
  Grid() { for (int a = 0; a < D; ++a) num_cells[a] = 1; }
  Grid(const Vec<R, D> &o, const Vec<R, D> &e, R h) : origin(o), end(e), cell_size(h) {
    inv_cell_size = R(1.0) / cell_size;
    num_cells_total = 1;
    for (int a = 0; a < D; ++a) {
      num_cells[a] = (int)((end[a] - origin[a]) / cell_size) + 1;
      num_cells_total *= num_cells[a];
    }
  }
  // NODE_MEM_INDEX: x fastest (types_common.h:98-121); plain int arithmetic,
  // reinterpreted as unsigned by callers that bound-check it.
  int linear(const VecI<D> &b) const {
    int h = b[0];
    if (D > 1) h += b[1] * num_cells[0];
    if (D > 2) h += b[2] * num_cells[0] * num_cells[1];
    return h;
  }
  VecI<D> ids_of(int idx) const {
    VecI<D> b;
    b[0] = idx % num_cells[0];
    if (D > 1) b[1] = (idx / num_cells[0]) % num_cells[1];
    if (D > 2) b[2] = idx / (num_cells[0] * num_cells[1]);
    return b;
  }
};

// ---------------------------------------------------------------------------
// Connectivity windows (global_settings.cpp:105-152, 232-281). Generated, not
// tabulated: offsets lo..lo+w-1 per axis, nesting y (outer), x, z (inner) in
// 3D and y (outer), x (inner) in 2D -- the order the reference's tables list,
// which fixes the per-particle summation order in G2P. The P2G window is the
// element-wise negation. Quadratic uses the 4-wide window of the cubic spline
// (its fourth node always has weight 0).
// ---------------------------------------------------------------------------
struct Windows {
  int nsn = 1;
  int g2p[64][3];
  int p2g[64][3];
};
inline Windows make_windows(int D, ShapeFn sf) {
  Windows w;
  const int lo = (sf == SF_LINEAR) ? 0 : -1;
  const int width = (sf == SF_LINEAR) ? 2 : 4;
  int n = 0;
  for (int i = 0; i < 64; ++i) w.g2p[i][0] = w.g2p[i][1] = w.g2p[i][2] = 0;
  if (D == 1) {
    for (int x = lo; x < lo + width; ++x) { w.g2p[n][0] = x; ++n; }
  } else if (D == 2) {
    for (int y = lo; y < lo + width; ++y)
      for (int x = lo; x < lo + width; ++x) { w.g2p[n][0] = x; w.g2p[n][1] = y; ++n; }
  } else {
    for (int y = lo; y < lo + width; ++y)
      for (int x = lo; x < lo + width; ++x)
        for (int z = lo; z < lo + width; ++z) {
          w.g2p[n][0] = x; w.g2p[n][1] = y; w.g2p[n][2] = z; ++n;
        }
  }
  w.nsn = n;
  for (int i = 0; i < 64; ++i)
    for (int a = 0; a < 3; ++a) w.p2g[i][a] = -w.g2p[i][a];
  return w;
}

// ---------------------------------------------------------------------------
// 1-D kernels.
// linear.cuh:28-52: N = 1-|r| (|r| evaluated in double), N' = -+inv_h, 0 at
// r = 0 and |r| > 1.
// ---------------------------------------------------------------------------
template <class R> R linear_N(R r) {
  const double a = std::fabs((double)r);
  if (a >= 1.0) return R(0);
  return (R)((double)(R)1.0 - a);
}
template <class R> R linear_dN(R r, R inv_h) {
  if ((double)r > 1.0 || (double)r < -1.0 || r == R(0)) return R(0);
  if (r > R(0)) return -inv_h;
  return inv_h;
}

// Boundary-modified cubic B-spline (cubic.cuh:28-197). Node type per axis:
// 1 = boundary node, 2 = first interior node, 3 = interior, 4 = last interior
// (never produced by the reference's node typing, nodes.cpp:160). Type 0 =
// untyped (node types not built) -> 0. Piece k covers r in [k-2, k-1);
// coefficients of c0 + c1 r + c2 r^2 + c3 r^3, evaluated in double as the
// reference does (double literals), then rounded to Real.
struct CubicPiece { double c0, c1, c2, c3; };
inline const CubicPiece &cubic_piece(int type, int k) {
  static const CubicPiece Z = {0, 0, 0, 0};
  static const CubicPiece OUT_L = {4.0 / 3.0, 2.0, 1.0, 1.0 / 6.0};   // (2+r)^3/6
  static const CubicPiece IN_L = {2.0 / 3.0, 0.0, -1.0, -0.5};        // 2/3 - r^2 - r^3/2
  static const CubicPiece IN_R = {2.0 / 3.0, 0.0, -1.0, 0.5};         // 2/3 - r^2 + r^3/2
  static const CubicPiece OUT_R = {4.0 / 3.0, -2.0, 1.0, -1.0 / 6.0}; // (2-r)^3/6
  static const CubicPiece B_L = {1.0, 1.0, 0.0, -1.0 / 6.0};
  static const CubicPiece B_R = {1.0, -1.0, 0.0, 1.0 / 6.0};
  static const CubicPiece T2_L = {2.0 / 3.0, 0.0, -1.0, -1.0 / 3.0};
  static const CubicPiece T4_R = {2.0 / 3.0, 0.0, -1.0, 1.0 / 3.0};
  static const CubicPiece *table[5][4] = {
      {&Z, &Z, &Z, &Z},                  // untyped
      {&OUT_L, &B_L, &B_R, &OUT_R},      // 1
      {&Z, &T2_L, &IN_R, &OUT_R},        // 2
      {&OUT_L, &IN_L, &IN_R, &OUT_R},    // 3
      {&OUT_L, &IN_L, &T4_R, &Z}};       // 4
  if (type < 0 || type > 4 || k < 0 || k > 3) return Z;
  return *table[type][k];
}
inline int cubic_piece_index(double r) {
  if (r >= -2.0 && r < -1.0) return 0;
  if (r >= -1.0 && r < 0.0) return 1;
  if (r >= 0.0 && r < 1.0) return 2;
  if (r >= 1.0 && r < 2.0) return 3;
  return -1;
}
template <class R> R cubic_N(R r, int type) {
  const double t = (double)r;
  const int k = cubic_piece_index(t);
  if (k < 0) return R(0);
  const CubicPiece &c = cubic_piece(type, k);
  return (R)(((c.c3 * t + c.c2) * t + c.c1) * t + c.c0);
}
template <class R> R cubic_dN(R r, R inv_h, int type) {
  const double t = (double)r;
  const int k = cubic_piece_index(t);
  if (k < 0) return R(0);
  const CubicPiece &c = cubic_piece(type, k);
  return (R)((double)inv_h * ((3.0 * c.c3 * t + 2.0 * c.c2) * t + c.c1));
}

// Quadratic B-spline (no reference implementation: quadratic.cuh:28 is a TODO).
// N = 3/4 - r^2 (|r| < 1/2), (3/2 - |r|)^2 / 2 (1/2 <= |r| < 3/2), else 0.
template <class R> R quad_N(R r) {
  const R a = std::fabs(r);
  if (a < R(0.5)) return R(0.75) - r * r;
  if (a < R(1.5)) { const R b = R(1.5) - a; return R(0.5) * b * b; }
  return R(0);
}
template <class R> R quad_dN(R r, R inv_h) {
  const R a = std::fabs(r);
  if (a < R(0.5)) return R(-2) * r * inv_h;
  if (a < R(1.5)) { const R b = R(1.5) - a; return (r > R(0) ? -b : b) * inv_h; }
  return R(0);
}

template <class R> R shape_N(ShapeFn sf, R r, int type) {
  if (sf == SF_LINEAR) return linear_N(r);
  if (sf == SF_QUADRATIC) return quad_N(r);
  return cubic_N(r, type);
}
template <class R> R shape_dN(ShapeFn sf, R r, R inv_h, int type) {
  if (sf == SF_LINEAR) return linear_dN(r, inv_h);
  if (sf == SF_QUADRATIC) return quad_dN(r, inv_h);
  return cubic_dN(r, inv_h, type);
}

// ---------------------------------------------------------------------------
// Materials. Parameter sets mirror the reference constructors.
// ---------------------------------------------------------------------------
enum MatKind { MAT_LINEAR_ELASTIC = 1, MAT_NEWTON_FLUID = 2, MAT_LOCAL_RHEO = 3, MAT_DP = 4, MAT_VON_MISES = 5,
               MAT_MOHR_COULOMB = 6, MAT_MCC = 7 };

template <class R> struct Material {
  MatKind kind = MAT_LINEAR_ELASTIC;
  R density = R(0);
  // elastic
  R E = R(0), pois = R(0), bulk_modulus = R(0), shear_modulus = R(0), lame_modulus = R(0);
  // Newtonian fluid
  R viscosity = R(0), gamma = R(7);
  // local mu(I)
  R I0 = R(0), mu_s = R(0), mu_2 = R(0), rho_c = R(0), particle_diameter = R(0),
    particle_density = R(0), EPS = R(0);
  // Drucker-Prager (DP_rheo)
  R M = R(0), lam = R(0), kap = R(0), Vs = R(0), Rratio = R(0), Pt = R(0), beta = R(0),
    dilatancy_angle = R(0), eta_overline = R(0), solid_volume_fraction = R(1);
  // von Mises / Mohr-Coulomb (hardening H; yield stress or cohesion,
  // friction angle in radians); modified Cam-Clay uses M, lam, kap, Vs, Rratio,
  // Pt above
  R yield_stress = R(0), H = R(0), cohesion = R(0), friction_angle = R(0);
  bool do_update_history = true;
  bool is_velgrad_strain_increment = false;
  // when the apex return divides by eta_overline == 0, compute its analytic
  // limit p = Pt instead of the reference's 0*inf = NaN
  bool apex_guard = true;
  // per-particle history (allocated by initialize): elastic strain (D x D
  // block used), reference stress, preconsolidation pressure, alpha = the
  // accumulated (von Mises, Mohr-Coulomb) or volumetric (Cam-Clay) plastic
  // strain, solid volume fraction (DP)
  std::vector<Mat<R, 3>> eps_e;
  std::vector<Mat<R, 3>> stress_ref;
  std::vector<R> pc, alpha, svf;

  // linearelastic.cpp:47-58
  static Material linear_elastic(R density, R E, R pois) {
    Material m; m.kind = MAT_LINEAR_ELASTIC; m.density = density; m.E = E; m.pois = pois;
    m.bulk_modulus = E / ((R)3.0 * ((R)1.0 - (R)2.0 * pois));
    m.shear_modulus = E / ((R)2.0 * ((R)1 + pois));
    m.lame_modulus = E * pois / (((R)1.0 + pois) * ((R)1.0 - (R)2.0 * pois));
    return m;
  }
  // newtonfluid.cpp:52-58
  static Material newton_fluid(R density, R viscosity, R bulk_modulus, R gamma) {
    Material m; m.kind = MAT_NEWTON_FLUID; m.density = density; m.viscosity = viscosity;
    m.bulk_modulus = bulk_modulus; m.gamma = gamma;
    return m;
  }
  // localrheo.cpp:50-67
  static Material local_rheo(R density, R E, R pois, R I0, R mu_s, R mu_2, R rho_c,
                             R particle_diameter, R particle_density) {
    Material m; m.kind = MAT_LOCAL_RHEO; m.density = density; m.E = E; m.pois = pois;
    m.I0 = I0; m.mu_s = mu_s; m.mu_2 = mu_2; m.rho_c = rho_c;
    m.particle_diameter = particle_diameter; m.particle_density = particle_density;
    m.bulk_modulus = ((R)1. / (R)3.) * (E / ((R)1. - (R)2. * pois));
    m.shear_modulus = ((R)1. / (R)2.) * E / ((R)1 + pois);
    m.lame_modulus = (pois * E) / (((R)1.0 + pois) * ((R)1 - (R)2.0 * pois));
    m.EPS = I0 / (R)std::sqrt(std::pow((double)particle_diameter, 2) * (double)particle_density);
    return m;
  }
  // vonmises.cpp:34-50
  static Material von_mises(R density, R E, R pois, R yield_stress, R H) {
    Material m; m.kind = MAT_VON_MISES; m.density = density; m.E = E; m.pois = pois;
    m.yield_stress = yield_stress; m.H = H;
    m.bulk_modulus = E / ((R)3.0 * ((R)1.0 - (R)2.0 * pois));
    m.shear_modulus = E / ((R)2.0 * ((R)1 + pois));
    return m;
  }
  // mohrcoulomb.cpp:33-52
  static Material mohr_coulomb(R density, R E, R pois, R cohesion, R friction_deg, R dilatancy_deg, R H) {
    Material m; m.kind = MAT_MOHR_COULOMB; m.density = density; m.E = E; m.pois = pois;
    m.cohesion = cohesion; m.H = H;
    m.bulk_modulus = ((R)1. / (R)3.) * (E / ((R)1. - (R)2. * pois));
    m.shear_modulus = ((R)1. / (R)2.) * E / ((R)1 + pois);
    m.lame_modulus = (pois * E) / (((R)1.0 + pois) * ((R)1 - (R)2.0 * pois));
    m.friction_angle = friction_deg * (R)(3.14159265358979323846 / 180.);
    m.dilatancy_angle = dilatancy_deg * (R)(3.14159265358979323846 / 180.);
    return m;
  }
  // modifiedcamclay.cpp:47-63
  static Material modified_cam_clay(R density, R E, R pois, R M, R lam, R kap, R Vs, R Rratio, R Pt, R beta) {
    Material m; m.kind = MAT_MCC; m.density = density; m.E = E; m.pois = pois; m.M = M; m.lam = lam;
    m.kap = kap; m.Vs = Vs; m.Rratio = Rratio; m.Pt = Pt; m.beta = beta;
    m.bulk_modulus = E / ((R)3.0 * ((R)1.0 - (R)2.0 * pois));
    m.shear_modulus = E / ((R)2.0 * ((R)1 + pois));
    return m;
  }
  // dp_rheo.cpp:45-70
  static Material dp_rheo(R density, R E, R pois, R M, R lam, R kap, R Vs, R Rratio, R Pt,
                          R beta, R dilatancy_angle_deg, R solid_volume_fraction) {
    Material m; m.kind = MAT_DP; m.density = density; m.E = E; m.pois = pois; m.M = M;
    m.lam = lam; m.kap = kap; m.Vs = Vs; m.Rratio = Rratio; m.Pt = Pt; m.beta = beta;
    m.solid_volume_fraction = solid_volume_fraction;
    m.bulk_modulus = E / ((R)3.0 * ((R)1.0 - (R)2.0 * pois));
    m.shear_modulus = E / ((R)2.0 * ((R)1 + pois));
    m.dilatancy_angle = dilatancy_angle_deg * (R)(3.14159265358979323846 / 180.);
    const double sp = std::sin((double)m.dilatancy_angle);
    m.eta_overline = (R)(6.0 * sp / (std::sqrt(3.0) * (3.0 + sp)));
    return m;
  }
};

// ---------------------------------------------------------------------------
// Boundary conditions.
// ---------------------------------------------------------------------------
enum BCKind { BC_GRAVITY = 1, BC_NODE_DOMAIN = 2, BC_RIGID_LEVELSET = 3, BC_RIGID_SDF = 4, BC_BODY_FORCE = 5,
              BC_PLANAR_DOMAIN = 6 };

template <class R, int D> struct BoundaryCondition {
  BCKind kind = BC_GRAVITY;
  // Gravity (gravity.cpp:53-88)
  Vec<R, D> gravity = Vec<R, D>::zero(), gravity_end = Vec<R, D>::zero();
  bool is_ramp = false;
  int ramp_step = 0;
  // NodeDomain (nodedomain.cpp:33-64): 0 = fixed, 1 = slip
  VecI<D> face0_mode{}, face1_mode{};
  // RigidBodyLevelSet (rigidbodylevelset.cpp:41-217)
  int rigid_mode = 0;  // 0 STATIC, 1 LOOP_ROTATE
  Vec<R, D> euler_angles = Vec<R, D>::zero(), euler_angles_per_second = Vec<R, D>::zero(),
            COM = Vec<R, D>::zero();
  Mat<R, D> rotation_matrix = Mat<R, D>::eye();
  // Analytic box SDF body (no reference implementation): half extents, centre,
  // body->world rotation (3x3 column-major, top-left D x D used), prescribed
  // velocity, contact band (node in contact when phi < band).
  Vec<R, 3> sdf_half{}, sdf_center{}, sdf_velocity{};
  Mat<R, 3> sdf_rot = Mat<R, 3>::eye();
  R sdf_band = R(0);
  // BodyForce (bodyforce.cpp:28-50): mode 0 "forces", 1 "moments", 2 "fixed";
  // one value and mask entry per node
  int body_mode = 0;
  std::vector<Vec<R, D>> body_values;
  std::vector<uint8_t> body_mask;
  // PlanarDomain (planardomain.cpp:27-33): friction per face, stored in
  // radians as the reference converts them
  Vec<R, D> face0_friction = Vec<R, D>::zero(), face1_friction = Vec<R, D>::zero();
};

// Box signed distance in the body frame and its world-frame gradient.
template <class R, int D>
R box_sdf(const BoundaryCondition<R, D> &bc, const Vec<R, D> &x, Vec<R, D> &grad) {
  R q[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) {
    R s = R(0);
    for (int a = 0; a < D; ++a) s += bc.sdf_rot(a, i) * (x[a] - bc.sdf_center[a]);
    q[i] = s;
  }
  R d[3];
  for (int i = 0; i < D; ++i) d[i] = std::fabs(q[i]) - bc.sdf_half[i];
  R out2 = R(0), inmax = d[0];
  int amax = 0;
  for (int i = 0; i < D; ++i) {
    if (d[i] > R(0)) out2 += d[i] * d[i];
    if (d[i] > inmax) { inmax = d[i]; amax = i; }
  }
  R gl[3] = {0, 0, 0};
  R phi;
  if (out2 > R(0)) {
    const R out = std::sqrt(out2);
    for (int i = 0; i < D; ++i)
      gl[i] = (d[i] > R(0) ? d[i] : R(0)) * (q[i] >= R(0) ? R(1) : R(-1)) / out;
    phi = out;
  } else {
    gl[amax] = q[amax] >= R(0) ? R(1) : R(-1);
    phi = inmax;
  }
  for (int a = 0; a < D; ++a) {
    R s = R(0);
    for (int i = 0; i < D; ++i) s += bc.sdf_rot(a, i) * gl[i];
    grad[a] = s;
  }
  return phi;
}

// ---------------------------------------------------------------------------
// The solver state (ParticlesContainer + SpatialPartition + NodesContainer +
// process globals of the reference, held per instance).
// ---------------------------------------------------------------------------
template <class R, int D> struct Sim {
  using V = Vec<R, D>;
  using M = Mat<R, D>;
  using M3 = Mat<R, 3>;
  using VI = VecI<D>;

  // globals (global_settings.cpp:49-98)
  R dt = R(0);
  int particles_per_cell = 1;
  ShapeFn shape_fn = SF_LINEAR;
  Windows win;
  int global_step = 0;

  // nodes (nodes.cpp:55-83)
  Grid<R, D> grid;
  std::vector<V> node_moments, node_moments_nt;
  std::vector<R> node_masses;
  std::vector<VI> node_ids, node_types;
  R small_mass_cutoff = (R)(0.000000001);

  // particles (particles.cpp:51-149)
  int num_particles = 0, num_rigid_particles = 0;
  std::vector<V> positions, velocities, forces_external;
  std::vector<M3> stresses;
  std::vector<M> velocity_gradient, F;
  std::vector<R> masses, volumes, volumes_original;
  std::vector<int> colors;
  std::vector<uint8_t> is_rigid, is_active;
  std::vector<R> psi;
  std::vector<V> dpsi;

  // spatial partition (spatialpartition.cpp:51-74)
  std::vector<VI> bins;
  std::vector<unsigned> hash_unsorted, hash_sorted;
  std::vector<int> sorted_index, cell_start, cell_end;

  // solver
  std::vector<Material<R>> materials;
  std::vector<BoundaryCondition<R, D>> bcs;
  bool musl = false;
  R alpha = R(0.99);
  // true: P2G bound-checks only the linear node index, so a cell index
  // outside the grid on some axis can wrap into range (usl_inline.h:122-126,
  // the reference's behaviour); false: such cells are skipped on every axis
  bool wrap_quirk = true;

  Sim(const V &origin, const V &end, R cell_size, ShapeFn sf, R dt_, int ppc)
      : dt(dt_), particles_per_cell(ppc), shape_fn(sf) {
    win = make_windows(D, sf);
    grid = Grid<R, D>(origin, end, cell_size);
    const int nn = grid.num_cells_total;
    node_moments.assign(nn, V::zero());
    node_moments_nt.assign(nn, V::zero());
    node_masses.assign(nn, R(0));
    node_ids.resize(nn);
    node_types.resize(nn);
    // calculate_bin_ids (nodes.cpp:89-129)
    for (int i = 0; i < nn; ++i) node_ids[i] = grid.ids_of(i);
    // calculate_bin_types (nodes.cpp:135-173); only for the cubic spline;
    // the reference's "left of boundary" test (ids == ids-2) never holds,
    // so type 4 is never assigned.
    for (int i = 0; i < nn; ++i)
      for (int a = 0; a < D; ++a) {
        int t = 0;
        if (sf == SF_CUBIC) {
          const int id = node_ids[i][a];
          if (id == 0 || id == grid.num_cells[a] - 1) t = 1;
          else if (id == 1) t = 2;
          else t = 3;
        }
        node_types[i][a] = t;
      }
  }

  // ParticlesContainer constructor: non-rigid first, then rigid.
  void set_particles(const std::vector<V> &pos, const std::vector<V> &vel,
                     const std::vector<int> &col, const std::vector<V> &rpos,
                     const std::vector<V> &rvel, const std::vector<int> &rcol) {
    const int nr = (int)rpos.size(), nn = (int)pos.size();
    num_rigid_particles = nr;
    num_particles = nn + nr;
    const int N = num_particles;
    positions.assign(N, V::zero());
    velocities.assign(N, V::zero());
    colors.assign(N, 0);
    is_rigid.assign(N, 0);
    for (int p = 0; p < N; ++p) {
      if (p < nn) {
        positions[p] = pos[p];
        if (!vel.empty()) velocities[p] = vel[p];
        if (!col.empty()) colors[p] = col[p];
      } else {
        const int r = p - nn;
        positions[p] = rpos[r];
        if (!rvel.empty()) velocities[p] = rvel[r];
        if (!rcol.empty()) colors[p] = rcol[r];
        is_rigid[p] = 1;
      }
    }
    stresses.assign(N, M3::zero());
    is_active.assign(N, 1);
    masses.assign(N, R(-1.0));
    volumes.assign(N, R(-1.0));
    volumes_original.assign(N, R(-1.0));
    velocity_gradient.assign(N, M::zero());
    F.assign(N, M::eye());
    forces_external.assign(N, V::zero());
    psi.assign((size_t)N * win.nsn, R(0));
    dpsi.assign((size_t)N * win.nsn, V::zero());
    bins.assign(N, VI{});
    hash_unsorted.assign(N, 0u);
    hash_sorted.assign(N, 0u);
    sorted_index.assign(N, -1);
    cell_start.assign(grid.num_cells_total, 0);
    cell_end.assign(grid.num_cells_total, 0);
    reset_spatial();
  }

  // ---------------- spatial partition ----------------
  void reset_spatial() {  // spatialpartition.cpp:77-86
    std::iota(sorted_index.begin(), sorted_index.end(), 0);
    std::fill(cell_start.begin(), cell_start.end(), 0);
    std::fill(cell_end.begin(), cell_end.end(), 0);
    std::fill(hash_sorted.begin(), hash_sorted.end(), 0u);
    std::fill(hash_unsorted.begin(), hash_unsorted.end(), 0u);
    for (auto &b : bins) for (int a = 0; a < D; ++a) b[a] = 0;
  }
  void calculate_hash() {  // spatialpartition_inline.h:49-62
    for (int p = 0; p < num_particles; ++p) {
      VI b;
      for (int a = 0; a < D; ++a)
        b[a] = (int)((positions[p][a] - grid.origin[a]) * grid.inv_cell_size);
      bins[p] = b;
      hash_unsorted[p] = (unsigned)grid.linear(b);
    }
    hash_sorted = hash_unsorted;
  }
  void sort_hashes() {  // spatialpartition.cpp:113-117 (stable_sort_by_key)
    std::vector<int> idx(sorted_index);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return hash_sorted[a] < hash_sorted[b]; });
    // thrust::stable_sort_by_key permutes keys and values together
    std::vector<unsigned> keys(num_particles);
    for (int j = 0; j < num_particles; ++j) keys[j] = hash_sorted[idx[j]];
    // sorted_index starts as 0..N-1, so value j of the permuted sequence is idx[j]
    std::vector<int> vals(num_particles);
    for (int j = 0; j < num_particles; ++j) vals[j] = sorted_index[idx[j]];
    hash_sorted = keys;
    sorted_index = vals;
  }
  void bin_particles() {  // spatialpartition_inline.h:85-115
    const int n = num_particles;
    const unsigned nc = (unsigned)grid.num_cells_total;
    for (int t = 0; t < n; ++t) {
      const unsigned h = hash_sorted[t];
      if (t < n - 1) {
        const unsigned nh = hash_sorted[t + 1];
        if (t == 0 && h < nc) cell_start[h] = t;
        if (h != nh) {
          if (h < nc) cell_end[h] = t + 1;
          if (nh < nc) cell_start[nh] = t + 1;
        }
      }
      if (t == n - 1 && h < nc) cell_end[h] = t + 1;
    }
  }
  void partition() { calculate_hash(); sort_hashes(); bin_particles(); }

  // ---------------- shape functions (shapefunction.cpp:70-142) ----------------
  void calculate_shape_function() {
    const int nsn = win.nsn;
    for (int p = 0; p < num_particles; ++p) {
      const V xp = positions[p];
      const VI bp = bins[p];
      for (int i = 0; i < nsn; ++i) {
        VI sb;
        for (int a = 0; a < D; ++a) sb[a] = bp[a] + win.g2p[i][a];
        const unsigned nmi = (unsigned)grid.linear(sb);
        if (nmi >= (unsigned)grid.num_cells_total) continue;
        V rel;
        for (int a = 0; a < D; ++a)
          rel[a] = (xp[a] - grid.origin[a]) * grid.inv_cell_size - (R)sb[a];
        const VI nt = node_types[nmi];
        R N[3], dN[3];
        for (int a = 0; a < D; ++a) {
          N[a] = shape_N<R>(shape_fn, rel[a], nt[a]);
          dN[a] = shape_dN<R>(shape_fn, rel[a], grid.inv_cell_size, nt[a]);
        }
        R w;
        V g;
        if (D == 1) { w = N[0]; g[0] = dN[0]; }
        else if (D == 2) { w = N[0] * N[1]; g[0] = dN[0] * N[1]; g[1] = dN[1] * N[0]; }
        else {
          w = N[0] * N[1] * N[2];
          g[0] = dN[0] * N[1] * N[2];
          g[1] = dN[1] * N[0] * N[2];
          g[2] = dN[2] * N[0] * N[1];
        }
        psi[(size_t)p * nsn + i] = w;
        dpsi[(size_t)p * nsn + i] = g;
      }
    }
  }

  // ---------------- volumes / masses (particles.cpp:345-393) ----------------
  void calculate_initial_volumes() {
    for (int p = 0; p < num_particles; ++p) {
      if (volumes[p] > R(0)) continue;
      R v = (R)std::pow((double)grid.cell_size, (double)D);
      v /= (R)particles_per_cell;
      volumes[p] = v;
    }
    volumes_original = volumes;
  }
  void calculate_initial_masses(int mat_id, R density) {
    for (int p = 0; p < num_particles; ++p) {
      if (colors[p] != mat_id || masses[p] > R(0)) continue;
      masses[p] = density * volumes[p];
    }
  }

  // ---------------- resets (usl.cpp:61-66) ----------------
  void reset() {
    std::fill(node_moments.begin(), node_moments.end(), V::zero());
    std::fill(node_moments_nt.begin(), node_moments_nt.end(), V::zero());
    std::fill(node_masses.begin(), node_masses.end(), R(0));
    std::fill(psi.begin(), psi.end(), R(0));
    std::fill(dpsi.begin(), dpsi.end(), V::zero());
    std::fill(velocity_gradient.begin(), velocity_gradient.end(), M::zero());
    std::fill(forces_external.begin(), forces_external.end(), V::zero());
    reset_spatial();
  }

  // ---------------- P2G (usl_inline.h:82-180) ----------------
  bool in_grid(const VI &b) const {
    for (int a = 0; a < D; ++a)
      if (b[a] < 0 || b[a] >= grid.num_cells[a]) return false;
    return true;
  }
  void p2g() {
    const int nsn = win.nsn;
    for (int n = 0; n < grid.num_cells_total; ++n) {
      const VI nb = node_ids[n];
      V mom = V::zero(), force = V::zero();
      R mass = R(0);
      for (int sid = 0; sid < nsn; ++sid) {
        VI sb;
        for (int a = 0; a < D; ++a) sb[a] = nb[a] + win.p2g[sid][a];
        if (!wrap_quirk && !in_grid(sb)) continue;
        const unsigned h = (unsigned)grid.linear(sb);
        if (h >= (unsigned)grid.num_cells_total) continue;
        const int cend = cell_end[h];
        if (cend == 0) continue;
        const int cstart = cell_start[h];
        for (int j = cstart; j < cend; ++j) {
          const int p = sorted_index[j];
          const R w = psi[(size_t)p * nsn + sid];
          const V g = dpsi[(size_t)p * nsn + sid];
          const R sm = w * masses[p];
          mass += sm;
          for (int a = 0; a < D; ++a) mom[a] += sm * velocities[p][a];
          for (int a = 0; a < D; ++a) force[a] += w * forces_external[p][a];
          const R mv = -volumes[p];
          for (int a = 0; a < D; ++a) {
            R s = R(0);
            for (int b = 0; b < D; ++b) s += mv * stresses[p](a, b) * g[b];
            force[a] += s;
          }
        }
      }
      if (mass <= small_mass_cutoff) continue;
      node_masses[n] = mass;
      node_moments[n] = mom;
      for (int a = 0; a < D; ++a) node_moments_nt[n][a] = mom[a] + force[a] * dt;
    }
  }

  // ---------------- G2P (usl_inline.h:245-332) ----------------
  // MUSL splits it: mode 1 = velocity/position only (first map), mode 2 =
  // velocity gradient / F / V from moments_nt (second map), mode 0 = USL.
  void g2p(int mode) {
    const int nsn = win.nsn;
    for (int p = 0; p < num_particles; ++p) {
      const VI bp = bins[p];
      V vinc = V::zero(), dvinc = V::zero();
      M L = M::zero();
      for (int i = 0; i < nsn; ++i) {
        VI sb;
        bool bad = false;
        for (int a = 0; a < D; ++a) {
          sb[a] = bp[a] + win.g2p[i][a];
          if (sb[a] < 0 || sb[a] >= grid.num_cells[a]) bad = true;
        }
        if (bad) continue;
        const int h = grid.linear(sb);
        const R nm = node_masses[h];
        if (nm <= small_mass_cutoff) continue;
        const R w = psi[(size_t)p * nsn + i];
        const V g = dpsi[(size_t)p * nsn + i];
        V nv, nvt;
        for (int a = 0; a < D; ++a) {
          nv[a] = node_moments[h][a] / nm;
          nvt[a] = node_moments_nt[h][a] / nm;
        }
        for (int a = 0; a < D; ++a) {
          dvinc[a] += w * (nvt[a] - nv[a]);
          vinc[a] += w * nvt[a];
        }
        for (int a = 0; a < D; ++a)
          for (int b = 0; b < D; ++b) L(a, b) += g[a] * nvt[b];
      }
      if (mode != 2) {
        for (int a = 0; a < D; ++a)
          velocities[p][a] = (R)(1. - alpha) * vinc[a] + alpha * (velocities[p][a] + dvinc[a]);
        for (int a = 0; a < D; ++a) positions[p][a] = positions[p][a] + dt * vinc[a];
      }
      if (mode != 1) {
        velocity_gradient[p] = L;
        M A = M::eye();
        for (int k = 0; k < D * D; ++k) A.m[k] += L.m[k] * dt;
        F[p] = matmul(A, F[p]);
        volumes[p] = det(F[p]) * volumes_original[p];
      }
    }
  }

  // MUSL re-map (no reference implementation; tests/unit/test_musl.h:118-141):
  // moments_nt := sum_p N m_p v_p over the first map's weights.
  void p2g_double_mapping() {
    const int nsn = win.nsn;
    for (int n = 0; n < grid.num_cells_total; ++n) {
      const VI nb = node_ids[n];
      V mom = V::zero();
      for (int sid = 0; sid < nsn; ++sid) {
        VI sb;
        for (int a = 0; a < D; ++a) sb[a] = nb[a] + win.p2g[sid][a];
        if (!wrap_quirk && !in_grid(sb)) continue;
        const unsigned h = (unsigned)grid.linear(sb);
        if (h >= (unsigned)grid.num_cells_total) continue;
        const int cend = cell_end[h];
        if (cend == 0) continue;
        for (int j = cell_start[h]; j < cend; ++j) {
          const int p = sorted_index[j];
          const R sm = psi[(size_t)p * nsn + sid] * masses[p];
          for (int a = 0; a < D; ++a) mom[a] += sm * velocities[p][a];
        }
      }
      if (node_masses[n] <= small_mass_cutoff) continue;
      node_moments_nt[n] = mom;
    }
  }

  // ---------------- materials ----------------
  // linearelastic_inline.h:40-107 -- sigma += lambda*deps*I + 2G*deps (the
  // matrix product with I makes it (lambda+2G)*deps).
  void update_linear_elastic(const Material<R> &mt, int mat_id) {
    for (int p = 0; p < num_particles; ++p) {
      if (!is_active[p] || colors[p] != mat_id) continue;
      const M Lp = velocity_gradient[p];
      const M LT = transpose(Lp);
      M deps;
      for (int k = 0; k < D * D; ++k) deps.m[k] = (R)0.5 * (Lp.m[k] + LT.m[k]) * dt;
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j)
          stresses[p](i, j) += mt.lame_modulus * deps(i, j) + (R)2. * mt.shear_modulus * deps(i, j);
    }
  }
  // newtonfluid_inline.h:49-96
  void update_newton_fluid(const Material<R> &mt, int mat_id) {
    for (int p = 0; p < num_particles; ++p) {
      if (!is_active[p] || colors[p] != mat_id) continue;
      const R rho = masses[p] / volumes[p];
      const R rho0 = masses[p] / volumes_original[p];
      const R pres = -mt.bulk_modulus * ((R)std::pow((double)(rho / rho0), (double)mt.gamma) - (R)1.0);
      const M Lp = velocity_gradient[p];
      const M LT = transpose(Lp);
      M D_;
      for (int k = 0; k < D * D; ++k) D_.m[k] = (R)0.5 * (Lp.m[k] + LT.m[k]);
      const R tr = trace(D_);
      M3 s = M3::zero();
      const R fac = (R)(2. * (double)mt.viscosity / 3.);
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          R dev = D_(i, j) - (i == j ? (R)((1. / 3.) * (double)tr) : R(0));
          s(i, j) = (i == j ? pres : R(0)) + dev * fac;
        }
      stresses[p] = s;
    }
  }
  // localrheo_inline.h:39-157 (grain constants 2450 / 0.005 hard-coded there)
  static double negative_root(double a, double b, double c) {
    if (b > 0) return (-b - std::sqrt(b * b - 4 * a * c)) / (2 * a);
    return (2 * c) / (-b + std::sqrt(b * b - 4 * a * c));
  }
  void update_local_rheo(const Material<R> &mt, int mat_id) {
    for (int p = 0; p < num_particles; ++p) {
      if (colors[p] != mat_id) continue;
      const M3 Lg = pad3<R, D>(velocity_gradient[p]);
      const R rho = masses[p] / volumes[p];
      const M3 LT = transpose(Lg);
      M3 Dm, W;
      for (int k = 0; k < 9; ++k) { Dm.m[k] = (R)0.5 * (Lg.m[k] + LT.m[k]); W.m[k] = (R)0.5 * (Lg.m[k] - LT.m[k]); }
      const M3 sp = stresses[p];
      const M3 sW = matmul(sp, W), Ws = matmul(W, sp);
      const R trD = trace(Dm);
      M3 st;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const R Gn = -sW(i, j) + Ws(i, j);
          st(i, j) = sp(i, j) + dt * ((R)2 * mt.shear_modulus * Dm(i, j) +
                                      (i == j ? mt.lame_modulus * trD : R(0)) + Gn);
        }
      const R ptr = (R)(-(1. / 3.) * (double)trace(st));
      M3 s0 = st;
      for (int i = 0; i < 3; ++i) s0(i, i) += ptr;
      const R tau = (R)std::sqrt(0.5 * (double)frob2(s0));
      if (ptr < R(0) || rho < mt.rho_c) { stresses[p] = M3::zero(); continue; }
      const R S0 = mt.mu_s * ptr;
      R scale = R(1);
      if (!(tau <= S0)) {
        const R S2 = mt.mu_2 * ptr;
        const R GRAIN_RHO = (R)2450.0, GRAIN_D = (R)0.005;
        const R al = mt.shear_modulus * mt.I0 * dt * (R)std::sqrt((double)(ptr / GRAIN_RHO)) / GRAIN_D;
        const R B = -(S2 + tau + al);
        const R H = S2 * tau + S0 * al;
        const R tn = (R)negative_root(1.0, (double)B, (double)H);
        scale = tn / tau;
      }
      M3 out;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out(i, j) = scale * s0(i, j) - (i == j ? ptr : R(0));
      stresses[p] = out;
    }
  }
  // dp_rheo.cpp:75-118
  void dp_initialize(Material<R> &mt) {
    const int N = num_particles;
    mt.alpha.assign(N, R(0));
    mt.pc.assign(N, R(0));
    mt.svf.assign(N, mt.solid_volume_fraction);
    mt.eps_e.assign(N, M3::zero());
    mt.stress_ref = stresses;
    for (int p = 0; p < N; ++p) mt.pc[p] = (-trace(stresses[p]) / (R)3.) * mt.Rratio;
  }
  // dp_rheo_inline.h:43-290
  void update_dp(Material<R> &mt, int mat_id) {
    const double K = mt.bulk_modulus, G = mt.shear_modulus;
    for (int p = 0; p < num_particles; ++p) {
      if (colors[p] != mat_id) continue;
      const M Lp = velocity_gradient[p];
      M eps_prev = block<R, D>(mt.eps_e[p]);
      M deps;
      if (mt.is_velgrad_strain_increment) {
        deps = Lp;
      } else {
        const M LT = transpose(Lp);
        M W;
        for (int k = 0; k < D * D; ++k) {
          deps.m[k] = (R)0.5 * (Lp.m[k] + LT.m[k]) * dt;
          W.m[k] = (R)0.5 * (Lp.m[k] - LT.m[k]) * dt;
        }
        const M We = matmul(W, eps_prev), eW = matmul(eps_prev, W);
        for (int k = 0; k < D * D; ++k) deps.m[k] += -We.m[k] + eW.m[k];
      }
      const M3 sref = mt.stress_ref[p];
      const R pref = (R)(1.0 / 3.0) * trace(sref);
      M3 s_ref = sref;
      for (int i = 0; i < 3; ++i) s_ref(i, i) -= pref;
      mt.svf[p] = mt.Vs / volumes[p];
      M etr;
      for (int k = 0; k < D * D; ++k) etr.m[k] = eps_prev.m[k] + deps.m[k];
      const R evtr = trace(etr);
      const R ptr = pref + mt.bulk_modulus * evtr;
      M3 edev = M3::zero();
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) edev(i, j) = etr(i, j);
      for (int i = 0; i < 3; ++i) edev(i, i) -= (R)(1 / 3.) * evtr;
      if (D < 3) edev(2, 2) = R(0);
      M3 str;
      for (int k = 0; k < 9; ++k) str.m[k] = s_ref.m[k] + (R)2. * mt.shear_modulus * edev.m[k];
      const R J2 = (R)0.5 * frob2(str);
      const R eta = (R)(std::sqrt(1. / 3.) * (double)mt.M);
      const R Phi = (R)std::sqrt((double)J2) + eta * (ptr - mt.Pt);
      if (Phi <= (R)0.0) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) stresses[p](i, j) = str(i, j) + (i == j ? ptr : R(0));
        if (mt.do_update_history) {
          M3 e = M3::zero();
          for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) e(i, j) = etr(i, j);
          mt.eps_e[p] = e;
        }
        continue;
      }
      // cone return (dp_rheo_inline.h:50-65)
      double dg = (double)Phi / (K * (double)eta * (double)mt.eta_overline + G);
      double pn = (double)ptr - K * (double)mt.eta_overline * dg;
      const double sqJ2 = std::sqrt((double)J2);
      M3 sn;
      {
        const double f = 1.0 - (G * dg) / sqJ2;
        for (int k = 0; k < 9; ++k) sn.m[k] = (R)f * str.m[k];
      }
      const bool apex = (double)(R)std::sqrt((double)J2) - G * dg < 0.0;
      if (apex) {  // dp_rheo_inline.h:67-78
        if (mt.apex_guard && mt.eta_overline == R(0)) {
          pn = (double)mt.Pt;
        } else {
          const double dga = ((double)ptr - (double)mt.Pt) / ((double)mt.eta_overline * K);
          pn = (double)ptr - K * (double)mt.eta_overline * dga;
        }
        sn = M3::zero();
      }
      // update_plasticity (dp_rheo_inline.h:80-96)
      M3 e3;
      const R ish = (R)(2.0 * G), vol = (R)((pn - (double)pref) / (3. * K));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          e3(i, j) = (sn(i, j) - s_ref(i, j)) / ish + (i == j ? vol : R(0));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) stresses[p](i, j) = sn(i, j) + (i == j ? (R)pn : R(0));
      if (mt.do_update_history) {
        M3 e = M3::zero();
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j) e(i, j) = e3(i, j);
        mt.eps_e[p] = e;
      }
    }
  }
  // vonmises.cpp:55-59 and mohrcoulomb.cpp:55-59: zero history
  void plastic_initialize(Material<R> &mt) {
    mt.alpha.assign(num_particles, R(0));
    mt.eps_e.assign(num_particles, M3::zero());
  }
  // modifiedcamclay.cpp:68-104: stress_ref = the stresses, pc = R p0
  void mcc_initialize(Material<R> &mt) {
    const int N = num_particles;
    mt.alpha.assign(N, R(0));
    mt.pc.assign(N, R(0));
    mt.eps_e.assign(N, M3::zero());
    mt.stress_ref = stresses;
    for (int p = 0; p < N; ++p) {
      const R p0 = -trace(stresses[p]) / (R)3.;
      mt.pc[p] = p0 * mt.Rratio;
    }
  }
  // (M M^T).trace(), each diagonal entry a row dot product
  static R trace_mmt(const M3 &m) {
    R t = R(0);
    for (int i = 0; i < 3; ++i) {
      const R r = m(i, 0) * m(i, 0) + m(i, 1) * m(i, 1) + m(i, 2) * m(i, 2);
      t = i == 0 ? r : t + r;
    }
    return t;
  }
  // Newton iterations stop here at the latest (the reference's loops have no
  // cap; see each)
  static constexpr int kMaxIters = 100;

  // vonmises_inline.h:41-140 (3D; 2D prints a message there and leaves the
  // stress alone)
  void update_von_mises(Material<R> &mt, int mat_id) {
    if (D != 3) return;
    const R K = mt.bulk_modulus, G = mt.shear_modulus;
    for (int p = 0; p < num_particles; ++p) {
      if (colors[p] != mat_id) continue;
      const M3 L3 = pad3<R, D>(velocity_gradient[p]);
      M3 deps;
      for (int k = 0; k < 9; ++k)
        deps.m[k] = mt.is_velgrad_strain_increment ? L3.m[k] : (R)0.5 * (L3.m[k] + transpose(L3).m[k]) * dt;
      M3 etr;
      for (int k = 0; k < 9; ++k) etr.m[k] = mt.eps_e[p].m[k] + deps.m[k];
      const R ev = trace(etr);
      const R ptr = K * ev;
      M3 dev = etr;
      for (int i = 0; i < 3; ++i) dev(i, i) -= (R)((1 / 3.) * (double)ev);
      M3 st;
      for (int k = 0; k < 9; ++k) st.m[k] = (R)(2. * (double)G) * dev.m[k];
      const R acc_tr = mt.alpha[p];
      const R sy_tr = mt.yield_stress + mt.H * acc_tr;
      const R q = (R)std::sqrt(3 * 0.5 * (double)trace_mmt(st));
      const R phi = q - sy_tr;
      if (phi <= 0) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) stresses[p](i, j) = st(i, j) + (i == j ? ptr : R(0));
        if (mt.do_update_history) mt.eps_e[p] = etr;
        continue;
      }
      double dg = 0.0, acc = 0.0, ph = 0.0;
      for (int it = 0; it < kMaxIters; ++it) {
        acc = (double)acc_tr + dg;
        const double sy = (double)mt.yield_stress + (double)mt.H * acc;
        ph = (double)q - 3.0 * (double)G * dg - sy;
        const double d = -3.0 * (double)G - (double)mt.H;
        dg = dg - ph / d;
        if (!(std::fabs(ph) > 1e-7)) break;
      }
      const R f = (R)(1.0 - (3.0 * (double)G * dg) / (double)q);
      M3 sc;
      for (int k = 0; k < 9; ++k) sc.m[k] = f * st.m[k];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) stresses[p](i, j) = sc(i, j) + (i == j ? ptr : R(0));
      if (mt.do_update_history) {
        const R g2 = (R)(2.0 * (double)G), v3 = (R)((1. / 3.) * (double)ev);
        M3 e;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) e(i, j) = sc(i, j) / g2 + (i == j ? v3 : R(0));
        mt.eps_e[p] = e;
        mt.alpha[p] = (R)acc;
      }
    }
  }

  // Symmetric 3 x 3 eigen-decomposition by cyclic Jacobi rotations in double
  // (the reference uses Eigen's closed form, SelfAdjointEigenSolver::
  // computeDirect); w descending, column k of V the eigenvector of w[k].
  static void sym_eig3(const double A[3][3], double w[3], double V[3][3]) {
    double a[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) { a[i][j] = A[i][j]; V[i][j] = i == j ? 1.0 : 0.0; }
    for (int sweep = 0; sweep < 12; ++sweep) {
      const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
      const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
      if (off <= 1e-32 * dia || off == 0.0) break;
      for (int p = 0; p < 2; ++p)
        for (int q = p + 1; q < 3; ++q) {
          if (a[p][q] == 0.0) continue;
          const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
          const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
          const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
          for (int k = 0; k < 3; ++k) {
            const double akp = a[k][p], akq = a[k][q];
            a[k][p] = c * akp - sn * akq;
            a[k][q] = sn * akp + c * akq;
          }
          for (int k = 0; k < 3; ++k) {
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
    int o[3] = {0, 1, 2};
    std::sort(o, o + 3, [&](int x, int y) { return a[x][x] > a[y][y]; });
    double Vs[3][3];
    for (int k = 0; k < 3; ++k) {
      w[k] = a[o[k]][o[k]];
      for (int i = 0; i < 3; ++i) Vs[i][k] = V[i][o[k]];
    }
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) V[i][k] = Vs[i][k];
  }

  // mohrcoulomb_inline.h:301-457 with the returns of lines 34-289: main plane,
  // right edge, left edge (each kept when the truncated principal stresses
  // stay ordered), apex
  void update_mohr_coulomb(Material<R> &mt, int mat_id) {
    const R K = mt.bulk_modulus, G = mt.shear_modulus, H = mt.H;
    for (int p = 0; p < num_particles; ++p) {
      if (colors[p] != mat_id) continue;
      const M Lp = velocity_gradient[p];
      M deps;
      for (int k = 0; k < D * D; ++k)
        deps.m[k] = mt.is_velgrad_strain_increment ? Lp.m[k] : (R)0.5 * (Lp.m[k] + transpose(Lp).m[k]) * dt;
      const M eps_prev = block<R, D>(mt.eps_e[p]);
      M etr;
      for (int k = 0; k < D * D; ++k) etr.m[k] = eps_prev.m[k] + deps.m[k];
      const R acc_tr = mt.alpha[p];
      const R ev = trace(etr);
      const R ptr = K * ev;
      M sdev = etr;
      for (int i = 0; i < D; ++i) sdev(i, i) -= (R)((1 / 3.) * (double)ev);
      for (int k = 0; k < D * D; ++k) sdev.m[k] = (R)(2. * (double)G) * sdev.m[k];
      const R coh_tr = mt.cohesion + H * acc_tr;
      M3 sig = M3::zero();
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) sig(i, j) = sdev(i, j);
      for (int i = 0; i < 3; ++i) sig(i, i) += ptr;
      double A[3][3], w[3], V[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = sig(i, j);
      sym_eig3(A, w, V);
      const R ps[3] = {(R)w[0], (R)w[1], (R)w[2]};
      const R sfa = std::sin(mt.friction_angle), cfa = std::cos(mt.friction_angle);
      const R phi_tr = (R)((double)(ps[0] - ps[2] + (ps[0] + ps[2]) * sfa) - 2. * (double)coh_tr * (double)cfa);
      if (phi_tr <= 0) {
        stresses[p] = sig;
        if (mt.do_update_history) {
          M3 e = M3::zero();
          for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) e(i, j) = etr(i, j);
          mt.eps_e[p] = e;
          mt.alpha[p] = acc_tr;
        }
        continue;
      }
      const R sda = std::sin(mt.dilatancy_angle);
      const double Kd = K, Gd = G, Hd = H, sf = sfa, cf = cfa, sd = sda;
      const R a = (R)(4. * Gd * (1. + (1. / 3.) * sf * sd) + (double)(4 * K * sfa * sda));
      auto finish = [&](const R pc[3], double acc_next) {
        M3 out;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            double t = 0.0;
            for (int k = 0; k < 3; ++k) t += V[i][k] * (double)pc[k] * V[j][k];
            out(i, j) = (R)t;
          }
        stresses[p] = out;
        if (mt.do_update_history) {
          const R pn = (R)((double)trace(out) / 3.0);
          const R i2g = (R)(1. / (2. * Gd)), i3k = (R)(1. / (3. * Kd));
          M3 e = M3::zero();
          for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) e(i, j) = (out(i, j) - (i == j ? pn : R(0))) * i2g + (i == j ? pn * i3k : R(0));
          mt.eps_e[p] = e;
          mt.alpha[p] = (R)acc_next;
        }
      };
      auto ordered = [](const R pc[3]) {
        return std::trunc(pc[0]) >= std::trunc(pc[1]) && std::trunc(pc[1]) >= std::trunc(pc[2]);
      };
      {  // main plane (34-98)
        double dg = 0.0, ph = phi_tr, acc_next = acc_tr;
        const double d = -(double)a - 4 * Hd * cf * cf;
        for (int it = 0; it < kMaxIters; ++it) {
          dg = dg - ph / d;
          acc_next = (double)acc_tr + 2 * cf * dg;
          const double ca = (double)mt.cohesion + Hd * acc_next;
          ph = (double)ps[0] - (double)ps[2] + ((double)ps[0] + (double)ps[2]) * sf - 2. * ca * cf - (double)a * dg;
          if (!(std::fabs(ph) > 1.e-6)) break;
        }
        const R pc[3] = {(R)((double)ps[0] - dg * (2. * Gd * (1. + (1. / 3.) * sd) + 2. * Kd * sd)),
                         (R)((double)ps[1] + dg * ((4. / 3.) * Gd - 2. * Kd) * sd),
                         (R)((double)ps[2] + dg * (2. * Gd * (1. - (1. / 3.) * sd) - 2. * Kd * sd))};
        if (ordered(pc)) { finish(pc, acc_next); continue; }
      }
      bool done = false;
      for (int edge = 0; edge < 2 && !done; ++edge) {  // right edge, then left (100-229)
        const bool right = edge == 0;
        const R s1 = ps[0] - ps[2] + (ps[0] + ps[2]) * sfa;
        const R s2 = right ? ps[0] - ps[1] + (ps[0] + ps[1]) * sfa : ps[1] - ps[2] + (ps[1] + ps[2]) * sfa;
        double ph0 = (double)s1 - 2. * (double)coh_tr * cf, ph1 = (double)s2 - 2. * (double)coh_tr * cf;
        const double b = right ? 2 * Gd * (1. + sf + sd - (1. / 3.) * sf * sd) + 4. * Kd * sf * sd
                               : 2 * Gd * (1. - sf - sd - (1 / 3.) * sf * sd) + 4. * Kd * sf * sd;
        double g0 = 0.0, g1 = 0.0, acc_next = acc_tr;
        for (int it = 0; it < kMaxIters; ++it) {
          const double d00 = -(double)a - 4. * Hd * cf * cf, d01 = -b - (double)(4 * H * cfa * cfa);
          const double invdet = 1.0 / (d00 * d00 - d01 * d01);
          const double i00 = d00 * invdet, i01 = -d01 * invdet;
          const double n0 = g0 - (i00 * ph0 + i01 * ph1), n1 = g1 - (i01 * ph0 + i00 * ph1);
          g0 = n0;
          g1 = n1;
          acc_next = (double)acc_tr + 2 * cf * (g0 + g1);
          const double ca = (double)mt.cohesion + Hd * acc_next;
          ph0 = (double)s1 - 2. * ca * cf - (double)a * g0 - b * g1;
          ph1 = (double)s2 - 2. * ca * cf - (double)a * g1 - b * g0;
          if (!(std::fabs(ph0) + std::fabs(ph1) > 1.e-7)) break;
        }
        const double c1 = 2. * Gd * (1. + (1. / 3.) * sd) + 2. * Kd * sd, c2 = ((4. / 3.) * Gd - 2. * Kd) * sd,
                     c3 = 2. * Gd * (1. - (1. / 3.) * sd) - 2. * Kd * sd;
        R pc[3];
        if (right) {
          pc[0] = (R)((double)ps[0] - c1 * (g0 + g1));
          pc[1] = (R)((double)ps[1] + c2 * g0 + c3 * g1);
          pc[2] = (R)((double)ps[2] + c3 * g0 + c2 * g1);
        } else {
          pc[0] = (R)((double)ps[0] - c1 * g0 + c2 * g1);
          pc[1] = (R)((double)ps[1] + c2 * g0 - c1 * g1);
          pc[2] = (R)((double)ps[2] + c3 * (g0 + g1));
        }
        if (ordered(pc)) { finish(pc, acc_next); done = true; }
      }
      if (done) continue;
      // apex (231-270)
      double dv = 0.0, ph = phi_tr, pn = 0.0, acc_next = acc_tr;
      const double al = cf / sd, cotf = cf / sf;
      for (int it = 0; it < kMaxIters; ++it) {
        acc_next = (double)acc_tr + al * dv;
        const double ca = (double)mt.cohesion + Hd * acc_next;
        pn = (double)ptr - Kd * dv;
        ph = ca * cotf - pn;
        dv = dv - ph / ((Hd * cf * cotf) / sd + Kd);
        if (!(std::fabs(ph) > 1.e-6)) break;
      }
      stresses[p] = M3::zero();
      for (int i = 0; i < 3; ++i) stresses[p](i, i) = (R)pn;
      if (mt.do_update_history) {
        M3 e = M3::zero();
        for (int i = 0; i < D; ++i) e(i, i) = (R)(pn * (1. / (3. * Kd)));
        mt.eps_e[p] = e;
        mt.alpha[p] = (R)acc_next;
      }
    }
  }

  // modifiedcamclay_inline.h:269-413 (update_modifiedcamclay_nonlinear, the
  // update modifiedcamclay.cpp:147-176 runs). p compression positive.
  static R mcc_yield(R p, R Pt, R ps, R q, R M) {  // compute_yield_function (:34-41)
    return (ps - p - Pt) * (ps - p - Pt) + (q / M) * (q / M) - ps * ps;
  }
  void update_mcc(Material<R> &mt, int mat_id) {
    const R K = mt.bulk_modulus, G = mt.shear_modulus, Mm = mt.M, Pt = mt.Pt;
    for (int p = 0; p < num_particles; ++p) {
      if (colors[p] != mat_id) continue;
      const M Lp = velocity_gradient[p];
      const M eps_prev = block<R, D>(mt.eps_e[p]);
      M deps;
      if (mt.is_velgrad_strain_increment) {
        deps = Lp;
      } else {
        const M LT = transpose(Lp);
        M W;
        for (int k = 0; k < D * D; ++k) {
          deps.m[k] = (R)0.5 * (Lp.m[k] + LT.m[k]) * dt;
          W.m[k] = (R)0.5 * (Lp.m[k] - LT.m[k]) * dt;
        }
        const M We = matmul(W, eps_prev), eW = matmul(eps_prev, W);
        for (int k = 0; k < D * D; ++k) deps.m[k] += -We.m[k] + eW.m[k];
      }
      const M3 sref = mt.stress_ref[p];
      const R pref = -(R)(1.0 / 3.0) * trace(sref);
      M3 s_ref = sref;
      for (int i = 0; i < 3; ++i) s_ref(i, i) += pref;
      M etr;
      for (int k = 0; k < D * D; ++k) etr.m[k] = eps_prev.m[k] + deps.m[k];
      const R evt = -trace(etr);
      M3 dev = M3::zero();
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) dev(i, j) = etr(i, j);
      for (int i = 0; i < 3; ++i) dev(i, i) += (R)((1.0 / 3.0) * (double)evt);
      if (D < 3) dev(2, 2) = R(0);
      const R vspec = volumes[p] / mt.Vs;
      const R p_prev = (R)(-(double)trace(stresses[p]) / 3.0);
      const R v_kap = vspec / mt.kap;
      const R dev_tr = -trace(deps);
      const R ptr = (R)((double)p_prev / (1.0 - (double)(v_kap * dev_tr)));
      M3 st;
      for (int k = 0; k < 9; ++k) st.m[k] = s_ref.m[k] + (R)(2. * (double)G) * dev.m[k];
      const R q_tr = (R)std::sqrt(3 * 0.5 * (double)trace_mmt(st));
      const R pc_prev = mt.pc[p];
      const R ps_tr = (R)(0.5 * (double)(pc_prev + Pt));
      const R f_tr = mcc_yield(ptr, Pt, ps_tr, q_tr, Mm);
      if (f_tr <= (R)0.0) {
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) stresses[p](i, j) = st(i, j) - (i == j ? ptr : R(0));
        if (mt.do_update_history) {
          M3 e = M3::zero();
          for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) e(i, j) = etr(i, j);
          mt.eps_e[p] = e;
        }
        continue;
      }
      const R epv_prev = mt.alpha[p];
      double sol0 = 0.0, sol1 = epv_prev;
      double p_next = 0.0, q_next = 0.0, ps_next = 0.0, tol = 1e-2, conv = 1e10, sfs = 1.0;
      R devn = 0;
      const double vlt = (double)vspec / (double)(mt.lam - mt.kap);
      const double MM = (double)(Mm * Mm), G6 = (double)(6 * G), G12 = (double)(-12 * G);
      int counter = 0;
      do {
        const double mult = sol0, epv = sol1;
        const R dpv = (R)(epv - (double)epv_prev);
        devn = dev_tr - dpv;
        p_next = (double)p_prev / (1.0 - (double)(v_kap * devn));
        q_next = (MM / (MM + G6 * mult)) * (double)q_tr;
        ps_next = 0.5 * ((double)pc_prev * (1.0 + vlt * (double)dpv) + (double)Pt);
        sfs = MM / (MM + 6.0 * (double)G * mult);
        const double r0 = mcc_yield((R)p_next, Pt, (R)ps_next, (R)q_next, Mm);
        const double r1 = epv - (double)epv_prev + 2.0 * mult * (ps_next - p_next - (double)Pt);
        const double Hh = 0.5 * vlt * ((double)pc_prev / std::pow(1.0 - vlt * (double)dpv, 2));
        const double pover = ps_next - p_next - (double)Pt;
        const double d00 = (G12 / (MM + G6 * mult)) * (q_next / (double)Mm) * (q_next / (double)Mm);
        const double d01 = (2.0 * pover) * ((double)K + Hh) - 2.0 * ps_next * Hh;
        const double d10 = 2.0 * pover;
        const double d11 = 1.0 + (2.0 * mult) * ((double)K + Hh);
        const double invdet = 1.0 / (d00 * d11 - d10 * d01);
        const double i00 = d11 * invdet, i10 = -d10 * invdet, i01 = -d01 * invdet, i11 = d00 * invdet;
        sol0 = sol0 - (i00 * r0 + i01 * r1);
        sol1 = sol1 - (i10 * r0 + i11 * r1);
        conv = std::sqrt(r0 * r0 + r1 * r1);
        if (counter > 1000) tol *= 10;
        counter += 1;
      } while (std::fabs(conv) > tol && counter < 4096);
      M3 sn;
      for (int k = 0; k < 9; ++k) sn.m[k] = (R)sfs * st.m[k];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) stresses[p](i, j) = (i == j ? (R)((double)sn(i, j) - p_next) : sn(i, j));
      if (mt.do_update_history) {
        const R vterm = (R)((1. / 3.) * (double)(trace(eps_prev) + devn));
        const R g2 = (R)(2.0 * (double)G);
        M3 e = M3::zero();
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j) e(i, j) = (sn(i, j) - s_ref(i, j)) / g2 - (i == j ? vterm : R(0));
        mt.eps_e[p] = e;
        mt.alpha[p] = (R)sol1;
        mt.pc[p] = (R)(2.0 * ps_next - (double)Pt);
      }
    }
  }
  void stress_update_material(int mat_id) {
    Material<R> &mt = materials[mat_id];
    switch (mt.kind) {
      case MAT_LINEAR_ELASTIC: update_linear_elastic(mt, mat_id); break;
      case MAT_NEWTON_FLUID: update_newton_fluid(mt, mat_id); break;
      case MAT_LOCAL_RHEO: update_local_rheo(mt, mat_id); break;
      case MAT_DP: update_dp(mt, mat_id); break;
      case MAT_VON_MISES: update_von_mises(mt, mat_id); break;
      case MAT_MOHR_COULOMB: update_mohr_coulomb(mt, mat_id); break;
      case MAT_MCC: update_mcc(mt, mat_id); break;
    }
  }
  void stress_update() {  // solver.cpp:110-121
    for (int m = 0; m < (int)materials.size(); ++m) stress_update_material(m);
  }

  // ---------------- boundary conditions ----------------
  // gravity.cpp:65-88 -- the ramp overwrites the stored gravity while
  // step < ramp_step and leaves the last ramped value afterwards.
  void apply_gravity(BoundaryCondition<R, D> &bc) {
    const R ramp = ((R)global_step) / static_cast<R>(bc.ramp_step);
    if (bc.is_ramp && global_step < bc.ramp_step)
      for (int a = 0; a < D; ++a) bc.gravity[a] = bc.gravity_end[a] * ramp;
    for (int n = 0; n < grid.num_cells_total; ++n) {
      const R m = node_masses[n];
      if (m <= (R)0.000000001) continue;
      for (int a = 0; a < D; ++a) node_moments_nt[n][a] += bc.gravity[a] * m * dt;
    }
  }
  // nodedomain_inline.h:40-87 (0 = fixed, 1 = slip; skip m <= 1e-5)
  void apply_node_domain(const BoundaryCondition<R, D> &bc) {
    for (int n = 0; n < grid.num_cells_total; ++n) {
      if (node_masses[n] <= (R)0.00001) continue;
      const VI nb = node_ids[n];
      for (int i = 0; i < D; ++i) {
        if (nb[i] < 1) {
          if (bc.face0_mode[i] == 0) {
            node_moments[n] = V::zero(); node_moments_nt[n] = V::zero();
          } else if (bc.face0_mode[i] == 1) {
            node_moments[n][i] = (R)std::max(0., (double)node_moments[n][i]);
            node_moments_nt[n][i] = (R)std::max(0., (double)node_moments_nt[n][i]);
          }
        } else if (nb[i] >= grid.num_cells[i] - 1) {
          if (bc.face1_mode[i] == 0) {
            node_moments[n] = V::zero(); node_moments_nt[n] = V::zero();
          } else if (bc.face1_mode[i] == 1) {
            node_moments[n][i] = (R)std::min(0., (double)node_moments[n][i]);
            node_moments_nt[n][i] = (R)std::min(0., (double)node_moments_nt[n][i]);
          }
        }
      }
    }
  }
  // rigidbodylevelset.cpp:152-217. STATIC leaves the reference's rotation
  // matrix uninitialised; here it is the identity (rigid velocity 0). In 3D
  // the reference's Rz is never filled, so R = Rz*Ry*Rx = 0.
  void rigid_update_animation(BoundaryCondition<R, D> &bc) {
    if (bc.rigid_mode != 1) return;
    const V prev = bc.euler_angles;
    for (int a = 0; a < D; ++a) bc.euler_angles[a] = bc.euler_angles[a] + bc.euler_angles_per_second[a] * dt;
    V de;
    for (int a = 0; a < D; ++a) de[a] = prev[a] - bc.euler_angles[a];
    M3 R3 = M3::zero();
    if (D == 2) {
      R3(0, 0) = (R)std::cos((double)de[0]); R3(0, 1) = (R)-std::sin((double)de[0]);
      R3(1, 0) = (R)std::sin((double)de[0]); R3(1, 1) = (R)std::cos((double)de[0]);
      R3(2, 2) = 1;
    }
    bc.rotation_matrix = block<R, D>(R3);
  }
  // rigidbodylevelset_inline.h:65-199 (normal from non-rigid dpsi*m with the
  // window index hard-coded to the linear size 2^D, nearest rigid particle in
  // cell-relative squared distance, stick contact if (v - v_b).n >= 0).
  void rigid_grid_normals() {
    const int lin = (D == 3) ? 8 : (D == 2 ? 4 : 2);
    const Windows lw = make_windows(D, SF_LINEAR);
    for (int n = 0; n < grid.num_cells_total; ++n) {
      const R nm = node_masses[n];
      if (nm <= (R)0.000000001) continue;
      const VI nb = node_ids[n];
      V normal = V::zero();
      R min_dist = (R)999999999999999.;
      int min_id = -1;
      for (int sid = 0; sid < lin; ++sid) {
        VI sb;
        for (int a = 0; a < D; ++a) sb[a] = nb[a] + lw.p2g[sid][a];
        const unsigned h = (unsigned)grid.linear(sb);
        if (h >= (unsigned)grid.num_cells_total) continue;
        const int cs = cell_start[h], ce = cell_end[h];
        if (cs < 0 || ce < 0) continue;
        for (int j = cs; j < ce; ++j) {
          const int p = sorted_index[j];
          if (is_rigid[p]) {
            V rp;
            R d2 = R(0);
            for (int a = 0; a < D; ++a) {
              rp[a] = (positions[p][a] - grid.origin[a]) * grid.inv_cell_size - (R)sb[a];
              d2 += rp[a] * rp[a];
            }
            if (d2 < min_dist) { min_dist = d2; min_id = p; }
          } else {
            const V g = dpsi[(size_t)p * lin + sid];
            for (int a = 0; a < D; ++a) normal[a] += g[a] * masses[p];
          }
        }
      }
      if (min_id == -1) continue;
      const V vb = velocities[min_id];
      R nn2 = R(0);
      for (int a = 0; a < D; ++a) nn2 += normal[a] * normal[a];
      if (nn2 > R(0)) { const R inv = (R)std::sqrt((double)nn2); for (int a = 0; a < D; ++a) normal[a] /= inv; }
      V nv, nvt, cv, cvt;
      R rc = R(0), rct = R(0);
      for (int a = 0; a < D; ++a) {
        nv[a] = node_moments[n][a] / nm;
        nvt[a] = node_moments_nt[n][a] / nm;
        cv[a] = nv[a] - vb[a];
        cvt[a] = nvt[a] - vb[a];
        rc += cv[a] * normal[a];
        rct += cvt[a] * normal[a];
      }
      if (rc >= 0) for (int a = 0; a < D; ++a) node_moments[n][a] = (nv[a] - cv[a]) * nm;
      if (rct >= 0) for (int a = 0; a < D; ++a) node_moments_nt[n][a] = (nvt[a] - cvt[a]) * nm;
    }
  }
  void apply_rigid_levelset(BoundaryCondition<R, D> &bc) {  // rigidbodylevelset.cpp:92-105
    rigid_update_animation(bc);
    for (int p = 0; p < num_particles; ++p) {  // update_rigid_velocity (inline.h:241-299)
      if (!is_rigid[p]) continue;
      V r, rr;
      for (int a = 0; a < D; ++a) r[a] = positions[p][a] - bc.COM[a];
      for (int a = 0; a < D; ++a) {
        R s = R(0);
        for (int b = 0; b < D; ++b) s += bc.rotation_matrix(a, b) * r[b];
        rr[a] = s;
      }
      for (int a = 0; a < D; ++a) velocities[p][a] = (rr[a] - r[a]) / dt;
    }
    rigid_grid_normals();
    for (int p = 0; p < num_particles; ++p) {  // update_rigid_position (inline.h:329-343)
      if (!is_rigid[p]) continue;
      for (int a = 0; a < D; ++a) positions[p][a] += velocities[p][a] * dt;
    }
  }
  // Analytic box SDF with prescribed velocity (no reference implementation):
  // the reference's contact rule with n = -grad(phi), applied to nodes with
  // phi < band; the body then advances by v_b * dt.
  void apply_rigid_sdf(BoundaryCondition<R, D> &bc, bool advance) {
    for (int n = 0; n < grid.num_cells_total; ++n) {
      const R nm = node_masses[n];
      if (nm <= (R)0.000000001) continue;
      V xn;
      for (int a = 0; a < D; ++a) xn[a] = grid.origin[a] + (R)node_ids[n][a] * grid.cell_size;
      V grad;
      const R phi = box_sdf<R, D>(bc, xn, grad);
      if (!(phi < bc.sdf_band)) continue;
      V nv, nvt;
      R rc = R(0), rct = R(0);
      for (int a = 0; a < D; ++a) {
        nv[a] = node_moments[n][a] / nm;
        nvt[a] = node_moments_nt[n][a] / nm;
        rc += (nv[a] - bc.sdf_velocity[a]) * -grad[a];
        rct += (nvt[a] - bc.sdf_velocity[a]) * -grad[a];
      }
      if (rc >= 0) for (int a = 0; a < D; ++a) node_moments[n][a] = bc.sdf_velocity[a] * nm;
      if (rct >= 0) for (int a = 0; a < D; ++a) node_moments_nt[n][a] = bc.sdf_velocity[a] * nm;
    }
    if (advance)
      for (int a = 0; a < D; ++a) bc.sdf_center[a] += bc.sdf_velocity[a] * dt;
  }
  // planardomain_inline.h:56-138 (one particle; apply_on_particles,
  // planardomain.cpp:40-66). The radius is half the D-th root of the volume
  // (pow of float arguments), the scalar m / dt^2 * overlap is a double that
  // scales the float vector n - mu (v - (n.v) n), as Eigen rounds a double
  // scalar to the vector's float before the product.
  void apply_planar_domain(const BoundaryCondition<R, D> &bc) {
    for (int p = 0; p < num_particles; ++p) {
      const V pos = positions[p], vel = velocities[p];
      const R vol = volumes[p];
      const R inv_dim = (R)1. / (R)D;
      const double pow_vol = std::pow(vol, inv_dim);
      const R radius = (R)0.5 * (R)pow_vol;
      const R mass = masses[p];
      for (int face = 0; face < 2; ++face) {
        for (int i = 0; i < D; ++i) {
          const R gap = face == 0 ? pos[i] - grid.origin[i] : grid.end[i] - pos[i];
          const R overlap = radius - gap;
          if (!(overlap > 0)) continue;
          V normal = V::zero();
          normal[i] = face == 0 ? R(1) : R(-1);
          R ndotv = R(0);
          for (int a = 0; a < D; ++a) ndotv += normal[a] * vel[a];
          const R fric = face == 0 ? bc.face0_friction[i] : bc.face1_friction[i];
          V fric_term;
          for (int a = 0; a < D; ++a) fric_term[a] = normal[a] - fric * (vel[a] - ndotv * normal[a]);
          const R scale = (R)(((double)mass / std::pow((double)dt, 2.)) * (double)overlap);
          for (int a = 0; a < D; ++a) forces_external[p][a] += scale * fric_term[a];
        }
      }
    }
  }
  // apply_on_particles hooks in list order (usl.cpp:79-84, 109-114)
  void apply_particle_bcs() {
    for (auto &bc : bcs)
      if (bc.kind == BC_PLANAR_DOMAIN) apply_planar_domain(bc);
  }
  // bodyforce_inline.h:41-59 (bodyforce.cpp:56-87): no mass test. In the
  // MUSL re-map pass only the kinematic "fixed" mode applies again.
  void apply_body_force(const BoundaryCondition<R, D> &bc, bool second) {
    for (int n = 0; n < grid.num_cells_total; ++n) {
      if (!bc.body_mask[n]) continue;
      if (bc.body_mode == 2) node_moments_nt[n] = bc.body_values[n];
      else if (!second)
        for (int a = 0; a < D; ++a) node_moments[n][a] += bc.body_values[n][a];
    }
  }

  // apply_on_nodes_moments hooks in list order (usl.cpp:96-103). The MUSL
  // re-map pass (second=true) applies only the kinematic constraints: gravity
  // is already in the particle velocities and the SDF body does not advance.
  void apply_node_bcs(bool second) {
    for (auto &bc : bcs) {
      switch (bc.kind) {
        case BC_GRAVITY: if (!second) apply_gravity(bc); break;
        case BC_NODE_DOMAIN: apply_node_domain(bc); break;
        case BC_RIGID_LEVELSET: if (!second) apply_rigid_levelset(bc); break;
        case BC_RIGID_SDF: apply_rigid_sdf(bc, !second); break;
        case BC_BODY_FORCE: apply_body_force(bc, second); break;
        case BC_PLANAR_DOMAIN: break;  // a particle condition
      }
    }
  }

  // ---------------- solver (solver.cpp:69-105, usl.cpp:69-115) ----------------
  void initialize() {
    partition();
    calculate_initial_volumes();
    for (int m = 0; m < (int)materials.size(); ++m) {
      calculate_initial_masses(m, materials[m].density);
      if (materials[m].kind == MAT_DP) dp_initialize(materials[m]);
      if (materials[m].kind == MAT_VON_MISES || materials[m].kind == MAT_MOHR_COULOMB)
        plastic_initialize(materials[m]);
      if (materials[m].kind == MAT_MCC) mcc_initialize(materials[m]);
    }
  }
  void solve() {
    reset();
    partition();
    calculate_shape_function();
    apply_particle_bcs();
    p2g();
    apply_node_bcs(false);
    if (!musl) {
      g2p(0);
    } else {
      g2p(1);
      p2g_double_mapping();
      apply_node_bcs(true);
      g2p(2);
    }
    stress_update();
    apply_particle_bcs();  // (forces the next reset clears)
  }
  void step() { solve(); global_step += 1; }
  // Solver::run executes total_steps + 1 solves (solver.cpp:140-174)
  void run(int total_steps) { for (int s = 0; s < total_steps + 1; ++s) step(); }
};

}  // namespace oracle
