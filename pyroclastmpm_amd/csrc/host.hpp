// host.hpp -- C++ host objects mirroring the reference's ParticlesContainer,
// NodesContainer, Solver/USL/MUSL, materials and boundary conditions
// (particles.h:94-271, nodes.h:73-144, solver.h:94-156, usl.h:85-114,
// materials/*.h, boundaryconditions/*.h). They own SoA device buffers and
// drive the kernels exclusively through the C ABI of include/pmpm.h. Getters
// return host copies in the reference's particle order, setters copy in.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "../../include/pmpm.h"

namespace pmpm_host {

// ---------------------------------------------------------------------------
// Process-wide settings of the reference API (global_settings.cpp:49-98,
// 161-281). The C ABI itself has no globals; these only feed the host
// objects, read at every solve() like the reference reads its globals.
// ---------------------------------------------------------------------------
struct Globals {
  float dt = 0.f;
  int particles_per_cell = 1;
  int shape_function = PMPM_SF_LINEAR;
  std::string output_directory = "./output/";
  int step = 0;
};
Globals &globals();
int shape_function_from_string(const std::string &s);
void set_globals(float dt, int ppc, const std::string &shape_function, const std::string &output_dir);
void set_global_shapefunction(const std::string &s);
void set_global_timestep(float dt);
void set_global_step(int step);
void set_global_output_directory(const std::string &d);
void set_global_particles_per_cell(int ppc);

void check(int rc, const char *what);
void hip_check(hipError_t e, const char *what);

// Owning device buffer with value semantics (the reference's containers are
// copied by value into solvers and through pybind).
template <class T> class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  DevBuf(const DevBuf &o) { *this = o; }
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf &operator=(const DevBuf &o) {
    if (this == &o) return *this;
    resize(o.n_);
    if (n_) hip_check(hipMemcpy(p_, o.p_, n_ * sizeof(T), hipMemcpyDeviceToDevice), "copy");
    return *this;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) hip_check(hipMalloc(&p_, n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void upload(const std::vector<T> &h) {
    resize(h.size());
    if (n_) hip_check(hipMemcpy(p_, h.data(), n_ * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
  std::vector<T> download() const {
    std::vector<T> h(n_);
    if (n_) hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "download");
    return h;
  }
  T *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T *p_ = nullptr;
  size_t n_ = 0;
};

template <int D> using VecD = std::array<float, D>;
template <int D> using VecI = std::array<int, D>;
template <int D> using MatD = std::array<float, D * D>;  // row-major
using Mat3 = std::array<float, 9>;                       // row-major

// ---------------------------------------------------------------------------
// Nodes (nodes.cpp:55-203).
// ---------------------------------------------------------------------------
template <int D> class NodesContainer {
 public:
  NodesContainer() = default;
  NodesContainer(const VecD<D> &origin, const VecD<D> &end, float cell_size);
  void reset();
  std::vector<VecD<D>> give_node_coords() const;
  std::vector<VecD<D>> moments() const;
  std::vector<VecD<D>> moments_nt() const;
  std::vector<float> masses() const;
  void set_moments(const std::vector<VecD<D>> &v);
  void set_moments_nt(const std::vector<VecD<D>> &v);
  void set_masses(const std::vector<float> &v);
  std::vector<VecI<D>> node_ids() const;
  std::vector<VecI<D>> node_types() const;
  pmpm_nodes view() const;
  // nodes.cpp:214-244: writes nodes<step>.vtp / .csv for each format
  void output_vtk() const;

  pmpm_grid grid{};
  VecD<D> origin{}, end{};
  float cell_size = 1.f;
  bool node_types_built = false;  // cubic spline selected when constructed
  float small_mass_cutoff = 0.000000001f;
  std::vector<std::string> output_formats;
  DevBuf<float> mass, mom, mom_nt, vel;  // vel: derived records for the gathers
};

// ---------------------------------------------------------------------------
// Particles (particles.cpp:51-421). Two buffer sets: the G2P kernels read one
// and write the other in cell order; id[] maps slots back to the reference's
// particle order for every getter and setter.
// ---------------------------------------------------------------------------
struct ParticleBuffers {
  DevBuf<float> x, v, stress, F, L, mass, volume, volume0, eps_e, stress_ref;
  DevBuf<float> hist;  // [2][N] plastic history scalars (von Mises, Mohr-Coulomb, Cam-Clay)
  DevBuf<uint8_t> color, flags;
  DevBuf<int32_t> id;
};

struct SpatialBuffers {
  DevBuf<uint32_t> hash, hash_sorted;
  DevBuf<int32_t> sorted, cell_start, cell_end, bins;
};

template <int D> class ParticlesContainer {
 public:
  ParticlesContainer() = default;
  ParticlesContainer(const std::vector<VecD<D>> &positions, const std::vector<VecD<D>> &velocities,
                     const std::vector<int> &colors, const std::vector<VecD<D>> &rigid_positions,
                     const std::vector<VecD<D>> &rigid_velocities, const std::vector<int> &rigid_colors);

  // reference-order host copies
  std::vector<VecD<D>> positions() const;
  std::vector<VecD<D>> velocities() const;
  std::vector<Mat3> stresses() const;
  std::vector<MatD<D>> F() const;
  std::vector<MatD<D>> velocity_gradient() const;
  std::vector<float> masses() const;
  std::vector<float> volumes() const;
  std::vector<float> volumes_original() const;
  std::vector<int> colors() const;
  std::vector<int> is_rigid() const;
  std::vector<int> is_active() const;
  std::vector<Mat3> eps_e() const;
  // [2][N]: accumulated or volumetric plastic strain, preconsolidation pressure
  std::vector<float> plastic_history() const;
  void set_positions(const std::vector<VecD<D>> &v);
  void set_velocities(const std::vector<VecD<D>> &v);
  void set_stresses(const std::vector<Mat3> &v);
  void set_F(const std::vector<MatD<D>> &v);
  void set_velocity_gradient(const std::vector<MatD<D>> &v);
  void set_masses(const std::vector<float> &v);
  void set_volumes(const std::vector<float> &v);
  void set_volumes_original(const std::vector<float> &v);
  void set_colors(const std::vector<int> &v);
  void set_is_active(const std::vector<int> &v);
  void set_eps_e(const std::vector<Mat3> &v);

  void set_output_formats(const std::vector<std::string> &f) { output_formats = f; }
  // particles.cpp:209-320: writes particles<step>.vtp / .csv for each format
  void output_vtk() const;
  void set_spawner(int rate, int volume);
  void spawn_particles(int global_step);
  void set_spatialpartition(const pmpm_grid &g);
  void partition();  // reference-order hash / stable sort / cell ranges
  void calculate_initial_volumes();
  void calculate_initial_masses(int mat_id, float density);
  void ensure_history();  // Drucker-Prager eps_e / stress_ref buffers
  void ensure_plastic_history();  // + the history scalars of the post-G2P materials

  // spatial partition of the last partition() call, reference semantics
  std::vector<VecI<D>> bins() const;
  std::vector<uint32_t> hash_unsorted() const;
  std::vector<uint32_t> hash_sorted() const;
  std::vector<int32_t> sorted_index() const;
  std::vector<int32_t> cell_start() const;
  std::vector<int32_t> cell_end() const;

  pmpm_particles view(int which) const;
  pmpm_spatial spatial_view(SpatialBuffers &sb) const;
  ParticleBuffers &cur() { return buf[cur_]; }
  const ParticleBuffers &cur() const { return buf[cur_]; }
  void swap_buffers() { cur_ ^= 1; }
  int cur_index() const { return cur_; }
  void ensure_second_set();

  int num_particles = 0;
  int num_rigid_particles = 0;
  bool isRestart = false;
  std::vector<std::string> output_formats;
  pmpm_grid grid{};
  bool has_grid = false;
  struct Spawner { int rate = 0, volume = 0, increment = 0; } spawner;
  ParticleBuffers buf[2];
  SpatialBuffers spatial;  // API partition (reference order)
  SpatialBuffers work;     // solver partition (slot order)
  pmpm_context *ctx() const;

 private:
  template <class T> std::vector<T> to_ref(const std::vector<T> &slot_order, int ncomp) const;
  template <class T> std::vector<T> from_ref(const std::vector<T> &ref_order, int ncomp) const;
  int cur_ = 0;
  mutable std::shared_ptr<pmpm_context> ctx_;
};

// ---------------------------------------------------------------------------
// Materials (materials/*.cpp constructors).
// ---------------------------------------------------------------------------
struct LinearElastic {
  LinearElastic() : LinearElastic(0.f, 0.f, 0.f) {}
  LinearElastic(float density, float E, float pois);
  float density, E, pois, shear_modulus, lame_modulus, bulk_modulus;
  pmpm_material device() const;
};
struct NewtonFluid {
  NewtonFluid(float density, float viscosity, float bulk_modulus, float gamma);
  float density, viscosity, bulk_modulus, gamma;
  pmpm_material device() const;
};
struct LocalGranularRheology {
  LocalGranularRheology(float density, float E, float pois, float I0, float mu_s, float mu_2, float rho_c,
                        float particle_diameter, float particle_density);
  float density, E, pois, I0, mu_s, mu_2, rho_c, particle_diameter, particle_density, EPS;
  float shear_modulus, lame_modulus, bulk_modulus;
  pmpm_material device() const;
};
struct DP_rheo {
  DP_rheo(float density, float E, float pois, float M, float lam, float kap, float Vs, float R, float Pt, float beta,
          float dilatancy_angle, float solid_volume_fraction);
  float density, E, pois, M, lam, kap, Vs, R, Pt, beta, solid_volume_fraction;
  float dilatancy_angle, eta_overline, shear_modulus, bulk_modulus, lame_modulus = 0.f;
  bool do_update_history = true;
  bool is_velgrad_strain_increment = false;
  bool apex_guard = true;
  bool stress_ref_nonzero = false;
  // history of a stand-alone initialize()/stress_update() (the reference keeps
  // it in the material: dp_rheo.h:139-160); solvers keep it with the particles
  std::vector<float> pc;
  pmpm_material device() const;
};
// vonmises.cpp:34-50 (3D only)
struct VonMises {
  VonMises(float density, float E, float pois, float yield_stress, float H);
  float density, E, pois, yield_stress, H, shear_modulus, bulk_modulus;
  bool do_update_history = true;
  bool is_velgrad_strain_increment = false;
  pmpm_material device() const;
};
// mohrcoulomb.cpp:33-52 (angles given in degrees, kept in radians)
struct MohrCoulomb {
  MohrCoulomb(float density, float E, float pois, float cohesion, float friction_angle, float dilatancy_angle,
              float H);
  float density, E, pois, cohesion, friction_angle, dilatancy_angle, H, shear_modulus, bulk_modulus, lame_modulus;
  bool do_update_history = true;
  bool is_velgrad_strain_increment = false;
  pmpm_material device() const;
};
// modifiedcamclay.cpp:47-63 (beta is accepted and unused, as there)
struct ModifiedCamClay {
  ModifiedCamClay(float density, float E, float pois, float M, float lam, float kap, float Vs, float R, float Pt,
                  float beta);
  float density, E, pois, M, lam, kap, Vs, R, Pt, beta, shear_modulus, bulk_modulus;
  bool do_update_history = true;
  bool is_velgrad_strain_increment = false;
  pmpm_material device() const;
};
using MaterialType = std::variant<LinearElastic, NewtonFluid, LocalGranularRheology, DP_rheo, VonMises, MohrCoulomb,
                                  ModifiedCamClay>;

// ---------------------------------------------------------------------------
// Boundary conditions.
// ---------------------------------------------------------------------------
template <int D> struct Gravity {
  Gravity() : gravity{}, is_ramp(false), ramp_step(0), gravity_end{} {}
  Gravity(const VecD<D> &g, bool is_ramp, int ramp_step, const VecD<D> &g_end)
      : gravity(g), is_ramp(is_ramp), ramp_step(ramp_step), gravity_end(g_end) {}
  VecD<D> gravity;
  bool is_ramp;
  int ramp_step;
  VecD<D> gravity_end;
  pmpm_bc device(int global_step);  // applies the ramp (gravity.cpp:65-76)
};
template <int D> struct NodeDomain {
  NodeDomain(const VecI<D> &f0, const VecI<D> &f1) : face0_mode(f0), face1_mode(f1) {}
  VecI<D> face0_mode, face1_mode;
  pmpm_bc device() const;
};
template <int D> struct RigidBodyLevelSet {
  RigidBodyLevelSet();
  void set_mode_loop_rotate(const VecD<D> &euler_angles_per_second, const VecD<D> &COM);
  void update_animation(float dt);  // rigidbodylevelset.cpp:152-217
  int mode = 0;                     // 0 STATIC, 1 LOOP_ROTATE
  VecD<D> euler_angles{}, euler_angles_per_second{}, COM{}, angular_velocities{};
  MatD<D> rotation_matrix{};
};
template <int D> struct RigidBodySDF {
  RigidBodySDF(const std::array<float, 3> &half_extents, const std::array<float, 3> &center,
               const std::array<float, 3> &velocity, const std::array<float, 9> &rotation, float band);
  std::array<float, 3> half_extents, center, velocity;
  std::array<float, 9> rotation;
  float band;
  pmpm_bc device() const;
};
// bodyforce.cpp:28-87: per-node values added to the moments ("forces",
// "moments") or imposed on the forward moments ("fixed") where mask is set.
template <int D> struct BodyForce {
  BodyForce(const std::string &mode, const std::vector<VecD<D>> &values, const std::vector<bool> &mask);
  int mode_id;
  bool isActive = true;
  std::vector<VecD<D>> values;
  std::vector<bool> mask;
  // device copies ([D][nn] values, [nn] mask), made on first use and shared
  // by copies of this object
  std::shared_ptr<DevBuf<float>> values_dev;
  std::shared_ptr<DevBuf<uint8_t>> mask_dev;
  pmpm_bc device(int num_nodes);
};
// planardomain.cpp:27-33: frictional walls on the faces of the domain box;
// the angles are given in degrees and kept in radians, which the contact uses
// as friction coefficients (planardomain_inline.h:108-136).
template <int D> struct PlanarDomain {
  PlanarDomain(const VecD<D> &face0_friction_deg, const VecD<D> &face1_friction_deg);
  VecD<D> face0_friction, face1_friction;
  bool isActive = true;
  pmpm_bc device(const VecD<D> &domain_end) const;
};
template <int D>
using BoundaryConditionType = std::variant<Gravity<D>, NodeDomain<D>, RigidBodyLevelSet<D>, RigidBodySDF<D>,
                                           BodyForce<D>, PlanarDomain<D>>;

// ---------------------------------------------------------------------------
// Solvers (solver.cpp:69-199, usl.cpp:48-199, MUSL from test_musl.h).
// ---------------------------------------------------------------------------
template <int D> class Solver {
 public:
  Solver(const ParticlesContainer<D> &particles, const NodesContainer<D> &nodes,
         const std::vector<MaterialType> &materials, const std::vector<BoundaryConditionType<D>> &bcs, float alpha,
         bool musl);
  virtual ~Solver();
  void solve();
  void solve_nsteps(int n);
  void run(int total_steps, int output_frequency);
  void stress_update();
  void output();
  // stepwise pieces of the reference's API (usl.cpp:118-199), on the current
  // partition of the current positions, in place
  void partition();
  void P2G();
  void G2P();
  void G2P_double_mapping();
  void P2G_double_mapping();
  void G2P_second();
  void apply_boundary_conditions(bool remap_pass);
  pmpm_params params() const;
  void synchronize() const;

  ParticlesContainer<D> particles;
  NodesContainer<D> nodes;
  std::vector<MaterialType> materials;
  std::vector<BoundaryConditionType<D>> boundaryconditions;
  float alpha;
  bool musl;
  bool wrap_quirk = true;
  hipStream_t stream = nullptr;

 private:
  void step_fused();
  void step_composed();
  void initialize_materials();
  void initialize_plastic_history();
  bool has_levelset() const;
  pmpm_params last_prm_{};
  DevBuf<float> dpsi_orig_, x_orig_, psi_tmp_, rigid_dx_;
};

}  // namespace pmpm_host
