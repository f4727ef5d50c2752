// host.cpp -- implementation of the host objects of host.hpp. Every piece of
// per-particle or per-node work runs in libpmpm_hip.so through the C ABI; the
// host code here only moves data in and out and sequences the calls.
#include "host.hpp"

#include <cmath>
#include <cstring>
#include <numeric>

namespace pmpm_host {

Globals &globals() {
  static Globals g;
  return g;
}
int shape_function_from_string(const std::string &s) {
  if (s == "linear") return PMPM_SF_LINEAR;
  if (s == "quadratic") return PMPM_SF_QUADRATIC;
  if (s == "cubic") return PMPM_SF_CUBIC;
  // the reference prints and calls exit(1) (global_settings.cpp:269-273)
  throw std::invalid_argument("Shape function not recognized: " + s);
}
void set_globals(float dt, int ppc, const std::string &sf, const std::string &out) {
  set_global_timestep(dt);
  set_global_shapefunction(sf);
  set_global_output_directory(out);
  set_global_particles_per_cell(ppc);
}
void set_global_shapefunction(const std::string &s) { globals().shape_function = shape_function_from_string(s); }
void set_global_timestep(float dt) { globals().dt = dt; }
void set_global_step(int step) { globals().step = step; }
void set_global_output_directory(const std::string &d) { globals().output_directory = d; }
void set_global_particles_per_cell(int ppc) { globals().particles_per_cell = ppc; }

void check(int rc, const char *what) {
  if (rc != PMPM_OK) throw std::runtime_error(std::string("pmpm: ") + what + " failed with code " + std::to_string(rc));
}
void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

namespace {
pmpm_grid make_grid(const float *origin, const float *end, float h, int D) {
  pmpm_grid g{};
  g.cell_size = h;
  g.inv_cell_size = (float)1.0 / h;
  g.num_cells_total = 1;
  for (int a = 0; a < 3; ++a) {
    g.origin[a] = a < D ? origin[a] : 0.f;
    g.num_cells[a] = 1;
  }
  for (int a = 0; a < D; ++a) {
    g.num_cells[a] = (int)((end[a] - origin[a]) / h) + 1;  // types_common.h:268-273
    g.num_cells_total *= g.num_cells[a];
  }
  return g;
}
std::shared_ptr<pmpm_context> new_context() {
  pmpm_context *c = nullptr;
  check(pmpm_context_create(&c), "pmpm_context_create");
  return std::shared_ptr<pmpm_context>(c, [](pmpm_context *p) { pmpm_context_destroy(p); });
}
}  // namespace

// ============================ NodesContainer ===============================
template <int D>
NodesContainer<D>::NodesContainer(const VecD<D> &o, const VecD<D> &e, float h) : origin(o), end(e), cell_size(h) {
  grid = make_grid(o.data(), e.data(), h, D);
  node_types_built = globals().shape_function == PMPM_SF_CUBIC;  // nodes.cpp:140-143
  const size_t nn = grid.num_cells_total;
  mass.resize(nn);
  mom.resize(nn * D);
  mom_nt.resize(nn * D);
  vel.resize(nn * PMPM_NODE_VEL_STRIDE(D));
  reset();
}
template <int D> void NodesContainer<D>::reset() {
  hip_check(hipMemset(mass.get(), 0, mass.size() * 4), "memset");
  hip_check(hipMemset(mom.get(), 0, mom.size() * 4), "memset");
  hip_check(hipMemset(mom_nt.get(), 0, mom_nt.size() * 4), "memset");
  hip_check(hipMemset(vel.get(), 0, vel.size() * 4), "memset");
}
template <int D> std::vector<VecI<D>> NodesContainer<D>::node_ids() const {
  std::vector<VecI<D>> ids(grid.num_cells_total);
  for (int i = 0; i < grid.num_cells_total; ++i) {
    ids[i][0] = i % grid.num_cells[0];
    if (D > 1) ids[i][1] = (i / grid.num_cells[0]) % grid.num_cells[1];
    if (D > 2) ids[i][2] = i / (grid.num_cells[0] * grid.num_cells[1]);
  }
  return ids;
}
template <int D> std::vector<VecI<D>> NodesContainer<D>::node_types() const {
  auto ids = node_ids();
  std::vector<VecI<D>> t(ids.size());
  for (size_t i = 0; i < ids.size(); ++i)
    for (int a = 0; a < D; ++a) {
      int v = 0;
      if (node_types_built) {
        const int id = ids[i][a];
        v = (id == 0 || id == grid.num_cells[a] - 1) ? 1 : (id == 1 ? 2 : 3);
      }
      t[i][a] = v;
    }
  return t;
}
template <int D> std::vector<VecD<D>> NodesContainer<D>::give_node_coords() const {
  auto ids = node_ids();
  std::vector<VecD<D>> c(ids.size());
  for (size_t i = 0; i < ids.size(); ++i)
    for (int a = 0; a < D; ++a) c[i][a] = origin[a] + (float)ids[i][a] * cell_size;
  return c;
}
template <int D> static std::vector<VecD<D>> soa_to_vecs(const std::vector<float> &s, size_t n) {
  std::vector<VecD<D>> v(n);
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < D; ++a) v[i][a] = s[a * n + i];
  return v;
}
template <int D> static std::vector<float> vecs_to_soa(const std::vector<VecD<D>> &v) {
  const size_t n = v.size();
  std::vector<float> s(n * D);
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < D; ++a) s[a * n + i] = v[i][a];
  return s;
}
template <int D> std::vector<VecD<D>> NodesContainer<D>::moments() const {
  return soa_to_vecs<D>(mom.download(), grid.num_cells_total);
}
template <int D> std::vector<VecD<D>> NodesContainer<D>::moments_nt() const {
  return soa_to_vecs<D>(mom_nt.download(), grid.num_cells_total);
}
template <int D> std::vector<float> NodesContainer<D>::masses() const { return mass.download(); }
template <int D> void NodesContainer<D>::set_moments(const std::vector<VecD<D>> &v) {
  if ((int)v.size() != grid.num_cells_total) throw std::invalid_argument("moments: wrong size");
  mom.upload(vecs_to_soa<D>(v));
}
template <int D> void NodesContainer<D>::set_moments_nt(const std::vector<VecD<D>> &v) {
  if ((int)v.size() != grid.num_cells_total) throw std::invalid_argument("moments_nt: wrong size");
  mom_nt.upload(vecs_to_soa<D>(v));
}
template <int D> void NodesContainer<D>::set_masses(const std::vector<float> &v) {
  if ((int)v.size() != grid.num_cells_total) throw std::invalid_argument("masses: wrong size");
  mass.upload(v);
}
template <int D> pmpm_nodes NodesContainer<D>::view() const {
  return pmpm_nodes{mass.get(), mom.get(), mom_nt.get(), vel.get()};
}

// ============================ ParticlesContainer ===========================
template <int D>
ParticlesContainer<D>::ParticlesContainer(const std::vector<VecD<D>> &pos, const std::vector<VecD<D>> &vel,
                                          const std::vector<int> &col, const std::vector<VecD<D>> &rpos,
                                          const std::vector<VecD<D>> &rvel, const std::vector<int> &rcol) {
  const int nn = (int)pos.size(), nr = (int)rpos.size();
  num_rigid_particles = nr;
  num_particles = nn + nr;
  const int N = num_particles;
  if (!vel.empty() && (int)vel.size() != nn) throw std::invalid_argument("velocities: wrong size");
  if (!col.empty() && (int)col.size() != nn) throw std::invalid_argument("colors: wrong size");
  std::vector<float> x(N * D, 0.f), v(N * D, 0.f);
  std::vector<uint8_t> color(N, 0), flags(N, PMPM_PF_ACTIVE);
  for (int p = 0; p < N; ++p) {
    const bool rigid = p >= nn;
    const int r = p - nn;
    for (int a = 0; a < D; ++a) {
      x[a * N + p] = rigid ? rpos[r][a] : pos[p][a];
      if (!rigid && !vel.empty()) v[a * N + p] = vel[p][a];
      if (rigid && !rvel.empty()) v[a * N + p] = rvel[r][a];
    }
    if (!rigid && !col.empty()) color[p] = (uint8_t)col[p];
    if (rigid && !rcol.empty()) color[p] = (uint8_t)rcol[r];
    if (rigid) flags[p] |= PMPM_PF_RIGID;
  }
  ParticleBuffers &b = buf[0];
  b.x.upload(x);
  b.v.upload(v);
  b.stress.upload(std::vector<float>(6 * N, 0.f));
  std::vector<float> F(N * D * D, 0.f);
  for (int p = 0; p < N; ++p)
    for (int a = 0; a < D; ++a) F[(a * D + a) * N + p] = 1.f;
  b.F.upload(F);
  b.L.upload(std::vector<float>(N * D * D, 0.f));
  b.mass.upload(std::vector<float>(N, -1.f));
  b.volume.upload(std::vector<float>(N, -1.f));
  b.volume0.upload(std::vector<float>(N, -1.f));
  b.color.upload(color);
  b.flags.upload(flags);
  std::vector<int32_t> id(N);
  std::iota(id.begin(), id.end(), 0);
  b.id.upload(id);
}

template <int D> pmpm_context *ParticlesContainer<D>::ctx() const {
  if (!ctx_) ctx_ = new_context();
  return ctx_.get();
}

template <int D> void ParticlesContainer<D>::ensure_second_set() {
  ParticleBuffers &o = buf[cur_ ^ 1];
  const ParticleBuffers &c = buf[cur_];
  const size_t N = num_particles;
  o.x.resize(N * D); o.v.resize(N * D); o.stress.resize(6 * N); o.F.resize(N * D * D); o.L.resize(N * D * D);
  o.mass.resize(N); o.volume.resize(N); o.volume0.resize(N); o.color.resize(N); o.flags.resize(N); o.id.resize(N);
  o.eps_e.resize(c.eps_e.size());
  o.stress_ref.resize(c.stress_ref.size());
  o.hist.resize(c.hist.size());
}

template <int D> void ParticlesContainer<D>::ensure_history() {
  const size_t N = num_particles;
  for (auto &b : buf) {
    if (b.x.size() == 0 && &b != &buf[cur_]) continue;
    if (b.eps_e.size() != 6 * N) { b.eps_e.resize(6 * N); hip_check(hipMemset(b.eps_e.get(), 0, 24 * N), "memset"); }
    if (b.stress_ref.size() != 6 * N) { b.stress_ref.resize(6 * N); hip_check(hipMemset(b.stress_ref.get(), 0, 24 * N), "memset"); }
  }
}

template <int D> void ParticlesContainer<D>::ensure_plastic_history() {
  ensure_history();
  const size_t N = num_particles;
  for (auto &b : buf) {
    if (b.x.size() == 0 && &b != &buf[cur_]) continue;
    if (b.hist.size() != 2 * N) { b.hist.resize(2 * N); hip_check(hipMemset(b.hist.get(), 0, 8 * N), "memset"); }
  }
}

template <int D> pmpm_particles ParticlesContainer<D>::view(int which) const {
  const ParticleBuffers &b = buf[which];
  pmpm_particles p{};
  p.n = num_particles;
  p.x = b.x.get(); p.v = b.v.get(); p.stress = b.stress.get(); p.F = b.F.get();
  p.L = b.L.size() ? b.L.get() : nullptr;
  p.mass = b.mass.get(); p.volume = b.volume.get(); p.volume0 = b.volume0.get();
  p.color = b.color.get(); p.flags = b.flags.get(); p.id = b.id.get();
  p.eps_e = b.eps_e.size() ? b.eps_e.get() : nullptr;
  p.stress_ref = b.stress_ref.size() ? b.stress_ref.get() : nullptr;
  p.hist = b.hist.size() ? b.hist.get() : nullptr;
  return p;
}

template <int D> pmpm_spatial ParticlesContainer<D>::spatial_view(SpatialBuffers &sb) const {
  const size_t N = num_particles;
  sb.hash.resize(N); sb.hash_sorted.resize(N); sb.sorted.resize(N); sb.bins.resize(N * D);
  sb.cell_start.resize(grid.num_cells_total); sb.cell_end.resize(grid.num_cells_total);
  return pmpm_spatial{sb.hash.get(), sb.hash_sorted.get(), sb.sorted.get(), sb.cell_start.get(), sb.cell_end.get(),
                      sb.bins.get()};
}

// slot order -> reference order (SoA with ncomp components)
template <int D> template <class T>
std::vector<T> ParticlesContainer<D>::to_ref(const std::vector<T> &s, int ncomp) const {
  const size_t N = num_particles;
  auto id = cur().id.download();
  std::vector<T> r(s.size());
  for (int c = 0; c < ncomp; ++c)
    for (size_t i = 0; i < N; ++i) r[c * N + id[i]] = s[c * N + i];
  return r;
}
template <int D> template <class T>
std::vector<T> ParticlesContainer<D>::from_ref(const std::vector<T> &r, int ncomp) const {
  const size_t N = num_particles;
  auto id = cur().id.download();
  std::vector<T> s(r.size());
  for (int c = 0; c < ncomp; ++c)
    for (size_t i = 0; i < N; ++i) s[c * N + i] = r[c * N + id[i]];
  return s;
}

template <int D> std::vector<VecD<D>> ParticlesContainer<D>::positions() const {
  return soa_to_vecs<D>(to_ref(cur().x.download(), D), num_particles);
}
template <int D> std::vector<VecD<D>> ParticlesContainer<D>::velocities() const {
  return soa_to_vecs<D>(to_ref(cur().v.download(), D), num_particles);
}
template <int D> std::vector<Mat3> ParticlesContainer<D>::stresses() const {
  auto s = to_ref(cur().stress.download(), 6);
  const size_t N = num_particles;
  std::vector<Mat3> m(N);
  for (size_t i = 0; i < N; ++i) {
    const float xx = s[i], yy = s[N + i], zz = s[2 * N + i], xy = s[3 * N + i], xz = s[4 * N + i], yz = s[5 * N + i];
    m[i] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
  }
  return m;
}
template <int D> static std::vector<float> mats_to_soa(const std::vector<MatD<D>> &v) {
  const size_t n = v.size();
  std::vector<float> s(n * D * D);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < D * D; ++k) s[k * n + i] = v[i][k];
  return s;
}
template <int D> static std::vector<MatD<D>> soa_to_mats(const std::vector<float> &s, size_t n) {
  std::vector<MatD<D>> v(n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < D * D; ++k) v[i][k] = s[k * n + i];
  return v;
}
template <int D> std::vector<MatD<D>> ParticlesContainer<D>::F() const {
  return soa_to_mats<D>(to_ref(cur().F.download(), D * D), num_particles);
}
template <int D> std::vector<MatD<D>> ParticlesContainer<D>::velocity_gradient() const {
  return soa_to_mats<D>(to_ref(cur().L.download(), D * D), num_particles);
}
template <int D> std::vector<float> ParticlesContainer<D>::masses() const { return to_ref(cur().mass.download(), 1); }
template <int D> std::vector<float> ParticlesContainer<D>::volumes() const { return to_ref(cur().volume.download(), 1); }
template <int D> std::vector<float> ParticlesContainer<D>::volumes_original() const {
  return to_ref(cur().volume0.download(), 1);
}
template <int D> std::vector<int> ParticlesContainer<D>::colors() const {
  auto c = to_ref(cur().color.download(), 1);
  return std::vector<int>(c.begin(), c.end());
}
template <int D> std::vector<int> ParticlesContainer<D>::is_rigid() const {
  auto f = to_ref(cur().flags.download(), 1);
  std::vector<int> r(f.size());
  for (size_t i = 0; i < f.size(); ++i) r[i] = (f[i] & PMPM_PF_RIGID) ? 1 : 0;
  return r;
}
template <int D> std::vector<int> ParticlesContainer<D>::is_active() const {
  auto f = to_ref(cur().flags.download(), 1);
  std::vector<int> r(f.size());
  for (size_t i = 0; i < f.size(); ++i) r[i] = (f[i] & PMPM_PF_ACTIVE) ? 1 : 0;
  return r;
}
template <int D> std::vector<float> ParticlesContainer<D>::plastic_history() const {
  const size_t N = num_particles;
  if (cur().hist.size() != 2 * N) return std::vector<float>(2 * N, 0.f);
  return to_ref(cur().hist.download(), 2);
}
template <int D> std::vector<Mat3> ParticlesContainer<D>::eps_e() const {
  const size_t N = num_particles;
  if (cur().eps_e.size() != 6 * N) return std::vector<Mat3>(N, Mat3{});
  auto s = to_ref(cur().eps_e.download(), 6);
  std::vector<Mat3> m(N);
  for (size_t i = 0; i < N; ++i) {
    const float xx = s[i], yy = s[N + i], zz = s[2 * N + i], xy = s[3 * N + i], xz = s[4 * N + i], yz = s[5 * N + i];
    m[i] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
  }
  return m;
}

template <int D> static void need(size_t got, size_t want, const char *what) {
  if (got != want) throw std::invalid_argument(std::string(what) + ": expected " + std::to_string(want) + " entries");
}
template <int D> void ParticlesContainer<D>::set_positions(const std::vector<VecD<D>> &v) {
  need<D>(v.size(), num_particles, "positions");
  cur().x.upload(from_ref(vecs_to_soa<D>(v), D));
}
template <int D> void ParticlesContainer<D>::set_velocities(const std::vector<VecD<D>> &v) {
  need<D>(v.size(), num_particles, "velocities");
  cur().v.upload(from_ref(vecs_to_soa<D>(v), D));
}
static std::vector<float> mat3_to_sym(const std::vector<Mat3> &m) {
  const size_t N = m.size();
  std::vector<float> s(6 * N);
  for (size_t i = 0; i < N; ++i) {
    const Mat3 &a = m[i];  // row-major; the symmetric part is stored
    s[i] = a[0]; s[N + i] = a[4]; s[2 * N + i] = a[8];
    s[3 * N + i] = 0.5f * (a[1] + a[3]);
    s[4 * N + i] = 0.5f * (a[2] + a[6]);
    s[5 * N + i] = 0.5f * (a[5] + a[7]);
  }
  return s;
}
template <int D> void ParticlesContainer<D>::set_stresses(const std::vector<Mat3> &v) {
  need<D>(v.size(), num_particles, "stresses");
  cur().stress.upload(from_ref(mat3_to_sym(v), 6));
}
template <int D> void ParticlesContainer<D>::set_eps_e(const std::vector<Mat3> &v) {
  need<D>(v.size(), num_particles, "eps_e");
  ensure_history();
  cur().eps_e.upload(from_ref(mat3_to_sym(v), 6));
}
template <int D> void ParticlesContainer<D>::set_F(const std::vector<MatD<D>> &v) {
  need<D>(v.size(), num_particles, "F");
  cur().F.upload(from_ref(mats_to_soa<D>(v), D * D));
}
template <int D> void ParticlesContainer<D>::set_velocity_gradient(const std::vector<MatD<D>> &v) {
  need<D>(v.size(), num_particles, "velocity_gradient");
  cur().L.upload(from_ref(mats_to_soa<D>(v), D * D));
}
template <int D> void ParticlesContainer<D>::set_masses(const std::vector<float> &v) {
  need<D>(v.size(), num_particles, "masses");
  cur().mass.upload(from_ref(v, 1));
}
template <int D> void ParticlesContainer<D>::set_volumes(const std::vector<float> &v) {
  need<D>(v.size(), num_particles, "volumes");
  cur().volume.upload(from_ref(v, 1));
}
template <int D> void ParticlesContainer<D>::set_volumes_original(const std::vector<float> &v) {
  need<D>(v.size(), num_particles, "volumes_original");
  cur().volume0.upload(from_ref(v, 1));
}
template <int D> void ParticlesContainer<D>::set_colors(const std::vector<int> &v) {
  need<D>(v.size(), num_particles, "colors");
  std::vector<uint8_t> c(v.begin(), v.end());
  cur().color.upload(from_ref(c, 1));
}
template <int D> void ParticlesContainer<D>::set_is_active(const std::vector<int> &v) {
  need<D>(v.size(), num_particles, "is_active");
  auto f = to_ref(cur().flags.download(), 1);
  for (size_t i = 0; i < f.size(); ++i) f[i] = (uint8_t)((f[i] & ~PMPM_PF_ACTIVE) | (v[i] ? PMPM_PF_ACTIVE : 0));
  cur().flags.upload(from_ref(f, 1));
}

// particles.cpp:209-216, 395-421
template <int D> void ParticlesContainer<D>::set_spawner(int rate, int volume) {
  spawner = Spawner{rate, volume, 0};
  set_is_active(std::vector<int>(num_particles, 0));
}
template <int D> void ParticlesContainer<D>::spawn_particles(int global_step) {
  if (spawner.rate <= 0 || global_step % spawner.rate != 0 || spawner.increment >= num_particles) return;
  auto act = is_active();
  for (int s = 0; s < spawner.volume && spawner.increment + s < num_particles; ++s) act[spawner.increment + s] = 1;
  spawner.increment += spawner.volume;
  set_is_active(act);
}

template <int D> void ParticlesContainer<D>::set_spatialpartition(const pmpm_grid &g) {
  grid = g;
  has_grid = true;
  partition();
}

// Reference-order partition (the reference sorts indices 0..N-1 of its
// unpermuted arrays): positions are gathered to reference order first.
template <int D> void ParticlesContainer<D>::partition() {
  if (!has_grid) throw std::runtime_error("partition: no grid (set_spatialpartition first)");
  const int N = num_particles;
  pmpm_context *c = ctx();
  check(pmpm_context_reserve(c, N, grid.num_cells_total), "pmpm_context_reserve");
  DevBuf<float> xr(N * D);
  check(pmpm_permute(N, D, 4, cur().x.get(), cur().id.get(), xr.get(), 0, nullptr), "pmpm_permute");
  pmpm_spatial sv = spatial_view(spatial);
  check(pmpm_partition(c, D, &grid, xr.get(), N, &sv, nullptr), "pmpm_partition");
  hip_check(hipDeviceSynchronize(), "partition");
}
template <int D> std::vector<VecI<D>> ParticlesContainer<D>::bins() const {
  auto b = spatial.bins.download();
  const size_t N = num_particles;
  std::vector<VecI<D>> r(N);
  for (size_t i = 0; i < N && b.size() == N * D; ++i)
    for (int a = 0; a < D; ++a) r[i][a] = b[a * N + i];
  return r;
}
template <int D> std::vector<uint32_t> ParticlesContainer<D>::hash_unsorted() const { return spatial.hash.download(); }
template <int D> std::vector<uint32_t> ParticlesContainer<D>::hash_sorted() const {
  return spatial.hash_sorted.download();
}
template <int D> std::vector<int32_t> ParticlesContainer<D>::sorted_index() const { return spatial.sorted.download(); }
template <int D> std::vector<int32_t> ParticlesContainer<D>::cell_start() const { return spatial.cell_start.download(); }
template <int D> std::vector<int32_t> ParticlesContainer<D>::cell_end() const { return spatial.cell_end.download(); }

// particles.cpp:345-365
template <int D> void ParticlesContainer<D>::calculate_initial_volumes() {
  if (isRestart) return;
  if (!has_grid) throw std::runtime_error("calculate_initial_volumes: no grid");
  auto v = cur().volume.download();
  for (auto &x : v) {
    if (x > 0.f) continue;
    x = (float)std::pow((double)grid.cell_size, (double)D);
    x /= (float)globals().particles_per_cell;
  }
  cur().volume.upload(v);
  cur().volume0.upload(v);
}
// particles.cpp:372-393
template <int D> void ParticlesContainer<D>::calculate_initial_masses(int mat_id, float density) {
  if (isRestart) return;
  auto c = cur().color.download();
  auto m = cur().mass.download();
  auto v = cur().volume.download();
  for (size_t i = 0; i < m.size(); ++i) {
    if ((int)c[i] != mat_id || m[i] > 0.f) continue;
    m[i] = density * v[i];
  }
  cur().mass.upload(m);
}

// =============================== materials =================================
LinearElastic::LinearElastic(float density, float E, float pois) : density(density), E(E), pois(pois) {
  bulk_modulus = E / ((float)3.0 * ((float)1.0 - (float)2.0 * pois));
  shear_modulus = E / ((float)2.0 * ((float)1 + pois));
  lame_modulus = E * pois / (((float)1.0 + pois) * ((float)1.0 - (float)2.0 * pois));
}
pmpm_material LinearElastic::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_LINEAR_ELASTIC;
  m.p[0] = lame_modulus; m.p[1] = shear_modulus; m.p[2] = bulk_modulus;
  return m;
}
NewtonFluid::NewtonFluid(float density, float viscosity, float bulk_modulus, float gamma)
    : density(density), viscosity(viscosity), bulk_modulus(bulk_modulus), gamma(gamma) {}
pmpm_material NewtonFluid::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_NEWTON_FLUID;
  m.p[0] = viscosity; m.p[1] = bulk_modulus; m.p[2] = gamma;
  return m;
}
LocalGranularRheology::LocalGranularRheology(float density, float E, float pois, float I0, float mu_s, float mu_2,
                                             float rho_c, float particle_diameter, float particle_density)
    : density(density), E(E), pois(pois), I0(I0), mu_s(mu_s), mu_2(mu_2), rho_c(rho_c),
      particle_diameter(particle_diameter), particle_density(particle_density) {
  bulk_modulus = ((float)1. / (float)3.) * (E / ((float)1. - (float)2. * pois));
  shear_modulus = ((float)1. / (float)2.) * E / ((float)1 + pois);
  lame_modulus = (pois * E) / (((float)1.0 + pois) * ((float)1 - (float)2.0 * pois));
  EPS = I0 / (float)std::sqrt(std::pow((double)particle_diameter, 2) * (double)particle_density);
}
pmpm_material LocalGranularRheology::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_LOCAL_RHEO;
  m.p[0] = shear_modulus; m.p[1] = lame_modulus; m.p[2] = bulk_modulus; m.p[3] = rho_c;
  m.p[4] = mu_s; m.p[5] = mu_2; m.p[6] = I0; m.p[7] = EPS;
  return m;
}
DP_rheo::DP_rheo(float density, float E, float pois, float M, float lam, float kap, float Vs, float R, float Pt,
                 float beta, float dilatancy_angle_deg, float svf)
    : density(density), E(E), pois(pois), M(M), lam(lam), kap(kap), Vs(Vs), R(R), Pt(Pt), beta(beta),
      solid_volume_fraction(svf) {
  bulk_modulus = E / ((float)3.0 * ((float)1.0 - (float)2.0 * pois));
  shear_modulus = E / ((float)2.0 * ((float)1 + pois));
  dilatancy_angle = dilatancy_angle_deg * (float)(3.14159265358979323846 / 180.);
  const double sp = std::sin((double)dilatancy_angle);
  eta_overline = (float)(6.0 * sp / (std::sqrt(3.0) * (3.0 + sp)));
}
pmpm_material DP_rheo::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_DRUCKER_PRAGER;
  m.flags = (do_update_history ? PMPM_MATF_UPDATE_HISTORY : 0) |
            (is_velgrad_strain_increment ? PMPM_MATF_VELGRAD_IS_INCREMENT : 0) |
            (apex_guard ? PMPM_MATF_APEX_GUARD : 0) | (stress_ref_nonzero ? PMPM_MATF_HAS_STRESS_REF : 0);
  m.p[0] = bulk_modulus; m.p[1] = shear_modulus; m.p[2] = M; m.p[3] = Pt; m.p[4] = eta_overline; m.p[5] = Vs;
  return m;
}

VonMises::VonMises(float density, float E, float pois, float ys, float H)
    : density(density), E(E), pois(pois), yield_stress(ys), H(H) {
  bulk_modulus = E / ((float)3.0 * ((float)1.0 - (float)2.0 * pois));
  shear_modulus = E / ((float)2.0 * ((float)1 + pois));
}
pmpm_material VonMises::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_VON_MISES;
  m.flags = (do_update_history ? PMPM_MATF_UPDATE_HISTORY : 0) |
            (is_velgrad_strain_increment ? PMPM_MATF_VELGRAD_IS_INCREMENT : 0);
  m.p[0] = bulk_modulus; m.p[1] = shear_modulus; m.p[2] = yield_stress; m.p[3] = H;
  return m;
}
MohrCoulomb::MohrCoulomb(float density, float E, float pois, float c, float phi_deg, float psi_deg, float H)
    : density(density), E(E), pois(pois), cohesion(c), H(H) {
  bulk_modulus = ((float)1. / (float)3.) * (E / ((float)1. - (float)2. * pois));
  shear_modulus = ((float)1. / (float)2.) * E / ((float)1 + pois);
  lame_modulus = (pois * E) / (((float)1.0 + pois) * ((float)1 - (float)2.0 * pois));
  friction_angle = phi_deg * (float)(3.14159265358979323846 / 180.);
  dilatancy_angle = psi_deg * (float)(3.14159265358979323846 / 180.);
}
pmpm_material MohrCoulomb::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_MOHR_COULOMB;
  m.flags = (do_update_history ? PMPM_MATF_UPDATE_HISTORY : 0) |
            (is_velgrad_strain_increment ? PMPM_MATF_VELGRAD_IS_INCREMENT : 0);
  m.p[0] = bulk_modulus; m.p[1] = shear_modulus; m.p[2] = cohesion; m.p[3] = friction_angle;
  m.p[4] = dilatancy_angle; m.p[5] = H;
  return m;
}
ModifiedCamClay::ModifiedCamClay(float density, float E, float pois, float M, float lam, float kap, float Vs, float R,
                                 float Pt, float beta)
    : density(density), E(E), pois(pois), M(M), lam(lam), kap(kap), Vs(Vs), R(R), Pt(Pt), beta(beta) {
  bulk_modulus = E / ((float)3.0 * ((float)1.0 - (float)2.0 * pois));
  shear_modulus = E / ((float)2.0 * ((float)1 + pois));
}
pmpm_material ModifiedCamClay::device() const {
  pmpm_material m{};
  m.kind = PMPM_MAT_MODIFIED_CAM_CLAY;
  // the reference stress is the stress at initialisation (modifiedcamclay.cpp:77-78)
  m.flags = (do_update_history ? PMPM_MATF_UPDATE_HISTORY : 0) |
            (is_velgrad_strain_increment ? PMPM_MATF_VELGRAD_IS_INCREMENT : 0) | PMPM_MATF_HAS_STRESS_REF;
  m.p[0] = bulk_modulus; m.p[1] = shear_modulus; m.p[2] = M; m.p[3] = lam; m.p[4] = kap; m.p[5] = Pt;
  m.p[6] = Vs;
  return m;
}

// ========================= boundary conditions =============================
template <int D> pmpm_bc Gravity<D>::device(int global_step) {
  const float ramp_factor = ((float)global_step) / static_cast<float>(ramp_step);
  if (is_ramp && global_step < ramp_step)
    for (int a = 0; a < D; ++a) gravity[a] = gravity_end[a] * ramp_factor;
  pmpm_bc b{};
  b.kind = PMPM_BC_GRAVITY;
  for (int a = 0; a < D; ++a) b.f[a] = gravity[a];
  return b;
}
template <int D> pmpm_bc NodeDomain<D>::device() const {
  pmpm_bc b{};
  b.kind = PMPM_BC_NODE_DOMAIN;
  for (int a = 0; a < 3; ++a) {
    b.i[a] = a < D ? face0_mode[a] : -1;
    b.i[3 + a] = a < D ? face1_mode[a] : -1;
  }
  return b;
}
template <int D> RigidBodyLevelSet<D>::RigidBodyLevelSet() {
  if (D == 1) throw std::invalid_argument("Rigid body level set not supported for 1D");
  for (int a = 0; a < D; ++a) rotation_matrix[a * D + a] = 1.f;  // STATIC: identity (reference: uninitialised)
}
template <int D> void RigidBodyLevelSet<D>::set_mode_loop_rotate(const VecD<D> &eps, const VecD<D> &com) {
  mode = 1;
  euler_angles_per_second = eps;
  COM = com;
}
// rigidbodylevelset.cpp:152-217; in 3D the reference's Rz stays zero, so the
// rotation is the zero matrix there.
template <int D> void RigidBodyLevelSet<D>::update_animation(float dt) {
  if (mode != 1) return;
  VecD<D> prev = euler_angles, de{};
  for (int a = 0; a < D; ++a) euler_angles[a] = euler_angles[a] + euler_angles_per_second[a] * dt;
  for (int a = 0; a < D; ++a) de[a] = prev[a] - euler_angles[a];
  for (int a = 0; a < D; ++a) angular_velocities[a] = de[a] / dt;
  rotation_matrix.fill(0.f);
  if (D == 2) {
    rotation_matrix[0] = (float)std::cos((double)de[0]);
    rotation_matrix[1] = (float)-std::sin((double)de[0]);
    rotation_matrix[2] = (float)std::sin((double)de[0]);
    rotation_matrix[3] = (float)std::cos((double)de[0]);
  }
}
template <int D>
RigidBodySDF<D>::RigidBodySDF(const std::array<float, 3> &h, const std::array<float, 3> &c,
                              const std::array<float, 3> &v, const std::array<float, 9> &r, float band)
    : half_extents(h), center(c), velocity(v), rotation(r), band(band) {}
template <int D> pmpm_bc RigidBodySDF<D>::device() const {
  pmpm_bc b{};
  b.kind = PMPM_BC_RIGID_SDF;
  for (int a = 0; a < 3; ++a) { b.f[a] = half_extents[a]; b.f[3 + a] = center[a]; b.f[6 + a] = velocity[a]; }
  for (int k = 0; k < 9; ++k) b.f[9 + k] = rotation[k];
  b.f[18] = band;
  return b;
}

template <int D>
BodyForce<D>::BodyForce(const std::string &mode, const std::vector<VecD<D>> &v, const std::vector<bool> &m)
    : values(v), mask(m) {
  if (mode == "forces") mode_id = 0;
  else if (mode == "moments") mode_id = 1;
  else if (mode == "fixed") mode_id = 2;
  else throw std::invalid_argument("BodyForce mode must be \"forces\", \"moments\" or \"fixed\"");
  if (values.size() != mask.size()) throw std::invalid_argument("BodyForce: values and mask differ in length");
}
template <int D> pmpm_bc BodyForce<D>::device(int num_nodes) {
  pmpm_bc b{};
  if (!isActive) return b;  // bodyforce.cpp:62-65
  if ((int)values.size() != num_nodes)
    throw std::invalid_argument("BodyForce: values and mask need one entry per grid node");
  if (!values_dev) {
    std::vector<float> vh((size_t)D * num_nodes);
    std::vector<uint8_t> mh((size_t)num_nodes);
    for (int i = 0; i < num_nodes; ++i) {
      for (int a = 0; a < D; ++a) vh[(size_t)a * num_nodes + i] = values[i][a];
      mh[i] = mask[i] ? 1 : 0;
    }
    values_dev = std::make_shared<DevBuf<float>>();
    mask_dev = std::make_shared<DevBuf<uint8_t>>();
    values_dev->upload(vh);
    mask_dev->upload(mh);
  }
  b.kind = PMPM_BC_BODY_FORCE;
  b.i[0] = mode_id;
  b.values = values_dev->get();
  b.mask = mask_dev->get();
  return b;
}
template <int D>
PlanarDomain<D>::PlanarDomain(const VecD<D> &f0, const VecD<D> &f1) {
  for (int a = 0; a < D; ++a) {
    face0_friction[a] = f0[a] * (float)(3.14159265358979323846 / 180.);
    face1_friction[a] = f1[a] * (float)(3.14159265358979323846 / 180.);
  }
}
template <int D> pmpm_bc PlanarDomain<D>::device(const VecD<D> &domain_end) const {
  pmpm_bc b{};
  if (!isActive) return b;
  b.kind = PMPM_BC_PLANAR_DOMAIN;
  for (int a = 0; a < D; ++a) {
    b.f[a] = face0_friction[a];
    b.f[3 + a] = face1_friction[a];
    b.f[6 + a] = domain_end[a];
  }
  return b;
}

// ================================ Solver ===================================
template <int D>
Solver<D>::Solver(const ParticlesContainer<D> &p, const NodesContainer<D> &n, const std::vector<MaterialType> &m,
                  const std::vector<BoundaryConditionType<D>> &b, float alpha, bool musl)
    : particles(p), nodes(n), materials(m), boundaryconditions(b), alpha(alpha), musl(musl) {
  if (materials.size() > PMPM_MAX_MATERIALS) throw std::invalid_argument("at most 8 materials");
  if (boundaryconditions.size() > PMPM_MAX_BCS) throw std::invalid_argument("at most 8 boundary conditions");
  particles.set_spatialpartition(nodes.grid);  // solver.cpp:77
  particles.calculate_initial_volumes();
  for (size_t k = 0; k < materials.size(); ++k)
    std::visit([&](auto &mat) { particles.calculate_initial_masses((int)k, mat.density); }, materials[k]);
  initialize_materials();
  particles.ensure_second_set();
}
template <int D> Solver<D>::~Solver() { set_global_step(0); }  // solver.cpp:199

// Drucker-Prager history (dp_rheo.cpp:75-118). The reference never calls it
// from its solver (its DP cannot be used there); here every DP material is
// initialised at construction so it can be.
template <int D> void Solver<D>::initialize_materials() {
  bool any_dp = false, any_ext = false;
  for (auto &m : materials) {
    if (std::holds_alternative<DP_rheo>(m)) any_dp = true;
    if (std::holds_alternative<VonMises>(m) || std::holds_alternative<MohrCoulomb>(m) ||
        std::holds_alternative<ModifiedCamClay>(m))
      any_ext = true;
    if (D != 3 && std::holds_alternative<VonMises>(m))
      throw std::invalid_argument("VonMises material only implemented for 3D");  // vonmises.cpp:46-49
  }
  if (any_ext) initialize_plastic_history();
  if (!any_dp) return;
  particles.ensure_history();
  ParticleBuffers &b = particles.cur();
  const size_t N = particles.num_particles;
  hip_check(hipMemset(b.eps_e.get(), 0, 24 * N), "memset");
  b.stress_ref = b.stress;
  auto s = b.stress.download();
  bool nonzero = false;
  for (float x : s) nonzero = nonzero || x != 0.f;
  auto col = b.color.download();
  for (size_t k = 0; k < materials.size(); ++k) {
    if (auto *dp = std::get_if<DP_rheo>(&materials[k])) {
      dp->stress_ref_nonzero = nonzero;
      dp->pc.assign(N, 0.f);
      for (size_t i = 0; i < N; ++i) dp->pc[i] = (-(s[i] + s[N + i] + s[2 * N + i]) / 3.f) * dp->R;
    }
  }
}

// History of the post-G2P materials: zero elastic strain and accumulated
// plastic strain (vonmises.cpp:55-59, mohrcoulomb.cpp:55-59); for the
// modified Cam-Clay the reference stress is the stress now and the
// preconsolidation pressure R p0 (modifiedcamclay.cpp:68-104), per particle of
// that material's colour.
template <int D> void Solver<D>::initialize_plastic_history() {
  particles.ensure_plastic_history();
  ParticleBuffers &b = particles.cur();
  const size_t N = particles.num_particles;
  hip_check(hipMemset(b.eps_e.get(), 0, 24 * N), "memset");
  b.stress_ref = b.stress;
  auto s = b.stress.download();
  auto col = b.color.download();
  std::vector<float> h(2 * N, 0.f);
  for (size_t i = 0; i < N; ++i) {
    const int c = col[i];
    if (c >= (int)materials.size()) continue;
    if (auto *mcc = std::get_if<ModifiedCamClay>(&materials[c])) {
      const float p0 = -(s[i] + s[N + i] + s[2 * N + i]) / 3.f;
      h[N + i] = p0 * mcc->R;
    }
  }
  b.hist.upload(h);
}

template <int D> bool Solver<D>::has_levelset() const {
  for (auto &b : boundaryconditions)
    if (std::holds_alternative<RigidBodyLevelSet<D>>(b)) return true;
  return false;
}

template <int D> pmpm_params Solver<D>::params() const {
  pmpm_params p{};
  p.dim = D;
  p.shape_fn = globals().shape_function;
  p.dt = globals().dt;
  p.alpha = alpha;
  p.small_mass_cutoff = nodes.small_mass_cutoff;
  p.node_types_built = nodes.node_types_built ? 1 : 0;
  p.wrap_quirk = wrap_quirk ? 1 : 0;
  p.num_materials = (int)materials.size();
  for (size_t k = 0; k < materials.size(); ++k)
    p.materials[k] = std::visit([](const auto &m) { return m.device(); }, materials[k]);
  p.num_bcs = 0;
  return p;
}

template <int D> void Solver<D>::synchronize() const { hip_check(hipStreamSynchronize(stream), "sync"); }

// Node BCs for this step in list order; gravity advances its ramp here, once
// per step, like the reference's apply_on_nodes_moments (gravity.cpp:65-76).
template <int D>
static void fill_bcs(pmpm_params &p, std::vector<BoundaryConditionType<D>> &bcs, int step,
                     const NodesContainer<D> &nodes) {
  p.num_bcs = (int)bcs.size();
  for (size_t k = 0; k < bcs.size(); ++k) {
    pmpm_bc b{};
    if (auto *g = std::get_if<Gravity<D>>(&bcs[k])) b = g->device(step);
    else if (auto *nd = std::get_if<NodeDomain<D>>(&bcs[k])) b = nd->device();
    else if (auto *sd = std::get_if<RigidBodySDF<D>>(&bcs[k])) b = sd->device();
    else if (auto *bf = std::get_if<BodyForce<D>>(&bcs[k])) b = bf->device(nodes.grid.num_cells_total);
    else if (auto *pd = std::get_if<PlanarDomain<D>>(&bcs[k])) b = pd->device(nodes.end);
    else b.kind = PMPM_BC_NONE;  // rigid-particle level set: own pass
    p.bcs[k] = b;
  }
}
template <int D> static void advance_sdf(std::vector<BoundaryConditionType<D>> &bcs, float dt) {
  for (auto &b : bcs)
    if (auto *sd = std::get_if<RigidBodySDF<D>>(&b))
      for (int a = 0; a < 3; ++a) sd->center[a] += sd->velocity[a] * dt;
}

// usl.cpp:69-115 (and the MUSL double mapping)
template <int D> void Solver<D>::solve() {
  particles.spawn_particles(globals().step);
  pmpm_params prm = params();
  fill_bcs<D>(prm, boundaryconditions, globals().step, nodes);
  last_prm_ = prm;
  if (has_levelset()) step_composed();
  else step_fused();
  advance_sdf<D>(boundaryconditions, prm.dt);
}

template <int D> void Solver<D>::step_fused() {
  const int N = particles.num_particles;
  pmpm_context *ctx = particles.ctx();
  check(pmpm_context_reserve(ctx, N, nodes.grid.num_cells_total), "pmpm_context_reserve");
  particles.ensure_second_set();
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_particles b = particles.view(particles.cur_index() ^ 1);
  pmpm_spatial sp = particles.spatial_view(particles.work);
  pmpm_nodes nv = nodes.view();
  check(pmpm_step(ctx, &last_prm_, &nodes.grid, musl ? 1 : 0, &a, &b, &sp, &nv, stream), "pmpm_step");
  particles.swap_buffers();
  // no wait here: the steps queue up on the stream, and every host read
  // (getters, output) copies on the same stream after them
}

// General sequence for lists with a rigid-particle level set: partition, P2G,
// node BCs in list order with the level set as its own pass, G2P.
template <int D> void Solver<D>::step_composed() {
  const int N = particles.num_particles;
  pmpm_context *ctx = particles.ctx();
  check(pmpm_context_reserve(ctx, N, nodes.grid.num_cells_total), "pmpm_context_reserve");
  particles.ensure_second_set();
  const pmpm_params &prm = last_prm_;
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_particles b = particles.view(particles.cur_index() ^ 1);
  pmpm_spatial sp = particles.spatial_view(particles.work);
  pmpm_nodes nv = nodes.view();
  check(pmpm_partition(ctx, D, &nodes.grid, a.x, N, &sp, stream), "pmpm_partition");
  check(pmpm_p2g(ctx, &prm, &nodes.grid, &a, &sp, &nv, stream), "pmpm_p2g");
  rigid_dx_.resize((size_t)N * D);  // kept between steps: no allocation per step
  int begin = 0;
  for (int k = 0; k < prm.num_bcs; ++k) {
    auto *ls = std::get_if<RigidBodyLevelSet<D>>(&boundaryconditions[k]);
    if (!ls) continue;
    check(pmpm_grid_bc(&prm, &nodes.grid, &nv, begin, k, 0, stream), "pmpm_grid_bc");
    begin = k + 1;
    ls->update_animation(prm.dt);
    // the reference reads dpsi from its stored table in reference order
    const int nsn = (prm.shape_fn == PMPM_SF_LINEAR) ? (1 << D) : (D == 1 ? 4 : (D == 2 ? 16 : 64));
    x_orig_.resize((size_t)N * D);
    psi_tmp_.resize((size_t)N * nsn);
    dpsi_orig_.resize((size_t)N * nsn * D);
    check(pmpm_permute(N, D, 4, a.x, a.id, x_orig_.get(), 0, stream), "pmpm_permute");
    check(pmpm_shape_function(D, prm.shape_fn, prm.node_types_built, &nodes.grid, x_orig_.get(), N, psi_tmp_.get(),
                              dpsi_orig_.get(), stream),
          "pmpm_shape_function");
    check(pmpm_rigid_levelset(&prm, &nodes.grid, &a, &sp, &nv, ls->rotation_matrix.data(), ls->COM.data(),
                              dpsi_orig_.get(), rigid_dx_.get(), stream),
          "pmpm_rigid_levelset");
  }
  check(pmpm_grid_bc(&prm, &nodes.grid, &nv, begin, prm.num_bcs, 0, stream), "pmpm_grid_bc");
  if (!musl) {
    check(pmpm_g2p(&prm, &nodes.grid, &a, &b, &sp, &nv, stream), "pmpm_g2p");
  } else {
    check(pmpm_musl_map(ctx, &prm, &nodes.grid, &a, &b, &sp, &nv, stream), "pmpm_musl_map");
    check(pmpm_grid_bc(&prm, &nodes.grid, &nv, 0, prm.num_bcs, 1, stream), "pmpm_grid_bc");
    check(pmpm_musl_g2p(&prm, &nodes.grid, &a, &b, &sp, &nv, stream), "pmpm_musl_g2p");
  }
  check(pmpm_rigid_advance(D, N, b.x, rigid_dx_.get(), sp.sorted, stream), "pmpm_rigid_advance");
  particles.swap_buffers();
}

template <int D> void Solver<D>::solve_nsteps(int n) {  // solver.cpp:125-135
  for (int s = 0; s < n; ++s) {
    solve();
    globals().step += 1;
  }
  output();
}
template <int D> void Solver<D>::run(int total_steps, int output_frequency) {  // solver.cpp:140-174
  output();
  solve();
  globals().step += 1;
  for (int s = 1; s < total_steps + 1; ++s) {
    solve();
    if (output_frequency > 0 && s % output_frequency == 0) output();
    globals().step += 1;
  }
}

template <int D> void Solver<D>::stress_update() {  // solver.cpp:110-121
  pmpm_params prm = params();
  pmpm_particles a = particles.view(particles.cur_index());
  for (int k = 0; k < prm.num_materials; ++k) check(pmpm_stress_update(&prm, k, &a, stream), "pmpm_stress_update");
  synchronize();
}

// Stepwise API (usl.cpp:118-199): operate in place on the current positions.
template <int D> void Solver<D>::partition() {
  particles.partition();
  pmpm_context *ctx = particles.ctx();
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_spatial sp = particles.spatial_view(particles.work);
  check(pmpm_partition(ctx, D, &nodes.grid, a.x, particles.num_particles, &sp, stream), "pmpm_partition");
  synchronize();
}
template <int D> void Solver<D>::P2G() {
  partition();
  pmpm_context *ctx = particles.ctx();
  pmpm_params prm = params();
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_spatial sp = particles.spatial_view(particles.work);
  pmpm_nodes nv = nodes.view();
  check(pmpm_p2g(ctx, &prm, &nodes.grid, &a, &sp, &nv, stream), "pmpm_p2g");
  synchronize();
}
template <int D> void Solver<D>::apply_boundary_conditions(bool remap_pass) {
  pmpm_params prm = params();
  fill_bcs<D>(prm, boundaryconditions, globals().step, nodes);
  pmpm_nodes nv = nodes.view();
  check(pmpm_grid_bc(&prm, &nodes.grid, &nv, 0, prm.num_bcs, remap_pass ? 1 : 0, stream), "pmpm_grid_bc");
  synchronize();
}
template <int D> void Solver<D>::G2P() {
  if (musl) { G2P_second(); return; }
  pmpm_params prm = params();
  prm.num_materials = 0;  // the reference's G2P has no stress update
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_spatial sp = particles.spatial_view(particles.work);
  pmpm_nodes nv = nodes.view();
  check(pmpm_grid_bc(&prm, &nodes.grid, &nv, 0, 0, 0, stream), "pmpm_grid_bc");  // node velocities
  check(pmpm_g2p(&prm, &nodes.grid, &a, &a, &sp, &nv, stream), "pmpm_g2p");
  synchronize();
}
// MUSL stepwise: first map (velocities, positions) + re-map in one pass; the
// positions of the step start are kept for the second map's weights.
template <int D> void Solver<D>::G2P_double_mapping() {
  const int N = particles.num_particles;
  pmpm_context *ctx = particles.ctx();
  pmpm_params prm = params();
  particles.ensure_second_set();
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_particles b = particles.view(particles.cur_index() ^ 1);
  pmpm_spatial sp = particles.spatial_view(particles.work);
  pmpm_nodes nv = nodes.view();
  check(pmpm_grid_bc(&prm, &nodes.grid, &nv, 0, 0, 0, stream), "pmpm_grid_bc");
  check(pmpm_musl_map(ctx, &prm, &nodes.grid, &a, &b, &sp, &nv, stream), "pmpm_musl_map");
  x_orig_.resize((size_t)N * D);
  hip_check(hipMemcpyAsync(x_orig_.get(), a.x, (size_t)N * D * 4, hipMemcpyDeviceToDevice, stream), "copy");
  // written in cell order j: scatter back to the slots
  check(pmpm_permute(N, D, 4, b.x, sp.sorted, a.x, 0, stream), "pmpm_permute");
  check(pmpm_permute(N, D, 4, b.v, sp.sorted, a.v, 0, stream), "pmpm_permute");
  synchronize();
}
template <int D> void Solver<D>::P2G_double_mapping() {
  pmpm_params prm = params();
  pmpm_nodes nv = nodes.view();
  check(pmpm_grid_bc(&prm, &nodes.grid, &nv, 0, 0, 1, stream), "pmpm_grid_bc");
  synchronize();
}
template <int D> void Solver<D>::G2P_second() {
  pmpm_params prm = params();
  prm.num_materials = 0;
  pmpm_particles a = particles.view(particles.cur_index());
  pmpm_particles w = a;
  if (x_orig_.size() == (size_t)particles.num_particles * D) w.x = x_orig_.get();
  pmpm_spatial sp = particles.spatial_view(particles.work);
  pmpm_nodes nv = nodes.view();
  check(pmpm_musl_g2p(&prm, &nodes.grid, &w, &a, &sp, &nv, stream), "pmpm_musl_g2p");
  synchronize();
}

template class NodesContainer<1>;
template class NodesContainer<2>;
template class NodesContainer<3>;
template class ParticlesContainer<1>;
template class ParticlesContainer<2>;
template class ParticlesContainer<3>;
template struct Gravity<1>;
template struct Gravity<2>;
template struct Gravity<3>;
template struct BodyForce<1>;
template struct BodyForce<2>;
template struct BodyForce<3>;
template struct PlanarDomain<1>;
template struct PlanarDomain<2>;
template struct PlanarDomain<3>;
template struct NodeDomain<1>;
template struct NodeDomain<2>;
template struct NodeDomain<3>;
template struct RigidBodyLevelSet<2>;
template struct RigidBodyLevelSet<3>;
template struct RigidBodySDF<1>;
template struct RigidBodySDF<2>;
template struct RigidBodySDF<3>;
template class Solver<1>;
template class Solver<2>;
template class Solver<3>;

}  // namespace pmpm_host
