// oracle_bind.cpp -- Python binding of the CPU oracle (oracle.h).
// TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg, never by the product package.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <map>
#include <string>

#include "oracle.h"

namespace py = pybind11;
using namespace oracle;

namespace {

template <class R> using Arr = py::array_t<R, py::array::c_style | py::array::forcecast>;

template <class R, int D> std::vector<Vec<R, D>> to_vecs(const py::object &o) {
  std::vector<Vec<R, D>> out;
  if (o.is_none()) return out;
  Arr<double> a(o);
  if (a.size() == 0) return out;
  if (a.ndim() == 1 && D == 1) {
    out.resize(a.shape(0));
    for (ssize_t i = 0; i < a.shape(0); ++i) out[i][0] = (R)a.at(i);
    return out;
  }
  if (a.ndim() != 2 || a.shape(1) != D) throw std::invalid_argument("expected (N, D) array");
  out.resize(a.shape(0));
  for (ssize_t i = 0; i < a.shape(0); ++i)
    for (int d = 0; d < D; ++d) out[i][d] = (R)a.at(i, d);
  return out;
}
template <class R, int D> py::array_t<double> from_vecs(const std::vector<Vec<R, D>> &v) {
  py::array_t<double> a({(ssize_t)v.size(), (ssize_t)D});
  auto m = a.template mutable_unchecked<2>();
  for (size_t i = 0; i < v.size(); ++i)
    for (int d = 0; d < D; ++d) m(i, d) = (double)v[i][d];
  return a;
}
template <class R, int N> std::vector<Mat<R, N>> to_mats(const py::object &o) {
  Arr<double> a(o);
  if (a.ndim() != 3 || a.shape(1) != N || a.shape(2) != N) throw std::invalid_argument("expected (N, n, n) array");
  std::vector<Mat<R, N>> out(a.shape(0));
  for (ssize_t p = 0; p < a.shape(0); ++p)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) out[p](i, j) = (R)a.at(p, i, j);
  return out;
}
template <class R, int N> py::array_t<double> from_mats(const std::vector<Mat<R, N>> &v) {
  py::array_t<double> a({(ssize_t)v.size(), (ssize_t)N, (ssize_t)N});
  auto m = a.template mutable_unchecked<3>();
  for (size_t p = 0; p < v.size(); ++p)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) m(p, i, j) = (double)v[p](i, j);
  return a;
}
template <class R> std::vector<R> to_scalars(const py::object &o) {
  Arr<double> a(o);
  std::vector<R> out(a.size());
  for (ssize_t i = 0; i < a.size(); ++i) out[i] = (R)a.data()[i];
  return out;
}
template <class T> py::array_t<double> from_scalars(const std::vector<T> &v) {
  py::array_t<double> a((ssize_t)v.size());
  for (size_t i = 0; i < v.size(); ++i) a.mutable_data()[i] = (double)v[i];
  return a;
}
template <int D> py::array_t<int64_t> from_veci(const std::vector<VecI<D>> &v) {
  py::array_t<int64_t> a({(ssize_t)v.size(), (ssize_t)D});
  auto m = a.template mutable_unchecked<2>();
  for (size_t i = 0; i < v.size(); ++i)
    for (int d = 0; d < D; ++d) m(i, d) = v[i][d];
  return a;
}
template <class T> py::array_t<int64_t> from_ints(const std::vector<T> &v) {
  py::array_t<int64_t> a((ssize_t)v.size());
  for (size_t i = 0; i < v.size(); ++i) a.mutable_data()[i] = (int64_t)v[i];
  return a;
}

double getp(const py::dict &d, const char *k, double def) {
  return d.contains(k) ? d[k].cast<double>() : def;
}

template <class R, int D> void bind_sim(py::module &m, const std::string &name) {
  using S = Sim<R, D>;
  using V = Vec<R, D>;
  py::class_<S>(m, name.c_str())
      .def(py::init([](py::object origin, py::object end, double h, const std::string &sf, double dt, int ppc) {
             auto o = to_vecs<R, D>(py::array_t<double>(Arr<double>(origin).reshape({1, D})));
             auto e = to_vecs<R, D>(py::array_t<double>(Arr<double>(end).reshape({1, D})));
             return new S(o[0], e[0], (R)h, shape_fn_from_string(sf), (R)dt, ppc);
           }),
           py::arg("origin"), py::arg("end"), py::arg("cell_size"), py::arg("shape_function"),
           py::arg("dt"), py::arg("particles_per_cell"))
      .def("set_particles",
           [](S &s, py::object pos, py::object vel, py::object col, py::object rpos, py::object rvel,
              py::object rcol) {
             std::vector<int> c, rc;
             if (!col.is_none()) for (auto x : to_scalars<double>(col)) c.push_back((int)x);
             if (!rcol.is_none()) for (auto x : to_scalars<double>(rcol)) rc.push_back((int)x);
             s.set_particles(to_vecs<R, D>(pos), to_vecs<R, D>(vel), c, to_vecs<R, D>(rpos),
                             to_vecs<R, D>(rvel), rc);
           },
           py::arg("positions"), py::arg("velocities") = py::none(), py::arg("colors") = py::none(),
           py::arg("rigid_positions") = py::none(), py::arg("rigid_velocities") = py::none(),
           py::arg("rigid_colors") = py::none())
      .def("add_material",
           [](S &s, const std::string &kind, const py::dict &p) {
             Material<R> mt;
             if (kind == "LinearElastic")
               mt = Material<R>::linear_elastic((R)getp(p, "density", 0), (R)getp(p, "E", 0), (R)getp(p, "pois", 0));
             else if (kind == "NewtonFluid")
               mt = Material<R>::newton_fluid((R)getp(p, "density", 0), (R)getp(p, "viscosity", 0),
                                              (R)getp(p, "bulk_modulus", 0), (R)getp(p, "gamma", 7));
             else if (kind == "LocalGranularRheology")
               mt = Material<R>::local_rheo((R)getp(p, "density", 0), (R)getp(p, "E", 0), (R)getp(p, "pois", 0),
                                            (R)getp(p, "I0", 0), (R)getp(p, "mu_s", 0), (R)getp(p, "mu_2", 0),
                                            (R)getp(p, "rho_c", 0), (R)getp(p, "particle_diameter", 0),
                                            (R)getp(p, "particle_density", 0));
             else if (kind == "DP_rheo") {
               mt = Material<R>::dp_rheo((R)getp(p, "density", 0), (R)getp(p, "E", 0), (R)getp(p, "pois", 0),
                                         (R)getp(p, "M", 0), (R)getp(p, "lam", 0), (R)getp(p, "kap", 0),
                                         (R)getp(p, "Vs", 0), (R)getp(p, "R", 0), (R)getp(p, "Pt", 0),
                                         (R)getp(p, "beta", 0), (R)getp(p, "dilatancy_angle", 0),
                                         (R)getp(p, "solid_volume_fraction", 1));
               mt.apex_guard = getp(p, "apex_guard", 1) != 0;
               mt.do_update_history = getp(p, "do_update_history", 1) != 0;
               mt.is_velgrad_strain_increment = getp(p, "is_velgrad_strain_increment", 0) != 0;
             } else if (kind == "VonMises") {
               mt = Material<R>::von_mises((R)getp(p, "density", 0), (R)getp(p, "E", 0), (R)getp(p, "pois", 0),
                                           (R)getp(p, "yield_stress", 0), (R)getp(p, "H", 0));
             } else if (kind == "MohrCoulomb") {
               mt = Material<R>::mohr_coulomb((R)getp(p, "density", 0), (R)getp(p, "E", 0), (R)getp(p, "pois", 0),
                                              (R)getp(p, "cohesion", 0), (R)getp(p, "friction_angle", 0),
                                              (R)getp(p, "dilatancy_angle", 0), (R)getp(p, "H", 0));
             } else if (kind == "ModifiedCamClay") {
               mt = Material<R>::modified_cam_clay((R)getp(p, "density", 0), (R)getp(p, "E", 0),
                                                   (R)getp(p, "pois", 0), (R)getp(p, "M", 0), (R)getp(p, "lam", 0),
                                                   (R)getp(p, "kap", 0), (R)getp(p, "Vs", 0), (R)getp(p, "R", 0),
                                                   (R)getp(p, "Pt", 0), (R)getp(p, "beta", 0));
             } else
               throw std::invalid_argument("unknown material " + kind);
             if (kind != "LinearElastic" && kind != "NewtonFluid" && kind != "LocalGranularRheology") {
               mt.do_update_history = getp(p, "do_update_history", 1) != 0;
               mt.is_velgrad_strain_increment = getp(p, "is_velgrad_strain_increment", 0) != 0;
             }
             s.materials.push_back(mt);
           })
      .def("add_bc",
           [](S &s, const std::string &kind, const py::dict &p) {
             BoundaryCondition<R, D> bc;
             auto vec = [&](const char *k) {
               V v = V::zero();
               if (p.contains(k)) {
                 auto a = to_scalars<double>(p[k]);
                 for (int i = 0; i < D && i < (int)a.size(); ++i) v[i] = (R)a[i];
               }
               return v;
             };
             auto veci = [&](const char *k) {
               VecI<D> v{};
               if (p.contains(k)) {
                 auto a = to_scalars<double>(p[k]);
                 for (int i = 0; i < D && i < (int)a.size(); ++i) v[i] = (int)a[i];
               }
               return v;
             };
             if (kind == "Gravity") {
               bc.kind = BC_GRAVITY;
               bc.gravity = vec("gravity");
               bc.is_ramp = getp(p, "is_ramp", 0) != 0;
               bc.ramp_step = (int)getp(p, "ramp_step", 0);
               bc.gravity_end = vec("gravity_end");
             } else if (kind == "NodeDomain") {
               bc.kind = BC_NODE_DOMAIN;
               bc.face0_mode = veci("face0_mode");
               bc.face1_mode = veci("face1_mode");
             } else if (kind == "RigidBodyLevelSet") {
               bc.kind = BC_RIGID_LEVELSET;
               bc.rigid_mode = (int)getp(p, "mode", 0);
               bc.euler_angles_per_second = vec("euler_angles_per_second");
               bc.COM = vec("COM");
             } else if (kind == "RigidBodySDF") {
               bc.kind = BC_RIGID_SDF;
               auto h = to_scalars<double>(p["half_extents"]);
               auto c = to_scalars<double>(p["center"]);
               auto v = to_scalars<double>(p["velocity"]);
               for (int i = 0; i < 3; ++i) {
                 bc.sdf_half[i] = i < (int)h.size() ? (R)h[i] : R(0);
                 bc.sdf_center[i] = i < (int)c.size() ? (R)c[i] : R(0);
                 bc.sdf_velocity[i] = i < (int)v.size() ? (R)v[i] : R(0);
               }
               if (p.contains("rotation")) {  // row-major 3x3
                 auto r = to_scalars<double>(p["rotation"]);
                 for (int i = 0; i < 3; ++i)
                   for (int j = 0; j < 3; ++j) bc.sdf_rot(i, j) = (R)r[i * 3 + j];
               }
               bc.sdf_band = (R)getp(p, "band", 0);
             } else if (kind == "BodyForce") {
               bc.kind = BC_BODY_FORCE;
               const std::string mode = p.contains("mode") ? py::cast<std::string>(p["mode"]) : "forces";
               bc.body_mode = mode == "fixed" ? 2 : (mode == "moments" ? 1 : 0);
               bc.body_values = to_vecs<R, D>(p["values"]);
               for (auto x : to_scalars<double>(p["mask"])) bc.body_mask.push_back(x != 0 ? 1 : 0);
               if ((int)bc.body_values.size() != s.grid.num_cells_total ||
                   (int)bc.body_mask.size() != s.grid.num_cells_total)
                 throw std::invalid_argument("BodyForce: one value and mask entry per node");
             } else if (kind == "PlanarDomain") {
               bc.kind = BC_PLANAR_DOMAIN;
               const V f0 = vec("face0_friction"), f1 = vec("face1_friction");
               for (int i = 0; i < D; ++i) {  // planardomain.cpp:27-33
                 bc.face0_friction[i] = f0[i] * (R)(3.14159265358979323846 / 180.);
                 bc.face1_friction[i] = f1[i] * (R)(3.14159265358979323846 / 180.);
               }
             } else
               throw std::invalid_argument("unknown bc " + kind);
             s.bcs.push_back(bc);
           })
      .def("set",
           [](S &s, const std::string &k, py::object a) {
             if (k == "positions") s.positions = to_vecs<R, D>(a);
             else if (k == "velocities") s.velocities = to_vecs<R, D>(a);
             else if (k == "masses") s.masses = to_scalars<R>(a);
             else if (k == "volumes") s.volumes = to_scalars<R>(a);
             else if (k == "volumes_original") s.volumes_original = to_scalars<R>(a);
             else if (k == "stresses") s.stresses = to_mats<R, 3>(a);
             else if (k == "F") s.F = to_mats<R, D>(a);
             else if (k == "velocity_gradient") s.velocity_gradient = to_mats<R, D>(a);
             else if (k == "colors") { s.colors.clear(); for (auto x : to_scalars<double>(a)) s.colors.push_back((int)x); }
             else if (k == "is_active") { s.is_active.clear(); for (auto x : to_scalars<double>(a)) s.is_active.push_back(x != 0); }
             else if (k == "node_masses") s.node_masses = to_scalars<R>(a);
             else if (k == "node_moments") s.node_moments = to_vecs<R, D>(a);
             else if (k == "node_moments_nt") s.node_moments_nt = to_vecs<R, D>(a);
             else throw std::invalid_argument("unknown field " + k);
           })
      .def("get",
           [](S &s, const std::string &k) -> py::object {
             if (k == "positions") return from_vecs<R, D>(s.positions);
             if (k == "velocities") return from_vecs<R, D>(s.velocities);
             if (k == "forces_external") return from_vecs<R, D>(s.forces_external);
             if (k == "masses") return from_scalars(s.masses);
             if (k == "volumes") return from_scalars(s.volumes);
             if (k == "volumes_original") return from_scalars(s.volumes_original);
             if (k == "stresses") return from_mats<R, 3>(s.stresses);
             if (k == "F") return from_mats<R, D>(s.F);
             if (k == "velocity_gradient") return from_mats<R, D>(s.velocity_gradient);
             if (k == "colors") return from_ints(s.colors);
             if (k == "bins") return from_veci<D>(s.bins);
             if (k == "hash_unsorted") return from_ints(s.hash_unsorted);
             if (k == "hash_sorted") return from_ints(s.hash_sorted);
             if (k == "sorted_index") return from_ints(s.sorted_index);
             if (k == "cell_start") return from_ints(s.cell_start);
             if (k == "cell_end") return from_ints(s.cell_end);
             if (k == "psi") return from_scalars(s.psi);
             if (k == "dpsi") return from_vecs<R, D>(s.dpsi);
             if (k == "node_masses") return from_scalars(s.node_masses);
             if (k == "node_moments") return from_vecs<R, D>(s.node_moments);
             if (k == "node_moments_nt") return from_vecs<R, D>(s.node_moments_nt);
             if (k == "node_ids") return from_veci<D>(s.node_ids);
             if (k == "node_types") return from_veci<D>(s.node_types);
             throw std::invalid_argument("unknown field " + k);
           })
      .def("get_material",
           [](S &s, int mat_id, const std::string &k) -> py::object {
             auto &mt = s.materials.at(mat_id);
             if (k == "eps_e") return from_mats<R, 3>(mt.eps_e);
             if (k == "stress_ref") return from_mats<R, 3>(mt.stress_ref);
             if (k == "pc") return from_scalars(mt.pc);
             if (k == "alpha") return from_scalars(mt.alpha);
             if (k == "solid_volume_fraction") return from_scalars(mt.svf);
             if (k == "bulk_modulus") return py::float_((double)mt.bulk_modulus);
             if (k == "shear_modulus") return py::float_((double)mt.shear_modulus);
             if (k == "lame_modulus") return py::float_((double)mt.lame_modulus);
             if (k == "eta_overline") return py::float_((double)mt.eta_overline);
             if (k == "M") return py::float_((double)mt.M);
             if (k == "Pt") return py::float_((double)mt.Pt);
             throw std::invalid_argument("unknown material field " + k);
           })
      .def("set_material",
           [](S &s, int mat_id, const std::string &k, py::object a) {
             auto &mt = s.materials.at(mat_id);
             if (k == "eps_e") mt.eps_e = to_mats<R, 3>(a);
             else if (k == "stress_ref") mt.stress_ref = to_mats<R, 3>(a);
             else if (k == "pc") mt.pc = to_scalars<R>(a);
             else if (k == "alpha") mt.alpha = to_scalars<R>(a);
             // scalar parameters (finite-difference checks of parameter gradients)
             else if (k == "bulk_modulus") mt.bulk_modulus = (R)a.cast<double>();
             else if (k == "shear_modulus") mt.shear_modulus = (R)a.cast<double>();
             else if (k == "lame_modulus") mt.lame_modulus = (R)a.cast<double>();
             else if (k == "M") mt.M = (R)a.cast<double>();
             else if (k == "Pt") mt.Pt = (R)a.cast<double>();
             else throw std::invalid_argument("unknown material field " + k);
           })
      .def("grid_info",
           [](S &s) {
             py::dict d;
             std::vector<int> nc(D);
             for (int a = 0; a < D; ++a) nc[a] = s.grid.num_cells[a];
             d["num_cells"] = nc;
             d["num_cells_total"] = s.grid.num_cells_total;
             d["inv_cell_size"] = (double)s.grid.inv_cell_size;
             d["nsn"] = s.win.nsn;
             return d;
           })
      .def("partition", &S::partition)
      .def("calculate_hash", &S::calculate_hash)
      .def("sort_hashes", &S::sort_hashes)
      .def("bin_particles", &S::bin_particles)
      .def("reset", &S::reset)
      .def("reset_spatial", &S::reset_spatial)
      .def("calculate_shape_function", &S::calculate_shape_function)
      .def("calculate_initial_volumes", &S::calculate_initial_volumes)
      .def("calculate_initial_masses", [](S &s, int m, double rho) { s.calculate_initial_masses(m, (R)rho); })
      .def("dp_initialize", [](S &s, int m) { s.dp_initialize(s.materials.at(m)); })
      .def("p2g", &S::p2g)
      .def("g2p", &S::g2p, py::arg("mode") = 0)
      .def("p2g_double_mapping", &S::p2g_double_mapping)
      .def("apply_node_bcs", &S::apply_node_bcs, py::arg("second") = false)
      .def("apply_particle_bcs", &S::apply_particle_bcs)
      .def("stress_update", &S::stress_update)
      .def("stress_update_material", &S::stress_update_material)
      .def("initialize", &S::initialize)
      .def("solve", &S::solve)
      .def("step", &S::step)
      .def("run", &S::run)
      .def_readwrite("musl", &S::musl)
      .def_readwrite("wrap_quirk", &S::wrap_quirk)
      .def_property("alpha", [](S &s) { return (double)s.alpha; }, [](S &s, double a) { s.alpha = (R)a; })
      .def_property("dt", [](S &s) { return (double)s.dt; }, [](S &s, double a) { s.dt = (R)a; })
      .def_readwrite("global_step", &S::global_step)
      .def_readonly("num_particles", &S::num_particles);
}

}  // namespace

PYBIND11_MODULE(_oracle, m) {
  m.doc() = "CPU oracle of the PyroclastMPM explicit MPM path (test infrastructure only)";
  bind_sim<float, 1>(m, "Sim1D_f32");
  bind_sim<float, 2>(m, "Sim2D_f32");
  bind_sim<float, 3>(m, "Sim3D_f32");
  bind_sim<double, 1>(m, "Sim1D_f64");
  bind_sim<double, 2>(m, "Sim2D_f64");
  bind_sim<double, 3>(m, "Sim3D_f64");
  m.def("cubic_N", [](double r, int type) { return (double)cubic_N<double>(r, type); });
  m.def("cubic_dN", [](double r, double inv_h, int type) { return (double)cubic_dN<double>(r, inv_h, type); });
  m.def("quad_N", [](double r) { return quad_N<double>(r); });
  m.def("quad_dN", [](double r, double inv_h) { return quad_dN<double>(r, inv_h); });
  m.def("linear_N", [](double r) { return linear_N<double>(r); });
  m.def("linear_dN", [](double r, double inv_h) { return linear_dN<double>(r, inv_h); });
}
