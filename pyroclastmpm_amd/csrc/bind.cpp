// bind.cpp -- Python module pyroclastmpm_amd._core with submodules MPM1D,
// MPM2D, MPM3D: the reference's pybind API (python/pybind/**,
// _pyroclastmpm.cpp:32-115) with the same class, method, property and keyword
// names, over the host objects of host.hpp.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>

#include "host.hpp"

namespace py = pybind11;
using namespace pmpm_host;

namespace {

template <int D> struct USL : Solver<D> {
  USL(const ParticlesContainer<D> &p, const NodesContainer<D> &n, const std::vector<MaterialType> &m,
      const std::vector<BoundaryConditionType<D>> &b, float alpha)
      : Solver<D>(p, n, m, b, alpha, false) {}
};
template <int D> struct MUSL : Solver<D> {
  MUSL(const ParticlesContainer<D> &p, const NodesContainer<D> &n, const std::vector<MaterialType> &m,
       const std::vector<BoundaryConditionType<D>> &b, float alpha)
      : Solver<D>(p, n, m, b, alpha, true) {}
};

using F32 = py::array_t<float, py::array::c_style | py::array::forcecast>;

template <int D> std::vector<VecD<D>> to_vecs(const py::object &o) {
  std::vector<VecD<D>> out;
  if (o.is_none()) return out;
  F32 a = py::array::ensure(o, py::array::c_style | py::array::forcecast);
  if (!a) a = F32(o);
  if (a.size() == 0) return out;
  if (a.ndim() == 1 && D == 1) {
    out.resize(a.shape(0));
    for (ssize_t i = 0; i < a.shape(0); ++i) out[i][0] = a.at(i);
    return out;
  }
  if (a.ndim() == 1 && a.shape(0) == D) {  // a single point
    out.resize(1);
    for (int d = 0; d < D; ++d) out[0][d] = a.at(d);
    return out;
  }
  if (a.ndim() != 2 || a.shape(1) != D)
    throw std::invalid_argument("expected an array of shape (N, " + std::to_string(D) + ")");
  out.resize(a.shape(0));
  for (ssize_t i = 0; i < a.shape(0); ++i)
    for (int d = 0; d < D; ++d) out[i][d] = a.at(i, d);
  return out;
}
template <int D> VecD<D> to_vec(const py::object &o) {
  F32 a(o);
  VecD<D> v{};
  if (a.size() == 1) { for (int d = 0; d < D; ++d) v[d] = a.data()[0]; return v; }
  if (a.size() != D) throw std::invalid_argument("expected a vector of length " + std::to_string(D));
  for (int d = 0; d < D; ++d) v[d] = a.data()[d];
  return v;
}
template <int D> VecI<D> to_veci(const py::object &o) {
  py::array_t<int, py::array::c_style | py::array::forcecast> a(o);
  VecI<D> v{};
  if (a.size() == 1) { for (int d = 0; d < D; ++d) v[d] = a.data()[0]; return v; }
  if (a.size() != D) throw std::invalid_argument("expected an integer vector of length " + std::to_string(D));
  for (int d = 0; d < D; ++d) v[d] = a.data()[d];
  return v;
}
template <int D> py::array from_vecs(const std::vector<VecD<D>> &v) {
  F32 a({(ssize_t)v.size(), (ssize_t)D});
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * D * sizeof(float));
  return a;
}
template <int N> std::vector<std::array<float, N * N>> to_mats(const py::object &o) {
  F32 a(o);
  std::vector<std::array<float, N * N>> out;
  if (a.size() == 0) return out;
  if (a.ndim() == 2 && a.shape(0) == N && a.shape(1) == N) {  // one matrix
    out.resize(1);
    std::memcpy(out[0].data(), a.data(), N * N * sizeof(float));
    return out;
  }
  if (a.ndim() != 3 || a.shape(1) != N || a.shape(2) != N)
    throw std::invalid_argument("expected an array of shape (M, " + std::to_string(N) + ", " + std::to_string(N) + ")");
  out.resize(a.shape(0));
  std::memcpy(out.data(), a.data(), out.size() * N * N * sizeof(float));
  return out;
}
template <int N> py::array from_mats(const std::vector<std::array<float, N * N>> &v) {
  F32 a({(ssize_t)v.size(), (ssize_t)N, (ssize_t)N});
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * N * N * sizeof(float));
  return a;
}
std::vector<float> to_floats(const py::object &o) {
  F32 a(o);
  return std::vector<float>(a.data(), a.data() + a.size());
}
std::vector<int> to_ints(const py::object &o) {
  if (o.is_none()) return {};
  py::array_t<int, py::array::c_style | py::array::forcecast> a(o);
  return std::vector<int>(a.data(), a.data() + a.size());
}
template <class T> py::array arr(const std::vector<T> &v) { return py::array_t<T>((ssize_t)v.size(), v.data()); }
template <int D> py::array veci_arr(const std::vector<VecI<D>> &v) {
  py::array_t<int> a({(ssize_t)v.size(), (ssize_t)D});
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * D * sizeof(int));
  return a;
}

// stand-alone stress update of one material (the reference's
// material.stress_update(particles, mat_id) binding: particles by value)
template <int D, class M> py::tuple standalone_update(M &mat, ParticlesContainer<D> p, int mat_id) {
  pmpm_params prm{};
  prm.dim = D;
  prm.shape_fn = globals().shape_function;
  prm.dt = globals().dt;
  prm.num_materials = mat_id + 1;
  if (mat_id < 0 || mat_id >= PMPM_MAX_MATERIALS) throw std::invalid_argument("mat_id out of range");
  prm.materials[mat_id] = mat.device();
  pmpm_particles v = p.view(p.cur_index());
  check(pmpm_stress_update(&prm, mat_id, &v, nullptr), "pmpm_stress_update");
  hip_check(hipDeviceSynchronize(), "sync");
  return py::make_tuple(p, mat_id);
}

template <int D> void bind_dim(py::module &m, py::module &root) {
  m.attr("global_dimension") = D;
  using P = ParticlesContainer<D>;
  using N = NodesContainer<D>;

  // ------------------------------------------------------------ particles
  py::class_<P> pc(m, "ParticlesContainer", py::dynamic_attr());
  pc.def(py::init([](py::object pos, py::object vel, py::object col, py::object rpos, py::object rvel,
                     py::object rcol) {
           return new P(to_vecs<D>(pos), to_vecs<D>(vel), to_ints(col), to_vecs<D>(rpos), to_vecs<D>(rvel),
                        to_ints(rcol));
         }),
         py::arg("positions"), py::arg("velocities") = py::none(), py::arg("colors") = py::none(),
         py::arg("rigid_positions") = py::none(), py::arg("rigid_velocities") = py::none(),
         py::arg("rigid_colors") = py::none());
  pc.def("set_output_formats", &P::set_output_formats, py::arg("output_formats"));
  pc.def("output_vtk", &P::output_vtk);
  pc.def("set_spawner", &P::set_spawner, py::arg("rate"), py::arg("volume"));
  pc.def_readwrite("isRestart", &P::isRestart);
  pc.def_readonly("num_particles", &P::num_particles);
  pc.def_readonly("num_rigid_particles", &P::num_rigid_particles);
  pc.def_readonly("output_formats", &P::output_formats);
  pc.def_property("positions", [](P &s) { return from_vecs<D>(s.positions()); },
                  [](P &s, py::object v) { s.set_positions(to_vecs<D>(v)); });
  pc.def_property("velocities", [](P &s) { return from_vecs<D>(s.velocities()); },
                  [](P &s, py::object v) { s.set_velocities(to_vecs<D>(v)); });
  pc.def_property("stresses", [](P &s) { return from_mats<3>(s.stresses()); },
                  [](P &s, py::object v) { s.set_stresses(to_mats<3>(v)); });
  pc.def_property("F", [](P &s) { return from_mats<D>(s.F()); },
                  [](P &s, py::object v) { s.set_F(to_mats<D>(v)); });
  pc.def_property("velocity_gradient", [](P &s) { return from_mats<D>(s.velocity_gradient()); },
                  [](P &s, py::object v) { s.set_velocity_gradient(to_mats<D>(v)); });
  pc.def_property("masses", [](P &s) { return arr(s.masses()); },
                  [](P &s, py::object v) { s.set_masses(to_floats(v)); });
  pc.def_property("volumes", [](P &s) { return arr(s.volumes()); },
                  [](P &s, py::object v) { s.set_volumes(to_floats(v)); });
  pc.def_property("volumes_original", [](P &s) { return arr(s.volumes_original()); },
                  [](P &s, py::object v) { s.set_volumes_original(to_floats(v)); });
  pc.def_property("colors", [](P &s) { return arr(s.colors()); },
                  [](P &s, py::object v) { s.set_colors(to_ints(v)); });
  pc.def_property("is_active", [](P &s) { return arr(s.is_active()); },
                  [](P &s, py::object v) { s.set_is_active(to_ints(v)); });
  pc.def_property_readonly("is_rigid", [](P &s) { return arr(s.is_rigid()); });
  // the plastic history the post-G2P materials keep per particle (the
  // reference keeps it in the material: acc_eps_p_gpu, alpha_gpu, pc_gpu)
  pc.def_property_readonly("plastic_history", [](P &s) {
    auto h = s.plastic_history();
    py::array_t<float> a({(ssize_t)2, (ssize_t)s.num_particles});
    if (!h.empty()) std::memcpy(a.mutable_data(), h.data(), h.size() * sizeof(float));
    return a;
  });
  pc.def_property("eps_e", [](P &s) { return from_mats<3>(s.eps_e()); },
                  [](P &s, py::object v) { s.set_eps_e(to_mats<3>(v)); });
  pc.def("set_spatialpartition", [](P &s, const N &nodes) { s.set_spatialpartition(nodes.grid); }, py::arg("nodes"));
  pc.def("partition", &P::partition);
  pc.def("calculate_initial_volumes", &P::calculate_initial_volumes);
  pc.def("calculate_initial_masses", &P::calculate_initial_masses, py::arg("mat_id"), py::arg("density"));
  pc.def_property_readonly("spatial_bins", [](P &s) { return veci_arr<D>(s.bins()); });
  pc.def_property_readonly("spatial_hash_unsorted", [](P &s) { return arr(s.hash_unsorted()); });
  pc.def_property_readonly("spatial_hash_sorted", [](P &s) { return arr(s.hash_sorted()); });
  pc.def_property_readonly("spatial_sorted_index", [](P &s) { return arr(s.sorted_index()); });
  pc.def_property_readonly("spatial_cell_start", [](P &s) { return arr(s.cell_start()); });
  pc.def_property_readonly("spatial_cell_end", [](P &s) { return arr(s.cell_end()); });
  // pickle (particles_bind.cpp:39-83): same tuple order; psi/dpsi are not
  // kept between steps here, so their entries are empty
  pc.def(py::pickle(
      [](P &s) {
        return py::make_tuple(from_vecs<D>(s.positions()), from_vecs<D>(s.velocities()), arr(s.colors()),
                              arr(s.is_rigid()), from_mats<3>(s.stresses()), arr(s.masses()), arr(s.volumes()),
                              s.output_formats, from_mats<D>(s.F()), from_mats<D>(s.velocity_gradient()), py::list(),
                              arr(s.volumes_original()), py::list());
      },
      [](py::tuple t) {
        auto pos = to_vecs<D>(t[0]), vel = to_vecs<D>(t[1]);
        auto col = to_ints(t[2]), rig = to_ints(t[3]);
        std::vector<VecD<D>> p0, v0, pr, vr;
        std::vector<int> c0, cr;
        for (size_t i = 0; i < pos.size(); ++i) {
          const bool r = i < rig.size() && rig[i];
          (r ? pr : p0).push_back(pos[i]);
          (r ? vr : v0).push_back(vel[i]);
          (r ? cr : c0).push_back(col[i]);
        }
        auto *p = new P(p0, v0, c0, pr, vr, cr);
        p->set_stresses(to_mats<3>(t[4]));
        p->set_masses(to_floats(t[5]));
        p->set_volumes(to_floats(t[6]));
        p->output_formats = t[7].cast<std::vector<std::string>>();
        p->set_F(to_mats<D>(t[8]));
        p->set_velocity_gradient(to_mats<D>(t[9]));
        p->set_volumes_original(to_floats(t[11]));
        return p;
      }));

  // ---------------------------------------------------------------- nodes
  py::class_<N> nc(m, "NodesContainer", py::dynamic_attr());
  nc.def(py::init([](py::object o, py::object e, float h) { return new N(to_vec<D>(o), to_vec<D>(e), h); }),
         py::arg("origin"), py::arg("end"), py::arg("cell_size"));
  nc.def("set_output_formats", [](N &s, const std::vector<std::string> &f) { s.output_formats = f; },
         py::arg("output_formats"));
  nc.def("give_coords", [](N &s) { return from_vecs<D>(s.give_node_coords()); });
  nc.def("output_vtk", &N::output_vtk);
  nc.def("reset", &N::reset);
  nc.def_property("moments", [](N &s) { return from_vecs<D>(s.moments()); },
                  [](N &s, py::object v) { s.set_moments(to_vecs<D>(v)); });
  nc.def_property("moments_nt", [](N &s) { return from_vecs<D>(s.moments_nt()); },
                  [](N &s, py::object v) { s.set_moments_nt(to_vecs<D>(v)); });
  nc.def_property("masses", [](N &s) { return arr(s.masses()); },
                  [](N &s, py::object v) { s.set_masses(to_floats(v)); });
  nc.def_property_readonly("node_ids", [](N &s) { return veci_arr<D>(s.node_ids()); });
  nc.def_property_readonly("node_types", [](N &s) { return veci_arr<D>(s.node_types()); });
  nc.def_property_readonly("num_cells", [](N &s) {
    std::vector<int> v(D);
    for (int a = 0; a < D; ++a) v[a] = s.grid.num_cells[a];
    return v;
  });
  nc.def_property_readonly("num_cells_total", [](N &s) { return s.grid.num_cells_total; });
  nc.def_readwrite("small_mass_cutoff", &N::small_mass_cutoff);

  // materials are dimension-free types, registered once on the root module
  for (const char *k : {"LinearElastic", "NewtonFluid", "LocalGranularRheology", "DP_rheo", "VonMises", "MohrCoulomb",
                        "ModifiedCamClay"})
    m.attr(k) = root.attr(k);
  // Friction-angle / cohesion form of DP_rheo (outer Mohr-Coulomb cone):
  // M = 6 sin(phi) / (3 - sin(phi)), Pt = c cot(phi). Not in the reference.
  m.def(
      "DruckerPrager",
      [](float density, float E, float pois, float phi_deg, float cohesion, float psi_deg) {
        const double phi = (double)phi_deg * 3.14159265358979323846 / 180.0;
        const double sp = std::sin(phi);
        const float M = (float)(6.0 * sp / (3.0 - sp));
        const float Pt = sp > 0 ? (float)((double)cohesion * std::cos(phi) / sp) : 0.f;
        return DP_rheo(density, E, pois, M, 0.f, 0.f, 0.f, 0.f, Pt, 0.f, psi_deg, 1.f);
      },
      py::arg("density"), py::arg("E"), py::arg("pois"), py::arg("friction_angle"), py::arg("cohesion") = 0.f,
      py::arg("dilatancy_angle") = 0.f);

  // -------------------------------------------------- boundary conditions
  py::class_<Gravity<D>>(m, "Gravity", py::dynamic_attr())
      .def(py::init([](py::object g, bool is_ramp, int ramp_step, py::object g_end) {
             VecD<D> ge{};
             if (!g_end.is_none()) ge = to_vec<D>(g_end);
             return new Gravity<D>(to_vec<D>(g), is_ramp, ramp_step, ge);
           }),
           py::arg("gravity"), py::arg("is_ramp") = false, py::arg("ramp_step") = 0,
           py::arg("gravity_end") = py::none())
      .def_property("gravity", [](Gravity<D> &s) { return arr(std::vector<float>(s.gravity.begin(), s.gravity.end())); },
                    [](Gravity<D> &s, py::object v) { s.gravity = to_vec<D>(v); })
      .def_readwrite("is_ramp", &Gravity<D>::is_ramp)
      .def_readwrite("ramp_step", &Gravity<D>::ramp_step);
  py::class_<NodeDomain<D>>(m, "NodeDomain", py::dynamic_attr())
      .def(py::init([](py::object f0, py::object f1) {
             VecI<D> a{}, b{};
             if (!f0.is_none()) a = to_veci<D>(f0);
             if (!f1.is_none()) b = to_veci<D>(f1);
             return new NodeDomain<D>(a, b);
           }),
           py::arg("face0_mode") = py::none(), py::arg("face1_mode") = py::none());
  // bodyforce_bind.cpp:40-77
  py::class_<BodyForce<D>>(m, "BodyForce", py::dynamic_attr())
      .def(py::init([](const std::string &mode, py::object values, py::object mask) {
             auto v = to_vecs<D>(values);
             std::vector<bool> mk;
             if (!mask.is_none()) {
               py::array_t<bool, py::array::c_style | py::array::forcecast> a(mask);
               mk.assign(a.data(), a.data() + a.size());
             }
             return new BodyForce<D>(mode, v, mk);
           }),
           py::arg("mode"), py::arg("values"), py::arg("mask"))
      .def_readonly("mode_id", &BodyForce<D>::mode_id)
      .def_readwrite("isActive", &BodyForce<D>::isActive);
  // planardomain_bind.cpp:41-73 (friction angles in degrees)
  py::class_<PlanarDomain<D>>(m, "PlanarDomain", py::dynamic_attr())
      .def(py::init([](py::object f0, py::object f1) {
             VecD<D> a{}, b{};
             if (!f0.is_none()) a = to_vec<D>(f0);
             if (!f1.is_none()) b = to_vec<D>(f1);
             return new PlanarDomain<D>(a, b);
           }),
           py::arg("face0_friction") = py::none(), py::arg("face1_friction") = py::none())
      .def_property_readonly("face0_friction", [](PlanarDomain<D> &s) {
        return arr(std::vector<float>(s.face0_friction.begin(), s.face0_friction.end()));
      })
      .def_property_readonly("face1_friction", [](PlanarDomain<D> &s) {
        return arr(std::vector<float>(s.face1_friction.begin(), s.face1_friction.end()));
      })
      .def_readwrite("isActive", &PlanarDomain<D>::isActive);
  if (D > 1) {
    py::class_<RigidBodyLevelSet<D>>(m, "RigidBodyLevelSet", py::dynamic_attr())
        .def(py::init<>())
        .def("set_mode_loop_rotate",
             [](RigidBodyLevelSet<D> &s, py::object eps, py::object com) {
               VecD<D> c{};
               if (!com.is_none()) c = to_vec<D>(com);
               s.set_mode_loop_rotate(to_vec<D>(eps), c);
             },
             py::arg("euler_angles_per_second"), py::arg("COM") = py::none());
  }
  // Analytic box rigid body with a prescribed velocity (no reference class).
  py::class_<RigidBodySDF<D>>(m, "RigidBodySDF", py::dynamic_attr())
      .def(py::init([](py::object half, py::object center, py::object vel, py::object rot, float band) {
             auto h = to_floats(half), c = to_floats(center), v = to_floats(vel);
             std::array<float, 3> H{}, C{}, V{};
             for (int a = 0; a < 3; ++a) {
               H[a] = a < (int)h.size() ? h[a] : 0.f;
               C[a] = a < (int)c.size() ? c[a] : 0.f;
               V[a] = a < (int)v.size() ? v[a] : 0.f;
             }
             std::array<float, 9> R{1, 0, 0, 0, 1, 0, 0, 0, 1};
             if (!rot.is_none()) {
               auto r = to_floats(rot);
               if (r.size() != 9) throw std::invalid_argument("rotation: 3x3 expected");
               for (int k = 0; k < 9; ++k) R[k] = r[k];
             }
             return new RigidBodySDF<D>(H, C, V, R, band);
           }),
           py::arg("half_extents"), py::arg("center"), py::arg("velocity"), py::arg("rotation") = py::none(),
           py::arg("band") = 0.f)
      .def_property_readonly("center", [](RigidBodySDF<D> &s) { return arr(std::vector<float>(s.center.begin(), s.center.end())); });

  // -------------------------------------------------------------- solvers
  using S = Solver<D>;
  py::class_<S>(m, "Solver")
      .def("run", &S::run, py::arg("total_steps"), py::arg("output_frequency"))
      .def("solve", &S::solve)
      .def("solve_nsteps", &S::solve_nsteps, py::arg("n_steps"))
      .def("stress_update", &S::stress_update)
      .def("partition", &S::partition)
      .def("P2G", &S::P2G)
      .def("G2P", &S::G2P)
      .def("G2P_double_mapping", &S::G2P_double_mapping)
      .def("P2G_double_mapping", &S::P2G_double_mapping)
      .def("apply_boundary_conditions", &S::apply_boundary_conditions, py::arg("remap_pass") = false)
      .def_readwrite("nodes", &S::nodes)
      .def_readwrite("particles", &S::particles)
      .def_readwrite("alpha", &S::alpha)
      .def_readwrite("wrap_quirk", &S::wrap_quirk)
      .def_readwrite("materials", &S::materials)
      .def_property(
          "boundaryconditions", [](S &s) { return s.boundaryconditions; },
          [](S &s, const std::vector<BoundaryConditionType<D>> &v) { s.boundaryconditions = v; });
  py::class_<USL<D>, S>(m, "USL").def(
      py::init<const P &, const N &, const std::vector<MaterialType> &, const std::vector<BoundaryConditionType<D>> &,
               float>(),
      py::arg("particles"), py::arg("nodes"), py::arg("materials") = std::vector<MaterialType>(),
      py::arg("boundaryconditions") = std::vector<BoundaryConditionType<D>>(), py::arg("alpha") = 0.99f);
  py::class_<MUSL<D>, S>(m, "MUSL").def(
      py::init<const P &, const N &, const std::vector<MaterialType> &, const std::vector<BoundaryConditionType<D>> &,
               float>(),
      py::arg("particles"), py::arg("nodes"), py::arg("materials") = std::vector<MaterialType>(),
      py::arg("boundaryconditions") = std::vector<BoundaryConditionType<D>>(), py::arg("alpha") = 0.99f);

  // --------------------------------------------------------- global settings
  m.def("set_globals", &set_globals, py::arg("time_step"), py::arg("particles_per_cell"),
        py::arg("shape_function") = "linear", py::arg("output_directory") = "./output/");
  m.def("set_global_shapefunction", &set_global_shapefunction, py::arg("shapefunction") = "linear");
  m.def("set_global_timestep", &set_global_timestep, py::arg("timestep"));
  m.def("set_global_step", &set_global_step, py::arg("step"));
  m.def("set_global_output_directory", &set_global_output_directory, py::arg("directory"));
  m.def("set_global_particles_per_cell", &set_global_particles_per_cell, py::arg("particles_per_cell"));
  m.def("get_global_step", []() { return globals().step; });
  // calculate_shape_function (shapefunction.cpp:167-190): psi [N, nsn] and
  // dpsi [N, nsn, D] in the reference's layout and particle order
  m.def(
      "calculate_shape_function",
      [](N &nodes, P &p) {
        const int n = p.num_particles;
        const int sf = globals().shape_function;
        const int w = (sf == PMPM_SF_LINEAR) ? 2 : 4;
        int nsn = 1;
        for (int a = 0; a < D; ++a) nsn *= w;
        DevBuf<float> xr((size_t)n * D), psi((size_t)n * nsn), dpsi((size_t)n * nsn * D);
        check(pmpm_permute(n, D, 4, p.cur().x.get(), p.cur().id.get(), xr.get(), 0, nullptr), "pmpm_permute");
        check(pmpm_shape_function(D, sf, nodes.node_types_built ? 1 : 0, &nodes.grid, xr.get(), n, psi.get(),
                                  dpsi.get(), nullptr),
              "pmpm_shape_function");
        auto hp = psi.download(), hd = dpsi.download();
        py::array_t<float> a({(ssize_t)n, (ssize_t)nsn}), b({(ssize_t)n, (ssize_t)nsn, (ssize_t)D});
        if (n) {
          std::memcpy(a.mutable_data(), hp.data(), hp.size() * 4);
          std::memcpy(b.mutable_data(), hd.data(), hd.size() * 4);
        }
        return py::make_tuple(a, b);
      },
      py::arg("nodes"), py::arg("particles"));
}


template <class M, int D> void add_update(py::class_<M> &c) {
  c.def("stress_update", [](M &s, ParticlesContainer<D> p, int id) { return standalone_update<D>(s, p, id); },
        py::arg("particles"), py::arg("mat_id") = 0);
}
// stress_update of a material whose history (elastic strain, plastic
// scalars) travels with the particles
template <class M, int D> void add_plastic_update(py::class_<M> &c) {
  c.def("stress_update",
        [](M &s, ParticlesContainer<D> p, int id) {
          p.ensure_plastic_history();
          return standalone_update<D>(s, p, id);
        },
        py::arg("particles"), py::arg("mat_id") = 0);
}
template <int D> void add_dp_methods(py::class_<DP_rheo> &dp) {
  using P = ParticlesContainer<D>;
  dp.def("initialize",
         [](DP_rheo &s, P p, int mat_id) {  // dp_rheo.cpp:75-118
           p.ensure_history();
           auto st = p.stresses();
           p.set_eps_e(std::vector<Mat3>(p.num_particles, Mat3{}));
           p.cur().stress_ref = p.cur().stress;
           bool nz = false;
           s.pc.assign(st.size(), 0.f);
           for (size_t i = 0; i < st.size(); ++i) {
             for (float x : st[i]) nz = nz || x != 0.f;
             s.pc[i] = (-(st[i][0] + st[i][4] + st[i][8]) / 3.f) * s.R;
           }
           s.stress_ref_nonzero = nz;
           return py::make_tuple(p, mat_id);
         },
         py::arg("particles"), py::arg("mat_id") = 0);
  dp.def("stress_update",
         [](DP_rheo &s, P p, int id) {
           p.ensure_history();
           return standalone_update<D>(s, p, id);
         },
         py::arg("particles"), py::arg("mat_id") = 0);
}

void bind_materials(py::module &m) {
  py::class_<LinearElastic> le(m, "LinearElastic", py::dynamic_attr());
  le.def(py::init<float, float, float>(), py::arg("density"), py::arg("E"), py::arg("pois") = 0.f)
      .def_readwrite("E", &LinearElastic::E)
      .def_readwrite("pois", &LinearElastic::pois)
      .def_readwrite("shear_modulus", &LinearElastic::shear_modulus)
      .def_readwrite("lame_modulus", &LinearElastic::lame_modulus)
      .def_readwrite("bulk_modulus", &LinearElastic::bulk_modulus)
      .def_readwrite("density", &LinearElastic::density)
      .def(py::pickle(
          [](const LinearElastic &a) {
            return py::make_tuple(a.density, a.E, a.pois, a.shear_modulus, a.lame_modulus, a.bulk_modulus);
          },
          [](py::tuple t) {
            LinearElastic a(t[0].cast<float>(), t[1].cast<float>(), t[2].cast<float>());
            a.shear_modulus = t[3].cast<float>();
            a.lame_modulus = t[4].cast<float>();
            a.bulk_modulus = t[5].cast<float>();
            return a;
          }));
  py::class_<NewtonFluid> nf(m, "NewtonFluid", py::dynamic_attr());
  nf.def(py::init<float, float, float, float>(), py::arg("density"), py::arg("viscosity"),
         py::arg("bulk_modulus") = 0.f, py::arg("gamma") = 7.f)
      .def_readwrite("density", &NewtonFluid::density)
      .def_readwrite("viscosity", &NewtonFluid::viscosity)
      .def_readwrite("bulk_modulus", &NewtonFluid::bulk_modulus)
      .def_readwrite("gamma", &NewtonFluid::gamma)
      .def(py::pickle([](const NewtonFluid &a) { return py::make_tuple(a.density, a.viscosity, a.bulk_modulus, a.gamma); },
                      [](py::tuple t) {
                        return NewtonFluid(t[0].cast<float>(), t[1].cast<float>(), t[2].cast<float>(),
                                           t[3].cast<float>());
                      }));
  py::class_<LocalGranularRheology> lr(m, "LocalGranularRheology", py::dynamic_attr());
  lr.def(py::init<float, float, float, float, float, float, float, float, float>(), py::arg("density"), py::arg("E"),
         py::arg("pois"), py::arg("I0"), py::arg("mu_s"), py::arg("mu_2"), py::arg("rho_c"),
         py::arg("particle_diameter"), py::arg("particle_density"))
      .def_readwrite("density", &LocalGranularRheology::density)
      .def_readwrite("E", &LocalGranularRheology::E)
      .def_readwrite("pois", &LocalGranularRheology::pois)
      .def_readwrite("shear_modulus", &LocalGranularRheology::shear_modulus)
      .def_readwrite("lame_modulus", &LocalGranularRheology::lame_modulus)
      .def_readwrite("bulk_modulus", &LocalGranularRheology::bulk_modulus)
      .def_readwrite("mu_s", &LocalGranularRheology::mu_s)
      .def_readwrite("mu_2", &LocalGranularRheology::mu_2)
      .def_readwrite("I0", &LocalGranularRheology::I0)
      .def_readwrite("rho_c", &LocalGranularRheology::rho_c)
      .def_readwrite("EPS", &LocalGranularRheology::EPS)
      .def_readwrite("particle_density", &LocalGranularRheology::particle_density)
      .def_readwrite("particle_diameter", &LocalGranularRheology::particle_diameter);
  py::class_<DP_rheo> dp(m, "DP_rheo", py::dynamic_attr());
  dp.def(py::init<float, float, float, float, float, float, float, float, float, float, float, float>(),
         py::arg("density"), py::arg("E"), py::arg("pois"), py::arg("M"), py::arg("lam"), py::arg("kap"),
         py::arg("Vs"), py::arg("R"), py::arg("Pt"), py::arg("beta"), py::arg("dilatancy_angle") = 0.f,
         py::arg("solid_volume_fraction") = 1.f)
      .def_readwrite("beta", &DP_rheo::beta)
      .def_readwrite("M", &DP_rheo::M)
      .def_readwrite("Pt", &DP_rheo::Pt)
      .def_readwrite("E", &DP_rheo::E)
      .def_readwrite("pois", &DP_rheo::pois)
      .def_readwrite("shear_modulus", &DP_rheo::shear_modulus)
      .def_readwrite("lame_modulus", &DP_rheo::lame_modulus)
      .def_readwrite("bulk_modulus", &DP_rheo::bulk_modulus)
      .def_readwrite("density", &DP_rheo::density)
      .def_readwrite("eta_overline", &DP_rheo::eta_overline)
      .def_readwrite("Vs", &DP_rheo::Vs)
      .def_readwrite("do_update_history", &DP_rheo::do_update_history)
      .def_readwrite("is_velgrad_strain_increment", &DP_rheo::is_velgrad_strain_increment)
      .def_readwrite("apex_guard", &DP_rheo::apex_guard)
      .def_property_readonly("pc_gpu", [](DP_rheo &s) { return arr(s.pc); });
  add_update<LinearElastic, 1>(le); add_update<LinearElastic, 2>(le); add_update<LinearElastic, 3>(le);
  add_update<NewtonFluid, 1>(nf); add_update<NewtonFluid, 2>(nf); add_update<NewtonFluid, 3>(nf);
  add_update<LocalGranularRheology, 1>(lr); add_update<LocalGranularRheology, 2>(lr);
  add_update<LocalGranularRheology, 3>(lr);
  add_dp_methods<1>(dp); add_dp_methods<2>(dp); add_dp_methods<3>(dp);
  // vonmises_bind.cpp:15-45
  py::class_<VonMises> vm(m, "VonMises", py::dynamic_attr());
  vm.def(py::init<float, float, float, float, float>(), py::arg("density"), py::arg("E"), py::arg("pois"),
         py::arg("yield_stress"), py::arg("H"))
      .def_readwrite("E", &VonMises::E)
      .def_readwrite("pois", &VonMises::pois)
      .def_readwrite("shear_modulus", &VonMises::shear_modulus)
      .def_readwrite("bulk_modulus", &VonMises::bulk_modulus)
      .def_readwrite("density", &VonMises::density)
      .def_readwrite("yield_stress", &VonMises::yield_stress)
      .def_readwrite("H", &VonMises::H)
      .def_readwrite("do_update_history", &VonMises::do_update_history)
      .def_readwrite("is_velgrad_strain_increment", &VonMises::is_velgrad_strain_increment);
  // mohr_coulomb_bind.cpp:16-52
  py::class_<MohrCoulomb> mc(m, "MohrCoulomb", py::dynamic_attr());
  mc.def(py::init<float, float, float, float, float, float, float>(), py::arg("density"), py::arg("E"),
         py::arg("pois") = 0.f, py::arg("cohesion") = 0.f, py::arg("friction_angle") = 0.f,
         py::arg("dilatancy_angle") = 0.f, py::arg("H") = 0.f)
      .def_readwrite("E", &MohrCoulomb::E)
      .def_readwrite("pois", &MohrCoulomb::pois)
      .def_readwrite("shear_modulus", &MohrCoulomb::shear_modulus)
      .def_readwrite("lame_modulus", &MohrCoulomb::lame_modulus)
      .def_readwrite("bulk_modulus", &MohrCoulomb::bulk_modulus)
      .def_readwrite("density", &MohrCoulomb::density)
      .def_readwrite("cohesion", &MohrCoulomb::cohesion)
      .def_readwrite("friction_angle", &MohrCoulomb::friction_angle)
      .def_readwrite("dilatancy_angle", &MohrCoulomb::dilatancy_angle)
      .def_readwrite("H", &MohrCoulomb::H)
      .def_readwrite("do_update_history", &MohrCoulomb::do_update_history)
      .def_readwrite("is_velgrad_strain_increment", &MohrCoulomb::is_velgrad_strain_increment);
  // modified_camclay_bind.cpp:15-49
  py::class_<ModifiedCamClay> mcc(m, "ModifiedCamClay", py::dynamic_attr());
  mcc.def(py::init<float, float, float, float, float, float, float, float, float, float>(), py::arg("density"),
          py::arg("E"), py::arg("pois"), py::arg("M"), py::arg("lam"), py::arg("kap"), py::arg("Vs"), py::arg("R"),
          py::arg("Pt"), py::arg("beta"))
      .def_readwrite("E", &ModifiedCamClay::E)
      .def_readwrite("pois", &ModifiedCamClay::pois)
      .def_readwrite("shear_modulus", &ModifiedCamClay::shear_modulus)
      .def_readwrite("bulk_modulus", &ModifiedCamClay::bulk_modulus)
      .def_readwrite("density", &ModifiedCamClay::density)
      .def_readwrite("M", &ModifiedCamClay::M)
      .def_readwrite("lam", &ModifiedCamClay::lam)
      .def_readwrite("kap", &ModifiedCamClay::kap)
      .def_readwrite("Vs", &ModifiedCamClay::Vs)
      .def_readwrite("R", &ModifiedCamClay::R)
      .def_readwrite("Pt", &ModifiedCamClay::Pt)
      .def_readwrite("beta", &ModifiedCamClay::beta)
      .def_readwrite("do_update_history", &ModifiedCamClay::do_update_history)
      .def_readwrite("is_velgrad_strain_increment", &ModifiedCamClay::is_velgrad_strain_increment);
  add_plastic_update<VonMises, 3>(vm);
  add_plastic_update<MohrCoulomb, 1>(mc); add_plastic_update<MohrCoulomb, 2>(mc); add_plastic_update<MohrCoulomb, 3>(mc);
  add_plastic_update<ModifiedCamClay, 1>(mcc); add_plastic_update<ModifiedCamClay, 2>(mcc);
  add_plastic_update<ModifiedCamClay, 3>(mcc);
}

}  // namespace

PYBIND11_MODULE(_core, m) {
  m.doc() = "MI355X explicit MPM (PyroclastMPM API) -- host objects over libpmpm_hip.so";
  m.attr("version") = pmpm_version();
  auto m1 = m.def_submodule("MPM1D");
  auto m2 = m.def_submodule("MPM2D");
  auto m3 = m.def_submodule("MPM3D");
  bind_materials(m);
  bind_dim<1>(m1, m);
  bind_dim<2>(m2, m);
  bind_dim<3>(m3, m);
  m.def("device_count", []() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    return n;
  });
}
