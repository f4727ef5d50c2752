// output.cpp -- particle and node output of the host API: the arrays of
// ParticlesContainer::output_vtk (particles.cpp:209-320) and
// NodesContainer::output_vtk (nodes.cpp:214-244), written as the reference's
// write_vtk_polydata writes them (output.cpp:132-190): "vtk" -> VTK XML
// PolyData <stem><step>.vtp, "csv" -> a table of the point data,
// <stem><step>.csv, multi-component arrays as columns "Name:0", "Name:1", ...
// (the column names the reference's acceptance scripts read,
// test_two_disks.py:137-145). The .vtp arrays are inline base64 ("binary"
// format, UInt64 headers, no compression), readable by VTK and ParaView.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "host.hpp"

namespace pmpm_host {
namespace {

struct PointArray {
  std::string name;
  int ncomp;
  std::vector<double> v;  // point-major: v[p * ncomp + c]
  bool integer;           // written without a fraction in the CSV
};

std::string base64(const unsigned char *p, size_t n) {
  static const char *tbl = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    const unsigned v = (p[i] << 16) | (p[i + 1] << 8) | p[i + 2];
    out += tbl[(v >> 18) & 63];
    out += tbl[(v >> 12) & 63];
    out += tbl[(v >> 6) & 63];
    out += tbl[v & 63];
  }
  if (i < n) {
    const unsigned v = (p[i] << 16) | (i + 1 < n ? p[i + 1] << 8 : 0);
    out += tbl[(v >> 18) & 63];
    out += tbl[(v >> 12) & 63];
    out += i + 1 < n ? tbl[(v >> 6) & 63] : '=';
    out += '=';
  }
  return out;
}

// One "binary" DataArray payload: UInt64 byte count, then the values.
template <class T> std::string encode(const std::vector<T> &v) {
  std::vector<unsigned char> buf(8 + v.size() * sizeof(T));
  const uint64_t nbytes = v.size() * sizeof(T);
  std::memcpy(buf.data(), &nbytes, 8);
  if (!v.empty()) std::memcpy(buf.data() + 8, v.data(), nbytes);
  return base64(buf.data(), buf.size());
}

std::string file_name(const std::string &stem, const char *ext) {
  const Globals &g = globals();
  std::filesystem::create_directories(g.output_directory);
  return g.output_directory + "/" + stem + std::to_string(g.step) + ext;
}

void write_vtp(const std::string &path, const std::vector<std::array<float, 3>> &points,
               const std::vector<PointArray> &arrays, double time) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  const size_t n = points.size();
  f << "<?xml version=\"1.0\"?>\n"
    << "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
    << "  <PolyData>\n    <FieldData>\n"
    << "      <DataArray type=\"Float64\" Name=\"Time\" NumberOfTuples=\"1\" format=\"binary\">"
    << encode(std::vector<double>{time}) << "</DataArray>\n    </FieldData>\n"
    << "    <Piece NumberOfPoints=\"" << n
    << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n"
    << "      <PointData>\n";
  for (const auto &a : arrays)
    f << "        <DataArray type=\"Float64\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.ncomp
      << "\" format=\"binary\">" << encode(a.v) << "</DataArray>\n";
  std::vector<float> xyz(3 * n);
  for (size_t p = 0; p < n; ++p)
    for (int c = 0; c < 3; ++c) xyz[3 * p + c] = points[p][c];
  f << "      </PointData>\n      <CellData>\n      </CellData>\n      <Points>\n"
    << "        <DataArray type=\"Float32\" Name=\"Points\" NumberOfComponents=\"3\" format=\"binary\">"
    << encode(xyz) << "</DataArray>\n      </Points>\n";
  for (const char *cells : {"Verts", "Lines", "Strips", "Polys"})
    f << "      <" << cells << ">\n"
      << "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"binary\">"
      << encode(std::vector<int64_t>{}) << "</DataArray>\n"
      << "        <DataArray type=\"Int64\" Name=\"offsets\" format=\"binary\">" << encode(std::vector<int64_t>{})
      << "</DataArray>\n      </" << cells << ">\n";
  f << "    </Piece>\n  </PolyData>\n</VTKFile>\n";
  if (!f) throw std::runtime_error("writing " + path + " failed");
}

void write_csv(const std::string &path, const std::vector<PointArray> &arrays, size_t n) {
  std::ofstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  bool first = true;
  for (const auto &a : arrays)
    for (int c = 0; c < a.ncomp; ++c) {
      f << (first ? "" : ",") << '"' << a.name;
      if (a.ncomp > 1) f << ':' << c;
      f << '"';
      first = false;
    }
  f << '\n';
  char num[40];
  for (size_t p = 0; p < n; ++p) {
    first = true;
    for (const auto &a : arrays)
      for (int c = 0; c < a.ncomp; ++c) {
        const double v = a.v[p * a.ncomp + c];
        if (a.integer) std::snprintf(num, sizeof(num), "%lld", (long long)v);
        else std::snprintf(num, sizeof(num), "%.9g", v);
        f << (first ? "" : ",") << num;
        first = false;
      }
    f << '\n';
  }
  if (!f) throw std::runtime_error("writing " + path + " failed");
}

// write_vtk_polydata (output.cpp:132-190) for the formats this build writes.
void write_point_set(const std::string &stem, const std::vector<std::string> &formats,
                     const std::vector<std::array<float, 3>> &points, const std::vector<PointArray> &arrays) {
  const double time = (double)globals().step * (double)globals().dt;
  for (const auto &fmt : formats) {
    if (fmt == "vtk") write_vtp(file_name(stem, ".vtp"), points, arrays, time);
    else if (fmt == "csv") write_csv(file_name(stem, ".csv"), arrays, points.size());
    else
      throw std::invalid_argument("output format '" + fmt + "' is not written by this build (\"vtk\" and \"csv\" are)");
  }
}

template <int D> std::array<float, 3> point3(const VecD<D> &x) {
  std::array<float, 3> p{0.f, 0.f, 0.f};
  for (int a = 0; a < D; ++a) p[a] = x[a];
  return p;
}

// A row-major N x N matrix as the reference's column-major Eigen data order.
template <int N> void push_matrix(std::vector<double> &v, const std::array<float, N * N> &m) {
  for (int k = 0; k < N * N; ++k) v.push_back(m[(k % N) * N + k / N]);
}

}  // namespace

template <int D> void ParticlesContainer<D>::output_vtk() const {
  if (output_formats.empty()) return;
  const auto pos = positions(), vel = velocities();
  const auto st = stresses();
  const auto L = velocity_gradient(), Fm = F();
  const auto m = masses(), vol = volumes(), vol0 = volumes_original();
  const auto col = colors(), rig = is_rigid();
  const bool exclude_rigid = true;  // particles.h:248, exclude_rigid_from_output
  std::vector<std::array<float, 3>> points;
  PointArray isr{"isRigid", 1, {}, true}, P{"Positions", D, {}, false}, V{"Velocity", D, {}, false},
      S{"Stress", 9, {}, false}, VG{"VelocityGradient", D * D, {}, false}, DF{"DeformationMatrix", D * D, {}, false},
      M{"Mass", 1, {}, false}, Vo{"Volume", 1, {}, false}, V0{"VolumeOriginal", 1, {}, false},
      C{"Color", 1, {}, true}, pr{"Pressure", 1, {}, false}, q{"q", 1, {}, false}, mu{"|q|/p", 1, {}, false},
      ds{"dev strain", 1, {}, false}, vs{"vol strain", 1, {}, false};
  for (int i = 0; i < num_particles; ++i) {
    if (exclude_rigid && rig[i]) continue;
    points.push_back(point3<D>(pos[i]));
    isr.v.push_back(rig[i]);
    for (int a = 0; a < D; ++a) { P.v.push_back(pos[i][a]); V.v.push_back(vel[i][a]); }
    push_matrix<3>(S.v, st[i]);
    push_matrix<D>(VG.v, L[i]);
    push_matrix<D>(DF.v, Fm[i]);
    M.v.push_back(m[i]);
    Vo.v.push_back(vol[i]);
    V0.v.push_back(vol0[i]);
    C.v.push_back(col[i]);
    // post-processed fields (particles.cpp:239-270), in the reference's Real
    float tr = 0.f;
    for (int a = 0; a < D; ++a) tr += st[i][a * 3 + a];
    const float p = tr / 3.f;
    float ss = 0.f;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const float s = st[i][a * 3 + b] - (a == b ? p : 0.f);
        ss += s * s;
      }
    const float qq = std::sqrt(3.f * 0.5f * ss);
    float e[D][D];
    float etr = 0.f;
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) e[a][b] = 0.5f * (Fm[i][b * D + a] + Fm[i][a * D + b]) - (a == b ? 1.f : 0.f);
    for (int a = 0; a < D; ++a) etr += e[a][a];
    const float evol = -etr;
    float es = 0.f;
    for (int a = 0; a < D; ++a)
      for (int b = 0; b < D; ++b) {
        const float s = e[a][b] + (a == b ? (float)(1. / 3) * evol : 0.f);
        es += s * s;
      }
    pr.v.push_back(p);
    q.v.push_back(qq);
    mu.v.push_back(std::fabs(qq) / (-p));
    ds.v.push_back(std::sqrt(0.5f * es));
    vs.v.push_back(evol);
  }
  write_point_set("particles", output_formats, points, {isr, P, V, S, VG, DF, M, Vo, V0, C, pr, q, mu, ds, vs});
}

template <int D> void NodesContainer<D>::output_vtk() const {
  if (output_formats.empty()) return;
  if (globals().step > 1) return;  // nodes.cpp:221-224 writes the first steps only
  const auto x = give_node_coords(), P = moments(), Q = moments_nt();
  const auto m = masses();
  std::vector<std::array<float, 3>> points;
  PointArray mo{"Moments", D, {}, false}, mn{"MomentsNT", D, {}, false}, ma{"Mass", 1, {}, false};
  for (size_t i = 0; i < x.size(); ++i) {
    points.push_back(point3<D>(x[i]));
    for (int a = 0; a < D; ++a) { mo.v.push_back(P[i][a]); mn.v.push_back(Q[i][a]); }
    ma.v.push_back(m[i]);
  }
  write_point_set("nodes", output_formats, points, {mo, mn, ma});
}

// Solver::output (solver.cpp:177-194): particles and nodes. (The reference
// also calls the boundary conditions' and materials' output_vtk; none of
// those the path keeps has output formats its Python API can set.)
template <int D> void Solver<D>::output() {
  particles.output_vtk();
  nodes.output_vtk();
}

template void ParticlesContainer<1>::output_vtk() const;
template void ParticlesContainer<2>::output_vtk() const;
template void ParticlesContainer<3>::output_vtk() const;
template void NodesContainer<1>::output_vtk() const;
template void NodesContainer<2>::output_vtk() const;
template void NodesContainer<3>::output_vtk() const;
template void Solver<1>::output();
template void Solver<2>::output();
template void Solver<3>::output();

}  // namespace pmpm_host
