// rigid.hip -- rigid-particle level set contact (rigidbodylevelset.cpp:92-275,
// rigidbodylevelset_inline.h:65-343) and the slot<->original permutation.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "device.hpp"

namespace pmpm {

struct RigidMotion {
  float R[9];  // row-major D x D
  float com[3];
};

// update_rigid_velocity: v = (R (x - COM) - (x - COM)) / dt for rigid particles.
template <int D>
__global__ void rigid_velocity_kernel(pmpm_particles p, RigidMotion mo, float dt) {
  const int n = p.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(p.flags[i] & PMPM_PF_RIGID)) return;
  float r[3];
#pragma unroll
  for (int a = 0; a < D; ++a) r[a] = p.x[a * n + i] - mo.com[a];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < D; ++b) s += mo.R[a * D + b] * r[b];
    p.v[a * n + i] = (s - r[a]) / dt;
  }
}

// update_rigid_position: x += v dt for rigid particles, or (dx != null) the
// displacement stored for a later pmpm_rigid_advance.
template <int D>
__global__ void rigid_position_kernel(pmpm_particles p, float dt, float *dx) {
  const int n = p.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool rigid = (p.flags[i] & PMPM_PF_RIGID) != 0;
  if (dx) {
#pragma unroll
    for (int a = 0; a < D; ++a) dx[a * n + i] = rigid ? p.v[a * n + i] * dt : 0.f;
    return;
  }
  if (!rigid) return;
#pragma unroll
  for (int a = 0; a < D; ++a) p.x[a * n + i] += p.v[a * n + i] * dt;
}

__global__ void rigid_advance_kernel(int dim, int n, float *x, const float *dx, const int32_t *sorted) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int src = sorted ? sorted[j] : j;
  for (int a = 0; a < dim; ++a) x[a * n + j] += dx[a * n + src];
}

// calculate_grid_normals_nn_rigid: per node with mass, scan the 2^D cells of
// the linear P2G window; nearest rigid particle by cell-relative squared
// distance (ties: earliest window entry, then lowest original index, as the
// reference's (window, stable-sort) iteration), normal = sum m dpsi of the
// non-rigid particles at table index id*2^D + entry, stick contact if
// (v - v_body).n >= 0, separately for mom and mom_nt.
template <int D>
__global__ void rigid_normals_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles p, pmpm_spatial part,
                                     pmpm_nodes nodes, const float *dpsi_orig) {
  const int nn = g.num_cells_total;
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nn) return;
  const float m = nodes.mass[node];
  if (m <= 0.000000001f) return;
  const int n = p.n;
  constexpr int LIN = (D == 3) ? 8 : (D == 2 ? 4 : 2);
  int ids[3] = {0, 0, 0};
  ids[0] = node % g.num_cells[0];
  if (D > 1) ids[1] = (node / g.num_cells[0]) % g.num_cells[1];
  if (D > 2) ids[2] = node / (g.num_cells[0] * g.num_cells[1]);
  float normal[3] = {0.f, 0.f, 0.f};
  float best = 999999999999999.f;
  int best_sid = 0x7fffffff, best_id = 0x7fffffff, best_slot = -1;
  for (int sid = 0; sid < LIN; ++sid) {
    // linear P2G window: negated G2P offsets; G2P order z inner, x, y outer (3D)
    int off[3] = {0, 0, 0};
    if (D == 1) off[0] = -(sid & 1);
    else if (D == 2) { off[0] = -(sid & 1); off[1] = -((sid >> 1) & 1); }
    else { off[2] = -(sid & 1); off[0] = -((sid >> 1) & 1); off[1] = -((sid >> 2) & 1); }
    int64_t h = ids[0] + off[0];
    if (D > 1) h += (int64_t)(ids[1] + off[1]) * g.num_cells[0];
    if (D > 2) h += (int64_t)(ids[2] + off[2]) * g.num_cells[0] * g.num_cells[1];
    if (h < 0 || h >= nn) continue;
    const int cs = part.cell_start[h], ce = part.cell_end[h];
    for (int j = cs; j < ce; ++j) {
      const int slot = part.sorted[j];
      const int oid = p.id[slot];
      if (p.flags[slot] & PMPM_PF_RIGID) {
        float d2 = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const float r = (p.x[a * n + slot] - g.origin[a]) * g.inv_cell_size - (float)(ids[a] + off[a]);
          d2 += r * r;
        }
        if (d2 < best || (d2 == best && sid == best_sid && oid < best_id)) {
          best = d2; best_sid = sid; best_id = oid; best_slot = slot;
        }
      } else {
        const float pm = p.mass[slot];
#pragma unroll
        for (int a = 0; a < D; ++a) normal[a] += dpsi_orig[((int64_t)oid * LIN + sid) * D + a] * pm;
      }
    }
  }
  if (best_slot < 0) return;
  float nn2 = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) nn2 += normal[a] * normal[a];
  if (nn2 > 0.f) {
    const float nrm = sqrtf(nn2);
#pragma unroll
    for (int a = 0; a < D; ++a) normal[a] /= nrm;
  }
  float vb[3], v[3], vt[3], rc = 0.f, rct = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    vb[a] = p.v[a * n + best_slot];
    v[a] = nodes.mom[a * nn + node] / m;
    vt[a] = nodes.mom_nt[a * nn + node] / m;
    rc += (v[a] - vb[a]) * normal[a];
    rct += (vt[a] - vb[a]) * normal[a];
  }
  if (rc >= 0.f) {
#pragma unroll
    for (int a = 0; a < D; ++a) nodes.mom[a * nn + node] = (v[a] - (v[a] - vb[a])) * m;
  }
  if (rct >= 0.f) {
#pragma unroll
    for (int a = 0; a < D; ++a) nodes.mom_nt[a * nn + node] = (vt[a] - (vt[a] - vb[a])) * m;
  }
}

template <class T>
__global__ void permute_kernel(int n, int ncomp, const T *src, const int32_t *id, T *dst, int inverse) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = id[i];
  for (int c = 0; c < ncomp; ++c) {
    if (inverse) dst[(int64_t)c * n + i] = src[(int64_t)c * n + o];
    else dst[(int64_t)c * n + o] = src[(int64_t)c * n + i];
  }
}

}  // namespace pmpm

using namespace pmpm;

extern "C" int pmpm_rigid_levelset(const pmpm_params *prm, const pmpm_grid *grid, pmpm_particles *p,
                                   const pmpm_spatial *part, pmpm_nodes *nodes, const float *rotation,
                                   const float *com, const float *dpsi_orig, float *rigid_dx,
                                   pmpm_stream_t stream) {
  if (!prm || !grid_ok(grid) || !p || !part || !nodes || !rotation || !com) return PMPM_ERR_INVALID;
  if (prm->dim < 2 || prm->dim > 3) return PMPM_ERR_INVALID;  // the reference exits in 1D
  if (p->n > 0 && !dpsi_orig) return PMPM_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  RigidMotion mo;
  const int D = prm->dim;
  for (int k = 0; k < 9; ++k) mo.R[k] = k < D * D ? rotation[k] : 0.f;
  for (int a = 0; a < 3; ++a) mo.com[a] = a < D ? com[a] : 0.f;
  const unsigned pb = (unsigned)ceil_div(p->n > 0 ? p->n : 1, 256);
  const unsigned nb = (unsigned)ceil_div(grid->num_cells_total, 256);
  if (D == 2) {
    PMPM_LAUNCH(rigid_velocity_kernel<2>, (::pmpm::kname("rigid_velocity_kernel", {(long)(2)})), dim3(pb), dim3(256), 0, s, *p, mo, prm->dt);
    PMPM_LAUNCH(rigid_normals_kernel<2>, (::pmpm::kname("rigid_normals_kernel", {(long)(2)})), dim3(nb), dim3(256), 0, s, *prm, *grid, *p, *part, *nodes, dpsi_orig);
    PMPM_LAUNCH(rigid_position_kernel<2>, (::pmpm::kname("rigid_position_kernel", {(long)(2)})), dim3(pb), dim3(256), 0, s, *p, prm->dt, rigid_dx);
  } else {
    PMPM_LAUNCH(rigid_velocity_kernel<3>, (::pmpm::kname("rigid_velocity_kernel", {(long)(3)})), dim3(pb), dim3(256), 0, s, *p, mo, prm->dt);
    PMPM_LAUNCH(rigid_normals_kernel<3>, (::pmpm::kname("rigid_normals_kernel", {(long)(3)})), dim3(nb), dim3(256), 0, s, *prm, *grid, *p, *part, *nodes, dpsi_orig);
    PMPM_LAUNCH(rigid_position_kernel<3>, (::pmpm::kname("rigid_position_kernel", {(long)(3)})), dim3(pb), dim3(256), 0, s, *p, prm->dt, rigid_dx);
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

extern "C" int pmpm_permute(int32_t n, int32_t ncomp, int32_t elem_bytes, const void *src, const int32_t *id,
                            void *dst, int32_t inverse, pmpm_stream_t stream) {
  if (n < 0 || ncomp < 0) return PMPM_ERR_INVALID;
  if (n == 0 || ncomp == 0) return PMPM_OK;
  if (!src || !id || !dst || src == dst) return PMPM_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const unsigned b = (unsigned)ceil_div(n, 256);
  switch (elem_bytes) {
    case 1: PMPM_LAUNCH(permute_kernel<uint8_t>, (::pmpm::kname("permute_kernel", {1})), dim3(b), dim3(256), 0, s, n, ncomp, (const uint8_t *)src, id, (uint8_t *)dst, inverse); break;
    case 4: PMPM_LAUNCH(permute_kernel<uint32_t>, (::pmpm::kname("permute_kernel", {4})), dim3(b), dim3(256), 0, s, n, ncomp, (const uint32_t *)src, id, (uint32_t *)dst, inverse); break;
    default: return PMPM_ERR_INVALID;
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

extern "C" int pmpm_rigid_advance(int32_t dim, int32_t n, float *x, const float *rigid_dx, const int32_t *sorted,
                                  pmpm_stream_t stream) {
  if (dim < 1 || dim > 3 || n < 0) return PMPM_ERR_INVALID;
  if (n == 0) return PMPM_OK;
  if (!x || !rigid_dx) return PMPM_ERR_INVALID;
  PMPM_LAUNCH(rigid_advance_kernel, (::pmpm::kname("rigid_advance_kernel", {})), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, dim, n,
                     x, rigid_dx, sorted);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}
