// transfer.hip -- particle/grid transfers on gfx950.
//
//   p2g_kernel        one workgroup per tile of cells, one lane per (cell, y
//                     node offset); each lane walks its cell's particles (cell
//                     order from the radix sort) and sums mass, momentum and
//                     force for its nodes in registers, then adds them into an
//                     LDS copy of the tile's nodes in conflict-free phases and
//                     the tile is stored to a per-tile slice. MODE 1 is the
//                     MUSL first map fused with the re-map of momentum.
//   grid_kernel       per node: sums the tile slices covering it, mass cutoff,
//                     mom_nt = mom + f dt, boundary conditions in list order,
//                     node velocities for G2P.
//   g2p_kernel        per particle (cell order): gather from the grid,
//                     FLIP/PIC update, F and volume, then the return map of the
//                     particle's material, written in cell order. MODE 2 is the
//                     MUSL second map.
// Weights are recomputed in registers from the positions; nothing per
// particle x node is stored.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "adjoint_math.hpp"
#include "context.hpp"
#include "device.hpp"

namespace pmpm {

// P2G workgroup. Lane = (cell, y node offset sy of the cell's node span); the
// CX = 64 / SY cells of a wave lie along x. 3D: 4 waves on 4 consecutive z
// layers (tile CX x 1 x 4 cells); ZH waves per layer would each take a share
// of the z node offsets (fewer registers, more redundant weight work; 1 is
// faster on gfx950). 1D/2D: 4 waves, 4 consecutive runs along x (tile
// 4 CX x 1 cells).

// LDS row pitch of the P2G node tile: at least nx, and an odd multiple of cx
// so the SY rows of cx lanes that one flush instruction covers fall into
// disjoint bank groups (64 banks of 4 bytes).
__host__ __device__ constexpr int p2g_pitch(int nx, int cx, int sy) {
  return sy <= 1 ? nx : (((nx + cx - 1) / cx) | 1) * cx;
}

// Tiles whose node block covers one node, at most, per axis: x and z tiles
// are wider than the span overlap (T >= SPAN - 1), y tiles are one cell.
constexpr int kCoverX = 2, kCoverY = 4, kCoverZ = 2;

template <int D, int SF> struct P2GGeom {
  static constexpr int SPAN = Spline<SF>::SPAN;  // node offsets a cell's particles reach, per axis
  static constexpr int NLO = Spline<SF>::NLO;    // the first of them, relative to the cell
  static constexpr int SY = D > 1 ? SPAN : 1;    // lanes per cell
  static constexpr int SZ = D > 2 ? SPAN : 1;    // z node offsets per cell
  static constexpr int ZH = 1;                   // waves sharing a z layer
  static constexpr int SZH = SZ / ZH;            // z node offsets per lane
  static constexpr int WAVES = 4;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int CX = 64 / SY;             // cells per wave
  static constexpr int TX = D == 3 ? CX : CX * WAVES;
  static constexpr int TY = 1;
  static constexpr int TZ = D == 3 ? WAVES / ZH : 1;
  static constexpr int NX = TX + SPAN - 1;       // tile nodes per axis
  static constexpr int NY = D > 1 ? TY + SPAN - 1 : 1;
  static constexpr int NZ = D > 2 ? TZ + SPAN - 1 : 1;
  static constexpr int NXP = p2g_pitch(NX, CX, SY);
  static constexpr int NT = NXP * NY * NZ;       // LDS nodes incl. padding
  static constexpr int NTN = NX * NY * NZ;       // nodes of one partial slice
  static_assert(CX * SY == 64 && SZH * ZH == SZ, "one wave = CX cells x SY lanes");
  static_assert((NX - 1) / TX + 1 <= kCoverX && (NY - 1) / TY + 1 <= kCoverY && (NZ - 1) / TZ + 1 <= kCoverZ,
                "the grid kernel's gather visits at most kCover tiles per axis");
};

__device__ inline int64_t lin_index(const pmpm_grid &g, const int *b, int D) {
  int64_t h = b[0];
  if (D > 1) h += (int64_t)b[1] * g.num_cells[0];
  if (D > 2) h += (int64_t)b[2] * g.num_cells[0] * g.num_cells[1];
  return h;
}

// Per-particle weight tables for the ENTRIES nodes per axis.
template <int D, int SF> struct PW {
  static constexpr int E = Spline<SF>::ENTRIES;
  float t[3];
  int bin[3];
  int first[3];
  float N[3][4];
  float dN[3][4];
};

template <int D, int SF>
__device__ inline void particle_weights(const pmpm_params &prm, const pmpm_grid &g, const float *x, int n, int slot,
                                        PW<D, SF> &w, float *xp) {
  // all coordinates in flight before the first weight
#pragma unroll
  for (int a = 0; a < 3; ++a) xp[a] = a < D ? x[a * n + slot] : 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < D) {
      w.t[a] = cell_coord(xp[a], g.origin[a], g.inv_cell_size);
      w.bin[a] = (int)w.t[a];
      axis_weights<SF>(w.t[a], w.bin[a], g.num_cells[a], prm.node_types_built, g.inv_cell_size, w.first[a], w.N[a],
                       w.dN[a]);
    } else {
      w.t[a] = 0.f; w.bin[a] = 0; w.first[a] = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) { w.N[a][e] = 1.f; w.dN[a][e] = 0.f; }
    }
  }
}

// Node velocity records (pmpm_nodes.vel): NP floats per node, v_nt then dv,
// so a gather entry is one 16-byte load (two for v_nt and dv in 3D) instead
// of one load per component.
template <int D> struct NodeVel {
  static constexpr int NP = D == 3 ? 8 : 2 * D;
};
// The byte offset of a record is a 32-bit value (grids hold fewer than
// PMPM_MAX_GRID_NODES nodes), so a load is the uniform base plus one VGPR.
template <int D, bool WANT_DV>
__device__ inline void load_node_vel(const float *__restrict__ vel, int node, float *v, float *dv) {
  const char *base = reinterpret_cast<const char *>(vel);
  const uint32_t at = (uint32_t)node * (uint32_t)(NodeVel<D>::NP * sizeof(float));
  if constexpr (D == 3) {
    const float4 a = *reinterpret_cast<const float4 *>(base + at);
    v[0] = a.x; v[1] = a.y; v[2] = a.z;
    if constexpr (WANT_DV) {
      const float4 b = *reinterpret_cast<const float4 *>(base + at + 16u);
      dv[0] = a.w; dv[1] = b.x; dv[2] = b.y;
    }
  } else if constexpr (D == 2) {
    const float4 a = *reinterpret_cast<const float4 *>(base + at);
    v[0] = a.x; v[1] = a.y;
    if constexpr (WANT_DV) { dv[0] = a.z; dv[1] = a.w; }
  } else {
    const float2 a = *reinterpret_cast<const float2 *>(base + at);
    v[0] = a.x;
    if constexpr (WANT_DV) dv[0] = a.y;
  }
}
template <int D> __device__ inline void store_node_vel(float *__restrict__ vel, int node, const float *v, const float *dv) {
  if constexpr (D == 3) {
    float4 *p = reinterpret_cast<float4 *>(vel + 8 * (int64_t)node);
    p[0] = make_float4(v[0], v[1], v[2], dv[0]);
    p[1] = make_float4(dv[1], dv[2], 0.f, 0.f);
  } else if constexpr (D == 2) {
    *reinterpret_cast<float4 *>(vel + 4 * (int64_t)node) = make_float4(v[0], v[1], dv[0], dv[1]);
  } else {
    *reinterpret_cast<float2 *>(vel + 2 * (int64_t)node) = make_float2(v[0], dv[0]);
  }
}
// Where a gather reads node velocities from: the global records
// (pmpm_nodes.vel) with the grid's node strides (node (i, j, k) at
// sum id_a stride_a), or a run's node rows staged in LDS (box4 = v_nt and
// dv.x, box2 = dv.y and dv.z per node): window entry (x, sy, sz) of a
// particle at x bin b sits at (b - 1 + x - lo[0]) + sy stride[1] + sz stride[2]
// with sy, sz = 0..3 the entry's node offset + 1 along y and z.
struct NodeSrc {
  const float *vel;
  const float4 *box4;
  const float2 *box2;
  int stride[3];
  int lo[3];
};
__device__ inline NodeSrc node_src(const pmpm_grid &g, const pmpm_nodes &nodes) {
  return NodeSrc{nodes.vel, nullptr, nullptr, {1, g.num_cells[0], g.num_cells[0] * g.num_cells[1]}, {0, 0, 0}};
}
template <int D, bool WANT_DV, bool BOX>
__device__ inline void load_src(const NodeSrc &src, int node, float *v, float *dv) {
  if constexpr (BOX) {
    const float4 a = src.box4[node];
    v[0] = a.x; v[1] = a.y; v[2] = a.z;
    if constexpr (WANT_DV) {
      const float2 b = src.box2[node];
      dv[0] = a.w; dv[1] = b.x; dv[2] = b.y;
    }
  } else {
    load_node_vel<D, WANT_DV>(src.vel, node, v, dv);
  }
}

// Grid gather (1D/2D in the reference's window order, y outer, x inner; 3D in
// the separable order below). The reference skips entries outside the grid
// on any axis (usl_inline.h:281-293); here such entries get zero weights and
// a clamped, valid node index, which leaves every sum unchanged and keeps the
// loop free of branches. Nodes below the mass cutoff carry zero velocities,
// which is equivalent to skipping them.
template <int D, int SF, bool WANT_V, bool WANT_L, bool BOX = false>
__device__ inline void gather_grid(const pmpm_grid &g, const NodeSrc &src, const PW<D, SF> &w, float *vinc,
                                   float *dvinc, float *L) {
  static_assert(!BOX || D == 3, "LDS node boxes are 3D");
  constexpr int E = Spline<SF>::ENTRIES;
  constexpr int EY = D > 1 ? E : 1;
  constexpr int EZ = D > 2 ? E : 1;
  float N[3][4], dN[3][4];
  int off[3][4];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int stride = src.stride[a];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (a < D && e < E) {
        const int i = w.bin[a] + w.first[a] + e;
        const bool ok = i >= 0 && i < g.num_cells[a];
        if (BOX) off[a][e] = (a == 0 ? i - src.lo[0] : w.first[a] + 1 + e) * stride;
        else off[a][e] = min(max(i, 0), g.num_cells[a] - 1) * stride;
        N[a][e] = ok ? w.N[a][e] : 0.f;
        dN[a][e] = ok ? w.dN[a][e] : 0.f;
      } else {
        off[a][e] = 0;
        N[a][e] = 1.f;
        dN[a][e] = 0.f;
      }
    }
  }
  if constexpr (D == 3) {
    // Separable form: the weights are products of per-axis factors, so the
    // sums contract x first (a row's x nodes are adjacent records), then y,
    // then z: per row S = sum_x Nx v, Sd = sum_x dNx v; per z layer
    // U0 = sum_y Ny Sd, U1 = sum_y dNy S, U2 = sum_y Ny S; then
    // L[x][.] += Nz U0, L[y][.] += Nz U1, L[z][.] += dNz U2, v += Nz U2.
    // The next row's records are loaded while the current row is summed.
    constexpr int NR = EY * EZ;
    float rv[2][E][3], rd[2][E][3];
    auto load_row = [&](int r, int buf) {
      const int row = off[1][r % EY] + off[2][r / EY];
#pragma unroll
      for (int ex = 0; ex < E; ++ex) load_src<D, WANT_V, BOX>(src, row + off[0][ex], rv[buf][ex], rd[buf][ex]);
    };
    load_row(0, 0);
    float U0[3] = {0.f, 0.f, 0.f}, U1[3] = {0.f, 0.f, 0.f}, U2[3] = {0.f, 0.f, 0.f}, Ud[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int ey = r % EY, ez = r / EY, cb = r & 1;
      if (r + 1 < NR) load_row(r + 1, cb ^ 1);
      float Sx[3] = {0.f, 0.f, 0.f}, Sdx[3] = {0.f, 0.f, 0.f}, Sxd[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int ex = 0; ex < E; ++ex)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          Sx[b] += N[0][ex] * rv[cb][ex][b];
          if (WANT_L) Sdx[b] += dN[0][ex] * rv[cb][ex][b];
          if (WANT_V) Sxd[b] += N[0][ex] * rd[cb][ex][b];
        }
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        U2[b] += N[1][ey] * Sx[b];
        if (WANT_L) {
          U0[b] += N[1][ey] * Sdx[b];
          U1[b] += dN[1][ey] * Sx[b];
        }
        if (WANT_V) Ud[b] += N[1][ey] * Sxd[b];
      }
      if (ey == EY - 1) {  // z layer complete
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          if (WANT_L) {
            L[b] += N[2][ez] * U0[b];
            L[3 + b] += N[2][ez] * U1[b];
            L[6 + b] += dN[2][ez] * U2[b];
          }
          if (WANT_V) {
            vinc[b] += N[2][ez] * U2[b];
            dvinc[b] += N[2][ez] * Ud[b];
          }
          U0[b] = U1[b] = U2[b] = Ud[b] = 0.f;
        }
      }
    }
    return;
  }
#pragma unroll
  for (int ey = 0; ey < EY; ++ey) {
#pragma unroll
    for (int ex = 0; ex < E; ++ex) {
      const float nxy = N[0][ex] * N[1][ey];
      const float dxy = dN[0][ex] * N[1][ey];
      const float dyx = dN[1][ey] * N[0][ex];
#pragma unroll
      for (int ez = 0; ez < EZ; ++ez) {
        const int node = off[0][ex] + off[1][ey] + off[2][ez];
        const float wgt = nxy * N[2][ez];
        float gr[3];
        gr[0] = dxy * N[2][ez];
        gr[1] = dyx * N[2][ez];
        gr[2] = (dN[2][ez] * N[0][ex]) * N[1][ey];
        float vnt[3], dv[3];
        load_node_vel<D, WANT_V>(src.vel, node, vnt, dv);
        if (WANT_V) {
#pragma unroll
          for (int a = 0; a < D; ++a) {
            dvinc[a] += wgt * dv[a];
            vinc[a] += wgt * vnt[a];
          }
        }
        if (WANT_L) {
#pragma unroll
          for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) L[3 * a + b] += gr[a] * vnt[b];
        }
      }
    }
  }
}

// Weight and gradient of one window entry with explicit node types (used when
// the node is reached through the linear-index wrap, and for materialised psi).
template <int D, int SF>
__device__ inline void entry_weight(const PW<D, SF> &w, const int *off, const int *types, float inv_h, float &wgt,
                                    float *gr) {
  float N[3] = {1.f, 1.f, 1.f}, dN[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float frac = __fsub_rn(w.t[a], (float)w.bin[a]);
    spline_eval<SF>(__fsub_rn(frac, (float)off[a]), types[a], inv_h, N[a], dN[a]);
  }
  wgt = (N[0] * N[1]) * N[2];
  gr[0] = (dN[0] * N[1]) * N[2];
  gr[1] = (dN[1] * N[0]) * N[2];
  gr[2] = (dN[2] * N[0]) * N[1];
}

template <int D>
__device__ inline void node_ids(const pmpm_grid &g, int64_t lin, int *ids) {
  ids[0] = (int)(lin % g.num_cells[0]);
  ids[1] = (D > 1) ? (int)((lin / g.num_cells[0]) % g.num_cells[1]) : 0;
  ids[2] = (D > 2) ? (int)(lin / ((int64_t)g.num_cells[0] * g.num_cells[1])) : 0;
}

// Slow path of the scatter: every entry goes straight to global memory,
// through the linear index (wrap-around reproduced when wrap_quirk is set).
// Used for particles whose bin is not their cell's coordinates and for the
// entries of boundary particles that leave the grid on some axis.
template <int D, int SF, int MODE>
__device__ inline void scatter_entry_global(const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes,
                                            int nn, const PW<D, SF> &w, const int *e, float m, const float *mv,
                                            const float A[3][3], const float *fext) {
  int nb[3] = {0, 0, 0}, off[3] = {0, 0, 0};
  bool inside = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    off[a] = w.first[a] + e[a];
    nb[a] = w.bin[a] + off[a];
    if (nb[a] < 0 || nb[a] >= g.num_cells[a]) inside = false;
  }
  const int64_t lin = lin_index(g, nb, D);
  if (lin < 0 || lin >= g.num_cells_total) return;
  if (!inside && !prm.wrap_quirk) return;
  int ids[3];
  node_ids<D>(g, lin, ids);
  int types[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) types[a] = cubic_node_type(ids[a], g.num_cells[a], prm.node_types_built);
  float wgt, gr[3];
  entry_weight<D, SF>(w, off, types, g.inv_cell_size, wgt, gr);
  const int node = (int)lin;
  if (MODE == 0) {
    if (wgt != 0.f) unsafeAtomicAdd(&nodes.mass[node], wgt * m);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float pm = wgt * mv[a];
      if (pm != 0.f) unsafeAtomicAdd(&nodes.mom[a * nn + node], pm);
      float f = wgt * fext[a];
#pragma unroll
      for (int b = 0; b < D; ++b) f += A[a][b] * gr[b];
      if (f != 0.f) unsafeAtomicAdd(&nodes.mom_nt[a * nn + node], f);
    }
  } else {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float pm = wgt * mv[a];
      if (pm != 0.f) unsafeAtomicAdd(&nodes.mom_nt[a * nn + node], pm);
    }
  }
}

// Weight of one particle at span offset S (node NLO + S relative to its cell)
// along one axis, frac = t - bin in [0, 1): the same value as the window
// weight (axis_weights), and 0 at an offset the window does not reach
// (quadratic). `type` is the node's cubic type (shapefunction.cpp:108-122).
// For the cubic spline (types 1-3; the caller handles unbuilt types),
// r = frac - (S - 1) lies in piece 3 - S of the table (cubic.cuh:28-197)
// except when frac = 1 - 2^-24, where rounding puts r on the next knot: at
// S = 0 (r = 2) the weight is 0 as in cubic_eval; at S = 3 (r = -1) both
// pieces meet there (the spline is C2), so the value agrees to rounding.
template <int SF, int S>
__device__ inline void span_eval(float frac, int type, float inv_h, float &N, float &dN) {
  if (SF != PMPM_SF_CUBIC) {
    spline_eval<SF>(__fsub_rn(frac, (float)(Spline<SF>::NLO + S)), 0, inv_h, N, dN);
    return;
  }
  cubic_span<S>(frac, type, inv_h, N, dN);
}

// span_eval for the NS consecutive offsets S0 .. S0 + NS - 1.
template <int SF, int S0, int NS>
__device__ inline void span_evals(float frac, const int *type, float inv_h, float *N, float *dN) {
  if constexpr (NS > 0) {
    span_eval<SF, S0>(frac, type[0], inv_h, N[0], dN[0]);
    span_evals<SF, S0 + 1, NS - 1>(frac, type + 1, inv_h, N + 1, dN + 1);
  }
}

// Per-particle inputs of a P2G deposit.
//   MODE 0: mass, velocity, -V sigma.
//   MODE 1 (the re-map's wrap-around deposits): mass and v', read from out.v
//   in cell order after the fused first map wrote it.
//   MODE 2 (adjoint scatter): v = adjoint of the PIC velocity, dv = adjoint of
//   the FLIP increment, Lb = adjoint of the velocity gradient (row-major 3x3),
//   read from the record array `rec` in cell order.
struct P2GRec {
  float x[3], v[3], m;
  Sym A;  // -V sigma (MODE 0)
  float f[3];  // external force (MODE 0: the planar-domain walls, else 0)
  float dv[3], Lb[9];
};
template <int D, int MODE>
__device__ inline void p2g_load(const pmpm_params &prm, const pmpm_grid &g, bool planar, const pmpm_particles &in,
                                const pmpm_particles &out, const float *rec, int n, int slot, int j, P2GRec &r) {
  r.f[0] = r.f[1] = r.f[2] = 0.f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    r.x[a] = in.x[a * n + slot];
    r.v[a] = MODE == 0 ? in.v[a * n + slot] : (MODE == 1 ? out.v[a * n + j] : rec[a * n + j]);
  }
  if (MODE == 2) {
    r.m = 1.f;
    r.A = sym_zero();
#pragma unroll
    for (int a = 0; a < D; ++a) r.dv[a] = rec[(3 + a) * n + j];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) r.Lb[3 * a + b] = rec[(6 + 3 * a + b) * n + j];
    return;
  }
  r.m = in.mass[slot];
  if (MODE == 0) {
    const Sym s = load_sym(in.stress, n, slot);
    const float vol = in.volume[slot];
    const float mvol = -vol;
    r.A = Sym{mvol * s.xx, mvol * s.yy, mvol * s.zz, mvol * s.xy, mvol * s.xz, mvol * s.yz};
    if (planar) planar_force<D>(prm, g, r.x, r.v, vol, r.m, r.f);
  } else {
    r.A = sym_zero();
  }
}

// P2G of a particle whose recomputed bin is not the cell it is listed in
// (only possible when the partition was built from other positions): this
// lane's y slice of its window goes straight to the node arrays.
template <int D, int SF, int MODE>
__device__ void p2g_stray(const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes, const float *x, int n,
                          int slot, int sy, float m, const float *v, const Sym &A, const float *fext) {
  constexpr int E = Spline<SF>::ENTRIES;
  constexpr int EZ = D > 2 ? E : 1;
  if (sy >= (D > 1 ? E : 1)) return;
  PW<D, SF> w;
  float xp[3];
  particle_weights<D, SF>(prm, g, x, n, slot, w, xp);
  float mv[3] = {0.f, 0.f, 0.f}, A3[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < D) mv[a] = m * v[a];
#pragma unroll
    for (int b = 0; b < 3; ++b) A3[a][b] = (a < D && b < D) ? sym_at(A, a, b) : 0.f;
  }
#pragma unroll 1
  for (int ex = 0; ex < E; ++ex)
#pragma unroll 1
    for (int ez = 0; ez < EZ; ++ez) {
      const int e3[3] = {ex, sy, ez};
      scatter_entry_global<D, SF, MODE>(prm, g, nodes, g.num_cells_total, w, e3, m, mv, A3, fext);
    }
}

// Per-particle P2G inputs staged in LDS, one float (or int bits) per field,
// field-major within a wave's region:
//   frac[D]  position in the cell (t - bin, as the weights use it); MODE 1
//            keeps the position itself (its first map moves it)
//   lin      linear index of the particle's own bin (-1 outside the grid)
//   m        mass (MODE 0/1)
//   v[D]     velocity (MODE 0 and 1), adjoint of the PIC velocity (MODE 2)
//   A        -V sigma, the D(D+1)/2 components used in D dimensions (MODE 0)
//   dv[D], Lb[D*D]  adjoints of the FLIP increment and of L (MODE 2)
// Fields per node of a P2G pass: MODE 0 mass, momentum, force; MODE 1 the
// re-mapped momentum; MODE 2 the adjoints of v_nt and dv.
template <int D, int MODE> constexpr int p2g_nv() { return MODE == 0 ? 1 + 2 * D : (MODE == 2 ? 2 * D : D); }

template <int D, int MODE> struct P2GStage {
  static constexpr int NA = D * (D + 1) / 2;
  static constexpr int F_LIN = D, F_M = D + 1, F_V = D + 2, F_X = 2 * D + 2;  // F_X: A (MODE 0) or dv (MODE 2)
  static constexpr int NF = MODE == 0 ? F_X + NA : (MODE == 2 ? F_X + D + D * D : F_X);
};
// Particles of one wave's run that fit its LDS region at once: the whole run
// of 16 cells at the usual 27 particles per cell (longer runs take more
// passes). MODE 0 leaves room for one workgroup per CU (the accumulators
// allow 2 waves per SIMD at most), MODE 1 for two.
template <int D, int SF, int MODE> constexpr int p2g_stage_cap() {
  constexpr int NV = p2g_nv<D, MODE>();
  constexpr int budget = MODE == 1 ? 18944 : 40448;  // floats per workgroup
  constexpr int room = (budget - NV * P2GGeom<D, SF>::NT) / (P2GGeom<D, SF>::WAVES * P2GStage<D, MODE>::NF);
  constexpr int cap = (room < 512 ? room : 512) / 64 * 64;
  static_assert(cap >= 64, "LDS budget too small for the staging area");
  return cap;
}

// LDS writes of a wave visible to all its lanes (and its earlier reads done)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ inline int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// v_nt and dv gathered over one particle's own window: the rare path of the
// fused MUSL map, kept as a compact loop (entries outside the grid skipped, as
// the reference does) so it costs the fused pass no registers.
template <int D, int SF>
__device__ inline void gather_one(const pmpm_params &prm, const pmpm_grid &g, const float *x, int n, int slot,
                                  const pmpm_nodes &nodes, float *vinc, float *dvinc) {
  constexpr int SPAN = Spline<SF>::SPAN, NLO = Spline<SF>::NLO;
  constexpr int NENT = D == 1 ? SPAN : (D == 2 ? SPAN * SPAN : SPAN * SPAN * SPAN);
  const float inv_h = g.inv_cell_size;
  float fr[3] = {0.f, 0.f, 0.f};
  int b[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float t = cell_coord(x[a * n + slot], g.origin[a], inv_h);
    b[a] = (int)t;
    fr[a] = __fsub_rn(t, (float)b[a]);
  }
#pragma unroll 1
  for (int e = 0; e < NENT; ++e) {
    int id[3] = {0, 0, 0};
    float w = 1.f;
    bool in = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int o = NLO + (a == 0 ? e % SPAN : (a == 1 ? (e / SPAN) % SPAN : e / (SPAN * SPAN)));
      id[a] = b[a] + o;
      in = in && id[a] >= 0 && id[a] < g.num_cells[a];
      float N, dN;
      spline_eval<SF>(__fsub_rn(fr[a], (float)o), cubic_node_type(id[a], g.num_cells[a], prm.node_types_built),
                      inv_h, N, dN);
      w *= N;
    }
    if (!in || w == 0.f) continue;
    float v[3], dv[3];
    load_node_vel<D, true>(nodes.vel, (int)lin_index(g, id, D), v, dv);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      vinc[a] += w * v[a];
      dvinc[a] += w * dv[a];
    }
  }
}

// Workgroups after the tiles of a MODE 1 launch: they run the first map of
// the particles outside the grid, which sort last and belong to no cell.
constexpr int kMuslTailBlocks = 4;
template <int D, int SF>
__device__ void musl_first_map_tail(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in,
                                    const pmpm_particles &out, const pmpm_spatial &part, const pmpm_nodes &nodes,
                                    int tail_block) {
  const int n = in.n;
  const uint32_t nc = (uint32_t)g.num_cells_total;
  if (n == 0 || part.hash_sorted[n - 1] < nc) return;  // none outside
  int lo = 0, hi = n - 1;  // first j with hash_sorted[j] == nc
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (part.hash_sorted[mid] >= nc) hi = mid;
    else lo = mid + 1;
  }
  const float am = (float)(1.0 - (double)prm.alpha);
  for (int j = lo + tail_block * (int)blockDim.x + (int)threadIdx.x; j < n; j += kMuslTailBlocks * (int)blockDim.x) {
    const int src = part.sorted[j];
    PW<D, SF> w;
    float xp[3];
    particle_weights<D, SF>(prm, g, in.x, n, src, w, xp);
    float vinc[3] = {0.f, 0.f, 0.f}, dvinc[3] = {0.f, 0.f, 0.f}, L[9];
    gather_grid<D, SF, true, false>(g, node_src(g, nodes), w, vinc, dvinc, L);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      out.v[a * n + j] = am * vinc[a] + prm.alpha * (in.v[a * n + src] + dvinc[a]);
      out.x[a * n + j] = xp[a] + prm.dt * vinc[a];
    }
  }
}

// P2G of one tile of cells.
//   MODE 0: mass, momentum and force (force accumulates in mom_nt).
//   MODE 1: MUSL first map and re-map in one pass: each particle's v', x'
//           from the node velocity records (written to slot j of `out`), then
//           m v' deposited with the same weights (the reference's
//           KERNEL_MUSL_G2P_DOUBLE_MAPPING and KERNEL_MUSL_P2G_DOUBLE_MAPPING,
//           musl_kernels.cuh:33-60). Particles outside the grid are mapped by
//           the kMuslTailBlocks workgroups after the tiles.
//   MODE 2: adjoint scatter of the G2P record adjoints.
// Lane (cell c, sy) walks the particles of c in sorted order and sums in
// registers their contributions to its SPAN x SZ nodes c + NLO + (sx, sy, sz).
// The sums then enter an LDS node tile in SPAN x SZ phases: within one phase
// the lanes of the workgroup address distinct nodes, so a plain read-add-write
// is race free (LDS float atomics are over an order of magnitude slower on
// gfx950). The tile goes with plain stores to this workgroup's slice of
// `partials` (flags[tile] marks the slice live) and the grid kernel sums the
// slices covering each node.
// A wave's cells are consecutive in the linear cell order, so their particles
// form one run of the sorted order. The wave loads the run into LDS first,
// lane = particle, with coalesced loads (fraction, own bin, and the products
// each lane needs), then the register loop reads only LDS: every particle line
// is fetched from HBM once instead of being re-fetched as the cells' lanes
// step through it together.
// Node offsets outside the grid: the reference bound-checks only the linear
// node index (usl_inline.h:122-126, shapefunction.cpp:96-98), so such a
// deposit lands on the wrapped node, weighted with that node's cubic types.
// Those sums go to the node arrays with one global atomic per cell, node
// offset and field (wrap_quirk), or are dropped.
// Which P2G passes stage their particles in LDS: the MUSL map (8 floats per
// particle, two workgroups per CU). The full P2G and the adjoint scatter
// (14 and 20 floats) would fit one workgroup per CU only, and measured slower
// that way than reading through L2 at two waves per SIMD.
template <int MODE> constexpr bool p2g_staged() { return MODE == 1; }
template <int D, int SF, int MODE> constexpr int p2g_lds_floats() {
  constexpr int NV = p2g_nv<D, MODE>();
  return NV * P2GGeom<D, SF>::NT +
         (p2g_staged<MODE>() ? P2GGeom<D, SF>::WAVES * P2GStage<D, MODE>::NF * p2g_stage_cap<D, SF, MODE>() : 1);
}
// Waves per SIMD the LDS allows (4 waves per workgroup, 160 KB per CU); at 1
// the register budget is the whole file.
template <int D, int SF, int MODE> constexpr int p2g_waves_per_simd() {
  return p2g_lds_floats<D, SF, MODE>() * 4 * 2 <= 163840 ? 2 : 1;
}

template <int D, int SF, int MODE>
__global__ __launch_bounds__((P2GGeom<D, SF>::THREADS)) __attribute__((amdgpu_waves_per_eu(p2g_waves_per_simd<D, SF, MODE>()))) void p2g_kernel(pmpm_params prm, pmpm_grid g,
                                                                      pmpm_particles in, pmpm_particles out,
                                                                      pmpm_spatial part, pmpm_nodes nodes,
                                                                      int tiles_x, int tiles_y,
                                                                      float *__restrict__ partials,
                                                                      int32_t *__restrict__ flags,
                                                                      const float *__restrict__ rec) {
  using G = P2GGeom<D, SF>;
  constexpr int SPAN = G::SPAN, NLO = G::NLO, SZ = G::SZ, SZH = G::SZH, NE = SPAN * SZH;
  constexpr int NV = p2g_nv<D, MODE>();
  constexpr int NXP = G::NXP, NY = G::NY, NT = G::NT;
  using ST = P2GStage<D, MODE>;
  constexpr bool kStaged = p2g_staged<MODE>();
  constexpr int NF = ST::NF, CW = p2g_stage_cap<D, SF, MODE>();
  __shared__ float tile[NV * NT];
  __shared__ float stage[kStaged ? G::WAVES * NF * CW : 1];

  const int n = in.n;
  const int nn = g.num_cells_total;
  const float inv_h = g.inv_cell_size;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int sy = D > 1 ? lane / G::CX : 0;
  const int lx = (D == 3 ? 0 : wave * G::CX) + lane % G::CX;  // tile-local cell
  const int lz = D == 3 ? wave / G::ZH : 0;
  const int zh = D == 3 ? wave % G::ZH : 0;  // which half of the z node offsets
  const int bx = blockIdx.x % tiles_x;
  const int by = (blockIdx.x / tiles_x) % tiles_y;
  const int bz = blockIdx.x / (tiles_x * tiles_y);
  const int c[3] = {bx * G::TX + lx, D > 1 ? by : 0, D > 2 ? bz * G::TZ + lz : 0};

  if constexpr (MODE == 1) {
    if ((int)blockIdx.x >= (int)gridDim.x - kMuslTailBlocks) {
      musl_first_map_tail<D, SF>(prm, g, in, out, part, nodes, (int)blockIdx.x - ((int)gridDim.x - kMuslTailBlocks));
      return;
    }
  }
  for (int i = t; i < NV * NT; i += G::THREADS) tile[i] = 0.f;

  bool in_grid = c[0] < g.num_cells[0];
  if (D > 1) in_grid = in_grid && c[1] < g.num_cells[1];
  if (D > 2) in_grid = in_grid && c[2] < g.num_cells[2];
  int cs = 0, ce = 0, cell = -2;
  if (in_grid) {
    cell = (int)lin_index(g, c, D);
    ce = part.cell_end[cell];
    cs = part.cell_start[cell];
  }
  // cubic weights without node types are zero (cubic.cuh: type 0): nothing
  // to deposit
  // (MODE 1 still moves the particles: its first map gathers zeros)
  const bool none = SF == PMPM_SF_CUBIC && !prm.node_types_built;
  const bool planar = MODE == 0 && has_planar(prm);  // wall forces on the particles
  if (!__syncthreads_or(ce > cs) || (none && MODE != 1)) {
    if (t == 0) flags[blockIdx.x] = 0;
    return;
  }

  // which of this lane's nodes lie inside the grid
  bool x_in[SPAN], z_in[SPAN];
  const int yn = c[1] + NLO + sy;
  const bool y_in = D < 2 || (yn >= 0 && yn < g.num_cells[1]);
#pragma unroll
  for (int s = 0; s < SPAN; ++s) {
    const int xn = c[0] + NLO + s, zn = c[2] + NLO + s;
    x_in[s] = xn >= 0 && xn < g.num_cells[0];
    z_in[s] = D < 3 || (zn >= 0 && zn < g.num_cells[2]);
  }

  float acc[NE][NV];
#pragma unroll
  for (int e = 0; e < NE; ++e)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[e][k] = 0.f;

  // cubic node types of this lane's span nodes (own particles share them)
  int tx[SPAN], tz[SZH];
#pragma unroll
  for (int s = 0; s < SPAN; ++s) tx[s] = cubic_node_type(c[0] + NLO + s, g.num_cells[0], prm.node_types_built);
#pragma unroll
  for (int s = 0; s < SZH; ++s)
    tz[s] = D > 2 ? cubic_node_type(c[2] + NLO + zh * SZH + s, g.num_cells[2], prm.node_types_built) : 0;
  // this lane's one y offset: its cubic piece (3 - sy) and the node's type
  // fix the coefficients for every particle of the cell
  const float ry0 = (float)(NLO + sy);
  float yc[4] = {0.f, 0.f, 0.f, 0.f};
  if (SF == PMPM_SF_CUBIC && D > 1) {
    const int ty = cubic_node_type(c[1] + NLO + sy, g.num_cells[1], prm.node_types_built);
#pragma unroll
    for (int k = 0; k < 4; ++k) yc[k] = kCubic[ty][3 - sy][k];
  }

  // MODE 1: the velocity records of this lane's SPAN x SZH nodes, loaded once
  // for all the cell's particles (zero outside the grid, as the gather skips
  // such entries)
  constexpr int NVN = MODE == 1 ? 2 * D : 1;
  float vnode[SPAN][SZH][NVN];
  if constexpr (MODE == 1) {
#pragma unroll
    for (int sx = 0; sx < SPAN; ++sx)
#pragma unroll
      for (int sz = 0; sz < SZH; ++sz) {
        float v[3] = {0.f, 0.f, 0.f}, dv[3] = {0.f, 0.f, 0.f};
        if (in_grid && x_in[sx] && y_in && z_in[zh * SZH + sz]) {
          const int nb[3] = {c[0] + NLO + sx, c[1] + NLO + sy, c[2] + NLO + zh * SZH + sz};
          load_node_vel<D, true>(nodes.vel, (int)lin_index(g, nb, D), v, dv);
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
          vnode[sx][sz][a] = v[a];
          vnode[sx][sz][D + a] = dv[a];
        }
      }
  }
  const float am = (float)(1.0 - (double)prm.alpha);

  // a particle's weights for this lane's nodes: Nx, Nz per span offset and
  // the lane's one Ny (with derivatives)
  struct Wts {
    float Nx[SPAN], dNx[SPAN], Nz[SZH], dNz[SZH], Ny, dNy;
  };
  auto weights = [&](const float *frac, Wts &W) {
    float *Nx = W.Nx, *dNx = W.dNx, *Nz = W.Nz, *dNz = W.dNz;
    float Ny = 1.f, dNy = 0.f;
    span_evals<SF, 0, SPAN>(frac[0], tx, inv_h, Nx, dNx);
    if constexpr (D > 2) {
      if constexpr (G::ZH == 1) {
        span_evals<SF, 0, SZH>(frac[2], tz, inv_h, Nz, dNz);
      } else {  // this wave's share of the z offsets (zh is uniform per wave)
        if (zh == 0) span_evals<SF, 0, SZH>(frac[2], tz, inv_h, Nz, dNz);
        else span_evals<SF, SZH, SZH>(frac[2], tz, inv_h, Nz, dNz);
      }
    } else {
      Nz[0] = 1.f;
      dNz[0] = 0.f;
    }
    if constexpr (D > 1) {
      const float r = __fsub_rn(frac[1], ry0);
      if (SF == PMPM_SF_CUBIC) {  // the Horner form of cubic_eval on this lane's piece
        Ny = ((yc[3] * r + yc[2]) * r + yc[1]) * r + yc[0];
        dNy = inv_h * ((3.f * yc[3] * r + 2.f * yc[2]) * r + yc[1]);
        if (sy == 0 && !(r < 2.f)) Ny = dNy = 0.f;  // frac = 1 - 2^-24 (see span_eval)
      } else {
        spline_eval<SF>(r, 0, inv_h, Ny, dNy);
      }
    }
    W.Ny = Ny;
    W.dNy = dNy;
  };
  // one particle's contribution to this lane's register sums
  auto deposit = [&](const P2GRec &cur, const Wts &W) {
    const float *Nx = W.Nx, *dNx = W.dNx, *Nz = W.Nz, *dNz = W.dNz;
    const float Ny = W.Ny, dNy = W.dNy;
    const float m = cur.m;
#pragma unroll
    for (int sx = 0; sx < SPAN; ++sx) {
      const float nxy = Nx[sx] * Ny;
      const float nxym = nxy * m;
      float mvxy[3], p[3], q[3];
      // f_a = sum_b A_ab g_b with g = (dNx Ny Nz, Nx dNy Nz, Nx Ny dNz):
      // f_a = p_a Nz + q_a dNz
      const float dxy = dNx[sx] * Ny, dyx = dNy * Nx[sx];
      float wd[3];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        mvxy[a] = nxym * cur.v[a];
        if (MODE == 2) {
          // adjoint: sum_c Lb_ca g_c = p_a Nz + q_a dNz
          wd[a] = nxy * cur.dv[a];
          p[a] = cur.Lb[a] * dxy;
          if (D > 1) p[a] += cur.Lb[3 + a] * dyx;
          q[a] = D > 2 ? cur.Lb[6 + a] * nxy : 0.f;
        } else {
          p[a] = sym_at(cur.A, a, 0) * dxy;
          if (D > 1) p[a] += sym_at(cur.A, a, 1) * dyx;
          q[a] = D > 2 ? sym_at(cur.A, a, 2) * nxy : 0.f;
          if (MODE == 0 && planar) p[a] += cur.f[a] * nxy;  // w f_ext (usl_inline.h:152-155)
        }
      }
#pragma unroll
      for (int sz = 0; sz < SZH; ++sz) {
        float *r = acc[sx * SZH + sz];
        if (MODE == 0) {
          r[0] += nxym * Nz[sz];
#pragma unroll
          for (int a = 0; a < D; ++a) {
            r[1 + a] += mvxy[a] * Nz[sz];
            r[1 + D + a] += p[a] * Nz[sz];
            if (D > 2) r[1 + D + a] += q[a] * dNz[sz];
          }
        } else if (MODE == 2) {
#pragma unroll
          for (int a = 0; a < D; ++a) {
            r[a] += (mvxy[a] + p[a]) * Nz[sz];
            if (D > 2) r[a] += q[a] * dNz[sz];
            r[D + a] += wd[a] * Nz[sz];
          }
        } else {
#pragma unroll
          for (int a = 0; a < D; ++a) r[a] += mvxy[a] * Nz[sz];
        }
      }
    }
  };

  if constexpr (!kStaged) {
    // the next particle's inputs are loaded while the current one is summed
    int j = cs;
    int slot = j < ce ? part.sorted[j] : 0;
    int slot_next = j + 1 < ce ? part.sorted[j + 1] : 0;
    P2GRec cur;
    if (j < ce) p2g_load<D, MODE>(prm, g, planar, in, out, rec, n, slot, j, cur);
    for (; j < ce; ++j) {
      const int jn = j + 1;
      P2GRec nxt = cur;
      if (jn < ce) p2g_load<D, MODE>(prm, g, planar, in, out, rec, n, slot_next, jn, nxt);
      const int slot_cur = slot;
      slot = slot_next;
      slot_next = jn + 1 < ce ? part.sorted[jn + 1] : 0;
      float frac[3] = {0.f, 0.f, 0.f};
      bool own = true;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const float ta = cell_coord(cur.x[a], g.origin[a], inv_h);
        const int ba = (int)ta;
        own = own && ba == c[a] && ta >= 0.f;  // (t < 0: table form, stray path)
        frac[a] = __fsub_rn(ta, (float)ba);
      }
      if (own) {
        Wts W;
        weights(frac, W);
        deposit(cur, W);
      } else if (MODE != 2 && zh == 0)  // (MODE 2 runs on a partition of the same positions)
        p2g_stray<D, SF, MODE>(prm, g, nodes, in.x, n, slot_cur, sy, cur.m, cur.v, cur.A, cur.f);
      cur = nxt;
    }
  } else {
    // this wave's run [run_lo, run_hi) of the sorted order
    const int run_lo = wave_min(ce > cs ? cs : 0x7fffffff);
    const int run_hi = wave_max(ce > cs ? ce : 0);
    float *st = stage + wave * (NF * CW);
    for (int w0 = run_lo; w0 < run_hi; w0 += CW) {  // wave-uniform
      const int wn = min(CW, run_hi - w0);
      // straight-line staging: every slot, then every field load, so the loads
      // of a lane's CW / 64 particles are in flight together (rounds past the
      // window read particle w0 and fill LDS slots nobody reads)
      constexpr int KR = CW / 64;
      int slots[KR], js[KR];
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int q = lane + 64 * r;
        js[r] = q < wn ? w0 + q : w0;
        slots[r] = part.sorted[js[r]];
      }
#pragma unroll
      for (int r = 0; r < KR; ++r) {
        const int q = lane + 64 * r;
        const int jj = js[r], slot = slots[r];
        int b[3] = {0, 0, 0};
        bool ing = true;
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const float ta = cell_coord(in.x[a * n + slot], g.origin[a], inv_h);
          b[a] = (int)ta;
          ing = ing && ta >= 0.f && b[a] < g.num_cells[a];  // (t < 0: stray path)
          // (MODE 1 keeps the position itself: the first map moves it)
          st[a * CW + q] = MODE == 1 ? in.x[a * n + slot] : __fsub_rn(ta, (float)b[a]);
        }
        st[ST::F_LIN * CW + q] = __int_as_float(ing ? (int)lin_index(g, b, D) : -1);
        if (MODE != 2) st[ST::F_M * CW + q] = in.mass[slot];
#pragma unroll
        for (int a = 0; a < D; ++a)
          st[(ST::F_V + a) * CW + q] = MODE != 2 ? in.v[a * n + slot] : rec[a * n + jj];
        if constexpr (MODE == 0) {
          const float mvol = -in.volume[slot];
          const Sym sg = load_sym(in.stress, n, slot);
          const float A6[6] = {mvol * sg.xx, mvol * sg.yy, mvol * sg.zz, mvol * sg.xy, mvol * sg.xz, mvol * sg.yz};
          // D components in sym order: 1D xx; 2D xx yy xy; 3D all six
          constexpr int pick[3][6] = {{0}, {0, 1, 3}, {0, 1, 2, 3, 4, 5}};
#pragma unroll
          for (int k = 0; k < ST::NA; ++k) st[(ST::F_X + k) * CW + q] = A6[pick[D - 1][k]];
        } else if constexpr (MODE == 2) {
#pragma unroll
          for (int a = 0; a < D; ++a) st[(ST::F_X + a) * CW + q] = rec[(3 + a) * n + jj];
#pragma unroll
          for (int a = 0; a < D; ++a)
#pragma unroll
            for (int bb = 0; bb < D; ++bb) st[(ST::F_X + D + a * D + bb) * CW + q] = rec[(6 + 3 * a + bb) * n + jj];
        }
      }
      wave_lds_sync();

      // this lane's particles in the staged window
      const int j0 = max(cs, w0), j1 = min(ce, w0 + wn);
      auto read = [&](int q, P2GRec &r, float *fr, int &lin) {
#pragma unroll
        for (int a = 0; a < D; ++a) fr[a] = st[a * CW + q];
        lin = __float_as_int(st[ST::F_LIN * CW + q]);
        r.m = MODE == 2 ? 1.f : st[ST::F_M * CW + q];
        r.f[0] = r.f[1] = r.f[2] = 0.f;  // (staged passes carry no force)
#pragma unroll
        for (int a = 0; a < D; ++a) r.v[a] = st[(ST::F_V + a) * CW + q];
        r.A = sym_zero();
        if constexpr (MODE == 0) {
          if (D == 1) r.A.xx = st[ST::F_X * CW + q];
          if (D == 2) {
            r.A.xx = st[ST::F_X * CW + q]; r.A.yy = st[(ST::F_X + 1) * CW + q]; r.A.xy = st[(ST::F_X + 2) * CW + q];
          }
          if (D == 3) {
            r.A.xx = st[ST::F_X * CW + q]; r.A.yy = st[(ST::F_X + 1) * CW + q]; r.A.zz = st[(ST::F_X + 2) * CW + q];
            r.A.xy = st[(ST::F_X + 3) * CW + q]; r.A.xz = st[(ST::F_X + 4) * CW + q]; r.A.yz = st[(ST::F_X + 5) * CW + q];
          }
        } else if constexpr (MODE == 2) {
#pragma unroll
          for (int a = 0; a < D; ++a) r.dv[a] = st[(ST::F_X + a) * CW + q];
#pragma unroll
          for (int a = 0; a < D; ++a)
#pragma unroll
            for (int bb = 0; bb < D; ++bb) r.Lb[3 * a + bb] = st[(ST::F_X + D + a * D + bb) * CW + q];
        }
      };
      // the next particle's inputs are read while the current one is summed
      // (MODE 1 reads each in turn: its node records take the registers)
      constexpr bool kPrefetch = MODE != 1;
      P2GRec cur, nxt;
      float fr[3] = {0.f, 0.f, 0.f}, fr_n[3] = {0.f, 0.f, 0.f};
      int lin = -1, lin_n = -1;
      if (kPrefetch && j0 < j1) read(j0 - w0, cur, fr, lin);
      for (int j = j0; j < j1; ++j) {
        if (!kPrefetch) read(j - w0, cur, fr, lin);
        else if (j + 1 < j1) read(j + 1 - w0, nxt, fr_n, lin_n);
        if constexpr (MODE == 1) {
          // first map of particle j (fr holds its position): the cell's lanes
          // each gather their y slice of the window, the butterfly over the
          // cell's lanes completes the sums, then the re-map deposits m v'
          float vinc[3] = {0.f, 0.f, 0.f}, dvinc[3] = {0.f, 0.f, 0.f};
          const bool own = lin == cell;
          Wts W;
          if (own) {
            float frac[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int a = 0; a < D; ++a) frac[a] = __fsub_rn(cell_coord(fr[a], g.origin[a], inv_h), (float)c[a]);
            weights(frac, W);
            if (!none) {
#pragma unroll
              for (int sx = 0; sx < SPAN; ++sx)
#pragma unroll
                for (int sz = 0; sz < SZH; ++sz) {
                  const float wxz = W.Nx[sx] * W.Nz[sz];
#pragma unroll
                  for (int a = 0; a < D; ++a) {
                    vinc[a] += wxz * vnode[sx][sz][a];
                    dvinc[a] += wxz * vnode[sx][sz][D + a];
                  }
                }
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
              vinc[a] *= W.Ny;
              dvinc[a] *= W.Ny;
            }
#pragma unroll
            for (int o = G::CX; o < 64; o <<= 1)
#pragma unroll
              for (int a = 0; a < D; ++a) {
                vinc[a] += __shfl_xor(vinc[a], o);
                dvinc[a] += __shfl_xor(dvinc[a], o);
              }
          } else if (sy == 0) {
            // a particle listed in this cell whose bin is another (only when
            // the partition was built from other positions): its own window
            gather_one<D, SF>(prm, g, in.x, n, part.sorted[j], nodes, vinc, dvinc);
          }
          float vn[3] = {0.f, 0.f, 0.f};
#pragma unroll
          for (int a = 0; a < D; ++a) {
            if (!own) {  // the cell's sy = 0 lane gathered
              vinc[a] = __shfl(vinc[a], lane % G::CX);
              dvinc[a] = __shfl(dvinc[a], lane % G::CX);
            }
            vn[a] = am * vinc[a] + prm.alpha * (cur.v[a] + dvinc[a]);
            if (a % G::SY == sy) {  // back into the particle's LDS slot (read above)
              st[(ST::F_V + a) * CW + (j - w0)] = vn[a];
              st[a * CW + (j - w0)] = fr[a] + prm.dt * vinc[a];
            }
          }
          P2GRec rv = cur;
#pragma unroll
          for (int a = 0; a < D; ++a) rv.v[a] = vn[a];
          if (own) {
            if (!none) deposit(rv, W);
          } else if (zh == 0) {
            p2g_stray<D, SF, 1>(prm, g, nodes, in.x, n, part.sorted[j], sy, rv.m, rv.v, rv.A, rv.f);
          }
        } else if (lin != cell) {
          // a particle listed in this cell whose bin is another (only when the
          // partition was built from other positions); MODE 2 is only ever
          // called on a partition of the same positions
          if (MODE != 2 && zh == 0)
            p2g_stray<D, SF, MODE>(prm, g, nodes, in.x, n, part.sorted[j], sy, cur.m, cur.v, cur.A, cur.f);
        } else {
          Wts W;
          weights(fr, W);
          deposit(cur, W);
        }
        if (kPrefetch) {
          cur = nxt;
          lin = lin_n;
#pragma unroll
          for (int a = 0; a < 3; ++a) fr[a] = fr_n[a];
        }
      }
      wave_lds_sync();
      if constexpr (MODE == 1) {
        // the window's x', v' leave with coalesced stores (written from the
        // cells' lanes they would land 16 lines per store, partly filled)
        for (int q = lane; q < wn; q += 64)
#pragma unroll
          for (int a = 0; a < D; ++a) {
            out.x[a * n + w0 + q] = st[a * CW + q];
            out.v[a * n + w0 + q] = st[(ST::F_V + a) * CW + q];
          }
        wave_lds_sync();
      }
    }
  }

  // sums into the LDS tile, one node offset (sx, sz) per phase; the waves
  // holding that z offset write. Node offsets outside the grid stay in the
  // tile too when the wrap-around is reproduced with weights that do not
  // depend on node types (linear, quadratic): the grid kernel adds such a
  // virtual node's sum to the node its linear index reaches, in a fixed
  // order. With typed cubic splines the wrapped node's types change the
  // weights, and p2g_wrap_kernel forms those sums instead. (MODE 2 transposes
  // G2P, which never reads wrapped nodes.)
  const bool keep_out = MODE != 2 && prm.wrap_quirk && !(SF == PMPM_SF_CUBIC && prm.node_types_built);
#pragma unroll
  for (int sx = 0; sx < SPAN; ++sx)
#pragma unroll
    for (int sz = 0; sz < SZ; ++sz) {
      __syncthreads();
      if (sz / SZH == zh && (keep_out || (x_in[sx] && y_in && z_in[sz]))) {
        const int li = (lx + sx) + NXP * (sy + NY * (lz + sz));
#pragma unroll
        for (int k = 0; k < NV; ++k) tile[k * NT + li] += acc[sx * SZH + sz % SZH][k];
      }
    }

  __syncthreads();
  constexpr int NX = G::NX, NTN = G::NTN;
  float *slice = partials + (int64_t)blockIdx.x * NV * NTN;
  if (t == 0) flags[blockIdx.x] = 1;
  for (int i = t; i < NTN; i += G::THREADS) {
    const int ix = i % NX;
    const int iy = (i / NX) % NY;
    const int iz = i / (NX * NY);
    const int li = ix + NXP * (iy + NY * iz);
#pragma unroll
    for (int k = 0; k < NV; ++k) slice[k * NTN + i] = tile[k * NT + li];
  }
}

// P2G of one tile on the matrix cores (3D, cubic splines with node types):
// the same tiles, slices and results as p2g_kernel MODE 0, with the per-cell
// sums done by v_mfma_f32_16x16x4_f32. For one cell and its 4 x 4 x 4 node
// window, with i, j, k the x, y, z node offsets and p the cell's particles,
//   C1[(i, f)][(j, k)] = sum_p (Nx_i F_f)(p) (Ny_j Nz_k)(p),  F = (m, m v)
//   C2[(i, a)][(j, k)] = sum_p (dNx_i A_ax)(Ny_j Nz_k) + (Nx_i A_ay)(dNy_j Nz_k)
//                              + (Nx_i A_az)(Ny_j dNz_k),       A = -V sigma
// which are the mass, momentum and force deposits (usl_inline.h:114-160) as
// 16 x 16 products over K = particles, four particles per instruction. In the
// instruction's layout lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and receives D[4 (l / 16) + r][l % 16]: it evaluates x and y weights at
// offset (l % 16) / 4 and the z weight at offset l % 4 for particle l / 16,
// and owns node (i = l / 16, j = (l % 16) / 4, k = l % 4) of the result.
// Wave w walks the 16 cells of z layer w one after another: it stages up to
// 32 of a cell's particles in LDS (lane = particle, coalesced loads; the 24
// span weights and derivatives and the products m v, -V sigma computed once
// per particle), so an instruction's operands are a record read and one
// multiply; it runs the products on four independent accumulators and adds
// each lane's node into the wave's own LDS node block; the four blocks
// overlap in z and are summed into the tile slice at the end.
struct MfmaGeom {
  // The tile: CELLS x 1 x 4 cells, one z layer per wave. Eight cells per
  // wave keep a workgroup short enough that the chip's workgroup slots are
  // refilled evenly (a tile is the unit of work, and a column whose x
  // extent ends inside a tile leaves it partly empty).
  static constexpr int CELLS = 8;                       // cells per wave (x)
  static constexpr int BX = CELLS + 3, BY = 4, BZ = 4;  // a wave's node block
  static constexpr int TZ = 4, NZ = TZ + 3;             // tile layers, slice node layers
  static constexpr int NTN = BX * BY * NZ;              // nodes of one partial slice
  // x pitch of the block in LDS: the 64 lanes of a flush, offsets
  // pk + 12 s1 + 48 s2, fall into 64 distinct banks
  static constexpr int BXP = BX + 1;
  static constexpr int BN = BXP * BY * BZ;
  static constexpr int CH = 32;  // particles staged per round
  // a staged particle's record: per span offset s (7 s + 0..6) the weights
  // Nx, dNx, Ny, dNy, Nz, dNz and mass / moment field s (m, m vx, m vy,
  // m vz), then A = -V sigma (28..33: xx, yy, zz, xy, xz, yz) and a zero
  // (34); the odd pitch puts the 16 records one read touches on distinct
  // banks
  static constexpr int RS = 35;
};
typedef float v4f __attribute__((ext_vector_type(4)));

// REMAP: the MUSL re-map instead (p2g_kernel MODE 1's deposit): m v' with v'
// from slot j of `out` (written by the first map, g2p_kernel MODE 1) and the
// weights of the old positions; the slice holds the 3 momentum fields.
// FEXT: the step has planar-domain walls; each particle's wall force f
// (planar_force) is staged too and w f joins the force product through its
// A operand (Nx f_a next to dNx A_a.x, both against Ny Nz).
template <int SF, bool REMAP = false, bool FEXT = false>
__global__ __launch_bounds__(256) void p2g_mfma_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles in,
                                                       pmpm_particles out, pmpm_spatial part, pmpm_nodes nodes,
                                                       int tiles_x, int tiles_y, float *__restrict__ partials,
                                                       int32_t *__restrict__ flags) {
  using G = P2GGeom<3, SF>;
  using M = MfmaGeom;
  static_assert(G::SPAN == 4 && G::NLO == -1, "cubic spans");
  static_assert((M::BX - 1) / M::CELLS + 1 <= kCoverX && (M::NZ - 1) / M::TZ + 1 <= kCoverZ,
                "the grid kernel's gather visits at most kCover tiles per axis");
  constexpr int NV = REMAP ? 3 : 7, NTN = M::NTN, NLO = G::NLO;
  static_assert(NV == p2g_nv<3, REMAP ? 1 : 0>(), "slice fields of the matching p2g_kernel mode");
  static_assert(!(REMAP && FEXT), "the re-map carries no force");
  constexpr int RS = M::RS + (FEXT ? 4 : 0);  // record pitch (35..37: wall force)
  __shared__ float block[4][NV][M::BN];
  __shared__ float stage[4][M::CH * RS];
  __shared__ __attribute__((aligned(16))) float s_coef[5][4][4];

  const int n = in.n;
  const float inv_h = g.inv_cell_size;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int bx = blockIdx.x % tiles_x;
  const int by = (blockIdx.x / tiles_x) % tiles_y;
  const int bz = blockIdx.x / (tiles_x * tiles_y);
  const int cy = by, cz = bz * 4 + wave;  // this wave's cell row
  // lanes 0..15 hold the ranges of the wave's 16 cells
  int cs = 0, ce = 0;
  {
    const int cx = bx * M::CELLS + (lane & 15);
    if (lane < 16 && cx < g.num_cells[0] && cy < g.num_cells[1] && cz < g.num_cells[2]) {
      const int cell = cx + g.num_cells[0] * (cy + g.num_cells[1] * cz);
      ce = part.cell_end[cell];
      cs = part.cell_start[cell];
    }
  }
  if (t < 80) (&s_coef[0][0][0])[t] = (&kCubic[0][0][0])[t];
  if (!__syncthreads_or(ce > cs)) {
    if (t == 0) flags[blockIdx.x] = 0;
    return;
  }
  float *blk = &block[wave][0][0];
  for (int i = lane; i < NV * M::BN; i += 64) blk[i] = 0.f;
  float *st = &stage[wave][0];

  // lane roles in the products
  const int s1 = (lane & 15) >> 2, s2 = lane & 3, pk = lane >> 4;
  // staged entries this lane reads (a particle's record, RS floats): the
  // weights of its offsets (x and y at s1, z at s2), mass / moment field s2,
  // row a = s2 of A = -V sigma (row 3 of the product is padding and reads the
  // record's zero) and, with walls, force component s2
  const int ex = 7 * s1, ez = 7 * s2;
  const int ax = s2 == 0 ? 28 : (s2 == 1 ? 31 : (s2 == 2 ? 32 : 34));
  const int ay = s2 == 0 ? 31 : (s2 == 1 ? 29 : (s2 == 2 ? 33 : 34));
  const int az = s2 == 0 ? 32 : (s2 == 1 ? 33 : (s2 == 2 ? 30 : 34));
  const int af = s2 < 3 ? 35 + s2 : 34;
  // span coefficients: piece 3 - s of the type of the node at offset s,
  // read from the LDS copy of the table (the types: y and z fixed for the
  // wave's row, x per cell; wave-uniform)
  int tys[4], tzs[4], txs[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    tys[s] = cubic_node_type(cy + NLO + s, g.num_cells[1], prm.node_types_built);
    tzs[s] = cubic_node_type(cz + NLO + s, g.num_cells[2], prm.node_types_built);
    txs[s] = 0;
  }
  auto coef = [&](int type, int s) { return reinterpret_cast<const float4 *>(&s_coef[type][3 - s][0])[0]; };
  // cubic piece at r (offset s of the span), value and derivative; offset 0
  // at frac = 1 - 2^-24 rounds to r = 2, where the spline is 0 (span_eval)
  auto piece = [&](const float4 c, float r, bool first, float &N, float &dN) {
    N = ((c.w * r + c.z) * r + c.y) * r + c.x;
    dN = inv_h * ((3.f * c.w * r + 2.f * c.z) * r + c.y);
    if (first && !(r < 2.f)) N = dN = 0.f;
  };

  // The wave walks its cells' particles in chunks of up to CH (lane =
  // particle), cell by cell. Loads run ahead of the work: while a chunk's
  // products run, the next chunk's particle fields and the slots of the one
  // after are in flight, so the latency of the dependent slot -> field loads
  // hides behind the products of the chunk before.
  // a chunk: cell lx of the row, particles [w0, w0 + CH) of its range; the
  // ranges are read from lanes 0..15 with scalar lane reads (the chunk walk
  // stays in scalar registers)
  auto cs_of = [&](int l) { return __builtin_amdgcn_readlane(cs, l); };
  auto ce_of = [&](int l) { return __builtin_amdgcn_readlane(ce, l); };
  auto first_chunk_from = [&](int l0, int &lx, int &w0) {  // wave-uniform
    lx = M::CELLS;
    for (int l = l0; l < M::CELLS; ++l) {
      const int a = cs_of(l), b = ce_of(l);
      if (b > a) { lx = l; w0 = a; return; }
    }
  };
  auto next_chunk = [&](int lx, int w0, int &nlx, int &nw0) {
    if (w0 + M::CH < ce_of(lx)) { nlx = lx; nw0 = w0 + M::CH; return; }
    first_chunk_from(lx + 1, nlx, nw0);
  };
  // raw fields of one particle of a chunk, held in registers until staged.
  // The loads are unconditional (lanes past the chunk read the chunk's last
  // particle) so that no branch or select waits for them: they stay in
  // flight while the chunk before is staged and multiplied.
  struct Raw {
    int slot;
    bool valid;
    float x[3], v[3], m, vol;
    Sym sg;
  };
  // (lx == CELLS: no chunk; its loads read particle 0 and stage nothing, so
  // that every step issues the same loads)
  auto chunk_len = [&](int lx, int w0) { return lx < M::CELLS ? min(M::CH, ce_of(lx) - w0) : 0; };
  // lanes l and l + CH hold the same particle (see stage_raw)
  static_assert(M::CH == 32, "two lanes per staged particle");
  const int prow = lane & (M::CH - 1);
  auto chunk_j = [&](int lx, int w0) {  // this lane's sorted index, clamped to the chunk
    return lx < M::CELLS ? w0 + min(prow, chunk_len(lx, w0) - 1) : 0;
  };
  auto chunk_slot = [&](int lx, int w0) { return part.sorted[chunk_j(lx, w0)]; };
  auto load_raw = [&](int lx, int w0, int slot, Raw &r) {
    r.slot = slot;
    r.valid = prow < chunk_len(lx, w0);
    const int j = chunk_j(lx, w0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r.x[a] = in.x[a * n + slot];
      r.v[a] = REMAP ? out.v[a * n + j] : in.v[a * n + slot];
    }
    r.m = in.mass[slot];
    r.vol = REMAP ? 0.f : in.volume[slot];
    r.sg = REMAP ? sym_zero() : load_sym(in.stress, n, slot);
  };
  // a chunk's particle records into the LDS stage: lanes l and l + 32 both
  // hold particle l (row l) and each writes half of its record, half h =
  // lane / 32 the weights of span offsets s = 2h, 2h + 1 with their fields,
  // and half of A (the same instructions on both halves, so a record costs
  // half the instructions of one lane per particle). The records past the
  // chunk's particles are zero.
  auto stage_raw = [&](int lx, const Raw &r) {
    const int h = lane >> 5;
    float *sp = st + prow * RS;
    if (!r.valid) {
      if (h == 0) {
#pragma unroll
        for (int e = 0; e < RS; ++e)
          if (!REMAP || (e < 28 && (e % 7) % 2 == 0)) sp[e] = 0.f;  // (the entries the products read)
      }
      return;
    }
    const int cx = bx * M::CELLS + lx;
    float fr[3];
    bool own = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float ta = cell_coord(r.x[a], g.origin[a], inv_h);
      const int ba = (int)ta;
      own = own && ta >= 0.f && ba == (a == 0 ? cx : (a == 1 ? cy : cz));
      fr[a] = __fsub_rn(ta, (float)ba);
    }
    Sym A = sym_zero();
    float fe[3] = {0.f, 0.f, 0.f};
    if (!REMAP) {
      const float mvol = -r.vol;
      A = Sym{mvol * r.sg.xx, mvol * r.sg.yy, mvol * r.sg.zz, mvol * r.sg.xy, mvol * r.sg.xz, mvol * r.sg.yz};
      if (FEXT) planar_force<3>(prm, g, r.x, r.v, r.vol, r.m, fe);
    }
    if (!own && h == 0) {
      // listed in this cell but binned elsewhere (a partition of other
      // positions): deposited over its own window, no product terms here
#pragma unroll 1
      for (int sy = 0; sy < 4; ++sy)
        p2g_stray<3, SF, REMAP ? 1 : 0>(prm, g, nodes, in.x, n, r.slot, sy, r.m, r.v, A, fe);
    }
    const float keep = own ? 1.f : 0.f;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int s = 2 * h + j;
      const int tx = h ? txs[2 + j] : txs[j], ty = h ? tys[2 + j] : tys[j], tz = h ? tzs[2 + j] : tzs[j];
      float N, dN;
      // (the re-map's products use the weights only: no derivatives)
      piece(coef(tx, s), __fsub_rn(fr[0], (float)(NLO + s)), s == 0, N, dN);
      sp[7 * s + 0] = N;
      if (!REMAP) sp[7 * s + 1] = dN;
      piece(coef(ty, s), __fsub_rn(fr[1], (float)(NLO + s)), s == 0, N, dN);
      sp[7 * s + 2] = N;
      if (!REMAP) sp[7 * s + 3] = dN;
      piece(coef(tz, s), __fsub_rn(fr[2], (float)(NLO + s)), s == 0, N, dN);
      sp[7 * s + 4] = N;
      if (!REMAP) sp[7 * s + 5] = dN;
      // field s: the mass (the re-map's field 0 is padding), then m v_a
      const float va = s == 1 ? r.v[0] : (s == 2 ? r.v[1] : r.v[2]);
      sp[7 * s + 6] = s == 0 ? (REMAP ? 0.f : keep * r.m) : keep * r.m * va;
    }
    if (!REMAP) {  // half 0: xx, yy, zz; half 1: xy, xz, yz
      sp[28 + 3 * h] = keep * (h ? A.xy : A.xx);
      sp[29 + 3 * h] = keep * (h ? A.xz : A.yy);
      sp[30 + 3 * h] = keep * (h ? A.yz : A.zz);
    }
    if (h == 0) {
      sp[34] = 0.f;
      if (FEXT) {
#pragma unroll
        for (int a = 0; a < 3; ++a) sp[35 + a] = keep * fe[a];
      }
    }
  };

  // The chunk walk: chunk k's fields in one register set while chunk k+1's
  // load into the other and chunk k+2's slots load (the two sets alternate,
  // so no register copy waits for a load in flight). Each step issues the
  // slots before the fields, so the in-order load counter lets a wait for
  // the slots leave the older fields in flight.
  int lx, w0 = 0;
  first_chunk_from(0, lx, w0);
  int nlx = M::CELLS, nw0 = 0;
  next_chunk(lx, w0, nlx, nw0);
  Raw ra, rb;
  int sa = chunk_slot(lx, w0), sb = chunk_slot(nlx, nw0);
  load_raw(lx, w0, sa, ra);
  int xc_cell = -1;
  // four independent accumulation chains: mass and moments, and the three
  // terms of the force (summed when the cell is done)
  v4f C1 = {0.f, 0.f, 0.f, 0.f}, C2x = C1, C2y = C1, C2z = C1;
  auto chunk = [&](Raw &cur, int &s_cur, Raw &nxt, int &s_nxt) {
    int nnlx = M::CELLS, nnw0 = 0;
    if (nlx < M::CELLS) next_chunk(nlx, nw0, nnlx, nnw0);
    s_cur = chunk_slot(nnlx, nnw0);      // the slots of chunk k+2
    load_raw(nlx, nw0, s_nxt, nxt);      // in flight during this chunk's products
    if (xc_cell != lx) {  // x node types of the new cell's span
#pragma unroll
      for (int s = 0; s < 4; ++s)
        txs[s] = cubic_node_type(bx * M::CELLS + lx + NLO + s, g.num_cells[0], prm.node_types_built);
      xc_cell = lx;
    }
    stage_raw(lx, cur);
    wave_lds_sync();
    const int wn = min(M::CH, ce_of(lx) - w0);
#pragma unroll
    for (int k0 = 0; k0 < M::CH; k0 += 4) {
      if (k0 >= wn) break;  // wave-uniform
      const float *sp = st + (k0 + pk) * RS;
      const float Nx = sp[ex], dNx = sp[ex + 1], Ny = sp[ex + 2], dNy = sp[ex + 3];
      const float Nz = sp[ez + 4], dNz = sp[ez + 5], Ff = sp[ez + 6];
      const float b1 = Ny * Nz;
      C1 = __builtin_amdgcn_mfma_f32_16x16x4f32(Nx * Ff, b1, C1, 0, 0, 0);
      if (!REMAP) {
        float a2x = dNx * sp[ax];
        if (FEXT) a2x += Nx * sp[af];  // + w f_a
        C2x = __builtin_amdgcn_mfma_f32_16x16x4f32(a2x, b1, C2x, 0, 0, 0);
        C2y = __builtin_amdgcn_mfma_f32_16x16x4f32(Nx * sp[ay], dNy * Nz, C2y, 0, 0, 0);
        C2z = __builtin_amdgcn_mfma_f32_16x16x4f32(Nx * sp[az], Ny * dNz, C2z, 0, 0, 0);
      }
    }
    wave_lds_sync();  // the stage is read; the next chunk may overwrite it
    if (nlx != lx) {
      // the cell is done: node (i, j, k) = (l / 16, (l % 16) / 4, l % 4) of
      // its window
      const int nb = (lx + pk) + M::BXP * (s1 + M::BY * s2);
      if (REMAP) {  // rows 1..3: m v'
#pragma unroll
        for (int r = 1; r < 4; ++r) blk[(r - 1) * M::BN + nb] += C1[r];
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) blk[r * M::BN + nb] += C1[r];
#pragma unroll
        for (int r = 0; r < 3; ++r) blk[(4 + r) * M::BN + nb] += (C2x[r] + C2y[r]) + C2z[r];
      }
      C1 = v4f{0.f, 0.f, 0.f, 0.f};
      C2x = C1;
      C2y = C1;
      C2z = C1;
    }
    lx = nlx;
    w0 = nw0;
    nlx = nnlx;
    nw0 = nnw0;
  };
  while (lx < M::CELLS) {  // wave-uniform
    chunk(ra, sa, rb, sb);
    if (lx >= M::CELLS) break;
    chunk(rb, sb, ra, sa);
  }
  __syncthreads();
  // slice node (ix, iy, iz): the blocks of waves iz - 3 .. iz
  float *slice = partials + (int64_t)blockIdx.x * NV * NTN;
  if (t == 0) flags[blockIdx.x] = 1;
  for (int i = t; i < NTN; i += 256) {
    const int ix = i % M::BX, iy = (i / M::BX) % M::BY, iz = i / (M::BX * M::BY);
    float sum[NV] = {};
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int z = iz - w;
      if (z < 0 || z >= M::BZ) continue;
#pragma unroll
      for (int f = 0; f < NV; ++f) sum[f] += block[w][f][ix + M::BXP * (iy + M::BY * z)];
    }
#pragma unroll
    for (int f = 0; f < NV; ++f) slice[f * NTN + i] = sum[f];
  }
}

// Floats per (cell, offset) record of the wrap-around sums: the NV fields
// padded to 16 or 32 bytes, so a record is one or two vector loads.
__host__ __device__ constexpr int wrap_rec_stride(int nv) { return nv <= 4 ? 4 : 8; }

// The wrap-around sums of one non-empty wrap cell c (index w), one lane per
// node offset; writes the records of the offsets the gather reads.
template <int D, int MODE, int NV, int NS>
__device__ inline void wrap_cell_sums(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in,
                                      const pmpm_particles &out, const pmpm_spatial &part, float *rec, int wcells,
                                      int w, const int *c, int cs, int ce, int lane, float (*s_frac)[64],
                                      float (*s_v)[64], float (*s_A)[64], float *s_m, int *s_own, int *s_list,
                                      uint8_t *touch, int32_t *wlist) {
  constexpr int SPAN = 4, NLO = -1;
  const int nlanes = D == 3 ? 64 : (D == 2 ? 16 : 4);
  const int nn = g.num_cells_total;
  // the node of span offset o (o = sx + 4 (sy + 4 sz)); active: outside the
  // grid with its linear index in range
  auto offset_node = [&](int o, int *sv, int64_t &lin) {
    sv[0] = o % SPAN;
    sv[1] = D > 1 ? (o / SPAN) % SPAN : 0;
    sv[2] = D > 2 ? o / (SPAN * SPAN) : 0;
    int nb[3] = {0, 0, 0};
    bool inside = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      nb[a] = c[a] + NLO + sv[a];
      inside = inside && nb[a] >= 0 && nb[a] < g.num_cells[a];
    }
    lin = lin_index(g, nb, D);
    return !inside && lin >= 0 && lin < nn;
  };
  int sv[3];
  int64_t lin;
  const bool active = lane < nlanes && offset_node(lane, sv, lin);
  // the node's flag (one byte of a 32-bit word); the first writer also
  // appends the node to the gather's list, so the list holds each reached
  // node once. The list has kWrapListShards parts (counts wlist[0..S), part
  // p's entries from wlist[S + p cap]); a wave appends to part w mod S, so
  // the appends of the wrap cells spread over S counters
  const int cap = wrap_target_count<D>(g), shard = w & (kWrapListShards - 1);
  bool fresh = false;
  if (active) {
    const uint32_t bit = 1u << (8 * (int)(lin & 3));
    fresh = (atomicOr(reinterpret_cast<uint32_t *>(touch) + (lin >> 2), bit) & bit) == 0u;
  }
  const uint64_t fm = __ballot(fresh);
  if (fm) {  // wave-uniform
    const int first = __ffsll((unsigned long long)fm) - 1;
    int at = 0;
    if (lane == first) at = atomicAdd(&wlist[shard], __popcll(fm));
    at = __shfl(at, first);
    const int idx = at + __popcll(fm & ((1ull << lane) - 1ull));
    if (fresh && idx < cap) wlist[kWrapListShards + shard * cap + idx] = (int)lin;  // (at most cap nodes in all)
  }
  // the active offsets in order (s_list[rank] = offset)
  const uint64_t act = __ballot(active);
  const int nact = __popcll(act);
  if (active) s_list[__popcll(act & ((1ull << lane) - 1ull))] = lane;
  wave_lds_sync();

  const int n = in.n;
  const float inv_h = g.inv_cell_size;
  const int q = lane & 3;  // a quarter of the cell's particles per lane
  // rounds of 16 active offsets, four lanes per offset
  for (int r0 = 0; r0 < nact; r0 += 16) {  // wave-uniform
    const int idx = r0 + (lane >> 2);
    const bool mine = idx < nact;
    const int off = mine ? s_list[idx] : 0;
    // r = frac - (NLO + s) lies in piece 3 - s of the wrapped node's spline
    // (the same piece argument as span_eval): its coefficients are fixed per
    // lane
    float cf[3][4];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int k = 0; k < 4; ++k) cf[a][k] = 0.f;
    int so[3] = {0, 0, 0};
    if (mine) {
      int64_t ln;
      (void)offset_node(off, so, ln);
      int ids[3];
      node_ids<D>(g, ln, ids);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const int type = cubic_node_type(ids[a], g.num_cells[a], prm.node_types_built);
#pragma unroll
        for (int k = 0; k < 4; ++k) cf[a][k] = kCubic[type][3 - so[a]][k];
      }
    }
    float acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.f;
    for (int base = cs; base < ce; base += 64) {
      const int cnt = min(64, ce - base);
      wave_lds_sync();  // the stage of the previous round is read
      if (lane < cnt) {
        const int j = base + lane;
        const int slot = part.sorted[j];
        bool own = true;
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const float ta = cell_coord(in.x[a * n + slot], g.origin[a], inv_h);
          const int ba = (int)ta;
          own = own && ta >= 0.f && ba == c[a];
          s_frac[a][lane] = __fsub_rn(ta, (float)ba);
        }
        s_own[lane] = own;
        P2GRec r;
        p2g_load<D, MODE>(prm, g, MODE == 0 && has_planar(prm), in, out, nullptr, n, slot, j, r);
        s_m[lane] = r.m;
#pragma unroll
        for (int a = 0; a < D; ++a) s_v[a][lane] = r.v[a];
        if constexpr (MODE == 0) {
          s_A[0][lane] = r.A.xx; s_A[1][lane] = r.A.yy; s_A[2][lane] = r.A.zz;
          s_A[3][lane] = r.A.xy; s_A[4][lane] = r.A.xz; s_A[5][lane] = r.A.yz;
          s_A[6][lane] = r.f[0]; s_A[7][lane] = r.f[1]; s_A[8][lane] = r.f[2];  // wall force
        }
      }
      wave_lds_sync();
      if (mine) {
        for (int k = q; k < cnt; k += 4) {
          if (!s_own[k]) continue;  // deposited whole by p2g_stray
          float N[3] = {1.f, 1.f, 1.f}, dN[3] = {0.f, 0.f, 0.f};
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const float r = __fsub_rn(s_frac[a][k], (float)(NLO + so[a]));
            N[a] = ((cf[a][3] * r + cf[a][2]) * r + cf[a][1]) * r + cf[a][0];
            dN[a] = inv_h * ((3.f * cf[a][3] * r + 2.f * cf[a][2]) * r + cf[a][1]);
            if (so[a] == 0 && !(r < 2.f)) N[a] = dN[a] = 0.f;  // frac = 1 - 2^-24
          }
          const float wk = (N[0] * N[1]) * N[2];
          const float wm = wk * s_m[k];
          const float gr[3] = {(dN[0] * N[1]) * N[2], (dN[1] * N[0]) * N[2], (dN[2] * N[0]) * N[1]};
          if constexpr (MODE == 0) {
            const Sym A{s_A[0][k], s_A[1][k], s_A[2][k], s_A[3][k], s_A[4][k], s_A[5][k]};
            acc[0] += wm;
#pragma unroll
            for (int a = 0; a < D; ++a) {
              acc[1 + a] += wm * s_v[a][k];
              acc[1 + D + a] += wk * s_A[6 + a][k];
#pragma unroll
              for (int b = 0; b < D; ++b) acc[1 + D + a] += sym_at(A, a, b) * gr[b];
            }
          } else {
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] += wm * s_v[a][k];
          }
        }
      }
    }
    // the four quarters of an offset, combined in a fixed order
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      acc[k] += __shfl_xor(acc[k], 1);
      acc[k] += __shfl_xor(acc[k], 2);
    }
    if (mine && q == 0) {  // one 16- or 32-byte record, fields past NV zero
      constexpr int RST = wrap_rec_stride(NV);
      float4 *dst = reinterpret_cast<float4 *>(rec + ((int64_t)w * 64 + off) * RST);
#pragma unroll
      for (int v = 0; v < RST / 4; ++v)
        dst[v] = float4{4 * v + 0 < NV ? acc[4 * v + 0] : 0.f, 4 * v + 1 < NV ? acc[4 * v + 1] : 0.f,
                        4 * v + 2 < NV ? acc[4 * v + 2] : 0.f, 4 * v + 3 < NV ? acc[4 * v + 3] : 0.f};
    }
  }
}

// Wrap-around sums of typed cubic splines: one wave per cell that can wrap
// (four per workgroup). A node offset (sx, sy, sz) of the cell's span whose
// node lies outside the grid but whose linear index is in range re-sums the
// cell's particles with the types of the node that index reaches
// (shapefunction.cpp:96-112); four lanes share such an offset, each summing a
// quarter of the particles, and the quarters are combined in a fixed order.
// The sums of a non-empty cell are stored per (cell, offset) in `rec`
// ([wcells][64] records of wrap_rec_stride(NV) floats; only the offsets the
// gather reads are written);
// wrap_gather_kernel adds them to their nodes in a fixed order, so the node
// sums do not depend on scheduling. Empty cells write nothing (the gather
// skips them by their cell ranges). The cell's particles are first loaded
// into LDS by the wave (lane = particle, up to 64 per round), so the per-node
// loops read LDS instead of walking a chain of dependent global loads.
constexpr int kWrapCellsPerBlock = 4;
template <int D, int MODE>
__global__ __launch_bounds__(64 * kWrapCellsPerBlock) void p2g_wrap_kernel(pmpm_params prm, pmpm_grid g,
                                                                           pmpm_particles in, pmpm_particles out,
                                                                           pmpm_spatial part, float *__restrict__ rec,
                                                                           uint8_t *__restrict__ touch,
                                                                           int32_t *__restrict__ wlist, int wcells) {
  constexpr int NV = MODE == 0 ? 1 + 2 * D : D;
  constexpr int NS = MODE == 0 ? 9 : 1;  // staged -V sigma (6) and wall force (3)
  constexpr int WPB = kWrapCellsPerBlock;
  __shared__ float s_frac_all[WPB][3][64], s_v_all[WPB][3][64], s_m_all[WPB][64], s_A_all[WPB][NS][64];
  __shared__ int s_own_all[WPB][64], s_list_all[WPB][64];
  // every wave works on its own cell; only wave-level synchronisation below
  const int wv = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int w = (int)blockIdx.x * WPB + wv;
  int c[3] = {0, 0, 0};
  if (w >= wcells || !wrap_cell<D>(g, w, c)) return;  // uniform over the wave
  const int cell = (int)lin_index(g, c, D);
  const int ce = part.cell_end[cell];
  const int cs = part.cell_start[cell];
  if (ce <= cs) return;
  wrap_cell_sums<D, MODE, NV, NS>(prm, g, in, out, part, rec, wcells, w, c, cs, ce, lane, s_frac_all[wv],
                                  s_v_all[wv], s_A_all[wv], s_m_all[wv], s_own_all[wv], s_list_all[wv], touch, wlist);
}

// The NV fields of the wrap record of (wrap cell w, offset rl).
template <int NV>
__device__ inline void wrap_rec_load(const float *wrec, int w, int rl, float *r) {
  constexpr int RST = wrap_rec_stride(NV);
  const float4 *src = reinterpret_cast<const float4 *>(wrec + ((int64_t)w * 64 + rl) * RST);
#pragma unroll
  for (int v = 0; v < RST / 4; ++v) {
    const float4 q = src[v];
    if (4 * v + 0 < NV) r[4 * v + 0] = q.x;
    if (4 * v + 1 < NV) r[4 * v + 1] = q.y;
    if (4 * v + 2 < NV) r[4 * v + 2] = q.z;
    if (4 * v + 3 < NV) r[4 * v + 3] = q.w;
  }
}

// Geometry of the P2G tiles for the grid kernel's gather.
struct TileGeom {
  int tiles[3];  // tiles per axis
  int T[3];      // cells per tile per axis
  int N[3];      // nodes per tile block per axis (T + SPAN - 1)
  int nlo;       // first node offset of a tile block
  int nv;        // fields per node in the partials
  int gather;    // 0: none, 1: P2G (mass, momentum, force), 2: MUSL re-map (momentum)
  // wrap-around sums (the reference bound-checks only the linear node index,
  // usl_inline.h:122-126): 0 none; 1 the out-of-grid nodes of the tile
  // partials (weights without node types); 2 the per-(cell, lane) records of
  // p2g_wrap_kernel (typed cubic)
  int wrap;
  int wcells;
  const float *wrec;
  const int32_t *cstart, *cend;  // cell ranges: a non-empty wrap cell wrote its records
};

// Sums the NV fields of the partial slices of the tiles whose node block
// covers node `ids`, in a fixed tile order (z, y, x ascending). The live
// flags of all candidate tiles are loaded first; then every candidate's loads
// are issued without a branch (a dead or out-of-range candidate reads tile
// 0's first entries and adds zero), so a node waits for two loads in a row,
// not one per tile. Returns whether any candidate tile was live (a node no
// live tile covers skips the loads).
template <int D, int NV>
__device__ inline bool gather_partials(const TileGeom &tg, const float *partials, const int32_t *flags,
                                       const int *ids, float *acc) {
  constexpr int MX = kCoverX, MY = D > 1 ? kCoverY : 1, MZ = D > 2 ? kCoverZ : 1, MC = MX * MY * MZ;
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    // tile b covers nodes [b*T + nlo, b*T + nlo + N - 1]
    const int r = ids[a] - tg.nlo;
    int b_hi = r / tg.T[a];
    int b_lo = (r - tg.N[a] + 1 + tg.T[a] - 1) >= 0 ? (r - tg.N[a] + tg.T[a]) / tg.T[a] : 0;
    if (b_hi > tg.tiles[a] - 1) b_hi = tg.tiles[a] - 1;
    lo[a] = b_lo;
    hi[a] = b_hi;
  }
  const int ntn = tg.N[0] * (D > 1 ? tg.N[1] : 1) * (D > 2 ? tg.N[2] : 1);
  bool live[MC];
  int64_t off[MC];
#pragma unroll
  for (int cz = 0; cz < MZ; ++cz)
#pragma unroll
    for (int cy = 0; cy < MY; ++cy)
#pragma unroll
      for (int cx = 0; cx < MX; ++cx) {
        const int c = (cz * MY + cy) * MX + cx;
        const int bx = lo[0] + cx, by = lo[1] + cy, bz = lo[2] + cz;
        const bool in = bx <= hi[0] && by <= hi[1] && bz <= hi[2];
        const int tile = in ? bx + tg.tiles[0] * (by + tg.tiles[1] * bz) : 0;
        live[c] = in && flags[tile] != 0;
        int li = ids[0] - (bx * tg.T[0] + tg.nlo);
        if (D > 1) li += tg.N[0] * (ids[1] - (by * tg.T[1] + tg.nlo));
        if (D > 2) li += tg.N[0] * tg.N[1] * (ids[2] - (bz * tg.T[2] + tg.nlo));
        off[c] = (int64_t)tile * NV * ntn + li;
      }
  bool any = false;
#pragma unroll
  for (int c = 0; c < MC; ++c) any = any || live[c];
  if (!any) return false;
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    const float *s = partials + (live[c] ? off[c] : 0);
    float v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = s[(int64_t)k * ntn];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += live[c] ? v[k] : 0.f;
  }
  return true;
}

// This lane's share of the wrap-around sums that land on node `node`: the
// node vectors outside the grid whose linear index is `node` (x out of range,
// or y in 3D), and for each the cells whose window holds it. Typed cubic
// (tg.wrap == 2): lane = the window offset (sx, sy, sz) of the source cell,
// whose record the wrap kernel left (empty cells wrote none and are skipped);
// other splines (tg.wrap == 1): lane 0 sums the tile partials at the outside
// node. Returns whether this lane added anything.
template <int D, int NV>
__device__ inline bool gather_wrapped_lane(const pmpm_grid &g, const TileGeom &tg, const float *partials,
                                           const int32_t *flags, int node, int lane, float *acc) {
  if (D < 2) return false;  // the only axis is the slowest: nothing wraps
  const int nx = g.num_cells[0], ny = g.num_cells[1], nz = D > 2 ? g.num_cells[2] : 1;
  const int span = tg.N[0] - tg.T[0] + 1;
  const int lo = tg.nlo, hi = tg.nlo + span - 1;
  // node = tx + nx (ty + ny tz); an outside node (bx, by, bz) with the same
  // linear index has bx = tx + kx nx and by + ny bz = rest - kx (32-bit
  // arithmetic: grids stay below 2^27 nodes)
  const int tx = node % nx, rest = node / nx;
  const int sx = lane & 3, sy = (lane >> 2) & 3, sz = lane >> 4;
  const bool my_lane = D > 2 || lane < 16;
  bool found = false;
#pragma unroll 1
  for (int kx = -2; kx <= 2; ++kx) {  // wave-uniform loops over the outside nodes
    const int bx = tx + kx * nx;
    if (bx < lo || bx > nx - 1 + hi) continue;
    const int q = rest - kx;
    const int qy = ((q % ny) + ny) % ny, qz0 = (q - qy) / ny;  // q = qy + ny qz0, exactly
#pragma unroll 1
    for (int ky = -2; ky <= 2; ++ky) {
      const int by = qy + ky * ny;
      if (by < lo || by > ny - 1 + hi) continue;
      const int qz = qz0 - ky;
      if (D == 2 && qz != 0) continue;
      if (D == 3 && (qz < lo || qz > nz - 1 + hi)) continue;
      const int nb[3] = {bx, by, D > 2 ? qz : 0};
      if (bx >= 0 && bx < nx && by >= 0 && by < ny && (D < 3 || (qz >= 0 && qz < nz))) continue;  // the node itself
      if (tg.wrap == 1) {
        if (lane == 0) found = gather_partials<D, NV>(tg, partials, flags, nb, acc) || found;
        continue;
      }
      if (!my_lane) continue;
      const int c[3] = {nb[0] + 1 - sx, nb[1] + 1 - sy, D > 2 ? nb[2] + 1 - sz : 0};
      if (c[0] < 0 || c[0] >= nx || c[1] < 0 || c[1] >= ny || (D > 2 && (c[2] < 0 || c[2] >= nz))) continue;
      const int w = wrap_index<D>(g, c);
      if (w < 0) continue;
      const int cl = c[0] + nx * (c[1] + ny * c[2]);
      if (tg.cend[cl] <= tg.cstart[cl]) continue;
      const int rl = sx + 4 * (sy + 4 * sz);
      float r[NV];
      wrap_rec_load<NV>(tg.wrec, w, rl, r);
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += r[k];
      found = true;
    }
  }
  return found;
}
// One wave per node the wrap-around can reach (wrap_target), four per
// workgroup: the lanes gather their candidates in parallel, a butterfly over
// the wave sums them (the same lanes in the same order every run) and lane 0
// adds the result to the node arrays, which hold nothing else but the (rare)
// stray deposits at this point; the grid kernel adds the tile partials next.
// Fixed order, no atomics. With typed cubic splines (touch != nullptr) only
// the nodes the wrap kernel flagged are gathered (and their flags cleared):
// the others receive nothing this pass.
template <int D, int NV>
__global__ __launch_bounds__(256) void wrap_gather_kernel(pmpm_grid g, TileGeom tg, const float *__restrict__ partials,
                                                          const int32_t *__restrict__ flags, pmpm_nodes nodes,
                                                          uint8_t *__restrict__ touch) {
  const int lane = (int)threadIdx.x & 63;
  const int t = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);  // wave-uniform
  int ids[3] = {0, 0, 0};
  if (t >= wrap_target_count<D>(g) || !wrap_target<D>(g, t, ids)) return;
  const int nn = g.num_cells_total;
  const int i = ids[0] + g.num_cells[0] * (ids[1] + (D > 2 ? g.num_cells[1] * ids[2] : 0));
  if (touch) {
    if (!touch[i]) return;  // uniform
    if (lane == 0) touch[i] = 0;
  }
  float acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.f;
  const bool found = gather_wrapped_lane<D, NV>(g, tg, partials, flags, i, lane, acc);
  if (!__ballot(found)) return;  // uniform: no wrap-around deposit reaches this node
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
  if (lane != 0) return;
  if constexpr (NV == 1 + 2 * D) {
    nodes.mass[i] += acc[0];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      nodes.mom[a * nn + i] += acc[1 + a];
      nodes.mom_nt[a * nn + i] += acc[1 + D + a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < D; ++a) nodes.mom_nt[a * nn + i] += acc[a];
  }
}

// gather_wrapped_lane for typed cubic records on grids of at least 3 cells
// per axis, with one round of loads: there the outside nodes whose linear
// index is `node` number at most four (at most two x images, x = tx or
// tx -/+ nx, and for each at most two y images), so they are found first by
// index arithmetic alone; then every lane issues the cell-range and record
// loads of its (image, offset) pairs together, and an empty cell's (stale)
// record is dropped by a select. The same pairs and summation order as
// gather_wrapped_lane.
template <int D, int NV>
__device__ inline bool gather_wrapped_typed(const pmpm_grid &g, const TileGeom &tg, int node, int lane, float *acc) {
  const int nx = g.num_cells[0], ny = g.num_cells[1], nz = D > 2 ? g.num_cells[2] : 1;
  constexpr int lo = -1, hi = 2;  // cubic spans: node offsets -1 .. +2
  const int tx = node % nx, rest = node / nx;
  int img[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, nimg = 0;
#pragma unroll
  for (int kx = -1; kx <= 1; ++kx) {
    const int bx = tx + kx * nx;
    if (bx < lo || bx > nx - 1 + hi) continue;
    const int q = rest - kx;
    const int qy = ((q % ny) + ny) % ny, qz0 = (q - qy) / ny;
#pragma unroll
    for (int ky = -1; ky <= 1; ++ky) {
      const int by = qy + ky * ny;
      if (by < lo || by > ny - 1 + hi) continue;
      const int qz = qz0 - ky;
      if (D == 2 && qz != 0) continue;
      if (D == 3 && (qz < lo || qz > nz - 1 + hi)) continue;
      if (bx >= 0 && bx < nx && by >= 0 && by < ny && (D < 3 || (qz >= 0 && qz < nz))) continue;  // the node itself
      // (slot by selects: no indexed register array)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (nimg == k) {
          img[k][0] = bx;
          img[k][1] = by;
          img[k][2] = D > 2 ? qz : 0;
        }
      ++nimg;
    }
  }
  const int sx = lane & 3, sy = (lane >> 2) & 3, sz = lane >> 4;
  const bool my_lane = D > 2 || lane < 16;
  int cl[4], wi[4];
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int c[3] = {img[k][0] + 1 - sx, img[k][1] + 1 - sy, D > 2 ? img[k][2] + 1 - sz : 0};
    ok[k] = k < nimg && my_lane && c[0] >= 0 && c[0] < nx && c[1] >= 0 && c[1] < ny &&
            (D < 3 || (c[2] >= 0 && c[2] < nz));
    wi[k] = ok[k] ? wrap_index<D>(g, c) : -1;
    ok[k] = ok[k] && wi[k] >= 0;
    cl[k] = ok[k] ? c[0] + nx * (c[1] + ny * c[2]) : 0;
    if (!ok[k]) wi[k] = 0;
  }
  const int rl = sx + 4 * (sy + 4 * sz);
  int cs[4], ce[4];
  float r[4][NV];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    cs[k] = tg.cstart[cl[k]];
    ce[k] = tg.cend[cl[k]];
    wrap_rec_load<NV>(tg.wrec, wi[k], rl, r[k]);
  }
  bool found = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool use = ok[k] && ce[k] > cs[k];
#pragma unroll
    for (int f = 0; f < NV; ++f) acc[f] += use ? r[k][f] : 0.f;
    found = found || use;
  }
  return found;
}

// The typed-cubic wrap-around sums, gathered for the nodes the wrap kernel
// listed (each once; the grid pass that follows empties the list): wave e
// takes list entry e (waves past the count exit after one load), gathers the
// node's
// candidates as wrap_gather_kernel does (same candidates, same order, so the
// node sums do not depend on the list order) and clears its flag.
template <int D, int NV>
__global__ __launch_bounds__(256) void wrap_gather_list_kernel(pmpm_grid g, TileGeom tg,
                                                               const float *__restrict__ partials,
                                                               const int32_t *__restrict__ flags, pmpm_nodes nodes,
                                                               uint8_t *__restrict__ touch, int32_t *__restrict__ wlist) {
  const int lane = (int)threadIdx.x & 63;
  const int wv = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6), nw = (int)gridDim.x * 4;
  const int nn = g.num_cells_total;
  // the list's parts: lane p < S holds part p's count; entry e of the whole
  // list is in the part whose inclusive count first exceeds e
  const int cap = wrap_target_count<D>(g);
  const int cnt = lane < kWrapListShards ? min(wlist[lane], cap) : 0;
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  const int count = __builtin_amdgcn_readlane(incl, 63);
  // grids below 3 cells on an axis have more images per node: generic path
  const bool small = g.num_cells[0] < 3 || g.num_cells[1] < 3 || (D > 2 && g.num_cells[2] < 3);
  for (int e = wv; e < count; e += nw) {  // wave-uniform
    const int part_e = __popcll(__ballot(lane < kWrapListShards && incl <= e));
    const int before = __shfl(incl - cnt, part_e);
    const int i = __builtin_amdgcn_readfirstlane(wlist[kWrapListShards + part_e * cap + (e - before)]);
    float acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.f;
    const bool found = small ? gather_wrapped_lane<D, NV>(g, tg, partials, flags, i, lane, acc)
                             : gather_wrapped_typed<D, NV>(g, tg, i, lane, acc);
    if (lane == 0) touch[i] = 0;
    if (!__ballot(found)) continue;  // uniform
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
    if (lane != 0) continue;
    if constexpr (NV == 1 + 2 * D) {
      nodes.mass[i] += acc[0];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        nodes.mom[a * nn + i] += acc[1 + a];
        nodes.mom_nt[a * nn + i] += acc[1 + D + a];
      }
    } else {
#pragma unroll
      for (int a = 0; a < D; ++a) nodes.mom_nt[a * nn + i] += acc[a];
    }
  }
}

// Per node: cutoff + forward moment (finalize), boundary conditions
// [bc_begin, bc_end) in list order, then the node velocities G2P reads.
// remap_pass: mom_nt holds the MUSL re-mapped momentum; only kinematic
// conditions apply (gravity already acted through the particle velocities).
// clear_mom_nt: mom_nt is left at zero instead of the forward moment (the
// MUSL step's first pass: the re-map accumulates into it next, and nothing
// reads the forward moment but the velocity records written here).
// clamp (optional): per node, the components a NodeDomain condition set to
// zero (bit a: moment, bit 3 + a: forward moment; a slip clamp counts when
// its input was not on the kept side), for the adjoint. wlist (optional): the
// typed-cubic wrap-around node list, emptied here for the next pass (its
// gather ran before this pass).
template <int D>
__global__ void grid_kernel(pmpm_params prm, pmpm_grid g, pmpm_nodes nodes, int finalize, int bc_begin, int bc_end,
                            int remap_pass, TileGeom tg, const float *__restrict__ partials,
                            const int32_t *__restrict__ flags, int clear_mom_nt, uint8_t *__restrict__ clamp,
                            int32_t *__restrict__ wlist) {
  const int nn = g.num_cells_total;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (wlist && i == 0)  // the wrap-around list of this pass is consumed
    for (int p2 = 0; p2 < kWrapListShards; ++p2) wlist[p2] = 0;
  if (i >= nn) return;
  float m = nodes.mass[i];
  float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f};
  if (tg.gather) {
    // node arrays hold the wrap-around deposits; add the tile partials
    int ids0[3];
    node_ids<D>(g, i, ids0);
    float acc[1 + 2 * 3] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (tg.gather == 1) gather_partials<D, 1 + 2 * D>(tg, partials, flags, ids0, acc);
    else gather_partials<D, D>(tg, partials, flags, ids0, acc);
    if (tg.gather == 1) {
      m += acc[0];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        nodes.mom[a * nn + i] += acc[1 + a];
        nodes.mom_nt[a * nn + i] += acc[1 + D + a];
      }
      nodes.mass[i] = m;
    } else {
#pragma unroll
      for (int a = 0; a < D; ++a) nodes.mom_nt[a * nn + i] += acc[a];
    }
  }
  if ((finalize || remap_pass) && m <= prm.small_mass_cutoff) {
    if (finalize) {
      nodes.mass[i] = 0.f;
#pragma unroll
      for (int a = 0; a < D; ++a) nodes.mom[a * nn + i] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) nodes.mom_nt[a * nn + i] = 0.f;
    // a body force acts whatever the node's mass (bodyforce_inline.h:41-59);
    // the gathers skip the node all the same
    for (int k = bc_begin; k < bc_end; ++k) {
      const pmpm_bc &bc = prm.bcs[k];
      if (bc.kind != PMPM_BC_BODY_FORCE || !bc.mask[i]) continue;
      if (bc.i[0] == 2) {
#pragma unroll
        for (int a = 0; a < D; ++a) nodes.mom_nt[a * nn + i] = bc.values[a * nn + i];
      } else if (!remap_pass) {
#pragma unroll
        for (int a = 0; a < D; ++a) nodes.mom[a * nn + i] += bc.values[a * nn + i];
      }
    }
    const float z[3] = {0.f, 0.f, 0.f};
    store_node_vel<D>(nodes.vel, i, z, z);
    return;
  }
#pragma unroll
  for (int a = 0; a < D; ++a) {
    P[a] = nodes.mom[a * nn + i];
    Q[a] = nodes.mom_nt[a * nn + i];
  }
  if (finalize) {
#pragma unroll
    for (int a = 0; a < D; ++a) Q[a] = P[a] + Q[a] * prm.dt;
  }
  int ids[3];
  node_ids<D>(g, i, ids);
  unsigned cbits = 0u;
  for (int k = bc_begin; k < bc_end; ++k) {
    const pmpm_bc &bc = prm.bcs[k];
    if (bc.kind == PMPM_BC_GRAVITY) {
      // gravity_inline.h:46-63
      if (remap_pass || m <= 0.000000001f) continue;
#pragma unroll
      for (int a = 0; a < D; ++a) Q[a] = __fadd_rn(Q[a], __fmul_rn(__fmul_rn(bc.f[a], m), prm.dt));
    } else if (bc.kind == PMPM_BC_NODE_DOMAIN) {
      // nodedomain_inline.h:40-87: 0 fixed, 1 slip (one-sided clamp)
      if (m <= 0.00001f) continue;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        if (ids[a] < 1) {
          if (bc.i[a] == 0) {
#pragma unroll
            for (int b = 0; b < D; ++b) { P[b] = 0.f; Q[b] = 0.f; }
            cbits |= 0x3fu;
          } else if (bc.i[a] == 1) {
            if (!(P[a] > 0.f)) cbits |= 1u << a;
            if (!(Q[a] > 0.f)) cbits |= 8u << a;
            P[a] = fmaxf(0.f, P[a]);
            Q[a] = fmaxf(0.f, Q[a]);
          }
        } else if (ids[a] >= g.num_cells[a] - 1) {
          if (bc.i[3 + a] == 0) {
#pragma unroll
            for (int b = 0; b < D; ++b) { P[b] = 0.f; Q[b] = 0.f; }
            cbits |= 0x3fu;
          } else if (bc.i[3 + a] == 1) {
            if (!(P[a] < 0.f)) cbits |= 1u << a;
            if (!(Q[a] < 0.f)) cbits |= 8u << a;
            P[a] = fminf(0.f, P[a]);
            Q[a] = fminf(0.f, Q[a]);
          }
        }
      }
    } else if (bc.kind == PMPM_BC_BODY_FORCE) {
      // bodyforce_inline.h:41-59: "fixed" sets the forward moments, the
      // other modes add to the moments; no mass test. The MUSL re-map pass
      // re-applies only the kinematic "fixed" mode.
      if (!bc.mask[i]) continue;
      if (bc.i[0] == 2) {
#pragma unroll
        for (int a = 0; a < D; ++a) Q[a] = bc.values[a * nn + i];
      } else if (!remap_pass) {
#pragma unroll
        for (int a = 0; a < D; ++a) P[a] = __fadd_rn(P[a], bc.values[a * nn + i]);
      }
    } else if (bc.kind == PMPM_BC_RIGID_SDF) {
      if (m <= 0.000000001f) continue;
      // box signed distance in the body frame; n = -grad(phi)
      float xn[3];
#pragma unroll
      for (int a = 0; a < D; ++a) xn[a] = g.origin[a] + (float)ids[a] * g.cell_size;
      float q[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < D; ++r) {
        float sacc = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) sacc += bc.f[9 + a * 3 + r] * (xn[a] - bc.f[3 + a]);
        q[r] = sacc;
      }
      float d[3], out2 = 0.f, inmax = -3.0e38f;
      int amax = 0;
#pragma unroll
      for (int r = 0; r < D; ++r) {
        d[r] = fabsf(q[r]) - bc.f[r];
        if (d[r] > 0.f) out2 += d[r] * d[r];
        if (d[r] > inmax) { inmax = d[r]; amax = r; }
      }
      float gl[3] = {0.f, 0.f, 0.f}, phi;
      if (out2 > 0.f) {
        const float o = sqrtf(out2);
#pragma unroll
        for (int r = 0; r < D; ++r) gl[r] = (d[r] > 0.f ? d[r] : 0.f) * (q[r] >= 0.f ? 1.f : -1.f) / o;
        phi = o;
      } else {
#pragma unroll
        for (int r = 0; r < D; ++r) gl[r] = (r == amax) ? (q[r] >= 0.f ? 1.f : -1.f) : 0.f;
        phi = inmax;
      }
      if (!(phi < bc.f[18])) continue;
      float nrm[3];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        float sacc = 0.f;
#pragma unroll
        for (int r = 0; r < D; ++r) sacc += bc.f[9 + a * 3 + r] * gl[r];
        nrm[a] = -sacc;
      }
      float rc = 0.f, rct = 0.f;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        rc += (P[a] / m - bc.f[6 + a]) * nrm[a];
        rct += (Q[a] / m - bc.f[6 + a]) * nrm[a];
      }
      if (rc >= 0.f) {
#pragma unroll
        for (int a = 0; a < D; ++a) P[a] = bc.f[6 + a] * m;
      }
      if (rct >= 0.f) {
#pragma unroll
        for (int a = 0; a < D; ++a) Q[a] = bc.f[6 + a] * m;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < D; ++a) {
    nodes.mom[a * nn + i] = P[a];
    nodes.mom_nt[a * nn + i] = clear_mom_nt ? 0.f : Q[a];
  }
  if (clamp) clamp[i] = (uint8_t)cbits;
  const bool live = m > prm.small_mass_cutoff;
  float vnt[3] = {0.f, 0.f, 0.f}, dv[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    vnt[a] = live ? Q[a] / m : 0.f;
    const float v = live ? P[a] / m : 0.f;
    dv[a] = vnt[a] - v;
  }
  store_node_vel<D>(nodes.vel, i, vnt, dv);
}

// MODE 0: USL G2P (usl_inline.h:245-332) + return map.
// MODE 2: MUSL second map: L, F, V from the re-mapped velocities + return map
//         (x, v of slot j of `out` were written by the first map, which
//         runs fused with the re-map: p2g_kernel MODE 1).
// 3D gathers read node rows staged in LDS per workgroup (see
// g2p_gather_staged); 1D/2D gathers read the node records through L2.
constexpr int kG2PThreads = 256;
// LDS node staging of the 3D G2P. A workgroup's 256 particles are
// consecutive in cell order, so they fall into a few runs: consecutive
// particles with the same y and z bins whose x bins do not decrease and stay
// within one 64-aligned block. A run's windows lie in 4 x 4 node rows (y, z
// bins - 1 .. + 2) over its x range - 1 .. + 2: at most 16 x 67 nodes. The
// runs are staged in order, as many per round as kG2PStageNodes holds (one
// round for the usual ten cells of one or two x rows at 27 particles per
// cell), and every particle gathers from LDS. A staged node is the clamped
// grid node its slot names: entries outside the grid have zero weight and
// read a finite value.
constexpr int kG2PStageNodes = 1088;  // one run of the widest x range (16 x 67 nodes) always fits
static_assert(kG2PStageNodes >= 16 * (64 + 3), "a run spans at most 64 x bins");
struct G2PRunTable {
  int y[kG2PThreads], z[kG2PThreads], x0[kG2PThreads], x1[kG2PThreads], base[kG2PThreads];
  int wave_sum[kG2PThreads / 64];
};
template <bool WANT_DV> constexpr int g2p_stage_bytes() {
  return kG2PStageNodes * (16 + (WANT_DV ? 8 : 0)) + (int)sizeof(G2PRunTable);
}

// Inclusive sum of v over the workgroup's threads in order; the total goes to
// *total. Synchronises the workgroup (wave_sum is rewritten on every call: a
// workgroup barrier must separate two calls).
__device__ inline int block_inclusive_sum(int v, int *wave_sum, int *total) {
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) wave_sum[wv] = v;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kG2PThreads / 64; ++k) {
    const int c = wave_sum[k];
    before += k < wv ? c : 0;
    all += c;
  }
  *total = all;
  return v + before;
}

// Gathers v_nt (and dv with WANT_DV, L with WANT_L) of the particles of lanes
// 0 .. nv - 1 (weights w) through the LDS staging `lds`
// (g2p_stage_bytes<WANT_DV>() bytes). Every thread of the workgroup calls it.
template <int D, int SF, bool WANT_DV, bool WANT_L = true>
__device__ inline void g2p_gather_staged(const pmpm_grid &g, const pmpm_nodes &nodes, const PW<D, SF> &w, int nv,
                                         char *lds, float *vinc, float *dvinc, float *L) {
  static_assert(D == 3, "3D staging");
  constexpr int CAP = kG2PStageNodes;
  float4 *box4 = reinterpret_cast<float4 *>(lds);
  float2 *box2 = reinterpret_cast<float2 *>(lds + CAP * 16);
  G2PRunTable &R = *reinterpret_cast<G2PRunTable *>(lds + CAP * (16 + (WANT_DV ? 8 : 0)));
  int *rb = reinterpret_cast<int *>(lds);  // lane bins [3][256], in box4's space before staging
  const int t = (int)threadIdx.x;
  const bool valid = t < nv;
  if (valid) {
#pragma unroll
    for (int a = 0; a < 3; ++a) rb[a * kG2PThreads + t] = w.bin[a];
  }
  __syncthreads();
  // does a run start at lane q (after lane q - 1)?
  auto starts_at = [&](int q) {
    const int p0 = rb[q - 1], q0 = rb[q];
    return rb[kG2PThreads + q - 1] != rb[kG2PThreads + q] || rb[2 * kG2PThreads + q - 1] != rb[2 * kG2PThreads + q] ||
           q0 < p0 || (q0 >> 6) != (p0 >> 6);
  };
  bool start = false, end = false;
  if (valid) {
    start = t == 0 || starts_at(t);
    end = t == nv - 1 || starts_at(t + 1);
  }
  int nruns = 0;
  const int k = block_inclusive_sum(start ? 1 : 0, R.wave_sum, &nruns) - 1;
  if (start) {
    R.y[k] = w.bin[1];
    R.z[k] = w.bin[2];
    R.x0[k] = w.bin[0];
  }
  if (end) R.x1[k] = w.bin[0];
  __syncthreads();
  const int cnt = t < nruns ? 16 * (R.x1[t] - R.x0[t] + 4) : 0;
  int all = 0;
  const int incl = block_inclusive_sum(cnt, R.wave_sum, &all);
  if (t < nruns) R.base[t] = incl - cnt;
  __syncthreads();
  const int myk = valid ? k : -1;
  for (int k0 = 0; k0 < nruns;) {
    const int b0 = R.base[k0];
    int k1 = k0 + 1;
    while (k1 < nruns && R.base[k1] + 16 * (R.x1[k1] - R.x0[k1] + 4) - b0 <= CAP) ++k1;
    const int tot = R.base[k1 - 1] + 16 * (R.x1[k1 - 1] - R.x0[k1 - 1] + 4) - b0;
    for (int q = t; q < tot; q += kG2PThreads) {
      int kk = k0;
      while (kk + 1 < k1 && R.base[kk + 1] - b0 <= q) ++kk;
      const int X = R.x1[kk] - R.x0[kk] + 4, qq = q - (R.base[kk] - b0);
      const int row = qq / X, x = qq - row * X;  // row = sz * 4 + sy
      const int ix = min(max(R.x0[kk] - 1 + x, 0), g.num_cells[0] - 1);
      const int iy = min(max(R.y[kk] - 1 + (row & 3), 0), g.num_cells[1] - 1);
      const int iz = min(max(R.z[kk] - 1 + (row >> 2), 0), g.num_cells[2] - 1);
      const uint32_t node = (uint32_t)ix + (uint32_t)g.num_cells[0] * ((uint32_t)iy + (uint32_t)g.num_cells[1] * (uint32_t)iz);
      const float4 *rec = reinterpret_cast<const float4 *>(nodes.vel + 8 * (size_t)node);
      box4[q] = rec[0];
      if (WANT_DV) {
        const float4 b = rec[1];
        box2[q] = make_float2(b.x, b.y);
      }
    }
    __syncthreads();
    if (myk >= k0 && myk < k1) {
      const int off = R.base[myk] - b0, X = R.x1[myk] - R.x0[myk] + 4;
      NodeSrc src{nodes.vel, box4 + off, box2 + off, {1, X, 4 * X}, {R.x0[myk] - 1, 0, 0}};
      gather_grid<D, SF, WANT_DV, WANT_L, true>(g, src, w, vinc, dvinc, L);
    }
    k0 = k1;
    __syncthreads();  // the next round overwrites the staging
  }
}

template <int D, int SF, int MODE>
__global__ __launch_bounds__(kG2PThreads) void g2p_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles in,
                                                          pmpm_particles out, pmpm_spatial part, pmpm_nodes nodes,
                                                          int reorder) {
  static_assert(MODE >= 0 && MODE <= 2, "G2P modes: USL (0), MUSL first map (1), MUSL second map (2)");
  constexpr bool kV = MODE != 2, kL = MODE != 1;  // what the gather sums
  const int n = in.n;
  const int j = blockIdx.x * kG2PThreads + threadIdx.x;
  const bool valid = j < n;
  if (D != 3 && !valid) return;  // the 3D workgroup synchronises below
  const int nn = g.num_cells_total;
  const int src = valid ? part.sorted[j] : 0;
  const int dst = reorder ? j : src;
  PW<D, SF> w;
  float xp[3];
  if (valid) particle_weights<D, SF>(prm, g, in.x, n, src, w, xp);
  float vinc[3] = {0.f, 0.f, 0.f}, dvinc[3] = {0.f, 0.f, 0.f};
  float L[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if constexpr (D == 3) {
    __shared__ __attribute__((aligned(16))) char lds[g2p_stage_bytes<kV>()];
    const int nv = min(n - (int)blockIdx.x * kG2PThreads, kG2PThreads);
    g2p_gather_staged<D, SF, kV, kL>(g, nodes, w, nv, lds, vinc, dvinc, L);
    if (!valid) return;
  } else {
    gather_grid<D, SF, kV, kL>(g, node_src(g, nodes), w, vinc, dvinc, L);
  }
  if (MODE != 2) {
    // MODE 1 writes slot j of `out` (cell order), as the second map reads it
    const int to = MODE == 1 ? j : dst;
    const float am = (float)(1.0 - (double)prm.alpha);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      out.v[a * n + to] = am * vinc[a] + prm.alpha * (in.v[a * n + src] + dvinc[a]);
      out.x[a * n + to] = xp[a] + prm.dt * vinc[a];
    }
    if (MODE == 1) return;
  }
  // F = (I + L dt) F, V = det(F) V0
  float Fo[9], Fn[9];
#pragma unroll
  for (int k = 0; k < D * D; ++k) Fo[k] = in.F[k * n + src];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) acc += ((a == k ? 1.f : 0.f) + L[3 * a + k] * prm.dt) * Fo[k * D + b];
      Fn[a * D + b] = acc;
    }
  const float vol0 = in.volume0[src];
  const float vol = det_dd<D>(Fn) * vol0;
#pragma unroll
  for (int k = 0; k < D * D; ++k) out.F[k * n + dst] = Fn[k];
  out.volume[dst] = vol;
  if (out.L) {
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) out.L[(a * D + b) * n + dst] = L[3 * a + b];
  }
  // return map of the particle's material
  const int color = in.color[src];
  const uint8_t pf = in.flags[src];
  const float mass = in.mass[src];
  Sym s = load_sym(in.stress, n, src);
  Sym eps = in.eps_e ? load_sym(in.eps_e, n, src) : sym_zero();
  Sym sref = sym_zero();
  if (in.stress_ref && color < prm.num_materials && (prm.materials[color].flags & PMPM_MATF_HAS_STRESS_REF))
    sref = load_sym(in.stress_ref, n, src);
  return_map<D>(prm, color, pf, L, mass, vol, vol0, sref, s, eps);
  store_sym(out.stress, n, dst, s);
  if (out.eps_e) store_sym(out.eps_e, n, dst, eps);
  if (reorder) {
    out.mass[dst] = mass;
    out.volume0[dst] = vol0;
    out.color[dst] = (uint8_t)color;
    out.flags[dst] = pf;
    out.id[dst] = in.id[src];
    // reference stresses only matter (and only travel) when a material reads them
    bool any_ref = false;
    for (int k = 0; k < prm.num_materials; ++k) any_ref = any_ref || (prm.materials[k].flags & PMPM_MATF_HAS_STRESS_REF);
    if (any_ref && in.stress_ref && out.stress_ref) store_sym(out.stress_ref, n, dst, load_sym(in.stress_ref, n, src));
  }
}

// The return maps of the materials that run after G2P (von Mises,
// Mohr-Coulomb, modified Cam-Clay), on particle j of `out` with the L, F, V,
// stress and elastic strain G2P wrote there; the history scalars follow the
// particles into cell order here (from slot sorted[j] of `in_hist` when G2P
// reordered, from slot j in place), for every particle.
template <int D>
__global__ __launch_bounds__(256) void ext_material_kernel(pmpm_params prm, const float *__restrict__ in_hist,
                                                           pmpm_particles out, const int32_t *__restrict__ sorted,
                                                           int reorder) {
  const int n = out.n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int src = reorder ? sorted[j] : j;
  float h0 = in_hist[src], h1 = in_hist[n + src];
  const int color = out.color[j];
  if (color < prm.num_materials && is_ext_material(prm.materials[color].kind)) {
    float L[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) L[3 * a + b] = out.L[(a * D + b) * n + j];
    Sym s = load_sym(out.stress, n, j), eps = load_sym(out.eps_e, n, j);
    Sym sref = sym_zero();
    if (out.stress_ref && (prm.materials[color].flags & PMPM_MATF_HAS_STRESS_REF)) sref = load_sym(out.stress_ref, n, j);
    return_map_ext<D>(prm, color, L, out.volume[j], sref, s, eps, h0, h1);
    store_sym(out.stress, n, j, s);
    store_sym(out.eps_e, n, j, eps);
  }
  out.hist[j] = h0;
  out.hist[n + j] = h1;
}

// Stand-alone return map of one material from the stored velocity gradient.
template <int D>
__global__ void stress_update_kernel(pmpm_params prm, int mat_id, pmpm_particles p) {
  const int n = p.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int color = p.color[i];
  if (color != mat_id) return;
  float L[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) L[3 * a + b] = p.L[(a * D + b) * n + i];
  Sym s = load_sym(p.stress, n, i);
  Sym eps = p.eps_e ? load_sym(p.eps_e, n, i) : sym_zero();
  Sym sref = sym_zero();
  if (p.stress_ref && (prm.materials[color].flags & PMPM_MATF_HAS_STRESS_REF)) sref = load_sym(p.stress_ref, n, i);
  return_map<D>(prm, color, p.flags[i], L, p.mass[i], p.volume[i], p.volume0[i], sref, s, eps);
  if (is_ext_material(prm.materials[color].kind)) {
    float h0 = p.hist[i], h1 = p.hist[n + i];
    return_map_ext<D>(prm, color, L, p.volume[i], sref, s, eps, h0, h1);
    p.hist[i] = h0;
    p.hist[n + i] = h1;
  }
  store_sym(p.stress, n, i, s);
  if (p.eps_e) store_sym(p.eps_e, n, i, eps);
}

// Materialised psi/dpsi in the reference's layout (shapefunction.cpp:70-142):
// window entries in the reference's table order, entries whose linear node
// index is out of range stay 0, node types of the (linear) node index.
template <int D, int SF>
__global__ void shape_function_kernel(pmpm_grid g, int types_built, const float *x, int n, float *psi, float *dpsi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int W = (SF == PMPM_SF_LINEAR) ? 2 : 4;
  constexpr int LO = (SF == PMPM_SF_LINEAR) ? 0 : -1;
  constexpr int NSN = (D == 1) ? W : (D == 2 ? W * W : W * W * W);
  float t[3];
  int bin[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    t[a] = cell_coord(x[a * n + i], g.origin[a], g.inv_cell_size);
    bin[a] = (int)t[a];
  }
  for (int k = 0; k < NSN; ++k) {
    int off[3] = {0, 0, 0};
    if (D == 1) off[0] = LO + k;
    else if (D == 2) { off[0] = LO + k % W; off[1] = LO + k / W; }
    else { off[2] = LO + k % W; off[0] = LO + (k / W) % W; off[1] = LO + k / (W * W); }
    int nb[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) nb[a] = bin[a] + off[a];
    const int64_t lin = lin_index(g, nb, D);
    float wgt = 0.f, gr[3] = {0.f, 0.f, 0.f};
    if (lin >= 0 && lin < g.num_cells_total) {
      int ids[3];
      node_ids<D>(g, lin, ids);
      float N[3] = {1.f, 1.f, 1.f}, dN[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const float frac = __fsub_rn(t[a], (float)bin[a]);
        spline_eval<SF>(__fsub_rn(frac, (float)off[a]), cubic_node_type(ids[a], g.num_cells[a], types_built),
                        g.inv_cell_size, N[a], dN[a]);
      }
      if (D == 1) { wgt = N[0]; gr[0] = dN[0]; }
      else if (D == 2) { wgt = N[0] * N[1]; gr[0] = dN[0] * N[1]; gr[1] = dN[1] * N[0]; }
      else {
        wgt = N[0] * N[1] * N[2];
        gr[0] = dN[0] * N[1] * N[2];
        gr[1] = dN[1] * N[0] * N[2];
        gr[2] = dN[2] * N[0] * N[1];
      }
    }
    psi[(int64_t)i * NSN + k] = wgt;
#pragma unroll
    for (int a = 0; a < D; ++a) dpsi[((int64_t)i * NSN + k) * D + a] = gr[a];
  }
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
#define PMPM_DISPATCH_DSF(dim, sf, FN, ...)                                       \
  switch ((dim) * 3 + (sf)) {                                                     \
    case 3: FN<1, PMPM_SF_LINEAR>(__VA_ARGS__); break;                            \
    case 4: FN<1, PMPM_SF_QUADRATIC>(__VA_ARGS__); break;                         \
    case 5: FN<1, PMPM_SF_CUBIC>(__VA_ARGS__); break;                             \
    case 6: FN<2, PMPM_SF_LINEAR>(__VA_ARGS__); break;                            \
    case 7: FN<2, PMPM_SF_QUADRATIC>(__VA_ARGS__); break;                         \
    case 8: FN<2, PMPM_SF_CUBIC>(__VA_ARGS__); break;                             \
    case 9: FN<3, PMPM_SF_LINEAR>(__VA_ARGS__); break;                            \
    case 10: FN<3, PMPM_SF_QUADRATIC>(__VA_ARGS__); break;                        \
    case 11: FN<3, PMPM_SF_CUBIC>(__VA_ARGS__); break;                            \
    default: return PMPM_ERR_INVALID;                                             \
  }

// How the wrap-around deposits of a P2G pass are summed (TileGeom::wrap).
static int wrap_mode(const pmpm_params &prm, const pmpm_grid &g) {
  if (!prm.wrap_quirk || prm.dim < 2) return 0;
  if (prm.shape_fn == PMPM_SF_CUBIC) {
    if (!prm.node_types_built) return 0;  // untyped cubic weights are zero
    return (prm.dim == 2 ? wrap_cell_count<2>(g) : wrap_cell_count<3>(g)) > 0 ? 2 : 0;
  }
  return 1;
}

template <int D, int SF> TileGeom tile_geom_of(const pmpm_grid &g, int gather) {
  using G = P2GGeom<D, SF>;
  TileGeom tg{};
  const int T[3] = {G::TX, G::TY, G::TZ}, N[3] = {G::NX, G::NY, G::NZ};
  tg.nlo = G::NLO;
  for (int a = 0; a < 3; ++a) {
    tg.T[a] = a < D ? T[a] : 1;
    tg.N[a] = a < D ? N[a] : 1;
    tg.tiles[a] = a < D ? (int)ceil_div(g.num_cells[a], T[a]) : 1;
  }
  tg.gather = gather;
  tg.nv = gather == 2 ? D : (gather == 3 ? 2 * D : 1 + 2 * D);  // 3: adjoint scatter
  return tg;
}
TileGeom tile_geom(int dim, int sf, const pmpm_grid &g, int gather) {
  switch (dim * 3 + sf) {
    case 3: return tile_geom_of<1, PMPM_SF_LINEAR>(g, gather);
    case 4: return tile_geom_of<1, PMPM_SF_QUADRATIC>(g, gather);
    case 5: return tile_geom_of<1, PMPM_SF_CUBIC>(g, gather);
    case 6: return tile_geom_of<2, PMPM_SF_LINEAR>(g, gather);
    case 7: return tile_geom_of<2, PMPM_SF_QUADRATIC>(g, gather);
    case 8: return tile_geom_of<2, PMPM_SF_CUBIC>(g, gather);
    case 9: return tile_geom_of<3, PMPM_SF_LINEAR>(g, gather);
    case 10: return tile_geom_of<3, PMPM_SF_QUADRATIC>(g, gather);
    default: return tile_geom_of<3, PMPM_SF_CUBIC>(g, gather);
  }
}
static size_t tile_count(const TileGeom &tg) { return (size_t)tg.tiles[0] * tg.tiles[1] * tg.tiles[2]; }
// The forward P2G and MUSL re-map of 3D typed cubic run on the matrix cores
// with their own tiles (MfmaGeom); every other pass uses p2g_kernel's.
static bool mfma_tiles(const pmpm_params &prm, int gather) {
  return prm.dim == 3 && prm.shape_fn == PMPM_SF_CUBIC && prm.node_types_built && (gather == 1 || gather == 2);
}
static TileGeom mfma_tile_geom(const pmpm_grid &g, int gather) {
  using M = MfmaGeom;
  TileGeom tg{};
  const int T[3] = {M::CELLS, 1, M::TZ}, N[3] = {M::BX, M::BY, M::NZ};
  tg.nlo = -1;
  for (int a = 0; a < 3; ++a) {
    tg.T[a] = T[a];
    tg.N[a] = N[a];
    tg.tiles[a] = (int)ceil_div(g.num_cells[a], T[a]);
  }
  tg.gather = gather;
  tg.nv = gather == 2 ? 3 : 7;
  return tg;
}
static TileGeom tile_geom(const pmpm_params &prm, const pmpm_grid &g, int gather) {
  return mfma_tiles(prm, gather) ? mfma_tile_geom(g, gather) : tile_geom(prm.dim, prm.shape_fn, g, gather);
}

// mode 0: P2G, 1: MUSL first map + re-map, 2: adjoint scatter of the
// records `rec`
template <int D, int SF>
void launch_p2g(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
                const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, float *partials, int32_t *flags,
                float *wrec, uint8_t *wtouch, int32_t *wlist, hipStream_t s,
                const float *rec = nullptr) {
  const TileGeom tg = tile_geom_of<D, SF>(g, 1);
  const dim3 grid((unsigned)tile_count(tg)), block(P2GGeom<D, SF>::THREADS);
  // 3D typed cubic P2G and MUSL re-map on the matrix cores (the re-map after
  // a G2P-style first map), on their own tiles; everything else on the VALU
  // kernel
  bool mfma = false;
  if constexpr (D == 3 && SF == PMPM_SF_CUBIC) {
    mfma = (mode == 0 || mode == 1) && prm.node_types_built;
    const TileGeom mg = mfma_tile_geom(g, mode + 1);
    const dim3 grid((unsigned)tile_count(mg));
    if (mfma && mode == 0 && has_planar(prm)) {
      PMPM_LAUNCH((p2g_mfma_kernel<SF, false, true>),
                  (::pmpm::kname("p2g_mfma_kernel", {(long)(SF), (long)(false), (long)(true)})), grid, dim3(256), 0, s,
                  prm, g, in, out, part, nodes, mg.tiles[0], mg.tiles[1], partials, flags);
    } else if (mfma && mode == 0) {
      PMPM_LAUNCH((p2g_mfma_kernel<SF, false>), (::pmpm::kname("p2g_mfma_kernel", {(long)(SF), (long)(false)})), grid, dim3(256), 0, s, prm, g, in, out, part, nodes,
                         mg.tiles[0], mg.tiles[1], partials, flags);
    } else if (mfma && mode == 1) {
      if (in.n > 0)
        PMPM_LAUNCH((g2p_kernel<D, SF, 1>), (::pmpm::kname("g2p_kernel", {(long)(D), (long)(SF), (long)(1)})), dim3((unsigned)ceil_div(in.n, kG2PThreads)), dim3(kG2PThreads), 0,
                           s, prm, g, in, out, part, nodes, 1);
      PMPM_LAUNCH((p2g_mfma_kernel<SF, true>), (::pmpm::kname("p2g_mfma_kernel", {(long)(SF), (long)(true)})), grid, dim3(256), 0, s, prm, g, in, out, part, nodes,
                         mg.tiles[0], mg.tiles[1], partials, flags);
    }
  }
  // (each mode launches exactly one tile kernel; the adjoint scatter reads
  // `rec` and is never launched without it)
  if (mode == 0) {
    if (!mfma)
      PMPM_LAUNCH((p2g_kernel<D, SF, 0>), (::pmpm::kname("p2g_kernel", {(long)(D), (long)(SF), (long)(0)})), grid, block, 0, s, prm, g, in, out, part, nodes, tg.tiles[0],
                         tg.tiles[1], partials, flags, rec);
  } else if (mode == 1) {
    if (!mfma)
      PMPM_LAUNCH((p2g_kernel<D, SF, 1>), (::pmpm::kname("p2g_kernel", {(long)(D), (long)(SF), (long)(1)})), dim3((unsigned)tile_count(tg) + kMuslTailBlocks), block, 0, s, prm,
                         g, in, out, part, nodes, tg.tiles[0], tg.tiles[1], partials, flags, rec);
  } else if (mode == 2 && rec) {
    PMPM_LAUNCH((p2g_kernel<D, SF, 2>), (::pmpm::kname("p2g_kernel", {(long)(D), (long)(SF), (long)(2)})), grid, block, 0, s, prm, g, in, out, part, nodes, tg.tiles[0],
                       tg.tiles[1], partials, flags, rec);
  }
  const int wcells = wrap_cell_count<D>(g);
  if (mode != 2 && wrap_mode(prm, g) == 2 && wcells > 0) {
    const dim3 wgrid((unsigned)ceil_div(wcells, kWrapCellsPerBlock)), wblock(64 * kWrapCellsPerBlock);
    if (mode == 0)
      PMPM_LAUNCH((p2g_wrap_kernel<D, 0>), (::pmpm::kname("p2g_wrap_kernel", {(long)(D), (long)(0)})), wgrid, wblock,
                  0, s, prm, g, in, out, part, wrec, wtouch, wlist, wcells);
    else
      PMPM_LAUNCH((p2g_wrap_kernel<D, 1>), (::pmpm::kname("p2g_wrap_kernel", {(long)(D), (long)(1)})), wgrid, wblock,
                  0, s, prm, g, in, out, part, wrec, wtouch, wlist, wcells);
  }
}
template <int D, int SF>
void launch_g2p(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
                const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, int reorder, hipStream_t s) {
  const unsigned blocks = (unsigned)ceil_div(in.n, 256);
  if (mode == 0)
    PMPM_LAUNCH((g2p_kernel<D, SF, 0>), (::pmpm::kname("g2p_kernel", {(long)(D), (long)(SF), (long)(0)})), dim3(blocks), dim3(256), 0, s, prm, g, in, out, part, nodes, reorder);
  else
    PMPM_LAUNCH((g2p_kernel<D, SF, 2>), (::pmpm::kname("g2p_kernel", {(long)(D), (long)(SF), (long)(2)})), dim3(blocks), dim3(256), 0, s, prm, g, in, out, part, nodes, reorder);
}
template <int D, int SF>
void launch_shape(const pmpm_grid &g, int types_built, const float *x, int n, float *psi, float *dpsi, hipStream_t s) {
  PMPM_LAUNCH((shape_function_kernel<D, SF>), (::pmpm::kname("shape_function_kernel", {(long)(D), (long)(SF)})), dim3((unsigned)ceil_div(n, 128)), dim3(128), 0, s, g, types_built,
                     x, n, psi, dpsi);
}

// gather: 0 = node arrays only, 1 = add the P2G tile partials, 2 = add the
// MUSL re-map tile partials (ctx holds them).
int launch_grid(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes, int finalize,
                int b0, int b1, int remap, int gather, hipStream_t s, int clear_mom_nt = 0) {
  const unsigned blocks = (unsigned)ceil_div(g.num_cells_total, 256);
  TileGeom tg = tile_geom(prm, g, gather);
  const float *partials = gather ? ctx->partials : nullptr;
  const int32_t *flags = gather ? ctx->tile_flags : nullptr;
  // the clamp record of a pass that applies node conditions (not the MUSL
  // re-map pass), for pmpm_adjoint_step
  int32_t *wl = (ctx && gather) ? ctx->wrap_list : nullptr;
  uint8_t *clamp = nullptr;
  if (!remap && b1 > b0)
    clamp = nodes.clamp ? nodes.clamp : (ctx ? reinterpret_cast<uint8_t *>(ctx->clamp_mask) : nullptr);

  switch (prm.dim) {
    case 1: PMPM_LAUNCH(grid_kernel<1>, (::pmpm::kname("grid_kernel", {(long)(1)}) + " gather " + std::to_string(gather)), dim3(blocks), dim3(256), 0, s, prm, g, nodes, finalize, b0, b1, remap, tg, partials, flags, clear_mom_nt, clamp, wl); break;
    case 2: PMPM_LAUNCH(grid_kernel<2>, (::pmpm::kname("grid_kernel", {(long)(2)}) + " gather " + std::to_string(gather)), dim3(blocks), dim3(256), 0, s, prm, g, nodes, finalize, b0, b1, remap, tg, partials, flags, clear_mom_nt, clamp, wl); break;
    case 3: PMPM_LAUNCH(grid_kernel<3>, (::pmpm::kname("grid_kernel", {(long)(3)}) + " gather " + std::to_string(gather)), dim3(blocks), dim3(256), 0, s, prm, g, nodes, finalize, b0, b1, remap, tg, partials, flags, clear_mom_nt, clamp, wl); break;
    default: return PMPM_ERR_INVALID;
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

bool params_ok(const pmpm_params *prm) {
  if (!(prm && prm->dim >= 1 && prm->dim <= 3 && prm->shape_fn >= 0 && prm->shape_fn <= 2 &&
        prm->num_materials >= 0 && prm->num_materials <= PMPM_MAX_MATERIALS && prm->num_bcs >= 0 &&
        prm->num_bcs <= PMPM_MAX_BCS))
    return false;
  for (int k = 0; k < prm->num_bcs; ++k) {  // a body force reads its node arrays
    const pmpm_bc &bc = prm->bcs[k];
    if (bc.kind == PMPM_BC_BODY_FORCE && (!bc.values || !bc.mask || bc.i[0] < 0 || bc.i[0] > 2)) return false;
  }
  return true;
}
static bool particles_ok(const pmpm_particles *p) {
  return p && p->n >= 0 &&
         (p->n == 0 || (p->x && p->v && p->stress && p->F && p->mass && p->volume && p->volume0 && p->color &&
                        p->flags && p->id));
}

int do_zero_nodes(const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes, hipStream_t s) {
  int rc;
  const int64_t nn = g.num_cells_total;
  ZeroLists z{};
  z.add(nodes.mass, nn);
  z.add(nodes.mom, nn * prm.dim);
  z.add(nodes.mom_nt, nn * prm.dim);
  if ((rc = zero_fill_lists(z, s)) != PMPM_OK) return rc;
  return PMPM_OK;
}

// Partials and tile flags for the larger of the two tilings a step can use
// (p2g_kernel's; the matrix-core passes' own).
static void partial_sizes(const pmpm_params &prm, const pmpm_grid &g, size_t *floats, size_t *tiles) {
  const TileGeom tg = tile_geom(prm.dim, prm.shape_fn, g, 1);
  *tiles = tile_count(tg);
  *floats = *tiles * tg.nv * (size_t)tg.N[0] * tg.N[1] * tg.N[2];
  if (mfma_tiles(prm, 1)) {
    const TileGeom mg = mfma_tile_geom(g, 1);
    const size_t mt = tile_count(mg);
    *tiles = std::max(*tiles, mt);
    *floats = std::max(*floats, mt * mg.nv * (size_t)mg.N[0] * mg.N[1] * mg.N[2]);
  }
}
// Words of the wrap-around node list: the part counts, then kWrapListShards
// parts of wrap_target_count entries each.
static size_t wrap_node_list_words(const pmpm_params &prm, const pmpm_grid &g) {
  const size_t cap = (size_t)(prm.dim == 3 ? wrap_target_count<3>(g) : (prm.dim == 2 ? wrap_target_count<2>(g) : 0));
  return (size_t)kWrapListShards * (1 + cap);
}
static int reserve_tiles(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g) {
  size_t floats = 0, tiles = 0;
  partial_sizes(prm, g, &floats, &tiles);
  int rc = ensure_partials(ctx, floats, tiles);
  if (rc != PMPM_OK) return rc;
  // the grid pass's clamp record (one byte per node, in words of 4)
  if ((rc = ensure_buffer(ctx->clamp_mask, ctx->clamp_mask_cap, ((size_t)g.num_cells_total + 3) / 4)) != PMPM_OK)
    return rc;
  if (wrap_mode(prm, g) != 2) return rc;
  const size_t wcells = (size_t)(prm.dim == 2 ? wrap_cell_count<2>(g) : wrap_cell_count<3>(g));
  if ((rc = ensure_buffer(ctx->wrap_rec, ctx->wrap_rec_cap, (size_t)wrap_rec_stride(1 + 2 * prm.dim) * wcells * 64)) != PMPM_OK)
    return rc;
  // node flags, one byte per node, cleared once here (the gather clears
  // what it visits; a stale flag would only cost the gather a visit)
  const size_t words = ((size_t)g.num_cells_total + 3) / 4;
  if (words > ctx->wrap_touch_cap) {
    if ((rc = ensure_buffer(ctx->wrap_touch, ctx->wrap_touch_cap, words)) != PMPM_OK) return rc;
    PMPM_HIP_CHECK(hipMemset(ctx->wrap_touch, 0, words * sizeof(float)));
  }
  // the gather's list: kWrapListShards counts, then as many parts of room
  // for every node the wrap-around can reach
  const size_t lwords = wrap_node_list_words(prm, g);
  if (lwords > ctx->wrap_list_cap) {
    float *lb = reinterpret_cast<float *>(ctx->wrap_list);
    size_t cap = ctx->wrap_list_cap;
    if ((rc = ensure_buffer(lb, cap, lwords)) != PMPM_OK) return rc;
    ctx->wrap_list = reinterpret_cast<int32_t *>(lb);
    ctx->wrap_list_cap = cap;
    PMPM_HIP_CHECK(hipMemset(ctx->wrap_list, 0, kWrapListShards * sizeof(int32_t)));
  }
  return PMPM_OK;
}
static uint8_t *wrap_touch(pmpm_context *ctx) { return reinterpret_cast<uint8_t *>(ctx->wrap_touch); }

void p2g_workspace(const pmpm_params &prm, const pmpm_grid &g, size_t *partial_floats, size_t *tiles,
                   size_t *wrap_floats, size_t *wrap_cells, size_t *wrap_touch_words, size_t *wrap_list_words,
                   size_t *clamp_words) {
  partial_sizes(prm, g, partial_floats, tiles);
  *wrap_cells = (size_t)(prm.dim == 3 ? wrap_cell_count<3>(g) : (prm.dim == 2 ? wrap_cell_count<2>(g) : 0));
  *wrap_floats = wrap_mode(prm, g) == 2 ? (size_t)wrap_rec_stride(1 + 2 * prm.dim) * *wrap_cells * 64 : 0;
  *wrap_touch_words = wrap_mode(prm, g) == 2 ? ((size_t)g.num_cells_total + 3) / 4 : 0;
  *wrap_list_words = wrap_mode(prm, g) == 2 ? wrap_node_list_words(prm, g) : 0;
  *clamp_words = ((size_t)g.num_cells_total + 3) / 4;
}

// The wrap-around sums of a P2G pass (gather 1) or MUSL re-map (gather 2)
// into the node arrays, after its tile kernel and wrap kernel.
static int launch_wrap_gather(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_spatial &part,
                              const pmpm_nodes &nodes, int gather, hipStream_t s) {
  const int mode = wrap_mode(prm, g);
  if (mode == 0) return PMPM_OK;
  TileGeom tg = tile_geom(prm, g, gather);
  tg.wrap = mode;
  tg.wcells = prm.dim == 2 ? wrap_cell_count<2>(g) : wrap_cell_count<3>(g);
  tg.wrec = ctx->wrap_rec;
  tg.cstart = part.cell_start;
  tg.cend = part.cell_end;
  const int targets = prm.dim == 2 ? wrap_target_count<2>(g) : wrap_target_count<3>(g);
  if (targets <= 0) return PMPM_OK;
  if (mode == 2) {  // typed cubic: the nodes the wrap kernel listed
    const dim3 lgrid((unsigned)ceil_div(targets, 4)), block(256);  // a wave per possible entry
    uint8_t *touch = wrap_touch(ctx);
    if (prm.dim == 2) {
      if (gather == 1)
        PMPM_LAUNCH((wrap_gather_list_kernel<2, 5>), (std::string("wrap_gather_list_kernel<2, 5>")), lgrid, block, 0,
                    s, g, tg, ctx->partials, ctx->tile_flags, nodes, touch, ctx->wrap_list);
      else
        PMPM_LAUNCH((wrap_gather_list_kernel<2, 2>), (std::string("wrap_gather_list_kernel<2, 2>")), lgrid, block, 0,
                    s, g, tg, ctx->partials, ctx->tile_flags, nodes, touch, ctx->wrap_list);
    } else {
      if (gather == 1)
        PMPM_LAUNCH((wrap_gather_list_kernel<3, 7>), (std::string("wrap_gather_list_kernel<3, 7>")), lgrid, block, 0,
                    s, g, tg, ctx->partials, ctx->tile_flags, nodes, touch, ctx->wrap_list);
      else
        PMPM_LAUNCH((wrap_gather_list_kernel<3, 3>), (std::string("wrap_gather_list_kernel<3, 3>")), lgrid, block, 0,
                    s, g, tg, ctx->partials, ctx->tile_flags, nodes, touch, ctx->wrap_list);
    }
    PMPM_LAUNCH_CHECK();
    return PMPM_OK;
  }
  uint8_t *touch = nullptr;
  const dim3 grid((unsigned)ceil_div(targets, 4)), block(256);  // one wave per target node
  if (prm.dim == 2) {
    if (gather == 1)
      PMPM_LAUNCH((wrap_gather_kernel<2, 5>), (std::string("wrap_gather_kernel<2, 5>")), grid, block, 0, s, g, tg,
                  ctx->partials, ctx->tile_flags, nodes, touch);
    else
      PMPM_LAUNCH((wrap_gather_kernel<2, 2>), (std::string("wrap_gather_kernel<2, 2>")), grid, block, 0, s, g, tg,
                  ctx->partials, ctx->tile_flags, nodes, touch);
  } else {
    if (gather == 1)
      PMPM_LAUNCH((wrap_gather_kernel<3, 7>), (std::string("wrap_gather_kernel<3, 7>")), grid, block, 0, s, g, tg,
                  ctx->partials, ctx->tile_flags, nodes, touch);
    else
      PMPM_LAUNCH((wrap_gather_kernel<3, 3>), (std::string("wrap_gather_kernel<3, 3>")), grid, block, 0, s, g, tg,
                  ctx->partials, ctx->tile_flags, nodes, touch);
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}


// Scatter into the tile partials (+ wrap-around deposits into the cleared
// node arrays); a grid_kernel pass with gather = 1 completes the P2G.
// (zeroed: the caller has already cleared mass, mom and mom_nt)
int do_p2g(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &p,
           const pmpm_spatial &part, const pmpm_nodes &nodes, hipStream_t s, bool zeroed = false) {
  int rc = reserve_tiles(ctx, prm, g);
  if (rc != PMPM_OK) return rc;
  if (!zeroed && (rc = do_zero_nodes(prm, g, nodes, s)) != PMPM_OK) return rc;
  PMPM_DISPATCH_DSF(prm.dim, prm.shape_fn, launch_p2g, prm, g, p, p, part, nodes, 0, ctx->partials, ctx->tile_flags,
                    ctx->wrap_rec, wrap_touch(ctx), ctx->wrap_list, s);
  PMPM_LAUNCH_CHECK();
  return launch_wrap_gather(ctx, prm, g, part, nodes, 1, s);
}

// True when a material of the list runs after G2P; then the particles must
// store L, the elastic strain and the history scalars.
static bool has_ext_material(const pmpm_params &prm) {
  for (int k = 0; k < prm.num_materials; ++k)
    if (is_ext_material(prm.materials[k].kind)) return true;
  return false;
}
static bool ext_buffers_ok(const pmpm_params &prm, const pmpm_particles &p) {
  return !has_ext_material(prm) || p.n == 0 || (p.L && p.eps_e && p.hist);
}

int do_g2p(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
           const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, hipStream_t s) {
  if (in.n == 0) return PMPM_OK;
  if (!ext_buffers_ok(prm, in) || !ext_buffers_ok(prm, out)) return PMPM_ERR_INVALID;
  const int reorder = (out.F != in.F) ? 1 : 0;  // in place when F is updated in place
  PMPM_DISPATCH_DSF(prm.dim, prm.shape_fn, launch_g2p, prm, g, in, out, part, nodes, mode, reorder, s);
  PMPM_LAUNCH_CHECK();
  if (has_ext_material(prm)) {
    const unsigned blocks = (unsigned)ceil_div(in.n, 256);
    switch (prm.dim) {
      case 1: PMPM_LAUNCH(ext_material_kernel<1>, (std::string("ext_material_kernel<1>")), dim3(blocks), dim3(256), 0, s, prm, in.hist, out, part.sorted, reorder); break;
      case 2: PMPM_LAUNCH(ext_material_kernel<2>, (std::string("ext_material_kernel<2>")), dim3(blocks), dim3(256), 0, s, prm, in.hist, out, part.sorted, reorder); break;
      default: PMPM_LAUNCH(ext_material_kernel<3>, (std::string("ext_material_kernel<3>")), dim3(blocks), dim3(256), 0, s, prm, in.hist, out, part.sorted, reorder); break;
    }
    PMPM_LAUNCH_CHECK();
  }
  return PMPM_OK;
}

// First map (x, v into `out` in cell order) + re-map of m v into the tile
// partials; a grid_kernel pass with gather = 2 completes mom_nt.
// (zeroed: mom_nt is already zero, left so by the grid pass before)
int do_musl_map(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in,
                const pmpm_particles &out, const pmpm_spatial &part, const pmpm_nodes &nodes, hipStream_t s,
                bool zeroed = false) {
  int rc = reserve_tiles(ctx, prm, g);
  if (rc != PMPM_OK) return rc;
  // the fused pass reads the node velocity records, never mom_nt: clear it
  // for the re-map's wrap-around atomics first
  if (!zeroed && (rc = zero_fill(nodes.mom_nt, (int64_t)g.num_cells_total * prm.dim, s)) != PMPM_OK) return rc;
  PMPM_DISPATCH_DSF(prm.dim, prm.shape_fn, launch_p2g, prm, g, in, out, part, nodes, 1, ctx->partials,
                    ctx->tile_flags, ctx->wrap_rec, wrap_touch(ctx), ctx->wrap_list, s);
  PMPM_LAUNCH_CHECK();
  return launch_wrap_gather(ctx, prm, g, part, nodes, 2, s);
}

}  // namespace pmpm

using namespace pmpm;

extern "C" int pmpm_shape_function(int32_t dim, int32_t shape_fn, int32_t node_types_built, const pmpm_grid *grid,
                                   const float *x, int32_t n, float *psi, float *dpsi, pmpm_stream_t stream) {
  if (!grid_ok(grid) || dim < 1 || dim > 3 || shape_fn < 0 || shape_fn > 2 || n < 0) return PMPM_ERR_INVALID;
  if (n == 0) return PMPM_OK;
  if (!x || !psi || !dpsi) return PMPM_ERR_INVALID;
  PMPM_DISPATCH_DSF(dim, shape_fn, launch_shape, *grid, node_types_built, x, n, psi, dpsi, (hipStream_t)stream);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

extern "C" int pmpm_p2g(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *p,
                        const pmpm_spatial *part, pmpm_nodes *nodes, pmpm_stream_t stream) {
  if (!ctx || !params_ok(prm) || !grid_ok(grid) || !particles_ok(p) || !part || !nodes) return PMPM_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  int rc = do_p2g(ctx, *prm, *grid, *p, *part, *nodes, s);
  if (rc != PMPM_OK) return rc;
  // the reference's P2G ends with the cutoff and mom_nt = mom + f dt
  return launch_grid(ctx, *prm, *grid, *nodes, 1, 0, 0, 0, 1, s);
}

extern "C" int pmpm_grid_bc(const pmpm_params *prm, const pmpm_grid *grid, pmpm_nodes *nodes, int32_t bc_begin,
                            int32_t bc_end, int32_t remap_pass, pmpm_stream_t stream) {
  if (!params_ok(prm) || !grid_ok(grid) || !nodes) return PMPM_ERR_INVALID;
  if (bc_begin < 0 || bc_end > prm->num_bcs || bc_begin > bc_end) return PMPM_ERR_INVALID;
  return launch_grid(nullptr, *prm, *grid, *nodes, 0, bc_begin, bc_end, remap_pass ? 1 : 0, 0, (hipStream_t)stream);
}

extern "C" int pmpm_g2p(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *in, pmpm_particles *out,
                        const pmpm_spatial *part, const pmpm_nodes *nodes, pmpm_stream_t stream) {
  if (!params_ok(prm) || !grid_ok(grid) || !particles_ok(in) || !particles_ok(out) || !part || !nodes) return PMPM_ERR_INVALID;
  if (out->n != in->n) return PMPM_ERR_INVALID;
  return do_g2p(*prm, *grid, *in, *out, *part, *nodes, 0, (hipStream_t)stream);
}

extern "C" int pmpm_musl_map(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid,
                             const pmpm_particles *in, pmpm_particles *out, const pmpm_spatial *part,
                             pmpm_nodes *nodes, pmpm_stream_t stream) {
  if (!ctx || !params_ok(prm) || !grid_ok(grid) || !particles_ok(in) || !particles_ok(out) || !part || !nodes)
    return PMPM_ERR_INVALID;
  if (out->n != in->n || out->x == in->x) return PMPM_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  int rc = do_musl_map(ctx, *prm, *grid, *in, *out, *part, *nodes, s);
  if (rc != PMPM_OK) return rc;
  // mom_nt := re-mapped momentum (mass cutoff as P2G), node velocities
  return launch_grid(ctx, *prm, *grid, *nodes, 0, 0, 0, 1, 2, s);
}

extern "C" int pmpm_musl_g2p(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *in,
                             pmpm_particles *out, const pmpm_spatial *part, const pmpm_nodes *nodes,
                             pmpm_stream_t stream) {
  if (!params_ok(prm) || !grid_ok(grid) || !particles_ok(in) || !particles_ok(out) || !part || !nodes) return PMPM_ERR_INVALID;
  if (out->n != in->n) return PMPM_ERR_INVALID;
  return do_g2p(*prm, *grid, *in, *out, *part, *nodes, 2, (hipStream_t)stream);
}

extern "C" int pmpm_stress_update(const pmpm_params *prm, int32_t mat_id, pmpm_particles *p, pmpm_stream_t stream) {
  if (!params_ok(prm) || !particles_ok(p) || !p || (p->n > 0 && !p->L)) return PMPM_ERR_INVALID;
  if (mat_id < 0 || mat_id >= prm->num_materials) return PMPM_ERR_INVALID;
  if (!ext_buffers_ok(*prm, *p)) return PMPM_ERR_INVALID;
  if (p->n == 0) return PMPM_OK;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = (unsigned)ceil_div(p->n, 256);
  switch (prm->dim) {
    case 1: PMPM_LAUNCH(stress_update_kernel<1>, (::pmpm::kname("stress_update_kernel", {(long)(1)})), dim3(blocks), dim3(256), 0, s, *prm, mat_id, *p); break;
    case 2: PMPM_LAUNCH(stress_update_kernel<2>, (::pmpm::kname("stress_update_kernel", {(long)(2)})), dim3(blocks), dim3(256), 0, s, *prm, mat_id, *p); break;
    default: PMPM_LAUNCH(stress_update_kernel<3>, (::pmpm::kname("stress_update_kernel", {(long)(3)})), dim3(blocks), dim3(256), 0, s, *prm, mat_id, *p); break;
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

// ===========================================================================
// Adjoint of one USL step.
//
// Forward (per step):  P2G   m_i = sum w m,  P_i = sum w m v,  f_i = -sum V sigma grad w
//                      grid  Q_i = P_i + dt f_i, BCs, vnt_i = Q_i / m_i, dv_i = vnt_i - P_i / m_i
//                      G2P   vpic = sum w vnt, dvp = sum w dv, L = sum grad w (x) vnt,
//                            v' = (1 - a) vpic + a (v + dvp), x' = x + dt vpic,
//                            F' = (I + dt L) F, (sigma', eps') = return map
// The adjoint runs these backwards: g2p_adjoint_kernel (per particle: return
// map VJP, F, v, and the records of the PIC/FLIP/L adjoints), the adjoint
// scatter (p2g_kernel MODE 2) of those records to the nodes,
// grid_adjoint_kernel (per node: velocities, BCs, forward moment), and
// p2g_adjoint_kernel (per particle: gather of the node adjoints into v, sigma,
// V = det(F) V0 and x, together with the position adjoint of the G2P weights
// over the same window). The weights' position dependence enters through
// grad w and its Hessian.
// ===========================================================================
namespace pmpm {

// Second derivatives of the window weights (same entries and node types as
// particle_weights).
template <int D, int SF>
__device__ inline void window_d2(const pmpm_params &prm, const pmpm_grid &g, const PW<D, SF> &w, float d2N[3][4]) {
  constexpr int E = Spline<SF>::ENTRIES;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float frac = __fsub_rn(w.t[a], (float)w.bin[a]);
    if (SF == PMPM_SF_CUBIC && a < D) {
      float N[4], dN[4];
      cubic_axis(frac, w.bin[a], g.num_cells[a], prm.node_types_built, g.inv_cell_size, N, dN, d2N[a]);
      continue;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      d2N[a][e] = 0.f;
      if (a < D && e < E) {
        const int o = w.first[a] + e;
        const int type = cubic_node_type(w.bin[a] + o, g.num_cells[a], prm.node_types_built);
        d2N[a][e] = spline_d2<SF>(__fsub_rn(frac, (float)o), type, g.inv_cell_size);
      }
    }
  }
}

// Weight, gradient and Hessian (xx, yy, zz, xy, xz, yz) of one window entry
// from per-axis values.
__device__ inline void entry_derivs(const float *N, const float *dN, const float *d2N, float &wgt, float *gr,
                                    float *H) {
  wgt = (N[0] * N[1]) * N[2];
  gr[0] = (dN[0] * N[1]) * N[2];
  gr[1] = (dN[1] * N[0]) * N[2];
  gr[2] = (dN[2] * N[0]) * N[1];
  H[0] = d2N[0] * N[1] * N[2];
  H[1] = N[0] * d2N[1] * N[2];
  H[2] = N[0] * N[1] * d2N[2];
  H[3] = dN[0] * dN[1] * N[2];
  H[4] = dN[0] * N[1] * dN[2];
  H[5] = N[0] * dN[1] * dN[2];
}

// Wave sum of v (all 64 lanes take part).
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Per particle (cell order j, state slot src): return-map VJP, F, v, x, and
// the records (vpic adjoint, FLIP adjoint, L adjoint) for the node scatter.
template <int D, int SF>
__global__ __launch_bounds__(256) void g2p_adjoint_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles st,
                                                          pmpm_adjoint ob, pmpm_adjoint ib, pmpm_spatial part,
                                                          pmpm_nodes nodes, float *__restrict__ rec,
                                                          float *__restrict__ pbar) {
  const int n = st.n;
  const int nn = g.num_cells_total;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = j < n;
  float pb[4] = {0.f, 0.f, 0.f, 0.f};
  int color = -1;
  const int src = valid ? part.sorted[j] : 0;
  PW<D, SF> w;
  float xp[3];
  if (valid) particle_weights<D, SF>(prm, g, st.x, n, src, w, xp);
  float vpic[3] = {0.f, 0.f, 0.f}, dvp[3] = {0.f, 0.f, 0.f};
  float L[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if constexpr (D == 3) {  // the forward G2P's gather, node rows staged in LDS
    static_assert(kG2PThreads == 256, "the workgroup of the staged gather");
    __shared__ __attribute__((aligned(16))) char lds[g2p_stage_bytes<true>()];
    const int nv = min(n - (int)blockIdx.x * kG2PThreads, kG2PThreads);
    g2p_gather_staged<D, SF, true, true>(g, nodes, w, nv, lds, vpic, dvp, L);
  } else if (valid) {
    gather_grid<D, SF, true, true>(g, node_src(g, nodes), w, vpic, dvp, L);
  }
  if (valid) {

    float xo[3] = {0.f, 0.f, 0.f}, vo[3] = {0.f, 0.f, 0.f}, Fo[9], F[9], so6[6], eo6[6], s6[6], e6[6], sr6[6];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      xo[a] = ob.x[a * n + j];
      vo[a] = ob.v[a * n + j];
    }
#pragma unroll
    for (int k = 0; k < D * D; ++k) {
      Fo[k] = ob.F[k * n + j];
      F[k] = st.F[k * n + src];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      so6[k] = ob.stress[k * n + j];
      eo6[k] = ob.eps_e ? ob.eps_e[k * n + j] : 0.f;
      s6[k] = st.stress[k * n + src];
      e6[k] = st.eps_e ? st.eps_e[k * n + src] : 0.f;
      sr6[k] = 0.f;
    }
    color = st.color[src];
    const uint8_t pf = st.flags[src];

    // return map (stress and elastic strain pass through unless updated)
    float Lb[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, sb6[6], eb6[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) { sb6[k] = 0.f; eb6[k] = 0.f; }
    bool s_through = true, e_through = true;
    if (color < prm.num_materials) {
      const pmpm_material &m = prm.materials[color];
      if (m.kind == PMPM_MAT_LINEAR_ELASTIC && (pf & PMPM_PF_ACTIVE)) {
        le_vjp<D>(m.p, prm.dt, L, s6, so6, sb6, Lb, pb);
        s_through = false;
      } else if (m.kind == PMPM_MAT_DRUCKER_PRAGER) {
        const bool history = (m.flags & PMPM_MATF_UPDATE_HISTORY) != 0;
        if (st.stress_ref && (m.flags & PMPM_MATF_HAS_STRESS_REF))
#pragma unroll
          for (int k = 0; k < 6; ++k) sr6[k] = st.stress_ref[k * n + src];
        dp_vjp_rev<D>(m.p, m.p[4], (m.flags & PMPM_MATF_VELGRAD_IS_INCREMENT) != 0, history,
                  (m.flags & PMPM_MATF_APEX_GUARD) != 0, prm.dt, L, sr6, e6, so6, eo6, eb6, Lb, pb);
        s_through = false;  // the stress is recomputed from the strain
        e_through = !history;
      }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (s_through) sb6[k] += so6[k];
      if (e_through) eb6[k] += eo6[k];
    }

    // F' = (I + dt L) F
    float Fb[9];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        float acc = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) acc += ((a == k ? 1.f : 0.f) + prm.dt * L[3 * a + k]) * Fo[a * D + b];
        Fb[k * D + b] = acc;
      }
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        float acc = 0.f;
#pragma unroll
        for (int b = 0; b < D; ++b) acc += Fo[a * D + b] * F[k * D + b];
        Lb[3 * a + k] += prm.dt * acc;
      }

    // v' = (1 - a) vpic + a (v + dvp), x' = x + dt vpic
    const float al = prm.alpha, am = (float)(1.0 - (double)prm.alpha);
    float vpb[3] = {0.f, 0.f, 0.f}, dvb[3] = {0.f, 0.f, 0.f}, xb[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; ++a) {
      vpb[a] = am * vo[a] + prm.dt * xo[a];
      dvb[a] = al * vo[a];
      xb[a] = xo[a];
    }
    // (the position adjoint through the weights is added by p2g_adjoint_kernel,
    // which walks the same window)

#pragma unroll
    for (int a = 0; a < D; ++a) {
      ib.x[a * n + src] = xb[a];
      ib.v[a * n + src] = al * vo[a];
      rec[a * n + j] = vpb[a];
      rec[(3 + a) * n + j] = dvb[a];
    }
#pragma unroll
    for (int k = 0; k < D * D; ++k) ib.F[k * n + src] = Fb[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      ib.stress[k * n + src] = sb6[k];
      if (ib.eps_e) ib.eps_e[k * n + src] = eb6[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) rec[(6 + k) * n + j] = Lb[k];
  }
  // material parameter adjoints: the workgroup's sum per material and
  // parameter (waves in a fixed order), reduced over workgroups in a fixed
  // order by param_bar_reduce_kernel, so the gradient does not depend on
  // scheduling
  __shared__ float wsum[4][4 * PMPM_MAX_MATERIALS];
  const int wave = threadIdx.x >> 6;
  for (int k = 0; k < prm.num_materials; ++k)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float s = wave_sum(color == k ? pb[q] : 0.f);
      if ((threadIdx.x & 63) == 0) wsum[wave][4 * k + q] = s;
    }
  __syncthreads();
  const int e = threadIdx.x;
  if (e < 4 * prm.num_materials)
    pbar[(size_t)blockIdx.x * (4 * prm.num_materials) + e] = ((wsum[0][e] + wsum[1][e]) + wsum[2][e]) + wsum[3][e];
}

// param_bar[e] += the sum over workgroups of part[b * entries + e]: one
// workgroup per entry, each lane summing a fixed stride of workgroups, then a
// fixed tree.
__global__ __launch_bounds__(256) void param_bar_reduce_kernel(const float *__restrict__ part, int groups,
                                                               int entries, float *__restrict__ param_bar) {
  __shared__ float red[256];
  const int e = blockIdx.x;
  float s = 0.f;
  for (int b = threadIdx.x; b < groups; b += 256) s += part[(size_t)b * entries + e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) param_bar[e] += red[0];
}

// Node adjoints (mass, momentum, force) as one record per node: 8 floats in
// 2D and 3D (m, P, f, padding), 4 in 1D, so a gather reads one or two
// 16-byte vectors per node instead of 1 + 2D scalars from separate arrays.
template <int D> struct NodeAdj { static constexpr int NP = D == 1 ? 4 : 8; };
template <int D>
__device__ inline void load_node_adj(const float *__restrict__ nbar, int node, float &mb, float *Pb, float *fb) {
  const char *base = reinterpret_cast<const char *>(nbar);
  const uint32_t at = (uint32_t)node * (uint32_t)(NodeAdj<D>::NP * sizeof(float));
  float r[8];
  const float4 a = *reinterpret_cast<const float4 *>(base + at);
  r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
  if constexpr (NodeAdj<D>::NP == 8) {
    const float4 b = *reinterpret_cast<const float4 *>(base + at + 16u);
    r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
  }
  mb = r[0];
#pragma unroll
  for (int a2 = 0; a2 < D; ++a2) {
    Pb[a2] = r[1 + a2];
    fb[a2] = r[1 + D + a2];
  }
}
template <int D>
__device__ inline void store_node_adj(float *__restrict__ nbar, int node, float mb, const float *Pb, const float *fb) {
  float r[8] = {mb, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    r[1 + a] = Pb[a];
    r[1 + D + a] = fb[a];
  }
  float4 *p = reinterpret_cast<float4 *>(nbar + (size_t)NodeAdj<D>::NP * node);
  p[0] = make_float4(r[0], r[1], r[2], r[3]);
  if constexpr (NodeAdj<D>::NP == 8) p[1] = make_float4(r[4], r[5], r[6], r[7]);
}

// Per node: adjoints of the node velocities (scattered in the tile partials)
// -> adjoints of mass, moment and force. Boundary conditions in reverse list
// order with the forward's active sets, which the forward grid pass recorded
// per node in `clamp` (components a NodeDomain condition set to zero).
template <int D>
__global__ void grid_adjoint_kernel(pmpm_params prm, pmpm_grid g, pmpm_nodes nodes, TileGeom tg,
                                    const float *__restrict__ partials, const int32_t *__restrict__ flags,
                                    const uint8_t *__restrict__ clamp, float *__restrict__ nbar) {
  const int nn = g.num_cells_total;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  int ids[3];
  node_ids<D>(g, i, ids);
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  gather_partials<D, 2 * D>(tg, partials, flags, ids, acc);
  const float m = nodes.mass[i];
  float mb = 0.f, Pb[3] = {0.f, 0.f, 0.f}, Qb[3] = {0.f, 0.f, 0.f}, fb[3] = {0.f, 0.f, 0.f};
  if (m > prm.small_mass_cutoff) {
    float vnt[3] = {0.f, 0.f, 0.f}, dvn[3] = {0.f, 0.f, 0.f};
    load_node_vel<D, true>(nodes.vel, i, vnt, dvn);
    const unsigned cm = clamp[i];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const float vb = acc[a], db = acc[D + a];
      Qb[a] = (vb + db) / m;
      Pb[a] = -db / m;
      mb -= (vb * vnt[a] + db * dvn[a]) / m;
    }
    for (int k = prm.num_bcs - 1; k >= 0; --k) {
      const pmpm_bc &bc = prm.bcs[k];
      if (bc.kind == PMPM_BC_NODE_DOMAIN) {
        // (one NodeDomain per list: pmpm_adjoint_step checks)
#pragma unroll
        for (int a = 0; a < D; ++a) {
          if (cm & (1u << a)) Pb[a] = 0.f;
          if (cm & (8u << a)) Qb[a] = 0.f;
        }
      } else if (bc.kind == PMPM_BC_GRAVITY) {
        if (m <= 0.000000001f) continue;
        // Q += g m dt
#pragma unroll
        for (int a = 0; a < D; ++a) mb += bc.f[a] * prm.dt * Qb[a];
      }
    }
    // Q = P + dt f
#pragma unroll
    for (int a = 0; a < D; ++a) {
      Pb[a] += Qb[a];
      fb[a] = prm.dt * Qb[a];
    }
  }
  store_node_adj<D>(nbar, i, mb, Pb, fb);
}

// d det(F) / dF (row-major D x D).
template <int D> __device__ inline void cofactor(const float *F, float *C) {
  if (D == 1) { C[0] = 1.f; return; }
  if (D == 2) { C[0] = F[3]; C[1] = -F[2]; C[2] = -F[1]; C[3] = F[0]; return; }
  C[0] = F[4] * F[8] - F[5] * F[7];
  C[1] = F[5] * F[6] - F[3] * F[8];
  C[2] = F[3] * F[7] - F[4] * F[6];
  C[3] = F[2] * F[7] - F[1] * F[8];
  C[4] = F[0] * F[8] - F[2] * F[6];
  C[5] = F[1] * F[6] - F[0] * F[7];
  C[6] = F[1] * F[5] - F[2] * F[4];
  C[7] = F[2] * F[3] - F[0] * F[5];
  C[8] = F[0] * F[4] - F[1] * F[3];
}

// Per particle: gather of the node adjoints through the P2G window (the same
// entries as the forward deposit, linear-index wrap-around included) into the
// adjoints of v, sigma, F (through V = det(F) V0) and x, plus the position
// adjoint through the G2P weights, which walk the same window:
//   xbar_c += sum_i grad_c w (vnt . vpb + dv . dvb) + sum_ab d2w/dx_c dx_a Lb_ab vnt_b
// with vpb, dvb, Lb the records g2p_adjoint_kernel wrote. The in-grid entries
// run branch-free (clamped node, zero weights outside, as gather_grid); the
// wrap-around entries of boundary particles, which only the deposit reaches,
// take a separate compact loop. Adds to ib.
// LDS staging of the node records the 3D cubic adjoint P2G reads (one lane
// per particle, kG2PThreads per workgroup, consecutive in cell order): as in
// g2p_gather_staged, the workgroup's particles form runs (same y and z bins,
// x bins not decreasing and within one 32-aligned block), whose windows lie
// in 4 x 4 node rows over the run's x range - 1 .. + 2, at most 16 x 35
// nodes; the runs are staged in rounds of at most kAdjStageNodes nodes, the
// node adjoint record (8 floats) and velocity record (8 floats) of each
// clamped node. gather(adj, vel, box0, X) is called by every live thread in
// the round that holds its run: the node of window entry (ex, sy, sz) is
// box0 + ex + X (sy + 4 sz) of adj and vel. Every thread of the workgroup
// calls this (barriers); `live` is false past the particles.
constexpr int kAdjStageNodes = 576;  // 16 x 36: one run always fits
static_assert(kAdjStageNodes >= 16 * (32 + 3), "a run spans at most 32 x bins");
struct AdjRunTable {
  int y[kG2PThreads], z[kG2PThreads], x0[kG2PThreads], x1[kG2PThreads], base[kG2PThreads];
  int wave_sum[kG2PThreads / 64];
};
template <class PWT, class F>
__device__ inline void adjoint_stage_rounds(const pmpm_grid &g, const pmpm_nodes &nodes,
                                            const float *__restrict__ nbar, const PWT &w, bool live, F &&gather) {
  constexpr int CAP = kAdjStageNodes;
  __shared__ float4 s_adj[CAP * 2], s_vel[CAP * 2];
  __shared__ AdjRunTable R;
  __shared__ int rb[3][kG2PThreads];
  const int t = (int)threadIdx.x;
#pragma unroll
  for (int a = 0; a < 3; ++a) rb[a][t] = w.bin[a];
  const int nv = __syncthreads_count(live);  // live threads are a prefix
  auto starts_at = [&](int q) {
    return rb[1][q - 1] != rb[1][q] || rb[2][q - 1] != rb[2][q] || rb[0][q] < rb[0][q - 1] ||
           (rb[0][q] >> 5) != (rb[0][q - 1] >> 5);
  };
  bool start = false, end = false;
  if (live) {
    start = t == 0 || starts_at(t);
    end = t == nv - 1 || starts_at(t + 1);
  }
  int nruns = 0;
  const int k = block_inclusive_sum(start ? 1 : 0, R.wave_sum, &nruns) - 1;
  if (start) {
    R.y[k] = w.bin[1];
    R.z[k] = w.bin[2];
    R.x0[k] = w.bin[0];
  }
  if (end) R.x1[k] = w.bin[0];
  __syncthreads();
  const int cnt = t < nruns ? 16 * (R.x1[t] - R.x0[t] + 4) : 0;
  int all = 0;
  const int incl = block_inclusive_sum(cnt, R.wave_sum, &all);
  if (t < nruns) R.base[t] = incl - cnt;
  __syncthreads();
  const int myk = live ? k : -1;
  for (int k0 = 0; k0 < nruns;) {  // uniform over the workgroup
    const int b0 = R.base[k0];
    int k1 = k0 + 1;
    while (k1 < nruns && R.base[k1] + 16 * (R.x1[k1] - R.x0[k1] + 4) - b0 <= CAP) ++k1;
    const int tot = R.base[k1 - 1] + 16 * (R.x1[k1 - 1] - R.x0[k1 - 1] + 4) - b0;
    for (int q = t; q < tot; q += kG2PThreads) {
      int kk = k0;
      while (kk + 1 < k1 && R.base[kk + 1] - b0 <= q) ++kk;
      const int X = R.x1[kk] - R.x0[kk] + 4, qq = q - (R.base[kk] - b0);
      const int row = qq / X, x = qq - row * X;  // row = sz * 4 + sy
      const int ix = min(max(R.x0[kk] - 1 + x, 0), g.num_cells[0] - 1);
      const int iy = min(max(R.y[kk] - 1 + (row & 3), 0), g.num_cells[1] - 1);
      const int iz = min(max(R.z[kk] - 1 + (row >> 2), 0), g.num_cells[2] - 1);
      const uint32_t node = (uint32_t)ix + (uint32_t)g.num_cells[0] * ((uint32_t)iy + (uint32_t)g.num_cells[1] * (uint32_t)iz);
      const float4 *ar = reinterpret_cast<const float4 *>(nbar + 8 * (size_t)node);
      const float4 *vr = reinterpret_cast<const float4 *>(nodes.vel + 8 * (size_t)node);
      s_adj[2 * q] = ar[0];
      s_adj[2 * q + 1] = ar[1];
      s_vel[2 * q] = vr[0];
      s_vel[2 * q + 1] = vr[1];
    }
    __syncthreads();
    if (myk >= k0 && myk < k1) {
      const int off = R.base[myk] - b0, X = R.x1[myk] - R.x0[myk] + 4;
      gather(reinterpret_cast<const float *>(s_adj), reinterpret_cast<const float *>(s_vel),
             off + (w.bin[0] - R.x0[myk]), X);
    }
    k0 = k1;
    __syncthreads();  // the next round overwrites the staging
  }
}

template <int D, int SF, int LANES>
__global__ __launch_bounds__(256) void p2g_adjoint_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles st,
                                                          pmpm_adjoint ib, pmpm_spatial part, pmpm_nodes nodes,
                                                          const float *__restrict__ nbar, const float *__restrict__ rec) {
  constexpr int E = Spline<SF>::ENTRIES;
  constexpr int EY = D > 1 ? E : 1;
  constexpr int EZ = D > 2 ? E : 1;
  const int n = st.n;
  const int nn = g.num_cells_total;
  // LANES = 1: one lane per particle, the window's y offsets in turn;
  // LANES = 4: four consecutive lanes per particle, one y offset each;
  // LANES = 16 (3D): sixteen lanes per particle, one (y, z) row of the window
  // each. The particle's sums are combined by shuffles; the wider layouts are
  // for launches too small to fill the chip otherwise.
  static_assert(LANES == 1 || LANES == 4 || (LANES == 16 && D == 3), "lanes per particle");
  // 3D cubic with one lane per particle: the workgroup's node records are
  // staged in LDS (adjoint_stage_rounds), so every thread takes part in the
  // staging barriers; a thread past the particles works on the last one and
  // writes nothing
  constexpr bool STAGED = D == 3 && LANES == 1 && SF == PMPM_SF_CUBIC;
  const int j = (int)blockIdx.x * ((int)blockDim.x / LANES) + (int)threadIdx.x / LANES;
  const int sub0 = (int)threadIdx.x % LANES;
  if (!STAGED && j >= n) return;  // (all lanes of a particle together)
  const bool live = j < n;
  const int jj = live ? j : n - 1;
  const int src = part.sorted[jj];
  PW<D, SF> w;
  float xp[3];
  particle_weights<D, SF>(prm, g, st.x, n, src, w, xp);
  float d2w[3][4];
  window_d2<D, SF>(prm, g, w, d2w);
  const float m = st.mass[src], V = st.volume[src];
  float v[3] = {0.f, 0.f, 0.f}, S[3][3];
#pragma unroll
  for (int a = 0; a < D; ++a) v[a] = st.v[a * n + src];
  const Sym s = load_sym(st.stress, n, src);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) S[a][b] = sym_at(s, a, b);
  float vpb[3] = {0.f, 0.f, 0.f}, dvb[3] = {0.f, 0.f, 0.f}, Lb[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    vpb[a] = rec[a * n + jj];
    dvb[a] = rec[(3 + a) * n + jj];
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) Lb[3 * a + b] = rec[(6 + 3 * a + b) * n + jj];
  // per node: the deposit adjoints (m, P, f) and, for the G2P weights, the
  // node velocities; G = sum f (x) grad w gives the stress and volume adjoints
  // after the loop
  float Sv[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) Sv[a][b] = V * S[a][b];
  float vb[3] = {0.f, 0.f, 0.f}, xb[3] = {0.f, 0.f, 0.f}, G[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f},
                                                                       {0.f, 0.f, 0.f}};
  // one entry's contributions (g2p = 1 adds the G2P weights' position term)
  // (adj, vel: the node adjoint and velocity records, in global memory or
  // staged in LDS, indexed by `node`)
  auto entry = [&](const float *adj, const float *vel, int node, const float *N3, const float *dN3,
                   const float *d2N3, bool g2p) {
    float wgt, gr[3], H[6];
    entry_derivs(N3, dN3, d2N3, wgt, gr, H);
    float mb, Pb[3] = {0.f, 0.f, 0.f}, fb[3] = {0.f, 0.f, 0.f};
    load_node_adj<D>(adj, node, mb, Pb, fb);
    // gs: the weight's scalar factor; wv = L_bar v_nt - V sigma^T f_bar
    float gs = mb, wv[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; ++a) gs += v[a] * Pb[a];
    gs *= m;
#pragma unroll
    for (int b = 0; b < D; ++b)
#pragma unroll
      for (int a = 0; a < D; ++a) wv[b] -= Sv[a][b] * fb[a];
    if (g2p) {
      float vnt[3] = {0.f, 0.f, 0.f}, dvn[3] = {0.f, 0.f, 0.f};
      load_node_vel<D, true>(vel, node, vnt, dvn);
#pragma unroll
      for (int b = 0; b < D; ++b) gs += vnt[b] * vpb[b] + dvn[b] * dvb[b];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) wv[a] += Lb[3 * a + b] * vnt[b];
    }
    const float mw = m * wgt;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      vb[a] += mw * Pb[a];
#pragma unroll
      for (int b = 0; b < D; ++b) G[a][b] += fb[a] * gr[b];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
      float hs = 0.f;
#pragma unroll
      for (int b = 0; b < D; ++b) hs += H[sym_idx(c, b)] * wv[b];
      xb[c] += gr[c] * gs + hs;
    }
  };
  // in-grid entries
  float N[3][4], dN[3][4], d2N[3][4];
  int off[3][4];
  bool leaves = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int stride = a == 0 ? 1 : (a == 1 ? g.num_cells[0] : g.num_cells[0] * g.num_cells[1]);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (a < D && e < E) {
        const int i = w.bin[a] + w.first[a] + e;
        const bool ok = i >= 0 && i < g.num_cells[a];
        leaves = leaves || !ok;
        off[a][e] = min(max(i, 0), g.num_cells[a] - 1) * stride;
        N[a][e] = ok ? w.N[a][e] : 0.f;
        dN[a][e] = ok ? w.dN[a][e] : 0.f;
        d2N[a][e] = ok ? d2w[a][e] : 0.f;
      } else {
        off[a][e] = 0;
        N[a][e] = 1.f;
        dN[a][e] = 0.f;
        d2N[a][e] = 0.f;
      }
    }
  }
  // one wrap-around entry (deposit only; the wrapped node's types)
  auto wrap_entry = [&](const int *e3) {
    int nb[3] = {0, 0, 0}, offs[3] = {0, 0, 0};
    bool inside = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      offs[a] = w.first[a] + e3[a];
      nb[a] = w.bin[a] + offs[a];
      inside = inside && nb[a] >= 0 && nb[a] < g.num_cells[a];
    }
    const int64_t lin = lin_index(g, nb, D);
    if (inside || lin < 0 || lin >= nn) return;
    int ids[3];
    node_ids<D>(g, lin, ids);
    float n3[3] = {1.f, 1.f, 1.f}, d3[3] = {0.f, 0.f, 0.f}, h3[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int type = cubic_node_type(ids[a], g.num_cells[a], prm.node_types_built);
      const float r = __fsub_rn(__fsub_rn(w.t[a], (float)w.bin[a]), (float)offs[a]);
      spline_eval<SF>(r, type, g.inv_cell_size, n3[a], d3[a]);
      h3[a] = spline_d2<SF>(r, type, g.inv_cell_size);
    }
    entry(nbar, nodes.vel, (int)lin, n3, d3, h3, false);
  };
  // the window's offsets along y (and z) are picked by selects, so the
  // in-grid entries stay unrolled without indexing registers at run time
  auto pick = [](const auto *a, int k) { return k == 0 ? a[0] : (k == 1 ? a[1] : (k == 2 ? a[2] : a[3])); };
  if constexpr (LANES == 16) {
    const int sy = sub0 & 3, sz = sub0 >> 2;
    if (sy < EY && sz < EZ) {
      const int oyz = pick(off[1], sy) + pick(off[2], sz);
      const float Ny = pick(N[1], sy), dNy = pick(dN[1], sy), d2Ny = pick(d2N[1], sy);
      const float Nz = pick(N[2], sz), dNz = pick(dN[2], sz), d2Nz = pick(d2N[2], sz);
#pragma unroll
      for (int ex = 0; ex < E; ++ex) {
        const float n3[3] = {N[0][ex], Ny, Nz}, d3[3] = {dN[0][ex], dNy, dNz}, h3[3] = {d2N[0][ex], d2Ny, d2Nz};
        entry(nbar, nodes.vel, off[0][ex] + oyz, n3, d3, h3, true);
      }
      if (leaves && prm.wrap_quirk) {
#pragma unroll 1
        for (int ex = 0; ex < E; ++ex) {
          const int e3[3] = {ex, sy, sz};
          wrap_entry(e3);
        }
      }
    }
  } else if constexpr (STAGED) {
    // in-grid entries from the LDS records of the workgroup's runs (the
    // clamped nodes, as off[]), then the wrap-around entries
    adjoint_stage_rounds(g, nodes, nbar, w, live, [&](const float *adj, const float *vel, int box0, int X) {
#pragma unroll 1
      for (int sy = 0; sy < 4; ++sy) {
        const float Ny = pick(N[1], sy), dNy = pick(dN[1], sy), d2Ny = pick(d2N[1], sy);
#pragma unroll
        for (int ex = 0; ex < 4; ++ex)
#pragma unroll
          for (int ez = 0; ez < 4; ++ez) {
            const float n3[3] = {N[0][ex], Ny, N[2][ez]}, d3[3] = {dN[0][ex], dNy, dN[2][ez]},
                        h3[3] = {d2N[0][ex], d2Ny, d2N[2][ez]};
            entry(adj, vel, box0 + ex + X * (sy + 4 * ez), n3, d3, h3, true);
          }
      }
    });
    if (leaves && prm.wrap_quirk) {
#pragma unroll 1
      for (int e = 0; e < 64; ++e) {
        const int e3[3] = {e & 3, (e >> 2) & 3, e >> 4};
        wrap_entry(e3);
      }
    }
  } else {
#pragma unroll 1
    for (int sub = sub0; sub < EY; sub += LANES) {
      const int oy = pick(off[1], sub);
      const float Ny = pick(N[1], sub), dNy = pick(dN[1], sub), d2Ny = pick(d2N[1], sub);
#pragma unroll
      for (int ex = 0; ex < E; ++ex)
#pragma unroll
        for (int ez = 0; ez < EZ; ++ez) {
          const float n3[3] = {N[0][ex], Ny, N[2][ez]}, d3[3] = {dN[0][ex], dNy, dN[2][ez]},
                      h3[3] = {d2N[0][ex], d2Ny, d2N[2][ez]};
          entry(nbar, nodes.vel, off[0][ex] + oy + off[2][ez], n3, d3, h3, true);
        }
      if (leaves && prm.wrap_quirk) {
#pragma unroll 1
        for (int e = 0; e < E * EZ; ++e) {
          const int e3[3] = {e % E, D > 1 ? sub : 0, D > 2 ? e / E : 0};
          wrap_entry(e3);
        }
      }
    }
  }
  if constexpr (LANES > 1) {  // the particle's sums over its lanes; lane 0 goes on
    auto sum_lanes = [](float v) {
#pragma unroll
      for (int o = 1; o < LANES; o <<= 1) v += __shfl_xor(v, o, 64);
      return v;
    };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      xb[a] = sum_lanes(xb[a]);
      vb[a] = sum_lanes(vb[a]);
#pragma unroll
      for (int b = 0; b < 3; ++b) G[a][b] = sum_lanes(G[a][b]);
    }
    if (sub0 != 0) return;
  }
  if (!live) return;  // (a staging helper past the particles)
  // M_bar = -V G, V_bar = -sigma : G
  float Mb[3][3], Vb = 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      Mb[a][b] = -V * G[a][b];
      Vb -= S[a][b] * G[a][b];
    }
  float F[9], C[9];
#pragma unroll
  for (int k = 0; k < D * D; ++k) F[k] = st.F[k * n + src];
  cofactor<D>(F, C);
  const float V0 = st.volume0[src];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    ib.x[a * n + src] += xb[a];
    ib.v[a * n + src] += vb[a];
  }
#pragma unroll
  for (int k = 0; k < D * D; ++k) ib.F[k * n + src] += Vb * V0 * C[k];
  // sigma is stored symmetric: an off-diagonal component feeds both entries
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = a; b < D; ++b) ib.stress[sym_idx(a, b) * n + src] += a == b ? Mb[a][a] : Mb[a][b] + Mb[b][a];
}

// Lanes per particle of the adjoint P2G: the fewest that still give three
// waves per SIMD (256 CUs x 4 SIMDs x 64 lanes x 3 lanes in all), 16 only in
// 3D. PMPM_ADJOINT_P2G_LANES=1, 4 or 16 in the environment overrides the
// choice (read at every launch), so the layouts can be checked against each
// other.
constexpr int kAdjointLanesWanted = 256 * 4 * 64 * 3;
inline int adjoint_p2g_lanes(int n, int dim) {
  const char *e = getenv("PMPM_ADJOINT_P2G_LANES");
  if (e) {
    const int v = atoi(e);
    if (v == 1 || v == 4 || (v == 16 && dim == 3)) return v;
  }
  if (n >= kAdjointLanesWanted) return 1;
  if ((int64_t)n * 4 >= kAdjointLanesWanted || dim != 3) return 4;
  return 16;
}

template <int D, int SF>
void launch_adjoint_particles(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &st,
                              const pmpm_adjoint &ob, const pmpm_adjoint &ib, const pmpm_spatial &part,
                              const pmpm_nodes &nodes, float *rec, float *pbar, int which, const float *nbar,
                              hipStream_t s) {
  const unsigned blocks = (unsigned)ceil_div(st.n, 256);
  if (which == 0)
    PMPM_LAUNCH((g2p_adjoint_kernel<D, SF>), (::pmpm::kname("g2p_adjoint_kernel", {(long)(D), (long)(SF)})), dim3(blocks), dim3(256), 0, s, prm, g, st, ob, ib, part, nodes,
                       rec, pbar);
  else {
    const int lanes = adjoint_p2g_lanes(st.n, D);
    const dim3 grid((unsigned)ceil_div(st.n, 256 / lanes));
    if (lanes == 1)
      PMPM_LAUNCH((p2g_adjoint_kernel<D, SF, 1>), (::pmpm::kname("p2g_adjoint_kernel", {(long)(D), (long)(SF), (long)(1)})), grid, dim3(256), 0, s, prm, g, st, ib, part, nodes, nbar,
                         rec);
    else if (lanes == 4 || D != 3)
      PMPM_LAUNCH((p2g_adjoint_kernel<D, SF, 4>), (::pmpm::kname("p2g_adjoint_kernel", {(long)(D), (long)(SF), (long)(4)})), grid, dim3(256), 0, s, prm, g, st, ib, part, nodes, nbar,
                         rec);
    else
      PMPM_LAUNCH((p2g_adjoint_kernel<D, SF, D == 3 ? 16 : 4>), (::pmpm::kname("p2g_adjoint_kernel", {(long)(D), (long)(SF), (long)(D == 3 ? 16 : 4)})), grid, dim3(256), 0, s, prm, g, st, ib, part,
                         nodes, nbar, rec);
  }
}

}  // namespace pmpm

extern "C" int pmpm_adjoint_step(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid,
                                 const pmpm_particles *state, const pmpm_adjoint *out_bar, pmpm_adjoint *in_bar,
                                 float *param_bar, int32_t recompute, pmpm_spatial *part, pmpm_nodes *nodes,
                                 pmpm_stream_t stream) {
  if (!ctx || !params_ok(prm) || !grid_ok(grid) || !particles_ok(state) || !out_bar || !in_bar || !param_bar || !part ||
      !nodes || (recompute != 0 && recompute != 1))
    return PMPM_ERR_INVALID;
  const int n = state->n;
  if (n == 0) return PMPM_OK;
  if (!out_bar->x || !out_bar->v || !out_bar->F || !out_bar->stress || !in_bar->x || !in_bar->v || !in_bar->F ||
      !in_bar->stress)
    return PMPM_ERR_INVALID;
  for (int k = 0; k < prm->num_materials; ++k) {
    const int kind = prm->materials[k].kind;
    if (kind != PMPM_MAT_LINEAR_ELASTIC && kind != PMPM_MAT_DRUCKER_PRAGER) return PMPM_ERR_UNSUPPORTED;
  }
  int walls = 0;
  for (int k = 0; k < prm->num_bcs; ++k) {
    const int kind = prm->bcs[k].kind;
    if (kind != PMPM_BC_GRAVITY && kind != PMPM_BC_NODE_DOMAIN) return PMPM_ERR_UNSUPPORTED;
    walls += kind == PMPM_BC_NODE_DOMAIN;
  }
  if (walls > 1) return PMPM_ERR_UNSUPPORTED;  // the clamp record holds one condition's set
  hipStream_t s = (hipStream_t)stream;
  const pmpm_grid &g = *grid;
  const int D = prm->dim;
  // the forward step's partition, P2G and grid update (kept by the caller
  // from its own pmpm_step when recompute == 0)
  int rc;
  if (recompute) {
    if ((rc = pmpm_partition(ctx, D, grid, state->x, n, part, stream)) != PMPM_OK) return rc;
    if ((rc = do_p2g(ctx, *prm, g, *state, *part, *nodes, s)) != PMPM_OK) return rc;
    if ((rc = launch_grid(ctx, *prm, g, *nodes, 1, 0, prm->num_bcs, 0, 1, s)) != PMPM_OK) return rc;
  } else if ((rc = reserve_tiles(ctx, *prm, g)) != PMPM_OK) {
    return rc;
  }
  if ((rc = ensure_buffer(ctx->adj_rec, ctx->adj_rec_cap, (size_t)15 * n)) != PMPM_OK) return rc;
  const size_t adj_np = D == 1 ? NodeAdj<1>::NP : NodeAdj<3>::NP;
  if ((rc = ensure_buffer(ctx->adj_nodes, ctx->adj_nodes_cap, adj_np * g.num_cells_total)) != PMPM_OK)
    return rc;
  const int groups = ceil_div(n, 256), entries = 4 * prm->num_materials;
  if ((rc = ensure_buffer(ctx->adj_pb, ctx->adj_pb_cap, (size_t)groups * entries)) != PMPM_OK) return rc;
  // G2P adjoint (material parameter adjoints per workgroup, then summed into
  // param_bar), then the scatter of its records to the nodes
  PMPM_DISPATCH_DSF(D, prm->shape_fn, launch_adjoint_particles, *prm, g, *state, *out_bar, *in_bar, *part, *nodes,
                    ctx->adj_rec, ctx->adj_pb, 0, nullptr, s);
  PMPM_LAUNCH_CHECK();
  if (entries > 0) {
    PMPM_LAUNCH(param_bar_reduce_kernel, (::pmpm::kname("param_bar_reduce_kernel", {})), dim3((unsigned)entries), dim3(256), 0, s, ctx->adj_pb, groups,
                       entries, param_bar);
    PMPM_LAUNCH_CHECK();
  }
  PMPM_DISPATCH_DSF(D, prm->shape_fn, launch_p2g, *prm, g, *state, *state, *part, *nodes, 2, ctx->partials,
                    ctx->tile_flags, nullptr, nullptr, nullptr, s, ctx->adj_rec);
  PMPM_LAUNCH_CHECK();
  // grid adjoint, with the forward grid pass's clamp record
  const uint8_t *clamp = nodes->clamp ? nodes->clamp : reinterpret_cast<const uint8_t *>(ctx->clamp_mask);
  const TileGeom tg = tile_geom(D, prm->shape_fn, g, 3);
  const unsigned nblocks = (unsigned)ceil_div(g.num_cells_total, 256);
  switch (D) {
    case 1: PMPM_LAUNCH(grid_adjoint_kernel<1>, (::pmpm::kname("grid_adjoint_kernel", {(long)(1)})), dim3(nblocks), dim3(256), 0, s, *prm, g, *nodes, tg, ctx->partials, ctx->tile_flags, clamp, ctx->adj_nodes); break;
    case 2: PMPM_LAUNCH(grid_adjoint_kernel<2>, (::pmpm::kname("grid_adjoint_kernel", {(long)(2)})), dim3(nblocks), dim3(256), 0, s, *prm, g, *nodes, tg, ctx->partials, ctx->tile_flags, clamp, ctx->adj_nodes); break;
    default: PMPM_LAUNCH(grid_adjoint_kernel<3>, (::pmpm::kname("grid_adjoint_kernel", {(long)(3)})), dim3(nblocks), dim3(256), 0, s, *prm, g, *nodes, tg, ctx->partials, ctx->tile_flags, clamp, ctx->adj_nodes); break;
  }
  PMPM_LAUNCH_CHECK();
  // P2G adjoint
  PMPM_DISPATCH_DSF(D, prm->shape_fn, launch_adjoint_particles, *prm, g, *state, *out_bar, *in_bar, *part, *nodes,
                    ctx->adj_rec, param_bar, 1, ctx->adj_nodes, s);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}
