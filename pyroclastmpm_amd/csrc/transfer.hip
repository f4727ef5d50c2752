// transfer.hip -- particle/grid transfers on gfx950.
//
//   p2g_tile_kernel   one workgroup per tile of cells, one lane per cell; each
//                     lane walks its cell's particles (cell order from the
//                     radix sort) and accumulates mass, momentum and force into
//                     an LDS copy of the tile's nodes. Lanes of one wave hold
//                     different cells, so a given window offset hits distinct
//                     LDS addresses; the tile is then flushed with one global
//                     atomic per node and field. The same kernel in MODE 1 is
//                     the MUSL first map + re-map.
//   grid_kernel       per node: mass cutoff, mom_nt = mom + f dt, boundary
//                     conditions in list order, node velocities for G2P.
//   g2p_kernel        per particle (cell order): gather from the grid,
//                     FLIP/PIC update, F and volume, then the return map of the
//                     particle's material, written in cell order.
// Weights are recomputed in registers from the positions; nothing per
// particle x node is stored.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "device.hpp"

namespace pmpm {

// Cell tile of one P2G workgroup (128 cells) and its thread count: every cell
// is shared by kTileThreads / 128 = 8 waves ("rounds").
constexpr int kTileThreads = 1024;
template <int D> struct Tile;
template <> struct Tile<3> { static constexpr int X = 16, Y = 4, Z = 2; };
template <> struct Tile<2> { static constexpr int X = 16, Y = 8, Z = 1; };
template <> struct Tile<1> { static constexpr int X = 128, Y = 1, Z = 1; };

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
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a < D) {
      xp[a] = x[a * n + slot];
      w.t[a] = cell_coord(xp[a], g.origin[a], g.inv_cell_size);
      w.bin[a] = (int)w.t[a];
      axis_weights<SF>(w.t[a], w.bin[a], g.num_cells[a], prm.node_types_built, g.inv_cell_size, w.first[a], w.N[a],
                       w.dN[a]);
    } else {
      xp[a] = 0.f; w.t[a] = 0.f; w.bin[a] = 0; w.first[a] = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) { w.N[a][e] = 1.f; w.dN[a][e] = 0.f; }
    }
  }
}

// Grid gather with the reference's per-axis bounds check (usl_inline.h:281-293)
// in its window order (y outer, x, z inner). Nodes below the mass cutoff carry
// zero velocities, which is equivalent to skipping them.
template <int D, int SF, bool WANT_V, bool WANT_L>
__device__ inline void gather_grid(const pmpm_grid &g, const pmpm_nodes &nodes, int nn, const PW<D, SF> &w,
                                   float *vinc, float *dvinc, float *L) {
  constexpr int E = Spline<SF>::ENTRIES;
  constexpr int EY = D > 1 ? E : 1;
  constexpr int EZ = D > 2 ? E : 1;
#pragma unroll
  for (int ey = 0; ey < EY; ++ey) {
    const int iy = (D > 1) ? w.bin[1] + w.first[1] + ey : 0;
    if (D > 1 && (iy < 0 || iy >= g.num_cells[1])) continue;
#pragma unroll
    for (int ex = 0; ex < E; ++ex) {
      const int ix = w.bin[0] + w.first[0] + ex;
      if (ix < 0 || ix >= g.num_cells[0]) continue;
      const float nxy = w.N[0][ex] * w.N[1][ey];
      const float dxy = w.dN[0][ex] * w.N[1][ey];
      const float dyx = w.dN[1][ey] * w.N[0][ex];
#pragma unroll
      for (int ez = 0; ez < EZ; ++ez) {
        const int iz = (D > 2) ? w.bin[2] + w.first[2] + ez : 0;
        if (D > 2 && (iz < 0 || iz >= g.num_cells[2])) continue;
        const int node = ix + iy * g.num_cells[0] + iz * g.num_cells[0] * g.num_cells[1];
        const float wgt = nxy * w.N[2][ez];
        float gr[3];
        gr[0] = dxy * w.N[2][ez];
        gr[1] = dyx * w.N[2][ez];
        gr[2] = (w.dN[2][ez] * w.N[0][ex]) * w.N[1][ey];
        float vnt[3];
#pragma unroll
        for (int a = 0; a < D; ++a) vnt[a] = nodes.vel_nt[a * nn + node];
        if (WANT_V) {
#pragma unroll
          for (int a = 0; a < D; ++a) {
            const float dv = nodes.dvel[a * nn + node];
            dvinc[a] += wgt * dv;
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
                                            const float A[3][3]) {
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
      float f = 0.f;
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

// LDS row pitch: at least the tile's node extent and 16 mod 32, so the two
// cell rows of one half-wave (16 lanes each) fall into opposite bank halves.
__host__ __device__ constexpr int lds_pitch(int nx, int rows_per_wave) {
  return rows_per_wave <= 1 ? nx : ((nx + 15) / 32) * 32 + 16;
}

// MODE 0: P2G of mass, momentum and force (force accumulates in mom_nt).
// MODE 1: MUSL first map (x, v of particle j written to slot j of `out`) and
//         re-map of momentum into mom_nt.
// Thread t of a workgroup owns cell (t % CELLS) of the tile and walks the
// particles k = t / CELLS, t / CELLS + ROUNDS, ... of that cell, so every wave
// holds 64 distinct cells (conflict-free LDS atomics at a fixed window offset)
// and a cell's particles spread over ROUNDS waves.
// The finished LDS block is written with plain stores to this tile's slice of
// `partials` (in-grid nodes; flags[tile] says whether the slice is live) and
// the grid kernel sums the <= 2^D slices covering each node. LDS positions
// outside the grid hold the reference's linear-index wrap-around deposits
// (weighted with the wrapped node's types); they are flushed with one global
// atomic per tile and field into the node arrays.
template <int D, int SF, int MODE>
__global__ __launch_bounds__(kTileThreads) void p2g_tile_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles in,
                                                                pmpm_particles out, pmpm_spatial part,
                                                                pmpm_nodes nodes, int tiles_x, int tiles_y,
                                                                float *__restrict__ partials,
                                                                int32_t *__restrict__ flags) {
  using SP = Spline<SF>;
  using T = Tile<D>;
  constexpr int CELLS = T::X * T::Y * T::Z;
  constexpr int ROUNDS = kTileThreads / CELLS;
  constexpr int NX = T::X + SP::SPAN - 1;
  constexpr int NY = D > 1 ? T::Y + SP::SPAN - 1 : 1;
  constexpr int NZ = D > 2 ? T::Z + SP::SPAN - 1 : 1;
  constexpr int NXP = lds_pitch(NX, D > 1 ? 64 / T::X : 1);
  constexpr int NT = NXP * NY * NZ;  // LDS nodes incl. padding
  constexpr int NV = MODE == 0 ? 1 + 2 * D : D;
  constexpr int E = SP::ENTRIES;
  constexpr int EY = D > 1 ? E : 1;
  constexpr int EZ = D > 2 ? E : 1;
  __shared__ float tile[NV * NT];

  const int n = in.n;
  const int nn = g.num_cells_total;
  const int t = threadIdx.x;
  const int cl = t % CELLS;
  const int round = t / CELLS;
  const int bx = blockIdx.x % tiles_x;
  const int by = (blockIdx.x / tiles_x) % tiles_y;
  const int bz = blockIdx.x / (tiles_x * tiles_y);
  const int c[3] = {bx * T::X + cl % T::X, by * T::Y + (cl / T::X) % T::Y, bz * T::Z + cl / (T::X * T::Y)};
  const int n0[3] = {bx * T::X + SP::NLO, D > 1 ? by * T::Y + SP::NLO : 0, D > 2 ? bz * T::Z + SP::NLO : 0};

  for (int i = t; i < NV * NT; i += kTileThreads) tile[i] = 0.f;

  bool in_grid = c[0] < g.num_cells[0];
  if (D > 1) in_grid = in_grid && c[1] < g.num_cells[1];
  if (D > 2) in_grid = in_grid && c[2] < g.num_cells[2];
  int cs = 0, ce = 0;
  if (in_grid) {
    const int cell = (int)lin_index(g, c, D);
    ce = part.cell_end[cell];
    cs = part.cell_start[cell];
  }
  if (!__syncthreads_or(ce - cs > round)) {
    if (t == 0) flags[blockIdx.x] = 0;
    return;
  }

  for (int j = cs + round; j < ce; j += ROUNDS) {
    const int slot = part.sorted[j];
    PW<D, SF> w;
    float xp[3];
    particle_weights<D, SF>(prm, g, in.x, n, slot, w, xp);
    const float m = in.mass[slot];
    float mv[3] = {0.f, 0.f, 0.f};
    float A[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (MODE == 0) {
#pragma unroll
      for (int a = 0; a < D; ++a) mv[a] = m * in.v[a * n + slot];
      const Sym s = load_sym(in.stress, n, slot);
      const float mvol = -in.volume[slot];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) A[a][b] = mvol * sym_at(s, a, b);
    } else {
      float vinc[3] = {0.f, 0.f, 0.f}, dvinc[3] = {0.f, 0.f, 0.f};
      gather_grid<D, SF, true, false>(g, nodes, nn, w, vinc, dvinc, nullptr);
      const float am = (float)(1.0 - (double)prm.alpha);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const float vn = am * vinc[a] + prm.alpha * (in.v[a * n + slot] + dvinc[a]);
        out.v[a * n + j] = vn;
        out.x[a * n + j] = xp[a] + prm.dt * vinc[a];
        mv[a] = m * vn;
      }
    }
    bool fast = true;  // own cell and the whole window inside the grid
#pragma unroll
    for (int a = 0; a < D; ++a)
      fast = fast && w.bin[a] == c[a] && w.bin[a] + w.first[a] >= 0 && w.bin[a] + w.first[a] + E - 1 < g.num_cells[a];
    int base = w.bin[0] + w.first[0] - n0[0];
    if (D > 1) base += NXP * (w.bin[1] + w.first[1] - n0[1]);
    if (D > 2) base += NXP * NY * (w.bin[2] + w.first[2] - n0[2]);
    const bool own = (w.bin[0] == c[0]) && (D < 2 || w.bin[1] == c[1]) && (D < 3 || w.bin[2] == c[2]);
#pragma unroll
    for (int ey = 0; ey < EY; ++ey) {
#pragma unroll
      for (int ex = 0; ex < E; ++ex) {
        const float nxy = w.N[0][ex] * w.N[1][ey];
        const float dxy = w.dN[0][ex] * w.N[1][ey];
        const float dyx = w.dN[1][ey] * w.N[0][ex];
#pragma unroll
        for (int ez = 0; ez < EZ; ++ez) {
          const int li = base + ex + NXP * ey + NXP * NY * ez;
          float wgt = nxy * w.N[2][ez];
          float g0 = dxy * w.N[2][ez];
          float g1 = dyx * w.N[2][ez];
          float g2 = w.dN[2][ez] * nxy;
          if (!fast) {
            const int e3[3] = {ex, ey, ez};
            if (!own) {  // particle outside its listed cell: straight to global memory
              scatter_entry_global<D, SF, MODE>(prm, g, nodes, nn, w, e3, m, mv, A);
              continue;
            }
            int nb[3] = {0, 0, 0}, off[3] = {0, 0, 0};
            bool inside = true;
#pragma unroll
            for (int a = 0; a < D; ++a) {
              off[a] = w.first[a] + e3[a];
              nb[a] = w.bin[a] + off[a];
              inside = inside && nb[a] >= 0 && nb[a] < g.num_cells[a];
            }
            if (!inside) {
              // the reference bound-checks only the linear node index
              const int64_t lin = lin_index(g, nb, D);
              if (!prm.wrap_quirk || lin < 0 || lin >= nn) continue;
              int ids[3], types[3] = {0, 0, 0};
              node_ids<D>(g, lin, ids);
#pragma unroll
              for (int a = 0; a < D; ++a) types[a] = cubic_node_type(ids[a], g.num_cells[a], prm.node_types_built);
              float gr[3];
              entry_weight<D, SF>(w, off, types, g.inv_cell_size, wgt, gr);
              g0 = gr[0];
              g1 = gr[1];
              g2 = gr[2];
            }
          }
          if (MODE == 0) {
            atomicAdd(&tile[li], wgt * m);
#pragma unroll
            for (int a = 0; a < D; ++a) {
              atomicAdd(&tile[(1 + a) * NT + li], wgt * mv[a]);
              float f = A[a][0] * g0;
              if (D > 1) f += A[a][1] * g1;
              if (D > 2) f += A[a][2] * g2;
              atomicAdd(&tile[(1 + D + a) * NT + li], f);
            }
          } else {
#pragma unroll
            for (int a = 0; a < D; ++a) atomicAdd(&tile[a * NT + li], wgt * mv[a]);
          }
        }
      }
    }
  }
  __syncthreads();
  constexpr int NTN = NX * NY * NZ;
  float *slice = partials + (int64_t)blockIdx.x * NV * NTN;
  if (t == 0) flags[blockIdx.x] = 1;
  for (int i = t; i < NTN; i += kTileThreads) {
    const int ix = i % NX;
    const int iy = (i / NX) % NY;
    const int iz = i / (NX * NY);
    const int nb[3] = {n0[0] + ix, n0[1] + iy, n0[2] + iz};
    const int li = ix + NXP * (iy + NY * iz);
    bool inside = nb[0] >= 0 && nb[0] < g.num_cells[0];
    if (D > 1) inside = inside && nb[1] >= 0 && nb[1] < g.num_cells[1];
    if (D > 2) inside = inside && nb[2] >= 0 && nb[2] < g.num_cells[2];
    if (inside) {
#pragma unroll
      for (int v = 0; v < NV; ++v) slice[v * NTN + i] = tile[v * NT + li];
      continue;
    }
    // wrap-around deposits, reduced over the tile: one atomic per field
    const int64_t lin = lin_index(g, nb, D);
    if (lin < 0 || lin >= nn) continue;
    const int node = (int)lin;
    if (MODE == 0) {
      const float mm = tile[li];
      if (mm != 0.f) unsafeAtomicAdd(&nodes.mass[node], mm);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const float pm = tile[(1 + a) * NT + li];
        if (pm != 0.f) unsafeAtomicAdd(&nodes.mom[a * nn + node], pm);
        const float f = tile[(1 + D + a) * NT + li];
        if (f != 0.f) unsafeAtomicAdd(&nodes.mom_nt[a * nn + node], f);
      }
    } else {
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const float pm = tile[a * NT + li];
        if (pm != 0.f) unsafeAtomicAdd(&nodes.mom_nt[a * nn + node], pm);
      }
    }
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
};

// Sums the partial slices of the (at most 2 per axis) tiles whose node block
// covers node `ids`, in a fixed tile order.
template <int D>
__device__ inline void gather_partials(const TileGeom &tg, const float *partials, const int32_t *flags,
                                       const int *ids, float *acc) {
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
  for (int bz = lo[2]; bz <= hi[2]; ++bz)
    for (int by = lo[1]; by <= hi[1]; ++by)
      for (int bx = lo[0]; bx <= hi[0]; ++bx) {
        const int tile = bx + tg.tiles[0] * (by + tg.tiles[1] * bz);
        if (!flags[tile]) continue;
        int li = ids[0] - (bx * tg.T[0] + tg.nlo);
        if (D > 1) li += tg.N[0] * (ids[1] - (by * tg.T[1] + tg.nlo));
        if (D > 2) li += tg.N[0] * tg.N[1] * (ids[2] - (bz * tg.T[2] + tg.nlo));
        const float *s = partials + (int64_t)tile * tg.nv * ntn + li;
        for (int v = 0; v < tg.nv; ++v) acc[v] += s[(int64_t)v * ntn];
      }
}

// Per node: cutoff + forward moment (finalize), boundary conditions
// [bc_begin, bc_end) in list order, then the node velocities G2P reads.
// remap_pass: mom_nt holds the MUSL re-mapped momentum; only kinematic
// conditions apply (gravity already acted through the particle velocities).
template <int D>
__global__ void grid_kernel(pmpm_params prm, pmpm_grid g, pmpm_nodes nodes, int finalize, int bc_begin, int bc_end,
                            int remap_pass, TileGeom tg, const float *__restrict__ partials,
                            const int32_t *__restrict__ flags) {
  const int nn = g.num_cells_total;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  float m = nodes.mass[i];
  float P[3] = {0.f, 0.f, 0.f}, Q[3] = {0.f, 0.f, 0.f};
  if (tg.gather) {
    // node arrays hold the wrap-around deposits; add the tile partials
    int ids0[3];
    node_ids<D>(g, i, ids0);
    float acc[1 + 2 * 3] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    gather_partials<D>(tg, partials, flags, ids0, acc);
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
    for (int a = 0; a < D; ++a) {
      nodes.mom_nt[a * nn + i] = 0.f;
      nodes.vel_nt[a * nn + i] = 0.f;
      nodes.dvel[a * nn + i] = 0.f;
    }
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
  for (int k = bc_begin; k < bc_end; ++k) {
    const pmpm_bc &bc = prm.bcs[k];
    if (bc.kind == PMPM_BC_GRAVITY) {
      // gravity_inline.h:46-63
      if (remap_pass || m <= 0.000000001f) continue;
#pragma unroll
      for (int a = 0; a < D; ++a) Q[a] += bc.f[a] * m * prm.dt;
    } else if (bc.kind == PMPM_BC_NODE_DOMAIN) {
      // nodedomain_inline.h:40-87: 0 fixed, 1 slip (one-sided clamp)
      if (m <= 0.00001f) continue;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        if (ids[a] < 1) {
          if (bc.i[a] == 0) {
#pragma unroll
            for (int b = 0; b < D; ++b) { P[b] = 0.f; Q[b] = 0.f; }
          } else if (bc.i[a] == 1) {
            P[a] = fmaxf(0.f, P[a]);
            Q[a] = fmaxf(0.f, Q[a]);
          }
        } else if (ids[a] >= g.num_cells[a] - 1) {
          if (bc.i[3 + a] == 0) {
#pragma unroll
            for (int b = 0; b < D; ++b) { P[b] = 0.f; Q[b] = 0.f; }
          } else if (bc.i[3 + a] == 1) {
            P[a] = fminf(0.f, P[a]);
            Q[a] = fminf(0.f, Q[a]);
          }
        }
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
    nodes.mom_nt[a * nn + i] = Q[a];
  }
  const bool live = m > prm.small_mass_cutoff;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float vnt = live ? Q[a] / m : 0.f;
    const float v = live ? P[a] / m : 0.f;
    nodes.vel_nt[a * nn + i] = vnt;
    nodes.dvel[a * nn + i] = vnt - v;
  }
}

// MODE 0: USL G2P (usl_inline.h:245-332) + return map.
// MODE 2: MUSL second map: L, F, V from the re-mapped velocities + return map
//         (x, v of slot j of `out` were written by the first map).
template <int D, int SF, int MODE>
__global__ __launch_bounds__(256) void g2p_kernel(pmpm_params prm, pmpm_grid g, pmpm_particles in, pmpm_particles out,
                                                  pmpm_spatial part, pmpm_nodes nodes, int reorder) {
  const int n = in.n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int nn = g.num_cells_total;
  const int src = part.sorted[j];
  const int dst = reorder ? j : src;
  PW<D, SF> w;
  float xp[3];
  particle_weights<D, SF>(prm, g, in.x, n, src, w, xp);
  float vinc[3] = {0.f, 0.f, 0.f}, dvinc[3] = {0.f, 0.f, 0.f};
  float L[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (MODE == 0) gather_grid<D, SF, true, true>(g, nodes, nn, w, vinc, dvinc, L);
  else gather_grid<D, SF, false, true>(g, nodes, nn, w, vinc, dvinc, L);
  if (MODE == 0) {
    const float am = (float)(1.0 - (double)prm.alpha);
#pragma unroll
    for (int a = 0; a < D; ++a) {
      out.v[a * n + dst] = am * vinc[a] + prm.alpha * (in.v[a * n + src] + dvinc[a]);
      out.x[a * n + dst] = xp[a] + prm.dt * vinc[a];
    }
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
    if (in.stress_ref && out.stress_ref) store_sym(out.stress_ref, n, dst, load_sym(in.stress_ref, n, src));
  }
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

TileGeom tile_geom(int dim, int sf, const pmpm_grid &g, int gather) {
  TileGeom tg{};
  const int T3[3] = {16, 4, 2}, T2[3] = {16, 8, 1}, T1[3] = {128, 1, 1};
  const int *T = dim == 3 ? T3 : (dim == 2 ? T2 : T1);
  const int span = (sf == PMPM_SF_LINEAR) ? 2 : 4;
  tg.nlo = (sf == PMPM_SF_LINEAR) ? 0 : -1;
  for (int a = 0; a < 3; ++a) {
    tg.T[a] = a < dim ? T[a] : 1;
    tg.N[a] = a < dim ? T[a] + span - 1 : 1;
    tg.tiles[a] = a < dim ? (int)ceil_div(g.num_cells[a], T[a]) : 1;
  }
  tg.gather = gather;
  tg.nv = gather == 2 ? dim : 1 + 2 * dim;
  return tg;
}
static size_t tile_count(const TileGeom &tg) { return (size_t)tg.tiles[0] * tg.tiles[1] * tg.tiles[2]; }

template <int D, int SF, int MODE>
void launch_p2g_tile(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
                     const pmpm_spatial &part, const pmpm_nodes &nodes, float *partials, int32_t *flags,
                     hipStream_t s) {
  const int tx = (int)ceil_div(g.num_cells[0], Tile<D>::X);
  const int ty = D > 1 ? (int)ceil_div(g.num_cells[1], Tile<D>::Y) : 1;
  const int tz = D > 2 ? (int)ceil_div(g.num_cells[2], Tile<D>::Z) : 1;
  hipLaunchKernelGGL((p2g_tile_kernel<D, SF, MODE>), dim3(tx * ty * tz), dim3(kTileThreads), 0, s, prm, g, in, out,
                     part, nodes, tx, ty, partials, flags);
}
template <int D, int SF>
void launch_p2g(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
                const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, float *partials, int32_t *flags,
                hipStream_t s) {
  if (mode == 0) launch_p2g_tile<D, SF, 0>(prm, g, in, out, part, nodes, partials, flags, s);
  else launch_p2g_tile<D, SF, 1>(prm, g, in, out, part, nodes, partials, flags, s);
}
template <int D, int SF>
void launch_g2p(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
                const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, int reorder, hipStream_t s) {
  const unsigned blocks = (unsigned)ceil_div(in.n, 256);
  if (mode == 0)
    hipLaunchKernelGGL((g2p_kernel<D, SF, 0>), dim3(blocks), dim3(256), 0, s, prm, g, in, out, part, nodes, reorder);
  else
    hipLaunchKernelGGL((g2p_kernel<D, SF, 2>), dim3(blocks), dim3(256), 0, s, prm, g, in, out, part, nodes, reorder);
}
template <int D, int SF>
void launch_shape(const pmpm_grid &g, int types_built, const float *x, int n, float *psi, float *dpsi, hipStream_t s) {
  hipLaunchKernelGGL((shape_function_kernel<D, SF>), dim3((unsigned)ceil_div(n, 128)), dim3(128), 0, s, g, types_built,
                     x, n, psi, dpsi);
}

// gather: 0 = node arrays only, 1 = add the P2G tile partials, 2 = add the
// MUSL re-map tile partials (ctx holds them).
int launch_grid(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes, int finalize,
                int b0, int b1, int remap, int gather, hipStream_t s) {
  const unsigned blocks = (unsigned)ceil_div(g.num_cells_total, 256);
  const TileGeom tg = tile_geom(prm.dim, prm.shape_fn, g, gather);
  const float *partials = gather ? ctx->partials : nullptr;
  const int32_t *flags = gather ? ctx->tile_flags : nullptr;
  switch (prm.dim) {
    case 1: hipLaunchKernelGGL(grid_kernel<1>, dim3(blocks), dim3(256), 0, s, prm, g, nodes, finalize, b0, b1, remap, tg, partials, flags); break;
    case 2: hipLaunchKernelGGL(grid_kernel<2>, dim3(blocks), dim3(256), 0, s, prm, g, nodes, finalize, b0, b1, remap, tg, partials, flags); break;
    case 3: hipLaunchKernelGGL(grid_kernel<3>, dim3(blocks), dim3(256), 0, s, prm, g, nodes, finalize, b0, b1, remap, tg, partials, flags); break;
    default: return PMPM_ERR_INVALID;
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

static bool params_ok(const pmpm_params *prm) {
  return prm && prm->dim >= 1 && prm->dim <= 3 && prm->shape_fn >= 0 && prm->shape_fn <= 2 &&
         prm->num_materials >= 0 && prm->num_materials <= PMPM_MAX_MATERIALS && prm->num_bcs >= 0 &&
         prm->num_bcs <= PMPM_MAX_BCS;
}
static bool particles_ok(const pmpm_particles *p) {
  return p && p->n >= 0 &&
         (p->n == 0 || (p->x && p->v && p->stress && p->F && p->mass && p->volume && p->volume0 && p->color &&
                        p->flags && p->id));
}

int do_zero_nodes(const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes, hipStream_t s) {
  int rc;
  const int64_t nn = g.num_cells_total;
  if ((rc = zero_fill(nodes.mass, nn, s)) != PMPM_OK) return rc;
  if ((rc = zero_fill(nodes.mom, nn * prm.dim, s)) != PMPM_OK) return rc;
  if ((rc = zero_fill(nodes.mom_nt, nn * prm.dim, s)) != PMPM_OK) return rc;
  return PMPM_OK;
}

static int reserve_tiles(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g) {
  const TileGeom tg = tile_geom(prm.dim, prm.shape_fn, g, 1);
  const size_t tiles = tile_count(tg);
  const size_t ntn = (size_t)tg.N[0] * tg.N[1] * tg.N[2];
  return ensure_partials(ctx, tiles * tg.nv * ntn, tiles);
}

// Scatter into the tile partials (+ wrap-around deposits into the cleared
// node arrays); a grid_kernel pass with gather = 1 completes the P2G.
int do_p2g(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &p,
           const pmpm_spatial &part, const pmpm_nodes &nodes, hipStream_t s) {
  int rc = reserve_tiles(ctx, prm, g);
  if (rc != PMPM_OK) return rc;
  if ((rc = do_zero_nodes(prm, g, nodes, s)) != PMPM_OK) return rc;
  PMPM_DISPATCH_DSF(prm.dim, prm.shape_fn, launch_p2g, prm, g, p, p, part, nodes, 0, ctx->partials, ctx->tile_flags,
                    s);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

int do_g2p(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
           const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, hipStream_t s) {
  if (in.n == 0) return PMPM_OK;
  const int reorder = (out.F != in.F) ? 1 : 0;  // in place when F is updated in place
  PMPM_DISPATCH_DSF(prm.dim, prm.shape_fn, launch_g2p, prm, g, in, out, part, nodes, mode, reorder, s);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

// First map + re-map into the tile partials; a grid_kernel pass with
// gather = 2 completes mom_nt.
int do_musl_map(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in,
                const pmpm_particles &out, const pmpm_spatial &part, const pmpm_nodes &nodes, hipStream_t s) {
  int rc = reserve_tiles(ctx, prm, g);
  if (rc != PMPM_OK) return rc;
  if ((rc = zero_fill(nodes.mom_nt, (int64_t)g.num_cells_total * prm.dim, s)) != PMPM_OK) return rc;
  PMPM_DISPATCH_DSF(prm.dim, prm.shape_fn, launch_p2g, prm, g, in, out, part, nodes, 1, ctx->partials,
                    ctx->tile_flags, s);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

}  // namespace pmpm

using namespace pmpm;

extern "C" int pmpm_shape_function(int32_t dim, int32_t shape_fn, int32_t node_types_built, const pmpm_grid *grid,
                                   const float *x, int32_t n, float *psi, float *dpsi, pmpm_stream_t stream) {
  if (!grid || dim < 1 || dim > 3 || shape_fn < 0 || shape_fn > 2 || n < 0) return PMPM_ERR_INVALID;
  if (n == 0) return PMPM_OK;
  if (!x || !psi || !dpsi) return PMPM_ERR_INVALID;
  PMPM_DISPATCH_DSF(dim, shape_fn, launch_shape, *grid, node_types_built, x, n, psi, dpsi, (hipStream_t)stream);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

extern "C" int pmpm_p2g(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *p,
                        const pmpm_spatial *part, pmpm_nodes *nodes, pmpm_stream_t stream) {
  if (!ctx || !params_ok(prm) || !grid || !particles_ok(p) || !part || !nodes) return PMPM_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  int rc = do_p2g(ctx, *prm, *grid, *p, *part, *nodes, s);
  if (rc != PMPM_OK) return rc;
  // the reference's P2G ends with the cutoff and mom_nt = mom + f dt
  return launch_grid(ctx, *prm, *grid, *nodes, 1, 0, 0, 0, 1, s);
}

extern "C" int pmpm_grid_bc(const pmpm_params *prm, const pmpm_grid *grid, pmpm_nodes *nodes, int32_t bc_begin,
                            int32_t bc_end, int32_t remap_pass, pmpm_stream_t stream) {
  if (!params_ok(prm) || !grid || !nodes) return PMPM_ERR_INVALID;
  if (bc_begin < 0 || bc_end > prm->num_bcs || bc_begin > bc_end) return PMPM_ERR_INVALID;
  return launch_grid(nullptr, *prm, *grid, *nodes, 0, bc_begin, bc_end, remap_pass ? 1 : 0, 0, (hipStream_t)stream);
}

extern "C" int pmpm_g2p(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *in, pmpm_particles *out,
                        const pmpm_spatial *part, const pmpm_nodes *nodes, pmpm_stream_t stream) {
  if (!params_ok(prm) || !grid || !particles_ok(in) || !particles_ok(out) || !part || !nodes) return PMPM_ERR_INVALID;
  if (out->n != in->n) return PMPM_ERR_INVALID;
  return do_g2p(*prm, *grid, *in, *out, *part, *nodes, 0, (hipStream_t)stream);
}

extern "C" int pmpm_musl_map(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid,
                             const pmpm_particles *in, pmpm_particles *out, const pmpm_spatial *part,
                             pmpm_nodes *nodes, pmpm_stream_t stream) {
  if (!ctx || !params_ok(prm) || !grid || !particles_ok(in) || !particles_ok(out) || !part || !nodes)
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
  if (!params_ok(prm) || !grid || !particles_ok(in) || !particles_ok(out) || !part || !nodes) return PMPM_ERR_INVALID;
  if (out->n != in->n) return PMPM_ERR_INVALID;
  return do_g2p(*prm, *grid, *in, *out, *part, *nodes, 2, (hipStream_t)stream);
}

extern "C" int pmpm_stress_update(const pmpm_params *prm, int32_t mat_id, pmpm_particles *p, pmpm_stream_t stream) {
  if (!params_ok(prm) || !particles_ok(p) || !p || (p->n > 0 && !p->L)) return PMPM_ERR_INVALID;
  if (mat_id < 0 || mat_id >= prm->num_materials) return PMPM_ERR_INVALID;
  if (p->n == 0) return PMPM_OK;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = (unsigned)ceil_div(p->n, 256);
  switch (prm->dim) {
    case 1: hipLaunchKernelGGL(stress_update_kernel<1>, dim3(blocks), dim3(256), 0, s, *prm, mat_id, *p); break;
    case 2: hipLaunchKernelGGL(stress_update_kernel<2>, dim3(blocks), dim3(256), 0, s, *prm, mat_id, *p); break;
    default: hipLaunchKernelGGL(stress_update_kernel<3>, dim3(blocks), dim3(256), 0, s, *prm, mat_id, *p); break;
  }
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}
