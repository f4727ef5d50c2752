// partition.hip -- spatial binning on gfx950: cell hash per particle, stable
// LSD radix sort of (hash, slot), cell start/end ranges. Replaces the
// reference's calculate_hashes / thrust::stable_sort_by_key / bin_particles
// (spatialpartition_inline.h:49-115, spatialpartition.cpp:77-136).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <new>

#include "context.hpp"
#include "device.hpp"

namespace pmpm {

thread_local char g_last_error[256] = "";
void note_error(const char *what) { snprintf(g_last_error, sizeof(g_last_error), "%s", what); }
void note_hip_error(hipError_t e, const char *what) {
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

__global__ void zero_kernel(uint32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int zero_fill(void *ptr, int64_t n_words, hipStream_t s) {
  if (n_words <= 0) return PMPM_OK;
  int64_t blocks = ceil_div(n_words, 256);
  if (blocks > 16384) blocks = 16384;
  PMPM_LAUNCH(zero_kernel, (::pmpm::kname("zero_kernel", {})), dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t *)ptr, n_words);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

// Up to kZeroLists buffers cleared by one launch (a fill kernel, so the
// sequence stays valid inside a captured graph).
__global__ void zero_lists_kernel(ZeroLists z) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int k = 0; k < z.count; ++k)
    for (int64_t i = t; i < z.n[k]; i += stride) z.p[k][i] = 0u;
}

int zero_fill_lists(const ZeroLists &z, hipStream_t s) {
  int64_t most = 0;
  for (int k = 0; k < z.count; ++k) most = z.n[k] > most ? z.n[k] : most;
  if (most <= 0) return PMPM_OK;
  int64_t blocks = ceil_div(most, 256);
  if (blocks > 16384) blocks = 16384;
  PMPM_LAUNCH(zero_lists_kernel, (::pmpm::kname("zero_lists_kernel", {})), dim3((unsigned)blocks), dim3(256), 0, s, z);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

// bins = trunc((x - origin) * inv_h) per axis (int cast truncates toward 0, as
// Eigen's cast<int>), hash = x + y*nx + z*nx*ny. Hashes outside
// [0, num_cells_total) are clamped to num_cells_total so they sort last and
// belong to no cell. The same launch clears the cell ranges (and the `clear`
// buffers) for the range pass that follows.
template <int D>
__global__ void hash_kernel(const float *__restrict__ x, int n, pmpm_grid g, uint32_t *__restrict__ hash,
                            int32_t *__restrict__ bins, int32_t *__restrict__ cstart, int32_t *__restrict__ cend,
                            ZeroLists clear) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t k = t; k < g.num_cells_total; k += stride) {
    cstart[k] = 0;
    cend[k] = 0;
  }
  for (int q = 0; q < clear.count; ++q)
    for (int64_t k = t; k < clear.n[q]; k += stride) clear.p[q][k] = 0u;
  const int i = (int)t;
  if (i >= n) return;
  int b[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float t = __fmul_rn(__fsub_rn(x[a * n + i], g.origin[a]), g.inv_cell_size);
    b[a] = (int)t;
  }
  int64_t h = b[0];
  if (D > 1) h += (int64_t)b[1] * g.num_cells[0];
  if (D > 2) h += (int64_t)b[2] * g.num_cells[0] * g.num_cells[1];
  const uint32_t key = (h < 0 || h >= g.num_cells_total) ? (uint32_t)g.num_cells_total : (uint32_t)h;
  hash[i] = key;
  if (bins) {
#pragma unroll
    for (int a = 0; a < D; ++a) bins[a * n + i] = b[a];
  }
}

// Per-block digit histogram (digits of RB bits), digit-major layout
// hist[digit * nblocks + block]. The keys arrive nearly sorted (the previous
// step's G2P wrote the particles in cell order), so a wave's 64 consecutive
// keys hold a few runs of one digit: the head lane of each run adds the
// run's length, one LDS atomic per run instead of one per key.
template <int RB>
__global__ __launch_bounds__(kSortBlock) void radix_hist_kernel(const uint32_t *__restrict__ keys, int n, int shift,
                                                                uint32_t *__restrict__ hist, int nblocks) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t h[R];
  const int t = threadIdx.x, lane = t & 63;
  for (int d = t; d < R; d += kSortBlock) h[d] = 0u;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortItems; ++r) {
    const int64_t i = base + (int64_t)r * kSortBlock + t;
    const bool valid = i < n;  // a prefix of the wave's lanes
    const uint32_t d = valid ? (keys[i] >> shift) & (R - 1) : 0u;
    const uint32_t prev = (uint32_t)__shfl_up((int)d, 1, 64);
    const bool head = valid && (lane == 0 || prev != d);
    const uint64_t heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    if (head) {
      const uint64_t later = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
      const int next = later ? __ffsll((unsigned long long)later) - 1 : nvalid;
      atomicAdd(&h[d], (uint32_t)(next - lane));
    }
  }
  __syncthreads();
  for (int d = t; d < R; d += kSortBlock) hist[(int64_t)d * nblocks + blockIdx.x] = h[d];
}

// Exclusive scan of each digit's column hist[d][0..nblocks) (one workgroup
// per digit); the column total goes to totals[d].
__global__ __launch_bounds__(256) void scan_columns_kernel(uint32_t *__restrict__ hist, int nblocks,
                                                           uint32_t *__restrict__ totals) {
  __shared__ uint32_t part[256];
  const int t = threadIdx.x;
  uint32_t *col = hist + (int64_t)blockIdx.x * nblocks;
  const int chunk = (nblocks + 255) / 256;
  const int b = t * chunk;
  const int e = min(b + chunk, nblocks);
  uint32_t s = 0;
  for (int i = b; i < e; ++i) s += col[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t v = (t >= off) ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint32_t run = (t > 0) ? part[t - 1] : 0u;
  for (int i = b; i < e; ++i) {
    const uint32_t v = col[i];
    col[i] = run;
    run += v;
  }
  if (t == 255) totals[blockIdx.x] = part[255];
}

// Stable scatter of one radix pass (digits of RB bits). Within a round of
// kSortBlock keys (thread order == input order) a lane's rank among
// same-digit lanes of its wave comes from RB ballots; per-wave counts are
// combined per digit in LDS (thread d owns digit d).
template <int RB>
__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const uint32_t *__restrict__ keys_in, const int32_t *__restrict__ vals_in, uint32_t *__restrict__ keys_out,
    int32_t *__restrict__ vals_out, int n, int shift, const uint32_t *__restrict__ hist, int nblocks,
    const uint32_t *__restrict__ totals) {
  constexpr int R = 1 << RB, W = kSortBlock / 64;
  static_assert(R <= kSortBlock, "one digit per thread");
  __shared__ uint32_t base[R];
  __shared__ uint32_t wcnt[W][R];
  __shared__ uint32_t wsum[W];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const bool owner = t < R;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // digit bases: exclusive scan of the R column totals (thread d owns digit
  // d): a shuffle scan within each wave, then the wave totals before it
  const uint32_t tot = owner ? totals[t] : 0u;
  uint32_t incl = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t before = 0u;
  for (int q = 0; q < w; ++q) before += wsum[q];
  if (owner) base[t] = before + incl - tot + hist[(int64_t)t * nblocks + blockIdx.x];
  const int64_t tile = (int64_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortItems; ++r) {
    if (owner) {
#pragma unroll
      for (int q = 0; q < W; ++q) wcnt[q][t] = 0u;
    }
    __syncthreads();
    const int64_t i = tile + (int64_t)r * kSortBlock + t;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t d = (key >> shift) & (R - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < RB; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t m = __ballot(on);
      peers &= on ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (owner) {
      uint32_t run = base[t];
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const uint32_t c = wcnt[q][t];
        wcnt[q][t] = run;
        run += c;
      }
      base[t] = run;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = wcnt[w][d] + rank;
      keys_out[pos] = key;
      vals_out[pos] = vals_in ? vals_in[i] : (int32_t)i;
    }
    __syncthreads();
  }
}

// Cell ranges over the sorted hashes (spatialpartition_inline.h:85-115);
// empty cells keep the 0 of the preceding clear (spatialpartition.cpp:81-82).
// The out-of-grid particles sort last: the first of them (or the last
// particle, when there are none) writes their count.
__global__ void cell_range_kernel(const uint32_t *__restrict__ hs, int n, uint32_t ncells,
                                  int32_t *__restrict__ cstart, int32_t *__restrict__ cend,
                                  uint32_t *__restrict__ out_of_grid) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t h = hs[j];
  if (h >= ncells) {
    if (j == 0 || hs[j - 1] < ncells) *out_of_grid = (uint32_t)(n - j);
    return;
  }
  if (j == n - 1) *out_of_grid = 0u;
  if (j == 0 || hs[j - 1] != h) cstart[h] = j;
  if (j == n - 1 || hs[j + 1] != h) cend[h] = j + 1;
}

}  // namespace pmpm

using namespace pmpm;

extern "C" const char *pmpm_last_error(void) { return g_last_error; }

extern "C" int pmpm_context_create(pmpm_context **ctx) {
  if (!ctx) return PMPM_ERR_INVALID;
  pmpm_context *c = new (std::nothrow) pmpm_context();
  if (!c) return PMPM_ERR_INVALID;
  if (!hip_ok(hipGetDevice(&c->device), "hipGetDevice")) { delete c; return PMPM_ERR_HIP; }
  if (!hip_ok(hipMalloc(&c->counters, 4 * sizeof(uint32_t)), "hipMalloc")) { delete c; return PMPM_ERR_HIP; }
  if (!hip_ok(hipMemset(c->counters, 0, 4 * sizeof(uint32_t)), "hipMemset")) {
    (void)hipFree(c->counters);
    delete c;
    return PMPM_ERR_HIP;
  }
  *ctx = c;
  return PMPM_OK;
}

int pmpm::ensure_partials(pmpm_context *ctx, size_t floats, size_t tiles) {
  if (floats > ctx->partials_cap) {
    (void)hipFree(ctx->partials);
    ctx->partials = nullptr;
    ctx->partials_cap = 0;
    PMPM_HIP_CHECK(hipMalloc(&ctx->partials, floats * sizeof(float)));
    ctx->partials_cap = floats;
  }
  if (tiles > ctx->tile_flags_cap) {
    (void)hipFree(ctx->tile_flags);
    ctx->tile_flags = nullptr;
    ctx->tile_flags_cap = 0;
    PMPM_HIP_CHECK(hipMalloc(&ctx->tile_flags, tiles * sizeof(int32_t)));
    ctx->tile_flags_cap = tiles;
  }
  return PMPM_OK;
}

int pmpm::ensure_buffer(float *&buf, size_t &cap, size_t floats) {
  if (floats > cap) {
    (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    PMPM_HIP_CHECK(hipMalloc(&buf, floats * sizeof(float)));
    cap = floats;
  }
  return PMPM_OK;
}

extern "C" int pmpm_context_destroy(pmpm_context *ctx) {
  if (!ctx) return PMPM_OK;
  (void)hipFree(ctx->adj_rec);
  (void)hipFree(ctx->adj_nodes);
  (void)hipFree(ctx->adj_pb);
  (void)hipFree(ctx->partials);
  (void)hipFree(ctx->tile_flags);
  (void)hipFree(ctx->wrap_rec);
  (void)hipFree(ctx->wrap_touch);
  (void)hipFree(ctx->wrap_list);
  (void)hipFree(ctx->clamp_mask);
  (void)hipFree(ctx->keys_tmp);
  (void)hipFree(ctx->vals_tmp);
  (void)hipFree(ctx->hist);
  (void)hipFree(ctx->counters);
  delete ctx;
  return PMPM_OK;
}

extern "C" int pmpm_context_reserve(pmpm_context *ctx, int32_t max_particles, int32_t max_cells) {
  if (!ctx || max_particles < 0 || max_cells < 0) return PMPM_ERR_INVALID;
  if (max_particles > ctx->cap_particles) {
    (void)hipFree(ctx->keys_tmp);
    (void)hipFree(ctx->vals_tmp);
    (void)hipFree(ctx->hist);
    ctx->keys_tmp = nullptr; ctx->vals_tmp = nullptr; ctx->hist = nullptr; ctx->cap_particles = 0;
    const size_t np = (size_t)(max_particles > 0 ? max_particles : 1);
    const size_t nblocks = (size_t)ceil_div((int64_t)np, kSortTile);
    ctx->hist_len = (size_t)kMaxRadix * nblocks + kMaxRadix;  // + digit totals
    PMPM_HIP_CHECK(hipMalloc(&ctx->keys_tmp, np * sizeof(uint32_t)));
    PMPM_HIP_CHECK(hipMalloc(&ctx->vals_tmp, np * sizeof(int32_t)));
    PMPM_HIP_CHECK(hipMalloc(&ctx->hist, ctx->hist_len * sizeof(uint32_t)));
    ctx->cap_particles = max_particles;
  }
  if (max_cells > ctx->cap_cells) ctx->cap_cells = max_cells;
  return PMPM_OK;
}

extern "C" int pmpm_context_out_of_grid(pmpm_context *ctx, pmpm_stream_t stream, int32_t *count) {
  if (!ctx || !count) return PMPM_ERR_INVALID;
  uint32_t v = 0;
  PMPM_HIP_CHECK(hipMemcpyAsync(&v, ctx->counters, sizeof(uint32_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
  PMPM_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  *count = (int32_t)v;
  return PMPM_OK;
}

int pmpm::partition_clearing(pmpm_context *ctx, int dim, const pmpm_grid *grid, const float *x, int n,
                             pmpm_spatial *part, hipStream_t s, const ZeroLists &extra) {
  if (!ctx || !grid_ok(grid) || !part || dim < 1 || dim > 3 || n < 0) return PMPM_ERR_INVALID;
  if (n > 0 && (!x || !part->hash || !part->hash_sorted || !part->sorted)) return PMPM_ERR_INVALID;
  if (!part->cell_start || !part->cell_end || grid->num_cells_total <= 0) return PMPM_ERR_INVALID;
  if (n > ctx->cap_particles) return PMPM_ERR_WORKSPACE;
  const pmpm_grid g = *grid;
  const uint32_t ncells = (uint32_t)g.num_cells_total;
  int rc;
  if (n == 0) {
    ZeroLists z = extra;
    z.add(ctx->counters, 1);
    z.add(part->cell_start, ncells);
    z.add(part->cell_end, ncells);
    return zero_fill_lists(z, s);
  }
  // enough threads for the particles, and for the clears at 4 words each
  int64_t words = ncells;
  for (int q = 0; q < extra.count; ++q) words = extra.n[q] > words ? extra.n[q] : words;
  int64_t nb = ceil_div(n, 256);
  const int64_t nb_clear = ceil_div(words, 256 * 4);
  nb = nb_clear > nb ? (nb_clear < 16384 ? nb_clear : (nb > 16384 ? nb : 16384)) : nb;
  const unsigned blocks = (unsigned)nb;
  switch (dim) {
    case 1: PMPM_LAUNCH(hash_kernel<1>, (::pmpm::kname("hash_kernel", {(long)(1)})), dim3(blocks), dim3(256), 0, s, x, n, g, part->hash, part->bins, part->cell_start, part->cell_end, extra); break;
    case 2: PMPM_LAUNCH(hash_kernel<2>, (::pmpm::kname("hash_kernel", {(long)(2)})), dim3(blocks), dim3(256), 0, s, x, n, g, part->hash, part->bins, part->cell_start, part->cell_end, extra); break;
    default: PMPM_LAUNCH(hash_kernel<3>, (::pmpm::kname("hash_kernel", {(long)(3)})), dim3(blocks), dim3(256), 0, s, x, n, g, part->hash, part->bins, part->cell_start, part->cell_end, extra); break;
  }
  PMPM_LAUNCH_CHECK();
  // keys are in [0, ncells] (ncells = out-of-grid bucket)
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)ncells) ++bits;
  // 10-bit digits: 2 passes up to 2^20 cells (8-bit digits when one pass
  // of them covers the keys)
  const int rb = bits <= 8 ? 8 : 10;
  const int passes = (bits + rb - 1) / rb;
  const int nblocks = (int)ceil_div(n, kSortTile);
  const uint32_t *kin = part->hash;
  const int32_t *vin = nullptr;  // identity on the first pass
  for (int p = 0; p < passes; ++p) {
    const bool to_dest = ((passes - 1 - p) % 2) == 0;
    uint32_t *kout = to_dest ? part->hash_sorted : ctx->keys_tmp;
    int32_t *vout = to_dest ? part->sorted : ctx->vals_tmp;
    const int shift = p * rb;
    const int radix = 1 << rb;
    uint32_t *totals = ctx->hist + (size_t)radix * nblocks;
    if (rb == 8) {
      PMPM_LAUNCH(radix_hist_kernel<8>, (::pmpm::kname("radix_hist_kernel", {(long)(8)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, n, shift, ctx->hist, nblocks);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(scan_columns_kernel, (::pmpm::kname("scan_columns_kernel", {})), dim3(radix), dim3(256), 0, s, ctx->hist, nblocks, totals);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(radix_scatter_kernel<8>, (::pmpm::kname("radix_scatter_kernel", {(long)(8)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                         shift, ctx->hist, nblocks, totals);
    } else {
      PMPM_LAUNCH(radix_hist_kernel<10>, (::pmpm::kname("radix_hist_kernel", {(long)(10)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, n, shift, ctx->hist,
                         nblocks);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(scan_columns_kernel, (::pmpm::kname("scan_columns_kernel", {})), dim3(radix), dim3(256), 0, s, ctx->hist, nblocks, totals);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(radix_scatter_kernel<10>, (::pmpm::kname("radix_scatter_kernel", {(long)(10)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                         shift, ctx->hist, nblocks, totals);
    }
    PMPM_LAUNCH_CHECK();
    kin = kout;
    vin = vout;
  }
  PMPM_LAUNCH(cell_range_kernel, (::pmpm::kname("cell_range_kernel", {})), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, part->hash_sorted, n, ncells,
                     part->cell_start, part->cell_end, ctx->counters);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

extern "C" int pmpm_partition(pmpm_context *ctx, int32_t dim, const pmpm_grid *grid, const float *x, int32_t n,
                              pmpm_spatial *part, pmpm_stream_t stream) {
  return partition_clearing(ctx, dim, grid, x, n, part, (hipStream_t)stream, ZeroLists{});
}
