// partition.hip -- spatial binning on gfx950: cell hash per particle, stable
// LSD radix sort of (hash, slot), cell start/end ranges. Replaces the
// reference's calculate_hashes / thrust::stable_sort_by_key / bin_particles
// (spatialpartition_inline.h:49-115, spatialpartition.cpp:77-136).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
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
// Sorts of at most this many keys with a device-side count (the hint
// partition's movers) run in one workgroup (small_sort_kernel); the
// multi-block radix kernels skip them.
constexpr int kSmallSort = 4096;

// n_dev (optional): the key count is min(n, *n_dev), read on the device (the
// grid is sized for n; the blocks past the count exit, and a count of at most
// kSmallSort is left to small_sort_kernel).
template <int RB>
__global__ __launch_bounds__(kSortBlock) void radix_hist_kernel(const uint32_t *__restrict__ keys, int n, int shift,
                                                                uint32_t *__restrict__ hist, int nblocks,
                                                                const uint32_t *__restrict__ n_dev) {
  constexpr int R = 1 << RB;
  __shared__ uint32_t h[R];
  const int t = threadIdx.x, lane = t & 63;
  if (n_dev) {
    n = min(n, (int)*n_dev);
    if (n <= kSmallSort || (int64_t)blockIdx.x * kSortTile >= n) return;  // block-uniform
  }
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
// With n_dev, only the columns' first ceil(*n_dev / kSortTile) entries (the
// blocks that hold keys) are scanned, and nothing for a count of at most
// kSmallSort.
__global__ __launch_bounds__(256) void scan_columns_kernel(uint32_t *__restrict__ hist, int nblocks,
                                                           uint32_t *__restrict__ totals,
                                                           const uint32_t *__restrict__ n_dev) {
  __shared__ uint32_t part[256];
  const int t = threadIdx.x;
  uint32_t *col = hist + (int64_t)blockIdx.x * nblocks;
  if (n_dev && *n_dev <= (uint32_t)kSmallSort) return;  // block-uniform
  const int nb = n_dev ? min(nblocks, (int)(((int64_t)*n_dev + kSortTile - 1) / kSortTile)) : nblocks;
  const int chunk = (nb + 255) / 256;
  const int b = t * chunk;
  const int e = min(b + chunk, nb);
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
    const uint32_t *__restrict__ totals, const uint32_t *__restrict__ n_dev) {
  constexpr int R = 1 << RB, W = kSortBlock / 64;
  static_assert(R <= kSortBlock, "one digit per thread");
  if (n_dev) {
    n = min(n, (int)*n_dev);
    if (n <= kSmallSort) return;  // block-uniform
  }
  if ((int64_t)blockIdx.x * kSortTile >= n) return;  // block-uniform: no keys here
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

// Stable LSD radix sort of at most kSmallSort keys in one workgroup, in LDS
// (10-bit digits, `passes` of them; values = input positions): the hint
// partition's movers when the count *n_dev is that small (the multi-block
// kernels take larger counts). Thread d owns digit d; ranking within a round
// of 1024 keys as in radix_scatter_kernel; only the rounds that hold keys run.
__global__ __launch_bounds__(kSortBlock) void small_sort_kernel(const uint32_t *__restrict__ keys_in,
                                                                const uint32_t *__restrict__ n_dev, int passes,
                                                                uint32_t *__restrict__ keys_out,
                                                                int32_t *__restrict__ vals_out) {
  constexpr int RB = 10, R = 1 << RB, W = kSortBlock / 64;
  static_assert(R == kSortBlock, "one digit per thread");
  __shared__ uint32_t sk[2][kSmallSort];
  __shared__ int32_t sv[2][kSmallSort];
  __shared__ uint32_t wcnt[W][R];
  __shared__ uint32_t wsum[W];
  const uint32_t cnt = *n_dev;
  if (cnt > (uint32_t)kSmallSort) return;  // block-uniform: the multi-block path
  const int m = (int)cnt;
  const int rounds = (m + kSortBlock - 1) / kSortBlock;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < rounds; ++r) {
    const int i = r * kSortBlock + t;
    if (i < m) {
      sk[0][i] = keys_in[i];
      sv[0][i] = i;
    }
  }
  int src = 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = RB * p, dst = src ^ 1;
    // digit totals (run lengths, as radix_hist_kernel), then their exclusive
    // scan: thread t owns digit t
#pragma unroll
    for (int q = 0; q < W; ++q) wcnt[q][t] = 0u;
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
      const int i = r * kSortBlock + t;
      const bool valid = i < m;
      const uint32_t d = valid ? (sk[src][i] >> shift) & (R - 1) : 0u;
      const uint32_t prev = (uint32_t)__shfl_up((int)d, 1, 64);
      const bool head = valid && (lane == 0 || prev != d);
      const uint64_t heads = __ballot(head);
      const int nvalid = __popcll(__ballot(valid));
      if (head) {
        const uint64_t later = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const int next = later ? __ffsll((unsigned long long)later) - 1 : nvalid;
        atomicAdd(&wcnt[0][d], (uint32_t)(next - lane));
      }
    }
    __syncthreads();
    const uint32_t tot = wcnt[0][t];
    uint32_t incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = (uint32_t)__shfl_up((int)incl, o, 64);
      if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0u;
    for (int q = 0; q < w; ++q) before += wsum[q];
    uint32_t base = before + incl - tot;  // this thread's digit
    for (int r = 0; r < rounds; ++r) {
#pragma unroll
      for (int q = 0; q < W; ++q) wcnt[q][t] = 0u;
      __syncthreads();
      const int i = r * kSortBlock + t;
      const bool valid = i < m;
      const uint32_t key = valid ? sk[src][i] : 0u;
      const int32_t val = valid ? sv[src][i] : 0;
      const uint32_t d = (key >> shift) & (R - 1);
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < RB; ++bit) {
        const bool on = (d >> bit) & 1u;
        const uint64_t mk = __ballot(on);
        peers &= on ? mk : ~mk;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
      if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
      __syncthreads();
      {
        uint32_t run = base;
#pragma unroll
        for (int q = 0; q < W; ++q) {
          const uint32_t c = wcnt[q][t];
          wcnt[q][t] = run;
          run += c;
        }
        base = run;
      }
      __syncthreads();
      if (valid) {
        const uint32_t pos = wcnt[w][d] + rank;
        if (pos < (uint32_t)m) {
          sk[dst][pos] = key;
          sv[dst][pos] = val;
        }
      }
      __syncthreads();
    }
    src = dst;
  }
  for (int r = 0; r < rounds; ++r) {
    const int i = r * kSortBlock + t;
    if (i < m) {
      keys_out[i] = sk[src][i];
      vals_out[i] = sv[src][i];
    }
  }
}

// ---- Partition from a sorted hint ----------------------------------------
// pmpm_step hands the particles to the next partition in the previous
// partition's (cell, slot) order, and that partition's sorted keys are still
// in hash_sorted. Used as a hint h[i], they equal the new key k[i] for every
// particle that kept its cell. If h is sorted, the particles with k[i] ==
// h[i] ("stayers") are in (key, slot) order among themselves already, so only
// the others ("movers") need sorting, and the stable order is a merge per
// cell c, with S[i] = stayers before slot i:
//   first slot of c:         out_start[c] = particles with key < c
//   stayer i:                out_start[c] + (S[i] - stayers with key < c)
//                            + movers of c before slot i
//   r-th mover (sorted):     out_start[c] + (r - movers with key < c)
//                            + clamp(S[i] - stayers with key < c, 0, stayers of c)
// (stayers are sorted by key, so the first S[i] of them are those before i).
// The movers are compacted in slot order and sorted by the radix passes; any
// hint gives the stable sort (an unsorted hint makes every particle a mover),
// a good one a short mover list.
constexpr int kHintBlock = 1024;   // particles, or cells, per workgroup
constexpr int kHintCellBits = 10;  // cells per block of the cell scan: 1 << 10
constexpr int kHintSlots = 4;      // cell blocks a classify workgroup sums in LDS
static_assert((1 << kHintCellBits) == kHintBlock, "one cell per thread in the cell scan");

// Exclusive scan over a workgroup of kHintBlock threads; every thread gets its
// prefix and the total. `w` holds kHintBlock / 64 words of LDS.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *w, uint32_t &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)incl, o, 64);
    if (lane >= o) incl += u;
  }
  __syncthreads();  // w may still be read by an earlier call
  if (lane == 63) w[wv] = incl;
  __syncthreads();
  uint32_t before = 0u, all = 0u;
#pragma unroll
  for (int q = 0; q < kHintBlock / 64; ++q) {
    const uint32_t x = w[q];
    before += q < wv ? x : 0u;
    all += x;
  }
  total = all;
  return before + incl - v;
}

// Sum over a workgroup of kHintBlock threads (every thread gets it).
__device__ inline uint32_t block_sum(uint32_t v, uint32_t *w) {
  uint32_t total;
  (void)block_exclusive_scan(v, w, total);
  return total;
}

// The same scan of two 32-bit counts at once, packed as (high << 32) | low
// (no carry crosses: each count stays below 2^31). `w` holds kHintBlock / 64
// 64-bit words of LDS.
__device__ inline uint64_t block_exclusive_scan2(uint64_t v, uint64_t *w, uint64_t &total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t u = (uint64_t)__shfl_up((unsigned long long)incl, o, 64);
    if (lane >= o) incl += u;
  }
  __syncthreads();  // w may still be read by an earlier call
  if (lane == 63) w[wv] = incl;
  __syncthreads();
  uint64_t before = 0u, all = 0u;
#pragma unroll
  for (int q = 0; q < kHintBlock / 64; ++q) {
    const uint64_t x = w[q];
    before += q < wv ? x : 0u;
    all += x;
  }
  total = all;
  return before + incl - v;
}

// Cell key per particle (as hash_kernel), its class against the hint, and the
// counts: per cell and per block of cells, stayers and movers (a run of lanes
// with one key and one class adds its length once); stayers per workgroup.
// flags[0] is set when the hint is not sorted. Blocks past the particles
// only clear `clear`.
template <int D>
__global__ __launch_bounds__(kHintBlock) void hint_classify_kernel(
    const float *__restrict__ x, int n, pmpm_grid g, const uint32_t *__restrict__ hint, uint32_t *__restrict__ hash,
    int32_t *__restrict__ bins, uint32_t *__restrict__ cnt, uint32_t *__restrict__ cbt, uint32_t *__restrict__ pblk,
    int npb, int ncb, uint32_t *__restrict__ flags, ZeroLists clear) {
  __shared__ uint32_t w[kHintBlock / 64];
  // per-class sums for the kHintSlots blocks of cells from the workgroup's
  // first key on (one global atomic per slot and class at the end: the
  // workgroups of a stretch of particles share these counters)
  __shared__ uint32_t slot_sum[kHintSlots][2];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int q = 0; q < clear.count; ++q)
    for (int64_t k = t; k < clear.n[q]; k += stride) clear.p[q][k] = 0u;
  if ((int)blockIdx.x >= npb) return;  // block-uniform
  if (threadIdx.x < 2 * kHintSlots) (&slot_sum[0][0])[threadIdx.x] = 0u;
  const int lane = threadIdx.x & 63;
  const int i = (int)t;
  const bool valid = i < n;  // a prefix of each wave
  uint32_t key = 0u;
  bool stay = false;
  if (valid) {
    int b[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) b[a] = (int)__fmul_rn(__fsub_rn(x[a * n + i], g.origin[a]), g.inv_cell_size);
    int64_t h = b[0];
    if (D > 1) h += (int64_t)b[1] * g.num_cells[0];
    if (D > 2) h += (int64_t)b[2] * g.num_cells[0] * g.num_cells[1];
    key = (h < 0 || h >= g.num_cells_total) ? (uint32_t)g.num_cells_total : (uint32_t)h;
    hash[i] = key;
    if (bins) {
#pragma unroll
      for (int a = 0; a < D; ++a) bins[a * n + i] = b[a];
    }
    const uint32_t p = hint[i];
    if (i + 1 < n && p > hint[i + 1]) atomicOr(&flags[0], 1u);
    stay = key == p;
  }
  // the cell block of the workgroup's first particle (i0 < n here)
  const int i0 = (int)blockIdx.x * kHintBlock;
  uint32_t k0;
  {
    int b0[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < D; ++a) b0[a] = (int)__fmul_rn(__fsub_rn(x[a * n + i0], g.origin[a]), g.inv_cell_size);
    int64_t h0 = b0[0];
    if (D > 1) h0 += (int64_t)b0[1] * g.num_cells[0];
    if (D > 2) h0 += (int64_t)b0[2] * g.num_cells[0] * g.num_cells[1];
    k0 = (h0 < 0 || h0 >= g.num_cells_total) ? (uint32_t)g.num_cells_total : (uint32_t)h0;
  }
  const uint32_t cb0 = k0 >> kHintCellBits;
  __syncthreads();  // slot_sum cleared
  const uint32_t pk = (uint32_t)__shfl_up((int)key, 1, 64);
  const int ps = __shfl_up((int)stay, 1, 64);
  const bool head = valid && (lane == 0 || pk != key || ps != (int)stay);
  const uint64_t heads = __ballot(head);
  const int nvalid = __popcll(__ballot(valid));
  if (head) {
    const uint64_t later = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
    const int next = later ? __ffsll((unsigned long long)later) - 1 : nvalid;
    const uint32_t len = (uint32_t)(next - lane);
    const int cls = stay ? 0 : 1;
    atomicAdd(&cnt[2 * (int64_t)key + cls], len);
    const uint32_t cb = key >> kHintCellBits;
    if (cb >= cb0 && cb - cb0 < (uint32_t)kHintSlots) atomicAdd(&slot_sum[cb - cb0][cls], len);
    else atomicAdd(&cbt[2 * (int64_t)cb + cls], len);
  }
  const uint32_t stayers = block_sum(stay ? 1u : 0u, w);  // (its barriers order the slot sums)
  if (threadIdx.x == 0) pblk[blockIdx.x] = stayers;
  if (threadIdx.x < 2 * kHintSlots) {
    const uint32_t v = (&slot_sum[0][0])[threadIdx.x];
    const int sl = threadIdx.x >> 1, cls = threadIdx.x & 1;
    const int64_t cb = (int64_t)cb0 + sl;
    if (v && cb < ncb) atomicAdd(&cbt[2 * cb + cls], v);
  }
}

// Two kinds of workgroups, both after the classify pass:
// - blocks [0, ncb): 1024 cells each (cells 0..ncells, the last holding the
//   out-of-grid particles): the exclusive scans of the stayer and mover
//   counts give each cell's record {first slot, stayers before, stayers,
//   movers}, its range (cell_start/cell_end, 0 for an empty cell), and for
//   the last cell the out-of-grid count (counters[0]) and the mover count
//   (counters[2]); the per-cell counts are cleared for the next call.
// - blocks [ncb, ncb + npb): 1024 particles each: S[i] (the stayers before
//   slot i) with the stayer bit, and each mover's key and S[i] at its place
//   i - S[i] in the slot-ordered mover list.
// With an unsorted hint (flags[0]) every particle counts as a mover.
__global__ __launch_bounds__(kHintBlock) void hint_scan_kernel(
    int n, uint32_t ncells, int ncb, const uint32_t *__restrict__ hash, const uint32_t *__restrict__ hint,
    uint32_t *__restrict__ cnt, const uint32_t *__restrict__ cbt, const uint32_t *__restrict__ pblk,
    uint32_t *__restrict__ counters, int4 *__restrict__ info, int32_t *__restrict__ cstart,
    int32_t *__restrict__ cend, int32_t *__restrict__ ss, uint32_t *__restrict__ mk, int32_t *__restrict__ sm) {
  __shared__ uint64_t w2[kHintBlock / 64];
  const bool unsorted = counters[1] != 0u;
  const int t = threadIdx.x;
  if ((int)blockIdx.x < ncb) {
    const int b = blockIdx.x;
    const int64_t c = (int64_t)b * kHintBlock + t;
    const bool valid = c <= (int64_t)ncells;
    uint32_t st = 0u, mv = 0u;
    if (valid) {  // (loaded before the prefix sums' barriers)
      st = cnt[2 * c];
      mv = cnt[2 * c + 1];
    }
    uint64_t pp = 0u;  // (stayers << 32) | movers of the cell blocks before this one
    for (int j = t; j < b; j += kHintBlock) pp += ((uint64_t)cbt[2 * j] << 32) | cbt[2 * j + 1];
    uint64_t pt;
    (void)block_exclusive_scan2(pp, w2, pt);
    uint32_t before_s = (uint32_t)(pt >> 32);
    uint32_t before_m = (uint32_t)pt;
    if (unsorted) {
      before_m += before_s;
      before_s = 0u;
    }
    if (valid) {
      cnt[2 * c] = 0u;
      cnt[2 * c + 1] = 0u;
    }
    if (unsorted) {
      mv += st;
      st = 0u;
    }
    uint64_t tot;
    const uint64_t smx = block_exclusive_scan2(((uint64_t)st << 32) | mv, w2, tot);
    const uint32_t sx = (uint32_t)(smx >> 32), mx = (uint32_t)smx;
    if (!valid) return;
    const uint32_t stay_before = before_s + sx, mover_before = before_m + mx;
    const uint32_t start = stay_before + mover_before, count = st + mv;
    info[c] = make_int4((int)start, (int)stay_before, (int)st, (int)mv);
    if (c < (int64_t)ncells) {
      cstart[c] = count ? (int32_t)start : 0;
      cend[c] = count ? (int32_t)(start + count) : 0;
    } else {
      counters[0] = count;
      counters[2] = mover_before + mv;
    }
    return;
  }
  const int b = (int)blockIdx.x - ncb;
  const int i = b * kHintBlock + t;
  const bool valid = i < n;
  const uint32_t k = valid ? hash[i] : 0u;  // (loaded before the prefix sum's barriers)
  const bool stay = valid && !unsorted && k == hint[i];
  // (stayers of the blocks before << 32) | this thread's stayer flag, in one scan
  uint64_t pv = stay ? 1u : 0u;
  if (!unsorted)
    for (int j = t; j < b; j += kHintBlock) pv += (uint64_t)pblk[j] << 32;
  uint64_t tot;
  const uint64_t ex = block_exclusive_scan2(pv, w2, tot);
  const uint32_t S = (uint32_t)(tot >> 32) + (uint32_t)ex;
  if (!valid) return;
  ss[i] = (int32_t)(S | (stay ? 0x80000000u : 0u));
  if (!stay) {
    const uint32_t q = (uint32_t)i - S;
    mk[q] = k;
    sm[q] = (int32_t)S;
  }
}

// The merge: blocks [0, npb) place the stayers (slot order), the blocks
// after them the sorted movers (mover r: key mkey[r], place mval[r] in the
// slot-ordered list; a loop over the movers); writes sorted and hash_sorted.
// Block 0 also clears the per-block cell counts and the unsorted flag for
// the next call.
__global__ __launch_bounds__(kHintBlock) void hint_place_kernel(
    int n, uint32_t ncells, int npb, int ncb, const uint32_t *__restrict__ hash, const int32_t *__restrict__ ss,
    const int4 *__restrict__ info, const uint32_t *__restrict__ mkey, const int32_t *__restrict__ mval,
    const int32_t *__restrict__ sm, uint32_t *__restrict__ counters, uint32_t *__restrict__ cbt,
    uint32_t *__restrict__ hash_sorted, int32_t *__restrict__ sorted) {
  const int t = threadIdx.x;
  if (blockIdx.x == 0) {
    for (int j = t; j < 2 * ncb; j += kHintBlock) cbt[j] = 0u;
    if (t == 0) counters[1] = 0u;
  }
  const int m = (int)min(counters[2], (uint32_t)n);
  if ((int)blockIdx.x < npb) {
    const int i = (int)blockIdx.x * kHintBlock + t;
    if (i >= n) return;
    const int32_t v = ss[i];
    if (v >= 0) return;  // a mover
    const int S = v & 0x7fffffff;
    const uint32_t k = hash[i];
    if (k > ncells) return;
    const int4 c = info[k];
    const int mb = c.x - c.y;  // movers with a smaller key
    // movers of this cell before slot i: their places in the slot-ordered
    // list are below i - S (ascending within the cell)
    const int lim = i - S;
    int lo = 0, hi = c.w;
    if (mb < 0 || mb + hi > m) return;  // (inconsistent counts: leave the slot)
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (mval[mb + mid] < lim) lo = mid + 1;
      else hi = mid;
    }
    const int pos = c.x + (S - c.y) + lo;
    if ((unsigned)pos >= (unsigned)n) return;
    sorted[pos] = i;
    hash_sorted[pos] = k;
    return;
  }
  const int nmw = ((int)gridDim.x - npb) * kHintBlock;  // mover threads
  for (int r = ((int)blockIdx.x - npb) * kHintBlock + t; r < m; r += nmw) {
    const int q = mval[r];
    const uint32_t k = mkey[r];
    if ((unsigned)q >= (unsigned)m || k > ncells) continue;
    const int4 c = info[k];
    const int S = sm[q];
    const int st = min(max(S - c.y, 0), c.z);
    const int pos = c.x + (r - (c.x - c.y)) + st;
    if ((unsigned)pos >= (unsigned)n) continue;
    sorted[pos] = S + q;
    hash_sorted[pos] = k;
  }
}

}  // namespace pmpm

using namespace pmpm;

extern "C" const char *pmpm_last_error(void) { return g_last_error; }

extern "C" int pmpm_context_create(pmpm_context **ctx) {
  if (!ctx) return PMPM_ERR_INVALID;
  pmpm_context *c = new (std::nothrow) pmpm_context();
  if (!c) return PMPM_ERR_INVALID;
  if (!hip_ok(hipGetDevice(&c->device), "hipGetDevice")) { delete c; return PMPM_ERR_HIP; }
  const char *radix_only = getenv("PMPM_PARTITION_RADIX");
  c->sort_by_hint = !(radix_only && radix_only[0] == '1');
  const char *hint_min = getenv("PMPM_PARTITION_HINT_MIN");
  if (hint_min && hint_min[0]) c->hint_min_particles = (int32_t)strtol(hint_min, nullptr, 10);
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
  (void)hipFree(ctx->hint_cnt);
  (void)hipFree(ctx->hint_cbt);
  (void)hipFree(ctx->hint_info);
  (void)hipFree(ctx->hint_part);
  (void)hipFree(ctx->keys_tmp);
  (void)hipFree(ctx->vals_tmp);
  (void)hipFree(ctx->hist);
  (void)hipFree(ctx->counters);
  delete ctx;
  return PMPM_OK;
}

static int reserve_hint(pmpm_context *ctx, int n, uint32_t ncells, hipStream_t s);

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
  if (ctx->sort_by_hint && max_particles >= ctx->hint_min_particles && max_particles > 0 && max_cells > 0) {
    // the hint partition's workspace too, so that stepping within these
    // sizes allocates nothing (the counts are cleared on the null stream,
    // which is waited for)
    int rc = reserve_hint(ctx, max_particles, (uint32_t)max_cells, nullptr);
    if (rc != PMPM_OK) return rc;
    PMPM_HIP_CHECK(hipStreamSynchronize(nullptr));
  }
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

// LSD radix sort of (key, position) pairs by keys in [0, ncells]: n keys
// from kin (n_dev: the count is min(n, *n_dev), read on the device), values
// = positions in kin; the passes alternate between (ktmp, vtmp) and (kdst,
// vdst), the last one writing (kdst, vdst). kin may alias neither.
static int radix_sort_pairs(pmpm_context *ctx, const uint32_t *kin, uint32_t *ktmp, int32_t *vtmp, uint32_t *kdst,
                            int32_t *vdst, int n, const uint32_t *n_dev, uint32_t ncells, hipStream_t s) {
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)ncells) ++bits;
  // 10-bit digits: 2 passes up to 2^20 cells (8-bit digits when one pass
  // of them covers the keys)
  const int rb = bits <= 8 ? 8 : 10;
  const int passes = (bits + rb - 1) / rb;
  const int nblocks = (int)ceil_div(n, kSortTile);
  const int32_t *vin = nullptr;  // identity on the first pass
  for (int p = 0; p < passes; ++p) {
    const bool to_dest = ((passes - 1 - p) % 2) == 0;
    uint32_t *kout = to_dest ? kdst : ktmp;
    int32_t *vout = to_dest ? vdst : vtmp;
    const int shift = p * rb;
    const int radix = 1 << rb;
    uint32_t *totals = ctx->hist + (size_t)radix * nblocks;
    if (rb == 8) {
      PMPM_LAUNCH(radix_hist_kernel<8>, (::pmpm::kname("radix_hist_kernel", {(long)(8)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, n, shift, ctx->hist, nblocks, n_dev);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(scan_columns_kernel, (::pmpm::kname("scan_columns_kernel", {})), dim3(radix), dim3(256), 0, s, ctx->hist, nblocks, totals, n_dev);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(radix_scatter_kernel<8>, (::pmpm::kname("radix_scatter_kernel", {(long)(8)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                  shift, ctx->hist, nblocks, totals, n_dev);
    } else {
      PMPM_LAUNCH(radix_hist_kernel<10>, (::pmpm::kname("radix_hist_kernel", {(long)(10)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, n, shift, ctx->hist,
                  nblocks, n_dev);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(scan_columns_kernel, (::pmpm::kname("scan_columns_kernel", {})), dim3(radix), dim3(256), 0, s, ctx->hist, nblocks, totals, n_dev);
      PMPM_LAUNCH_CHECK();
      PMPM_LAUNCH(radix_scatter_kernel<10>, (::pmpm::kname("radix_scatter_kernel", {(long)(10)})), dim3(nblocks), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                  shift, ctx->hist, nblocks, totals, n_dev);
    }
    PMPM_LAUNCH_CHECK();
    kin = kout;
    vin = vout;
  }
  return PMPM_OK;
}

// Grows a context buffer of 32-bit words; a new buffer is zero-filled on s
// when `zero` (the counts that must be zero between calls).
static int ensure_words(float *&buf, size_t &cap, size_t words, bool zero, hipStream_t s) {
  if (words <= cap) return PMPM_OK;
  int rc = ensure_buffer(buf, cap, words);
  if (rc != PMPM_OK) return rc;
  return zero ? zero_fill(buf, (int64_t)cap, s) : PMPM_OK;
}

// Words of the per-particle hint workspace: the per-block stayer counts, then
// five arrays of n words, each starting on a 256-byte boundary.
static size_t hint_part_words(int n) {
  const size_t npb = (size_t)ceil_div(n, kHintBlock);
  return (npb + 63) / 64 * 64 + 4 * (((size_t)n + 63) / 64 * 64) + (size_t)n;
}

// Grows the hint workspace to n particles on ncells cells (the counts are
// zero-filled on s when new).
static int reserve_hint(pmpm_context *ctx, int n, uint32_t ncells, hipStream_t s) {
  const int ncb = (int)ceil_div((int64_t)ncells + 1, kHintBlock);
  const int npb = (int)ceil_div(n, kHintBlock);
  int rc;
  if ((rc = ensure_words(ctx->hint_cnt, ctx->hint_cnt_cap, 2 * ((size_t)ncells + 1), true, s)) != PMPM_OK) return rc;
  if ((rc = ensure_words(ctx->hint_cbt, ctx->hint_cbt_cap, 2 * (size_t)ncb, true, s)) != PMPM_OK) return rc;
  if ((rc = ensure_words(ctx->hint_info, ctx->hint_info_cap, 4 * ((size_t)ncells + 1), false, s)) != PMPM_OK)
    return rc;
  return ensure_words(ctx->hint_part, ctx->hint_part_cap, hint_part_words(n), false, s);
}

// The partition from the hint in part->hash_sorted (see hint_classify_kernel).
static int partition_by_hint(pmpm_context *ctx, int dim, const pmpm_grid &g, const float *x, int n,
                             pmpm_spatial *part, hipStream_t s, const ZeroLists &extra) {
  const uint32_t ncells = (uint32_t)g.num_cells_total;
  const int ncb = (int)ceil_div((int64_t)ncells + 1, kHintBlock);
  const int npb = (int)ceil_div(n, kHintBlock);
  int rc;
  if ((rc = reserve_hint(ctx, n, ncells, s)) != PMPM_OK) return rc;
  uint32_t *pblk = reinterpret_cast<uint32_t *>(ctx->hint_part);
  const size_t o = ((size_t)npb + 63) / 64 * 64;  // 256-byte aligned arrays of n words
  const size_t na = ((size_t)n + 63) / 64 * 64;
  int32_t *ss = reinterpret_cast<int32_t *>(ctx->hint_part) + o;
  uint32_t *mk = reinterpret_cast<uint32_t *>(ss + na);
  int32_t *sm = reinterpret_cast<int32_t *>(mk + na);
  uint32_t *k2 = reinterpret_cast<uint32_t *>(sm + na);
  int32_t *v2 = reinterpret_cast<int32_t *>(k2 + na);
  if ((size_t)((v2 + n) - reinterpret_cast<int32_t *>(ctx->hint_part)) != hint_part_words(n) ||
      hint_part_words(n) > ctx->hint_part_cap) {
    note_error("partition: hint workspace layout");
    return PMPM_ERR_INVALID;
  }
  uint32_t *cnt = reinterpret_cast<uint32_t *>(ctx->hint_cnt);
  uint32_t *cbt = reinterpret_cast<uint32_t *>(ctx->hint_cbt);
  int4 *info = reinterpret_cast<int4 *>(ctx->hint_info);
  // classify: enough workgroups for the particles and for the clears
  int64_t words = 0;
  for (int q = 0; q < extra.count; ++q) words = extra.n[q] > words ? extra.n[q] : words;
  int64_t nb = npb;
  const int64_t nb_clear = ceil_div(words, (int64_t)kHintBlock * 4);
  if (nb_clear > nb) nb = nb_clear < 4096 ? nb_clear : (nb > 4096 ? nb : 4096);
  const unsigned blocks = (unsigned)nb;
  switch (dim) {
    case 1: PMPM_LAUNCH(hint_classify_kernel<1>, (::pmpm::kname("hint_classify_kernel", {(long)(1)})), dim3(blocks), dim3(kHintBlock), 0, s, x, n, g, part->hash_sorted, part->hash, part->bins, cnt, cbt, pblk, npb, ncb, ctx->counters + 1, extra); break;
    case 2: PMPM_LAUNCH(hint_classify_kernel<2>, (::pmpm::kname("hint_classify_kernel", {(long)(2)})), dim3(blocks), dim3(kHintBlock), 0, s, x, n, g, part->hash_sorted, part->hash, part->bins, cnt, cbt, pblk, npb, ncb, ctx->counters + 1, extra); break;
    default: PMPM_LAUNCH(hint_classify_kernel<3>, (::pmpm::kname("hint_classify_kernel", {(long)(3)})), dim3(blocks), dim3(kHintBlock), 0, s, x, n, g, part->hash_sorted, part->hash, part->bins, cnt, cbt, pblk, npb, ncb, ctx->counters + 1, extra); break;
  }
  PMPM_LAUNCH_CHECK();
  PMPM_LAUNCH(hint_scan_kernel, (::pmpm::kname("hint_scan_kernel", {})), dim3((unsigned)(ncb + npb)), dim3(kHintBlock), 0, s, n, ncells, ncb, part->hash,
              part->hash_sorted, cnt, cbt, pblk, ctx->counters, info, part->cell_start, part->cell_end, ss, mk, sm);
  PMPM_LAUNCH_CHECK();
  // the movers sorted by (key, slot): one workgroup when at most kSmallSort
  // of them, else the radix passes (each launch leaves the other case alone)
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)ncells) ++bits;
  PMPM_LAUNCH(small_sort_kernel, (::pmpm::kname("small_sort_kernel", {})), dim3(1), dim3(kSortBlock), 0, s, mk, ctx->counters + 2, (bits + 9) / 10, k2, v2);
  PMPM_LAUNCH_CHECK();
  if ((rc = radix_sort_pairs(ctx, mk, ctx->keys_tmp, ctx->vals_tmp, k2, v2, n, ctx->counters + 2, ncells, s)) !=
      PMPM_OK)
    return rc;
  const int nmb = npb < 64 ? npb : 64;  // workgroups of the mover loop
  PMPM_LAUNCH(hint_place_kernel, (::pmpm::kname("hint_place_kernel", {})), dim3((unsigned)(npb + nmb)), dim3(kHintBlock), 0, s, n, ncells, npb, ncb, part->hash,
              ss, info, k2, v2, sm, ctx->counters, cbt, part->hash_sorted, part->sorted);
  PMPM_LAUNCH_CHECK();
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
  if (ctx->sort_by_hint && n >= ctx->hint_min_particles) return partition_by_hint(ctx, dim, g, x, n, part, s, extra);
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
  if ((rc = radix_sort_pairs(ctx, part->hash, ctx->keys_tmp, ctx->vals_tmp, part->hash_sorted, part->sorted, n, nullptr,
                             ncells, s)) != PMPM_OK)
    return rc;
  PMPM_LAUNCH(cell_range_kernel, (::pmpm::kname("cell_range_kernel", {})), dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, part->hash_sorted, n, ncells,
                     part->cell_start, part->cell_end, ctx->counters);
  PMPM_LAUNCH_CHECK();
  return PMPM_OK;
}

extern "C" int pmpm_partition(pmpm_context *ctx, int32_t dim, const pmpm_grid *grid, const float *x, int32_t n,
                              pmpm_spatial *part, pmpm_stream_t stream) {
  return partition_clearing(ctx, dim, grid, x, n, part, (hipStream_t)stream, ZeroLists{});
}
