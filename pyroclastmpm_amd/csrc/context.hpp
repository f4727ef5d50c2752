// context.hpp -- private definition of pmpm_context and host-side helpers
// shared by the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <string>

#include "../../include/pmpm.h"

struct pmpm_context {
  int device = 0;
  int32_t cap_particles = 0;
  int32_t cap_cells = 0;
  uint32_t *keys_tmp = nullptr;  // radix sort ping-pong keys
  int32_t *vals_tmp = nullptr;   // radix sort ping-pong values
  uint32_t *hist = nullptr;      // [radix][blocks] digit histograms
  size_t hist_len = 0;
  uint32_t *counters = nullptr;  // [4]: out-of-grid particles, hint unsorted flag, movers, -
  // P2G tile partials: per tile, the LDS node block of each accumulated field
  // (plain stores; the grid kernel sums the <= 2^D tiles covering a node)
  float *partials = nullptr;
  size_t partials_cap = 0;       // floats
  int32_t *tile_flags = nullptr;  // 1 = tile had particles this launch
  size_t tile_flags_cap = 0;
  // wrap-around sums of typed cubic splines: per (wrap cell, lane) records
  // ([fields][cells][64] floats)
  float *wrap_rec = nullptr;
  size_t wrap_rec_cap = 0;
  // a byte per node: the typed-cubic wrap-around sums reach it this pass
  // (allocated in words of 4 bytes)
  float *wrap_touch = nullptr;
  size_t wrap_touch_cap = 0;
  // the nodes the typed-cubic wrap-around sums reach this pass (each once):
  // kWrapListShards counts, then as many parts of node indices
  int32_t *wrap_list = nullptr;
  size_t wrap_list_cap = 0;
  // a byte per node: the components the last grid pass's NodeDomain
  // condition set to zero (words of 4 bytes), read by the adjoint
  float *clamp_mask = nullptr;
  size_t clamp_mask_cap = 0;
  // adjoint step: per-particle scatter records (15 floats, cell order) and
  // node adjoints (one record of 8 floats per node, 4 in 1D)
  float *adj_rec = nullptr;
  size_t adj_rec_cap = 0;
  float *adj_nodes = nullptr;
  size_t adj_nodes_cap = 0;
  // per-workgroup partial sums of the material parameter adjoints
  float *adj_pb = nullptr;
  size_t adj_pb_cap = 0;
  // partition from a sorted hint (the previous call's sorted keys): false
  // always runs the full radix sort (PMPM_PARTITION_RADIX=1 at creation);
  // below hint_min_particles particles the radix sort runs too (its fewer
  // launches win there; PMPM_PARTITION_HINT_MIN at creation)
  bool sort_by_hint = true;
  int32_t hint_min_particles = 131072;
  // per cell (ncells + 1, the last for the out-of-grid particles): counts
  // of stayers and movers (zero between calls), and per block of 1024 cells
  float *hint_cnt = nullptr;
  size_t hint_cnt_cap = 0;
  float *hint_cbt = nullptr;
  size_t hint_cbt_cap = 0;
  // per cell: first slot, stayers before the cell, stayers, movers
  float *hint_info = nullptr;
  size_t hint_info_cap = 0;
  // stayers per block of 1024 particles, then five arrays of n words:
  // stayers before each particle (sign bit: a stayer), the movers' keys in
  // slot order, the stayers before each mover, the sorted movers' keys and
  // positions in the mover list
  float *hint_part = nullptr;
  size_t hint_part_cap = 0;
};

namespace pmpm {
// Grows the tile-partial workspace (hipMalloc on first use of a larger grid;
// run one step before capturing a graph).
int ensure_partials(pmpm_context *ctx, size_t floats, size_t tiles);
// Grows a context-owned float buffer to at least `floats`.
int ensure_buffer(float *&buf, size_t &cap, size_t floats);
}  // namespace pmpm

namespace pmpm {

constexpr int kSortBlock = 1024;  // 16 waves rank 1024 keys per round
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kMaxRadixBits = 10;  // digits of 8 bits (keys < 2^8) or 10 bits
constexpr int kMaxRadix = 1 << kMaxRadixBits;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A grid the kernels can address: at least one node, fewer than
// PMPM_MAX_GRID_NODES (32-bit byte offsets into the node records).
inline bool grid_ok(const pmpm_grid *g) {
  return g && g->num_cells_total > 0 && g->num_cells_total < PMPM_MAX_GRID_NODES;
}

// Records the failing HIP call of the calling thread (pmpm_last_error).
void note_hip_error(hipError_t e, const char *what);
// Records a non-HIP failure of the calling thread (pmpm_last_error).
void note_error(const char *what);
inline bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  note_hip_error(e, what);
  return false;
}

#define PMPM_HIP_CHECK(expr)                              \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) {                               \
      ::pmpm::note_hip_error(e_, #expr);                  \
      return PMPM_ERR_HIP;                                \
    }                                                     \
  } while (0)

#define PMPM_LAUNCH_CHECK()                               \
  do {                                                    \
    hipError_t e_ = hipGetLastError();                    \
    if (e_ != hipSuccess) {                               \
      ::pmpm::note_hip_error(e_, "kernel launch");        \
      return PMPM_ERR_HIP;                                \
    }                                                     \
  } while (0)

// Kernel timing (pmpm_timing_enable): a TimedRegion around a launch records
// a HIP event pair on the launch stream when the calling thread has timing on;
// the name is only formed then.
bool timing_enabled();
std::string kname(const char *base, std::initializer_list<long> targs);
class TimedRegion {
 public:
  template <class F> TimedRegion(hipStream_t s, F &&name) {
    if (timing_enabled()) start(s, name());
  }
  ~TimedRegion() {
    if (on_) finish();
  }
  TimedRegion(const TimedRegion &) = delete;
  TimedRegion &operator=(const TimedRegion &) = delete;

 private:
  void start(hipStream_t s, std::string name);
  void finish();
  bool on_ = false;
  hipStream_t s_ = nullptr;
  hipEvent_t a_ = nullptr, b_ = nullptr;
  std::string name_;
};
// hipLaunchKernelGGL with the launch timed under `name` (an expression giving
// a std::string, in parentheses) when timing is on.
#define PMPM_LAUNCH(kernel, name, grid, block, shm, stream, ...)                 \
  do {                                                                           \
    ::pmpm::TimedRegion pmpm_timed_((stream), [&]() { return std::string name; }); \
    hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);           \
  } while (0)

// Fill n floats/ints with zero (a kernel, not hipMemsetAsync, so the step can
// be captured in a HIP graph).
int zero_fill(void *ptr, int64_t n_words, hipStream_t s);
constexpr int kZeroLists = 6;
constexpr int kWrapListShards = 32;  // parts of the wrap-around node list (a power of two)
struct ZeroLists {  // buffers of 32-bit words to clear in one launch
  uint32_t *p[kZeroLists];
  int64_t n[kZeroLists];
  int count;
  void add(void *ptr, int64_t words) {
    if (ptr && words > 0 && count < kZeroLists) { p[count] = (uint32_t *)ptr; n[count] = words; ++count; }
  }
};
int zero_fill_lists(const ZeroLists &z, hipStream_t s);
// One step on the workspace of ctx (arguments already checked by
// step_args_ok); the body of pmpm_step.
int step_on(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid, int32_t musl,
            const pmpm_particles *a, pmpm_particles *b, pmpm_spatial *part, pmpm_nodes *nodes, hipStream_t s);
// Dimension, shape function, list sizes, and the node arrays of body forces.
bool params_ok(const pmpm_params *prm);
bool step_args_ok(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *a, const pmpm_particles *b,
                  const pmpm_spatial *part, const pmpm_nodes *nodes);
// Workspace the P2G passes of one step use on grid g: tile partials (floats)
// and tile flags, wrap-around records (floats) and wrap cells.
void p2g_workspace(const pmpm_params &prm, const pmpm_grid &g, size_t *partial_floats, size_t *tiles,
                   size_t *wrap_floats, size_t *wrap_cells, size_t *wrap_touch_words, size_t *wrap_list_words,
                   size_t *clamp_words);
// pmpm_partition whose first kernel also clears the buffers of `extra`.
int partition_clearing(pmpm_context *ctx, int dim, const pmpm_grid *grid, const float *x, int n, pmpm_spatial *part,
                       hipStream_t s, const ZeroLists &extra);

}  // namespace pmpm
