// noalloc.hip -- pmpm_step_noalloc: the whole step without a context and
// without allocating. Its workspace (the radix sort's ping-pong keys, slots and
// digit histograms, the P2G tile partials and flags, the wrap-around records
// and node flags, the grid pass's clamp record, the counters) is carved out of one file-scope device array of
// PMPM_NOALLOC_SCRATCH_BYTES, which each GPU holds for the life of the process.
// Every call on a GPU uses that GPU's copy: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "context.hpp"

namespace pmpm {
namespace {

static __device__ __align__(256) unsigned char scratch[PMPM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch[kMaxDevices] = {};

// The device address of this GPU's `scratch` (looked up once per device, with
// the device current; the array's name in host code is not that address).
int scratch_of_current_device(unsigned char **ptr) {
  int dev = 0;
  PMPM_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return PMPM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    PMPM_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<unsigned char *>(p);
  }
  *ptr = g_scratch[dev];
  return PMPM_OK;
}

// Bump allocator over the scratch bytes, 256-byte aligned pieces.
struct Carver {
  size_t used = 0;
  size_t take(size_t bytes) {
    const size_t at = used;
    used += (bytes + 255) / 256 * 256;
    return at;
  }
};

}  // namespace
}  // namespace pmpm

using namespace pmpm;

extern "C" int pmpm_step_noalloc(const pmpm_params *prm, const pmpm_grid *grid, int32_t musl,
                                 const pmpm_particles *a, pmpm_particles *b, pmpm_spatial *part, pmpm_nodes *nodes,
                                 pmpm_stream_t stream) {
  if (!step_args_ok(prm, grid, a, b, part, nodes)) return PMPM_ERR_INVALID;
  const size_t n = (size_t)(a->n > 0 ? a->n : 1);
  const size_t nblocks = (size_t)ceil_div((int64_t)n, kSortTile);
  const size_t hist_len = (size_t)kMaxRadix * nblocks + kMaxRadix;
  size_t partial_floats = 0, tiles = 0, wrap_floats = 0, wrap_cells = 0, touch_words = 0, list_words = 0,
         clamp_words = 0;
  p2g_workspace(*prm, *grid, &partial_floats, &tiles, &wrap_floats, &wrap_cells, &touch_words, &list_words,
                &clamp_words);
  Carver c;
  const size_t o_keys = c.take(n * sizeof(uint32_t));
  const size_t o_vals = c.take(n * sizeof(int32_t));
  const size_t o_hist = c.take(hist_len * sizeof(uint32_t));
  const size_t o_counters = c.take(4 * sizeof(uint32_t));
  const size_t o_partials = c.take(partial_floats * sizeof(float));
  const size_t o_flags = c.take(tiles * sizeof(int32_t));
  const size_t o_wrec = c.take(wrap_floats * sizeof(float));
  const size_t o_touch = c.take(touch_words * sizeof(float));
  const size_t o_list = c.take(list_words * sizeof(int32_t));
  const size_t o_clamp = c.take(clamp_words * sizeof(float));
  if (c.used > (size_t)PMPM_NOALLOC_SCRATCH_BYTES) {
    note_error("pmpm_step_noalloc: the step needs more workspace than PMPM_NOALLOC_SCRATCH_BYTES "
               "(fewer particles or a smaller grid, or pmpm_step with a context)");
    return PMPM_ERR_WORKSPACE;
  }
  unsigned char *base = nullptr;
  int rc = scratch_of_current_device(&base);
  if (rc != PMPM_OK) return rc;
  // a context over the scratch; every capacity is what this call needs, so no
  // entry point below grows (allocates) anything
  pmpm_context ctx;
  PMPM_HIP_CHECK(hipGetDevice(&ctx.device));
  ctx.cap_particles = (int32_t)n;
  ctx.cap_cells = grid->num_cells_total;
  ctx.keys_tmp = reinterpret_cast<uint32_t *>(base + o_keys);
  ctx.vals_tmp = reinterpret_cast<int32_t *>(base + o_vals);
  ctx.hist = reinterpret_cast<uint32_t *>(base + o_hist);
  ctx.hist_len = hist_len;
  ctx.counters = reinterpret_cast<uint32_t *>(base + o_counters);
  ctx.sort_by_hint = false;  // the full radix sort: its workspace is all carved above
  ctx.partials = reinterpret_cast<float *>(base + o_partials);
  ctx.partials_cap = partial_floats;
  ctx.tile_flags = reinterpret_cast<int32_t *>(base + o_flags);
  ctx.tile_flags_cap = tiles;
  ctx.wrap_rec = wrap_floats ? reinterpret_cast<float *>(base + o_wrec) : nullptr;
  ctx.wrap_rec_cap = wrap_floats;
  // (the step's first kernel clears the node flags and the list counters:
  // the scratch holds whatever an earlier call left)
  ctx.wrap_touch = touch_words ? reinterpret_cast<float *>(base + o_touch) : nullptr;
  ctx.wrap_touch_cap = touch_words;
  ctx.wrap_list = list_words ? reinterpret_cast<int32_t *>(base + o_list) : nullptr;
  ctx.wrap_list_cap = list_words;
  ctx.clamp_mask = reinterpret_cast<float *>(base + o_clamp);
  ctx.clamp_mask_cap = clamp_words;
  // (the context is a plain struct on this stack frame: nothing frees the
  // scratch pieces it points at)
  return step_on(&ctx, prm, grid, musl, a, b, part, nodes, (hipStream_t)stream);
}
