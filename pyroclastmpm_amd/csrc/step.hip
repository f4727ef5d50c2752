// step.hip -- one full USL / MUSL step through the C ABI (usl.cpp:69-115):
// partition -> P2G -> node BCs -> G2P + return maps, all enqueued on one
// stream with no host synchronisation, so the sequence can be captured in a
// HIP graph by the caller.
#include <hip/hip_runtime.h>

#include "context.hpp"

namespace pmpm {
int do_p2g(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &p,
           const pmpm_spatial &part, const pmpm_nodes &nodes, hipStream_t s, bool zeroed);
int do_g2p(const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in, const pmpm_particles &out,
           const pmpm_spatial &part, const pmpm_nodes &nodes, int mode, hipStream_t s);
int do_musl_map(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_particles &in,
                const pmpm_particles &out, const pmpm_spatial &part, const pmpm_nodes &nodes, hipStream_t s,
                bool zeroed);
int launch_grid(pmpm_context *ctx, const pmpm_params &prm, const pmpm_grid &g, const pmpm_nodes &nodes, int finalize,
                int b0, int b1, int remap, int gather, hipStream_t s, int clear_mom_nt);
}  // namespace pmpm

using namespace pmpm;

// The step on the workspace of `ctx` (pmpm_step, pmpm_step_noalloc).
int pmpm::step_on(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid, int32_t musl,
                  const pmpm_particles *a, pmpm_particles *b, pmpm_spatial *part, pmpm_nodes *nodes,
                  hipStream_t s) {
  // the partition's first kernel also clears the node arrays P2G sums into,
  // and in MUSL the first grid pass leaves mom_nt cleared for the re-map
  const int64_t nn = grid->num_cells_total;
  ZeroLists z{};
  z.add(nodes->mass, nn);
  z.add(nodes->mom, nn * prm->dim);
  z.add(nodes->mom_nt, nn * prm->dim);
  // the wrap-around node flags and list counters (once allocated): every
  // step starts from an empty list
  z.add(ctx->wrap_touch, (int64_t)ctx->wrap_touch_cap);
  z.add(ctx->wrap_list, ctx->wrap_list ? kWrapListShards : 0);
  int rc = partition_clearing(ctx, prm->dim, grid, a->x, a->n, part, s, z);
  if (rc != PMPM_OK) return rc;
  if ((rc = do_p2g(ctx, *prm, *grid, *a, *part, *nodes, s, true)) != PMPM_OK) return rc;
  if ((rc = launch_grid(ctx, *prm, *grid, *nodes, 1, 0, prm->num_bcs, 0, 1, s, musl ? 1 : 0)) != PMPM_OK) return rc;
  if (!musl) return do_g2p(*prm, *grid, *a, *b, *part, *nodes, 0, s);
  if ((rc = do_musl_map(ctx, *prm, *grid, *a, *b, *part, *nodes, s, true)) != PMPM_OK) return rc;
  if ((rc = launch_grid(ctx, *prm, *grid, *nodes, 0, 0, prm->num_bcs, 1, 2, s, 0)) != PMPM_OK) return rc;
  return do_g2p(*prm, *grid, *a, *b, *part, *nodes, 2, s);
}

bool pmpm::step_args_ok(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *a,
                        const pmpm_particles *b, const pmpm_spatial *part, const pmpm_nodes *nodes) {
  if (!params_ok(prm) || !grid_ok(grid) || !a || !b || !part || !nodes) return false;
  return a->n >= 0 && a->n == b->n && a->x != b->x;
}

extern "C" int pmpm_step(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid, int32_t musl,
                         const pmpm_particles *a, pmpm_particles *b, pmpm_spatial *part, pmpm_nodes *nodes,
                         pmpm_stream_t stream) {
  if (!ctx || !step_args_ok(prm, grid, a, b, part, nodes)) return PMPM_ERR_INVALID;
  return step_on(ctx, prm, grid, musl, a, b, part, nodes, (hipStream_t)stream);
}

extern "C" const char *pmpm_version(void) { return "pyroclastmpm_amd 0.1 (gfx950)"; }
