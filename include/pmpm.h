/* pmpm.h -- C ABI of the MI355X explicit-MPM hot path (libpmpm_hip.so).
 *
 * All pointers are device pointers unless stated. All work is enqueued
 * asynchronously on the given HIP stream (NULL = default stream); nothing in
 * an entry point synchronises the device unless it says so. Every entry point
 * returns PMPM_OK (0) or a negative PMPM_ERR_* code and never exits the
 * process. There are no process globals: time step, shape function, grid and
 * material parameters travel in the arguments, so independent contexts can run
 * side by side (one per scene).
 *
 * HBM layout (structure of arrays, component-major):
 *   particle vector field  float[dim][n]        v[a*n + i]
 *   particle D x D tensor  float[dim*dim][n]    F[(a*dim+b)*n + i] = F(a,b)
 *   particle stress        float[6][n]          xx, yy, zz, xy, xz, yz
 *                                               (symmetric Cauchy stress; the
 *                                               reference's 3x3 in all dims)
 *   node vector field      float[dim][nn]       x-fastest node index
 * Particles are stored in "slots"; id[slot] is the particle's original index.
 * Solvers keep slots physically sorted by cell (the G2P kernels write their
 * output in cell order), and the host API maps back through id[].
 */
#ifndef PMPM_H_
#define PMPM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *pmpm_stream_t; /* == hipStream_t */

enum {
  PMPM_OK = 0,
  PMPM_ERR_INVALID = -1,     /* bad argument (dimension, shape function, size) */
  PMPM_ERR_HIP = -2,         /* a HIP runtime call failed */
  PMPM_ERR_WORKSPACE = -3,   /* context workspace too small; call pmpm_context_reserve */
  PMPM_ERR_UNSUPPORTED = -4  /* valid request this build does not implement */
};

enum { PMPM_SF_LINEAR = 0, PMPM_SF_QUADRATIC = 1, PMPM_SF_CUBIC = 2 };
enum { PMPM_MAT_NONE = 0, PMPM_MAT_LINEAR_ELASTIC = 1, PMPM_MAT_NEWTON_FLUID = 2,
       PMPM_MAT_LOCAL_RHEO = 3, PMPM_MAT_DRUCKER_PRAGER = 4, PMPM_MAT_VON_MISES = 5,
       PMPM_MAT_MOHR_COULOMB = 6, PMPM_MAT_MODIFIED_CAM_CLAY = 7 };
enum { PMPM_BC_NONE = 0, PMPM_BC_GRAVITY = 1, PMPM_BC_NODE_DOMAIN = 2, PMPM_BC_RIGID_SDF = 4,
       PMPM_BC_BODY_FORCE = 5, PMPM_BC_PLANAR_DOMAIN = 6 };

/* material flags (DP) */
enum { PMPM_MATF_UPDATE_HISTORY = 1, PMPM_MATF_VELGRAD_IS_INCREMENT = 2,
       PMPM_MATF_APEX_GUARD = 4, PMPM_MATF_HAS_STRESS_REF = 8 };
/* particle flag bits */
enum { PMPM_PF_RIGID = 1, PMPM_PF_ACTIVE = 2 };

#define PMPM_MAX_MATERIALS 8
#define PMPM_MAX_BCS 8

/* Largest node count of a grid (num_cells_total < 2^27): the gathers address
 * the node velocity records (32 B per node in 3D) with 32-bit byte offsets.
 * Entry points given a larger grid return PMPM_ERR_INVALID. */
#define PMPM_MAX_GRID_NODES (1 << 27)

/* Background grid (types_common.h:250-291 Grid). num_cells = nodes per axis. */
typedef struct pmpm_grid {
  float origin[3];
  float cell_size;
  float inv_cell_size;
  int32_t num_cells[3];
  int32_t num_cells_total;
} pmpm_grid;

/* Particle state (particles.h:181-232 ParticlesContainer fields, as SoA). */
typedef struct pmpm_particles {
  int32_t n;
  float *x;          /* [dim][n] positions */
  float *v;          /* [dim][n] velocities */
  float *stress;     /* [6][n] */
  float *F;          /* [dim*dim][n] deformation gradient */
  float *L;          /* [dim*dim][n] velocity gradient (may be NULL: not stored) */
  float *mass;       /* [n] */
  float *volume;     /* [n] current volume */
  float *volume0;    /* [n] original volume */
  uint8_t *color;    /* [n] material id */
  uint8_t *flags;    /* [n] PMPM_PF_* */
  int32_t *id;       /* [n] original particle index of each slot */
  float *eps_e;      /* [6][n] elastic strain history (Drucker-Prager, von Mises,
                        Mohr-Coulomb, modified Cam-Clay) or NULL */
  float *stress_ref; /* [6][n] reference stress (Drucker-Prager, modified
                        Cam-Clay) or NULL */
  float *hist;       /* [2][n] plastic history scalars or NULL: row 0 the
                        accumulated plastic strain (von Mises, Mohr-Coulomb) or
                        the volumetric plastic strain (modified Cam-Clay), row 1
                        the preconsolidation pressure (modified Cam-Clay) */
} pmpm_particles;

/* Floats per node of the derived velocity records pmpm_nodes.vel. */
#define PMPM_NODE_VEL_STRIDE(dim) ((dim) == 3 ? 8 : 2 * (dim))

/* Node state (nodes.h:113-140 NodesContainer fields, as SoA). */
typedef struct pmpm_nodes {
  float *mass;   /* [nn] */
  float *mom;    /* [dim][nn] moments */
  float *mom_nt; /* [dim][nn] forward moments (USL), re-mapped moments (MUSL) */
  float *vel;    /* [nn][PMPM_NODE_VEL_STRIDE(dim)] derived per node, written by
                    the grid update for the gathers: v_nt = mom_nt / mass
                    (dim floats), then dv = (mom_nt - mom) / mass (dim floats),
                    zero padded; all zero where mass <= cutoff. No counterpart
                    in the reference (it divides per gather). */
  uint8_t *clamp; /* [nn] or NULL: the components a NodeDomain condition set to
                     zero in the last grid pass that applied node conditions
                     (bit a: mom, bit 3 + a: mom_nt), written by that pass and
                     read by pmpm_adjoint_step. NULL: the context keeps one
                     such record (enough when the adjoint recomputes the
                     forward). No counterpart in the reference. */
} pmpm_nodes;

/* Spatial partition (spatialpartition.h:77-132 SpatialPartition fields). */
typedef struct pmpm_spatial {
  uint32_t *hash;        /* [n] cell hash per slot (unsorted) */
  uint32_t *hash_sorted; /* [n] */
  int32_t *sorted;       /* [n] slot of the j-th particle in (hash, slot) order */
  int32_t *cell_start;   /* [num_cells_total] */
  int32_t *cell_end;     /* [num_cells_total]; empty cells hold 0 in both */
  int32_t *bins;         /* [dim][n] integer cell coordinates per slot, or NULL */
} pmpm_spatial;

typedef struct pmpm_material {
  int32_t kind;  /* PMPM_MAT_* */
  int32_t flags; /* PMPM_MATF_* */
  /* LINEAR_ELASTIC: lame, shear, bulk
   * NEWTON_FLUID:   viscosity, bulk_modulus, gamma
   * LOCAL_RHEO:     shear, lame, bulk, rho_c, mu_s, mu_2, I0, EPS
   * DRUCKER_PRAGER: bulk, shear, M, Pt, eta_overline, Vs
   * VON_MISES:      bulk, shear, yield_stress, H  (3D only)
   * MOHR_COULOMB:   bulk, shear, cohesion, friction angle, dilatancy angle
   *                 (radians), H
   * MODIFIED_CAM_CLAY: bulk, shear, M, lam, kap, Pt, Vs
   * The last three need L, eps_e and hist stored; their return maps run in a
   * pass after G2P (the reference's stress_update, solver.cpp:110-121). */
  float p[16];
} pmpm_material;

typedef struct pmpm_bc {
  int32_t kind; /* PMPM_BC_* */
  /* NODE_DOMAIN: face0_mode[3], face1_mode[3] (0 fixed, 1 slip, other none)
   * BODY_FORCE:  mode (0 "forces", 1 "moments": moments += values on masked
   *              nodes; 2 "fixed": moments_nt = values on masked nodes;
   *              bodyforce.cpp:28-87, bodyforce_inline.h:41-59) */
  int32_t i[6];
  /* GRAVITY: gravity[3] applied this step
   * RIGID_SDF: half[3], center[3], velocity[3], rotation[9] (row-major,
   *            body->world), band, advance (1 = move body by velocity*dt)
   * PLANAR_DOMAIN: face0 friction[3], face1 friction[3] in radians, used as
   *            the friction coefficients of the walls at the grid's origin
   *            and at the domain end (planardomain.cpp:27-33), end[3] = the
   *            domain end the walls sit at (the reference's Grid::end) */
  float f[24];
  /* BODY_FORCE: values [dim][num_cells_total] (device), mask [num_cells_total]
   * (device, nonzero = apply); NULL for the other kinds */
  const float *values;
  const uint8_t *mask;
} pmpm_bc;

/* Everything one step needs besides the buffers. */
typedef struct pmpm_params {
  int32_t dim;              /* 1, 2, 3 */
  int32_t shape_fn;         /* PMPM_SF_* */
  float dt;
  float alpha;              /* FLIP/PIC blend (usl.h:85-114) */
  float small_mass_cutoff;  /* nodes.h:140, 1e-9 */
  int32_t node_types_built; /* 0 reproduces a NodesContainer built before the
                               cubic spline was selected (all node types 0) */
  int32_t wrap_quirk;       /* 1: P2G bounds-checks only the linear node index,
                               as the reference does (usl_inline.h:122-126) */
  int32_t num_materials;
  pmpm_material materials[PMPM_MAX_MATERIALS];
  int32_t num_bcs;
  pmpm_bc bcs[PMPM_MAX_BCS];
} pmpm_params;

typedef struct pmpm_context pmpm_context;

/* ---- context / workspace -------------------------------------------------
 * A context owns the scratch memory of the spatial sort for up to
 * max_particles particles and max_cells cells on the current device (and the
 * P2G, wrap-around and adjoint workspace, grown on first use). Its work is
 * stream-ordered: calls on one context must be issued to one stream, or
 * ordered between streams by the caller; overlapping calls on one context
 * race on its workspace. Independent scenes take a context each. */
int pmpm_context_create(pmpm_context **ctx);
int pmpm_context_destroy(pmpm_context *ctx);
int pmpm_context_reserve(pmpm_context *ctx, int32_t max_particles, int32_t max_cells);
/* Particles found outside [0, num_cells_total) by the last pmpm_spatial
 * (synchronises the stream). The reference writes out of bounds there
 * (spatialpartition_inline.h:102-113); here they sort last and join no cell. */
int pmpm_context_out_of_grid(pmpm_context *ctx, pmpm_stream_t stream, int32_t *count);

/* ---- spatial binning: replaces SpatialPartition::calculate_hash,
 * sort_hashes and bin_particles (spatialpartition.cpp:90-136,
 * spatialpartition_inline.h:49-115). sorted[] equals
 * thrust::stable_sort_by_key over slots 0..n-1. On entry, part->hash_sorted
 * is read as a hint: when it holds the previous call's sorted keys and the
 * particles come in that call's cell order (as pmpm_step leaves them), only
 * the particles that changed cell are sorted. Any content gives the same
 * result; only the time differs. Below 131,072 particles (or the value of
 * the environment variable PMPM_PARTITION_HINT_MIN when the context was
 * created) and for a context created with PMPM_PARTITION_RADIX=1 it runs
 * the full LSD radix sort. */
int pmpm_partition(pmpm_context *ctx, int32_t dim, const pmpm_grid *grid, const float *x,
                   int32_t n, pmpm_spatial *part, pmpm_stream_t stream);

/* ---- shape functions: replaces calculate_shape_function
 * (shapefunction.cpp:70-190). Materialises psi [n][nsn] and dpsi [n][nsn][dim]
 * in the reference's layout and window order, per slot. The solver kernels
 * never call this: they recompute weights in registers. */
int pmpm_shape_function(int32_t dim, int32_t shape_fn, int32_t node_types_built,
                        const pmpm_grid *grid, const float *x, int32_t n, float *psi,
                        float *dpsi, pmpm_stream_t stream);

/* ---- P2G: replaces USL::P2G / usl_p2g_kernel (usl.cpp:118-161,
 * usl_inline.h:82-180). Clears the node arrays, scatters mass, momentum and
 * force, then writes mom_nt = mom + force*dt (nodes with mass <= cutoff are
 * left at 0). Needs a partition of the same positions. The context holds the
 * per-tile partial sums (allocated on the first call for a larger grid). */
int pmpm_p2g(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid,
             const pmpm_particles *p, const pmpm_spatial *part, pmpm_nodes *nodes,
             pmpm_stream_t stream);

/* ---- node boundary conditions: replaces the apply_on_nodes_moments hooks of
 * Gravity, NodeDomain (gravity.cpp:65-88, nodedomain.cpp:42-64) and the
 * analytic box body, applied in list order to prm->bcs[bc_begin, bc_end).
 * remap_pass = 1 applies only kinematic constraints (MUSL second map).
 * Always recomputes the node velocity records (vel) afterwards. */
int pmpm_grid_bc(const pmpm_params *prm, const pmpm_grid *grid, pmpm_nodes *nodes,
                 int32_t bc_begin, int32_t bc_end, int32_t remap_pass, pmpm_stream_t stream);

/* ---- rigid-particle level set contact: replaces
 * RigidBodyLevelSet::apply_on_nodes_moments (rigidbodylevelset.cpp:92-275,
 * rigidbodylevelset_inline.h:65-343): rigid particle velocities from the
 * body rotation, nearest-rigid-particle stick contact on the nodes, then the
 * rigid particles advance. rotation (row-major D x D, host-side animation)
 * and com are HOST pointers. dpsi_orig is the materialised dpsi of
 * pmpm_shape_function in ORIGINAL particle order: the reference indexes it
 * with the linear window size 2^D whatever the spline
 * (rigidbodylevelset_inline.h:147-148), so the stored table is what it reads. */
int pmpm_rigid_levelset(const pmpm_params *prm, const pmpm_grid *grid, pmpm_particles *p,
                        const pmpm_spatial *part, pmpm_nodes *nodes, const float *rotation,
                        const float *com, const float *dpsi_orig, float *rigid_dx,
                        pmpm_stream_t stream);
/* rigid_dx (optional, [dim][n]): when given, the rigid particles are not moved
 * here; their displacement v*dt is stored (0 for the others) and applied after
 * G2P with pmpm_rigid_advance, so that G2P still weights with the positions
 * the step started from (the reference's stored psi). */
int pmpm_rigid_advance(int32_t dim, int32_t n, float *x, const float *rigid_dx,
                       const int32_t *sorted, pmpm_stream_t stream);

/* ---- utility: dst[c*n + id[i]] = src[c*n + i] for c < ncomp (slot order ->
 * original order), or the inverse when inverse != 0. No reference
 * counterpart: the reference never reorders particle storage. */
int pmpm_permute(int32_t n, int32_t ncomp, int32_t elem_bytes, const void *src,
                 const int32_t *id, void *dst, int32_t inverse, pmpm_stream_t stream);

/* ---- G2P fused with the return maps: replaces USL::G2P + Solver::stress_update
 * (usl_inline.h:245-332, solver.cpp:110-121, linearelastic_inline.h:40-107,
 * newtonfluid_inline.h:49-96, localrheo_inline.h:74-157,
 * dp_rheo_inline.h:123-290). Reads particle j of the cell order from slot
 * part->sorted[j] of `in` and writes it to slot j of `out` when out->F !=
 * in->F (the particles end up sorted by cell), or back to its own slot when
 * the F buffers are the same (in place). Weights use in->x. */
int pmpm_g2p(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *in,
             pmpm_particles *out, const pmpm_spatial *part, const pmpm_nodes *nodes,
             pmpm_stream_t stream);

/* ---- MUSL (no reference implementation; semantics of test_musl.h:37-156,
 * musl.cuh:39-72). First map + re-map, fused: velocities and positions from
 * the grid, then mom_nt := sum N m v with the new velocities (mass cutoff as
 * P2G) and the node velocities from it. Writes x, v of particle j to slot j
 * of `out` (out != in required). */
int pmpm_musl_map(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid,
                  const pmpm_particles *in, pmpm_particles *out, const pmpm_spatial *part,
                  pmpm_nodes *nodes, pmpm_stream_t stream);
/* Second map: velocity gradient, F, volume from the re-mapped moments, then
 * the return maps (weights from in->x, the positions the step started from);
 * slot placement as pmpm_g2p. */
int pmpm_musl_g2p(const pmpm_params *prm, const pmpm_grid *grid, const pmpm_particles *in,
                  pmpm_particles *out, const pmpm_spatial *part, const pmpm_nodes *nodes,
                  pmpm_stream_t stream);

/* ---- stand-alone return map: replaces Material::stress_update for one
 * material id (linearelastic.cpp:63-90 etc.), in place, from p->L. */
int pmpm_stress_update(const pmpm_params *prm, int32_t mat_id, pmpm_particles *p,
                       pmpm_stream_t stream);

/* ---- one full step: USL::solve (usl.cpp:69-115) or its MUSL counterpart
 * for configurations without a rigid-particle level set: partition, P2G,
 * node BCs, G2P + return maps. Reads `a`, writes `b` (cell-sorted). */
int pmpm_step(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid, int32_t musl,
              const pmpm_particles *a, pmpm_particles *b, pmpm_spatial *part,
              pmpm_nodes *nodes, pmpm_stream_t stream);

/* ---- one full step that allocates nothing. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It runs the
 * step pmpm_step runs, with the same arguments minus the context: the
 * workspace (sort keys and slots, digit histograms, P2G tile partials and
 * flags, wrap-around records and cell list, counters) comes from a file-scope
 * device array of PMPM_NOALLOC_SCRATCH_BYTES that each GPU holds until the
 * process ends. A step that needs more returns PMPM_ERR_WORKSPACE (see
 * pmpm_last_error); it calls no allocator. All calls on one GPU share that
 * GPU's array: calls that overlap on one GPU race. */
#define PMPM_NOALLOC_SCRATCH_BYTES (64u << 20)
int pmpm_step_noalloc(const pmpm_params *prm, const pmpm_grid *grid, int32_t musl, const pmpm_particles *a,
                      pmpm_particles *b, pmpm_spatial *part, pmpm_nodes *nodes, pmpm_stream_t stream);

/* ---- adjoint of one USL step (differentiable form; the reference has no
 * counterpart: it is the reverse mode of the step pmpm_step(musl = 0) runs,
 * usl.cpp:69-115, for the hand-written adjoint kernels of P2G, the grid
 * update, G2P and the return maps).
 * Adjoints of the particle state, SoA as pmpm_particles. */
typedef struct pmpm_adjoint {
  float *x;       /* [dim][n] */
  float *v;       /* [dim][n] */
  float *stress;  /* [6][n] */
  float *F;       /* [dim*dim][n] row-major */
  float *eps_e;   /* [6][n] */
} pmpm_adjoint;

/* Given the state `state` at the start of a USL step and the adjoint of the
 * step's output `out_bar` (in the output's order: slot j is the j-th particle
 * of the cell order of `state`, as pmpm_step writes it), writes the adjoint of
 * `state` into `in_bar` (state's order) and adds the gradient with respect to
 * each material's p[0..3] to param_bar[4 * material + k]. With recompute = 1
 * the step's partition, P2G and grid update are recomputed into `part` and
 * `nodes`; with recompute = 0 they are read as they are, and must be what
 * pmpm_step(musl = 0) left there when it stepped `state` with the same
 * parameters and grid (a caller that keeps one part/nodes set per step of a
 * recomputed segment saves that work). The active sets (return-map branch,
 * slip clamps, mass cutoff) are those of the forward step; no smoothing.
 * Supported: linear elastic and Drucker-Prager materials, gravity and
 * node-domain conditions; anything else returns PMPM_ERR_UNSUPPORTED. */
int pmpm_adjoint_step(pmpm_context *ctx, const pmpm_params *prm, const pmpm_grid *grid,
                      const pmpm_particles *state, const pmpm_adjoint *out_bar, pmpm_adjoint *in_bar,
                      float *param_bar, int32_t recompute, pmpm_spatial *part, pmpm_nodes *nodes,
                      pmpm_stream_t stream);

/* ---- batched scenes (no reference counterpart: the reference runs one
 * scene per process through its globals, global_settings.cpp; this is the
 * multi-scene form of the step pmpm_step runs, for parameter sweeps and
 * optimisation over many independent columns).
 *
 * A scene is everything one pmpm_step call takes, with its own context and
 * both particle buffer sets: a step reads set[cur] and writes set[cur ^ 1]. */
typedef struct pmpm_scene {
  pmpm_context *ctx;
  pmpm_params prm;
  pmpm_grid grid;
  pmpm_particles set[2];
  pmpm_spatial part;
  pmpm_nodes nodes;
  int32_t musl;
} pmpm_scene;

#define PMPM_MAX_LANES 8

typedef struct pmpm_batch pmpm_batch;

/* How a batch runs: `lanes` streams forked from the caller's stream (scenes
 * placed on them by pmpm_batch_assign with `weight` per scene), eager
 * launches or a replayed HIP graph. */
typedef struct pmpm_batch_tuning {
  int32_t lanes;     /* 1 .. PMPM_MAX_LANES */
  int32_t use_graph; /* 0: enqueue every step; 1: capture once per parity, replay */
  float step_ms;     /* measured time per batched step of the chosen setting */
} pmpm_batch_tuning;

/* A batch on the current device with `lanes` streams (default priority). */
int pmpm_batch_create(pmpm_batch **batch, int32_t lanes);
int pmpm_batch_destroy(pmpm_batch *batch);

/* Longest-processing-time split (host only): scenes in decreasing weight,
 * each onto the lane with the least total so far (ties: lowest lane). Writes
 * lane_of[count]; returns PMPM_OK. */
int pmpm_batch_assign(const float *weight, int32_t count, int32_t lanes, int32_t *lane_of);

/* Per-scene weights used by pmpm_batch_step (NULL: particle counts). */
int pmpm_batch_set_weights(pmpm_batch *batch, const float *weight, int32_t count);

/* One step of every scene (reads set[cur], writes set[cur ^ 1]); the lanes
 * are forked from and joined back into `stream` with events, so the call is
 * ordered like a single launch on `stream`. Each scene's workspace must
 * already be sized (one eager step does it) before a graph is captured. */
int pmpm_batch_step(pmpm_batch *batch, const pmpm_scene *scenes, int32_t count, int32_t cur,
                    pmpm_stream_t stream);

/* Captures the step of parity `cur` into the batch's graph for that parity
 * (replacing any earlier one). The scenes' buffers must stay where they are
 * for as long as the graph is replayed. */
int pmpm_batch_capture(pmpm_batch *batch, const pmpm_scene *scenes, int32_t count, int32_t cur,
                       pmpm_stream_t stream);
/* Replays the captured step of parity `cur` on `stream`. */
int pmpm_batch_replay(pmpm_batch *batch, int32_t cur, pmpm_stream_t stream);

/* Autotuner: times each scene's step alone (these become the LPT weights),
 * then `steps` batched steps for every lane count up to the batch's streams,
 * eager and graph-replayed, and keeps the fastest setting in the batch
 * (writes it to *out if non-NULL). No graph is kept: when the setting
 * replays graphs, call pmpm_batch_capture for both parities. It advances
 * every scene by the steps it runs (two alone, then `steps` per candidate,
 * rounded up to even) so the buffer parity `cur` is the same on return. */
int pmpm_batch_autotune(pmpm_batch *batch, const pmpm_scene *scenes, int32_t count, int32_t cur, int32_t steps,
                        pmpm_stream_t stream, pmpm_batch_tuning *out);
/* The batch's current setting. */
int pmpm_batch_get_tuning(const pmpm_batch *batch, pmpm_batch_tuning *out);

/* One scene's arguments to pmpm_adjoint_step (see there), for the batched
 * reverse pass of an ensemble of differentiated scenes. */
typedef struct pmpm_adjoint_scene {
  pmpm_context *ctx;
  pmpm_params prm;
  pmpm_grid grid;
  pmpm_particles state;
  pmpm_adjoint out_bar;
  pmpm_adjoint in_bar;
  float *param_bar;
  int32_t recompute;
  pmpm_spatial part;
  pmpm_nodes nodes;
} pmpm_adjoint_scene;

/* The adjoint step of every scene, each exactly as pmpm_adjoint_step runs
 * it, on the batch's lanes (its current lane count; scenes placed by its
 * weights, else by particle count), forked from and joined back into
 * `stream`. Scenes must not share a context, param_bar or buffers. Eager
 * launches only. Returns the first scene's error, if any. */
int pmpm_batch_adjoint_step(pmpm_batch *batch, const pmpm_adjoint_scene *scenes, int32_t count,
                            pmpm_stream_t stream);

/* ---- kernel timing (instrumentation; the reference's only timing hook is
 * the "Done Elapsed" stopwatch around Solver::run, solver.cpp:153-169).
 * With timing enabled, every kernel that the calling thread's entry points
 * launch is bracketed by a pair of HIP events recorded on its launch stream
 * (thread-local state; leave it off while capturing a HIP graph).
 * pmpm_timing_read waits for those events and returns, per kernel name in
 * first-launch order, the launch count and the summed event time in
 * milliseconds, then forgets them. `names` receives the names '\n'-separated
 * (truncated to names_len - 1 bytes); the first max_kernels counts and times
 * go to counts[] and ms[]; *num_kernels is the number of names. */
int pmpm_timing_enable(int32_t enable);
int pmpm_timing_read(char *names, int32_t names_len, int32_t *counts, float *ms, int32_t max_kernels,
                     int32_t *num_kernels);

/* The last HIP failure an entry point met on the calling thread, as
 * "call: hipGetErrorString (code)", or "" if none (thread-local storage). */
const char *pmpm_last_error(void);

/* Build identification string (static storage). */
const char *pmpm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PMPM_H_ */
