/*
 * rainbow_fmm.h -- C ABI of the MI355X (gfx950) p=1 Cartesian uniform-octree
 * FMM step.
 *
 * Part 1 restates the reference planner's exported symbols one for one
 * (src/lib.rs:88-157), with the same names, argument meaning and clamping
 * behaviour: the functions are pure, never allocate, have no error codes and
 * return 0 for out-of-range levels.
 *
 * Part 2 restates the headless harness's plan construction
 * (native/main.rs:38-128) and FNV-1a digest (native/main.rs:152-163).
 *
 * Part 3 is the solver, which in the reference is host code driving WGSL
 * entry points (native/main.rs:355-421, app.js:609-648).  Here it is a
 * library: a handle owns device memory and one HIP stream, and every call
 * returns 0 on success or a negative status, with a thread-local message in
 * rainbow_last_error().  Nothing aborts and no exception crosses the ABI.
 *
 * Part 4 (batch) advances many independent systems in one launch sequence per
 * step; Part 5 (serve) is the continuous-batching request loop built on it.
 *
 * Buffer layouts seen through this ABI are the reference's
 * (src/rainbow_fmm.wgsl:18-73):
 *   Particle 32 B = {pos.xyz, hue} {vel.xyz, excitation}       (f32 x 8)
 *   Moment   16 B = {count u32, sum_x, sum_y, sum_z i32}       fixed point
 *   Field    48 B = {a.xyz, phi} {Jxx,Jxy,Jxz,Jyy} {Jyz,Jzz,0,0}
 * Cell g of level l is level_offset(l) + x + 2^l (y + 2^l z).
 */
#ifndef RAINBOW_FMM_H
#define RAINBOW_FMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Part 1. Planner ABI (src/lib.rs:88-157).                                  */
/* ------------------------------------------------------------------------ */

/* == 1.  Replaces rainbow_api_version, src/lib.rs:88-91. */
uint32_t rainbow_api_version(void);

/* Static payload of the solver kernels (the gfx950 HIP source compiled into
 * this library), no NUL terminator, never freed.  Replaces
 * rainbow_shader_ptr/len, src/lib.rs:93-101, which point at the WGSL text. */
const uint8_t* rainbow_shader_ptr(void);
size_t rainbow_shader_len(void);

/* Static algorithm name.  Replaces src/lib.rs:103-111. */
const uint8_t* rainbow_algorithm_ptr(void);
size_t rainbow_algorithm_len(void);

/* depth_for_particles(clamp_particle_count(n)).  Replaces src/lib.rs:113-116. */
uint32_t rainbow_depth_for_particles(uint32_t particles);
/* 8^level, or 0 when level > 6.  Replaces src/lib.rs:118-125. */
uint32_t rainbow_level_cells(uint32_t level);
/* (8^level - 1) / 7, or 0 when level > 7.  Replaces src/lib.rs:127-134. */
uint32_t rainbow_level_offset(uint32_t level);
/* total_cells(clamp(depth, 4, 6)).  Replaces src/lib.rs:136-139. */
uint32_t rainbow_total_cells(uint32_t depth);
/* clamp(n, 2^14, 2^23).  Replaces src/lib.rs:141-144. */
uint32_t rainbow_clamp_particle_count(uint32_t requested);
/* fixed_point_scale(clamp_particle_count(n)).  Replaces src/lib.rs:146-149. */
float rainbow_fixed_point_scale(uint32_t particles);
/* Low 32 bits of 3 n + 199 total_cells(d), n and d clamped.
 * Replaces src/lib.rs:151-157. */
uint32_t rainbow_linear_work_lo(uint32_t particles, uint32_t depth);

/* The unclamped planner functions the native harness calls directly
 * (native/main.rs:90-92): depth_for_particles (src/lib.rs:45-52),
 * fixed_point_scale (src/lib.rs:59-69) and linear_work_bound
 * (src/lib.rs:83-86).  Solves use these, as the harness does. */
uint32_t rainbow_plan_depth(uint32_t particles);
float rainbow_plan_fixed_point_scale(uint32_t particles);
uint64_t rainbow_plan_linear_work(uint32_t particles, uint32_t depth);

/* The browser's per-frame time step (app.js:42-43, 716-719): BASE_DT =
 * 0.0009 scaled by the frame interval over a 60 Hz frame, the interval
 * capped at 100 ms, computed in double and stored as the f32 Params.dt. */
float rainbow_frame_dt(double frame_interval_ms);

/* ------------------------------------------------------------------------ */
/* Part 2. Params image, plan table and digest (native/main.rs).            */
/* ------------------------------------------------------------------------ */

/* The 128-byte uniform block, field for field as struct Params in
 * src/rainbow_fmm.wgsl:18-50 (eight u32, then twenty-four f32). */
typedef struct rainbow_params {
  uint32_t particle_count;
  uint32_t depth;
  uint32_t level;
  uint32_t level_offset;
  uint32_t level_dim;
  uint32_t parent_offset;
  uint32_t parent_dim;
  uint32_t total_cells;
  float dt;
  float gravity;
  float softening;
  float point_size;
  float time;
  float aspect;
  float camera_yaw;
  float camera_pitch;
  float viewport_width;
  float viewport_height;
  float exposure;
  float speed;
  float control_point[4];     /* xyz brush centre, w > 0.5 = active      */
  float control_axis_mode[4]; /* xyz vortex axis, w: 1 pull 2 push 3 spin */
  float control_settings[4];  /* x strength, y radius, z fixed-point scale */
} rainbow_params_t;

#define RAINBOW_PARAM_BYTES 128u
#define RAINBOW_PARAM_STRIDE 256u
#define RAINBOW_PARAM_SLOTS 16u
#define RAINBOW_PARTICLE_BYTES 32u
#define RAINBOW_MOMENT_BYTES 16u
#define RAINBOW_FIELD_BYTES 48u

/* Fills one Params block with the harness constants (dt 0.0009, gravity 0.34,
 * softening 0.012, speed 1, control off, scale in control_settings.z).
 * Replaces write_params, native/main.rs:38-87. */
void rainbow_write_params(uint32_t count, uint32_t depth, uint32_t level,
                          uint32_t level_offset, uint32_t level_dim,
                          uint32_t parent_offset, uint32_t parent_dim,
                          uint32_t cells, float scale, rainbow_params_t* out);

/* Fills the 16 x 256-byte slot table: slot l (l = 0..depth) is the level plan
 * with parent = l-1 (saturating), slot 8+l (l = 0..depth-1) is the upward
 * plan whose "parent" is the child level l+1.  Unused bytes are zero.
 * Uses the unclamped planner.  Returns the depth.
 * Replaces build_plans, native/main.rs:89-128. */
uint32_t rainbow_build_plans(uint32_t count, uint8_t out_slots[4096]);

/* FNV-1a-32 over bytes (basis 2166136261, prime 16777619).
 * Replaces digest, native/main.rs:152-163. */
uint32_t rainbow_fnv1a(const void* data, size_t bytes);

/* Seeded initial conditions (host).  The reference seeds its galaxy on the
 * GPU (initialize, src/rainbow_fmm.wgsl:232-374); these produce the
 * synthetic inputs of the benchmark workloads.  SplitMix64 stream, f64
 * generation rounded to f32, equal masses 1/N, hue uniform in [0,1),
 * excitation 0.  `out` receives 32 N bytes in the reference layout.
 * Return 0 or RAINBOW_E_INVALID (with rainbow_last_error() naming the
 * violated bound: n >= 1, a positive radius or scale, 1 <= clumps <= n). */
/* Uniform in a ball of `radius`, at rest. */
int rainbow_ic_uniform_ball(uint32_t n, uint64_t seed, double radius, void* out);
/* Plummer sphere of scale radius a and total mass 1 truncated at rmax, with
 * isotropic virial velocities for gravity G (Aarseth-Henon-Wielen). */
int rainbow_ic_plummer(uint32_t n, uint64_t seed, double a, double rmax, double gravity,
                       void* out);
/* `clumps` Plummer clumps of scale radius a truncated at rcut_over_a * a,
 * clump masses 10^U(0,2) normalised, bodies per clump in proportion,
 * centres uniform in [-0.5, 0.5]^3, virial velocities within each clump. */
int rainbow_ic_clumps(uint32_t n, uint64_t seed, uint32_t clumps, double a, double rcut_over_a,
                      double gravity, void* out);

/* ------------------------------------------------------------------------ */
/* Part 3. Solver (one system).                                              */
/* ------------------------------------------------------------------------ */

typedef struct rainbow_hip rainbow_hip_t;

enum {
  RAINBOW_OK = 0,
  RAINBOW_E_INVALID = -1, /* bad argument, size or handle          */
  RAINBOW_E_HIP = -2,     /* a HIP runtime call failed             */
  RAINBOW_E_NOMEM = -3,   /* device or host allocation failed      */
  RAINBOW_E_BUDGET = -4,  /* request exceeds the admission budget  */
  RAINBOW_E_STATE = -5,   /* call not valid in the current state   */
  RAINBOW_E_CANCELLED = -6, /* a served request was cancelled       */
  RAINBOW_E_TIMEOUT = -7    /* a served request ran out of time     */
};

enum { RAINBOW_PARTICLES = 0, RAINBOW_MOMENTS = 1, RAINBOW_FIELDS = 2 };

/* Solver options (bit flags). */
enum {
  /* Neumaier-compensated FP32 sums in M2L, near field and the own-leaf term
   * (the precision form).  Off by default: the plain step follows the
   * reference's straight FP32 sums bit for bit. */
  RAINBOW_OPT_COMPENSATED = 1u
};

/* Creates a solver for `particles` bodies.  Depth, cell count and scale come
 * from the unclamped planner (as native/main.rs:90-92); dt, gravity,
 * softening, speed and the control block come from `params` (NULL = the
 * harness constants of native/main.rs:57-82).  A nonzero
 * params->control_settings[2] overrides the planner's fixed-point scale.
 * Particles start zeroed; fields start zeroed. */
int rainbow_hip_create(uint32_t particles, const rainbow_params_t* params,
                       rainbow_hip_t** out);
/* Same, with option flags. */
int rainbow_hip_create_ex(uint32_t particles, const rainbow_params_t* params,
                          uint32_t options, rainbow_hip_t** out);
void rainbow_hip_destroy(rainbow_hip_t* h);

/* Physics parameters for later steps (same meaning as in create). */
int rainbow_hip_set_params(rainbow_hip_t* h, const rainbow_params_t* params);

/* Plan chosen at create: depth, total cells, fixed-point scale. */
int rainbow_hip_plan(rainbow_hip_t* h, uint32_t* depth, uint32_t* cells,
                     float* scale);

/* Copies 32 * particles bytes of reference-layout particles in (host memory,
 * caller-owned, copied before return). */
int rainbow_hip_upload_particles(rainbow_hip_t* h, const void* particles,
                                 size_t bytes);

/* Seeds the bodies on the device with the reference's default scene, in
 * place of an upload: the initialize dispatch of native/main.rs:423-434
 * (src/rainbow_fmm.wgsl:205-374: 5% bulge, 70% halo, 18% four-arm disc, 7%
 * polar lanes, speeds from the solver's gravity).  Body i is a function of
 * i alone.  The device's sin/cos/pow/atan2 are used, so the scene equals
 * the reference's within a tolerance (the WGSL leaves their accuracy to
 * the implementation), not bit for bit. */
int rainbow_hip_seed_galaxy(rainbow_hip_t* h);

/* One field solve without integrate: clear, deposit, upward, M2L, downward,
 * near (the digest point of native/main.rs:436-454).  Asynchronous. */
int rainbow_hip_solve_fields(rainbow_hip_t* h);

/* `steps` full steps (native/main.rs:412-421), replayed from a captured HIP
 * graph.  Asynchronous. */
int rainbow_hip_step(rainbow_hip_t* h, uint32_t steps);

/* One step of a browser frame (app.js frame(), :711-722): dt =
 * rainbow_frame_dt(frame_interval_ms) for this and later steps, the other
 * physics unchanged, then one step. */
int rainbow_hip_step_frame(rainbow_hip_t* h, double frame_interval_ms);

/* Waits for all work on the handle's stream. */
int rainbow_hip_synchronize(rainbow_hip_t* h);

/* Copies a buffer out in the reference layout: particles 32 N bytes, moments
 * 16 C bytes, fields 48 C bytes.  `bytes` must equal that size.  Synchronous. */
int rainbow_hip_download(rainbow_hip_t* h, int which, void* dst, size_t bytes);

/* FNV-1a-32 of a downloaded buffer (as native/main.rs:450-454). */
int rainbow_hip_digest(rainbow_hip_t* h, int which, uint32_t* out);

/* Runs `steps` steps, timed with HIP events on the handle's stream; writes
 * the elapsed milliseconds.  The timed region holds only the step graph
 * replays; inputs are already resident. */
int rainbow_hip_time_steps(rainbow_hip_t* h, uint32_t steps, float* ms);

/* Runs `steps` steps without the graph, timing every kernel launch with an
 * event pair recorded by the dispatch itself, and writes the mean duration
 * in ms of each phase: out[0] deposit, out[1] upward (subtrees), out[2]
 * upward (top levels, with the level-1 fields), out[3 + i] tree launch i
 * (as listed by rainbow_hip_tree_launches),
 * then integrate, then the body sort (the steps sort when due, as
 * rainbow_hip_step does; the entry is the sort time divided by `steps`).
 * `n` is the capacity of out (at least 5 + the number of tree launches; 16
 * always suffices).  Returns the number of entries written (>= 0) or a
 * negative status. */
int rainbow_hip_phase_ms(rainbow_hip_t* h, uint32_t steps, float* out,
                         uint32_t n);

/* Bodies are stored in Morton order of their finest cell: every
 * `interval`-th step (0 = never; default 8) counting-sorts them by the cell
 * its deposit found, before its tree pass.  The order is internal:
 * downloads return caller order and no result depends on it. */
int rainbow_hip_set_sort_interval(rainbow_hip_t* h, uint32_t interval);

/* The tree pass runs as a few launches, in one of two forms.
 * Runs of levels: each launch covers levels lo..hi, and a block of it
 * recomputes the fields of its item's ancestors down to level lo instead of
 * waiting for another launch.  rainbow_hip_set_tree_split: bit l of
 * `starts` (2 <= l <= 6) starts a launch at level l (bit 2 is implied).
 * rainbow_hip_set_tree_fusion(h, hi): levels 2..hi together, each deeper
 * level alone.
 * Two launches (`starts` = 1, bit 0): the M2L sums of every level above the
 * finest (they depend on no parent's field), then every level's fields:
 * each parent's field from the level-1 field and its ancestors' sums (the
 * L2L additions in order), the finest level's M2L, the L2L terms and the
 * near fields.
 * Until one of the calls above (or autotune) is made, the split is
 * measured: the first step after a configuration records how many internal
 * parents each level has, and a level starts its own launch when they fill
 * two thirds of the device's resident blocks; when that leaves every level
 * but the finest in one run, the two-launch form is used instead (a
 * 1M-body Plummer system: two launches; 256 systems of up to 262144 bodies:
 * one launch per level).  Results do not depend on the split. */
int rainbow_hip_set_tree_split(rainbow_hip_t* h, uint32_t starts);
int rainbow_hip_set_tree_fusion(rainbow_hip_t* h, uint32_t hi);

/* The tree launches of a step, as (lo | form << 8, hi) pairs in launch
 * order; form 0: a run of levels lo..hi, 1: the M2L sums of levels lo..hi
 * (two-launch form), 2: the fields of levels lo..hi from them, 3: form 1
 * for a single system, also running the top levels (so the step has no
 * launch of its own for them) and the levels 2..depth-3 whose parents
 * those list (lo..hi are the levels listed by upward).  `n` is the
 * capacity of lo_hi (2 entries per launch).  Returns the number of
 * launches or a negative status. */
int rainbow_hip_tree_launches(rainbow_hip_t* h, uint32_t* lo_hi, uint32_t n);

/* Block shapes of the step's kernels (no counterpart in the reference,
 * whose workgroup sizes are fixed in the WGSL: src/rainbow_fmm.wgsl:394,
 * 408, 424, 457, 521, 585, 613).  integrate_threads: the width of an
 * integrate+deposit block over its 1024-body deposit range, 1024 (one body
 * per thread, the default) or 512 (two per thread).  tree_chunks: the tree
 * items' pair tile, chunks of 8 targets x 8 sources evaluated per round,
 * 8 (the default; two per wave, 6 blocks per CU), 4 (one per wave, half
 * the LDS of terms, 8 blocks per CU), 1 (wave items: each wave walks
 * items of its own, one chunk at a time, without block barriers; faster
 * when many items share a launch, as in large batches; fused runs of
 * levels keep 8) or 0 (per launch: wave items for a launch that held at
 * least 8192 items in the last measured step, 8 otherwise; the server's
 * shape).  Results do not depend on the shape; RAINBOW_E_INVALID for other
 * values. */
int rainbow_hip_set_shape(rainbow_hip_t* h, uint32_t integrate_threads, uint32_t tree_chunks);
int rainbow_hip_get_shape(rainbow_hip_t* h, uint32_t* integrate_threads, uint32_t* tree_chunks);

/* Autotuner: from the same starting state, times `steps` steps at each
 * block shape (the four combinations above, each at the two-launch form
 * and at one launch per level) and keeps the fastest, then at each tree
 * split (every partition of levels 2..depth into runs, and the two-launch
 * form), then at each candidate sort interval (1, 2, 4, 8, 16, 32, never);
 * keeps the fastest of each and writes the interval to *interval (the
 * shape: rainbow_hip_get_shape).  The split is fixed afterwards.  The
 * bodies are restored afterwards. */
int rainbow_hip_autotune(rainbow_hip_t* h, uint32_t steps, uint32_t* interval);

/* Number of cells active in the last solve at each level (M2L targets:
 * non-empty cells whose parent holds more than 16 bodies), plus the
 * number of near-field leaves.  out[l] for l = 0..6, out[7] = leaves. */
int rainbow_hip_tree_stats(rainbow_hip_t* h, uint32_t out[8]);

/* Non-empty source pairs of the last solve per level: out[l] M2L pairs,
 * out[8 + l] near-field pairs (l = 1..6), computed on the host from the
 * moments.  Each pair costs 40 f32 operations of pair_field plus 10
 * additions into the target's field. */
int rainbow_hip_pair_counts(rainbow_hip_t* h, uint64_t out[16]);

/* FP64 diagnostics of the current particle state (precision form):
 * out[0] kinetic energy, out[1] potential energy (the tree's own potential
 * at each body, halved), out[2..4] linear momentum, out[5..7] angular
 * momentum.  Uses the fields of the last solve; call after
 * rainbow_hip_solve_fields for a state that matches the particles. */
int rainbow_hip_diagnostics(rainbow_hip_t* h, double out[8]);

/* FP64 direct sum over all bodies (mass 1/N each, softening from the
 * physics) at `ntargets` bodies given by caller index: out[4 t + 0..3] =
 * {ax, ay, az, phi} = G/N sum_{j != target} {r, -|r|} / (|r|^2 + eps^2)^{3/2
 * or 1/2}, r = x_j - x_target.  The exact reference for force-error
 * measurements; O(N ntargets) work on the device. */
int rainbow_hip_direct_sum_f64(rainbow_hip_t* h, const uint32_t* targets, uint32_t ntargets, double* out);

/* The same FP64 direct sum without allocating: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It runs the kernels of rainbow_hip_direct_sum_f64 in a fixed 17 MiB
 * device array that each GPU the library runs on holds until the process
 * ends, so its output equals that function's bit for bit.  At most 4096
 * targets (more: RAINBOW_E_INVALID).  Calls that overlap on one GPU share
 * the array and race; the call returns after its own stream has finished. */
int rainbow_hip_direct_sum_f64_noalloc(rainbow_hip_t* h, const uint32_t* targets, uint32_t ntargets,
                                       double* out);

/* Per-body acceleration the integrator applies (local expansion plus the
 * own-leaf monopole), as float4 {ax, ay, az, phi} per body, for the current
 * particles and the fields of the last solve.  `bytes` = 16 N. */
int rainbow_hip_accelerations(rainbow_hip_t* h, void* dst, size_t bytes);

/* Error message of the last failing call on this thread ("" if none). */
const char* rainbow_last_error(void);

/* ------------------------------------------------------------------------ */
/* Part 4. Batch: many independent systems advanced together.               */
/* ------------------------------------------------------------------------ */

typedef struct rainbow_batch rainbow_batch_t;

/* Creates `systems` independent systems with the given body counts (each
 * >= 1, planner depth per system), sharing the physics of `params` (NULL =
 * harness constants).  Each system's scale is its own planner scale. */
int rainbow_batch_create(uint32_t systems, const uint32_t* counts,
                         const rainbow_params_t* params, uint32_t options,
                         rainbow_batch_t** out);
void rainbow_batch_destroy(rainbow_batch_t* b);
int rainbow_batch_upload(rainbow_batch_t* b, uint32_t system,
                         const void* particles, size_t bytes);
/* As rainbow_hip_seed_galaxy, for one system of the batch. */
int rainbow_batch_seed_galaxy(rainbow_batch_t* b, uint32_t system);
int rainbow_batch_download(rainbow_batch_t* b, uint32_t system, int which,
                           void* dst, size_t bytes);
/* Solve without integrate, for every system. */
int rainbow_batch_solve_fields(rainbow_batch_t* b);
/* `steps` steps of every system, one captured graph replay per step. */
int rainbow_batch_step(rainbow_batch_t* b, uint32_t steps);
int rainbow_batch_time_steps(rainbow_batch_t* b, uint32_t steps, float* ms);
int rainbow_batch_synchronize(rainbow_batch_t* b);
/* Total bodies over all systems. */
uint64_t rainbow_batch_particles(rainbow_batch_t* b);
/* As rainbow_hip_set_sort_interval / rainbow_hip_set_tree_fusion /
 * rainbow_hip_set_tree_split / rainbow_hip_set_shape / rainbow_hip_get_shape
 * / rainbow_hip_autotune, for the batch. */
int rainbow_batch_set_shape(rainbow_batch_t* b, uint32_t integrate_threads, uint32_t tree_chunks);
int rainbow_batch_get_shape(rainbow_batch_t* b, uint32_t* integrate_threads, uint32_t* tree_chunks);
int rainbow_batch_set_sort_interval(rainbow_batch_t* b, uint32_t interval);
int rainbow_batch_set_tree_fusion(rainbow_batch_t* b, uint32_t hi);
int rainbow_batch_set_tree_split(rainbow_batch_t* b, uint32_t starts);
int rainbow_batch_autotune(rainbow_batch_t* b, uint32_t steps, uint32_t* interval);

/* ------------------------------------------------------------------------ */
/* Part 5. Serving: continuous batching of N-body solve requests.           */
/* ------------------------------------------------------------------------ */

typedef struct rainbow_server rainbow_server_t;

/* A request: body count, steps to run, and its particles (32 N bytes,
 * reference layout, caller-owned, copied at submit).  timeout_s > 0: the
 * request is dropped (status RAINBOW_E_TIMEOUT) if it has not completed
 * that many seconds after submit; 0 = no limit. */
typedef struct rainbow_request {
  uint64_t id;
  uint32_t particles;
  uint32_t steps;
  const void* data;
  double timeout_s;
  uint32_t flags;  /* RAINBOW_REQUEST_BORROW */
  uint32_t reserved;
} rainbow_request_t;

/* Request flag: the server reads `data` in place instead of copying it at
 * submit; the caller keeps it valid and unchanged until the request's
 * completion record has been popped.  With pinned memory from
 * rainbow_host_alloc the upload is a DMA straight from the caller's buffer. */
enum { RAINBOW_REQUEST_BORROW = 1u };

/* Page-locked host memory for zero-copy requests (NULL on failure), and its
 * release. */
void* rainbow_host_alloc(size_t bytes);
void rainbow_host_free(void* p);

/* Completion record.  status RAINBOW_OK: the result is ready for
 * rainbow_server_result; RAINBOW_E_CANCELLED / RAINBOW_E_TIMEOUT: the
 * request stopped after steps_run of its steps and has no result. */
typedef struct rainbow_completion {
  uint64_t id;
  uint32_t particles;
  uint32_t steps;
  double submit_s;   /* host clock at submit                          */
  double admit_s;    /* host clock when it got a slot (0: never)       */
  double finish_s;   /* host clock when its completion was observed    */
  int32_t status;
  uint32_t steps_run;
} rainbow_completion_t;

/* A server with `budget_bytes` of device memory for resident systems and at
 * most `max_slots` systems in flight.  Resident systems live in an arena of
 * large device chunks grown as the working set grows (kept until destroy),
 * so steady-state admission allocates nothing; the batched step's work
 * tables are reserved for the whole budget at create.  Uploads and result
 * copies run on a second stream of the same priority, ordered by events. */
int rainbow_server_create(uint64_t budget_bytes, uint32_t max_slots,
                          const rainbow_params_t* params, uint32_t options,
                          rainbow_server_t** out);
void rainbow_server_destroy(rainbow_server_t* s);
/* Device bytes a request of `particles` bodies holds while resident. */
uint64_t rainbow_server_request_bytes(uint32_t particles);
/* Queues a request (copies its particles).  RAINBOW_E_BUDGET if it could
 * never fit the budget; RAINBOW_E_INVALID for an empty request, a negative
 * timeout or an id still in use (queued, running or with an untaken
 * result). */
int rainbow_server_submit(rainbow_server_t* s, const rainbow_request_t* r);
/* Cancels a queued or running request: it stops at the next step boundary
 * and completes with status RAINBOW_E_CANCELLED.  RAINBOW_E_STATE if the
 * request is not queued or running. */
int rainbow_server_cancel(rainbow_server_t* s, uint64_t id);
/* One scheduling pass, without waiting for the device unless the step
 * window is full: retire requests whose results have arrived, drop
 * cancelled and expired ones, admit queued requests while budget and slots
 * allow, and enqueue batched steps of every resident system (up to 3 steps
 * ahead of the device, up to 16 per pass); a request leaves the batch right
 * after its own last step.  Writes the number of steps enqueued. */
int rainbow_server_poll(rainbow_server_t* s, uint32_t* stepped);
/* Polls until every submitted request has completed. */
int rainbow_server_drain(rainbow_server_t* s);
/* Cumulative: out[0] host seconds admitting, out[1] reconfiguring the
 * batch, out[2] enqueuing steps (and waiting on a full window), out[3]
 * retiring; out[4] batched steps run, out[5] polls; out[6] device seconds
 * of the steps (events around each); out[7] host seconds blocked on the
 * device. */
int rainbow_server_stats(rainbow_server_t* s, double out[8]);
/* Pops up to `cap` completions; writes how many. */
int rainbow_server_completions(rainbow_server_t* s, rainbow_completion_t* out,
                               uint32_t cap, uint32_t* n);
/* Copies a completed request's final particles out (32 N bytes), once. */
int rainbow_server_result(rainbow_server_t* s, uint64_t id, void* dst,
                          size_t bytes);
/* Zero-copy form: the completed result's pinned buffer (32 N bytes, the
 * reference layout), valid until rainbow_server_release(s, id), which ends
 * the request (as rainbow_server_result does). */
int rainbow_server_result_view(rainbow_server_t* s, uint64_t id, const void** data, size_t* bytes);
int rainbow_server_release(rainbow_server_t* s, uint64_t id);
/* Requests queued, resident (running or finishing), device bytes resident. */
int rainbow_server_state(rainbow_server_t* s, uint32_t* queued,
                         uint32_t* resident, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* RAINBOW_FMM_H */
