/*
 * pcisph.h - C ABI of the MI355X (gfx950) PCISPH simulation step.
 *
 * Drop-in for the reference's GPU step `sim::Fluid` (PCISPH/src/sim/Fluid.h:13-102,
 * PCISPH/src/sim/Fluid.cpp:32-466).  One opaque handle replaces the class; one
 * `pcisph_set_*` per reference setter with the same semantics; `pcisph_step`
 * replaces `Fluid::update()`.  Plain pointers and sizes only: no HIP or torch
 * types cross this boundary, no C++ exception escapes it (every entry point
 * returns a pcisph_status and `pcisph_last_error` gives the message).
 *
 * Particle buffers use the reference layout (PCISPH/src/scenes.cpp:192-208,
 * PCISPH/src/sim/Fluid.h:40-50): vectors are packed float[3N] (x0 y0 z0 x1 ...),
 * scalars float[N].  After every step the fluid arrays are in the step's sorted
 * (cell-hash) order, exactly as the reference permutes them in place each step
 * (PCISPH/data/kernels/sort_utils.cl:51-78).
 *
 * Threading: a handle is used by one host thread at a time; its work is ordered
 * on its own HIP stream.  Not re-entrant (neither is the reference,
 * PCISPH/src/sim/Fluid.cpp:108-327).
 */
#ifndef PCISPH_H
#define PCISPH_H

#include <stddef.h>
#include <stdint.h>

#include "pcisph_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pcisph_status {
    PCISPH_OK = 0,
    PCISPH_ERR_INVALID_ARGUMENT = 1, /* null handle/pointer, unknown field, bad option */
    PCISPH_ERR_SIZE = 2,             /* buffer size inconsistent with the particle counts
                                        (reference: std::runtime_error from
                                        checkBuffersConsistent, Fluid.cpp:90-106) */
    PCISPH_ERR_DEVICE = 3,           /* HIP runtime error (reference: cl::Error) */
    PCISPH_ERR_NO_DEVICE = 4,        /* no usable gfx950 device / kernels missing */
    PCISPH_ERR_OUT_OF_MEMORY = 5,
    PCISPH_ERR_NUMERIC = 6,          /* a non-finite value appeared in the step */
    PCISPH_ERR_INTERNAL = 7          /* a device-side bound was hit (e.g. sort spin limit) */
} pcisph_status;

/* Every particle array the step reads or writes.  Counts are in elements of the
 * element type (float or uint32), i.e. 3*N for vector fields. */
typedef enum pcisph_field {
    PCISPH_FLUID_POSITIONS = 0,          /* float[3N]  Fluid.h:43 */
    PCISPH_FLUID_VELOCITIES = 1,         /* float[3N]  Fluid.h:48 */
    PCISPH_FLUID_NORMALS = 2,            /* float[3N]  Fluid.h:44 */
    PCISPH_FLUID_PREDICTED_POSITIONS = 3,/* float[3N]  Fluid.h:45 */
    PCISPH_FLUID_DENSITIES = 4,          /* float[N]   Fluid.h:46 */
    PCISPH_FLUID_OTHER_FORCES = 5,       /* float[3N]  Fluid.h:47 */
    PCISPH_FLUID_PRESSURES = 6,          /* float[N]   Fluid.h:49 */
    PCISPH_FLUID_PRESSURE_FORCES = 7,    /* float[3N]  Fluid.h:50 */
    PCISPH_FLUID_DENSITY_VARIATIONS = 8, /* float[N]   Fluid.h:99 (internal there) */
    PCISPH_BOUNDARY_POSITIONS = 9,       /* float[3Nb] Fluid.h:40 */
    PCISPH_BOUNDARY_PRESSURES = 10,      /* float[Nb]  Fluid.h:41 */
    /* sort products, exported for parity checks (internal buffers in the reference,
       Fluid.h:88-98) */
    PCISPH_FLUID_KEYS = 11,              /* uint32[N]  sorted cell-hash keys */
    PCISPH_FLUID_SRC_LOCATIONS = 12,     /* uint32[N]  permutation: sorted slot -> previous slot */
    PCISPH_FLUID_CELL_OFFSETS = 13,      /* uint32[2B] (start, end) per bucket, (0,0) if empty */
    PCISPH_BOUNDARY_KEYS = 14,           /* uint32[Nb] */
    PCISPH_BOUNDARY_SRC_LOCATIONS = 15,  /* uint32[Nb] */
    PCISPH_BOUNDARY_CELL_OFFSETS = 16,   /* uint32[2B] */
    PCISPH_BOUNDARY_INIT_PRED_DENSITIES = 17, /* float[Nb]  (PCISPH/data/kernels/pcisph.cl:195-210) */
    /* per-particle neighbour statistics of the last step's density pass:
       (fluid candidates, boundary candidates, fluid neighbours with r^2 <= h^2,
        boundary neighbours with r^2 <= h^2).  Candidates = particles in the 27
       looked-up buckets (PCISPH/data/kernels/grid_utils.cl:65-79). */
    PCISPH_FLUID_NEIGHBOR_COUNTS = 18,   /* uint32[4N] */
    PCISPH_FIELD_COUNT = 19
} pcisph_field;

/* Options the reference hard-codes (PCISPH/src/sim/Fluid.cpp:250-251,118). */
typedef struct pcisph_solver_options {
    uint32_t min_iterations;        /* reference: 3 (termination tested from i >= 2) */
    uint32_t max_iterations;        /* reference: 7 */
    uint32_t launch_boundary_init;  /* 1: run the boundary-boundary predicted density pass
                                       (pcisph.cl:195-210) after each boundary sort; the
                                       reference sets its arguments but never launches it
                                       (Fluid.cpp:166-171), leaving the buffer as allocated:
                                       0 (default) keeps it zero. */
    uint32_t resort_boundary;       /* 1: re-sort boundary particles every step (moving walls).
                                       The reference sorts them only after a parameter change
                                       (Fluid.cpp:135,173). */
    uint32_t block_size;            /* threads per block of the per-particle kernels (a
                                       multiple of 64, at most 512); 0 = by size: 256 from
                                       65,536 fluid particles in the batch, 128 from 32,768,
                                       else 64 */
    uint32_t sort_digit_bits;       /* radix digit width, 4..8; 0 = autotuned/default */
    float neighbor_skin;            /* neighbour lists: each step the density pass records, in
                                       the reference's summation order, every candidate within
                                       (1 + skin) h; later passes walk those lists instead of
                                       the 27 buckets.  Results are identical (the skipped
                                       candidates contribute exact zeros).  In the pressure
                                       pass, where distances are taken at predicted positions,
                                       a wavefront walks lists only if for each of its
                                       particles its displacement plus the largest in its 27
                                       buckets is at most 0.9 skin h, else its buckets.  A
                                       scene whose pressure passes walked their buckets in a
                                       step (lists in under 1% of their fluid wavefronts) builds
                                       lists of radius 1.01 h (shorter lists for the passes at
                                       x) until a probe step every 32 steps finds the full
                                       radius used again.  Batches of 2^24 or more fluid or
                                       boundary particles step without lists (entries are
                                       stored in 24 bits).  0 = off; default 0.2 */
    uint32_t graph_iterations;      /* > 0: the step's launches are replayed from HIP graphs --
                                       one for everything before the correction loop and one
                                       per `graph_iterations` loop iterations (iterations past
                                       the end of every scene's loop return at once on the
                                       device).  Graphs are captured on first use and again
                                       whenever a launch argument changes; results are
                                       identical to 0 (launch by launch).  Default 4.  Steps
                                       timed by the stage profiler launch kernel by kernel. */
    uint32_t work_order;            /* bit 0: fluid particles of batches of at least 131,072
                                       are stored in a Morton layout
                                       rebuilt every step -- each bucket whole and in the
                                       reference's order, the buckets along a Morton curve of
                                       their cells -- so a wavefront covers a compact block of
                                       cells and its gathers lie together in memory; sums,
                                       downloads and sort products are unchanged (fields are
                                       presented in the reference's order); bit 1: workgroups are
                                       placed so that each of the 8 XCDs (own L2) works on
                                       compact chunks of that order; bit 2: FP32 pressure
                                       passes with 256-thread workgroups stage each
                                       workgroup's neighbour buckets in LDS and walk them
                                       there (measured slower on the dam breaks); bit 3: the
                                       Morton layout (bit 0) at any size; bits 4 and 5: the
                                       FP32 fluid pressure passes take their candidates from a
                                       per-step LDS staging (each workgroup's particles in the
                                       Morton order of their own cells, the bucket ranges of
                                       their rows copied once per pass and walked in the
                                       reference's order: identical results) -- automatically
                                       when no scene's pressure passes may walk neighbour
                                       lists in the step, always with bit 4, never with bit 5;
                                       bit 6: the correction loop runs as one cooperative
                                       launch whose workgroups meet at a grid barrier between
                                       the passes (identical results; not in steps the stage
                                       profiler times, warm-started steps or with bit 2);
                                       bit 7: the step's radix sorts run the one-sweep variant
                                       (one read of the keys for every digit histogram, then
                                       one scatter kernel per digit whose tile offsets come
                                       from a decoupled look-back; identical results; a look-
                                       back wait past 2 s fails the step with
                                       PCISPH_ERR_DEVICE).  Default 3 (bits 0 and 1). */
    uint32_t stage_cap;             /* LDS-staged pressure pass (work_order bits 4/5): fluid
                                       candidates a workgroup stages, 64..2048; 0 = 2048 (the
                                       kernel's LDS).  A workgroup with more, up to twice this,
                                       stages the rows that fit and walks its other rows'
                                       buckets in global memory; one with more still walks
                                       every bucket there.  Results are identical. */
} pcisph_solver_options;

/* Per-step report (the reference only has a compiled-out timer, Fluid.cpp:319-326). */
typedef struct pcisph_step_stats {
    uint32_t iterations;            /* correction iterations run in this step */
    uint32_t converged;             /* 1: stopped on max dv / rho0 < threshold */
    float max_density_variation;    /* max over fluid of max(0, rho* - rho0), last iteration */
    float max_density_error;        /* max_density_variation / rest_density (FP32, as tested) */
    uint32_t nonfinite;             /* 1 if a non-finite position/velocity was produced */
    uint32_t warm_start;            /* bit 0: the network's start was evaluated; bit 1: it was
                                       rejected for the plain start (larger residual); bit 2:
                                       the scene is left out of the next step's warm start
                                       (see pcisph_set_warm_start) */
    uint64_t fluid_candidates;      /* sum over fluid particles of fluid candidates */
    uint64_t boundary_candidates;   /* sum over fluid particles of boundary candidates */
    float max_density_expansion;    /* max of max(0, rho0 - rho*) over fluid particles whose
                                       step-start density was >= rho0 (free-surface particles
                                       sit below rho0 regardless), last iteration: the side of
                                       the residual the loop does not bound */
    uint32_t list_waves;            /* fluid wavefronts (64 particles) of this step's pressure
                                       passes that walked neighbour lists; the others walked
                                       their 27 buckets (sum over the step's iterations) */
    float first_density_error;      /* max density error of the plain (zero-pressure) start
                                       after the first correction iteration; in a warm-started
                                       step the dual first iteration evaluates it beside the
                                       network's start */
} pcisph_step_stats;

typedef struct pcisph_solver pcisph_solver;

/* Fluid::Fluid(ctx, device, queue) (Fluid.cpp:32-88).  `hip_stream` is a hipStream_t
 * or NULL (the solver then creates its own stream on `device_ordinal`). */
int pcisph_create(int device_ordinal, void* hip_stream, pcisph_solver** out);
int pcisph_destroy(pcisph_solver* s);
/* Message of the last failing call on this handle ("" if none).  With a NULL
 * handle: the last pcisph_create failure of this thread. */
const char* pcisph_last_error(const pcisph_solver* s);

/* Setters (Fluid.h:21-29, Fluid.cpp:333-375).  All but the threshold mark the
 * parameters dirty: the next step recomputes the derived parameters
 * (Fluid.cpp:377-443), reallocates internal buffers, re-sorts the boundary and
 * ZEROES the fluid velocities (Fluid.cpp:446-466), as the reference does. */
int pcisph_set_boundary_count(pcisph_solver* s, uint32_t boundary_count);
int pcisph_set_fluid_count(pcisph_solver* s, uint32_t fluid_count);
int pcisph_set_delta_t(pcisph_solver* s, float delta_t);
int pcisph_set_rest_density(pcisph_solver* s, float rest_density);
int pcisph_set_particle_radius(pcisph_solver* s, float particle_radius);
int pcisph_set_gravity(pcisph_solver* s, float gravity);
int pcisph_set_viscosity(pcisph_solver* s, float viscosity);
int pcisph_set_surface_tension(pcisph_solver* s, float surface_tension_coefficient);
int pcisph_set_density_variation_threshold(pcisph_solver* s, float threshold);

/* Fluid::get_params (Fluid.cpp:329-331): the byte-identical 80-byte struct.
 * Derived fields are those of the last step (as in the reference). */
int pcisph_get_params(const pcisph_solver* s, Simulation_Params* out);
int pcisph_set_solver_options(pcisph_solver* s, const pcisph_solver_options* opt);
int pcisph_get_solver_options(const pcisph_solver* s, pcisph_solver_options* out);

/* Fluid::checkBuffersConsistent (Fluid.cpp:90-106): PCISPH_ERR_SIZE if a buffer
 * does not match the counts. */
int pcisph_check_buffers_consistent(const pcisph_solver* s);

/* Arithmetic of the step (no counterpart in the reference, whose precision is the
 * OpenCL device's float):
 *   0 PCISPH_PRECISION_FAST  FP32 with fused multiply-adds and hardware reciprocal /
 *                            square root (default; the accuracy class OpenCL allows
 *                            for `/`, distance() and fast_length()); a step whose
 *                            scenes include a kernel radius above 1 m runs as IEEE;
 *   1 PCISPH_PRECISION_IEEE  FP32, one rounding per operation, the reference's order
 *                            of operations (matches the CPU restatement bit for bit);
 *   2 PCISPH_PRECISION_FP64  the same operations in double precision.
 * Switching to or from FP64 reallocates the particle arrays: set it before uploading. */
enum { PCISPH_PRECISION_FAST = 0, PCISPH_PRECISION_IEEE = 1, PCISPH_PRECISION_FP64 = 2 };
int pcisph_set_precision(pcisph_solver* s, int precision);

/* Element count the solver expects for `field` under the current counts. */
int pcisph_field_size(const pcisph_solver* s, int field, size_t* out_count);

/* Copy a field in or out in the reference layout.  `count` must equal
 * pcisph_field_size (PCISPH_ERR_SIZE otherwise).  Host pointers; downloads wait
 * for the step to finish.  Uploading PCISPH_BOUNDARY_POSITIONS schedules a
 * boundary re-sort at the next step (in a batch: of every scene's walls, in one
 * sort).  Sort products and neighbour counts are read-only. */
int pcisph_upload(pcisph_solver* s, int field, const void* host, size_t count);
int pcisph_download(pcisph_solver* s, int field, void* host, size_t count);
/* Same, with a device pointer on the solver's device (ordered on its stream). */
int pcisph_upload_device(pcisph_solver* s, int field, const void* device_ptr, size_t count);
int pcisph_download_device(pcisph_solver* s, int field, void* device_ptr, size_t count);

/* Float fields widened to double (exact for FP32 solvers; the FP64 solver's own
 * values otherwise), for error reports against an FP64 run. */
int pcisph_download_f64(pcisph_solver* s, int field, double* host, size_t count);

/* Fluid::update() (Fluid.cpp:108-327): one simulation step.  `stats` may be NULL.
 * Returns when the correction loop's iteration count is known; the final
 * integration may still be running (pcisph_sync waits). */
int pcisph_step(pcisph_solver* s, pcisph_step_stats* stats);
/* n_steps consecutive steps; per-step stats into stats[0..n_steps) if non-NULL. */
int pcisph_run(pcisph_solver* s, uint32_t n_steps, pcisph_step_stats* stats);
int pcisph_sync(pcisph_solver* s);

/* Apply pending setter changes now (derive parameters, reallocate, zero velocities,
 * schedule the boundary sort) instead of at the next step, so that data uploaded
 * afterwards -- e.g. the velocities of a resumed checkpoint -- survive the next
 * step.  No counterpart in the reference, where this happens inside update()
 * (Fluid.cpp:109-112). */
int pcisph_prepare(pcisph_solver* s);

/* Per-stage timing with HIP events on the solver's stream (the reference's
 * compiled-out step timer, Fluid.cpp:123,182,316,319-326, per stage).  Enabling
 * resets the totals; iterations enqueued ahead of the loop's end and then skipped
 * are not counted. */
typedef enum pcisph_stage {
    PCISPH_STAGE_BOUNDARY_SORT = 0, /* hash + sort + reorder of the boundary */
    PCISPH_STAGE_HASH = 1,          /* cell-table reset + hash keys (sort_utils.cl:7-23) */
    PCISPH_STAGE_SORT = 2,          /* radix sort, all passes */
    PCISPH_STAGE_REORDER = 3,       /* sort_utils.cl:51-78 */
    PCISPH_STAGE_DENSITY = 4,       /* pcisph.cl:65-92 (+ boundary pressure reset) */
    PCISPH_STAGE_NORMAL = 5,        /* pcisph.cl:94-112 */
    PCISPH_STAGE_FORCE_INIT = 6,    /* pcisph.cl:121-173 + first prediction */
    PCISPH_STAGE_PRESSURE = 7,      /* pcisph.cl:212-263, boundary and fluid */
    PCISPH_STAGE_PRESSURE_FORCE = 8,/* pcisph.cl:265-300 + prediction / final integration */
    PCISPH_STAGE_WARM_START = 9,    /* network inference + its pressure force */
    PCISPH_STAGE_COUNT = 10
} pcisph_stage;

typedef struct pcisph_stage_times {
    double ms[16];        /* summed device time per stage */
    uint64_t launches[16];/* timed launches per stage */
} pcisph_stage_times;

/* enable = 0: off; 1: time every step; n > 1: time every n-th step (the others replay
 * their HIP graphs untimed).  Timed steps launch kernel by kernel. */
int pcisph_profile_enable(pcisph_solver* s, int enable);
/* Waits for the stream, then returns the totals since profiling was enabled. */
int pcisph_profile_read(pcisph_solver* s, pcisph_stage_times* out);

/* ---- learned warm start (no counterpart in the reference) ------------------------
 * A network 8 -> 64 SiLU -> 64 SiLU -> 1 (BF16 weights and activations on the matrix
 * cores, FP32 accumulation) predicts each fluid particle's converged pressure from
 * local features (see PCISPH_MLP_INPUTS).  With the warm start on, the correction
 * loop starts from that pressure instead of zero; its first iteration evaluates both
 * starts and keeps the one with the smaller residual max |rho* - rho0|, so a bad
 * prediction falls back to the plain loop, and termination is the unchanged test.
 * Parameters: W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3 (FP32). */
#define PCISPH_MLP_INPUTS 8
#define PCISPH_MLP_PARAMS 4801
int pcisph_mlp_set_params(pcisph_solver* s, const float* params, size_t count);
int pcisph_mlp_get_params(pcisph_solver* s, float* params, size_t count);
/* y[i] = net(x[i]) for n host rows x[n][8]. */
int pcisph_mlp_forward(pcisph_solver* s, const float* x, float* y, size_t n);
/* Gradient of 0.5 * sum (net(x) - t)^2 / n w.r.t. the parameters, and that loss. */
int pcisph_mlp_gradients(pcisph_solver* s, const float* x, const float* t, size_t n, float* grad, float* loss);
/* pcisph_mlp_gradients without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under graph capture, tight serving or
 * training loops, harnesses that forbid hipMalloc).  Its device scratch (the staged rows
 * and the accumulators) is a static per-GPU array of ~2.4 MB, so n is at most 65,536
 * rows (PCISPH_ERR_SIZE beyond); the network's parameters must already be on the device
 * (pcisph_mlp_set_params).  Calls on one GPU share that array: overlapping calls race. */
int pcisph_mlp_gradients_noalloc(pcisph_solver* s, const float* x, const float* t, size_t n, float* grad,
                                 float* loss);
/* Device sample store: (features, converged pressure / p_ref) pairs. */
int pcisph_samples_reserve(pcisph_solver* s, size_t capacity);  /* clears the store */
/* Every following step appends the samples of every `stride`-th fluid particle
 * (rotating the phase each step); 0 stops collecting. */
int pcisph_samples_collect(pcisph_solver* s, uint32_t stride);
int pcisph_samples_count(const pcisph_solver* s, size_t* count);
int pcisph_samples_download(pcisph_solver* s, float* x, float* t, size_t count);
int pcisph_samples_upload(pcisph_solver* s, const float* x, const float* t, size_t count);
/* `steps` Adam steps on random batches of the store; mean loss of the last step. */
int pcisph_mlp_train(pcisph_solver* s, uint32_t steps, uint32_t batch, float learning_rate, uint64_t seed,
                     float* last_loss);
/* Steps start the correction loop from the network (max(0, p_prev + y p_ref) per
 * particle; the plain zero-pressure start is evaluated beside it in the first
 * iteration and kept where its residual is smaller).  A scene is left out of a
 * step's warm start, and runs exactly the plain step, after a step that ran to the
 * iteration cap without converging, or that converged at the minimum iteration count
 * with the plain start (or with the network's start while the plain start's first
 * iteration was already within the bound): no start shortens such a loop.  It is
 * taken back after a step that needed more than the minimum.  Enabling clears every
 * scene's left-out mark. */
int pcisph_set_warm_start(pcisph_solver* s, int enable);

/* ---- batches of scenes (no counterpart in the reference) -------------------------
 * A handle holds `count` independent scenes (default 1), each what one sim::Fluid
 * holds (Fluid.h:13-102): its own parameters, particle counts, threshold and
 * buffers.  The setters, pcisph_get_params, pcisph_set_density_variation_threshold,
 * pcisph_field_size, the uploads and downloads and pcisph_translate_boundary address
 * the selected scene (scene 0 until pcisph_select_scene); solver options, precision
 * and the warm-start network are shared.  One step advances every scene with one
 * launch per stage; each scene's correction loop ends on its own test, and the step
 * ends when the last scene's loop ends.  Sort products downloaded per scene are in
 * the scene's own numbering (keys < its bucket count, indices < its counts), exactly
 * as a single-scene handle gives them.  Changing the count marks every scene's
 * parameters changed (velocities zeroed at the next step, as after any setter);
 * positions and velocities of the scenes that remain are kept. */
int pcisph_set_scene_count(pcisph_solver* s, uint32_t count);
int pcisph_get_scene_count(const pcisph_solver* s, uint32_t* count);
int pcisph_select_scene(pcisph_solver* s, uint32_t index);
/* One step of every scene with per-scene statistics (count = scene count).
 * pcisph_step / pcisph_run report the batch as one: the largest iteration count and
 * error, converged only if every scene converged, candidates summed. */
int pcisph_step_scenes(pcisph_solver* s, pcisph_step_stats* stats, size_t count);

/* Rigid motion of the walls (moving tanks): adds (dx, dy, dz) to every boundary
 * position on the device and re-sorts the boundary at the next step.  No counterpart
 * in the reference, whose boundary never moves. */
int pcisph_translate_boundary(pcisph_solver* s, float dx, float dy, float dz);
/* Every scene's walls in one launch: scene k moves by shifts[3k .. 3k+2]
 * (count = 3 * scene count).  If any shift is non-zero the next step re-sorts the walls
 * of the batch, all scenes in one sort: an unmoved scene's walls keep their positions
 * and order, and its BOUNDARY_SRC_LOCATIONS become the identity (the permutation of
 * sorting walls that are already sorted), as after an upload of its walls. */
int pcisph_translate_boundaries(pcisph_solver* s, const float* shifts, size_t count);

/* Autotuner: times `trial_steps` steps of the current scene(s) for each threads-per-block
 * choice (64..512), radix digit width (5..8 bits) and tiling of the pressure pass
 * (work_order bits 4/5: LDS-staged workgroup tiles when no lists serve it, always, or
 * never) from the same saved starting state,
 * keeps the fastest in the solver options (and returns them in `chosen` if non-NULL),
 * and restores the state (particles, pending boundary sorts, warm-start eligibility,
 * sample collection; the options too if it fails).  It takes the snapshot slot of
 * pcisph_snapshot: a snapshot taken before the call is replaced by the state at the
 * call.  No counterpart in the reference (clogs has its own offline tuner,
 * exec_clogs_tune.bat). */
int pcisph_autotune(pcisph_solver* s, uint32_t trial_steps, pcisph_solver_options* chosen);

/* Device-side checkpoint of every scene's state at a step boundary: fluid positions and
 * velocities (with the previous step's pressures the warm start reads) and boundary
 * positions.  pcisph_restore puts it back (same counts), so runs can be repeated from
 * one state; pending setter changes are applied before the snapshot is taken.  No
 * counterpart in the reference. */
int pcisph_snapshot(pcisph_solver* s);
int pcisph_restore(pcisph_solver* s);

/* Derived parameters for a parameter set without a device (host only). */
int pcisph_compute_params(float particle_radius, float rest_density, uint32_t boundary_count,
                          uint32_t fluid_count, float viscosity, float surface_tension,
                          float gravity, float delta_t, Simulation_Params* out);

#ifdef __cplusplus
}
#endif

#endif /* PCISPH_H */
