/*
 * libplume_hip: MI355X (gfx950) kernels for the Plume simulation substep and volume ray march, behind a
 * plain C ABI.
 *
 * The reference (chadmv/plume) exposes this path as a Python operator API over Warp arrays; it has no
 * FFI. Each entry point below replaces one reference operator or kernel launch, cited as file:line
 * relative to the reference repository (src/plume/...). The Python package `plume_amd` binds these with
 * ctypes, passing device pointers of contiguous float32 torch tensors and torch's current HIP stream.
 *
 * Conventions
 *   - Arrays are C-order [i][j][k] with k contiguous. Cells are dims = (nx, ny, nz); faces are
 *     u (nx+1, ny, nz), v (nx, ny+1, nz), w (nx, ny, nz+1). vec3 arrays are AoS, 3 floats per cell.
 *   - `stream` is a hipStream_t (NULL = the null stream). Every call is asynchronous on it, except where
 *     noted; no call synchronises the device.
 *   - Every call returns 0 on success, PLM_EINVAL for a bad argument (nothing launched), or the
 *     hipError_t of a failed launch. plm_last_error() gives this thread's last message.
 *   - The library allocates no device memory behind the caller's back: every buffer, scratch included,
 *     comes in as an argument. The one exception is plm_smoke_budget_noalloc, whose scratch is a static
 *     device array of fixed size, loaded with the library.
 */
#ifndef PLUME_HIP_H
#define PLUME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLM_EINVAL (-1)

/* Per-level grid geometry, passed by value to kernels. Same meaning as GridInfo, sim/grid.py:21-33:
   open_lo/open_hi are 0/1 per axis (lo = left/bottom/back, hi = right/top/front, sim/grid.py:78-79);
   fill_hi is the share of the last cell on each axis inside the domain and ghost_gap the distance of the
   open-face pressure ghost, both (1,1,1)/0.5 on the simulation grid (sim/grid.py:80-81). */
typedef struct plm_grid_info {
    int32_t dims[3];
    float cell_m;
    float origin_m[3];
    int32_t open_lo[3];
    int32_t open_hi[3];
    float fill_hi[3];
    float ghost_gap;
} plm_grid_info;

/* One emitter, EmitterParams, sim/emitters.py:23-38 (shape 0 sphere, 1 box). */
typedef struct plm_emitter {
    int32_t shape;
    float center_m[3];
    float radius_m;
    float half_size_m[3];
    float edge_softness;
    float start_s;
    float end_s;
    float smoke_per_s;
    float temperature_k;
    float velocity_mps[3];
    int32_t outward;
    float noise_scale_m;
    float noise_strength;
    int32_t stream;
} plm_emitter;

/* ViewParams, render/raymarch.py:30-42. perspective: PLM_VIEW_ORTHO (0) and PLM_VIEW_PERSPECTIVE (1) as the
   reference; PLM_VIEW_ORTHO_NEAR (2, no reference counterpart) is the orthographic view with every ray started
   on the camera plane through the world origin instead of FAR_M = 1e4 m toward the camera. The rays and their
   samples are the same lines and points; FP32 positions then carry no rounding of a 1e4 m offset (~5e-4 m),
   which the differentiable rollout needs on centimetre cells. */
#define PLM_VIEW_ORTHO 0
#define PLM_VIEW_PERSPECTIVE 1
#define PLM_VIEW_ORTHO_NEAR 2
typedef struct plm_view {
    float right[3];
    float up[3];
    float toward[3];
    float left_m;
    float top_m;
    float pixel_m;
    int32_t perspective;
    float eye[3];
} plm_view;

/* LookParams, render/raymarch.py:45-52. */
typedef struct plm_look {
    float opacity_per_density;
    float glow_start_k;
    float smoke_hides_heat;
    float normal_fade_thickness;
    int32_t shade_layout;
    float shade_fill;
} plm_look;

/* ------------------------------------------------------------------ library */

/* Human-readable message of this thread's last failed call ("" if none). */
const char *plm_last_error(void);
/* ABI version of this header (bumped on any signature change). */
int plm_abi_version(void);

/* ------------------------------------------------------------------ advection: sim/advect.py */

/* semi_lagrangian_kernel, sim/advect.py:24-31 (advect_semi_lagrangian :53-56, _launch_trace :71-76):
   dst = sample_field(src, trace(index + offset, dt_over_cell)). `shape` is src's shape. */
int plm_advect_sl(const plm_grid_info *info, const float *src, float *dst, const int32_t shape[3],
                  const float offset[3], float outside, int32_t uses_outside, const float *u, const float *v,
                  const float *w, float dt_over_cell, void *stream);

/* maccormack_correct_kernel, sim/advect.py:34-50. */
int plm_maccormack_correct(const plm_grid_info *info, const float *original, const float *hat, const float *back,
                           float *dst, const int32_t shape[3], const float offset[3], float outside,
                           int32_t uses_outside, const float *u, const float *v, const float *w,
                           float dt_over_cell, void *stream);

/* advect_maccormack, sim/advect.py:59-68: data -> hat (SL +dt), hat -> back (SL -dt), correct -> next.
   With fuse_back != 0 the backward pass is evaluated inside the correction (same arithmetic, `back` is
   neither written nor read, one launch fewer). */
int plm_advect_maccormack(const plm_grid_info *info, const float *data, float *hat, float *back, float *next,
                          const int32_t shape[3], const float offset[3], float outside, int32_t uses_outside,
                          const float *u, const float *v, const float *w, float dt_over_cell, int32_t fuse_back,
                          void *stream);

/* Stepper._advect_velocity, sim/stepper.py:46-52: MacCormack of u, v and w by the same (old) velocity,
   in one launch per pass over all three face arrays. *_next receive the results (callers swap). */
int plm_advect_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w, float *u_hat,
                        float *v_hat, float *w_hat, float *u_next, float *v_next, float *w_next,
                        float dt_over_cell, void *stream);

/* The same velocity advection with the clamp range carried from the forward pass: the forward pass also
   writes the min/max of the 8 taps it blended into *_bounds (2 x the field's size: lo, then hi), and the
   correction reads them instead of retracing (sim/advect.py:43-49 takes the range at that same source).
   The trace's first, grid-aligned velocity sample skips its zero-weight taps. Bit-identical to
   plm_advect_velocity for finite fields. */
int plm_advect_velocity_bounded(const plm_grid_info *info, const float *u, const float *v, const float *w,
                                float *u_hat, float *v_hat, float *w_hat, float *u_bounds, float *v_bounds,
                                float *w_bounds, float *u_next, float *v_next, float *w_next, float dt_over_cell,
                                void *stream);

/* Stepper._advect_scalars, sim/stepper.py:54-60: MacCormack of two cell-centred fields (smoke and
   temperature) that move with the same velocity, sharing one trace per cell; each field reads its
   `outside` value beyond open faces. The second field's pointers may be NULL (one field). Bounds as in
   plm_advect_velocity_bounded. The smoke budget (plm_smoke_budget) runs after this on the first field. */
int plm_advect_scalars(const plm_grid_info *info, const float *a_data, float *a_hat, float *a_bounds, float *a_next,
                       float a_outside, const float *b_data, float *b_hat, float *b_bounds, float *b_next,
                       float b_outside, const float *u, const float *v, const float *w, float dt_over_cell,
                       void *stream);

/* ------------------------------------------------------------------ forces: sim/forces.py */

/* buoyancy_kernel, sim/forces.py:13-31, on v. */
int plm_buoyancy(const plm_grid_info *info, float *v, const float *smoke, const float *temperature, float gravity,
                 float heat_buoyancy, float smoke_weight, float ambient_k, float dt, void *stream);

/* drag_kernel on u, v, w, sim/forces.py:34-48 (keep = exp(-drag*dt) formed by the caller in double). */
int plm_drag(const plm_grid_info *info, float *u, float *v, float *w, const float wind[3], float keep,
             void *stream);

/* ------------------------------------------------------------------ vorticity: sim/vorticity.py */

/* center_velocity_kernel, sim/vorticity.py:17-22 (center: vec3 per cell). */
int plm_center_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w, float *center,
                        void *stream);
/* curl_kernel, sim/vorticity.py:38-45. */
int plm_curl(const plm_grid_info *info, const float *center, float *curl, void *stream);
/* confinement_push_kernel, sim/vorticity.py:53-71. */
int plm_confinement_push(const plm_grid_info *info, const float *curl, float *push, float strength,
                         void *stream);
/* add_face_push_kernel on u, v, w, sim/vorticity.py:74-84. */
int plm_add_face_push(const plm_grid_info *info, float *u, float *v, float *w, const float *push, float dt,
                      void *stream);
/* VorticityConfinement.apply, sim/vorticity.py:95-106: the four passes; nothing when strength is 0.
   With fuse != 0 the centre-velocity pass is folded into the curl pass (center is not written). */
int plm_confinement(const plm_grid_info *info, float *u, float *v, float *w, float *center, float *curl,
                    float *push, float strength, float dt, int32_t fuse, void *stream);

/* ------------------------------------------------------------------ emitters, noise, turbulence */

/* emit, sim/emitters.py:147-153: emit_scalars_kernel (:90-102) and emit_velocity_kernel (:105-121) on all
   three face arrays. `emitters` is a device array of n_emitters structs. */
int plm_emit(const plm_grid_info *info, float *smoke, float *temperature, float *u, float *v, float *w,
             const plm_emitter *emitters, int32_t n_emitters, int32_t seed, float time_s, float dt, void *stream);
/* The same two emit kernels (sim/emitters.py:90-121, 147-153) launched only over the cells [lo, hi) and their
   faces, in one launch. Every cell and face outside that box must lie at least an emitter's extent (radius or
   half size) from each active emitter's centre, where coverage is 0 and the kernels change nothing: the
   caller passes a box that holds every active emitter's support (plume_amd.sim.emitters.emit_box). The result
   is then bit for bit plm_emit's. 0 <= lo <= hi <= dims per axis. */
int plm_emit_box(const plm_grid_info *info, float *smoke, float *temperature, float *u, float *v, float *w,
                 const plm_emitter *emitters, int32_t n_emitters, int32_t seed, float time_s, float dt,
                 const int32_t lo[3], const int32_t hi[3], void *stream);

/* value_noise, sim/noise.py:40-57, at `count` points (xyz AoS) -> out; lattice_value (:24-32) when
   points is NULL and `lattice` holds (i,j,k) int triples. For parity tests of the integer hash. */
int plm_noise_probe(int32_t seed, int32_t stream_id, const float *points, const int32_t *lattice, int32_t count,
                    float *out, void *stream);

/* apply_turbulence, sim/turbulence.py:59-67 (turbulence_kernel :35-53) on u, v, w;
   pattern_time = time_s * speed / scale formed by the caller in double. */
int plm_turbulence(const plm_grid_info *info, float *u, float *v, float *w, int32_t seed, float scale_m,
                   float strength, float pattern_time, float dt, void *stream);

/* dissipate_kernel, sim/dissipate.py:10-24 (keeps formed by the caller in double). */
int plm_dissipate(int64_t count, float *smoke, float *temperature, float smoke_keep, float heat_keep,
                  float ambient_k, void *stream);

/* ------------------------------------------------------------------ pressure: sim/pressure.py */

/* zero_closed_faces_kernel on u, v, w, sim/pressure.py:21-28 (only boundary planes are touched). */
int plm_zero_closed_faces(const plm_grid_info *info, float *u, float *v, float *w, void *stream);
/* divergence_kernel, sim/pressure.py:31-36. */
int plm_divergence(const plm_grid_info *info, const float *u, const float *v, const float *w, float *divergence,
                   void *stream);
/* project's first two steps, sim/pressure.py:146-149 (zero_closed_faces_kernel, :21-28, then divergence_kernel,
   :31-36, on the same faces) in one launch: the faces and the divergence end bit for bit as the two calls above
   leave them. */
int plm_close_divergence(const plm_grid_info *info, float *u, float *v, float *w, float *divergence, void *stream);
/* The per-cell boundary flags behind neighbor_term, sim/pressure.py:39-69: for neighbours x-,x+,y-,y+,z-,z+
   two bits each (0 in-domain, 1 open-face ghost, 2 closed), bits 12..14 = last cell on x, y, z. */
int plm_build_flags(const plm_grid_info *info, uint16_t *flags, void *stream);
/* red_black_sweep_kernel, sim/pressure.py:80-92: one colour, in place on q. */
int plm_rb_sweep(const plm_grid_info *info, float *q, const float *rhs, int32_t color, float relaxation,
                 void *stream);
/* red_black_sweeps, sim/pressure.py:135-141: `sweeps` red+black sweeps, in place on q. With `scratch` (a
   level-sized float buffer, or NULL) levels of 4096+ cells run both colours of a sweep in one LDS-tiled
   launch that ping-pongs q -> scratch -> q; the result is bit-identical to the colour-by-colour launches
   (each colour reads only the other). */
int plm_rb_sweeps(const plm_grid_info *info, float *q, const float *rhs, float *scratch, int32_t sweeps,
                  float relaxation, void *stream);
/* subtract_gradient_kernel on u, v, w, sim/pressure.py:95-116. */
int plm_subtract_gradient(const plm_grid_info *info, float *u, float *v, float *w, const float *q, void *stream);

/* Tile shapes of the two-colour sweep (i x j x 64 cells per block): fills up to `capacity` entries of each
   (NULL skips an array) and returns how many shapes there are. */
int32_t plm_sweep_tile_shapes(int32_t *ti, int32_t *tj, int32_t *threads, int32_t capacity);
/* The project's autotuner for the sweep tile (no reference counterpart; the sweep is bit-identical under
   every shape): times each shape on this level (2 * repeats sweeps per shape, min(2 * repeats, 64), the
   shapes taking turns, ping-ponging q <-> scratch, which are overwritten; each sweep timed by its own
   dispatch's begin and end), writes each shape's median microseconds per sweep into times_us (may be NULL),
   pins the fastest shape for this level size for the rest of the process and returns it in *chosen.
   Synchronises the stream. */
int plm_autotune_sweep(const plm_grid_info *info, float *q, const float *rhs, float *scratch, int32_t repeats,
                       float *times_us, int32_t *chosen, void *stream);
/* Pins a tile shape for a level size (-1 restores the default). */
int plm_set_sweep_tile(const int32_t dims[3], int32_t shape);

/* ------------------------------------------------------------------ multigrid: sim/multigrid.py */

/* residual_kernel, sim/multigrid.py:36-43. */
int plm_residual(const plm_grid_info *info, const float *q, const float *rhs, float *residual, void *stream);
/* restrict_kernel, sim/multigrid.py:46-60. */
int plm_restrict(const int32_t fine_shape[3], const float *fine, const int32_t coarse_shape[3], float *coarse,
                 void *stream);
/* prolong_add_kernel, sim/multigrid.py:63-68. */
int plm_prolong_add(const int32_t coarse_shape[3], const float *coarse, const int32_t fine_shape[3], float *fine,
                    void *stream);
/* Fills `count` floats with 0 by a kernel (coarse.q.zero_(), sim/multigrid.py:143; graph-capture safe). */
int plm_zero(int64_t count, float *f, void *stream);

/* Multigrid.__call__ / _v_cycle, sim/multigrid.py:124-147, over caller-owned level arrays: level 0 is the
   simulation grid (q[0] = pressure, rhs[0] = divergence), level l+1 = coarse_cells(level l). Runs
   `cycles` V(sweeps, sweeps) cycles; the coarsest level gets max(20, 4*longest) sweeps at 1.5.
   Levels with at most PLM_LDS_CELLS cells run as one workgroup in LDS (one launch for that whole tail).
   If tolerance > 0 and norm_scratch is given, the cycle count becomes a cap: after each cycle the
   FP64 relative residual |rhs - Aq| / |rhs| of level 0 is reduced on the device, and once it is at or
   below `tolerance` the remaining cycles' kernels exit at once (no host round trip; graph-safe).
   norm_scratch: device doubles, at least plm_norm_scratch_doubles(info[0]); cycles_run (device int32[2],
   may be NULL): [0] receives the number of cycles this solve ran, [1] is a running total across solves
   that the library only increments; norm_scratch[0] receives the last relative residual.
   resume != 0 continues the previous tolerance-mode solve (no initial norm; the gate and the cycle
   count carry on), so a caller can run a predicted number of cycles, read the gate, and add more. */
#define PLM_LDS_CELLS 8192
int plm_vcycle(int32_t n_levels, const plm_grid_info *infos, float *const *q, float *const *rhs,
               float *const *residual, int32_t cycles, int32_t sweeps, double tolerance, int32_t resume,
               double *norm_scratch, int32_t *cycles_run, void *stream);
int64_t plm_norm_scratch_doubles(const plm_grid_info *info);

/* Batched multigrid for an ensemble of independent solves of uneven size (no reference counterpart: the
   reference solves one grid per run, sim/run.py:57-58; this serves the batched ensemble). Each member is a
   plm_vcycle level list (level 0 its simulation grid), with its own norm_scratch (zero-filled doubles,
   plm_norm_scratch_doubles of its level 0), cycles_run (device int32[2]) and cap on cycles. A solve checks
   every member's FP64 relative residual, then runs up to `cycles` V(sweeps, sweeps) cycles in lockstep, each
   stage of a cycle one launch for all members (a persistent grid of blocks over the members' tiles, blocks of
   one workgroup per member for the LDS tails), and the norms after each cycle. A member stops once its residual
   is at or below `tolerance` or it has run max_cycles[m] cycles, so each member ends exactly as its own
   plm_vcycle(tolerance, max_cycles) would (bit for bit) when `cycles` covers what it needs; otherwise its gate
   (norm_scratch[1]) is still 1 afterwards. create validates, builds the stage tables and copies them to the
   device (the only allocation: freed by destroy); solve allocates nothing and reads nothing back (graph-safe).
   Requirements: 1 to 64 members, tolerance > 0, sweeps even, every member's coarse levels form an LDS tail.
   The arrays must stay where they are for the batch's lifetime. */
typedef struct plm_mg_batch plm_mg_batch;
int plm_mg_batch_create(int32_t members, const int32_t *n_levels, const plm_grid_info *const *infos,
                        float *const *const *q, float *const *const *rhs, float *const *const *residual,
                        int32_t sweeps, double tolerance, const int32_t *max_cycles, double *const *norm_scratch,
                        int32_t *const *cycles_run, plm_mg_batch **out);
int plm_mg_batch_solve(plm_mg_batch *batch, int32_t cycles, void *stream);
int plm_mg_batch_destroy(plm_mg_batch *batch);
/* The batch's autotuner (project-written; the sweep is bit-identical under every shape): at each depth of the
   V-cycle that runs tile sweeps, times every tile shape for a launch over all members (2 * repeats launches of
   a first sweep from 0 into the members' scratch levels, which are overwritten; q is not touched), writes the
   microseconds into times_us[depth * shapes + shape] (may be NULL; plm_mg_batch_depth(batch) depths and
   plm_sweep_tile_shapes shapes), keeps the fastest shape per depth and rebuilds the batch's table.
   Synchronises the stream; not for graph capture. */
int plm_mg_batch_autotune(plm_mg_batch *batch, int32_t repeats, float *times_us, void *stream);
/* Depths of the batch that run tile sweeps (the largest first level of any member's LDS tail); -1 for NULL. */
int32_t plm_mg_batch_depth(const plm_mg_batch *batch);

/* |rhs - A q|^2 and |rhs|^2 of one level in FP64, deterministic (fixed-order tree), into out[0..1]
   (device). For the convergence tests of sim/multigrid.py's V-cycle. */
int plm_residual_norms(const plm_grid_info *info, const float *q, const float *rhs, double *scratch, double *out,
                       void *stream);

/* ------------------------------------------------------------------ mixed-precision pressure solve (no
   reference counterpart: the reference solves in FP32 only). Storage of the correction solve's q and rhs. */
#define PLM_STORAGE_F32 0
#define PLM_STORAGE_BF16 1
#define PLM_STORAGE_F16 2
/* FP16 storage with the sweeps of flat levels (the simulation level, and coarser levels whose fill_hi is
   (1, 1, 1)) in packed FP16 arithmetic (two cells per lane: v_pk_add_f16, v_pk_fma_f16; the per-cell factors
   1 / t1 and -cell_m^2 rhs / t1 formed in FP32 and rounded once). Other levels, residuals and refinement are
   those of PLM_STORAGE_F16; the entry points below that only store take it as PLM_STORAGE_F16. A level sweeps
   packed where its k extent is a multiple of 4 and every extent is at least 2 (the packed tile's aligned
   4-cell groups); elsewhere as PLM_STORAGE_F16. */
#define PLM_STORAGE_F16_PACKED 3

/* `cycles` V(sweeps, sweeps) cycles of Multigrid._v_cycle (sim/multigrid.py:130-147) with q, rhs and scratch of
   every level stored as `storage` (FP32 arithmetic in the sweeps and the FP32 residual inside the fused
   residual + restrict), and the coarsest level (at most 4096 cells) solved in FP64 with the reference's
   max(20, 4 * longest) sweeps at 1.5. zero_start: the first cycle starts from q = 0 (q is not read). q/rhs/
   scratch: device arrays of the storage type; the coarsest level needs no scratch. sweeps >= 1. */
int plm_vcycle_lp(int32_t n_levels, const plm_grid_info *infos, int32_t storage, void *const *q, void *const *rhs,
                  void *const *scratch, int32_t cycles, int32_t sweeps, int32_t zero_start, void *stream);
/* `sweeps` two-colour sweeps at relaxation 1 of one level in `storage` (red_black_sweeps, sim/pressure.py:80-92,
   as the V-cycle of plm_vcycle_lp runs them on that level): q and scratch ping-pong, the result in q (an odd
   count ends with a copy back). zero_start: the first sweep starts from q = 0 (q is not read). */
int plm_rb_sweeps_lp(const plm_grid_info *info, int32_t storage, void *q, const void *rhs, void *scratch,
                     int32_t sweeps, int32_t zero_start, void *stream);
/* residual_kernel + restrict_kernel (sim/multigrid.py:36-60) of one level in `storage`, as the V-cycle of
   plm_vcycle_lp forms the next level's right-hand side: coarse_rhs (the level's dims halved, rounded up) =
   restrict(rhs - A q), the residual in FP32 and stored rounded to nearest even. */
int plm_residual_restrict_lp(const plm_grid_info *info, int32_t storage, const void *q, const void *rhs,
                             void *coarse_rhs, void *stream);
/* q += prolongation(coarse) (prolongation_kernel, sim/multigrid.py:63-80) on one level in `storage`, as the
   V-cycle of plm_vcycle_lp applies the coarse correction: the sum in FP32, stored rounded to nearest even.
   coarse has the level's dims halved, rounded up. */
int plm_prolong_add_lp(const plm_grid_info *info, int32_t storage, const void *coarse, void *q, void *stream);
/* The tile shape PLM_STORAGE_F16_PACKED sweeps launch with, an index into {8 x 8 on 256 threads, 8 x 8 on 512,
   4 x 8 on 256} (i x j, 64 cells along k); process-wide. */
int plm_set_f16x2_shape(int32_t shape);
/* dst = storage(src * scale), rounded to nearest even. */
int plm_to_storage(int64_t count, const float *src, float scale, int32_t storage, void *dst, void *stream);
/* x = x + float(e) * inv_scale. */
int plm_add_storage(int64_t count, float *x, const void *e, int32_t storage, float inv_scale, void *stream);
/* One refinement step, fused (no reference counterpart; it replaces plm_add_storage, plm_residual,
   plm_to_storage and plm_residual_norms in a row): x_out = x + float(e) * inv_scale (x alone when e is NULL),
   r = rhs - A x_out in FP32 (residual_kernel, sim/multigrid.py:36-43), r_out = storage(r * r_scale) when r_out
   is not NULL, out[0..1] = |r|^2 and |rhs|^2 in FP64 (device, deterministic). x and x_out must not overlap;
   x_out may be NULL when e is (nothing written there). scratch holds plm_norm_scratch_doubles(info) doubles. */
int plm_refine_residual(const plm_grid_info *info, const float *x, const void *e, float inv_scale, int32_t storage,
                        const float *rhs, float *x_out, void *r_out, float r_scale, double *scratch, double *out,
                        void *stream);

/* ------------------------------------------------------------------ adjoints (no reference counterpart: the
   reference runs no backward pass, sim/device.py:28). Reverse-mode derivatives of the differentiable substep
   (emission by per-cell sources, buoyancy, confinement, MacCormack advection, projection) and of the ray
   march's heat and thickness channels. Each *_bar array is an adjoint of the same shape as its field; unless
   stated otherwise the entry points add into them. */

/* buoyancy_kernel (sim/forces.py:13-31): smoke_bar and temperature_bar += the adjoint of v's update given v_bar;
   weight_terms (v faces) = v_bar * d v / d smoke_weight, for the caller's sum (the forward's v passes through). */
int plm_adj_buoyancy(const plm_grid_info *info, const float *v_bar, const float *smoke, float gravity,
                     float heat_buoyancy, float smoke_weight, float ambient_k, float dt, float *smoke_bar,
                     float *temperature_bar, float *weight_terms, void *stream);
/* VorticityConfinement.apply (sim/vorticity.py:95-106) at the forward input velocity u, v, w: the face
   adjoints (in: of the output, out: of the input). curl (vec3), push_bar (vec3), curl_bar (vec3), spin_bar
   (cells) and center_bar (vec3) are scratch. */
int plm_adj_confinement(const plm_grid_info *info, const float *u, const float *v, const float *w, float strength,
                        float dt, float *u_bar, float *v_bar, float *w_bar, float *curl, float *push_bar,
                        float *curl_bar, float *spin_bar, float *center_bar, void *stream);
/* MacCormack of u, v, w by the old velocity (sim/stepper.py:46-52, sim/advect.py:24-68): next_*_bar in,
   *_bar (zeroed by the caller) out, through both the advected fields and the velocity that traces them.
   hat_* and hat_*_bar are scratch of the face shapes; dto = dt / cell_m. */
int plm_adj_advect_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w, float *hat_u,
                            float *hat_v, float *hat_w, const float *next_u_bar, const float *next_v_bar,
                            const float *next_w_bar, float *u_bar, float *v_bar, float *w_bar, float *hat_u_bar,
                            float *hat_v_bar, float *hat_w_bar, float dto, void *stream);
/* MacCormack of two cell-centred fields with outside values (smoke, temperature) by u, v, w: a_bar and b_bar
   (zeroed by the caller) from the next adjoints; u_bar, v_bar, w_bar accumulate the velocity's part. */
int plm_adj_advect_scalars(const plm_grid_info *info, const float *u, const float *v, const float *w,
                           const float *a_data, const float *b_data, float a_outside, float b_outside, float *a_hat,
                           float *b_hat, const float *a_next_bar, const float *b_next_bar, float *a_bar,
                           float *b_bar, float *u_bar, float *v_bar, float *w_bar, float *a_hat_bar,
                           float *b_hat_bar, float dto, void *stream);
/* Projection transposes (sim/pressure.py:95-121, 144-154): rhs = G^T f_bar (cells, written), and
   f_bar = Z (f_bar - D^T y) (closed faces end at 0). With y the Poisson solve of rhs, the two give the
   transpose of the projection. */
int plm_adj_gradient_t(const plm_grid_info *info, const float *u_bar, const float *v_bar, const float *w_bar,
                       float *rhs, void *stream);
int plm_adj_divergence_t(const plm_grid_info *info, const float *y, float *u_bar, float *v_bar, float *w_bar,
                         void *stream);
/* raymarch_kernel's heat and thickness (render/raymarch.py:86-140), orthographic: smoke_bar and
   temperature_bar += the adjoint of image_bar (height, width, 5; channels 0 and 1 are read). max_steps as in
   plm_raymarch (0: half-cell steps). record: scratch of height * width * stride * 4 floats, stride from
   plm_adj_raymarch_record_stride. */
int plm_adj_raymarch_record_stride(const plm_grid_info *info, int32_t max_steps);
int plm_adj_raymarch(const plm_grid_info *info, const float *smoke, const float *temperature, const plm_view *view,
                     const plm_look *look, const float *image_bar, int32_t height, int32_t width, int32_t max_steps,
                     float *record, float *smoke_bar, float *temperature_bar, void *stream);

/* ------------------------------------------------------------------ budget: sim/reduce.py, conserve.py */

/* line_sums_kernel, sim/reduce.py:12-25: lines[3][ny][nz] = per-(j,k) sums over i, sequential FP32. */
int plm_line_sums(const int32_t shape[3], const float *a, const float *b, const float *c, float *lines,
                  void *stream);
/* SmokeBudget.hold, sim/conserve.py:45-52, on the device with no host round trip: line sums, their FP64
   totals (fixed-order tree, in place of the host NumPy sum at sim/reduce.py:37), held_total (:33-35),
   factor = target / maccormack (1 if maccormack <= 0) and, when factor != 1, next *= (float)factor.
   lines: 3*ny*nz floats; totals (device doubles, 4): before, semi-Lagrangian, MacCormack, factor. */
int plm_smoke_budget(const int32_t shape[3], const float *data, const float *hat, float *next, float *lines,
                     double *totals, void *stream);
/* No counterpart in the reference: plm_smoke_budget for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same kernels and results,
   but the line sums and FP64 totals live in a static device array of the library (each GPU holds its own, of
   3 * PLM_BUDGET_NOALLOC_MAX_LINES floats plus 4 doubles, until the process ends), so nothing is allocated.
   Grids with ny * nz above PLM_BUDGET_NOALLOC_MAX_LINES are rejected with PLM_EINVAL. totals (device, 4
   doubles, or NULL) receives before, semi-Lagrangian, MacCormack and factor. Calls on one GPU share the
   array: calls that overlap on one GPU (other streams or threads) race. Call with the stream's device
   current. */
#define PLM_BUDGET_NOALLOC_MAX_LINES 262144
int plm_smoke_budget_noalloc(const int32_t shape[3], const float *data, const float *hat, float *next,
                             double *totals, void *stream);
/* scale_kernel, sim/conserve.py:27-30. */
int plm_scale(int64_t count, float *f, float factor, void *stream);

/* ------------------------------------------------------------------ CFL cap: sim/step_cap.py */

/* StepCap.grid_speed, sim/step_cap.py:50-55: max |f| over u, v, w into *out (device float). */
int plm_max_abs3(const plm_grid_info *info, const float *u, const float *v, const float *w, float *out,
                 void *stream);

/* ------------------------------------------------------------------ rendering: render/ */

/* light_volume_kernel, render/light_volume.py:18-33 (toward_light: unit, cell units). max_steps > 0 fixes
   the march at that many steps per cell (a fixed-count bench mode: step = length / max_steps); 0 keeps the
   reference's half-cell steps. */
int plm_light_volume(const int32_t shape[3], const float *smoke, const float toward_light[3], float cell_m,
                     float opacity_per_density, int32_t max_steps, float *light, void *stream);
/* blur_axis_kernel, render/normal_field.py:12-22 (weights: device, 2*radius+1 floats). */
int plm_blur_axis(const int32_t shape[3], const float *src, float *dst, const float *weights, int32_t radius,
                  int32_t axis, void *stream);
/* gradient_kernel, render/normal_field.py:25-38 (gradient: vec3 per cell). */
int plm_gradient(const int32_t shape[3], const float *src, float *gradient, void *stream);
/* raymarch_kernel, render/raymarch.py:86-140: image (height, width, 5). gradient / light are the layout's
   volume (the unused one may be a 1-cell stand-in, render/frame_renderer.py:29-35). max_steps > 0 fixes
   the march at that many samples per ray (bench mode); 0 keeps half-cell steps. */
int plm_raymarch(const plm_grid_info *info, const float *smoke, const float *temperature, const float *gradient,
                 const int32_t gradient_shape[3], const float *light, const int32_t light_shape[3],
                 const plm_view *view, const plm_look *look, float *image, int32_t height, int32_t width,
                 int32_t max_steps, void *stream);

/* ------------------------------------------------------------------ viewer: view/ */

/* narrow_kernel, view/frame_cache.py:26-29: count FP32 values rounded to IEEE binary16 (round to nearest
   even, subnormals kept) into dst (count halves). src 16-byte and dst 8-byte aligned. */
int plm_narrow_f16(int64_t count, const float *src, void *dst, void *stream);
/* widen_kernel and widen_vector_kernel, view/frame_cache.py:32-35,45-49: count binary16 values to FP32 (a vec3h
   array is 3N halves). src 8-byte and dst 16-byte aligned. */
int plm_widen_f16(int64_t count, const void *src, float *dst, void *stream);
/* center_velocity_kernel (sim/vorticity.py:17-22) then narrow_vector_kernel (view/frame_cache.py:38-42), in one
   kernel: per cell (0.5 (u_i + u_i+1), 0.5 (v_j + v_j+1), 0.5 (w_k + w_k+1)) in FP32, rounded to binary16, into
   center as 3 halves per cell (nx, ny, nz, 3). */
int plm_narrow_center_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w,
                               void *center, void *stream);
/* boundary_layer_kernel, view/live_run.py:41-57: layers (6, extent_a, extent_b), face = 2 axis + side in
   left, right, bottom, top, back, front order, a and b the face's other two axes in x, y, z order; entries
   past a face's size are 0. */
int plm_boundary_layers(const int32_t shape[3], const float *smoke, float *layers, int32_t extent_a,
                        int32_t extent_b, void *stream);

/* ------------------------------------------------------------------ timing */

/* Kernel timing for the benchmark: while enabled, the library brackets every launch of the named kernel
   family ("rb_sweeps", "advect", ...) with HIP events on the launch stream and accumulates their elapsed
   time; a single-kernel family ("rb_sweep_tile_top", "rb_sweep_tile", "narrow_f16", "widen_f16",
   "narrow_center_velocity") binds the events to the kernel's own
   dispatch (hipExtLaunchKernelGGL), so they time the kernel alone. plm_timing_read synchronises those
   events and returns total milliseconds and launch count. No counterpart in the reference. */
int plm_timing_enable(const char *family);
int plm_timing_disable(void);
int plm_timing_read(double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* PLUME_HIP_H */
