/*
 * ffk -- C ABI of the MI355X (gfx950) Monte Carlo form-factor kernels.
 *
 * Plain C: opaque handles, plain pointers and sizes, int return codes, and a
 * hipStream_t passed as void*. "d_" pointers are device (HBM) pointers,
 * "h_" pointers are host pointers. Every call returns FFK_OK (0) or a
 * negative code; HIP failures never abort the process.
 *
 * The reference's plugin API (ray_caster / emission / form_factors /
 * thermal_equation / thermal_solution system_t method tables) is implemented
 * on top of these calls, see include/ff_plugin.h. Each entry below names the
 * reference function whose work it performs.
 *
 * Streams: work is queued on the stream passed to each call. Different
 * scenes (and batches) keep their own device buffers and may run at the same
 * time on different streams; calls on one scene are issued one stream at a
 * time (they share that scene's argument and work buffers).
 *
 * Layouts in HBM:
 *   faces : float[9*n]  p0 p1 p2 (AoS, math/types.h:51-54)
 *   rays  : float[6*n]  origin, END POINT (math/types.h:72-76; the direction
 *                        is end - origin, math/triangle.cpp:41)
 *   hit   : int32[n]    face index or -1;  points: float[3*n] (hit only)
 *   tally : uint64[N*N] per (source mesh, hit mesh) exact ray counts
 *   F     : float[N*N]  row-major form factors (two-sided convention)
 */
#ifndef FFK_H
#define FFK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes -------------------------------------------------------- */
#define FFK_OK 0
#define FFK_ERR_ARG (-1)      /* invalid argument or state                     */
#define FFK_ERR_HIP (-2)      /* a HIP runtime call failed                     */
#define FFK_ERR_NOMEM (-3)    /* device allocation failed                      */
#define FFK_ERR_LIMIT (-4)    /* size beyond a documented limit                */
#define FFK_ERR_EMPTY (-5)    /* empty scene (no faces / no meshes)            */

/* ---- enums ------------------------------------------------------------------ */
#define FFK_CAST_NAIVE 0      /* brute force, triangle SoA tiles staged in LDS  */
#define FFK_CAST_GRID 1       /* uniform 3-D grid built on the GPU, DDA walk    */

#define FFK_FRAME_REFERENCE 0   /* rotate_towards(z, unnormalised normal),     */
                                /* as emission/malley_cpu.cpp:125-129           */
#define FFK_FRAME_ORTHONORMAL 1 /* rotate_towards(z, unit normal)              */

#define FFK_F32 0
#define FFK_BF16 1
#define FFK_F16 2

typedef struct ffk_scene ffk_scene;
typedef struct ffk_batch ffk_batch;

/* Last error text for this thread (static storage, never NULL). */
const char* ffk_last_error(void);

/* Device properties used for launch sizing and roofline reporting. */
int ffk_device_info(int* n_cu, int* clock_khz, int* mem_clock_khz, int* bus_width_bits,
                    int* l2_bytes, char* name, int name_len);

/* ---- scene ------------------------------------------------------------------ */
/* A device-resident scene: faces, per-face intersection terms, face->mesh map,
 * emission layout and the cast grid. */
int ffk_scene_create(ffk_scene** out);
void ffk_scene_destroy(ffk_scene* s);

/* Upload n_faces faces (host AoS) and precompute per-face terms on the device.
 * Replaces the faces half of ray_caster::system_set_scene + prepare
 * (ray_caster/system.h:94-98; cuda_ray_caster/system.cu:93-108). */
int ffk_scene_set_faces(ffk_scene* s, const float* h_faces, int n_faces, void* stream);

/* Mesh table (first_idx, n_faces, pad) triples in scene order
 * (form_factors/system.h:33-38); builds face->mesh on the device
 * (cpuFactorsCalculator/cpu_system.cpp:108-118). */
int ffk_scene_set_meshes(ffk_scene* s, const int* h_meshes, int n_meshes, void* stream);

/* Emission layout: per-face front/rear ray counts (emitted_front/rear,
 * emission/malley_emission.cpp:50-60, evaluated by the caller or by
 * ffk_emitted_counts) and the rotation frame. Rays are packed face by face,
 * front block then rear block (emission/system.h:48). Also records the
 * mesh-order ray ranges that the tally walks (cpu_system.cpp:163-191).
 * Returns the total ray count through *n_rays_out. */
int ffk_scene_set_emission(ffk_scene* s, const int* h_front, const int* h_rear, int frame,
                           long long* n_rays_out, void* stream);

/* Build the uniform grid on the device (count, sort by cell, cell starts).
 * nx*ny*nz == 0 picks a resolution from the triangle count. Replaces the
 * host-built 2-D zgrid (ray_caster/zgrid_cpu.cpp:57-68). */
int ffk_scene_build_grid(ffk_scene* s, int nx, int ny, int nz, void* stream);
int ffk_scene_grid_info(const ffk_scene* s, int* dims3, float* lo3, float* hi3,
                        long long* n_refs);
/* Host copy of the grid's cell words, two int32 per cell, cell
 * c = x + nx*(y + ny*z): (first reference, count > 0) for an occupied cell;
 * (boxes, 0) for an empty one, where bits 4o..4o+3 of boxes hold the size k
 * (1..15) of the empty box [c, c + s*(k-1)] toward direction octant o (bit 0
 * of o: x decreasing, bit 1: y, bit 2: z), clipped to the grid, cells outside
 * counting as empty. Synchronises the device. An inspection call with no
 * counterpart in the reference (its zgrid keeps no empty-space data). */
int ffk_scene_grid_cells(const ffk_scene* s, int* h_cells);
int ffk_scene_counts(const ffk_scene* s, int* n_faces, int* n_meshes, long long* n_rays);
/* Host copies of the scene's faces (9 per face), meshes (3 per mesh), emitted
 * front/rear ray counts and emission frame; any pointer may be NULL. */
int ffk_scene_layout(const ffk_scene* s, float* h_faces, int* h_meshes, int* h_front, int* h_rear, int* frame);

/* ---- casting ---------------------------------------------------------------- */
/* Nearest hit for n_rays device rays: naive_cpu::cast semantics
 * (ray_caster/naive_cpu.cpp:65-97): the reference triangle test bit for bit,
 * nearest by squared distance, ties to the lowest face index. d_hit_pts may
 * be NULL. d_stats (uint64[2], may be NULL) accumulates triangle tests and
 * grid cells visited. */
int ffk_cast(ffk_scene* s, int method, const float* d_rays, long long n_rays, int* d_hit_idx,
             float* d_hit_pts, unsigned long long* d_stats, void* stream);

/* ---- emission --------------------------------------------------------------- */
/* emitted_front/rear for every face: max(1, (int)(n_rays * (w / total))) in
 * float for w > 0, else 0 (emission/malley_emission.cpp:50-60). Host arrays. */
int ffk_emitted_counts(const float* h_weights, int n_faces, int n_rays, float total,
                       int* h_front, int* h_rear);

/* Two-sided area weights (w[2f] = w[2f+1] = area_f) and their float total,
 * accumulated face by face as total += 2.f * area
 * (cpuFactorsCalculator/cpu_system.cpp:120-129). Host arrays. */
int ffk_area_weights(const float* h_faces, int n_faces, float* h_weights, float* total);

/* Malley cosine-weighted rays for the whole emission layout
 * (emission/malley_cpu.cpp:142-189) from Philox4x32-10 keyed by seed, counter
 * (ray-in-face j, face, 0, 0). d_rays: float[6*n_rays]. */
int ffk_emit_malley(ffk_scene* s, unsigned long long seed, float* d_rays, void* stream);

/* Parallel rays from a disk of radius `radius` centred at center - distance*dir
 * (emission/parallel_rays_cpu.cpp:88-106), counter (ray, 0, 1, 0). */
int ffk_emit_parallel(const float* h_dir3, const float* h_center3, float distance, float radius,
                      long long n_rays, unsigned long long seed, float* d_rays, void* stream);

/* ---- tally and form factors ------------------------------------------------- */
/* Zero the N*N uint64 tally (fill kernel; graph-capture safe). */
int ffk_tally_zero(const ffk_scene* s, unsigned long long* d_tally, void* stream);
/* Exact integer form of cpu_system.cpp:163-191: ray g is counted for the mesh
 * whose mesh-order range holds g, against the mesh of the face it hit. */
int ffk_tally(const ffk_scene* s, const int* d_hit_idx, long long n_rays,
              unsigned long long* d_tally, void* stream);
/* F[m,j] = (float)tally[m,j] * (1.f / (float)rays_of_mesh_m)
 * (cpu_system.cpp:193-195). */
int ffk_normalise(const ffk_scene* s, const unsigned long long* d_tally, float* d_F, void* stream);

/* Fused emit -> cast -> tally -> normalise in one pass over the emission
 * layout (form_factors::system_calculate, form_factors/system.h:110-116).
 * d_tally: uint64[N*N] scratch owned by the caller. */
int ffk_ff_fused(ffk_scene* s, int method, unsigned long long seed, unsigned long long* d_tally,
                 float* d_F, unsigned long long* d_stats, void* stream);

/* Ray scheduling of the fused kernel. lanes: 0 = each lane casts one ray at
 * a time (the wavefront waits for its slowest ray); 1..64 = lanes walk one
 * grid cell per iteration and a wavefront starts new rays on its idle lanes
 * once at least `lanes` of its 64 lanes are idle (default 32). pool (with
 * lanes > 0): 0 = each lane tests its own cells' triangles; 1..64 = lanes
 * that reach an occupied cell wait, and once waiting lanes x pool >= active
 * lanes the wavefront spreads all their (ray, triangle) tests over its 64
 * lanes (default 2). Results are identical in every setting (integer
 * tallies, the same nearest-hit rule). */
int ffk_set_refill(ffk_scene* s, int lanes, int pool);
/* The scene's current refill threshold and pool ratio (either pointer may be NULL). */
int ffk_get_refill(const ffk_scene* s, int* lanes, int* pool);
int ffk_batch_set_refill(ffk_batch* b, int lanes, int pool);
/* Walk mode of the pooled schedule, two bits. Bit 0, how lanes leave an
 * empty grid cell: past its empty box toward the ray's direction octant (the
 * jump) or, when that box is the cell itself, by one DDA step; clear
 * (default): a wavefront runs the jump code for all its empty-cell lanes when
 * any of them can jump, the DDA step otherwise; set: always the jump code
 * (fastest when rays cross large empty regions). Bit 1, how pooled tests are
 * dealt: clear (default): every waiting lane's tests in each round; set: a
 * round of at least 64 tests deals whole chunks of 64 only, and the tests
 * past the last whole chunk wait for the next round (fuller chunks, later
 * exits: faster on scenes whose rounds are large). Bit 2, how a pooled test
 * reads the face's 64 bytes of terms: clear (default): all four 16-byte parts
 * at once; set: the plane parts first and the edge parts only in the lanes
 * whose ray meets the plane in front (fewer L1 accesses, one more load round
 * trip: faster where many tests fail at the plane). Modes 0..7; same cells,
 * same hits in every mode. Default FFK_WALK_DEFAULT (bits 0 and 1). */
#define FFK_WALK_JUMP 1
#define FFK_WALK_WHOLE 2
#define FFK_WALK_STAGED 4
#define FFK_WALK_DEFAULT (FFK_WALK_JUMP | FFK_WALK_WHOLE)
int ffk_set_walk(ffk_scene* s, int mode);
int ffk_batch_set_walk(ffk_batch* b, int mode);

/* 1 when ffk_ff_fused needs d_tally for this scene and tuning (the rays of a
 * mesh span several tiles, or the row histogram is tallied globally: by
 * default when a 16-bit histogram would take more than 48 KB of LDS, i.e.
 * more than 24,576 meshes), else 0. */
int ffk_ff_needs_tally(const ffk_scene* s);

/* Launch tuning of the fused kernel (0 = automatic; hist_mode -1 = automatic:
 * uint32 LDS up to 16,384 meshes, uint16 LDS up to 24,576, global atomics
 * above): threads per block, rays per tile (<= 65535) and the row histogram
 * (0 global atomics, 1 uint32 LDS, 2 uint16 LDS, whenever it fits). */
int ffk_set_tuning(ffk_scene* s, int block_size, int rays_per_tile, int hist_mode);

/* ---- batched form factors ---------------------------------------------------- */
/* Many scenes of uneven size through one persistent launch over a queue of
 * (scene, mesh, ray-range) tiles. Scenes must have faces, meshes, emission
 * and grid set, and must outlive the batch. The batch holds the scenes'
 * device tables and ray layouts as they are at creation: once a scene's grid
 * is rebuilt or its faces, meshes or emission are set again (directly, by
 * ffk_autotune or by a radiance step), ffk_batch_run and ffk_batch_set_tuning
 * return FFK_ERR_ARG before launching anything; create the batch again. */
int ffk_batch_create(ffk_batch** out, ffk_scene* const* scenes, int n_scenes, void* stream);
void ffk_batch_destroy(ffk_batch* b);
/* seeds[i] keys scene i; d_tallies[i] and d_Fs[i] are that scene's buffers.
 * use_graph != 0 records the launch sequence into a hipGraph on first use and
 * replays it afterwards while the tally and F buffers stay the same; new
 * seeds are copied ahead of the replay in stream order, so successive runs of
 * one batch go on one stream (or the caller orders them). */
int ffk_batch_run(ffk_batch* b, const unsigned long long* seeds, unsigned long long* const* d_tallies,
                  float* const* d_Fs, int use_graph, unsigned long long* d_stats, void* stream);
int ffk_batch_set_tuning(ffk_batch* b, int block_size, int rays_per_tile, int n_blocks);
/* Number of scenes in the batch. */
int ffk_batch_n_scenes(const ffk_batch* b);
/* The batch-wide settings of ffk_batch_set_tuning (0 = automatic; any pointer may be NULL). */
int ffk_batch_get_tuning(const ffk_batch* b, int* block_size, int* rays_per_tile, int* n_blocks);
/* Scenes whose 16-bit row histogram would exceed `bytes` of LDS are tallied
 * in the global uint64 tally instead (more resident wavefronts per CU, at
 * the cost of N^2 tally traffic); 0 restores the default (48 KB, the rule
 * of the single-scene launch). Changes which scenes need a tally buffer
 * (ffk_batch_tally_flags). */
int ffk_batch_set_lds_limit(ffk_batch* b, int bytes);
/* on != 0: the launches of the LDS classes go on forked streams and may
 * overlap; 0 (default): one after another on the launch stream. */
int ffk_batch_set_concurrent(ffk_batch* b, int on);
/* Tile queue order inside an LDS class: 0 longest tiles first across all
 * scenes; 1 (default) scene by scene (most rays first), longest tiles first
 * within a scene, which keeps one scene's grid and face tables in L2. */
int ffk_batch_set_order(ffk_batch* b, int order);
/* The batch's LDS classes (one persistent launch each, run in this order):
 * ids[c] (0..4: 16-bit row histogram of at most 8, 24, 48, 96, 160 KB; 5:
 * global tally), their rays and queued tiles; returns the number of classes
 * (arrays filled up to capacity; any may be NULL). */
int ffk_batch_classes(const ffk_batch* b, int* ids, long long* rays, int* n_tiles, int capacity);
/* A 64-bit signature of the batch (every scene's faces, meshes, ray counts,
 * frame and grid shape, and the batch-wide settings): the key of its tuning
 * results. */
int ffk_batch_signature(const ffk_batch* b, unsigned long long* key);
/* Launch settings of one class (by id), overriding the batch-wide ones; -1 =
 * the batch-wide value (block_size <= 0 likewise). None changes results. */
int ffk_batch_set_class_tuning(ffk_batch* b, int class_id, int block_size, int refill, int pool, int walk);
int ffk_batch_get_class_tuning(const ffk_batch* b, int class_id, int* block_size, int* refill, int* pool,
                               int* walk);
/* Tuning probe: the persistent launch of one class over the first tiles of
 * its queue holding at least max_rays rays (<= 0: all), with the buffers of
 * the last ffk_batch_run. It leaves partial sums in F and the tallies: the
 * next ffk_batch_run recomputes them. */
int ffk_batch_run_class(ffk_batch* b, int class_id, long long max_rays, void* stream);
/* Per-class autotuning of the batch: one untuned ffk_batch_run with these
 * buffers (its results are valid), then for each class the walk mode (3
 * against 0, 1, 2 and 7), the ray scheduling (refill/pool 32/2, 24/3, 16/4, 40/3, 40/4) and the block size (256,
 * 512, 1024 where the histogram allows), timed greedily with
 * ffk_batch_run_class on about sample_rays of the class; a setting replaces
 * the defaults only when it is 3 % faster there and then 1 % faster over the
 * whole class. The winners are set with ffk_batch_set_class_tuning. Then,
 * batch-wide, the tile size (rays_per_tile 32768, 16384, 8192 against the
 * current one), timed on whole runs and kept when 1 % faster; only sizes that
 * leave every scene's tally flag as it is are tried, so the caller's buffers
 * stay valid. (The persistent grid is not tuned: it is the blocks the device
 * holds at once; more only find the tile queue empty, fewer leave CUs idle.)
 * Results are cached per scene set (the scenes' signatures and the device)
 * in memory and in the tuning cache file (ffk_tune_cache_path). Returns 0
 * when tuned, 1 when taken from a cache. */
int ffk_batch_autotune(ffk_batch* b, const unsigned long long* seeds, unsigned long long* const* d_tallies,
                       float* const* d_Fs, long long sample_rays, void* stream);
/* h_flags[i] = 1 when scene i needs a tally buffer in ffk_batch_run, else 0. */
int ffk_batch_tally_flags(const ffk_batch* b, int* h_flags);

/* ---- autotuning ------------------------------------------------------------------
 * Picks the grid resolution (a ladder of 0.25 .. 16 cells per face, each
 * timed by casting the scene's own Malley rays with the emission layout
 * scaled to about sample_rays; when a full run is estimated under 250 ms and
 * F fits in half the free memory, the best candidates are re-timed with the
 * fused kernel on the full layout), the fused kernel's block size (256, 512
 * or 1024), the walk mode (ffk_set_walk, all eight; another than the default
 * must be 1 % faster) and, timed the same way, the ray
 * scheduling (ffk_set_refill: 32/2, 24/3, 16/4, 40/3 or 40/4), then leaves the scene built
 * and tuned with them (ffk_get_refill reads the scheduling back). Results are
 * cached per scene signature (faces, meshes, ray counts, frame) for the life
 * of the process. Returns 0 when tuned, 1 when taken from the cache. None of
 * the choices changes hit indices. out_dims3 / out_block / out_walk may be
 * NULL (block 0 = the automatic choice was kept). */
int ffk_autotune(ffk_scene* s, long long sample_rays, int* out_dims3, int* out_block, int* out_walk,
                 void* stream);
/* The tuning cache file shared by ffk_autotune and ffk_batch_autotune: the
 * path in FFK_TUNE_CACHE if set ("" or "off": no file), else
 * $XDG_CACHE_HOME/form_factors_amd/autotune.txt or
 * $HOME/.cache/form_factors_amd/autotune.txt. Text, one result per line,
 * appended as results are found; the last line for a key wins. Writes the
 * path (empty when there is no file) into buf; returns its length. */
int ffk_tune_cache_path(char* buf, int capacity);
/* Forgets the in-memory results, so that the next autotune call reads the
 * file again. */
void ffk_tune_cache_reset(void);

/* ---- radiance equation --------------------------------------------------------
 * radiance_equation::calculate (thermal_equation/radiance_cpu.cpp:132-190) on
 * the device, over the scene's current emission layout and grid. The caller
 * sets the layout from the radiance weights T * eps * A per face side
 * (radiance_cpu.cpp:96-122, e.g. with ffk_emitted_counts and
 * ffk_scene_set_emission) and passes h_ray_power[2f], h_ray_power[2f+1] = the
 * power of one front / rear ray of face f (front_emission / front_rays).
 * Malley rays come from Philox keyed by seed (the stream of ffk_emit_malley);
 * ray g is attributed to the face and side of the reference's mesh-order walk
 * at position g, and its power is added to the mesh of the face it hits.
 * h_absorption (double[n_meshes], overwritten) is the exact sum of the ray
 * powers, accumulated in 64-bit fixed point (order-independent). d_stats as
 * in ffk_cast. */
int ffk_radiance(ffk_scene* s, unsigned long long seed, const float* h_ray_power, double* h_absorption,
                 unsigned long long* d_stats, void* stream);
/* Host helpers with the reference's float arithmetic. Weights
 * (radiance_cpu.cpp:96-122): w[2f] = T_m * eps_front_m * area_f,
 * w[2f+1] = T_m * eps_rear_m * area_f for f in mesh m (0 outside every mesh),
 * total accumulated in mesh order. Ray powers and emission
 * (radiance_cpu.cpp:151-178): power[2f] = front_emission / front_rays (0 when
 * the face emits no front rays), emission[m] += front_emission +
 * rear_emission (added to the caller's values). h_eps2 holds the front and
 * rear emissivity of each mesh. */
int ffk_radiance_weights(const float* h_faces, int n_faces, const int* h_meshes, int n_meshes,
                         const float* h_eps2, const float* h_T, float* h_weights, float* total);
int ffk_radiance_powers(const float* h_faces, int n_faces, const int* h_meshes, int n_meshes,
                        const float* h_eps2, const float* h_T, const int* h_front, const int* h_rear,
                        float* h_ray_power, float* h_emission);
/* One whole radiance_equation::calculate step (radiance_cpu.cpp:132-190) on
 * the scene: the weights and their total, the emitted ray counts
 * (ffk_emitted_counts with n_rays), the emission layout in `frame` (set again
 * only when the counts or the frame change), the ray powers and emission, and
 * the device pass of ffk_radiance. The same arithmetic as the calls above, with
 * the face areas computed once per scene. h_emission (float[n_meshes]) is
 * added to, h_absorption (double[n_meshes]) overwritten. The scene's grid must
 * be built. */
int ffk_radiance_step(ffk_scene* s, const float* h_eps2, const float* h_T, int n_rays, int frame,
                      unsigned long long seed, float* h_emission, double* h_absorption,
                      unsigned long long* d_stats, void* stream);

/* ---- device-resident step equations ------------------------------------------
 * The per-step equations of a thermal run (radiance_cpu.cpp:132-190,
 * conductive_cpu.cpp:83-141, heat_source_cpu.cpp:48-61,
 * parallel_rays_cpu.cpp:92-131, cpu_adams.cpp:107-182) for a caller that keeps
 * the step's float emission and absorption (thermal_equation::task_t's
 * arrays) in device memory. Each adds its terms to d_emission / d_absorption
 * (float[n_meshes]) with the float arithmetic and in the order of the
 * host-array forms above, so the sums are theirs bit for bit; nothing is
 * copied back and the stream is not synchronised. No counterpart in the
 * reference (whose equations take host arrays): they let a thermal solution
 * run a step's equations and its Adams update without a host round trip.
 *
 * ffk_radiance_step_dev: ffk_radiance_step with device outputs. The weights,
 * their float total in mesh order and the emitted counts stay on the host
 * (h_T is the step's float temperature on the host); the emission, the pass
 * and the absorption (the exact double sum rounded to float) run on the
 * device. ffk_radiance_status, called once the stream has drained, reports
 * FFK_ERR_ARG when a walked ray power was negative or not finite (the outputs
 * are then not meaningful). */
int ffk_radiance_step_dev(ffk_scene* s, const float* h_eps2, const float* h_T, int n_rays, int frame,
                          unsigned long long seed, float* d_emission, float* d_absorption, void* stream);
int ffk_radiance_status(ffk_scene* s);
/* ffk_sun and ffk_sun_deposit on device arrays. ffk_sun_count_dev overwrites
 * d_counts (uint64[2 n_meshes], caller scratch) with the hit counts per mesh
 * side; the counts depend on the scene, the aperture and the seed only, so a
 * caller may count on another stream while the step's other equations run.
 * ffk_sun_deposit_dev then adds (float)(counts * (absorbance * ray_power)),
 * the double sum from zero, to d_absorption, with d_absorb2 (float[2
 * n_meshes], front and rear absorbance) on the device. */
int ffk_sun_count_dev(ffk_scene* s, const float* h_dir3, const float* h_center3, float distance, float radius,
                      int aperture, long long n_rays, unsigned long long seed, unsigned long long* d_counts,
                      void* stream);
int ffk_sun_deposit_dev(int n_meshes, const unsigned long long* d_counts, const float* d_absorb2, float ray_power,
                        float* d_absorption, void* stream);
/* ffk_conduction from the step's float temperatures: the per-mesh double sums
 * (from zero) rounded to float and added to d_emission / d_absorption. */
int ffk_conduction_dev(int n, const int* d_row_ptr, const int* d_cols, const double* d_G, const float* d_T,
                       float* d_emission, float* d_absorption, void* stream);
/* ffk_sbff from the step's float temperatures d_T: promoted into d_T64 (n
 * doubles of caller scratch), the exchange into d_em64 / d_ab64 (n doubles
 * each, zeroed here), and those double sums rounded to float and added to
 * d_emission / d_absorption. */
int ffk_sbff_dev(int n, const void* d_S, int dtype, const double* d_row_scale, const float* d_T,
                 const double* d_area, const double* d_eps_front, const double* d_eps_rear, double* d_T64,
                 double* d_em64, double* d_ab64, float* d_work, float* d_emission, float* d_absorption,
                 void* stream);
/* Heat sources in list order: power > 0 is added to d_absorption[mesh], else
 * -power to d_emission[mesh] (d_mesh, d_power: device arrays of n_sources). */
int ffk_heat_sources_dev(int n_sources, const int* d_mesh, const float* d_power, float* d_emission,
                         float* d_absorption, void* stream);
/* Starts a step: d_T_f = (float)d_T_n, d_emission = d_absorption = 0. */
int ffk_step_begin_dev(int n, const double* d_T_n, float* d_T_f, float* d_emission, float* d_absorption,
                       void* stream);
/* ffk_adams_step on the step's float emission and absorption (each promoted
 * to double, as the host form's caller does). */
int ffk_adams_step_f(int n, int n_step, double dt, const double* d_C, const float* d_emission,
                     const float* d_absorption, double* d_E_hist, const double* d_T_n, double* d_T_next,
                     double* d_balance, void* stream);

/* ---- distant source (sun) absorption ------------------------------------------
 * parallel_rays_cpu::calculate (thermal_equation/parallel_rays_cpu.cpp:92-131).
 * ffk_bsphere is math::triangles_bsphere (math/triangle.cpp:361-383, whose
 * radius loop skips triangle 0). The caller places the aperture as the
 * reference does: radius = 1.1 * bsphere radius, distance = 1.1 + radius,
 * centre = bsphere centre. ffk_sun casts n_rays parallel rays (aperture 0 =
 * disk of `radius`, the reference; 1 = square of half side `radius`; Philox
 * counter (i, 0, 1, i >> 32) keyed by seed, as ffk_emit_parallel) on the
 * scene's grid and counts hits per mesh side: h_counts[2m] front
 * (dot(normal, dir) < 0), h_counts[2m+1] rear. ffk_sun_deposit adds
 * counts * (absorbance * ray_power) to h_absorption (double), the exact sum
 * of the reference's per-ray float values; ray_power = power * (pi r^2) /
 * n_rays in float. */
int ffk_bsphere(const float* h_faces, int n_faces, float* center3, float* radius);
int ffk_sun(ffk_scene* s, const float* h_dir3, const float* h_center3, float distance, float radius,
            int aperture, long long n_rays, unsigned long long seed, unsigned long long* h_counts,
            unsigned long long* d_stats, void* stream);
int ffk_sun_deposit(int n_meshes, const unsigned long long* h_counts, const float* h_absorb2, float ray_power,
                    double* h_absorption);

/* ---- thermal ---------------------------------------------------------------- */
/* Elementwise conversion of F (float32) to the storage type
 * (FFK_BF16/FFK_F16/FFK_F32), round to nearest even. */
int ffk_convert_f(const float* d_F, long long n_elems, int dtype, void* d_out, void* stream);
/* Exchange storage of an N x N form-factor matrix: S[m, :] = dtype(F[m, :] * mul_m)
 * and d_row_scale[m] = 1 / mul_m (device, double[N]); the exchange then uses
 * q_j = sum_m S[m, j] * (E_m * scale_m), which equals sum_m F[m, j] E_m.
 * d_out is TRANSPOSED: d_out[j*N + m] = S[m, j], one contiguous row per
 * target mesh (the layout ffk_sbff and ffk_thermal_set_exchange read).
 * mul_m = d_row_mul[m] when given (e.g. the rays emitted by mesh m, which makes
 * S the integer hit counts: exact in BF16 up to 256 and in FP16 up to 2048).
 * With d_row_mul == NULL a count-valued row (every entry an integer multiple
 * of its smallest nonzero entry 1/n) gets mul_m = n; other rows mul_m = 1.
 * FP16 rows whose scaled maximum would reach 32768 use instead the power of
 * two that puts the row maximum in [16384, 32768). FP32 keeps mul_m = 1. */
int ffk_pack_exchange(const float* d_F, int n, int dtype, const double* d_row_mul, void* d_out,
                      double* d_row_scale, void* stream);
/* Radiative exchange of thermal_equation/form_factors.cpp:109-128:
 *   E_m = (eps_f_m + eps_r_m) * sigma * T_m^4 * A_m        (emission[m] += E_m)
 *   absorption[n] += sum_m E_m * F[m*N + n]
 * d_S is the packed storage of ffk_pack_exchange (d_row_scale from the same
 * call), accumulated in FP32; T/area/eps/outputs in FP64. d_work: n floats
 * of device scratch (the scaled E in FP32), owned by the caller. */
int ffk_sbff(int n, const void* d_S, int dtype, const double* d_row_scale, const double* d_T,
             const double* d_area, const double* d_eps_front, const double* d_eps_rear,
             double* d_emission, double* d_absorption, float* d_work, void* stream);
/* ffk_sbff without d_work, allocating nothing. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * scaled-emission scratch is a static device array of FFK_NOALLOC_MAX_MESHES
 * floats (4 MiB), held by each device for the life of the process and shared
 * by every call on that device: calls that overlap on one device race, so
 * order them (one stream). n above the limit returns FFK_ERR_LIMIT before
 * any launch. The address of the array is looked up on the first call made
 * on a device, with that device current; call it with the stream's device
 * current. Results equal ffk_sbff's bit for bit. */
#define FFK_NOALLOC_MAX_MESHES (1 << 20)
int ffk_sbff_noalloc(int n, const void* d_S, int dtype, const double* d_row_scale, const double* d_T,
                     const double* d_area, const double* d_eps_front, const double* d_eps_rear,
                     double* d_emission, double* d_absorption, void* stream);
/* Conduction (thermal_equation/conductive_cpu.cpp:83-134) as a CSR operator
 * over mesh adjacency. For each neighbour k of mesh m, f = G_mk (T_k - T_m):
 * f > 0 is added to absorption[m], else -f to emission[m] (the reference's
 * per-pair split). */
int ffk_conduction(int n, const int* d_row_ptr, const int* d_cols, const double* d_G,
                   const double* d_T, double* d_emission, double* d_absorption, void* stream);
/* Adams-Bashforth step of thermal_solution/cpu_adams.cpp:107-182 in FP64:
 *   E_n[m] = (absorption - emission) / C_m; C_m = A_m * thickness * rho * c
 *   T_next = T_n + dt * sum_{l<=k} u[k][l] E_{n-l},  k = min(n_step, 4).
 * d_E_hist is a ring of 5 rows of n doubles (row = step % 5); d_balance
 * (double[2], may be NULL) receives sum(absorption - emission) and
 * sum(C * (T_next - T_n)) / dt for the energy-balance check. */
int ffk_adams_step(int n, int n_step, double dt, const double* d_C, const double* d_emission,
                   const double* d_absorption, double* d_E_hist, const double* d_T_n,
                   double* d_T_next, double* d_balance, void* stream);

/* ---- mesh adjacency for conduction (host, set-up time) ----------------------
 * thermal_equation/conductive_cpu.cpp:83-141 over the face graph of
 * subject/mesh_graph.cpp:116-279: every pair of faces of different meshes
 * that share an edge adds G = 1/R to the symmetric mesh operator, with
 * R = (d_l/k_l + d_r/k_r) / (min(thickness) * edge length) evaluated in float
 * as the reference does. h_meshes are (first_idx, n_faces, pad) triples;
 * conductivity and thickness are per mesh. Call with h_row_ptr == NULL to get
 * *nnz_out, then with buffers of that capacity (CSR sorted by row, column). */
int ffk_conduction_graph(const float* h_faces, int n_faces, const int* h_meshes, int n_meshes,
                         const float* h_conductivity, const float* h_thickness, int* h_row_ptr,
                         int* h_cols, double* h_G, int capacity, int* nnz_out);

/* ---- device-resident transient run -------------------------------------------
 * thermal_solution::calculate (cpu_adams.cpp:147-182) over the SB_FF equation
 * (F on the device in FFK_F32/BF16/F16, FP32 exchange accumulation), CSR
 * conduction and fixed heat sources, all on the device with the temperature
 * update in FP64. One step is one kernel: a wavefront per target mesh
 * streams its row of the packed (transposed) exchange storage against E in
 * LDS, then applies emission, conduction, sources and the Adams update into a
 * second temperature buffer. Steps are replayed from a hipGraph of 10 steps
 * when use_graph != 0. h_net_stored (double[2*n_steps], may be NULL) receives per
 * step the net power sum(absorption - emission) and the stored power
 * sum(C (T_next - T_n) / dt). h_T0 == NULL continues the previous run. */
typedef struct ffk_thermal ffk_thermal;
int ffk_thermal_create(ffk_thermal** out, int n_meshes, const double* h_C, const double* h_area,
                       const double* h_eps_front, const double* h_eps_rear);
void ffk_thermal_destroy(ffk_thermal* t);
/* d_S: packed exchange storage from ffk_pack_exchange (device, N*N of dtype,
 * stays owned by the caller; NULL = no radiative exchange); d_row_scale
 * (device, double[N], may be NULL = 1) is copied. */
int ffk_thermal_set_exchange(ffk_thermal* t, const void* d_S, int dtype, const double* d_row_scale);
/* Step-kernel launch shape: 16-byte loads in flight per lane (2, 4 or 8;
 * 0 = 4) and an upper bound on the grid (0 = every block that fits). */
int ffk_thermal_set_tuning(ffk_thermal* t, int unroll, int max_blocks);
int ffk_thermal_set_conduction(ffk_thermal* t, const int* h_row_ptr, const int* h_cols, const double* h_G);
int ffk_thermal_set_sources(ffk_thermal* t, int n_sources, const int* h_mesh, const double* h_power);
int ffk_thermal_run(ffk_thermal* t, const double* h_T0, int n_steps, double dt, double* h_T_out,
                    double* h_net_stored, int use_graph, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FFK_H */
