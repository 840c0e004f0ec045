/*
 * CPU oracle for the Monte Carlo form-factor path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it restates, in
 * plain C++, the reference's CPU algorithm for emission, casting, tallying
 * and the thermal step, so that the HIP path can be compared with it on the
 * same inputs. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it. The product library (libff_amd.so) never links,
 * loads or calls it, and has no CPU fallback.
 *
 * Arithmetic contract: compiled with g++ -O2 -ffp-contract=off and no
 * -march flags, so every float operation is a single IEEE binary32 SSE op in
 * the order the reference writes it. Comparisons against 0.0, 1.0 and 1e-8
 * are done in double, as in the reference (math/operations.h:31).
 *
 * Layouts (all row-major, contiguous):
 *   face  = float[9]  : p0.xyz p1.xyz p2.xyz      (math/types.h:51-54)
 *   ray   = float[6]  : origin.xyz endpoint.xyz   (math/types.h:72-76)
 *   mesh  = int[3]    : first_idx n_faces pad      (form_factors/system.h:33-38)
 *   mat33 = float[9]  : p[row][col]                (math/types.h:79-82)
 *
 * Pinning (see tests/test_oracle.py): every function is checked against
 * golden vectors produced from the compiled reference (tests/golden/), the
 * reference's own known-answer tests, and closed forms. Functions without
 * such a check are marked "parity unpinned" below.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- math (math/triangle.cpp, math/mat.cpp) ---------------------------- */
int   or_triangle_intersect(const float* ray6, const float* tri9, float* point3);
float or_triangle_area(const float* tri9);
void  or_rotate_towards(const float* from3, const float* to3, float* m9);
void  or_mat_vec(const float* m9, const float* v3, float* out3);

/* ---- counter RNG and trig used by the HIP emission (own design) -------- */
void  or_philox4x32_10(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4);
void  or_sincos_turns(float u, float* s, float* c);

/* ---- casters (ray_caster/naive_cpu.cpp, ray_caster/zgrid_cpu.cpp) ------ */
/* hit_idx[i] = face index or -1; hit_pts written only on a hit.
 * n_threads > 1 splits the rays into contiguous slices (one per thread). */
void or_naive_cast(const float* faces, int n_faces, const float* rays, long n_rays,
                   int* hit_idx, float* hit_pts, int n_threads);

typedef struct or_zgrid or_zgrid;
or_zgrid* or_zgrid_build(const float* faces, int n_faces);
void or_zgrid_free(or_zgrid* g);
void or_zgrid_info(const or_zgrid* g, int* nx, int* ny, float* base3, float* side3,
                   float* size3, int* n_refs);
void or_zgrid_cells(const or_zgrid* g, int* counts, int* offsets, int* refs);
void or_zgrid_cast(const or_zgrid* g, const float* rays, long n_rays, int* hit_idx,
                   float* hit_pts, int n_threads);

/* 2-D grid primitives on an explicit grid (math/grid.cpp), for the grid
 * known-answer tests. Cells are written as (x, y) pairs in visit order;
 * returns the number of cells (capped at max_cells). */
int or_grid_traverse(const float* base3, const float* size3, int nx, int ny,
                     const float* ray6, int is_segment, int* cells, int max_cells);
int or_grid_rasterize(const float* base3, const float* size3, int nx, int ny,
                      const float* tri9, int* cells, int max_cells);
int or_ray_intersect_segment(const float* p6, const float* q6, float* s);

/* ---- emission (emission/malley_emission.cpp, emission/malley_cpu.cpp) -- */
#define OR_RNG_MT19937 0   /* four std::mt19937 seeded 1..4, as the reference */
#define OR_RNG_PHILOX  1   /* Philox4x32-10 keyed by seed, counter (j, face, stream) */
#define OR_FRAME_REFERENCE   0  /* rotate_towards(z, unnormalised normal) */
#define OR_FRAME_ORTHONORMAL 1  /* rotate_towards(z, unit normal) */

void or_emitted_counts(const float* weights, int n_faces, int n_rays, float total,
                       int* front, int* rear);
/* Two-sided area weights and their float total, accumulated in face order
 * (cpuFactorsCalculator/cpu_system.cpp:122-129). Returns the total. */
float or_area_weights(const float* faces, int n_faces, float* weights);
long or_malley_count(const float* weights, int n_faces, int n_rays, float total);
long or_malley_emit(const float* faces, int n_faces, const float* weights, float total,
                    int n_rays, int rng, uint64_t seed, int frame, float* rays_out);

/* parallel "sun" rays (emission/parallel_rays_cpu.cpp:88-106) */
void or_parallel_emit(const float* dir3, const float* center3, float distance, float radius,
                      int n_rays, int rng, uint64_t seed, float* rays_out);

/* ---- form factors (cpuFactorsCalculator/cpu_system.cpp) ---------------- */
/* Tally in mesh order with float counts and per-row normalisation. */
int or_ff_tally(const int* meshes, int n_meshes, int n_faces, const int* front,
                const int* rear, long n_total, const int* hit_idx, float* F);
/* Whole pipeline: weights -> emission -> cast (0 naive, 1 zgrid) -> tally.
 * Returns 0 or the reference's negative error code. */
int or_ff_calculate(const float* faces, int n_faces, const int* meshes, int n_meshes,
                    int n_rays, int rng, uint64_t seed, int frame, int caster,
                    int n_threads, float* F);

/* ---- thermal (thermal_equation/*.cpp, thermal_solution/cpu_adams.cpp) -- */
/* q_n += sum_m E_m F[m*N+n] with E_m = (eps_f+eps_r) sigma T^4 A_m, in double. */
void or_parallel_emit2(const float* dir3, const float* center3, float distance, float radius,
                       int aperture, int n_rays, int rng, uint64_t seed, float* out);
void or_bsphere(const float* faces, int n_faces, float* center3, float* radius);
void or_sun(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* absorb2,
            float power, const float* dir3, int n_rays, int rng, uint64_t seed, int aperture, int caster,
            int n_threads, float* absorption, double* absorption64);
float or_radiance_weights(const float* faces, int n_faces, const int* meshes, int n_meshes,
                          const float* eps2, const float* T, float* weights);
void or_radiance_walk(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* eps2,
                      const float* T, const int* front, const int* rear, long n_total, const int* hit_idx,
                      float* emission, float* absorption, double* absorption64);
int or_radiance(const float* faces, int n_faces, const int* meshes, int n_meshes, const float* eps2,
                const float* T, int n_rays, int rng, uint64_t seed, int frame, int caster, int n_threads,
                float* emission, float* absorption, double* absorption64);
void or_sbff_f64(int n, const double* F, const double* T, const double* area,
                 const double* eps_front, const double* eps_rear,
                 double* emission, double* absorption);
/* One Adams-Bashforth step in double with the reference's u-matrix:
 * T_next = T_n + dt * sum_{l<=k} u[k][l] * E_{n-l}, k = min(n_step, 4).
 * E_hist holds 5 rows of n values in ring order (row = step % 5). */
void or_adams_combine_f64(int n, int n_step, double dt, const double* T_n,
                          const double* E_hist, double* T_next);

#ifdef __cplusplus
}
#endif
