/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * A serial FP64 CPU restatement of the reference nabla_solve hot path, used
 * as the checker for the HIP implementation.  Only tests/, the smoke() entry
 * and bench.py's CPU-baseline leg load it; the product never links it.
 *
 * Every function follows the reference loop structure and arithmetic order so
 * that, built with -ffp-contract=off, it reproduces the reference bit for bit
 * (this is pinned against the reference compiled from its own sources, see
 * oracle/ref_shim.cpp, and against the committed golden fixtures).
 *
 * Reference files restated (paths relative to the reference repository):
 *   core/src/flow/ns_solver.cpp, core/src/flow/poisson.cpp,
 *   core/src/flow/multigrid.cpp, core/src/flow/octree_poisson.cpp,
 *   core/src/mesh/octree.cpp, core/include/nabla/mesh/morton.hpp.
 */
#ifndef NBL_ORACLE_H
#define NBL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "nbl_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MAC Navier-Stokes (ns_solver.cpp). */
void* orc_ns_create(const nbl_grid_desc* g, const nbl_flow_cfg* cfg);
void orc_ns_destroy(void* h);
void orc_ns_initialize(void* h);
int orc_ns_set_state(void* h, const double* u, size_t nu, const double* v, size_t nv,
                     const double* w, size_t nw, const double* p, size_t np, int step, double t,
                     double dt_last);
void orc_ns_get_state(void* h, double* u, double* v, double* w, double* p, int* step, double* t,
                      double* dt_last);
double orc_ns_compute_dt(void* h);
void orc_ns_step(void* h, double dt, nbl_step_report* rep);
void orc_ns_rhs(void* h, double* ru, double* rv, double* rw);
void orc_ns_cell_velocity(void* h, int comp, double* out);

/* Uniform pressure Poisson (poisson.cpp, multigrid.cpp). */
void orc_apply_a(const nbl_grid_desc* g, const double* x, double* out);
void orc_diagonal(const nbl_grid_desc* g, double* diag);
void orc_mg_solve(const nbl_grid_desc* g, const double* b, double* x, double tol, int max_cycles,
                  nbl_poisson_report* rep);
void orc_cg_solve(const nbl_grid_desc* g, const double* b, double* x, double tol, int max_iters,
                  nbl_poisson_report* rep);
/* Multigrid pieces on levels 0/1 (no null-space projection). */
int orc_mg_levels(const nbl_grid_desc* g, int* dims);
void orc_mg_smooth_once(const nbl_grid_desc* g, const double* b, double* x);
void orc_mg_restrict(const nbl_grid_desc* g, const double* rf, double* rc);
void orc_mg_prolong(const nbl_grid_desc* g, const double* ec, double* xf);

/* Octree (octree.cpp) and OctreePoisson (octree_poisson.cpp). */
uint64_t orc_morton_encode(uint32_t x, uint32_t y, uint32_t z);
void orc_morton_decode(uint64_t key, uint32_t* x, uint32_t* y, uint32_t* z);
void* orc_oct_create(const double origin[3], const double extent[3], int max_level);
void orc_oct_destroy(void* h);
size_t orc_oct_count(void* h);
void orc_oct_refine_uniform(void* h, int times);
void orc_oct_refine_box(void* h, const double lo[3], const double hi[3], int times);
int orc_oct_refine_key(void* h, uint64_t key);
int orc_oct_level_of(void* h, uint64_t key);
void orc_oct_set_mask(void* h, size_t i, uint8_t m);
void orc_oct_dump(void* h, uint64_t* keys, uint8_t* levels, uint8_t* masks);
int orc_oct_face_neighbors(void* h, size_t i, int dir, uint64_t* out);
// cnt[6i+d]: neighbours of leaf i across face d; nbr[24i+4d+m]: their indices (-1 unused)
void orc_oct_face_table(void* h, uint8_t* cnt, int64_t* nbr);
int orc_oct_is_balanced(void* h);
void orc_oct_centers(void* h, double* xyz);
void orc_oct_apply(void* h, const double* x, double* out);
void orc_oct_diagonal(void* h, double* diag);
void orc_oct_solve(void* h, const double* b, double* x, double tol, int max_iters,
                   nbl_poisson_report* rep);
/* Octree multigrid-preconditioned flexible CG (no reference counterpart: the
 * restatement of the device design, pinned by its reduction to the uniform
 * multigrid on uniform trees and by the OctreePoisson system it solves). */
void orc_oct_mg_solve(void* h, const double* b, double* x, double tol, int max_iters,
                      nbl_poisson_report* rep);
int orc_oct_mg_levels(void* h, long long* counts, int cap);
/* octreeTransportStep (adaptive/transport.cpp:24-60) on scalar phi (in/out),
 * cell-centred u, v, w; returns the integral sum phi_i vol_i. */
double orc_oct_transport(void* h, const double* u, const double* v, const double* w, double* phi, double dt);
// writeGridVtu (case.cpp:175-241) of the solver's state; 0 or -1 (cannot open).
int orc_ns_write_vtu(void* h, const char* path);
// PhysicalSensorEvaluator::evaluate (adaptive_core.cpp:73-176): out[8*n] field-major, flags[3*n].
void orc_oct_sensors(void* h, const double* u, const double* v, const double* w, const double* p, double nu,
                     double* speedPrev, double* residPrev, double* out, uint8_t* flags);

#ifdef __cplusplus
}
#endif
#endif
