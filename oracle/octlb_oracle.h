/*
 * octlb_oracle.h -- CPU restatement of the reference's per-time-step AMR
 * D3Q19 LBM update (gezengran/octlb, TimeLoop::advance_one and everything it
 * calls).  TEST INFRASTRUCTURE ONLY: tests, the smoke check and the bench's
 * CPU-baseline leg use it as the checker; the product never links it.
 *
 * Storage mirrors the reference BlockLattice exactly: per block an AoS array
 * of (n+2h)^3 halo-inclusive slots x 19 shifted FP64 populations
 * (block_lattice.h:159-163, 294-304), so block dumps compare 1:1 with the
 * reference harness in oracle/_ref and with the GPU path's downloads.
 */
#ifndef OCTLB_ORACLE_H_
#define OCTLB_ORACLE_H_

#include <stdint.h>

#include "../include/octlb_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ol_sim ol_sim;

ol_sim* ol_create(const octlb_topology* topo, const octlb_params* params);
void ol_destroy(ol_sim* sim);
const char* ol_last_error(void);

int ol_set_bc_kinds(ol_sim* sim, const uint8_t* kinds);
int ol_set_links(ol_sim* sim, int64_t count, const int32_t* link, const double* q);
int ol_init_equilibrium(ol_sim* sim, double rho0, const double u0[3], uint32_t kind_mask);
int ol_init_field(ol_sim* sim, const double* rho, const double* u);
int ol_set_block(ol_sim* sim, int32_t block, const double* aos);
int ol_get_block(const ol_sim* sim, int32_t block, double* aos);
int ol_get_core(const ol_sim* sim, double* out);
int ol_get_snapshot_block(const ol_sim* sim, int32_t block, double* aos);

int ol_advance(ol_sim* sim, int32_t n_coarse_steps);
int ol_set_threads(ol_sim* sim, int32_t nthreads);
int ol_set_step(ol_sim* sim, int32_t step);

int ol_drag(const ol_sim* sim, const double bbox[6], double out[3]);
double ol_mass(const ol_sim* sim);
double ol_average_rho(const ol_sim* sim);

/* TimeLoopCounters (time_loop.h:19-34): per-level collide/stream counts and
 * the coupler call log as (phase, coarse_level) pairs. */
int ol_counters(const ol_sim* sim, int32_t* collide, int32_t* stream, int32_t max_levels,
                int32_t* calls, int32_t max_calls, int32_t* num_calls);

/* Coupling plan (level_coupler.cpp:70-149): rows of 12 ints
 * coarse_id,ci,cj,ck,fine_id,fi,fj,fk,normal,coarse_level,fine_level,fine_slot. */
int32_t ol_plan_size(const ol_sim* sim);
int ol_plan(const ol_sim* sim, int32_t* rows);

/* Same-rank edge exchange list (ghost_schedule.h:56-91): rows local,d1,d2,remote. */
int32_t ol_num_edges(const ol_sim* sim);
int ol_edges(const ol_sim* sim, int32_t* rows);

/* Single-block primitives used by the unit checks against the reference
 * harness: one phase of the loop on one block. */
int ol_block_collide(ol_sim* sim, int32_t block, double t);
int ol_block_stream(ol_sim* sim, int32_t block);
int ol_block_post_stream(ol_sim* sim, int32_t block, double t);

/* Coupler phases on the current state, for the unit checks of the Lagrava
 * coupling: capture prev from the coarse cells (LevelCoupler ctor), one
 * prolongation of the plan points whose coarse level is `level` (half_time
 * 1 = apply_half_time, 0 = apply_full_time), one restriction onto the
 * coarse cells of the points whose fine level is `fine_level`. */
int ol_coupler_init(ol_sim* sim);
int ol_apply_prolongation(ol_sim* sim, int32_t level, int32_t half_time);
int ol_apply_restriction(ol_sim* sim, int32_t fine_level);

/* BouzidiLinkData::Build (bouzidi_link_data.cpp:139-233): Bouzidi wall
 * fractions of every fluid cell whose in-block neighbour x + c_opp(i) is
 * solid-like, from the nearest point of a triangle soup (triangle edges
 * only).  bounds: 6 per block; material: MaterialKind bytes per block in
 * MaterialField order (k*n + j)*n + i; triangles: v0, v1, v2 (9 doubles)
 * each.  Links come out in (block, (i*n + j)*n + k, pop) order, the dense
 * q_frac order.  Returns the number of links; writes at most `capacity`. */
int64_t ol_build_links(int32_t num_blocks, int32_t n, const double* bounds, const uint8_t* material,
                       int64_t num_triangles, const double* triangles, int64_t capacity, int32_t* links,
                       double* q);

/* Discrete adjoint of ol_advance for one level (concrete handler, plain
 * BGK, Bouzidi streaming, snapshots on, no FD boundary kinds).  Runs
 * `steps` steps from feq(rho0, u0) on the cells whose kind is in kind_mask
 * (all other slots zero) and returns J = mean over steps win_begin+1 ..
 * win_end of dir . drag(bbox), dJ/du0 and dJ/dq of each listed link
 * (block, cell, pop).  Not part of the reference: it checks the device
 * adjoint and is pinned by central differences of ol_advance. */
int ol_adjoint(ol_sim* sim, int32_t steps, int32_t win_begin, int32_t win_end, const double dir[3],
               const double bbox[6], double rho0, const double u0[3], uint32_t kind_mask, int64_t nlinks,
               const int32_t* link, double* J, double dJ_du0[3], double* dJ_dq);

/* State of one block of an adapted forest from the old forest's
 * halo-inclusive views (AoS [(n+2h)^3][19] in slot order
 * ((x+h)(n+2h) + y+h)(n+2h) + z+h).  mode 0: copy views[0]; mode 1:
 * refine, views[0] = the parent, oct = this block's octant in it; mode 2:
 * coarsen, views[c] = the child at octant c = ox + 2 oy + 4 oz.  Writes
 * every slot of out.  Not part of the reference (its dynamic-AMR rebuild
 * is a plan, doc/prd/octlb-framework.md:180,521): the rules apply the
 * coupler's rescaling (level_coupler.cpp:187, 226-228) to whole blocks. */
int ol_migrate_block(int32_t mode, int32_t n, int32_t h, const int32_t oct[3], const double* const* views,
                     double tau, double* out);

/* voxelize_part (part_voxelizer.cpp:71-114) with validate_geometry_part
 * (:51-69): MaterialKind per cell (MaterialField order (k*n + j)*n + i per
 * block) of one part.  role 0 = kInternalChannel (inner soup required),
 * 1 = kExternalObstacle.  Every geometric test is decided exactly on the
 * double inputs (expansion arithmetic, no floating-point filter), as CGAL's
 * exact predicates decide them: cell box vs closed triangle, and ray
 * parity along +x, +y, +z from the cell centre (CgalSurfaceMesh::is_inside,
 * cgal_surface_mesh.cpp).  Returns 0, or -1 with ol_last_error() set for
 * the reference's GeometryBuildError cases. */
int ol_voxelize_part(int32_t num_blocks, int32_t n, const double* bounds, int64_t num_triangles,
                     const double* triangles, int64_t num_inner, const double* inner, int32_t role,
                     uint8_t* material);

#ifdef __cplusplus
}
#endif

#endif /* OCTLB_ORACLE_H_ */
