/*
 * octlb_gpu.h -- C ABI of the MI355X (gfx950) per-time-step AMR D3Q19
 * lattice-Boltzmann update.
 *
 * Everything here is plain C: integers, doubles and pointers.  A C++ host
 * (the reference's TimeLoop / GhostSchedule / LevelCoupler / BlockLattice
 * call sites), a cgo/ctypes binding or a Python harness can all bind it.
 *
 * Conventions shared with the reference (gezengran/octlb):
 *   - Face numbering  x-,x+,y-,y+,z-,z+ = 0..5      (src/common/types.h:15-22)
 *   - BcKind byte per core cell, index (ix*n+iy)*n+iz (block_lattice.cpp:288-296)
 *   - Populations are SHIFTED (f_i - w_i); rho = 1 + sum(f)  (block_lattice.h:113-125)
 *   - D3Q19 velocity order of OpenLB  (descriptor/definition/common.h:166-179)
 *   - Block population layout on upload/download is the reference's
 *     halo-inclusive AoS: [(hx*NP+hy)*NP+hz][19], NP = n+2*halo
 *     (block_lattice.h:294-304).
 *
 * Every function returns 0 on success or a negative OCTLB_E* code; the
 * message of the last failure on the calling thread is octlb_last_error().
 * One device handle owns one HIP stream; a handle is not thread-safe.
 */
#ifndef OCTLB_GPU_H_
#define OCTLB_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCTLB_Q 19

/* error codes */
#define OCTLB_OK 0
#define OCTLB_EINVAL (-1)  /* bad argument / inconsistent tables          */
#define OCTLB_ENOMEM (-2)  /* host or device allocation failed            */
#define OCTLB_EHIP (-3)    /* HIP runtime error (message has the details) */
#define OCTLB_ESTATE (-4)  /* call not valid in the current state         */
#define OCTLB_EIO (-5)     /* output file could not be written            */

/* FaceDir (reference src/common/types.h:15-22) */
enum { OCTLB_XMIN = 0, OCTLB_XMAX, OCTLB_YMIN, OCTLB_YMAX, OCTLB_ZMIN, OCTLB_ZMAX };

/* BcKind (reference src/solver/lbm/bc_kind.h:14-23) */
enum {
  OCTLB_KIND_BULK = 0,
  OCTLB_KIND_SOLID,
  OCTLB_KIND_BOUNCE_BACK,
  OCTLB_KIND_MOVING_BOUNCE_BACK,
  OCTLB_KIND_BOUZIDI,
  OCTLB_KIND_VELOCITY_DIRICHLET,
  OCTLB_KIND_PRESSURE_DIRICHLET,
  OCTLB_KIND_OUTFLOW
};

/* DomainBcType (reference domain_boundary_handler.h:16-22) */
enum {
  OCTLB_BC_NO_SLIP = 0,
  OCTLB_BC_MOVING_LID,
  OCTLB_BC_INTERP_VELOCITY,
  OCTLB_BC_OUTFLOW,
  OCTLB_BC_INTERP_PRESSURE
};

/* InletVelocityField subclasses (reference boundary/inlet_velocity_field.h) */
enum {
  OCTLB_INLET_NONE = 0,
  OCTLB_INLET_UNIFORM,     /* UniformInletProfile            :131-176 */
  OCTLB_INLET_POISEUILLE,  /* PoiseuilleInletProfile (index) :47-129  */
  OCTLB_INLET_CHANNEL      /* ChannelPoiseuilleInletProfile  :178-270 */
};

/* Which DomainBoundaryHandler the TimeLoop is given (time_loop.h:49-54). */
enum {
  OCTLB_HANDLER_NOOP = 0,     /* NoOpDomainBoundaryHandler (domain_boundary_handler.h:94-99) */
  OCTLB_HANDLER_CONCRETE = 1  /* ConcreteDomainBoundaryHandler (.cpp:31-89)                  */
};

/* One DomainBcSpec (domain_boundary_handler.h:31-42) with its inlet field. */
typedef struct {
  int32_t face;         /* FaceDir the spec applies to                    */
  int32_t type;         /* OCTLB_BC_*                                     */
  double u_wall[3];     /* constant wall velocity                         */
  double rho_target;    /* pressure-outlet density (default 1.0)          */
  int32_t inlet;        /* OCTLB_INLET_*; NONE -> use u_wall              */
  int32_t inlet_face;   /* face_dir of the profile (in-plane axes)        */
  double inlet_u[3];    /* UNIFORM: u_const;  POISEUILLE/CHANNEL: flow_dir */
  double u_peak;        /* POISEUILLE/CHANNEL peak velocity               */
  int32_t n_inplane0;   /* POISEUILLE cells along in-plane axis 0         */
  int32_t n_inplane1;   /* POISEUILLE cells along in-plane axis 1         */
  double lo0, hi0;      /* CHANNEL physical extent, in-plane axis 0       */
  double lo1, hi1;      /* CHANNEL physical extent, in-plane axis 1       */
  double ramp_end_t;    /* smoothstep ramp end time (0 = no ramp)         */
} octlb_bc_spec;

/* One CoarseFineFace (reference src/mesh/topology/face_pair_list.h:22-38). */
typedef struct {
  int32_t coarse_id;
  int32_t fine_ids[4];
  int32_t normal;            /* coarse-side FaceDir pointing at the fine side */
  int32_t coarse_level;
  int32_t fine_level;
  double coarse_bounds[6];   /* x_min,y_min,z_min,x_max,y_max,z_max          */
  double fine_bounds[4][6];
} octlb_cf_face;

/*
 * Flattened forest: what p4est's OctreeForest + FacePairList hand to the
 * solver once per adapt.  All arrays are borrowed for the duration of the
 * call that receives the struct (they are copied).
 */
typedef struct {
  int32_t num_blocks;
  int32_t n;                      /* cells per block edge (even)           */
  int32_t halo;                   /* halo width h of BlockLattice(n,n,n,h) */
  const int32_t* level;           /* [num_blocks]                          */
  const double* bounds;           /* [num_blocks*6]                        */
  int32_t num_same_level;
  const int32_t* same_level;      /* [num_same_level*3] local,dir,remote   */
  int32_t num_coarse_fine;
  const octlb_cf_face* coarse_fine;
  int32_t num_tree_boundary;
  const int32_t* tree_boundary;   /* [num_tree_boundary*2] octant,dir      */
} octlb_topology;

/* Solver configuration: the TimeLoop constructor arguments plus the
 * per-block stream mode and the domain BC specs. */
typedef struct {
  double omega;          /* single relaxation frequency (all levels)           */
  int32_t precision;     /* 64: FP64 populations; 32: FP32 shifted populations */
  int32_t handler;       /* OCTLB_HANDLER_*                                     */
  int32_t const_rho;     /* use_const_rho_bgk (time_loop.h:52)                  */
  int32_t stats_scope;   /* 0 kFluidAndBoundary, 1 kFluidOnly (time_loop.h:39)  */
  int32_t bouzidi;       /* 1: blocks carry Bouzidi links -> per-cell stream    */
  int32_t snapshot;      /* set_snapshot_post_collide (time_loop.h:84)          */
  int32_t num_specs;
  const octlb_bc_spec* specs;
  int32_t fast_stats;    /* ConstRho average: 0 = the reference's sequential
                            cell-order sum (bit-reproducible), 1 = parallel
                            tree reduction (faster, order differs)          */
} octlb_params;

const char* octlb_last_error(void);

/* ------------------------------------------------------------------ */
/* Host forest: p4est-free brick forest with refine / 2:1 face balance */
/* and FacePairList enumeration.  Replaces, for brick domains,          */
/*   OctreeForest(comm, domain, bx,by,bz)   octree_forest.cpp:204-229   */
/*   OctreeForest::refine(bounds pred)      octree_forest.cpp:248-261   */
/*   OctreeForest::balance()                octree_forest.cpp:263-271   */
/*   FacePairList(forest)                   face_pair_list.cpp:480-570  */
/* ------------------------------------------------------------------ */
typedef struct octlb_forest octlb_forest;

int octlb_forest_create(const double domain[6], int32_t bricks_x, int32_t bricks_y,
                        int32_t bricks_z, octlb_forest** out);
void octlb_forest_destroy(octlb_forest* forest);
/* Refine every octant below max_level whose bounds overlap box (closed
 * intervals), recursively, like refine(criterion=overlap(box), max_level). */
int octlb_forest_refine_box(octlb_forest* forest, const double box[6], int32_t max_level);
/* Refine every octant whose mark[id] != 0 by one level (ids of the current
 * octant order). */
int octlb_forest_refine_marked(octlb_forest* forest, const uint8_t* mark, int64_t count);
int octlb_forest_balance(octlb_forest* forest);
int32_t octlb_forest_num_octants(const octlb_forest* forest);
/* Build the face lists and expose them as a topology for blocks of n^3
 * cells with halo h.  The pointers stay valid until the forest changes. */
int octlb_forest_topology(octlb_forest* forest, int32_t n, int32_t halo,
                          octlb_topology* out);

/* ------------------------------------------------------------------ */
/* Device solver                                                       */
/* ------------------------------------------------------------------ */
typedef struct octlb_dev octlb_dev;

/* Allocate device tables and lattices for a topology ("once per adapt").
 * Replaces the construction of BlockCollection + GhostSchedule +
 * LevelCoupler + DomainBoundaryHandler + TimeLoop (time_loop.h:123-164).
 * All populations start at zero (shifted rest state). */
int octlb_dev_create(const octlb_topology* topo, const octlb_params* params,
                     octlb_dev** out);
int octlb_dev_destroy(octlb_dev* dev);

/* BlockLattice::set_bc_kind for all core cells: kinds[num_blocks*n^3]. */
int octlb_dev_set_bc_kinds(octlb_dev* dev, const uint8_t* kinds);

/* Bouzidi wall fractions (BouzidiLinkData::set_q_frac, bouzidi_link_data.h:21).
 * link[3*i] = block, link[3*i+1] = cell (ix*n+iy)*n+iz, link[3*i+2] = iPop.
 * Links not listed keep q = 1.0 (half-way bounce-back). */
int octlb_dev_set_links(octlb_dev* dev, int64_t count, const int32_t* link,
                        const double* q);

/* BlockLattice::initialize / initialize_from_material (lattice_material_init.cpp:11-33):
 * feq(rho0,u0) on core cells whose kind is in kind_mask (bit k = BcKind k);
 * all other slots are left untouched. */
int octlb_dev_init_equilibrium(octlb_dev* dev, double rho0, const double u0[3],
                               uint32_t kind_mask);
/* feq(rho, u) per core cell: rho[num_blocks*n^3], u[num_blocks*n^3*3]. */
int octlb_dev_init_field(octlb_dev* dev, const double* rho, const double* u);

/* Reference AoS halo-inclusive layout, FP64 on the host side. */
int octlb_dev_upload_block(octlb_dev* dev, int32_t block, const double* aos);
int octlb_dev_download_block(octlb_dev* dev, int32_t block, double* aos);
/* Whole lattice, core cells only: out[num_blocks*n^3*19], (block, cell, pop). */
int octlb_dev_download_core(octlb_dev* dev, double* out);

/* TimeLoop::advance_one() n times (time_loop.h:242-246). */
int octlb_dev_advance(octlb_dev* dev, int32_t n_coarse_steps);
/* Same as octlb_dev_advance but replays a captured hipGraph of one coarse step. */
int octlb_dev_advance_graph(octlb_dev* dev, int32_t n_coarse_steps);
/* Captures and instantiates every hipGraph octlb_dev_advance_graph replays
 * (one step and 16 steps, from either level-0 parity) without running them,
 * so no capture or instantiation happens inside a timed or latency-bound
 * loop.  No reference counterpart (the reference has no graphs). */
int octlb_dev_prepare_graphs(octlb_dev* dev);
int octlb_dev_synchronize(octlb_dev* dev);
int octlb_dev_set_step(octlb_dev* dev, int32_t step);

/* TimeLoop::set_snapshot_post_collide (time_loop.h:84): take the post-collide
 * snapshot of the drag links' outgoing populations in every following step
 * (octlb_params.snapshot sets it at create). */
int octlb_dev_set_snapshot(octlb_dev* dev, int32_t on);

/* Update scheme of the fused collide+stream kernel (no reference
 * counterpart; the reference has one CPU loop).  0: every 8^3 block pulls
 * through an LDS box of its cells and one-cell apron, re-colliding the apron
 * cells; 1 (default): blocks whose cells all stream plainly and whose 18
 * halo parts are all exchanged at their level push their post-collide
 * values to the destination cells (no apron), the others pull.  Both give
 * bit-identical states. */
int octlb_dev_set_kernel_variant(octlb_dev* dev, int32_t variant);
/* The variant in use and how many blocks it updates by push. */
int octlb_dev_kernel_variant(octlb_dev* dev, int32_t* variant, int32_t* push_blocks);
/* Autotuner: times the fused kernel of every level under each variant on
 * the current state (left unchanged), weights level l by its 2^l launches
 * per coarse step, and keeps the fastest.  step_ms[2] (optional) receives
 * the weighted kernel time per coarse step of variants 0 and 1 (0 when a
 * variant does not apply).  Then, per kind, it decides the coupling
 * epilogues (below) the same way. */
int octlb_dev_autotune(octlb_dev* dev, int32_t reps, double* step_ms);
/* Coupling epilogues (no counterpart in the reference; same results bit for
 * bit either way): bit 0, the half-time prolongation (LevelCoupler::
 * apply_half_time, level_coupler.cpp:438-478) that follows an update runs in
 * that update's launch, in the workgroups of the blocks holding its coarse
 * cells; bit 1, the restriction (level_coupler.cpp:613-654) after the
 * finest level's second update runs in the workgroups of the blocks holding
 * its fine cells.  Default 3; applies only when no FD block is coupled and
 * the coupled blocks pull (else the separate coupling kernels run).
 * octlb_dev_coupling_epilogues reports the mode in effect and the
 * autotuner's weighted times per coarse step of the affected launches,
 * ms[0] without and ms[1] with its choice. */
int octlb_dev_set_coupling_epilogues(octlb_dev* dev, int32_t mode);
int octlb_dev_coupling_epilogues(octlb_dev* dev, int32_t* mode, double ms[2]);

/* MomentumExchangeDrag::force_on_fluid_if summed over blocks with the
 * obstacle predicate "neighbour cell centre inside bbox"
 * (drag_postprocessor.h:66-115, cylinder3d_case.h:473-503). FP64 sum. */
int octlb_dev_drag(octlb_dev* dev, const double bbox[6], double out[3]);
/* Sum of rho over kBulk cells (cylinder3d_case.h:601-622). FP64 sum. */
int octlb_dev_mass(octlb_dev* dev, double* out);
/* TimeLoop::average_rho() (time_loop.h:88). */
int octlb_dev_average_rho(octlb_dev* dev, double* out);

/* Fine-equivalent and actual cell updates per coarse step. */
int octlb_dev_update_counts(octlb_dev* dev, double* actual, double* fine_equivalent);

/* Timing of the dominant kernel (fused collide+stream) over the last
 * octlb_dev_advance* call, measured with HIP events on the handle's stream:
 * total milliseconds and number of launches. */
int octlb_dev_kernel_timing(octlb_dev* dev, double* ms, int64_t* launches,
                            double* cell_updates);
int octlb_dev_enable_timing(octlb_dev* dev, int32_t on);
/* Mean duration of the fused collide+stream kernel of one level, measured by
 * replaying it reps times back to back between two HIP events on the
 * handle's stream.  The replays read the current state and write only the
 * idle ping-pong buffer (no snapshot, no ConstRho statistics), so the
 * simulation state is unchanged.  cell_updates = cells updated per launch.
 * (No reference counterpart: the measurement hook of the roofline figure.) */
int octlb_dev_time_kernel(octlb_dev* dev, int32_t level, int32_t reps, double* avg_ms,
                          double* cell_updates);

/* ---- Ensembles ---------------------------------------------------------
 * No counterpart in the reference, which runs one case per TimeLoop: many
 * independent cases (members) in one device state, stepped by the same
 * launches and the same hipGraph.  topo holds every member's blocks, member
 * by member: blocks [member_begin[m], member_begin[m+1]) are member m, no
 * block face joins two members, and every member is a single-level forest
 * of 8^3 blocks with halo 1.  params[m] gives member m's relaxation rate and
 * boundary specs; precision, handler, ConstRho, Bouzidi, snapshot and
 * statistics mode must be the same for all members.  Every other call works
 * on the whole ensemble; octlb_dev_average_rho reports member 0. */
int octlb_dev_create_ensemble(const octlb_topology* topo, int32_t num_members, const int32_t* member_begin,
                              const octlb_params* params, octlb_dev** out);
/* ConstRho average of one member (TimeLoop::average_rho of that case). */
int octlb_dev_member_average_rho(octlb_dev* dev, int32_t member, double* out);

/* ---- Material field of a geometry part --------------------------------
 * Replaces voxelize_part(forest, part, config) (part_voxelizer.cpp:71-114)
 * with its checks (validate_geometry_part, :51-69), run on the current HIP
 * device.  The merge of several parts (merge_material_field, :116-146, in
 * priority order, geometry_engine.cpp:28-39) is byte work left to the
 * caller.
 *   bounds, n   as for octlb_build_links (cell_width = n)
 *   triangles   the part's soup, 9 doubles each
 *   inner_*     the inner cavity soup (role INTERNAL_CHANNEL only)
 *   role        OCTLB_PART_INTERNAL_CHANNEL / _EXTERNAL_OBSTACLE
 *               (GeometryPartRole values)
 * Output: material, n^3 MaterialKind bytes per block in MaterialField order
 * (k*n + j)*n + i; *num_exact (optional) = cells whose tests needed the
 * exact-arithmetic fallback.  Inside tests are ray parity along +x, +y, +z
 * (majority of three) and cell-surface tests closed triangle-box
 * intersection, both decided exactly on the double inputs as CGAL's exact
 * predicates decide them.  OCTLB_EINVAL for an empty soup, a channel
 * without inner soup, or a channel whose outer mesh leaves less than 5 % of
 * 5^3 samples outside. */
#define OCTLB_PART_INTERNAL_CHANNEL 0
#define OCTLB_PART_EXTERNAL_OBSTACLE 1
int octlb_voxelize_part(int32_t num_blocks, int32_t n, const double* bounds, int64_t num_triangles,
                        const double* triangles, int64_t num_inner_triangles, const double* inner_triangles,
                        int32_t role, uint8_t* material, int64_t* num_exact);

/* ---- Bouzidi wall fractions from a triangle soup ----------------------
 * Replaces BouzidiLinkData::Build(forest, material, assembly)
 * (bouzidi_link_data.cpp:139-233), run on the current HIP device.
 *   bounds      6 doubles per block (x_min, y_min, z_min, x_max, y_max,
 *               z_max) = OctreeForest::quadrant_bounds
 *   material    MaterialKind bytes (0 fluid, 1 solid, 2 boundary) per block,
 *               n^3 each in MaterialField order (k*n + j)*n + i
 *   triangles   9 doubles each (v0, v1, v2) of the merged soup, parts in
 *               assembly order (MergeAssemblySoup); normals are not used
 * Output: *num_links = number of links, and when links and q are both
 * non-null (capacity >= *num_links, else OCTLB_EINVAL) the links as
 * (block, cell (ix*n + iy)*n + iz, population) int32 triples with their q,
 * in (block, cell, population) order -- the entries of the dense q_frac
 * array that differ from 1.0, ready for octlb_dev_set_links.  Pass null
 * links and q to query the count. */
int octlb_build_links(int32_t num_blocks, int32_t n, const double* bounds, const uint8_t* material,
                      int64_t num_triangles, const double* triangles, int64_t capacity,
                      int64_t* num_links, int32_t* links, double* q);

/* No counterpart in the reference: octlb_build_links for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results,
 * but every array (bounds, material, triangles in; links, q out) is a
 * device pointer the caller owns, the work is queued on `stream` (null: the
 * default stream) and the call allocates nothing: its scratch is a static
 * device array of about 10.9 MB per GPU, held until the process ends.
 * Limits: num_blocks <= 16384 and at most 65536 candidate cells (fluid cells
 * with a solid-like in-block neighbour); beyond either, OCTLB_EINVAL (the
 * block limit before any device work, the candidate limit after the
 * counting pass, before the scratch could overflow).  The call synchronizes `stream` twice
 * (the candidate and link counts come back to the host).  Calls that
 * overlap on one GPU share the scratch and race: serialize them. */
int octlb_build_links_noalloc(int32_t num_blocks, int32_t n, const double* bounds, const uint8_t* material,
                              int64_t num_triangles, const double* triangles, int64_t capacity,
                              int64_t* num_links, int32_t* links, double* q, void* stream);

/* Re-adaptation (the reference plans a dynamic-AMR rebuild,
 * doc/prd/octlb-framework.md:180,521, and has
 * no migration code): after the forest changes and a new handle dst is
 * built for it, move src's state onto dst.  Same block edge, halo,
 * precision and relaxation rate.  Unchanged blocks (same level and
 * bounds) are copied whole; a block refined from a src block takes
 * feq + s_p fneq of its parent cells (s_p = (tau - 1/4)/tau, the
 * coupler's prolongation rule, level_coupler.cpp:163-214); a block
 * coarsened from eight src blocks takes the mean rho, the
 * momentum-conserving mean u and s_r times the mean fneq of its fine cells
 * (s_r = tau/(tau - 1/4)).  The host matches blocks by hashed (level,
 * integer origin) keys and uploads one task per new block; one device
 * launch (a workgroup per new block) moves the populations from src's
 * buffers into dst's, FP64 arithmetic in both precisions. */
int octlb_dev_migrate(octlb_dev* dst, octlb_dev* src);

/* ------------------------------------------------------------------ */
/* Output from device buffers in the reference's VTK formats           */
/* ------------------------------------------------------------------ */
/* Density and velocity of every core cell from the device populations,
 * CellProxy::computeRhoU (block_lattice.h:113-125) as the reference's VTK
 * samplers take them (vtk_lbm_fields.h:42-80), in VTK cell order:
 * rho[block][k][j][i], u[block][k][j][i][3] (x fastest).  4 doubles per cell
 * cross to the host instead of 19 populations. */
int octlb_dev_macro_fields(octlb_dev* dev, double* rho, double* u);
/* Cylinder3dCase::write_vtk_timestep (cylinder3d_case.h:878-909): per block
 * output_dir/base_block<id>_T<iT>.vts with CellData "velocity" (3) and
 * "pressure" (rho/3), the index output_dir/base_T<iT as %05d>.vtm, and a
 * DataSet entry base_T<iT>.vtm appended to output_dir/base.pvd (the
 * reference's names, including the unpadded time step in the .pvd entry).
 * output_dir must exist. */
int octlb_dev_write_vtk(octlb_dev* dev, int32_t iT, const char* output_dir, const char* base_name);
/* WriteStructuredGridVts (structured_grid_writer.cpp:61-117): one block of
 * nx*ny*nz cells over bounds (x_min, y_min, z_min, x_max, y_max, z_max);
 * field f has components[f] (1..3) doubles per cell in VTK cell order. */
int octlb_vtk_write_vts(const char* path, int32_t nx, int32_t ny, int32_t nz, const double bounds[6],
                        int32_t num_fields, const char* const* names, const int32_t* components,
                        const double* const* data);
/* WriteVtmIndex (parallel_vtk_index.cpp:132-158), one rank. */
int octlb_vtk_write_vtm(const char* path, int32_t count, const char* const* relative_paths);
/* AppendPvdTimestep (parallel_vtk_index.cpp:160-186): creates the
 * collection if absent, then appends one DataSet. */
int octlb_vtk_append_pvd(const char* pvd_path, int32_t iT, const char* vtm_relative_path);

/* ------------------------------------------------------------------ */
/* Discrete adjoint (no counterpart in the reference, which has none): */
/* gradients of a time-averaged drag objective through the step that   */
/* octlb_dev_advance runs, with the concrete handler, plain BGK and     */
/* Bouzidi streaming on 8^3 blocks.  One level without FD boundary      */
/* kinds runs the single-level engine (fused kernels); several levels   */
/* (Lagrava coupling, level_coupler.cpp) and/or FD inlet / outlet cells */
/* (interpolated_velocity.h) run the multi-level engine.  One adjoint   */
/* entry octlb_adj_run(dev, ...) is split into create / run so the      */
/* transposed stream operator is built once per adapt.                  */
/* ------------------------------------------------------------------ */
typedef struct octlb_adj octlb_adj;

/* Build the adjoint of dev's step from its topology, BcKinds and Bouzidi
 * links as currently set (OCTLB_EINVAL if dev is outside the scope above). */
int octlb_adj_create(octlb_dev* dev, octlb_adj** out);
int octlb_adj_destroy(octlb_adj* adj);
/* From feq(rho0, u0) on the core cells whose BcKind is in kind_mask (all
 * other slots zero, step 0), run `steps` steps -- single-level engine:
 * updates of the blocks' level (a forest whose blocks all sit on level L
 * updates 2^L times per coarse step of octlb_dev_advance); multi-level
 * engine: coarse steps, time t = 0 for the first --, keeping a checkpoint every
 * `checkpoint` steps, and return J = mean over steps win_begin+1 .. win_end
 * of dir . drag(bbox) (octlb_dev_drag's sum), dJ/du0[3] and, for each of the
 * num_links (block, cell, pop) triples in link, dJ/dq of that link's wall
 * fraction (0 for a triple that is not a Bouzidi row).  The forward sweep
 * and its recompute run in the handle's precision (an FP32 handle: FP32
 * states and collisions, the FP32 path's arithmetic); adjoint variables,
 * vector-Jacobian products and all sums are FP64.  Device memory ~
 * (steps/checkpoint + checkpoint + 3) states of 19 x live slots x 4 or 8 B
 * plus 3 FP64 adjoint states. */
int octlb_adj_run(octlb_adj* adj, int32_t steps, int32_t win_begin, int32_t win_end, int32_t checkpoint,
                  const double dir[3], const double bbox[6], double rho0, const double u0[3],
                  uint32_t kind_mask, int64_t num_links, const int32_t* link, double* J,
                  double dJ_du0[3], double* dJ_dq);
/* Core populations after the last run's forward sweep, in download_core
 * order [block][cell][19] (for checks against octlb_dev_advance). */
int octlb_adj_final_core(octlb_adj* adj, double* out);
/* dJ/d(inlet velocity) of the last run: the derivative with respect to
 * the inlet_u vector of the velocity boundary's inlet field (uniform,
 * Poiseuille or channel profile, inlet_velocity_field.h:47-270) through the
 * velocity-Dirichlet collide and the FD pass; zeros for the single-level
 * engine (no FD cells). */
int octlb_adj_inlet_grad(const octlb_adj* adj, double out[3]);
/* Device time of the last run's forward sweep and reverse sweep (ms). */
int octlb_adj_timing(const octlb_adj* adj, double* forward_ms, double* reverse_ms);
/* Multi-level engine: device memory allowed for one checkpoint segment's
 * tapes (bytes; -1, the default: 60 % of the free device memory).  When a
 * segment's K steps of tapes fit, the reverse sweep runs one taping pass per
 * segment (two forward passes per step in all); otherwise it recomputes the
 * segment's states and re-runs each step before reversing it (three).  Both
 * give the same gradients bit for bit.  *packed_last_run (may be null)
 * reports which mode the last run used. */
int octlb_adj_set_tape_budget(octlb_adj* adj, int64_t bytes, int32_t* packed_last_run);

#ifdef __cplusplus
}
#endif

#endif /* OCTLB_GPU_H_ */
