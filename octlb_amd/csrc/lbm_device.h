// Device data structures, boundary helpers and kernel launchers of the
// per-time-step AMR D3Q19 update (gfx950).
//
// Two storage modes, both with two ping-pong copies and one parity bit per
// refinement level:
//  * blocked (8^3 blocks, halo 1 -- the AMR path):
//      core  [block][19][512]  SoA, cell (ix*8+iy)*8+iz, shifted populations
//      shell [shell][19][488]  halo slots of blocks that own a halo face or
//            edge the same-level ghost exchange does not refresh (coarse-fine
//            faces, domain faces, unexchanged edges).  Exchanged halo slots
//            are never stored: the fused kernel reads the neighbour's cell
//            and collides it on the fly, which is the value
//            GhostSchedule::exchange copies there (ghost_schedule.h:121-255).
//  * padded (one large block, any halo -- the single-block cavity path):
//      [block][19][NP^3] halo-inclusive SoA, so the reference's rotation of
//      the whole padded storage (block_lattice.cpp:551-577) is reproduced.
#ifndef OCTLB_AMD_CSRC_LBM_DEVICE_H_
#define OCTLB_AMD_CSRC_LBM_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19.h"
#include "octlb_gpu.h"

namespace octlb {

// Pointer qualifier for read-only kernel tables: in device code the
// constant address space, whose loads never alias the kernel's stores, so
// the compiler keeps them as hoistable scalar loads like kernel arguments.
#if defined(__HIP_DEVICE_COMPILE__)
#define OCTLB_CONST_TABLE __attribute__((address_space(4)))
#else
#define OCTLB_CONST_TABLE
#endif

constexpr int kTile = 8;                       // tile edge (cells)
constexpr int kBox = kTile + 2;                // tile + one-cell apron
constexpr int kBoxCells = kBox * kBox * kBox;  // 1000
constexpr int kTileCells = kTile * kTile * kTile;
constexpr int kApron = kBoxCells - kTileCells;  // 488
constexpr int kMaxLevels = 16;
// Apron slots (the halo of the 10^3 box, padded coordinates 0..9) in
// workgroup-thread and shell-store order: the six faces x-, x+, y-, y+, z-,
// z+ (64 slots each, the later in-plane axis fastest, so a wave reads one
// neighbour block's boundary layer in its memory order), then the 12 edges
// (8 slots each), then the 8 corners.
struct ApronTable {
  short idx[kApron];
};
constexpr ApronTable MakeApron() {
  ApronTable t{};
  int k = 0;
  for (int f = 0; f < 6; ++f) {
    const int axis = f / 2, p = (f & 1) ? kBox - 1 : 0;
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    for (int a = 1; a <= kTile; ++a)
      for (int c = 1; c <= kTile; ++c) {
        int v[3] = {0, 0, 0};
        v[axis] = p;
        v[o1] = a;
        v[o2] = c;
        t.idx[k++] = static_cast<short>((v[0] * kBox + v[1]) * kBox + v[2]);
      }
  }
  for (int a1 = 0; a1 < 3; ++a1)
    for (int a2 = a1 + 1; a2 < 3; ++a2)
      for (int s1 = 0; s1 < 2; ++s1)
        for (int s2 = 0; s2 < 2; ++s2)
          for (int r = 1; r <= kTile; ++r) {
            int v[3] = {0, 0, 0};
            v[a1] = s1 ? kBox - 1 : 0;
            v[a2] = s2 ? kBox - 1 : 0;
            v[3 - a1 - a2] = r;
            t.idx[k++] = static_cast<short>((v[0] * kBox + v[1]) * kBox + v[2]);
          }
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y)
      for (int z = 0; z < 2; ++z)
        t.idx[k++] = static_cast<short>(((x ? kBox - 1 : 0) * kBox + (y ? kBox - 1 : 0)) * kBox + (z ? kBox - 1 : 0));
  return t;
}

constexpr int kMacro = 4 + kQ;  // rho, u[3], f_neq[19]

// Resolved DomainBcSpec of one face (FindSpec, interpolated_velocity.h:38-45).
struct DevSpec {
  int type;
  int inlet;
  int inlet_face;
  int n0, n1;
  double u_wall[3];
  double rho_target;
  double inlet_u[3];
  double u_peak;
  double lo0, hi0, lo1, hi1;
  double ramp_end_t;
};

struct SimConst {
  DevSpec spec[6];
  double omega;
  int n, h, NP, ncore, NS;
  int plane;      // elements per population plane of one block (512 or NP^3)
  int padded;     // storage mode
  int handler;    // 0 NoOp, 1 concrete
  int const_rho;
  int bouzidi;    // per-cell stream path (else padded-block rotate)
};

enum : int {
  kFlagLinks = 1,      // block has Bouzidi links
  kFlagShellLive = 2,  // block's shell is read by its core: evolve and persist it
  kFlagDrag = 4,       // block has drag-snapshot links
  kFlagPlain = 8,      // every cell streams by the plain pull (no per-cell rule applies)
  kFlagProlong = 16,   // block holds fine-side targets of a prolongation
  kFlagPush = 32,      // push-class: updated by the push body (KBlock), see push_to
  kFlagPushOut = 64,   // pull-class block with push-class blocks pulling from it
};

// Halo parts of a block: faces 0..5 (x-, x+, y-, y+, z-, z+), then the 12
// edges 6 + pair * 4 + s1 * 2 + s2 for the axis pairs (x,y), (x,z), (y,z)
// with s1, s2 the side (0 low, 1 high) on each axis.
struct BlockTab {
  int level;
  int shell;    // shell index or -1
  int flags;
  int member;   // ensemble member whose constants (SimConst) the block uses
  int nbr[18];  // block whose cells this block's halo part holds; -1 = not exchanged
  // push_to[s]: the push-class block whose halo part opposite to s is this
  // block (it pulls this block's cells that stream across side s), or -1
  int push_to[18];
  double origin[3];
  double cw;
};

// blocked mode: block = the tile's block; ox = second block of a push tile
// (-1: none), oy = 1 for a push tile (two push-class blocks, one per half
// of the workgroup), 0 for a pull tile.  padded mode: ox, oy, oz = origin
// of the tile in the padded box.
struct Tile {
  int block;
  int ox, oy, oz;
};

struct ParArray {
  int v[kMaxLevels];
};

template <typename T>
struct FdArgs {
  const int2* cells;     // (block, cell) per FD entry, phases plane|edge|corner
  const int* smap;       // [fd block][ncore] -> entry or -1
  const int* fd_index;   // block -> fd block or -1
  const int* level;      // block -> level
  const BlockTab* bt;
  const uint8_t* kind;
  T* pops0;
  T* pops1;
  T* post;               // [entries][19]
  const int* step;
  int* step_next;
  const OCTLB_CONST_TABLE SimConst* scm;  // per ensemble member, indexed by BlockTab::member
  const short* fd_list;  // blocked mode: [fd block][512] FD cells, plane|edge|corner order
  const int4* fd_count;  // blocked mode: [fd block] end offsets of the three phases, total
  ParArray par;
  SimConst sc;
};

template <typename T>
struct CoupleArgs {
  const int* plan;  // 12 ints per coupling point
  T* pops0;
  T* pops1;
  T* prev;          // [points][kMacro]
  ParArray par;
  T prolong_scale;
  T restrict_scale;
  const int* step;
  int* step_next;
};

template <typename T>
struct TileArgs {
  const Tile* tiles;
  int ntiles;
  const BlockTab* bt;
  const uint8_t* kind;
  const uint8_t* rule;    // per core cell: 1 = some pull-stream rule beyond the plain pull applies
  const T* src;
  T* dst;
  const T* shell_src;
  T* shell_dst;
  const uint32_t* qmask;
  const int* qoff;
  const double* qval;
  const uint32_t* dmask;
  const int* doff;
  T* snap;
  double* stats;          // [blocks*2] partial sums per block of the level (want_stats == 1)
  const int* block_rank;  // block -> its index among its level's blocks in visiting order
  double* stat_val;       // per-cell statistic in reference order (want_stats == 2)
  uint8_t* stat_use;
  int tiles_per_block;
  const double* avg_rho;  // previous step's ConstRho average, per ensemble member
  const OCTLB_CONST_TABLE SimConst* scm;  // per ensemble member (relaxation rate, boundary specs), by BlockTab::member
  const int* step;        // coarse-step counter (time of the inlet ramps)
  int* step_next;         // non-null in the first kernel of a step: *step_next = *step + 1
  int want_stats;         // 0 none, 1 parallel partial sums, 2 reference order
  int snap_on;            // write the post-collide drag snapshot
  // FD passes of boundary blocks of other levels, run by the first fd_wgs
  // workgroups of the launch (fd_cnt of them busy, the rest pad to 8)
  FdArgs<T> fd;
  const int* fd_blocks;
  const short* fd_slot;
  const int* fd_idx;
  int fd_cnt;
  int fd_wgs;
  int fd_rep_level;       // blocks of this level get two passes (one deferred from earlier), or -1
  // full-time prolongation in the same launch, when it is the last update
  // before it: pc_wgs workgroups after the FD ones run the coupling points,
  // and the cell threads leave the targeted cells to them (pfull non-null)
  const int* pfull;       // [block][512] active coupling point targeting the cell, or -1
  CoupleArgs<T> pc;
  const int* pc_list;
  int pc_count;
  int pc_wgs;
  // coupling steps that follow this update, done by the coupled blocks' own
  // workgroups after their stores (non-null: per block b, the active points
  // ep_*_pts[ep_*_beg[b] .. ep_*_beg[b + 1])): the half-time prolongation
  // whose coarse cells this launch updates, the restriction whose fine cells
  // it updates
  const int* ep_pro_beg;
  const int* ep_pro_pts;
  CoupleArgs<T> ep_pro;
  const int* ep_res_beg;
  const int* ep_res_pts;
  CoupleArgs<T> ep_res;
  SimConst sc;
};



template <typename T>
struct DragArgs {
  const int4* links;  // block, cell, pop, snapshot index (reference loop order)
  const BlockTab* bt;
  const T* pops0;
  const T* pops1;
  const T* snap;      // nullptr: no snapshot, read the live populations
  double* term;
  uint8_t* use;
  double bbox[6];
  ParArray par;
  SimConst sc;
};

__host__ __device__ inline int CoreSlot(const SimConst& sc, int c) {
  if (!sc.padded) return c;
  const int n = sc.n;
  const int x = c / (n * n), y = (c / n) % n, z = c % n;
  return ((x + sc.h) * sc.NP + (y + sc.h)) * sc.NP + z + sc.h;
}

// ------------------------------------------------------------- BC helpers
// smoothstep ramp x^3(10+x(-15+6x)) (inlet_velocity_field.h:77-90)
__device__ inline double RampD(double t, double ramp_end) {
  if (ramp_end <= 0.0) return 1.0;
  double x = t / ramp_end;
  if (x < 0.0) {
    x = 0.0;
  } else if (x > 1.0) {
    x = 1.0;
  }
  const double x2 = x * x;
  const double x3 = x2 * x;
  return x3 * (10.0 + x * (-15.0 + 6.0 * x));
}

__device__ inline double Clamp01(double c) { return c < 0.0 ? 0.0 : (c > 1.0 ? 1.0 : c); }

// InletVelocityField::velocity / velocity_phys (inlet_velocity_field.h:19-270)
__device__ inline void InletU(const DevSpec& sp, int ix, int iy, int iz, double t, const double* org,
                              double cw, double* u) {
  const bool xf = sp.inlet_face == OCTLB_XMIN || sp.inlet_face == OCTLB_XMAX;
  const bool zf = sp.inlet_face == OCTLB_ZMIN || sp.inlet_face == OCTLB_ZMAX;
  if (sp.inlet == OCTLB_INLET_CHANNEL) {
    const double px = org[0] + (ix + 0.5) * cw;
    const double py = org[1] + (iy + 0.5) * cw;
    const double pz = org[2] + (iz + 0.5) * cw;
    const double r = RampD(t, sp.ramp_end_t);
    const double a0 = xf ? py : px;
    const double a1 = zf ? py : pz;
    const double c0 = Clamp01((a0 - sp.lo0) / (sp.hi0 - sp.lo0));
    const double c1 = Clamp01((a1 - sp.lo1) / (sp.hi1 - sp.lo1));
    const double p0 = 4.0 * c0 * (1.0 - c0);
    const double p1 = 4.0 * c1 * (1.0 - c1);
    const double mag = r * sp.u_peak * p0 * p1;
    for (int d = 0; d < 3; ++d) u[d] = mag * sp.inlet_u[d];
    return;
  }
  if (sp.inlet == OCTLB_INLET_POISEUILLE) {
    const double r = RampD(t, sp.ramp_end_t);
    const int i0 = xf ? iy : ix;
    const int i1 = zf ? iy : iz;
    const double c0 = (static_cast<double>(i0) + 0.5) / static_cast<double>(sp.n0);
    const double c1 = (static_cast<double>(i1) + 0.5) / static_cast<double>(sp.n1);
    const double p0 = 4.0 * c0 * (1.0 - c0);
    const double p1 = 4.0 * c1 * (1.0 - c1);
    const double mag = r * sp.u_peak * p0 * p1;
    for (int d = 0; d < 3; ++d) u[d] = mag * sp.inlet_u[d];
    return;
  }
  const double r = RampD(t, sp.ramp_end_t);
  for (int d = 0; d < 3; ++d) u[d] = r * sp.inlet_u[d];
}

// Scale factor of the prescribed velocity with respect to the inlet
// velocity vector of the cell's inlet face (InletU: u = factor * inlet_u);
// 0 where the velocity does not come from an inlet field.
__device__ inline double InletFactor(const SimConst& sc, const double* org, double cw, int ix, int iy, int iz,
                                     double t) {
  const int n = sc.n;
  const bool on[6] = {ix == 0, ix == n - 1, iy == 0, iy == n - 1, iz == 0, iz == n - 1};
  for (int f = 0; f < 6; ++f) {
    if (!on[f]) continue;
    const DevSpec& sp = sc.spec[f];
    if (sp.inlet == OCTLB_INLET_NONE) continue;
    const double r = RampD(t, sp.ramp_end_t);
    if (sp.inlet == OCTLB_INLET_UNIFORM) return r;
    const bool xf = sp.inlet_face == OCTLB_XMIN || sp.inlet_face == OCTLB_XMAX;
    const bool zf = sp.inlet_face == OCTLB_ZMIN || sp.inlet_face == OCTLB_ZMAX;
    if (sp.inlet == OCTLB_INLET_CHANNEL) {
      const double px = org[0] + (ix + 0.5) * cw, py = org[1] + (iy + 0.5) * cw, pz = org[2] + (iz + 0.5) * cw;
      const double c0 = Clamp01(((xf ? py : px) - sp.lo0) / (sp.hi0 - sp.lo0));
      const double c1 = Clamp01(((zf ? py : pz) - sp.lo1) / (sp.hi1 - sp.lo1));
      return r * sp.u_peak * 4.0 * c0 * (1.0 - c0) * 4.0 * c1 * (1.0 - c1);
    }
    const double c0 = (static_cast<double>(xf ? iy : ix) + 0.5) / static_cast<double>(sp.n0);
    const double c1 = (static_cast<double>(zf ? iy : iz) + 0.5) / static_cast<double>(sp.n1);
    return r * sp.u_peak * 4.0 * c0 * (1.0 - c0) * 4.0 * c1 * (1.0 - c1);
  }
  return 0.0;
}

// detail::PrescribedBoundaryU (interpolated_velocity.h:69-117), block-local faces.
__device__ inline void PrescribedU(const SimConst& sc, const double* org, double cw, int ix, int iy,
                                   int iz, double t, double* u) {
  const int n = sc.n;
  u[0] = u[1] = u[2] = 0.0;
  const bool on[6] = {ix == 0, ix == n - 1, iy == 0, iy == n - 1, iz == 0, iz == n - 1};
  for (int f = 0; f < 6; ++f) {
    if (!on[f]) continue;
    if (sc.spec[f].inlet != OCTLB_INLET_NONE) {
      InletU(sc.spec[f], ix, iy, iz, t, org, cw, u);
      return;
    }
  }
  if (iy == n - 1 && ix > 0 && ix < n - 1 && iz > 0 && iz < n - 1) {
    for (int d = 0; d < 3; ++d) u[d] = sc.spec[OCTLB_YMAX].u_wall[d];
  }
}

// detail::FlatBoundaryFaceInfo (interpolated_velocity.h:134-167)
__device__ inline bool FlatInfo(int n, int ix, int iy, int iz, int& dir, int& ori) {
  int count = 0;
  dir = 0;
  ori = 0;
  if (ix == 0) {
    ++count;
    dir = 0;
    ori = -1;
  } else if (ix == n - 1) {
    ++count;
    dir = 0;
    ori = 1;
  }
  if (iy == 0) {
    ++count;
    dir = 1;
    ori = -1;
  } else if (iy == n - 1) {
    ++count;
    dir = 1;
    ori = 1;
  }
  if (iz == 0) {
    ++count;
    dir = 2;
    ori = -1;
  } else if (iz == n - 1) {
    ++count;
    dir = 2;
    ori = 1;
  }
  return count == 1;
}

// BcDispatcher::collide (bc_dispatcher.h:32-81) for the concrete handler;
// BlockLattice::collide / collide_const_rho (block_lattice.cpp:376-499) for
// the NoOp handler.  Statistics follow CollideRhoStats (block_lattice.h:22-32).
struct Origin {
  double x, y, z, cw;
};

template <typename T>
__device__ __forceinline__ void CollideCell(T* f, int k, const SimConst& sc, const Origin& og, int ix, int iy,
                                   int iz, double t, T avg, bool want, double& ssum, int& scnt) {
  const double org[3] = {og.x, og.y, og.z};
  const double cw = og.cw;
  const T omega = static_cast<T>(sc.omega);
  if (k == kBulk) {
    if (sc.const_rho) {
      T rho, u[3];
      ProxyRhoU(f, rho, u);
      if (want) {
        if constexpr (kShiftedArith<T>) {
          ssum += 1.0 + (static_cast<double>(rho) - static_cast<double>(avg));  // rho + (1 - avg)
        } else {
          ssum += static_cast<double>(rho + (T(1) - avg));
        }
        scnt += 1;
      }
      ConstRhoBgk(f, omega, avg, rho, u);
    } else if (sc.handler == OCTLB_HANDLER_CONCRETE) {
      T rho, u[3];
      ProxyRhoU(f, rho, u);
      OlbBgk(f, rho, u, omega);
    } else {
      CollideBgkAt(f, omega);
    }
    return;
  }
  if (sc.handler != OCTLB_HANDLER_CONCRETE) return;
  if (k == kBounceBack) {
    T tmp[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) tmp[i] = f[i];
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = tmp[Opp(i)];
    return;
  }
  if (k == kVelocity) {
    // CollideDirichletBoundaryCellAt (interpolated_velocity.h:858-884)
    double ud[3];
    PrescribedU(sc, org, cw, ix, iy, iz, t, ud);
    T u[3] = {static_cast<T>(ud[0]), static_cast<T>(ud[1]), static_cast<T>(ud[2])};
    const int n = sc.n;
    const int fc = (ix == 0) + (ix == n - 1) + (iy == 0) + (iy == n - 1) + (iz == 0) + (iz == n - 1);
    T rho;
    if (fc >= 2) {
      T ut[3];
      ProxyRhoU(f, rho, ut);
    } else {
      int dir, ori;
      FlatInfo(n, ix, iy, iz, dir, ori);
      rho = VelRhoFromPop(dir, ori, f, u);
    }
    OlbBgk(f, rho, u, omega);
    if (want) {
      ssum += RhoDouble(rho);
      scnt += 1;
    }
  }
}

// --------------------------------------------------------------- launchers
template <typename T>
void LaunchTile(const TileArgs<T>& a, hipStream_t s);
// ConstRho averages of the members: member m owns tiles [tile_begin[m],
// tile_begin[m+1]) of the partial sums, or statistic slots [cell_begin[m],
// cell_begin[m+1]) in the reference's order.
template <typename T>
void LaunchStatsFinalize(const double* partial, const int* tile_begin, int members, double* avg, hipStream_t s);
void LaunchStatsSequential(const double* val, const uint8_t* use, const int* cell_begin, int members, double* avg,
                           hipStream_t s);
void LaunchStepInc(const int* step, int* step_next, hipStream_t s);
template <typename T>
void LaunchInitEq(T* pops, const uint8_t* kind, int nb, const SimConst& sc, double rho0,
                  const double* u0, uint32_t mask, hipStream_t s);
template <typename T>
void LaunchInitField(T* pops, int nb, const SimConst& sc, const double* rho, const double* u,
                     hipStream_t s);
template <typename T>
void LaunchFd(const FdArgs<T>& a, const int* phase_begin, hipStream_t s);
// FD passes by FD-cell count: class c holds the blocks with at most
// kFdClassCells[c] FD cells (c < 2) and the rest (c = 2); ncls[c] blocks
// each, listed class by class in fd_idx, one launch per class with the LDS
// it needs (more workgroups share a CU).
constexpr int kFdClasses = 3;
constexpr int kFdClassCells[2] = {64, 128};
// fd_idx entries: (FD block, block, its level, 0).
template <typename T>
void LaunchFdBlocks(const FdArgs<T>& a, const short* fd_slot, const int4* fd_idx, const int ncls[kFdClasses],
                    hipStream_t s);
// FD pass + the coupling step after it in one launch; mode 0/1 = prolongation
// (full / half time), 2 = restriction.
template <typename T>
void LaunchFdCouple(const FdArgs<T>& fa, const short* fd_slot, const int4* fd_idx, int nfd, const CoupleArgs<T>& ca,
                    const int* list, int count, int mode, hipStream_t s);
template <typename T>
void LaunchCouplerInit(const CoupleArgs<T>& a, int npts, hipStream_t s);
template <typename T>
void LaunchProlong(const CoupleArgs<T>& a, const int* list, int count, int half, hipStream_t s);
template <typename T>
void LaunchRestrict(const CoupleArgs<T>& a, const int* list, int count, hipStream_t s);
template <typename T>
void LaunchDrag(const DragArgs<T>& a, const int* block_begin, int nblk, double* block_out, unsigned int* done,
                double* result, hipStream_t s);
template <typename T>
void LaunchMass(const T* p0, const T* p1, const ParArray& par, const BlockTab* bt, const uint8_t* kind,
                const SimConst& sc, int nb, double* partial, double* out, hipStream_t s);
template <typename T>
void LaunchMacro(const T* p0, const T* p1, const ParArray& par, const BlockTab* bt, const SimConst& sc, int nb,
                 double* rho, double* u, hipStream_t s);

// ------------------------------------------------------ state migration
// What one block of the adapted forest takes from the old forest.
enum MigMode : int { kMigCopy = 0, kMigRefine = 1, kMigCoarsen = 2 };
struct MigTask {
  int mode;
  // copy: src[0] = the block with the same level and bounds; refine:
  // src[0] = the parent, oct = this block's octant in it (0/1 per axis);
  // coarsen: src[c] = the child at octant c = ox + 2 oy + 4 oz
  int src[8];
  int oct[3];
};
// One forest's populations as the kernel reads (old) or writes (new) them:
// block b lives in buffer par[bt[b].level].
template <typename T>
struct MigSide {
  const BlockTab* bt;
  T* pops[2];
  T* shells[2];
  ParArray par;
};
template <typename T>
struct MigArgs {
  const MigTask* task;
  MigSide<T> src, dst;
  int n, h, NP, NS, plane, padded;
  double sp, sr;  // (tau - 1/4)/tau and its inverse
};
// list: the new blocks of each rule, count[0] copies, then count[1]
// refinements, then count[2] coarsenings.
template <typename T>
void LaunchMigrate(const MigArgs<T>& a, const int* list, const int count[3], hipStream_t s);

}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_LBM_DEVICE_H_
