// MI355X (gfx950) per-time-step AMR D3Q19 lattice-Boltzmann update: device
// kernels.  The C ABI that drives them is in api.hip.
//
// One coarse step replays TimeLoop::advance_one (time_loop.h:201-246):
//   advance(l): fused collide + halo + stream of the level-l blocks
//               -> [ConstRho statistics] -> post-stream FD boundaries
//               -> if l < max: prolong(half) advance(l+1) prolong(full)
//                              advance(l+1) restrict(l+1)
// The fused kernel keeps the reference's phase semantics: it reads the
// post-stream state S of step n-1, collides the block's cells and its
// one-cell apron (neighbour cells are re-collided from their own pre-collide
// state; halo slots the reference's ghost exchange does not refresh come
// from the block's shell store), pull-streams through an LDS copy of the
// whole 10^3 box and writes S of step n to the other buffer.
#include "lbm_device.h"
#include "adjoint_ml.h"

namespace octlb {

// ------------------------------------------------------------------ apron map
__constant__ ApronTable kApronTab = MakeApron();

__device__ inline int EdgeSlot(int a1, int s1, int a2, int s2) {
  // a1 < a2; axis pair (x,y)=0, (x,z)=1, (y,z)=2
  const int pair = (a1 == 0) ? (a2 == 1 ? 0 : 1) : 2;
  return 6 + pair * 4 + s1 * 2 + s2;
}

__device__ inline int XcdTile(int w, int nt) {
  // Consecutive tiles (Morton-ordered blocks) go to the same XCD, so the
  // apron re-reads of neighbouring blocks hit that XCD's L2.
  const int q8 = nt >> 3;
  return (w < (q8 << 3)) ? (w & 7) * q8 + (w >> 3) : w;
}

// A launch's read-only inputs seen through the constant address space:
// their loads cannot alias the launch's stores, so the compiler may hoist
// and batch them (and keep uniform ones as scalar loads).
template <typename P>
__device__ __forceinline__ const OCTLB_CONST_TABLE P* Ro(const P* p) {
  return (const OCTLB_CONST_TABLE P*)p;
}

// The coarse-step counter is double-buffered: a step's kernels read *step;
// the first kernel of the step also writes the next step's slot, which
// nothing in this step reads.
__device__ inline void BumpStep(const int* step, int* step_next) {
  if (step_next != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *step_next = *step + 1;
}

template <typename T>
__device__ inline void BlockReduceStats(double s, int c, double* red_s, int* red_c, double* out) {
  const int tid = threadIdx.x;
  __syncthreads();  // red_s / red_c alias the streaming planes
  red_s[tid] = s;
  red_c[tid] = c;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (tid < w) {
      red_s[tid] += red_s[tid + w];
      red_c[tid] += red_c[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = red_s[0];
    out[1] = static_cast<double>(red_c[0]);
  }
}

// Statistics of a push tile, one partial per block: each half reduces its
// 512 values with the tree a pull tile applies to its cells (whose other
// half holds zeros), so a block's partial does not depend on the tiling.
template <typename T>
__device__ inline void HalfReduceStats(double s, int c, double* red_s, int* red_c, double* out0, double* out1) {
  const int tid = threadIdx.x;
  __syncthreads();  // red_s / red_c alias the streaming planes
  red_s[tid] = s;
  red_c[tid] = c;
  __syncthreads();
  for (int w = kTileCells / 2; w > 0; w >>= 1) {
    if ((tid & (kTileCells - 1)) < w) {
      red_s[tid] += red_s[tid + w];
      red_c[tid] += red_c[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out0[0] = red_s[0];
    out0[1] = static_cast<double>(red_c[0]);
  }
  if (tid == kTileCells && out1 != nullptr) {
    out1[0] = red_s[kTileCells];
    out1[1] = static_cast<double>(red_c[kTileCells]);
  }
}

// Pull-stream rule of one core cell (block_lattice.cpp:579-674) for
// population i (a constant after the caller's unrolling), reading the LDS
// box of post-collide values: pl = plane of population i, po = plane of its
// opposite, own_l = the cell's box index.
template <typename T, bool PERCELL>
__device__ __forceinline__ T StreamPop(int i, const T* pl, const T* po, const uint8_t* kbox, int own_l, int k,
                                       bool src_interior, uint32_t qm, const double* qv) {
  const int off = (Cx(i) * kBox + Cy(i)) * kBox + Cz(i);
  const T pulled = pl[own_l - off];
  if (!PERCELL) return pulled;
  if (k == kSolid) return pl[own_l];
  if (ReflectsOnPull(k) && !src_interior) return po[own_l];
  const int sk = kbox[own_l - off];
  if (sk == kSolid || sk == kBouzidi) {
    const T own_opp = po[own_l];
    double q = 1.0;
    if ((qm >> i) & 1u) q = qv[__popc(qm & ((1u << i) - 1u))];
    if (q > 1.0e-10 && q < 1.0 - 1.0e-10) {
      const T own_i = pl[own_l];
      const int nl = own_l + off;
      const double f_int = (kbox[nl] == kBulk) ? static_cast<double>(pl[nl]) : static_cast<double>(own_i);
      return static_cast<T>(BouzidiPull(static_cast<double>(own_opp), f_int, static_cast<double>(own_i), q));
    }
    return own_opp;
  }
  return pulled;
}

// ------------------------------------------------------------- push streaming
// Push form of the plain pull (block_lattice.cpp:579-674 read from the source
// side): the post-collide value f of population i at core cell (x, y, z) of
// block b streams into cell (x, y, z) + c_i -- in b, or in the push-class
// block bt.push_to[part] whose halo part opposite to `part` holds b's cells
// (that block would have pulled it from there).  Destinations in pull-class
// blocks are skipped: those blocks pull (and re-collide) the value themselves.
template <typename T>
__device__ __forceinline__ void PushPop(int i, T f, const BlockTab& bt, int b, int x, int y, int z, T* dst) {
  constexpr int n = kTile, ncore = kTileCells;
  const int dx = x + Cx(i), dy = y + Cy(i), dz = z + Cz(i);
  const int ox = dx < 0 ? -1 : (dx >= n ? 1 : 0);
  const int oy = dy < 0 ? -1 : (dy >= n ? 1 : 0);
  const int oz = dz < 0 ? -1 : (dz >= n ? 1 : 0);
  int db = b;
  if (ox != 0 || oy != 0 || oz != 0) {
    int part;
    if ((ox != 0) + (oy != 0) + (oz != 0) == 1) {
      part = ox != 0 ? (ox > 0 ? 1 : 0) : (oy != 0 ? 2 + (oy > 0 ? 1 : 0) : 4 + (oz > 0 ? 1 : 0));
    } else {
      const int a1 = ox != 0 ? 0 : 1, a2 = oz != 0 ? 2 : 1;
      const int s1 = (a1 == 0 ? ox : oy) > 0 ? 1 : 0, s2 = (a2 == 2 ? oz : oy) > 0 ? 1 : 0;
      part = EdgeSlot(a1, s1, a2, s2);
    }
    db = bt.push_to[part];
    if (db < 0) return;
  }
  const int wx = dx - ox * n, wy = dy - oy * n, wz = dz - oz * n;
  dst[static_cast<size_t>(db) * kQ * ncore + static_cast<size_t>(i) * ncore + (wx * n + wy) * n + wz] = f;
}

// Pushes the populations of a boundary cell that cross the block's surface
// into push-class neighbours (pull-class blocks with kFlagPushOut).
template <typename T>
__device__ __forceinline__ void PushOut(const T* f, const BlockTab& bt, int b, int x, int y, int z, T* dst) {
  constexpr int n = kTile;
  if (x > 0 && x < n - 1 && y > 0 && y < n - 1 && z > 0 && z < n - 1) return;
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    const int dx = x + Cx(i), dy = y + Cy(i), dz = z + Cz(i);
    if (dx >= 0 && dx < n && dy >= 0 && dy < n && dz >= 0 && dz < n) continue;
    PushPop<T>(i, f[i], bt, b, x, y, z, dst);
  }
}

// Push tile body: two push-class blocks, one per half of the 1,024-thread
// workgroup, one thread per cell: load the cell's 19 populations
// (coalesced), collide, push.  No apron, no LDS exchange, no barrier (except
// for the statistics reduction).
template <typename T>
__device__ __forceinline__ void PushTile(const TileArgs<T>& a, const Tile& tl, double* red_s, int* red_c) {
  constexpr int n = kTile, ncore = kTileCells;
  const int tid = threadIdx.x;
  const int half = tid >= ncore ? 1 : 0;
  const int b = half ? tl.ox : tl.block;
  const int c = tid - half * ncore;
  double ssum = 0.0;
  int scnt = 0;
  if (b >= 0) {
    const BlockTab& bt = *(const BlockTab*)&Ro(a.bt)[b];
    const SimConst& sc = *(const SimConst*)&a.scm[bt.member];
    const int x = c >> 6, y = (c >> 3) & 7, z = c & 7;
    T f[kQ];
    const size_t base = static_cast<size_t>(b) * kQ * ncore + c;
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = Ro(a.src)[base + static_cast<size_t>(i) * ncore];
    const int k = Ro(a.kind)[static_cast<size_t>(b) * ncore + c];
    const double t = static_cast<double>(*Ro(a.step));
    const T avg = AvgArg<T>(sc.const_rho ? Ro(a.avg_rho)[bt.member] : 1.0);
    const Origin og{bt.origin[0], bt.origin[1], bt.origin[2], bt.cw};
    CollideCell<T>(f, k, sc, og, x, y, z, t, avg, a.want_stats != 0, ssum, scnt);
#pragma unroll
    for (int i = 0; i < kQ; ++i) PushPop<T>(i, f[i], bt, b, x, y, z, a.dst);
  }
  if (a.want_stats == 1)
    HalfReduceStats<T>(ssum, scnt, red_s, red_c, a.stats + 2 * a.block_rank[tl.block],
                       tl.ox >= 0 ? a.stats + 2 * a.block_rank[tl.ox] : nullptr);
}

// ------------------------------------------------------- blocked fused kernel
// One 1024-thread workgroup per 8^3 block.  Threads 0..511 own the block's
// cells; threads 512..999 own its 488 apron slots (the halo of the 10^3 box,
// in shell order).  Every thread loads its 19 pre-collide populations -- own
// cell, neighbour block's boundary cell (the ghost exchange, done in place)
// or, for slots no exchange fills, the block's shell -- collides real cells,
// and writes all 19 post-collide values into a 19 x 10^3 LDS box (76 KB FP32,
// 152 KB FP64; gfx950 gives a workgroup up to 160 KB).  After one barrier the
// cell threads pull-stream and store, and apron threads of unexchanged slots
// do the halo-shell torus pull (block_lattice.cpp:595-611).
// Lagrava macroscopic read / prolongation write, defined with the coupler below.
template <typename T>
__device__ inline void ReadMacro(const T* base, int plane, int c, T* m);
template <typename T>
__device__ inline void WriteMacro(T* base, int plane, int c, const T* m, T scaling);

template <typename T>
__device__ inline void ProlongPoint(const CoupleArgs<T>& a, const int* list, int count, int half, int j);
template <typename T>
struct RestrictLds;
template <typename T, int Groups>
__device__ inline void RestrictBodyIn(const CoupleArgs<T>& a, const int* list, int count, int wg,
                                      RestrictLds<T>* sm);

// FD boundary pass, defined with the FD kernels below.
constexpr int kMaxFdCells = 6 * 36 + 12 * 6 + 8;  // face, edge and corner cells of an 8^3 block
template <typename T, int MaxFd = kMaxFdCells>
struct FdSmem;
template <typename T, int MaxFd = kMaxFdCells, bool kEarly = false>
__device__ inline void FdBlockBody(const FdArgs<T>& a, const short* fd_slot, int4 d, FdSmem<T, MaxFd>& sm);
template <typename T>
__device__ inline void FdBlockBodyAt(const FdArgs<T>& a, const int* fd_blocks, const short* fd_slot, int fb,
                                     FdSmem<T>& sm);

// FP32: two workgroups per CU (2 x 77 KB LDS) need <= 64 VGPRs (8 waves per
// SIMD); FP64: one workgroup per CU (153 KB LDS), 4 waves per SIMD.
template <typename T>
struct BlockOccupancy {
  static constexpr int kWavesPerSimd = sizeof(T) == 4 ? 8 : 4;
};

template <typename T, bool PERCELL>
__global__ __launch_bounds__(1024, BlockOccupancy<T>::kWavesPerSimd) void KBlock(TileArgs<T> a) {
  __shared__ T plane[kQ][kBoxCells];
  __shared__ uint8_t kbox[kBoxCells];
  static_assert(sizeof(plane) >= 1024 * (sizeof(double) + sizeof(int)), "stats alias");
  double* red_s = reinterpret_cast<double*>(&plane[0][0]);
  int* red_c = reinterpret_cast<int*>(red_s + 1024);

  BumpStep(a.step, a.step_next);
  if (static_cast<int>(blockIdx.x) < a.fd_wgs) {
    // FD pass of a boundary block of another level: independent of this
    // level's blocks, so it runs beside them instead of in a launch of its own
    static_assert(sizeof(FdSmem<T>) <= sizeof(plane), "FD pass LDS must fit the box");
    if (static_cast<int>(blockIdx.x) < a.fd_cnt) {
      const int fb = a.fd_idx[blockIdx.x];
      const int reps = (a.fd_rep_level >= 0 && a.fd.level[a.fd_blocks[fb]] == a.fd_rep_level) ? 2 : 1;
      for (int r = 0; r < reps; ++r)  // each pass ends on a barrier: the next one sees its stores
        FdBlockBodyAt<T>(a.fd, a.fd_blocks, a.fd_slot, fb, *reinterpret_cast<FdSmem<T>*>(&plane[0][0]));
    }
    return;
  }
  if (static_cast<int>(blockIdx.x) < a.fd_wgs + a.pc_wgs) {
    // LevelCoupler::apply_full_time (level_coupler.cpp:438-478) on the fine
    // cells this launch's blocks leave unwritten
    ProlongPoint<T>(a.pc, a.pc_list, a.pc_count, 0,
                    (static_cast<int>(blockIdx.x) - a.fd_wgs) * static_cast<int>(blockDim.x) +
                        static_cast<int>(threadIdx.x));
    return;
  }
  const int ti = XcdTile(static_cast<int>(blockIdx.x) - a.fd_wgs - a.pc_wgs, a.ntiles);
  const Tile tl = Ro(a.tiles)[ti];
  if (tl.oy == 1) {
    PushTile<T>(a, tl, red_s, red_c);
    return;
  }
  const int b = tl.block;
  constexpr int n = kTile, ncore = kTileCells;
  const int tid = threadIdx.x;
  const bool own = tid < ncore;
  const int ap = tid - ncore;  // apron slot (shell index) when !own
  const BlockTab& bt = *(const BlockTab*)&Ro(a.bt)[b];
  const SimConst& sc = *(const SimConst*)&a.scm[bt.member];  // apron cells belong to the same member
  const int bflags = bt.flags;
  const int bshell = bt.shell;

  // ---- where this thread's cell lives
  int l = -1;       // 10^3 box index
  int src_b = -1;   // block whose core holds the cell (-1: shell or nothing)
  int cx = 0, cy = 0, cz = 0;
  int nb = -1;
  if (own) {
    cx = tid >> 6;
    cy = (tid >> 3) & 7;
    cz = tid & 7;
    l = ((cx + 1) * kBox + cy + 1) * kBox + cz + 1;
    src_b = b;
  } else if (ap < kApron) {
    l = kApronTab.idx[ap];
    const int px = l / (kBox * kBox), py = (l / kBox) % kBox, pz = l % kBox;
    const int x = px - 1, y = py - 1, z = pz - 1;
    const int ox = (x < 0 || x >= n), oy = (y < 0 || y >= n), oz = (z < 0 || z >= n);
    const int nout = ox + oy + oz;
    if (nout == 1) {
      const int axis = ox ? 0 : (oy ? 1 : 2);
      const int v = axis == 0 ? x : (axis == 1 ? y : z);
      nb = bt.nbr[axis * 2 + (v >= n ? 1 : 0)];
    } else if (nout == 2) {
      const int a1 = ox ? 0 : 1;
      const int a2 = oz ? 2 : 1;
      const int v1 = a1 == 0 ? x : y;
      const int v2 = a2 == 2 ? z : y;
      nb = bt.nbr[EdgeSlot(a1, v1 >= n ? 1 : 0, a2, v2 >= n ? 1 : 0)];
    }
    if (nb >= 0) {
      cx = x < 0 ? n - 1 : (x >= n ? 0 : x);
      cy = y < 0 ? n - 1 : (y >= n ? 0 : y);
      cz = z < 0 ? n - 1 : (z >= n ? 0 : z);
      src_b = nb;
    }
  }
  const int cc = (cx * n + cy) * n + cz;

  // ---- load: every global read is issued before any arithmetic
  T f[kQ];
  int k = kBulk;
  if (src_b >= 0) {
    const size_t base = static_cast<size_t>(src_b) * kQ * ncore + cc;
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = Ro(a.src)[base + static_cast<size_t>(i) * ncore];
    k = Ro(a.kind)[static_cast<size_t>(src_b) * ncore + cc];
  } else if (l >= 0 && bshell >= 0) {
    const size_t base = static_cast<size_t>(bshell) * kQ * kApron + ap;
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = Ro(a.shell_src)[base + static_cast<size_t>(i) * kApron];
  } else {
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = T(0);
  }
  const bool rules = PERCELL && own && !(bflags & kFlagPlain) && Ro(a.rule)[static_cast<size_t>(b) * ncore + tid] != 0;
  uint32_t qm = 0;
  const double* qv = a.qval;
  if (PERCELL && own && (bflags & kFlagLinks)) {
    qm = Ro(a.qmask)[static_cast<size_t>(b) * ncore + tid];
    qv = a.qval + Ro(a.qoff)[static_cast<size_t>(b) * ncore + tid];
  }

  // ---- collide real cells (own and exchanged apron); halo-shell slots are
  // streamed but never collided
  double ssum = 0.0;
  int scnt = 0;
  if (src_b >= 0) {
    const double t = static_cast<double>(*Ro(a.step));
    const T avg = AvgArg<T>(sc.const_rho ? Ro(a.avg_rho)[bt.member] : 1.0);
    const BlockTab& st = *(const BlockTab*)&Ro(a.bt)[src_b];
    const Origin og{st.origin[0], st.origin[1], st.origin[2], st.cw};
    CollideCell<T>(f, k, sc, og, cx, cy, cz, t, avg, own && a.want_stats != 0, ssum, scnt);
  }
  if (own && (bflags & kFlagPushOut)) PushOut<T>(f, bt, b, cx, cy, cz, a.dst);
  if (l >= 0) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) plane[i][l] = f[i];
    kbox[l] = static_cast<uint8_t>(own ? k : kBulk);  // kind_at_halo -> kBulk (block_lattice.cpp:510-518)
  }
  if (own && (bflags & kFlagDrag) && a.snap_on) {
    // post-collide snapshot of the outgoing populations on wall links
    // (take_post_collide_snapshot, block_lattice.cpp:309-316)
    const uint32_t dm = Ro(a.dmask)[static_cast<size_t>(b) * ncore + tid];
    if (dm) {
      T* sp = a.snap + Ro(a.doff)[static_cast<size_t>(b) * ncore + tid];
      int j = 0;
#pragma unroll
      for (int i = 0; i < kQ; ++i)
        if ((dm >> i) & 1u) sp[j++] = f[i];
    }
  }
  __syncthreads();

  // a cell that the full-time prolongation of this launch overwrites is
  // left to the coupling workgroups
  const bool keep = !(own && a.pfull != nullptr && (bflags & kFlagProlong) &&
                      Ro(a.pfull)[static_cast<size_t>(b) * ncore + tid] >= 0);
  if (own) {
    T* dst = a.dst + static_cast<size_t>(b) * kQ * ncore + tid;
    if (!keep) {
      // overwritten by the full-time prolongation of this launch
    } else if (!PERCELL || !rules) {
      // plain pull: no rule of the per-cell path changes any population
#pragma unroll
      for (int i = 0; i < kQ; ++i)
        dst[static_cast<size_t>(i) * ncore] = plane[i][l - ((Cx(i) * kBox + Cy(i)) * kBox + Cz(i))];
    } else {
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        const int sx = cx - Cx(i), sy = cy - Cy(i), sz = cz - Cz(i);
        const bool src_int = sx >= 0 && sx < n && sy >= 0 && sy < n && sz >= 0 && sz < n;
        dst[static_cast<size_t>(i) * ncore] =
            StreamPop<T, true>(i, plane[i], plane[Opp(i)], kbox, l, k, src_int, qm, qv);
      }
    }
  } else if (l >= 0 && nb < 0 && bshell >= 0 && (bflags & kFlagShellLive)) {
    // padded-block torus pull of a halo slot that no exchange refills;
    // in-halo and wrapped sources are both (d - off) mod NP^3
    T* sdst = a.shell_dst + static_cast<size_t>(bshell) * kQ * kApron + ap;
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const int off = (Cx(i) * kBox + Cy(i)) * kBox + Cz(i);
      int sidx = l - off;
      if (sidx < 0) sidx += kBoxCells;
      if (sidx >= kBoxCells) sidx -= kBoxCells;
      sdst[static_cast<size_t>(i) * kApron] = plane[i][sidx];
    }
  }
  if (a.want_stats == 1) {
    BlockReduceStats<T>(ssum, scnt, red_s, red_c, a.stats + 2 * a.block_rank[b]);
  } else if (a.want_stats == 2 && own) {
    const size_t ord = static_cast<size_t>(a.block_rank[b]) * ncore + tid;
    a.stat_val[ord] = ssum;
    a.stat_use[ord] = static_cast<uint8_t>(scnt);
  }
  // ---- coupling steps on this block's fresh output (the barrier makes the
  // workgroup's stores visible to all its threads and ends every read of the
  // box that the restriction reuses)
  if (a.ep_pro_beg != nullptr) {
    // LevelCoupler::apply_half_time (level_coupler.cpp:438-478) of the points
    // whose coarse cell lies in this block
    const int p0 = a.ep_pro_beg[b], np = a.ep_pro_beg[b + 1] - p0;
    if (np > 0) {
      __syncthreads();
      for (int j = tid; j < np; j += static_cast<int>(blockDim.x))
        ProlongPoint<T>(a.ep_pro, a.ep_pro_pts + p0, np, 1, j);
    }
  }
  if (a.ep_res_beg != nullptr) {
    // LevelCoupler::restrict (level_coupler.cpp:613-654) of the points whose
    // fine cell lies in this block, 32 points per pass
    const int p0 = a.ep_res_beg[b], np = a.ep_res_beg[b + 1] - p0;
    if (np > 0) {
      constexpr int kGroups = 1024 / 32;
      static_assert(sizeof(plane) >= kGroups * sizeof(RestrictLds<T>), "restriction LDS must fit the box");
      RestrictLds<T>* rs = reinterpret_cast<RestrictLds<T>*>(&plane[0][0]);
      for (int w = 0; w * kGroups < np; ++w) {
        __syncthreads();
        RestrictBodyIn<T, kGroups>(a.ep_res, a.ep_res_pts + p0, np, w, rs);
      }
    }
  }
}

// -------------------------------------------------------- padded fused kernel
// Tiles of 8^3 padded slots cover the whole padded block; every slot is
// updated.  Core slots stream by the cell rule, padding slots by the padded
// torus pull (block_lattice.cpp:595-611 / 551-577), whose wrapped sources
// are padding slots (h >= 1) read straight from the previous state.
template <typename T, bool PERCELL>
__global__ __launch_bounds__(512) void KPadded(TileArgs<T> a) {
  __shared__ T plane[2][kBoxCells];
  __shared__ uint8_t kbox[kBoxCells];  // 0..7 core kinds, 254 padding, 255 outside
  static_assert(sizeof(plane) >= 512 * (sizeof(double) + sizeof(int)), "stats alias");
  double* red_s = reinterpret_cast<double*>(&plane[0][0]);
  int* red_c = reinterpret_cast<int*>(red_s + 512);

  BumpStep(a.step, a.step_next);
  const int ti = XcdTile(blockIdx.x, a.ntiles);
  const Tile tl = a.tiles[ti];
  const int b = tl.block;
  const SimConst& sc = a.sc;
  const int n = sc.n, h = sc.h, NP = sc.NP, NS = sc.NS;
  const int tid = threadIdx.x;
  const size_t bbase = static_cast<size_t>(b) * kQ * NS;
  const double t = static_cast<double>(*a.step);
  const T avg = AvgArg<T>(sc.const_rho ? *a.avg_rho : 1.0);
  const BlockTab& bt = a.bt[b];
  const Origin borg{bt.origin[0], bt.origin[1], bt.origin[2], bt.cw};
  double ssum = 0.0;
  int scnt = 0;

  const int lx = tid >> 6, ly = (tid >> 3) & 7, lz = tid & 7;
  const int hx = tl.ox + lx, hy = tl.oy + ly, hz = tl.oz + lz;
  const bool active = hx < NP && hy < NP && hz < NP;
  const int own_l = ((lx + 1) * kBox + ly + 1) * kBox + lz + 1;
  const int d = (hx * NP + hy) * NP + hz;
  const int x = hx - h, y = hy - h, z = hz - h;
  const bool core = active && x >= 0 && x < n && y >= 0 && y < n && z >= 0 && z < n;

  T f[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) f[i] = T(0);
  int k = kBulk;
  if (active) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = a.src[bbase + static_cast<size_t>(i) * NS + d];
    if (core) {
      k = a.kind[static_cast<size_t>(b) * sc.ncore + (x * n + y) * n + z];
      CollideCell<T>(f, k, sc, borg, x, y, z, t, avg, a.want_stats != 0, ssum, scnt);
      kbox[own_l] = static_cast<uint8_t>(k);
    } else {
      kbox[own_l] = 254;
    }
  } else {
    kbox[own_l] = 255;
  }

  T g[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) g[i] = T(0);
  int ap_l = -1;
  if (tid < kApron) {
    ap_l = kApronTab.idx[tid];
    const int px = ap_l / (kBox * kBox), py = (ap_l / kBox) % kBox, pz = ap_l % kBox;
    const int gx = tl.ox - 1 + px, gy = tl.oy - 1 + py, gz = tl.oz - 1 + pz;
    if (gx >= 0 && gx < NP && gy >= 0 && gy < NP && gz >= 0 && gz < NP) {
      const int dg = (gx * NP + gy) * NP + gz;
#pragma unroll
      for (int i = 0; i < kQ; ++i) g[i] = a.src[bbase + static_cast<size_t>(i) * NS + dg];
      const int cx = gx - h, cy = gy - h, cz = gz - h;
      if (cx >= 0 && cx < n && cy >= 0 && cy < n && cz >= 0 && cz < n) {
        const int kk = a.kind[static_cast<size_t>(b) * sc.ncore + (cx * n + cy) * n + cz];
        double d0 = 0.0;
        int d1 = 0;
        CollideCell<T>(g, kk, sc, borg, cx, cy, cz, t, avg, false, d0, d1);
        kbox[ap_l] = static_cast<uint8_t>(kk);
      } else {
        kbox[ap_l] = 254;
      }
    } else {
      kbox[ap_l] = 255;
    }
  }

  T out[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    T* pl = plane[i & 1];
    pl[own_l] = f[i];
    if (ap_l >= 0) pl[ap_l] = g[i];
    __syncthreads();
    const int cxi = Cx(i), cyi = Cy(i), czi = Cz(i);
    const int box_off = (cxi * kBox + cyi) * kBox + czi;
    const int sx = hx - cxi, sy = hy - cyi, sz = hz - czi;
    T v = T(0);
    if (active) {
      if (core && PERCELL) {
        const bool src_int = sx - h >= 0 && sx - h < n && sy - h >= 0 && sy - h < n && sz - h >= 0 &&
                             sz - h < n;
        if (k == kSolid) {
          v = f[i];
        } else if (ReflectsOnPull(k) && !src_int) {
          v = f[Opp(i)];
        } else {
          const int sk = kbox[own_l - box_off];
          if (src_int && (sk == kSolid || sk == kBouzidi)) {
            v = f[Opp(i)];  // padded blocks carry no Bouzidi links: q = 1 -> half-way bounce-back
          } else {
            v = pl[own_l - box_off];
          }
        }
      } else if (sx >= 0 && sx < NP && sy >= 0 && sy < NP && sz >= 0 && sz < NP) {
        v = pl[own_l - box_off];
      } else {
        const int off = (cxi * NP + cyi) * NP + czi;
        int s = (d - off) % NS;
        if (s < 0) s += NS;
        v = a.src[bbase + static_cast<size_t>(i) * NS + s];
      }
    }
    out[i] = v;
  }
  if (active) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) a.dst[bbase + static_cast<size_t>(i) * NS + d] = out[i];
    if (core && (bt.flags & kFlagDrag) && a.snap_on) {
      const int c = (x * n + y) * n + z;
      const uint32_t dm = a.dmask[static_cast<size_t>(b) * sc.ncore + c];
      if (dm) {
        T* sp = a.snap + a.doff[static_cast<size_t>(b) * sc.ncore + c];
        int j = 0;
#pragma unroll
        for (int i = 0; i < kQ; ++i)
          if ((dm >> i) & 1u) sp[j++] = f[i];
      }
    }
  }
  if (a.want_stats == 1) {
    BlockReduceStats<T>(ssum, scnt, red_s, red_c, a.stats + 2 * ti);
  } else if (a.want_stats == 2 && core) {
    const size_t ord = static_cast<size_t>(ti / a.tiles_per_block) * sc.ncore + (x * n + y) * n + z;
    a.stat_val[ord] = ssum;
    a.stat_use[ord] = static_cast<uint8_t>(scnt);
  }
}

// ------------------------------------------------------------ small kernels
// Tree reduction of the per-tile partial sums (fast_stats: not the
// reference's summation order, but a fixed one): one workgroup per member,
// threads take strided tiles in order, then a pairwise reduction.
constexpr int kStatsThreads = 256;
__global__ __launch_bounds__(kStatsThreads) void KStatsFinalize(const double* partial, const int* tile_begin,
                                                                double* avg) {
  __shared__ double rs[kStatsThreads], rc[kStatsThreads];
  const int m = blockIdx.x, tid = threadIdx.x;
  double s = 0.0, c = 0.0;
  for (int i = tile_begin[m] + tid; i < tile_begin[m + 1]; i += kStatsThreads) {
    s += partial[2 * i];
    c += partial[2 * i + 1];
  }
  rs[tid] = s;
  rc[tid] = c;
  __syncthreads();
  for (int w = kStatsThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      rs[tid] += rs[tid + w];
      rc[tid] += rc[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) avg[m] = rc[0] > 0.0 ? rs[0] / rc[0] : 1.0;  // CollideRhoStats::average (block_lattice.h:31)
}

// CollideRhoStats accumulated in the reference's cell order (blocks of the
// level in order, cells (ix,iy,iz) lexicographic); one thread per member.
__global__ void KStatsSequential(const double* val, const uint8_t* use, const int* cell_begin, int members,
                                 double* avg) {
  const int m = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (m >= members) return;
  double s = 0.0;
  int c = 0;
  for (int k = cell_begin[m]; k < cell_begin[m + 1]; ++k) {
    if (use[k]) {
      s += val[k];
      ++c;
    }
  }
  avg[m] = c > 0 ? s / static_cast<double>(c) : 1.0;
}

__global__ void KStepInc(const int* step, int* step_next) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step_next = *step + 1;
}

template <typename T>
__global__ void KInitEq(T* pops, const uint8_t* kind, int nb, SimConst sc, double rho0, double u0x,
                        double u0y, double u0z, uint32_t mask) {
  const size_t gid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = static_cast<size_t>(nb) * sc.ncore;
  if (gid >= total) return;
  const int b = static_cast<int>(gid / sc.ncore), c = static_cast<int>(gid % sc.ncore);
  if (!((mask >> kind[gid]) & 1u)) return;
  const double u[3] = {u0x, u0y, u0z};
  const double usqr = NormSqr(u);
  const int slot = CoreSlot(sc, c);
  T* base = pops + static_cast<size_t>(b) * kQ * sc.plane;
  for (int i = 0; i < kQ; ++i)
    base[static_cast<size_t>(i) * sc.plane + slot] = static_cast<T>(Feq<double>(i, rho0, u, usqr));
}

template <typename T>
__global__ void KInitField(T* pops, int nb, SimConst sc, const double* rho, const double* u) {
  const size_t gid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = static_cast<size_t>(nb) * sc.ncore;
  if (gid >= total) return;
  const int b = static_cast<int>(gid / sc.ncore), c = static_cast<int>(gid % sc.ncore);
  const double uc[3] = {u[3 * gid], u[3 * gid + 1], u[3 * gid + 2]};
  const double usqr = NormSqr(uc);
  const int slot = CoreSlot(sc, c);
  T* base = pops + static_cast<size_t>(b) * kQ * sc.plane;
  for (int i = 0; i < kQ; ++i)
    base[static_cast<size_t>(i) * sc.plane + slot] = static_cast<T>(Feq<double>(i, rho[gid], uc, usqr));
}

// ------------------------------------------------------- post-stream FD
// ApplyInterpolatedVelocityBoundaryCells (interpolated_velocity.h:950-1036)
// processes FD cells in (ix,iy,iz) order in place.  Plane cells read only
// pre-pass populations, edge cells read plane cells' densities and corners
// read edge cells' densities (n >= 4), so three phases over all FD cells
// reproduce the sequential order: a read of an FD cell that precedes the
// reader in (ix,iy,iz) order takes its new value from `post`.
template <typename T>
struct FdView {
  const T* cur;
  const T* post;
  const int* smap;      // this block's FD cell -> entry (or -1)
  const T* lpre;        // LDS copy of the block ([19][512]) or nullptr
  const T* lpost;       // LDS new values per FD slot ([slot][19])
  const short* lsmap;   // LDS FD cell -> slot (or -1)
  const uint8_t* kind;
  const SimConst* sc;
  const double* org;
  double cw;
  double t;
  int self;

  __device__ bool InDomain(int x, int y, int z) const {
    const int n = sc->n;
    return x >= 0 && x < n && y >= 0 && y < n && z >= 0 && z < n;
  }
  // IsPaddingOverlapCell (interpolated_velocity.h:238-240): outside the block
  // but within min(h, 2) of it.  Farther reads (halo 1, mixed-sign edges of
  // ApplyEdgeBoundary) take the mirror path (:254-279, 407-412).
  __device__ bool Padding(int x, int y, int z) const {
    const int n = sc->n, st = sc->h < 2 ? sc->h : 2;
    return !InDomain(x, y, z) && x >= -st && x < n + st && y >= -st && y < n + st && z >= -st &&
           z < n + st;
  }
  __device__ static int Mirror(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i); }
  __device__ int Kind(int x, int y, int z) const { return kind[(x * sc->n + y) * sc->n + z]; }
  __device__ void Pops(int x, int y, int z, T* f) const {
    const int c = (x * sc->n + y) * sc->n + z;
    if (lpre) {
      const int e = lsmap[c];
      if (e >= 0 && c < self) {
#pragma unroll
        for (int i = 0; i < kQ; ++i) f[i] = lpost[e * kQ + i];
      } else {
#pragma unroll
        for (int i = 0; i < kQ; ++i) f[i] = lpre[i * kTileCells + c];
      }
      return;
    }
    const int e = smap[c];
    if (e >= 0 && c < self) {
#pragma unroll
      for (int i = 0; i < kQ; ++i) f[i] = post[static_cast<size_t>(e) * kQ + i];
    } else {
      const int s = CoreSlot(*sc, c);
#pragma unroll
      for (int i = 0; i < kQ; ++i) f[i] = cur[static_cast<size_t>(i) * sc->plane + s];
    }
  }
  __device__ void Prescribed(int x, int y, int z, T* u) const {
    double ud[3];
    PrescribedU(*sc, org, cw, x, y, z, t, ud);
    u[0] = static_cast<T>(ud[0]);
    u[1] = static_cast<T>(ud[1]);
    u[2] = static_cast<T>(ud[2]);
  }
  // BoundaryLatticeView::ComputeU (interpolated_velocity.h:376-401) with
  // MirrorExteriorVelocity (:266-279) and ExtrapolatedU (:297-316) unrolled
  // into one loop (no device recursion).
  __device__ void ComputeU(int x, int y, int z, T* u) const {
    const int n = sc->n;
    bool fx = false, fy = false, fz = false;
    bool check = true;
    for (;;) {
      if (check) {
        if (Padding(x, y, z)) {
          u[0] = u[1] = u[2] = T(0);
          break;
        }
        if (!InDomain(x, y, z)) {
          fx = x < 0 || x >= n;
          fy = y < 0 || y >= n;
          fz = z < 0 || z >= n;
          x = Mirror(x, n);
          y = Mirror(y, n);
          z = Mirror(z, n);
        }
      }
      const int k = Kind(x, y, z);
      if (k == kVelocity) {
        Prescribed(x, y, z, u);
        break;
      }
      if (k == kPressure) {
        int dir, ori;
        if (!FlatInfo(n, x, y, z, dir, ori)) {
          u[0] = u[1] = u[2] = T(0);
          break;
        }
        if (dir == 0) {
          x -= ori;
        } else if (dir == 1) {
          y -= ori;
        } else {
          z -= ori;
        }
        check = true;
        continue;
      }
      T f[kQ], rho;
      Pops(x, y, z, f);
      ProxyRhoU(f, rho, u);
      break;
    }
    if (fx) u[0] = -u[0];
    if (fy) u[1] = -u[1];
    if (fz) u[2] = -u[2];
  }
  __device__ T PrescribedRho(int x, int y, int z) const {
    int dir, ori;
    if (FlatInfo(sc->n, x, y, z, dir, ori)) return ROf<T>(sc->spec[dir * 2 + (ori < 0 ? 0 : 1)].rho_target);
    return RestR<T>();
  }
  __device__ T ComputeRho(int x, int y, int z) const {
    if (Padding(x, y, z)) return RestR<T>();
    if (!InDomain(x, y, z)) {
      const int n = sc->n;
      x = Mirror(x, n);
      y = Mirror(y, n);
      z = Mirror(z, n);
    }
    const int k = Kind(x, y, z);
    if (k == kPressure) return PrescribedRho(x, y, z);
    T f[kQ];
    Pops(x, y, z, f);
    if (k == kVelocity) {
      int dir, ori;
      if (FlatInfo(sc->n, x, y, z, dir, ori)) {
        T u[3];
        Prescribed(x, y, z, u);
        return VelRhoFromPop(dir, ori, f, u);
      }
    }
    T rho, u[3];
    ProxyRhoU(f, rho, u);
    return rho;
  }
  __device__ void ComputeRhoU(int x, int y, int z, T& rho, T* u) const {
    const int k = Kind(x, y, z);
    if (k == kVelocity) {
      Prescribed(x, y, z, u);
      rho = ComputeRho(x, y, z);
      return;
    }
    if (k == kPressure) {
      rho = PrescribedRho(x, y, z);
      int dir, ori;
      if (!FlatInfo(sc->n, x, y, z, dir, ori)) {
        u[0] = u[1] = u[2] = T(0);
        return;
      }
      ComputeU(x - (dir == 0 ? ori : 0), y - (dir == 1 ? ori : 0), z - (dir == 2 ? ori : 0), u);
      return;
    }
    T f[kQ];
    Pops(x, y, z, f);
    ProxyRhoU(f, rho, u);
  }
  __device__ void NormalGrad(int x, int y, int z, int axis, int normal, T* g) const {
    const int ox = axis == 0 ? -normal : 0, oy = axis == 1 ? -normal : 0, oz = axis == 2 ? -normal : 0;
    T u0[3], u1[3], u2[3];
    ComputeU(x, y, z, u0);
    ComputeU(x + ox, y + oy, z + oz, u1);
    ComputeU(x + 2 * ox, y + 2 * oy, z + 2 * oz, u2);
    for (int d = 0; d < 3; ++d)
      g[d] = T(-normal) * ((-T(3) * u0[d] + T(4) * u1[d] - T(1) * u2[d]) / T(2));
  }
  __device__ void TangGrad(int x, int y, int z, int tx, int ty, int tz, T* g) const {
    T up[3], um[3];
    ComputeU(x + tx, y + ty, z + tz, up);
    ComputeU(x - tx, y - ty, z - tz, um);
    for (int d = 0; d < 3; ++d) g[d] = (up[d] - um[d]) / T(2);
  }
};

template <typename T>
__device__ inline void WriteFromPi(T* f, T rho, const T* u, const T* pi) {
  const T usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
  for (int i = 0; i < kQ; ++i) f[i] = Feq(i, rho, u, usqr) + PiToFneq(i, pi);
}

// rho is r (rho for FP64, rho - 1 for FP32)
template <typename T>
__device__ inline void PiFromGrad(T rho, T omega, const T* dx, const T* dy, const T* dz, T* pi) {
  const T s = -RhoOf(rho) / T(3) / omega;
  pi[0] = T(2) * dx[0] * s;
  pi[3] = T(2) * dy[1] * s;
  pi[5] = T(2) * dz[2] * s;
  pi[1] = (dx[1] + dy[0]) * s;
  pi[2] = (dx[2] + dz[0]) * s;
  pi[4] = (dy[2] + dz[1]) * s;
}

// One FD boundary cell (ApplyInterpolatedVelocityBoundaryCells body).
template <typename T, typename V>
__device__ inline void FdCell(const V& v, int ix, int iy, int iz, T* fnew) {
  const SimConst& sc = *v.sc;
  const int n = sc.n;
  const T omega = static_cast<T>(sc.omega);
  const int xm = ix == 0, xp = ix == n - 1, ym = iy == 0, yp = iy == n - 1, zm = iz == 0, zp = iz == n - 1;
  const int fc = xm + xp + ym + yp + zm + zp;
  const int xn = xm ? -1 : (xp ? 1 : 0), yn = ym ? -1 : (yp ? 1 : 0), zn = zm ? -1 : (zp ? 1 : 0);
  T pi[6];
  if (fc == 3) {
    // ApplyCornerBoundary (:638-685)
    const T r100 = v.ComputeRho(ix - xn, iy, iz);
    const T r010 = v.ComputeRho(ix, iy - yn, iz);
    const T r001 = v.ComputeRho(ix, iy, iz - zn);
    const T r200 = v.ComputeRho(ix - 2 * xn, iy, iz);
    const T r020 = v.ComputeRho(ix, iy - 2 * yn, iz);
    const T r002 = v.ComputeRho(ix, iy, iz - 2 * zn);
    const T rho = T(4) / T(9) * (r001 + r010 + r100) - T(1) / T(9) * (r002 + r020 + r200);
    T dx[3], dy[3], dz[3], u[3];
    v.NormalGrad(ix, iy, iz, 0, xn, dx);
    v.NormalGrad(ix, iy, iz, 1, yn, dy);
    v.NormalGrad(ix, iy, iz, 2, zn, dz);
    v.ComputeU(ix, iy, iz, u);
    PiFromGrad(rho, omega, dx, dy, dz, pi);
    WriteFromPi(fnew, rho, u, pi);
  } else if (fc == 2) {
    // ApplyEdgeBoundary (:687-752)
    int plane, n1, n2;
    if (xn != 0 && yn != 0) {
      plane = 2;
      n1 = xn;
      n2 = yn;
    } else if (xn != 0 && zn != 0) {
      plane = 1;
      n1 = xn;
      n2 = zn;
    } else {
      plane = 0;
      n1 = yn;
      n2 = zn;
    }
    const int d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
    T rs[4];
    const int st[4][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}};
    for (int q = 0; q < 4; ++q) {
      int o[3] = {0, 0, 0};
      o[d1] = -n1 * st[q][0];
      o[d2] = -n2 * st[q][1];
      rs[q] = v.ComputeRho(ix + o[0], iy + o[1], iz + o[2]);
    }
    const T rho = T(2) / T(3) * (rs[1] + rs[0]) - T(1) / T(6) * (rs[3] + rs[2]);
    T g[3][3];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) g[p][q] = T(0);
    v.TangGrad(ix, iy, iz, plane == 0, plane == 1, plane == 2, g[plane]);
    v.NormalGrad(ix, iy, iz, d1, n1, g[d1]);
    v.NormalGrad(ix, iy, iz, d2, n2, g[d2]);
    T u[3];
    v.ComputeU(ix, iy, iz, u);
    const T s = -RhoOf(rho) / T(3) / omega;
    pi[0] = T(2) * g[0][0] * s;
    pi[3] = T(2) * g[1][1] * s;
    pi[5] = T(2) * g[2][2] * s;
    pi[1] = (g[0][1] + g[1][0]) * s;
    pi[2] = (g[0][2] + g[2][0]) * s;
    pi[4] = (g[1][2] + g[2][1]) * s;
    WriteFromPi(fnew, rho, u, pi);
  } else {
    // ApplyPlaneFdBoundary (:570-636)
    int face = OCTLB_XMIN;
    if (xp) {
      face = OCTLB_XMAX;
    } else if (yp) {
      face = OCTLB_YMAX;
    } else if (ym) {
      face = OCTLB_YMIN;
    } else if (zp) {
      face = OCTLB_ZMAX;
    } else if (zm) {
      face = OCTLB_ZMIN;
    }
    const int direction = face / 2, orientation = (face & 1) ? 1 : -1;
    T rho, u[3];
    v.ComputeRhoU(ix, iy, iz, rho, u);
    T dx[3] = {T(0), T(0), T(0)}, dy[3] = {T(0), T(0), T(0)}, dz[3] = {T(0), T(0), T(0)};
    if (direction == 0) {
      v.NormalGrad(ix, iy, iz, 0, orientation, dx);
      v.TangGrad(ix, iy, iz, 0, 1, 0, dy);
      v.TangGrad(ix, iy, iz, 0, 0, 1, dz);
    } else if (direction == 1) {
      v.TangGrad(ix, iy, iz, 1, 0, 0, dx);
      v.NormalGrad(ix, iy, iz, 1, orientation, dy);
      v.TangGrad(ix, iy, iz, 0, 0, 1, dz);
    } else {
      v.TangGrad(ix, iy, iz, 1, 0, 0, dx);
      v.TangGrad(ix, iy, iz, 0, 1, 0, dy);
      v.NormalGrad(ix, iy, iz, 2, orientation, dz);
    }
    PiFromGrad(rho, omega, dx, dy, dz, pi);
    WriteFromPi(fnew, rho, u, pi);
  }
}

template <typename T>
__global__ void KFd(FdArgs<T> a, int begin, int count) {
  BumpStep(a.step, a.step_next);
  const int e = begin + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= begin + count) return;
  const int b = a.cells[e].x, c = a.cells[e].y;
  const SimConst& sc = a.sc;
  const int n = sc.n;
  const int ix = c / (n * n), iy = (c / n) % n, iz = c % n;
  const int lvl = a.level[b];
  const T* cur = (a.par.v[lvl] ? a.pops1 : a.pops0) + static_cast<size_t>(b) * kQ * sc.plane;
  FdView<T> v;
  v.cur = cur;
  v.post = a.post;
  v.smap = a.smap + static_cast<size_t>(a.fd_index[b]) * sc.ncore;
  v.lpre = nullptr;
  v.lpost = nullptr;
  v.lsmap = nullptr;
  v.kind = a.kind + static_cast<size_t>(b) * sc.ncore;
  v.sc = &a.sc;
  v.org = a.bt[b].origin;
  v.cw = a.bt[b].cw;
  v.t = static_cast<double>(*a.step);
  v.self = c;
  T fnew[kQ];
  FdCell<T>(v, ix, iy, iz, fnew);
  T* po = a.post + static_cast<size_t>(e) * kQ;
  for (int i = 0; i < kQ; ++i) po[i] = fnew[i];
}

// Blocked mode: one workgroup per boundary block.  Every cell's density and
// velocity as the FD stencils see them (BoundaryLatticeView::ComputeRho /
// ComputeU, interpolated_velocity.h:352-434) are computed once into LDS, the
// pressure-extrapolation chains are resolved once, and the plane / edge /
// corner phases read only those arrays.  A phase stores the density of each
// cell it rewrote, which later phases read for cells that precede them in
// (ix,iy,iz) order.
template <typename T>
struct FdLds {
  const T* rho_pre;
  const T* rho_post;
  const T* ux;
  const T* uy;
  const T* uz;
  const short* slot;
  int self;

  __device__ bool InDomain(int x, int y, int z) const {
    constexpr int n = kTile;
    return x >= 0 && x < n && y >= 0 && y < n && z >= 0 && z < n;
  }
  __device__ bool Padding(int x, int y, int z) const {
    constexpr int n = kTile, st = 1;  // blocked mode: 8^3 cells, halo 1 -> min(h, 2) = 1
    return !InDomain(x, y, z) && x >= -st && x < n + st && y >= -st && y < n + st && z >= -st &&
           z < n + st;
  }
  __device__ static int Mirror(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * (n - 1) - i : i); }
  __device__ void ComputeU(int x, int y, int z, T* u) const {
    constexpr int n = kTile;
    if (Padding(x, y, z)) {
      u[0] = u[1] = u[2] = T(0);
      return;
    }
    const bool fx = x < 0 || x >= n, fy = y < 0 || y >= n, fz = z < 0 || z >= n;
    const int c = (Mirror(x, n) * n + Mirror(y, n)) * n + Mirror(z, n);
    u[0] = fx ? -ux[c] : ux[c];
    u[1] = fy ? -uy[c] : uy[c];
    u[2] = fz ? -uz[c] : uz[c];
  }
  __device__ T ComputeRho(int x, int y, int z) const {
    constexpr int n = kTile;
    if (Padding(x, y, z)) return RestR<T>();
    const int c = (Mirror(x, n) * n + Mirror(y, n)) * n + Mirror(z, n);
    return (slot[c] >= 0 && c < self) ? rho_post[c] : rho_pre[c];
  }
  __device__ void ComputeRhoU(int x, int y, int z, T& rho, T* u) const {
    const int c = (x * kTile + y) * kTile + z;
    rho = ComputeRho(x, y, z);
    u[0] = ux[c];
    u[1] = uy[c];
    u[2] = uz[c];
  }
  __device__ void NormalGrad(int x, int y, int z, int axis, int normal, T* g) const {
    const int ox = axis == 0 ? -normal : 0, oy = axis == 1 ? -normal : 0, oz = axis == 2 ? -normal : 0;
    T u0[3], u1[3], u2[3];
    ComputeU(x, y, z, u0);
    ComputeU(x + ox, y + oy, z + oz, u1);
    ComputeU(x + 2 * ox, y + 2 * oy, z + 2 * oz, u2);
    for (int d = 0; d < 3; ++d)
      g[d] = T(-normal) * ((-T(3) * u0[d] + T(4) * u1[d] - T(1) * u2[d]) / T(2));
  }
  __device__ void TangGrad(int x, int y, int z, int tx, int ty, int tz, T* g) const {
    T up[3], um[3];
    ComputeU(x + tx, y + ty, z + tz, up);
    ComputeU(x - tx, y - ty, z - tz, um);
    for (int d = 0; d < 3; ++d) g[d] = (up[d] - um[d]) / T(2);
  }
};

// ComputeRho of an in-domain cell from its populations (:403-434).
template <typename T>
__device__ inline T FdRhoOf(int k, const T* f, const T* pu, const SimConst& sc, int ix, int iy, int iz) {
  int dir, ori;
  const bool flat = FlatInfo(sc.n, ix, iy, iz, dir, ori);
  if (k == kPressure) return flat ? ROf<T>(sc.spec[dir * 2 + (ori < 0 ? 0 : 1)].rho_target) : RestR<T>();
  if (k == kVelocity && flat) return VelRhoFromPop(dir, ori, f, pu);
  T rho, u[3];
  ProxyRhoU(f, rho, u);
  return rho;
}

// Runtime-index forms of Feq / PiToFneq for the (cell, population) task
// loops; same operations in the same order as the compile-time versions.
__constant__ int kCxT[kQ] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
__constant__ int kCyT[kQ] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
__constant__ int kCzT[kQ] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
__constant__ int kWDenT[kQ] = {3, 18, 18, 18, 36, 36, 36, 36, 36, 36, 18, 18, 18, 36, 36, 36, 36, 36, 36};

template <typename T>
__device__ inline T FdPopulation(const int* c, T t, T rho, const T* u, T usqr, const T* pi) {
  T cu = T(0);
  cu += T(c[0]) * u[0];
  cu += T(c[1]) * u[1];
  cu += T(c[2]) * u[2];
  T feq;
  if constexpr (kShiftedArith<T>) {  // rho is rho - 1 (shifted FP32 arithmetic)
    feq = t * rho + (T(1) + rho) * t * (T(3) * cu + T(3) * T(3) * T(0.5) * cu * cu - T(3) * T(0.5) * usqr);
  } else {
    feq = rho * t * (T(1) + T(3) * cu + T(3) * T(3) * T(0.5) * cu * cu - T(3) * T(0.5) * usqr) - t;
  }
  T fneq = T(0);
  int ipi = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a; b < 3; ++b) {
      T to_add = T(c[a] * c[b]);
      if (a == b) {
        to_add -= T(1) / T(3);
      } else {
        to_add *= T(2);
      }
      to_add *= pi[ipi++];
      fneq += to_add;
    }
  }
  fneq *= t * T(3) * T(3) / T(2);
  return feq + fneq;
}

// Blocked mode: one workgroup per boundary block.  Every cell's density and
// velocity as the FD stencils see them (BoundaryLatticeView::ComputeRho /
// ComputeU, interpolated_velocity.h:352-434) are computed once into LDS and
// the pressure-extrapolation chains resolved once.  FD cells are then spread
// over the threads (one per cell, plane|edge|corner order): all gradients are
// taken at once; per phase the cells of that phase get their density (edges
// and corners extrapolate densities that earlier phases rewrote, as the
// reference's in-place (ix,iy,iz) sweep does), and the 19 populations of every
// cell of the phase are written by all 512 threads as (cell, population) tasks.
// LDS of one FD block pass over at most MaxFd FD cells; carved from the
// caller's shared buffer so the pass can also run inside the fused kernel's
// workgroups (aliasing its box).
template <typename T, int MaxFd>
struct FdSmem {
  T rho_pre[kTileCells], rho_post[kTileCells];
  T ux[kTileCells], uy[kTileCells], uz[kTileCells];
  T crho[MaxFd], cusq[MaxFd];
  T cu[MaxFd][3], cpi[MaxFd][6];
  T fbuf[MaxFd][kQ];
  T wtab[kQ];
  int ctab[kQ][3];
  short slot[kTileCells];
  short lst[MaxFd];
  uint8_t kind[kTileCells];
};

// One FD block pass by a workgroup of 512 or more threads (threads beyond
// the block's 512 cells take part in the barriers and the task loop only).
// The pass over FD block fb (block b) whose populations are
// cur[i * plane + cell] (plane = 512 in the production layout), at time t.
// kEarly: every global input is loaded before the first barrier (the FD
// cell list and phase ends with the populations); off where the pass runs
// inside the fused kernel, whose register budget the longer live ranges
// would exceed.
template <typename T, int MaxFd = kMaxFdCells, bool kEarly = false>
__device__ inline void FdBlockCore(const FdArgs<T>& a, const short* fd_slot, int fb, int b, T* cur, int plane,
                                   double t, FdSmem<T, MaxFd>& sm) {
  T* rho_pre = sm.rho_pre;
  T* rho_post = sm.rho_post;
  T* ux = sm.ux;
  T* uy = sm.uy;
  T* uz = sm.uz;
  short* slot = sm.slot;
  uint8_t* kind = sm.kind;
  T* crho = sm.crho;
  T* cusq = sm.cusq;
  T(*cu)[3] = sm.cu;
  T(*cpi)[6] = sm.cpi;
  T(*fbuf)[kQ] = sm.fbuf;
  short* lst = sm.lst;
  int(*ctab)[3] = sm.ctab;
  T* wtab = sm.wtab;
  const bool act = threadIdx.x < kTileCells;
  if (threadIdx.x < kQ) {
    // velocity set and weights for the (cell, population) task loop
    const int i = threadIdx.x;
    ctab[i][0] = kCxT[i];
    ctab[i][1] = kCyT[i];
    ctab[i][2] = kCzT[i];
    wtab[i] = T(1) / T(kWDenT[i]);
  }
  const int tid = threadIdx.x;
  const SimConst& sc = *(const SimConst*)&a.scm[a.bt[b].member];
  // the FD cells in plane|edge|corner order and the phase ends, issued with
  // the population loads; the list reaches LDS after them
  int4 cnt = make_int4(0, 0, 0, 0);
  short lraw = -1;
  if constexpr (kEarly) {
    cnt = a.fd_count[fb];
    if (tid < MaxFd) lraw = a.fd_list[static_cast<size_t>(fb) * kTileCells + tid];
  }
  if (act) {
    const int ix = tid >> 6, iy = (tid >> 3) & 7, iz = tid & 7;
    const int k = a.kind[static_cast<size_t>(b) * kTileCells + tid];
    kind[tid] = static_cast<uint8_t>(k);
    slot[tid] = fd_slot[static_cast<size_t>(fb) * kTileCells + tid];
    const Origin og{a.bt[b].origin[0], a.bt[b].origin[1], a.bt[b].origin[2], a.bt[b].cw};
    T f[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = cur[static_cast<size_t>(i) * plane + tid];
    T pu[3] = {T(0), T(0), T(0)};
    if (k == kVelocity) {
      const double org[3] = {og.x, og.y, og.z};
      double ud[3];
      PrescribedU(sc, org, og.cw, ix, iy, iz, t, ud);
      pu[0] = static_cast<T>(ud[0]);
      pu[1] = static_cast<T>(ud[1]);
      pu[2] = static_cast<T>(ud[2]);
    }
    rho_pre[tid] = FdRhoOf(k, f, pu, sc, ix, iy, iz);
    if (k == kVelocity) {
      ux[tid] = pu[0];
      uy[tid] = pu[1];
      uz[tid] = pu[2];
    } else {
      T rho, u[3];
      ProxyRhoU(f, rho, u);
      ux[tid] = u[0];
      uy[tid] = u[1];
      uz[tid] = u[2];
    }
  }
  if constexpr (kEarly) {
    if (tid < MaxFd) lst[tid] = lraw;
  }
  __syncthreads();
  {
    // ExtrapolatedU (:297-316): a pressure cell takes the velocity of the
    // first non-pressure cell inward along its face normal; only pressure
    // entries are rewritten, and walks only read non-pressure entries.
    const int k = act ? kind[tid] : kBulk;
    T pux = T(0), puy = T(0), puz = T(0);
    if (k == kPressure) {
      int x = tid >> 6, y = (tid >> 3) & 7, z = tid & 7;
      for (;;) {
        const int c = (x * kTile + y) * kTile + z;
        if (kind[c] != kPressure) {
          pux = ux[c];
          puy = uy[c];
          puz = uz[c];
          break;
        }
        int dir, ori;
        if (!FlatInfo(kTile, x, y, z, dir, ori)) break;
        if (dir == 0) {
          x -= ori;
        } else if (dir == 1) {
          y -= ori;
        } else {
          z -= ori;
        }
        if (x < 0 || x >= kTile || y < 0 || y >= kTile || z < 0 || z >= kTile) break;
      }
    }
    __syncthreads();
    if (k == kPressure) {
      ux[tid] = pux;
      uy[tid] = puy;
      uz[tid] = puz;
    }
  }
  __syncthreads();

  // ---- gradients of the FD cell this thread owns
  if constexpr (!kEarly) cnt = a.fd_count[fb];
  const int ph_end[3] = {cnt.x, cnt.y, cnt.z};
  // plane, edge and corner cells start on separate waves, so the three
  // stencil shapes run side by side instead of one after another in a wave
  const int t1 = (cnt.x + 63) & ~63, t2 = (t1 + (cnt.y - cnt.x) + 63) & ~63;
  int e = -1;
  if (tid < cnt.x) {
    e = tid;
  } else if (tid >= t1 && tid < t1 + (cnt.y - cnt.x)) {
    e = cnt.x + (tid - t1);
  } else if (tid >= t2 && tid < t2 + (cnt.z - cnt.y)) {
    e = cnt.y + (tid - t2);
  }
  int my = -1, myph = 3;
  T G[3][3];
  T myu[3] = {T(0), T(0), T(0)};
  int fc = 0, xn = 0, yn = 0, zn = 0, ix = 0, iy = 0, iz = 0;
  if (e >= 0) {
    if constexpr (kEarly) {
      my = lst[e];
    } else {
      my = a.fd_list[static_cast<size_t>(fb) * kTileCells + e];
      lst[e] = static_cast<short>(my);
    }
    myph = e < cnt.x ? 0 : (e < cnt.y ? 1 : 2);
    ix = my >> 6;
    iy = (my >> 3) & 7;
    iz = my & 7;
    FdLds<T> v{rho_pre, rho_post, ux, uy, uz, slot, my};
    const int xm = ix == 0, xp = ix == kTile - 1, ym = iy == 0, yp = iy == kTile - 1, zm = iz == 0,
              zp = iz == kTile - 1;
    fc = xm + xp + ym + yp + zm + zp;
    xn = xm ? -1 : (xp ? 1 : 0);
    yn = ym ? -1 : (yp ? 1 : 0);
    zn = zm ? -1 : (zp ? 1 : 0);
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) G[p][q] = T(0);
    if (fc == 3) {
      // ApplyCornerBoundary (:638-685)
      v.NormalGrad(ix, iy, iz, 0, xn, G[0]);
      v.NormalGrad(ix, iy, iz, 1, yn, G[1]);
      v.NormalGrad(ix, iy, iz, 2, zn, G[2]);
    } else if (fc == 2) {
      // ApplyEdgeBoundary (:687-752)
      int plane, n1, n2;
      if (xn != 0 && yn != 0) {
        plane = 2;
        n1 = xn;
        n2 = yn;
      } else if (xn != 0 && zn != 0) {
        plane = 1;
        n1 = xn;
        n2 = zn;
      } else {
        plane = 0;
        n1 = yn;
        n2 = zn;
      }
      const int d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
      v.TangGrad(ix, iy, iz, plane == 0, plane == 1, plane == 2, G[plane]);
      v.NormalGrad(ix, iy, iz, d1, n1, G[d1]);
      v.NormalGrad(ix, iy, iz, d2, n2, G[d2]);
    } else {
      // ApplyPlaneFdBoundary (:570-636)
      int face = OCTLB_XMIN;
      if (xp) {
        face = OCTLB_XMAX;
      } else if (yp) {
        face = OCTLB_YMAX;
      } else if (ym) {
        face = OCTLB_YMIN;
      } else if (zp) {
        face = OCTLB_ZMAX;
      } else if (zm) {
        face = OCTLB_ZMIN;
      }
      const int direction = face / 2, orientation = (face & 1) ? 1 : -1;
      if (direction == 0) {
        v.NormalGrad(ix, iy, iz, 0, orientation, G[0]);
        v.TangGrad(ix, iy, iz, 0, 1, 0, G[1]);
        v.TangGrad(ix, iy, iz, 0, 0, 1, G[2]);
      } else if (direction == 1) {
        v.TangGrad(ix, iy, iz, 1, 0, 0, G[0]);
        v.NormalGrad(ix, iy, iz, 1, orientation, G[1]);
        v.TangGrad(ix, iy, iz, 0, 0, 1, G[2]);
      } else {
        v.TangGrad(ix, iy, iz, 1, 0, 0, G[0]);
        v.TangGrad(ix, iy, iz, 0, 1, 0, G[1]);
        v.NormalGrad(ix, iy, iz, 2, orientation, G[2]);
      }
    }
    v.ComputeU(ix, iy, iz, myu);
  }

  const T omega = static_cast<T>(sc.omega);
  for (int ph = 0; ph < 3; ++ph) {
    const int begin = ph == 0 ? 0 : ph_end[ph - 1], end = ph_end[ph];
    if (begin == end) continue;  // uniform: the counts are per block
    if (myph == ph) {
      FdLds<T> v{rho_pre, rho_post, ux, uy, uz, slot, my};
      T rho;
      if (fc == 3) {
        const T r100 = v.ComputeRho(ix - xn, iy, iz);
        const T r010 = v.ComputeRho(ix, iy - yn, iz);
        const T r001 = v.ComputeRho(ix, iy, iz - zn);
        const T r200 = v.ComputeRho(ix - 2 * xn, iy, iz);
        const T r020 = v.ComputeRho(ix, iy - 2 * yn, iz);
        const T r002 = v.ComputeRho(ix, iy, iz - 2 * zn);
        rho = T(4) / T(9) * (r001 + r010 + r100) - T(1) / T(9) * (r002 + r020 + r200);
      } else if (fc == 2) {
        int plane, n1, n2;
        if (xn != 0 && yn != 0) {
          plane = 2;
          n1 = xn;
          n2 = yn;
        } else if (xn != 0 && zn != 0) {
          plane = 1;
          n1 = xn;
          n2 = zn;
        } else {
          plane = 0;
          n1 = yn;
          n2 = zn;
        }
        const int d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
        T rs[4];
        const int st[4][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}};
        for (int q = 0; q < 4; ++q) {
          int o[3] = {0, 0, 0};
          o[d1] = -n1 * st[q][0];
          o[d2] = -n2 * st[q][1];
          rs[q] = v.ComputeRho(ix + o[0], iy + o[1], iz + o[2]);
        }
        rho = T(2) / T(3) * (rs[1] + rs[0]) - T(1) / T(6) * (rs[3] + rs[2]);
      } else {
        rho = v.ComputeRho(ix, iy, iz);
      }
      T pi[6];
      PiFromGrad(rho, omega, G[0], G[1], G[2], pi);
      crho[e] = rho;
      cusq[e] = myu[0] * myu[0] + myu[1] * myu[1] + myu[2] * myu[2];
      cu[e][0] = myu[0];
      cu[e][1] = myu[1];
      cu[e][2] = myu[2];
#pragma unroll
      for (int q = 0; q < 6; ++q) cpi[e][q] = pi[q];
    }
    __syncthreads();
    // WriteFromPi (:446-453) as (cell, population) tasks
    const int tasks = (end - begin) * kQ;
    for (int w = tid; w < tasks; w += static_cast<int>(blockDim.x)) {
      const int te = begin + w / kQ, i = w % kQ;
      const T f = FdPopulation<T>(ctab[i], wtab[i], crho[te], cu[te], cusq[te], cpi[te]);
      fbuf[te][i] = f;
      cur[static_cast<size_t>(i) * plane + lst[te]] = f;
    }
    __syncthreads();
    if (myph == ph) {
      const int k = kind[my];
      const T pu[3] = {ux[my], uy[my], uz[my]};
      rho_post[my] = FdRhoOf(k, fbuf[e], pu, sc, ix, iy, iz);
    }
    __syncthreads();
  }
}

// The pass over FD block fb as the fused kernel runs it (its register
// budget is tight: the block and its level are loaded here, in this order).
template <typename T>
__device__ inline void FdBlockBodyAt(const FdArgs<T>& a, const int* fd_blocks, const short* fd_slot, int fb,
                                     FdSmem<T>& sm) {
  const int b = fd_blocks[fb];
  T* cur = (a.par.v[a.level[b]] ? a.pops1 : a.pops0) + static_cast<size_t>(b) * kQ * kTileCells;
  FdBlockCore<T>(a, fd_slot, fb, b, cur, kTileCells, static_cast<double>(*a.step), sm);
}

// d = (FD block, block, its level, 0): one load gives everything the
// population loads depend on.
template <typename T, int MaxFd, bool kEarly>
__device__ inline void FdBlockBody(const FdArgs<T>& a, const short* fd_slot, int4 d, FdSmem<T, MaxFd>& sm) {
  T* cur = (a.par.v[d.z] ? a.pops1 : a.pops0) + static_cast<size_t>(d.y) * kQ * kTileCells;
  FdBlockCore<T, MaxFd, kEarly>(a, fd_slot, d.x, d.y, cur, kTileCells, static_cast<double>(*a.step), sm);
}

// MaxFd: the most FD cells any block of the launch holds (the LDS the pass
// needs, and so how many workgroups share a CU).
// Small classes: FP32 at 64 VGPRs (8 waves per SIMD, four 512-thread
// workgroups per CU, 20-28 KB LDS each); FP64 at 4 waves per SIMD, the most
// it reaches without spilling (two workgroups per CU).
template <typename T, int MaxFd>
struct FdOccupancy {
  static constexpr int kWavesPerSimd = MaxFd == kMaxFdCells ? 1 : (sizeof(T) == 4 ? 8 : 4);
};

template <typename T, int MaxFd>
__global__ __launch_bounds__(512, (FdOccupancy<T, MaxFd>::kWavesPerSimd)) void KFdBlock(
    FdArgs<T> a, const short* fd_slot, const int4* fd_idx) {
  __shared__ FdSmem<T, MaxFd> sm;
  BumpStep(a.step, a.step_next);
  FdBlockBody<T, MaxFd, true>(a, fd_slot, fd_idx[blockIdx.x], sm);
}

template <typename T>
__global__ void KFdCommit(FdArgs<T> a, int count) {
  const int e = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= count) return;
  const int b = a.cells[e].x, c = a.cells[e].y;
  const SimConst& sc = a.sc;
  T* cur = (a.par.v[a.level[b]] ? a.pops1 : a.pops0) + static_cast<size_t>(b) * kQ * sc.plane;
  const int s = CoreSlot(sc, c);
  const T* po = a.post + static_cast<size_t>(e) * kQ;
  for (int i = 0; i < kQ; ++i) cur[static_cast<size_t>(i) * sc.plane + s] = po[i];
}

// ------------------------------------------------------ Lagrava coupling
// ReadMacro (level_coupler.cpp:183-189): olb computeRhoU + computeFneq.
template <typename T>
__device__ inline void ReadMacro(const T* base, int plane, int c, T* m) {
  T f[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) f[i] = base[static_cast<size_t>(i) * plane + c];
  T rho, u[3];
  OlbRhoU(f, rho, u);
  const T usqr = NormSqr(u);
  m[0] = rho;
  m[1] = u[0];
  m[2] = u[1];
  m[3] = u[2];
#pragma unroll
  for (int i = 0; i < kQ; ++i) m[4 + i] = f[i] - Feq(i, rho, u, usqr);
}

// WriteProlongation (level_coupler.cpp:191-201)
template <typename T>
__device__ inline void WriteMacro(T* base, int plane, int c, const T* m, T scaling) {
  const T u[3] = {m[1], m[2], m[3]};
  const T usqr = NormSqr(u);
#pragma unroll
  for (int i = 0; i < kQ; ++i) base[static_cast<size_t>(i) * plane + c] = Feq(i, m[0], u, usqr) + scaling * m[4 + i];
}

template <typename T>
__global__ void KCouplerInit(CoupleArgs<T> a, int npts) {
  const int p = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= npts) return;
  const int* r = a.plan + 12 * p;
  const int cl = r[9];
  const T* cb = (a.par.v[cl] ? a.pops1 : a.pops0) + static_cast<size_t>(r[0]) * kQ * kTileCells;
  ReadMacro(cb, kTileCells, (r[1] * kTile + r[2]) * kTile + r[3], a.prev + static_cast<size_t>(p) * kMacro);
}

// ApplyProlongation (level_coupler.cpp:438-478) over the deduplicated points
// of one coarse level (a later point writing the same fine cell wins).
template <typename T>
__device__ inline void ProlongPoint(const CoupleArgs<T>& a, const int* list, int count, int half, int j) {
  if (j >= count) return;
  const int p = list[j];
  const int* r = a.plan + 12 * p;
  const int cl = r[9], fl = r[10];
  const T* cb = (a.par.v[cl] ? a.pops1 : a.pops0) + static_cast<size_t>(r[0]) * kQ * kTileCells;
  T* fb = (a.par.v[fl] ? a.pops1 : a.pops0) + static_cast<size_t>(r[4]) * kQ * kTileCells;
  T m[kMacro];
  const int cc = (r[1] * kTile + r[2]) * kTile + r[3];
  ReadMacro(cb, kTileCells, cc, m);
  T* pv = a.prev + static_cast<size_t>(p) * kMacro;
  if (half) {
    for (int q = 0; q < kMacro; ++q) m[q] = T(0.5) * (pv[q] + m[q]);
  }
  WriteMacro(fb, kTileCells, (r[5] * kTile + r[6]) * kTile + r[7], m, a.prolong_scale);
  if (!half) {
    for (int q = 0; q < kMacro; ++q) pv[q] = m[q];
  }
}

template <typename T>
__global__ void KProlong(CoupleArgs<T> a, const int* list, int count, int half) {
  BumpStep(a.step, a.step_next);
  ProlongPoint<T>(a, list, count, half, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x));
}

// LevelCoupler::restrict (level_coupler.cpp:613-654) over the deduplicated
// points of one fine level (a later point writing the same coarse cell wins).
// Restriction (level_coupler.cpp:613-654).  A 32-lane group per coupling
// point: lane j < 19 reads the macroscopic state of the point's fine cell
// (j = 0) or of its D3Q19 neighbour j when that lies inside the fine block;
// after a barrier lane q < 19 averages f_neq population q over those cells in
// the reference's neighbour order and writes population q of the coarse cell.
constexpr int kRestrictGroups = 8;

// Groups = threads per workgroup / 32; wg = this workgroup's index among the
// restriction workgroups.
template <typename T>
struct RestrictLds {
  T fneq[kQ][kQ + 1];
  T mac[4];
};

// Groups 32-lane groups, group g working in sm[g].
template <typename T, int Groups>
__device__ inline void RestrictBodyIn(const CoupleArgs<T>& a, const int* list, int count, int wg,
                                      RestrictLds<T>* sm) {
  const int g = threadIdx.x >> 5, lane = threadIdx.x & 31;
  T(*fneq)[kQ + 1] = sm[g].fneq;
  T* mac = sm[g].mac;
  const int j = wg * Groups + g;
  const bool active = j < count && lane < kQ;
  int fi = 0, fj = 0, fk = 0;
  const int* r = nullptr;
  if (active) {
    r = a.plan + 12 * list[j];
    fi = r[5];
    fj = r[6];
    fk = r[7];
    const int x = fi + Cx(lane), y = fj + Cy(lane), z = fk + Cz(lane);
    if (x >= 0 && x < kTile && y >= 0 && y < kTile && z >= 0 && z < kTile) {
      const T* fb = (a.par.v[r[10]] ? a.pops1 : a.pops0) + static_cast<size_t>(r[4]) * kQ * kTileCells;
      T m[kMacro];
      ReadMacro(fb, kTileCells, (x * kTile + y) * kTile + z, m);
#pragma unroll
      for (int q = 0; q < kQ; ++q) fneq[lane][q] = m[4 + q];
      if (lane == 0) {
        mac[0] = m[0];
        mac[1] = m[1];
        mac[2] = m[2];
        mac[3] = m[3];
      }
    }
  }
  __syncthreads();
  if (!active) return;
  const int q = lane;
  T sum = T(0);
  int cnt = 0;
  for (int jp = 1; jp < kQ; ++jp) {
    const int x = fi + Cx(jp), y = fj + Cy(jp), z = fk + Cz(jp);
    if (x < 0 || x >= kTile || y < 0 || y >= kTile || z < 0 || z >= kTile) continue;
    sum += fneq[jp][q];
    ++cnt;
  }
  const T avg = (fneq[0][q] + sum) / static_cast<T>(cnt + 1);
  const T u[3] = {mac[1], mac[2], mac[3]};
  const T usqr = NormSqr(u);
  T* cb = (a.par.v[r[9]] ? a.pops1 : a.pops0) + static_cast<size_t>(r[0]) * kQ * kTileCells;
  cb[static_cast<size_t>(q) * kTileCells + (r[1] * kTile + r[2]) * kTile + r[3]] =
      Feq(q, mac[0], u, usqr) + a.restrict_scale * avg;
}

template <typename T, int Groups>
__device__ inline void RestrictBody(const CoupleArgs<T>& a, const int* list, int count, int wg) {
  __shared__ RestrictLds<T> sm[Groups];
  RestrictBodyIn<T, Groups>(a, list, count, wg, sm);
}

template <typename T>
__global__ __launch_bounds__(256) void KRestrict(CoupleArgs<T> a, const int* list, int count) {
  BumpStep(a.step, a.step_next);
  RestrictBody<T, kRestrictGroups>(a, list, count, static_cast<int>(blockIdx.x));
}

// FD boundary pass of the level just streamed, merged into one launch with
// the coupling step that follows it (prolongation or restriction) when the
// host has checked that the FD blocks and the coupled blocks are disjoint:
// workgroups [0, nfd) run the FD pass, the rest the coupling points.
constexpr int kMergedGroups = 16;  // 512 threads / 32-lane restriction groups

template <typename T>
__global__ __launch_bounds__(512) void KFdCouple(FdArgs<T> fa, const short* fd_slot, const int4* fd_idx, int nfd,
                                                 CoupleArgs<T> ca, const int* list,
                                                 int count, int mode) {
  __shared__ FdSmem<T> sm;
  BumpStep(fa.step, fa.step_next);
  const int w = static_cast<int>(blockIdx.x);
  if (w < nfd) {
    FdBlockBody<T, kMaxFdCells, true>(fa, fd_slot, fd_idx[w], sm);
  } else if (mode == 2) {
    RestrictBody<T, kMergedGroups>(ca, list, count, w - nfd);
  } else {
    ProlongPoint<T>(ca, list, count, mode, (w - nfd) * 512 + static_cast<int>(threadIdx.x));
  }
}

// ------------------------------------------------------ drag and mass
// MomentumExchangeDrag::force_on_fluid_if (drag_postprocessor.h:66-115):
// one term per wall link, summed per block in the reference's loop order
// (iz, iy, ix, iPop) and over blocks in block order, all in FP64.
// MomentumExchangeDrag::force_on_fluid_if of one block per workgroup
// (drag_postprocessor.h:66-115): the threads evaluate the block's link terms
// (f_out + f_bar times c_i, the bbox predicate of cylinder3d_case.h:483-490)
// into LDS, then one thread sums them in the reference's loop order --
// bit-equal to the sequential loop, with only LDS reads (no table lookup)
// on the dependent summation chain.
constexpr int kDragThreads = 256;
constexpr int kDragChunk = 2048;

// The cylinder case's sum of the per-block forces in block order
// (cylinder3d_case.h:492-497), staged through LDS so the sequential chain
// reads no global memory.  Blocks without links add -0.0, which leaves any
// sum unchanged, so only blocks with links are listed.
// st: kDragChunk doubles of LDS.
__device__ inline void SumBlocks3(const double* v, int count, double* out, double* st) {
  constexpr int kPer = kDragChunk / 3;
  const int tid = threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int c0 = 0; c0 < count; c0 += kPer) {
    const int m = min(kPer, count - c0);
    for (int e = tid; e < 3 * m; e += kDragThreads) st[e] = v[3 * c0 + e];
    __syncthreads();
    if (tid == 0)
      for (int e = 0; e < m; ++e) {
        s0 += st[3 * e];
        s1 += st[3 * e + 1];
        s2 += st[3 * e + 2];
      }
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
  }
}

// One workgroup per block with drag links; the last workgroup to finish
// (a device-scope counter, reset by it) sums the block forces and writes
// the result to `result`, host-visible pinned memory.
template <typename T>
__global__ __launch_bounds__(kDragThreads) void KDragBlock(DragArgs<T> a, const int* block_begin, double* out,
                                                           unsigned int* done, int nblk, double* result) {
  // per link: the three products t * c_i (the reference's operands, so the
  // sequential sum below adds exactly what its loop adds), +0.0 for a link
  // outside the bbox predicate: a sum that starts at +0.0 is never -0.0
  // under round-to-nearest, so adding +0.0 leaves it exactly as skipping does
  __shared__ double tx[kDragChunk], ty[kDragChunk], tz[kDragChunk];
  const int j = blockIdx.x, tid = threadIdx.x;
  const int begin = block_begin[j], end = block_begin[j + 1];
  const SimConst& sc = a.sc;
  const int n = sc.n;
  double fo0 = 0.0, fo1 = 0.0, fo2 = 0.0;
  for (int c0 = begin; c0 < end; c0 += kDragChunk) {
    const int m = min(kDragChunk, end - c0);
    for (int e = tid; e < m; e += kDragThreads) {
      const int4 L = a.links[c0 + e];  // block, cell, pop, snapshot index
      const int b = L.x, c = L.y, i = L.z;
      const int ix = c / (n * n), iy = (c / n) % n, iz = c % n;
      const double* org = a.bt[b].origin;
      const double cw = a.bt[b].cw;
      const double px = org[0] + (ix + kCxT[i] + 0.5) * cw;
      const double py = org[1] + (iy + kCyT[i] + 0.5) * cw;
      const double pz = org[2] + (iz + kCzT[i] + 0.5) * cw;
      const bool inside = px >= a.bbox[0] && px <= a.bbox[3] && py >= a.bbox[1] && py <= a.bbox[4] &&
                          pz >= a.bbox[2] && pz <= a.bbox[5];
      const T* live = (a.par.v[a.bt[b].level] ? a.pops1 : a.pops0) + static_cast<size_t>(b) * kQ * sc.plane;
      const int sl = CoreSlot(sc, c);
      const int opp = i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9);
      const double fbar = static_cast<double>(live[static_cast<size_t>(opp) * sc.plane + sl]);
      const double fout =
          a.snap ? static_cast<double>(a.snap[L.w]) : static_cast<double>(live[static_cast<size_t>(i) * sc.plane + sl]);
      const double t = fout + fbar;
      tx[e] = inside ? t * static_cast<double>(kCxT[i]) : 0.0;
      ty[e] = inside ? t * static_cast<double>(kCyT[i]) : 0.0;
      tz[e] = inside ? t * static_cast<double>(kCzT[i]) : 0.0;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll 8
      for (int e = 0; e < m; ++e) {
        fo0 += tx[e];
        fo1 += ty[e];
        fo2 += tz[e];
      }
    }
    __syncthreads();
  }
  __shared__ int last;
  if (tid == 0) {
    out[3 * j] = -fo0;
    out[3 * j + 1] = -fo1;
    out[3 * j + 2] = -fo2;
    __threadfence();  // the block force before the count that publishes it
    last = atomicAdd(done, 1u) == static_cast<unsigned int>(nblk - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // every block force is visible to the last workgroup
  SumBlocks3(out, nblk, result, tx);
  if (tid == 0) {
    *done = 0u;
    __threadfence_system();
  }
}


// Cylinder3dCase::total_mass (cylinder3d_case.h:601-622): rho over kBulk
// cells, FP64 partial sums per block.
template <typename T>
__global__ __launch_bounds__(256) void KMass(const T* pops0, const T* pops1, ParArray par,
                                             const BlockTab* bt, const uint8_t* kind, SimConst sc,
                                             double* partial) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const T* base = (par.v[bt[b].level] ? pops1 : pops0) + static_cast<size_t>(b) * kQ * sc.plane;
  double s = 0.0;
  for (int c = threadIdx.x; c < sc.ncore; c += blockDim.x) {
    if (kind[static_cast<size_t>(b) * sc.ncore + c] != kBulk) continue;
    T f[kQ];
    const int sl = CoreSlot(sc, c);
    for (int i = 0; i < kQ; ++i) f[i] = base[static_cast<size_t>(i) * sc.plane + sl];
    T rho, u[3];
    ProxyRhoU(f, rho, u);
    s += RhoDouble(rho);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[b] = red[0];
}

// Sum of the per-block masses: each thread sums a fixed strided subset,
// then a fixed tree -- deterministic, one workgroup.
__global__ __launch_bounds__(256) void KSum1(const double* v, int count, double* out) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int j = tid; j < count; j += 256) s += v[j];
  red[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) *out = red[0];
}

// ---------------------------------------------------------- instantiations
template <typename T>
void LaunchTile(const TileArgs<T>& a, hipStream_t s) {
  if (a.ntiles <= 0) return;
  if (a.sc.padded) {
    if (a.sc.bouzidi) {
      hipLaunchKernelGGL((KPadded<T, true>), dim3(a.ntiles), dim3(512), 0, s, a);
    } else {
      hipLaunchKernelGGL((KPadded<T, false>), dim3(a.ntiles), dim3(512), 0, s, a);
    }
  } else if (a.sc.bouzidi) {
    hipLaunchKernelGGL((KBlock<T, true>), dim3(a.ntiles + a.fd_wgs + a.pc_wgs), dim3(1024), 0, s, a);
  } else {
    hipLaunchKernelGGL((KBlock<T, false>), dim3(a.ntiles + a.fd_wgs + a.pc_wgs), dim3(1024), 0, s, a);
  }
}

static inline dim3 Grid(long long n, int bs) { return dim3(static_cast<unsigned>((n + bs - 1) / bs)); }

template <typename T>
void LaunchStatsFinalize(const double* partial, const int* tile_begin, int members, double* avg, hipStream_t s) {
  hipLaunchKernelGGL(KStatsFinalize, dim3(members), dim3(kStatsThreads), 0, s, partial, tile_begin, avg);
}

void LaunchStatsSequential(const double* val, const uint8_t* use, const int* cell_begin, int members, double* avg,
                           hipStream_t s) {
  hipLaunchKernelGGL(KStatsSequential, Grid(members, 64), dim3(64), 0, s, val, use, cell_begin, members, avg);
}

void LaunchStepInc(const int* step, int* step_next, hipStream_t s) {
  hipLaunchKernelGGL(KStepInc, dim3(1), dim3(1), 0, s, step, step_next);
}

template <typename T>
void LaunchInitEq(T* pops, const uint8_t* kind, int nb, const SimConst& sc, double rho0,
                  const double* u0, uint32_t mask, hipStream_t s) {
  const long long total = static_cast<long long>(nb) * sc.ncore;
  hipLaunchKernelGGL(KInitEq<T>, Grid(total, 256), dim3(256), 0, s, pops, kind, nb, sc, rho0, u0[0],
                     u0[1], u0[2], mask);
}

template <typename T>
void LaunchInitField(T* pops, int nb, const SimConst& sc, const double* rho, const double* u,
                     hipStream_t s) {
  const long long total = static_cast<long long>(nb) * sc.ncore;
  hipLaunchKernelGGL(KInitField<T>, Grid(total, 256), dim3(256), 0, s, pops, nb, sc, rho, u);
}

template <typename T>
void LaunchFdBlocks(const FdArgs<T>& a, const short* fd_slot, const int4* fd_idx, const int ncls[kFdClasses],
                    hipStream_t s) {
  const int4* idx = fd_idx;
  // one launch per FD-cell class (fd_idx lists the blocks of class 0 first):
  // the small classes fit more workgroups per CU
  if (ncls[0] > 0)
    hipLaunchKernelGGL((KFdBlock<T, kFdClassCells[0]>), dim3(ncls[0]), dim3(kTileCells), 0, s, a, fd_slot, idx);
  idx += ncls[0];
  if (ncls[1] > 0)
    hipLaunchKernelGGL((KFdBlock<T, kFdClassCells[1]>), dim3(ncls[1]), dim3(kTileCells), 0, s, a, fd_slot, idx);
  idx += ncls[1];
  if (ncls[2] > 0)
    hipLaunchKernelGGL((KFdBlock<T, kMaxFdCells>), dim3(ncls[2]), dim3(kTileCells), 0, s, a, fd_slot, idx);
}

template <typename T>
void LaunchFd(const FdArgs<T>& a, const int* phase_begin, hipStream_t s) {
  for (int ph = 0; ph < 3; ++ph) {
    const int cnt = phase_begin[ph + 1] - phase_begin[ph];
    if (cnt > 0) hipLaunchKernelGGL(KFd<T>, Grid(cnt, 128), dim3(128), 0, s, a, phase_begin[ph], cnt);
  }
  const int total = phase_begin[3];
  if (total > 0) hipLaunchKernelGGL(KFdCommit<T>, Grid(total, 256), dim3(256), 0, s, a, total);
}

template <typename T>
void LaunchCouplerInit(const CoupleArgs<T>& a, int npts, hipStream_t s) {
  if (npts > 0) hipLaunchKernelGGL(KCouplerInit<T>, Grid(npts, 128), dim3(128), 0, s, a, npts);
}

template <typename T>
void LaunchProlong(const CoupleArgs<T>& a, const int* list, int count, int half, hipStream_t s) {
  if (count > 0) hipLaunchKernelGGL(KProlong<T>, Grid(count, 128), dim3(128), 0, s, a, list, count, half);
}

template <typename T>
void LaunchFdCouple(const FdArgs<T>& fa, const short* fd_slot, const int4* fd_idx, int nfd, const CoupleArgs<T>& ca,
                    const int* list, int count, int mode, hipStream_t s) {
  const int per = mode == 2 ? kMergedGroups : 512;
  const int ncw = (count + per - 1) / per;
  if (nfd + ncw <= 0) return;
  hipLaunchKernelGGL(KFdCouple<T>, dim3(nfd + ncw), dim3(512), 0, s, fa, fd_slot, fd_idx, nfd, ca, list, count, mode);
}

template <typename T>
void LaunchRestrict(const CoupleArgs<T>& a, const int* list, int count, hipStream_t s) {
  if (count > 0)
    hipLaunchKernelGGL(KRestrict<T>, dim3((count + kRestrictGroups - 1) / kRestrictGroups), dim3(32 * kRestrictGroups),
                       0, s, a, list, count);
}

template <typename T>
void LaunchDrag(const DragArgs<T>& a, const int* block_begin, int nblk, double* block_out, unsigned int* done,
                double* result, hipStream_t s) {
  if (nblk > 0)
    hipLaunchKernelGGL(KDragBlock<T>, dim3(nblk), dim3(kDragThreads), 0, s, a, block_begin, block_out, done, nblk,
                       result);
}

template <typename T>
void LaunchMass(const T* p0, const T* p1, const ParArray& par, const BlockTab* bt, const uint8_t* kind,
                const SimConst& sc, int nb, double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL(KMass<T>, dim3(nb), dim3(256), 0, s, p0, p1, par, bt, kind, sc, partial);
  hipLaunchKernelGGL(KSum1, dim3(1), dim3(256), 0, s, partial, nb, out);
}

// Density and velocity per core cell for output (CellProxy::computeRhoU,
// block_lattice.h:113-125), written in VTK cell order: x fastest.
template <typename T>
__global__ __launch_bounds__(256) void KMacro(const T* pops0, const T* pops1, ParArray par, const BlockTab* bt,
                                              SimConst sc, double* rho, double* u) {
  const int b = blockIdx.x;
  const T* base = (par.v[bt[b].level] ? pops1 : pops0) + static_cast<size_t>(b) * kQ * sc.plane;
  const int n = sc.n;
  for (int t = threadIdx.x; t < sc.ncore; t += blockDim.x) {
    const int i = t % n, j = (t / n) % n, k = t / (n * n);  // VTK order
    const int sl = CoreSlot(sc, (i * n + j) * n + k);
    T f[kQ];
    for (int q = 0; q < kQ; ++q) f[q] = base[static_cast<size_t>(q) * sc.plane + sl];
    T r, v[3];
    ProxyRhoU(f, r, v);
    const size_t o = static_cast<size_t>(b) * sc.ncore + t;
    rho[o] = RhoDouble(r);
    for (int d = 0; d < 3; ++d) u[3 * o + d] = static_cast<double>(v[d]);
  }
}

template <typename T>
void LaunchMacro(const T* p0, const T* p1, const ParArray& par, const BlockTab* bt, const SimConst& sc, int nb,
                 double* rho, double* u, hipStream_t s) {
  hipLaunchKernelGGL(KMacro<T>, dim3(nb), dim3(256), 0, s, p0, p1, par, bt, sc, rho, u);
}

#define OCTLB_INSTANTIATE(T)                                                                      \
  template void LaunchTile<T>(const TileArgs<T>&, hipStream_t);                                   \
  template void LaunchStatsFinalize<T>(const double*, const int*, int, double*, hipStream_t);     \
  template void LaunchInitEq<T>(T*, const uint8_t*, int, const SimConst&, double, const double*,  \
                                uint32_t, hipStream_t);                                           \
  template void LaunchInitField<T>(T*, int, const SimConst&, const double*, const double*,        \
                                   hipStream_t);                                                  \
  template void LaunchFd<T>(const FdArgs<T>&, const int*, hipStream_t);                           \
  template void LaunchFdBlocks<T>(const FdArgs<T>&, const short*, const int4*, const int*, hipStream_t);  \
  template void LaunchCouplerInit<T>(const CoupleArgs<T>&, int, hipStream_t);                     \
  template void LaunchProlong<T>(const CoupleArgs<T>&, const int*, int, int, hipStream_t);        \
  template void LaunchRestrict<T>(const CoupleArgs<T>&, const int*, int, hipStream_t);            \
  template void LaunchFdCouple<T>(const FdArgs<T>&, const short*, const int4*, int, const CoupleArgs<T>&, \
                                  const int*, int, int, hipStream_t);                               \
  template void LaunchDrag<T>(const DragArgs<T>&, const int*, int, double*, unsigned int*, double*,  \
                              hipStream_t);                                                       \
  template void LaunchMass<T>(const T*, const T*, const ParArray&, const BlockTab*,               \
                              const uint8_t*, const SimConst&, int, double*, double*, hipStream_t); \
  template void LaunchMacro<T>(const T*, const T*, const ParArray&, const BlockTab*, const SimConst&, \
                               int, double*, double*, hipStream_t);

OCTLB_INSTANTIATE(double)
OCTLB_INSTANTIATE(float)

#include "adjoint_ml.inc"

}  // namespace octlb
