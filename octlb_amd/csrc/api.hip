// C ABI of the device solver (include/octlb_gpu.h): table flattening once per
// adapt, the recursive level subcycle of TimeLoop (time_loop.h:201-246) as a
// launch sequence on one HIP stream, hipGraph capture of a coarse step, and
// FP64 drag / mass reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <unordered_map>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_device.h"
#include "octlb_gpu.h"
#include "octlb_internal.h"
#include "adjoint_topo.h"

namespace octlb {

namespace {
thread_local std::string g_last_error;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      throw HipError(std::string(#expr) + ": " + hipGetErrorString(_e) + " (api.hip:" +         \
                     std::to_string(__LINE__) + ")");                                            \
  } while (0)

constexpr int kOpposite[6] = {1, 0, 3, 2, 5, 4};

// Halo slot order of the blocked shell: padded (10^3) slots that are not core,
// in padded linear order (same enumeration as the device apron table).
std::vector<int> ShellSlots() {
  constexpr ApronTable t = MakeApron();
  return std::vector<int>(t.idx, t.idx + kApron);
}

// Block that supplies halo cell (x,y,z) (core coordinates, one of them
// outside [0,8)) through a face or edge exchange, or -1; *cc = its cell.
int ApronNeighbour(const BlockTab& r, int x, int y, int z, int* cc) {
  constexpr int n = kTile;
  const int ox = (x < 0 || x >= n), oy = (y < 0 || y >= n), oz = (z < 0 || z >= n);
  const int nout = ox + oy + oz;
  int nbk = -1;
  if (nout == 1) {
    const int axis = ox ? 0 : (oy ? 1 : 2);
    const int v = axis == 0 ? x : (axis == 1 ? y : z);
    nbk = r.nbr[axis * 2 + (v >= n ? 1 : 0)];
  } else if (nout == 2) {
    const int a1 = ox ? 0 : 1, a2 = oz ? 2 : 1;
    const int v1 = a1 == 0 ? x : y, v2 = a2 == 2 ? z : y;
    const int pair = (a1 == 0) ? (a2 == 1 ? 0 : 1) : 2;
    nbk = r.nbr[6 + pair * 4 + (v1 >= n ? 1 : 0) * 2 + (v2 >= n ? 1 : 0)];
  }
  const int wx = x < 0 ? n - 1 : (x >= n ? 0 : x);
  const int wy = y < 0 ? n - 1 : (y >= n ? 0 : y);
  const int wz = z < 0 ? n - 1 : (z >= n ? 0 : z);
  *cc = (wx * n + wy) * n + wz;
  return nbk;
}

template <typename P>
P* DevAlloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, count * sizeof(P)));
  return static_cast<P*>(p);
}

template <typename P>
void Upload(P* dst, const std::vector<P>& src, hipStream_t s) {
  if (!src.empty()) HIP_CHECK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(P), hipMemcpyHostToDevice, s));
}

}  // namespace

int SetError(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace octlb

using namespace octlb;

struct octlb_dev {
  // ---- configuration
  int nb = 0, n = 0, h = 0, NP = 0, ncore = 0, NS = 0, plane = 0;
  bool padded = false;
  int prec = 64;
  size_t elem = 8;
  SimConst sc{};
  int max_level = 0;
  bool snapshot = false;
  std::vector<int> level;
  std::vector<double> bounds;
  std::vector<int> same;
  std::vector<int> tb;
  std::vector<octlb_cf_face> cf;
  std::vector<int> tb_blocks;
  std::vector<std::array<int, 6>> fn;
  std::vector<BlockTab> bt;
  std::vector<uint8_t> kinds;
  std::map<int64_t, double> links;  // (block*ncore + cell)*19 + pop -> q
  int nshell = 0;
  // ---- coupling plan (level_coupler.cpp:70-149)
  std::vector<int> plan;  // 12 ints per point
  int npts = 0;
  std::vector<std::vector<int>> prolong_active;   // per coarse level
  std::vector<std::vector<int>> restrict_active;  // per fine level
  // ---- tiles per level
  std::vector<Tile> tiles;
  std::vector<int> tile_begin, tile_count;
  std::vector<int> level_blocks;  // blocks per level (a push tile holds two)
  std::vector<int> block_order;   // visiting order of the blocks of a level
  std::vector<int> block_rank_host;
  // Update scheme of the blocked fused kernel: 0 every block pulls through
  // its LDS box (re-colliding the apron); 1 plain interior blocks push
  // (push tiles), the rest pull.  Chosen by octlb_dev_set_kernel_variant or
  // octlb_dev_autotune.
  int variant = 1;
  int npush_blocks = 0;
  // ---- device
  hipStream_t stream = nullptr;
  void* pops[2] = {nullptr, nullptr};
  void* shells[2] = {nullptr, nullptr};
  uint8_t* d_kind = nullptr;
  uint8_t* d_rule = nullptr;  // per core cell: needs the per-cell stream rules
  int* d_pfull = nullptr;     // per core cell: active prolongation point targeting it, or -1
  int pfull_fuse_level = -1;  // the next update launch of this level applies the full-time prolongation
  bool pfull_fused = false;
  // Coupling epilogues (TileArgs::ep_*): the half-time prolongation after an
  // update of level l and the restriction after the second update of level l
  // in its parent's visit run in the coupled blocks' own workgroups of that
  // update launch.  Per level, CSR over all blocks of the active points.
  bool ep_ok = false;
  int* d_ep_pro_beg[kMaxLevels] = {};
  int* d_ep_pro_pts[kMaxLevels] = {};
  int* d_ep_res_beg[kMaxLevels] = {};
  int* d_ep_res_pts[kMaxLevels] = {};
  bool second_visit[kMaxLevels] = {};   // the next update of this level is its parent's second visit
  bool pro_done = false, res_done = false;  // the coupling step that follows ran in the update launch
  int ep_mode = 3;                      // bit 0: half-time prolongation epilogue, bit 1: restriction epilogue
  double ep_ms[2] = {0.0, 0.0};         // autotuner: fused + coupling kernels per coarse step, off / on
  template <typename T>
  CoupleArgs<T> EpilogueArgs(int l, bool timing_only, int other_level) const;
  template <typename T>
  double TimeLaunches(int l, int reps, int what);
  BlockTab* d_bt = nullptr;
  Tile* d_tiles = nullptr;
  int* d_block_rank = nullptr;  // block -> index among its level's blocks in visiting order
  int* d_level = nullptr;
  uint32_t* d_qmask = nullptr;
  int* d_qoff = nullptr;
  double* d_qval = nullptr;
  uint32_t* d_dmask = nullptr;
  int* d_doff = nullptr;
  void* d_snap = nullptr;
  int4* d_dlinks = nullptr;
  int* d_dblock_begin = nullptr;
  double* d_dterm = nullptr;
  uint8_t* d_duse = nullptr;
  double* d_dblock_out = nullptr;
  unsigned int* d_drag_done = nullptr;  // drag workgroups finished (the last one resets it)
  int ndlinks = 0, ndblocks = 0;
  double* d_stats = nullptr;
  double* d_stat_val = nullptr;
  uint8_t* d_stat_use = nullptr;
  bool fast_stats = false;
  double* d_avg = nullptr;  // ConstRho average per ensemble member
  // ---- ensemble members (one member unless octlb_dev_create_ensemble)
  int nmembers = 1;
  std::vector<int> member_begin;  // [nmembers + 1] block offsets
  std::vector<SimConst> scm;      // constants per member
  SimConst* d_scm = nullptr;
  int* d_stat_tiles = nullptr;    // [nmembers + 1] stats-level tile offsets per member
  int* d_stat_cells = nullptr;    // [nmembers + 1] statistic slots per member
  int* d_step = nullptr;   // [2]: double-buffered coarse-step counter
  int step_slot = 0;       // slot holding the current step's time
  int* inc_next = nullptr; // next-step slot, handed to the step's first kernel
  int* TakeInc(bool launching) {
    if (!launching) return nullptr;
    int* p = inc_next;
    inc_next = nullptr;
    return p;
  }
  double* d_out = nullptr;
  double* h_out = nullptr;  // pinned host copy of d_out: small results come back without a staging copy
  double* d_mass_partial = nullptr;
  int* d_plan = nullptr;
  void* d_prev = nullptr;
  int* d_prolong = nullptr;
  std::vector<int> prolong_off;
  int* d_restrict = nullptr;
  std::vector<int> restrict_off;
  int2* d_fd_cells = nullptr;
  int* d_fd_smap = nullptr;
  int* d_fd_index = nullptr;
  void* d_fd_post = nullptr;
  int fd_phase[4] = {0, 0, 0, 0};
  int* d_fd_blocks = nullptr;  // blocked mode: boundary blocks with FD cells
  short* d_fd_list = nullptr;  // blocked mode: FD cells per FD block, phase order
  int4* d_fd_count = nullptr;  // blocked mode: phase end offsets per FD block
  // host copies of the blocked FD tables (the multi-level adjoint takes them)
  std::vector<int> fd_blocks_host;
  std::vector<short> fd_slot_host, fd_list_host;
  std::vector<int4> fd_count_host;
  short* d_fd_slot = nullptr;  // [fd block][512] FD slot of each cell or -1
  int nfd_blocks = 0;
  std::vector<void*> owned;
  std::vector<void*> prep_owned;  // tables rebuilt by Prepare()
  // ---- state
  int par[kMaxLevels] = {0};
  bool prepared = false;
  bool coupler_ready = false;
  int step = 0;
  // ---- timing of the fused kernel
  bool timing = false;
  bool capturing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  double tile_ms = 0.0;
  long long tile_launches = 0;
  double tile_updates = 0.0;
  // ---- graphs (one per level-0 parity)
  // captured coarse steps: [0..1] one step from level-0 parity p,
  // [2..3] kGraphSteps steps (an even count, so parities come back)
  static constexpr int kGraphSteps = 16;
  hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};

  template <typename P>
  P* Alloc(size_t count) {
    P* p = DevAlloc<P>(count);
    owned.push_back(p);
    return p;
  }
  void* AllocBytes(size_t bytes) {
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes == 0 ? 1 : bytes));
    owned.push_back(p);
    return p;
  }
  template <typename P>
  P* PAlloc(size_t count) {
    P* p = DevAlloc<P>(count);
    prep_owned.push_back(p);
    return p;
  }
  void* PAllocBytes(size_t bytes) {
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes == 0 ? 1 : bytes));
    prep_owned.push_back(p);
    return p;
  }
  // Teardown: failures cannot be reported from here and must not stop the
  // remaining releases, so their status is dropped on purpose.
  void FreePrepared() {
    for (void* p : prep_owned) (void)hipFree(p);
    prep_owned.clear();
  }
  void FreeAll() {
    FreePrepared();
    for (auto& g : gexec)
      if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
    if (h_out) (void)hipHostFree(h_out);
    h_out = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
  ParArray Par() const {
    ParArray p{};
    for (int l = 0; l < kMaxLevels; ++l) p.v[l] = par[l];
    return p;
  }
  char* Core(int buf, int b) const {
    return static_cast<char*>(pops[buf]) + static_cast<size_t>(b) * kQ * plane * elem;
  }
  void InvalidateGraphs() {
    for (auto& g : gexec) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }

  void Build(const octlb_topology* t, const octlb_params* p);
  void SetMembers(int m, const int32_t* begin, const octlb_params* params);
  void Prepare();
  template <typename T>
  TileArgs<T> MakeTileArgs(int l);
  template <typename T>
  double TimeKernel(int l, int reps);
  template <typename T>
  void Level(int l);
  template <typename T>
  FdArgs<T> MakeFdArgs();
  template <typename T>
  void FlushFd();
  template <typename T>
  void Couple(CoupleArgs<T> c, const int* list, int count, int mode, int next_level);
  bool fd_pending = false;   // blocked mode: FD pass of the last streamed level not launched yet
  const int4* fd_pend_idx = nullptr;  // its FD blocks as (FD block, block, level, 0) and count
  int fd_pend_cnt = 0;
  int fd_pend_cls[kFdClasses] = {};  // its blocks per FD-cell class (listed class by class)
  int fd_pend_level = -1;            // level whose own FD blocks are pending, -1 = all levels
  // per level: FD blocks of that level / of the other levels, as index runs in d_fd_idx
  int* d_fd_idx = nullptr;
  int4* d_fd_desc = nullptr;  // d_fd_idx's entries as (FD block, block, level, 0)
  int fd_same_off[kMaxLevels] = {}, fd_same_cnt[kMaxLevels] = {};
  int fd_other_off[kMaxLevels] = {}, fd_other_cnt[kMaxLevels] = {};
  int fd_all_off = 0;
  int fd_same_cls[kMaxLevels][kFdClasses] = {}, fd_all_cls[kFdClasses] = {};
  bool fd_merge_ok = false;  // FD blocks and coupled blocks are disjoint
  template <typename T>
  void CoarseStep();
  void Step();
  void Sync() { HIP_CHECK(hipStreamSynchronize(stream)); }
};

namespace {

DevSpec ResolveSpec(const octlb_params* p, int face) {
  DevSpec d{};
  d.type = OCTLB_BC_NO_SLIP;
  d.rho_target = 1.0;
  d.inlet = OCTLB_INLET_NONE;
  d.inlet_face = face;
  for (int i = 0; i < p->num_specs; ++i) {
    const octlb_bc_spec& s = p->specs[i];
    if (s.face != face) continue;  // FindSpec: first match (interpolated_velocity.h:38-45)
    d.type = s.type;
    d.inlet = s.inlet;
    d.inlet_face = s.inlet_face;
    d.n0 = s.n_inplane0;
    d.n1 = s.n_inplane1;
    for (int k = 0; k < 3; ++k) {
      d.u_wall[k] = s.u_wall[k];
      d.inlet_u[k] = s.inlet_u[k];
    }
    d.rho_target = s.rho_target;
    d.u_peak = s.u_peak;
    d.lo0 = s.lo0;
    d.hi0 = s.hi0;
    d.lo1 = s.lo1;
    d.hi1 = s.hi1;
    d.ramp_end_t = s.ramp_end_t;
    break;
  }
  return d;
}

}  // namespace

void octlb_dev::Build(const octlb_topology* t, const octlb_params* p) {
  nb = t->num_blocks;
  n = t->n;
  h = t->halo;
  if (nb < 1) throw ArgError("topology has no blocks");
  if (n < 4) throw ArgError("block edge n must be >= 4");
  if ((n % 2) != 0 && t->num_coarse_fine > 0)  // LevelCoupler (level_coupler.cpp:72-74)
    throw ArgError("coarse-fine coupling needs an even block edge n");
  if (h < 1) throw ArgError("halo width must be >= 1");
  if (p->precision != 64 && p->precision != 32) throw ArgError("precision must be 64 or 32");
  if (!(p->omega > 0.0)) throw ArgError("omega must be positive");
  NP = n + 2 * h;
  ncore = n * n * n;
  NS = NP * NP * NP;
  padded = !(n == kTile && h == 1);
  if (padded && (t->num_same_level > 0 || t->num_coarse_fine > 0))
    throw ArgError("blocks other than 8^3 with halo 1 are supported only for forests without block faces");
  plane = padded ? NS : ncore;
  prec = p->precision;
  elem = prec == 64 ? sizeof(double) : sizeof(float);
  snapshot = p->snapshot != 0;
  for (int f = 0; f < 6; ++f) sc.spec[f] = ResolveSpec(p, f);
  sc.omega = p->omega;
  sc.n = n;
  sc.h = h;
  sc.NP = NP;
  sc.ncore = ncore;
  sc.NS = NS;
  sc.plane = plane;
  sc.padded = padded ? 1 : 0;
  sc.handler = p->handler;
  sc.const_rho = p->const_rho;
  sc.bouzidi = p->bouzidi;

  level.assign(t->level, t->level + nb);
  bounds.assign(t->bounds, t->bounds + static_cast<size_t>(nb) * 6);
  max_level = 0;
  for (int b = 0; b < nb; ++b) {
    if (level[b] < 0 || level[b] >= kMaxLevels) throw ArgError("block level out of range");
    max_level = std::max(max_level, level[b]);
  }
  same.assign(t->same_level, t->same_level + static_cast<size_t>(t->num_same_level) * 3);
  tb.assign(t->tree_boundary, t->tree_boundary + static_cast<size_t>(t->num_tree_boundary) * 2);
  cf.assign(t->coarse_fine, t->coarse_fine + t->num_coarse_fine);

  // same-level face neighbours (GhostSchedule face_nbr, ghost_schedule.h:61-68)
  fn.assign(nb, {-1, -1, -1, -1, -1, -1});
  for (size_t e = 0; e < same.size() / 3; ++e) {
    const int a = same[3 * e], d = same[3 * e + 1], r = same[3 * e + 2];
    if (a < 0 || a >= nb || r < 0 || r >= nb || d < 0 || d > 5) throw ArgError("same-level face out of range");
    if (level[a] != level[r]) throw ArgError("same-level face joins blocks of different levels");
    fn[a][d] = r;
  }
  for (size_t e = 0; e < tb.size() / 2; ++e) {
    const int b = tb[2 * e];
    if (b < 0 || b >= nb || tb[2 * e + 1] < 0 || tb[2 * e + 1] > 5) throw ArgError("tree-boundary face out of range");
    if (std::find(tb_blocks.begin(), tb_blocks.end(), b) == tb_blocks.end()) tb_blocks.push_back(b);
  }

  // Block table: exchanged face and edge neighbours.  The edge ghost (e1,e2)
  // of block B is filled from fn[fn[B][e2]][e1] or fn[fn[B][e1]][e2]
  // (composition of two same-level faces, ghost_schedule.h:69-91).
  bt.assign(nb, BlockTab{});
  nshell = 0;
  for (int b = 0; b < nb; ++b) {
    BlockTab& r = bt[b];
    r.level = level[b];
    r.flags = 0;
    for (int d = 0; d < 6; ++d) r.nbr[d] = fn[b][d];
    for (int k = 0; k < 18; ++k) r.push_to[k] = -1;
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int pi = 0; pi < 3; ++pi)
      for (int s1 = 0; s1 < 2; ++s1)
        for (int s2 = 0; s2 < 2; ++s2) {
          const int e1 = pairs[pi][0] * 2 + s1, e2 = pairs[pi][1] * 2 + s2;
          int a = -1;
          // the later of the two composition orders wins when both exist
          const int via2 = fn[b][e2];
          if (via2 >= 0 && fn[via2][e1] >= 0) a = fn[via2][e1];
          const int via1 = fn[b][e1];
          if (via1 >= 0 && fn[via1][e2] >= 0) a = fn[via1][e2];
          r.nbr[6 + pi * 4 + s1 * 2 + s2] = a;
        }
    bool needs_shell = false;
    for (int k = 0; k < 18; ++k) needs_shell |= (r.nbr[k] < 0);
    r.shell = (!padded && needs_shell) ? nshell++ : -1;
    const double* bb = &bounds[static_cast<size_t>(b) * 6];
    r.origin[0] = bb[0];
    r.origin[1] = bb[1];
    r.origin[2] = bb[2];
    r.cw = (bb[3] - bb[0]) / static_cast<double>(n);  // cylinder3d_case.h:404-409
  }

  // coupling plan
  plan.clear();
  for (const octlb_cf_face& f : cf) {
    if (f.coarse_id < 0 || f.coarse_id >= nb) throw ArgError("coarse-fine face out of range");
    if (f.normal < 0 || f.normal > 5) throw ArgError("coarse-fine face normal out of range");
    if (f.coarse_level < 0 || f.coarse_level + 1 >= kMaxLevels || f.coarse_level != level[f.coarse_id] ||
        f.fine_level != f.coarse_level + 1)
      throw ArgError("coarse-fine face levels disagree with the block levels");
    const double* cb = f.coarse_bounds;
    const double cc[3] = {0.5 * (cb[0] + cb[3]), 0.5 * (cb[1] + cb[4]), 0.5 * (cb[2] + cb[5])};
    const int axis = f.normal / 2, is_max = f.normal & 1;
    const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;
    const int cni = is_max ? n - 1 : 0, fni = is_max ? 0 : n - 1;
    const int half = n / 2;
    for (int sl = 0; sl < 4; ++sl) {
      if (f.fine_ids[sl] < 0 || f.fine_ids[sl] >= nb) throw ArgError("coarse-fine face out of range");
      if (level[f.fine_ids[sl]] != f.fine_level) throw ArgError("coarse-fine face levels disagree with the block levels");
      const double* fb = f.fine_bounds[sl];
      const double fc[3] = {0.5 * (fb[0] + fb[3]), 0.5 * (fb[1] + fb[4]), 0.5 * (fb[2] + fb[5])};
      const int b0 = (fc[t0] < cc[t0] ? 0 : 1) * half, b1 = (fc[t1] < cc[t1] ? 0 : 1) * half;
      for (int a0 = b0; a0 < b0 + half; ++a0)
        for (int a1 = b1; a1 < b1 + half; ++a1) {
          int ci[3], fi[3];
          ci[axis] = cni;
          fi[axis] = fni;
          ci[t0] = a0;
          ci[t1] = a1;
          fi[t0] = a0 - b0;
          fi[t1] = a1 - b1;
          const int row[12] = {f.coarse_id, ci[0], ci[1], ci[2], f.fine_ids[sl], fi[0], fi[1],
                               fi[2],       f.normal, f.coarse_level, f.fine_level, sl};
          plan.insert(plan.end(), row, row + 12);
        }
    }
  }
  npts = static_cast<int>(plan.size() / 12);
  // LevelRangeBegin/End (level_coupler.cpp:309-323): one [min,max+1) range
  // per coarse level; points inside a range are prolonged in plan order, so
  // the last point writing a fine cell wins.
  prolong_active.assign(kMaxLevels, {});
  restrict_active.assign(kMaxLevels, {});
  {
    int maxc = -1;
    for (int i = 0; i < npts; ++i) maxc = std::max(maxc, plan[12 * i + 9]);
    for (int l = 0; l <= maxc; ++l) {
      int beg = npts, end = 0;
      for (int i = 0; i < npts; ++i)
        if (plan[12 * i + 9] == l) {
          beg = std::min(beg, i);
          end = std::max(end, i + 1);
        }
      std::map<int64_t, int> last;
      for (int i = beg; i < end; ++i) {
        const int* r = &plan[12 * i];
        last[static_cast<int64_t>(r[4]) * ncore + (r[5] * n + r[6]) * n + r[7]] = i;
      }
      for (int i = beg; i < end; ++i) {
        const int* r = &plan[12 * i];
        if (last[static_cast<int64_t>(r[4]) * ncore + (r[5] * n + r[6]) * n + r[7]] == i)
          prolong_active[l].push_back(i);
      }
    }
    for (int fl = 1; fl <= max_level; ++fl) {
      std::map<int64_t, int> last;
      for (int i = 0; i < npts; ++i) {
        const int* r = &plan[12 * i];
        if (r[10] != fl) continue;
        last[static_cast<int64_t>(r[0]) * ncore + (r[1] * n + r[2]) * n + r[3]] = i;
      }
      for (int i = 0; i < npts; ++i) {
        const int* r = &plan[12 * i];
        if (r[10] != fl) continue;
        if (last[static_cast<int64_t>(r[0]) * ncore + (r[1] * n + r[2]) * n + r[3]] == i)
          restrict_active[fl].push_back(i);
      }
    }
  }

  // tiles per level (octants_by_level_, time_loop.h:155-160)
  tile_begin.assign(kMaxLevels, 0);
  tile_count.assign(kMaxLevels, 0);
  level_blocks.assign(kMaxLevels, 0);
  tiles.clear();
  const int tpa = padded ? (NP + kTile - 1) / kTile : 1;
  // Blocks of a level are visited in lexicographic order of their origin, z
  // fastest: z-neighbours (whose shared face is the strided one in the
  // [19][x][y][z] layout) are then consecutive tiles on the same XCD, so the
  // apron re-read hits L2 (4 % faster than octant order on the cylinder).
  // With ConstRho the octant order is kept: the average is summed in tile
  // order, which must be the reference's.
  std::vector<int>& order = block_order;
  order.resize(nb);
  for (int b = 0; b < nb; ++b) order[b] = b;
  if (!padded && !sc.const_rho) {
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
      const double* pa = &bounds[6 * static_cast<size_t>(a)];
      const double* pc = &bounds[6 * static_cast<size_t>(c)];
      if (pa[0] != pc[0]) return pa[0] < pc[0];
      if (pa[1] != pc[1]) return pa[1] < pc[1];
      return pa[2] < pc[2];
    });
  }
  for (int l = 0; l <= max_level; ++l) {
    tile_begin[l] = static_cast<int>(tiles.size());
    for (int bi = 0; bi < nb; ++bi) {
      const int b = order[bi];
      if (level[b] != l) continue;
      for (int x = 0; x < tpa; ++x)
        for (int y = 0; y < tpa; ++y)
          for (int z = 0; z < tpa; ++z) tiles.push_back(Tile{b, x * kTile, y * kTile, z * kTile});
    }
    tile_count[l] = static_cast<int>(tiles.size()) - tile_begin[l];
    level_blocks[l] = tile_count[l] / (tpa * tpa * tpa);
  }

  kinds.assign(static_cast<size_t>(nb) * ncore, 0);

  // ---- device allocation
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const size_t pop_bytes = static_cast<size_t>(nb) * kQ * plane * elem;
  for (int k = 0; k < 2; ++k) {
    pops[k] = AllocBytes(pop_bytes);
    HIP_CHECK(hipMemsetAsync(pops[k], 0, pop_bytes, stream));
    const size_t sb = static_cast<size_t>(nshell) * kQ * kApron * elem;
    shells[k] = AllocBytes(sb);
    if (sb) HIP_CHECK(hipMemsetAsync(shells[k], 0, sb, stream));
  }
  d_kind = Alloc<uint8_t>(kinds.size());
  HIP_CHECK(hipMemsetAsync(d_kind, 0, kinds.size(), stream));
  d_bt = Alloc<BlockTab>(nb);
  Upload(d_bt, bt, stream);
  d_tiles = Alloc<Tile>(tiles.size());
  Upload(d_tiles, tiles, stream);
  {
    std::vector<int> rank(nb, 0), next(kMaxLevels, 0);
    for (int bi = 0; bi < nb; ++bi) rank[block_order[bi]] = next[level[block_order[bi]]]++;
    d_block_rank = Alloc<int>(nb);
    Upload(d_block_rank, rank, stream);
    block_rank_host = rank;
  }
  d_level = Alloc<int>(nb);
  Upload(d_level, level, stream);
  d_stats = Alloc<double>(2 * tiles.size());
  fast_stats = p->fast_stats != 0;
  if (sc.const_rho && !fast_stats) {
    d_stat_val = Alloc<double>(static_cast<size_t>(nb) * ncore);
    d_stat_use = Alloc<uint8_t>(static_cast<size_t>(nb) * ncore);
  }
  d_avg = Alloc<double>(1);
  const double one = 1.0;
  HIP_CHECK(hipMemcpyAsync(d_avg, &one, sizeof(double), hipMemcpyHostToDevice, stream));
  nmembers = 1;
  member_begin = {0, nb};
  scm.assign(1, sc);
  d_scm = Alloc<SimConst>(1);
  Upload(d_scm, scm, stream);
  d_step = Alloc<int>(2);
  HIP_CHECK(hipMemsetAsync(d_step, 0, 2 * sizeof(int), stream));
  d_out = Alloc<double>(8);
  HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_out), 8 * sizeof(double), hipHostMallocDefault));
  d_mass_partial = Alloc<double>(nb);
  d_plan = Alloc<int>(plan.size());
  Upload(d_plan, plan, stream);
  d_prev = AllocBytes(static_cast<size_t>(std::max(npts, 1)) * kMacro * elem);
  std::vector<int> cat;
  prolong_off.assign(kMaxLevels + 1, 0);
  for (int l = 0; l < kMaxLevels; ++l) {
    prolong_off[l] = static_cast<int>(cat.size());
    cat.insert(cat.end(), prolong_active[l].begin(), prolong_active[l].end());
  }
  prolong_off[kMaxLevels] = static_cast<int>(cat.size());
  d_prolong = Alloc<int>(cat.size());
  Upload(d_prolong, cat, stream);
  cat.clear();
  restrict_off.assign(kMaxLevels + 1, 0);
  for (int l = 0; l < kMaxLevels; ++l) {
    restrict_off[l] = static_cast<int>(cat.size());
    cat.insert(cat.end(), restrict_active[l].begin(), restrict_active[l].end());
  }
  restrict_off[kMaxLevels] = static_cast<int>(cat.size());
  d_restrict = Alloc<int>(cat.size());
  Upload(d_restrict, cat, stream);
  Sync();
}

// Tables that depend on the BcKind bytes and Bouzidi links.
void octlb_dev::Prepare() {
  Sync();
  FreePrepared();
  nfd_blocks = 0;
  fd_blocks_host.clear();
  fd_slot_host.clear();
  fd_list_host.clear();
  fd_count_host.clear();
  fd_merge_ok = false;
  fd_pending = false;
  fd_pend_cnt = 0;
  d_fd_idx = nullptr;
  d_fd_desc = nullptr;
  for (int lv = 0; lv < kMaxLevels; ++lv) fd_same_cnt[lv] = fd_other_cnt[lv] = 0;
  d_fd_blocks = nullptr;
  d_fd_slot = nullptr;
  d_fd_list = nullptr;
  d_fd_count = nullptr;
  if (!links.empty() && !sc.bouzidi) throw StateError("Bouzidi links set on a solver built without them");
  if (!links.empty() && padded) throw StateError("Bouzidi links need 8^3 blocks with halo 1");
  for (BlockTab& r : bt) r.flags = 0;
  // ---- Bouzidi links -> per-cell mask + compact q list (block, cell, pop order)
  std::vector<uint32_t> qmask(static_cast<size_t>(nb) * ncore, 0);
  std::vector<int> qoff(static_cast<size_t>(nb) * ncore, 0);
  std::vector<double> qval;
  {
    auto it = links.begin();
    for (size_t cellg = 0; cellg < qmask.size(); ++cellg) {
      qoff[cellg] = static_cast<int>(qval.size());
      while (it != links.end() && it->first / kQ == static_cast<int64_t>(cellg)) {
        const int pop = static_cast<int>(it->first % kQ);
        qmask[cellg] |= 1u << pop;
        qval.push_back(it->second);
        bt[cellg / ncore].flags |= kFlagLinks;
        ++it;
      }
    }
  }
  // ---- per-cell stream rules (block_lattice.cpp:579-674): a cell needs more
  // than the plain pull when it is kSolid (not streamed), reflects on pull
  // (velocity, pressure, Bouzidi kinds), carries Bouzidi links, or pulls from
  // an in-block kSolid / kBouzidi cell (halo slots count as kBulk).  Wall
  // kinds that act in the collide (bounce-back, moving wall, outflow) stream
  // like kBulk.
  std::vector<uint8_t> rule(static_cast<size_t>(nb) * ncore, 0);
  if (!padded) {
    for (int b = 0; b < nb; ++b) {
      const uint8_t* kb = &kinds[static_cast<size_t>(b) * ncore];
      bool plain = true;
      for (int c = 0; c < ncore; ++c) {
        const size_t cg = static_cast<size_t>(b) * ncore + c;
        bool need = kb[c] == kSolid || ReflectsOnPull(kb[c]) || qmask[cg] != 0;
        const int ix = c / (n * n), iy = (c / n) % n, iz = c % n;
        for (int i = 1; i < kQ && !need; ++i) {
          const int x = ix - Cx(i), y = iy - Cy(i), z = iz - Cz(i);
          if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) continue;
          const int sk = kb[(x * n + y) * n + z];
          need = sk == kSolid || sk == kBouzidi;
        }
        rule[cg] = need ? 1 : 0;
        plain = plain && !need;
      }
      if (plain) bt[b].flags |= kFlagPlain;
    }
  }
  d_rule = PAlloc<uint8_t>(rule.size());
  Upload(d_rule, rule, stream);
  // fine-side targets of the (deduplicated) prolongation points: the update
  // launch just before a full-time prolongation applies it to its own cells
  {
    std::vector<int> pfull(static_cast<size_t>(nb) * ncore, -1);
    if (!padded) {
      for (int lv = 0; lv < max_level; ++lv)
        for (int i : prolong_active[lv]) {
          const int* r = &plan[12 * i];
          pfull[static_cast<size_t>(r[4]) * ncore + (r[5] * n + r[6]) * n + r[7]] = i;
          bt[r[4]].flags |= kFlagProlong;
        }
    }
    d_pfull = PAlloc<int>(pfull.size());
    Upload(d_pfull, pfull, stream);
  }
  // ---- drag links: kBulk cells with an in-block kSolid/kBouzidi neighbour at
  // x + c_i, in the order of drag_postprocessor.h:77-110 (iz, iy, ix, iPop)
  std::vector<uint32_t> dmask(static_cast<size_t>(nb) * ncore, 0);
  std::vector<int> doff(static_cast<size_t>(nb) * ncore, 0);
  std::vector<int4> dlinks;
  std::vector<int> dblock_begin;
  {
    std::vector<std::vector<std::pair<int, int>>> per_cell(static_cast<size_t>(nb) * ncore);
    int nsnap = 0;
    for (int b = 0; b < nb; ++b) {
      const uint8_t* kb = &kinds[static_cast<size_t>(b) * ncore];
      bool any = false;
      const size_t first = dlinks.size();
      for (int iz = 0; iz < n; ++iz)
        for (int iy = 0; iy < n; ++iy)
          for (int ix = 0; ix < n; ++ix) {
            const int c = (ix * n + iy) * n + iz;
            if (kb[c] != kBulk) continue;
            for (int i = 1; i < kQ; ++i) {
              const int x = ix + Cx(i), y = iy + Cy(i), z = iz + Cz(i);
              if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) continue;
              const int nk = kb[(x * n + y) * n + z];
              if (nk != kSolid && nk != kBouzidi) continue;
              dmask[static_cast<size_t>(b) * ncore + c] |= 1u << i;
              dlinks.push_back(make_int4(b, c, i, -1));
              any = true;
            }
          }
      if (any) {
        dblock_begin.push_back(static_cast<int>(first));
        // snapshot slots: per cell, pops ascending (matches the kernel's packing)
        for (int c = 0; c < ncore; ++c) {
          const uint32_t m = dmask[static_cast<size_t>(b) * ncore + c];
          doff[static_cast<size_t>(b) * ncore + c] = nsnap;
          nsnap += __builtin_popcount(m);
        }
        for (size_t j = first; j < dlinks.size(); ++j) {
          const int c = dlinks[j].y, i = dlinks[j].z;
          const uint32_t m = dmask[static_cast<size_t>(b) * ncore + c];
          dlinks[j].w = doff[static_cast<size_t>(b) * ncore + c] + __builtin_popcount(m & ((1u << i) - 1u));
        }
        if (snapshot) bt[b].flags |= kFlagDrag;
      }
    }
    dblock_begin.push_back(static_cast<int>(dlinks.size()));
    ndlinks = static_cast<int>(dlinks.size());
    ndblocks = static_cast<int>(dblock_begin.size()) - 1;
    d_snap = PAllocBytes(static_cast<size_t>(std::max(nsnap, 1)) * elem);
    HIP_CHECK(hipMemsetAsync(d_snap, 0, static_cast<size_t>(std::max(nsnap, 1)) * elem, stream));
  }
  // ---- shell liveness: a non-exchanged halo slot is read by some core cell
  if (!padded) {
    for (int b = 0; b < nb; ++b) {
      BlockTab& r = bt[b];
      if (r.shell < 0) continue;
      const uint8_t* kb = &kinds[static_cast<size_t>(b) * ncore];
      bool live = false;
      for (int c = 0; c < ncore && !live; ++c) {
        const int ix = c / (n * n), iy = (c / n) % n, iz = c % n;
        if (ix > 0 && ix < n - 1 && iy > 0 && iy < n - 1 && iz > 0 && iz < n - 1) continue;
        const int k = kb[c];
        if (sc.bouzidi && (k == kSolid || ReflectsOnPull(k))) continue;
        for (int i = 1; i < kQ && !live; ++i) {
          const int x = ix - Cx(i), y = iy - Cy(i), z = iz - Cz(i);
          const int ox = (x < 0 || x >= n), oy = (y < 0 || y >= n), oz = (z < 0 || z >= n);
          const int nout = ox + oy + oz;
          if (nout == 0) continue;
          int nbk;
          if (nout == 1) {
            const int axis = ox ? 0 : (oy ? 1 : 2);
            const int v = axis == 0 ? x : (axis == 1 ? y : z);
            nbk = r.nbr[axis * 2 + (v >= n ? 1 : 0)];
          } else {
            const int a1 = ox ? 0 : 1, a2 = oz ? 2 : 1;
            const int v1 = a1 == 0 ? x : y, v2 = a2 == 2 ? z : y;
            const int pair = (a1 == 0) ? (a2 == 1 ? 0 : 1) : 2;
            nbk = r.nbr[6 + pair * 4 + (v1 >= n ? 1 : 0) * 2 + (v2 >= n ? 1 : 0)];
          }
          if (nbk < 0) live = true;
        }
      }
      if (live) r.flags |= kFlagShellLive;
    }
  }
  // ---- push-class blocks and the tile list.  A block is push-class when
  // every cell streams by the plain pull, its 18 halo parts are all exchanged
  // at its level (no shell, no coupling, no FD face), nothing reads its
  // post-collide values between collide and stream (no drag snapshot) and no
  // coupling point writes it.  Reference-order ConstRho statistics keep one
  // block per tile, so they keep the pull scheme.
  npush_blocks = 0;
  if (!padded) {
    const bool push_ok = variant == 1 && !(sc.const_rho && !fast_stats);
    std::vector<char> is_push(nb, 0);
    for (int b = 0; b < nb; ++b) {
      BlockTab& r = bt[b];
      for (int k = 0; k < 18; ++k) r.push_to[k] = -1;
      if (!push_ok) continue;
      bool ok = (r.flags & kFlagPlain) && !(r.flags & (kFlagShellLive | kFlagDrag | kFlagProlong | kFlagLinks));
      for (int k = 0; k < 18 && ok; ++k) ok = r.nbr[k] >= 0;
      is_push[b] = ok ? 1 : 0;
    }
    auto opposite = [](int part) {
      if (part < 6) return part ^ 1;
      const int e = part - 6;
      return 6 + (e & ~3) + (3 - (e & 3));  // both sides flipped
    };
    for (int b = 0; b < nb; ++b) {
      if (!is_push[b]) continue;
      bt[b].flags |= kFlagPush;
      ++npush_blocks;
      for (int k = 0; k < 18; ++k) {
        const int src = bt[b].nbr[k];
        const int back = opposite(k);
        if (bt[src].push_to[back] >= 0 && bt[src].push_to[back] != b)
          throw StateError("push tables: two blocks pull the same halo part");
        bt[src].push_to[back] = b;
      }
    }
    for (int b = 0; b < nb; ++b) {
      if (is_push[b]) continue;
      for (int k = 0; k < 18; ++k)
        if (bt[b].push_to[k] >= 0) bt[b].flags |= kFlagPushOut;
    }
    // tiles: per level in visiting order; consecutive push-class blocks of
    // one ensemble member pair up in a push tile
    tiles.clear();
    for (int l = 0; l <= max_level; ++l) {
      tile_begin[l] = static_cast<int>(tiles.size());
      int pending = -1;
      for (int bi = 0; bi < nb; ++bi) {
        const int b = block_order[bi];
        if (level[b] != l) continue;
        if (!is_push[b]) {
          tiles.push_back(Tile{b, 0, 0, 0});
          continue;
        }
        if (pending >= 0 && bt[pending].member == bt[b].member) {
          tiles.push_back(Tile{pending, b, 1, 0});
          pending = -1;
          continue;
        }
        if (pending >= 0) tiles.push_back(Tile{pending, -1, 1, 0});
        pending = b;
      }
      if (pending >= 0) tiles.push_back(Tile{pending, -1, 1, 0});
      tile_count[l] = static_cast<int>(tiles.size()) - tile_begin[l];
    }
    Upload(d_tiles, tiles, stream);
  }
  // ---- post-stream FD cell lists (concrete handler, tree-boundary blocks)
  std::vector<int2> fd_cells[3];
  std::vector<int> fd_index(nb, -1);
  std::vector<int> smap;
  if (sc.handler == OCTLB_HANDLER_CONCRETE) {
    int nfd = 0;
    for (int b : tb_blocks) fd_index[b] = nfd++;
    smap.assign(static_cast<size_t>(std::max(nfd, 1)) * ncore, -1);
    for (int b : tb_blocks) {
      const uint8_t* kb = &kinds[static_cast<size_t>(b) * ncore];
      for (int c = 0; c < ncore; ++c) {
        if (!IsFd(kb[c])) continue;
        const int ix = c / (n * n), iy = (c / n) % n, iz = c % n;
        const int fc = (ix == 0) + (ix == n - 1) + (iy == 0) + (iy == n - 1) + (iz == 0) + (iz == n - 1);
        const int ph = fc == 3 ? 2 : (fc == 2 ? 1 : 0);
        fd_cells[ph].push_back(make_int2(b, c));
      }
    }
    fd_phase[0] = 0;
    fd_phase[1] = static_cast<int>(fd_cells[0].size());
    fd_phase[2] = fd_phase[1] + static_cast<int>(fd_cells[1].size());
    fd_phase[3] = fd_phase[2] + static_cast<int>(fd_cells[2].size());
    std::vector<int2> all;
    for (int ph = 0; ph < 3; ++ph) all.insert(all.end(), fd_cells[ph].begin(), fd_cells[ph].end());
    for (size_t e = 0; e < all.size(); ++e)
      smap[static_cast<size_t>(fd_index[all[e].x]) * ncore + all[e].y] = static_cast<int>(e);
    d_fd_cells = PAlloc<int2>(all.size());
    Upload(d_fd_cells, all, stream);
    d_fd_smap = PAlloc<int>(smap.size());
    Upload(d_fd_smap, smap, stream);
    d_fd_index = PAlloc<int>(nb);
    Upload(d_fd_index, fd_index, stream);
    d_fd_post = PAllocBytes(std::max<size_t>(all.size(), 1) * kQ * elem);
    if (!padded) {
      std::vector<int> fdb;
      std::vector<short> slots, lists;
      std::vector<int4> counts;
      for (int b : tb_blocks) {
        const uint8_t* kb = &kinds[static_cast<size_t>(b) * ncore];
        std::vector<short> sl(ncore, -1);
        std::vector<short> byph[3];
        short k = 0;
        for (int c = 0; c < ncore; ++c) {
          if (!IsFd(kb[c])) continue;
          sl[c] = k++;
          const int ix = c / (n * n), iy = (c / n) % n, iz = c % n;
          const int fc = (ix == 0) + (ix == n - 1) + (iy == 0) + (iy == n - 1) + (iz == 0) + (iz == n - 1);
          byph[fc == 3 ? 2 : (fc == 2 ? 1 : 0)].push_back(static_cast<short>(c));
        }
        if (k == 0) continue;
        fdb.push_back(b);
        slots.insert(slots.end(), sl.begin(), sl.end());
        std::vector<short> ls;
        for (int ph = 0; ph < 3; ++ph) ls.insert(ls.end(), byph[ph].begin(), byph[ph].end());
        if (ls.size() > static_cast<size_t>(6 * 36 + 12 * 6 + 8))
          throw ArgError("FD cells must lie on block faces");
        const int e0 = static_cast<int>(byph[0].size()), e1 = e0 + static_cast<int>(byph[1].size());
        counts.push_back(make_int4(e0, e1, static_cast<int>(ls.size()), 0));
        ls.resize(ncore, -1);
        lists.insert(lists.end(), ls.begin(), ls.end());
      }
      nfd_blocks = static_cast<int>(fdb.size());
      fd_blocks_host = fdb;
      fd_slot_host = slots;
      fd_list_host = lists;
      fd_count_host = counts;
      d_fd_blocks = PAlloc<int>(fdb.size());
      Upload(d_fd_blocks, fdb, stream);
      d_fd_slot = PAlloc<short>(slots.size());
      Upload(d_fd_slot, slots, stream);
      d_fd_list = PAlloc<short>(lists.size());
      Upload(d_fd_list, lists, stream);
      d_fd_count = PAlloc<int4>(counts.size());
      Upload(d_fd_count, counts, stream);
      {
        // FD-cell class of each FD block: the pass's LDS is sized by it
        auto cls = [&](int k) {
          const int cells = counts[static_cast<size_t>(k)].z;
          return cells <= kFdClassCells[0] ? 0 : (cells <= kFdClassCells[1] ? 1 : 2);
        };
        std::vector<int> idx;
        for (int lv = 0; lv <= max_level; ++lv) {
          fd_same_off[lv] = static_cast<int>(idx.size());
          for (int c = 0; c < kFdClasses; ++c) {
            const size_t before = idx.size();
            for (int k = 0; k < static_cast<int>(fdb.size()); ++k)
              if (level[fdb[k]] == lv && cls(k) == c) idx.push_back(k);
            fd_same_cls[lv][c] = static_cast<int>(idx.size() - before);
          }
          fd_same_cnt[lv] = static_cast<int>(idx.size()) - fd_same_off[lv];
          fd_other_off[lv] = static_cast<int>(idx.size());
          for (int k = 0; k < static_cast<int>(fdb.size()); ++k)
            if (level[fdb[k]] != lv) idx.push_back(k);
          fd_other_cnt[lv] = static_cast<int>(idx.size()) - fd_other_off[lv];
        }
        fd_all_off = static_cast<int>(idx.size());
        for (int c = 0; c < kFdClasses; ++c) {
          const size_t before = idx.size();
          for (int k = 0; k < static_cast<int>(fdb.size()); ++k)
            if (cls(k) == c) idx.push_back(k);
          fd_all_cls[c] = static_cast<int>(idx.size() - before);
        }
        d_fd_idx = PAlloc<int>(idx.size());
        Upload(d_fd_idx, idx, stream);
        std::vector<int4> desc(idx.size());
        for (size_t j = 0; j < idx.size(); ++j) desc[j] = make_int4(idx[j], fdb[idx[j]], level[fdb[idx[j]]], 0);
        d_fd_desc = PAlloc<int4>(desc.size());
        Upload(d_fd_desc, desc, stream);
      }
      // the FD pass may share a launch with a coupling step only if no FD
      // block is read or written by a coupling point
      std::vector<char> is_fd(nb, 0);
      for (int b : fdb) is_fd[b] = 1;
      fd_merge_ok = true;
      for (int i = 0; i < npts && fd_merge_ok; ++i)
        fd_merge_ok = !is_fd[plan[12 * i]] && !is_fd[plan[12 * i + 4]];
    }
  } else {
    fd_phase[0] = fd_phase[1] = fd_phase[2] = fd_phase[3] = 0;
  }

  // ---- coupling epilogues: the coupled blocks must be pull tiles of their
  // level's launch, and no FD block may be coupled (the FD pass of a level
  // then commutes with its coupling steps)
  ep_ok = !padded && npts > 0 && (nfd_blocks == 0 || fd_merge_ok);
  for (int l = 0; l < kMaxLevels && ep_ok; ++l) {
    for (int i : prolong_active[l]) ep_ok = ep_ok && !(bt[plan[12 * i]].flags & kFlagPush);
    for (int i : restrict_active[l]) ep_ok = ep_ok && !(bt[plan[12 * i + 4]].flags & kFlagPush);
  }
  for (int l = 0; l < kMaxLevels; ++l) {
    d_ep_pro_beg[l] = d_ep_pro_pts[l] = d_ep_res_beg[l] = d_ep_res_pts[l] = nullptr;
    if (!ep_ok) continue;
    auto csr = [&](const std::vector<int>& pts, int field, int** d_beg, int** d_pts) {
      if (pts.empty()) return;
      std::vector<int> beg(static_cast<size_t>(nb) + 1, 0), out(pts.size());
      for (int i : pts) ++beg[plan[12 * i + field] + 1];
      for (int b = 0; b < nb; ++b) beg[b + 1] += beg[b];
      std::vector<int> at(beg.begin(), beg.end() - 1);
      for (int i : pts) out[at[plan[12 * i + field]]++] = i;
      *d_beg = PAlloc<int>(beg.size());
      Upload(*d_beg, beg, stream);
      *d_pts = PAlloc<int>(out.size());
      Upload(*d_pts, out, stream);
    };
    csr(prolong_active[l], 0, &d_ep_pro_beg[l], &d_ep_pro_pts[l]);   // by coarse block
    csr(restrict_active[l], 4, &d_ep_res_beg[l], &d_ep_res_pts[l]);  // by fine block
  }

  d_qmask = PAlloc<uint32_t>(qmask.size());
  Upload(d_qmask, qmask, stream);
  d_qoff = PAlloc<int>(qoff.size());
  Upload(d_qoff, qoff, stream);
  d_qval = PAlloc<double>(qval.size());
  Upload(d_qval, qval, stream);
  d_dmask = PAlloc<uint32_t>(dmask.size());
  Upload(d_dmask, dmask, stream);
  d_doff = PAlloc<int>(doff.size());
  Upload(d_doff, doff, stream);
  d_dlinks = PAlloc<int4>(dlinks.size());
  Upload(d_dlinks, dlinks, stream);
  d_dblock_begin = PAlloc<int>(dblock_begin.size());
  Upload(d_dblock_begin, dblock_begin, stream);
  d_dterm = PAlloc<double>(dlinks.size());
  d_duse = PAlloc<uint8_t>(dlinks.size());
  d_dblock_out = PAlloc<double>(3 * static_cast<size_t>(std::max(ndblocks, 1)));
  d_drag_done = PAlloc<unsigned int>(1);
  HIP_CHECK(hipMemsetAsync(d_drag_done, 0, sizeof(unsigned int), stream));
  // ConstRho statistics of the finest level, per member: the member's
  // tiles are contiguous in tile order (octant order under ConstRho)
  {
    const int l = max_level;
    const int tpb = padded ? ((NP + kTile - 1) / kTile) * ((NP + kTile - 1) / kTile) * ((NP + kTile - 1) / kTile) : 1;
    // partial-sum slots of the member's blocks (blocked mode: one per block,
    // in visiting order; padded mode: one per tile)
    std::vector<int> tiles_of(nmembers + 1, 0);
    int prev = 0;
    if (padded) {
      for (int i = 0; i < tile_count[l]; ++i) {
        const int m = bt[tiles[tile_begin[l] + i].block].member;
        if (m < prev) throw ArgError("ensemble members must be contiguous in block order");
        prev = m;
        tiles_of[m + 1] += 1;
      }
    } else {
      for (int bi = 0; bi < nb; ++bi) {
        const int b = block_order[bi];
        if (level[b] != l) continue;
        const int m = bt[b].member;
        if (m < prev) throw ArgError("ensemble members must be contiguous in block order");
        prev = m;
        tiles_of[m + 1] += 1;
      }
    }
    std::vector<int> cells(nmembers + 1, 0);
    for (int m = 0; m < nmembers; ++m) {
      tiles_of[m + 1] += tiles_of[m];
      cells[m + 1] = tiles_of[m + 1] / tpb * ncore;
    }
    d_stat_tiles = PAlloc<int>(nmembers + 1);
    Upload(d_stat_tiles, tiles_of, stream);
    d_stat_cells = PAlloc<int>(nmembers + 1);
    Upload(d_stat_cells, cells, stream);
  }
  Upload(d_bt, bt, stream);
  Sync();
  prepared = true;
  InvalidateGraphs();
}

// Independent cases in one device state: blocks [begin[m], begin[m+1]) are
// member m, whose relaxation rate and boundary specs come from params[m].
// Members share the solver's structure (precision, handler, ConstRho,
// Bouzidi, snapshot, statistics mode) and are single-level, with no block
// face between members.
void octlb_dev::SetMembers(int m, const int32_t* begin, const octlb_params* params) {
  if (m < 1) throw ArgError("an ensemble needs at least one member");
  if (padded) throw ArgError("ensembles need 8^3 blocks with halo 1");
  if (begin[0] != 0 || begin[m] != nb) throw ArgError("member block ranges must cover the blocks");
  for (int i = 0; i < m; ++i)
    if (begin[i + 1] <= begin[i]) throw ArgError("member block ranges must be increasing and non-empty");
  if (m > 1 && (max_level != 0 || !cf.empty())) throw ArgError("ensemble members must be single-level forests");
  const octlb_params& p0 = params[0];
  std::vector<SimConst> v(m, sc);
  for (int i = 0; i < m; ++i) {
    const octlb_params& p = params[i];
    if (p.precision != p0.precision || p.handler != p0.handler || p.const_rho != p0.const_rho ||
        p.stats_scope != p0.stats_scope || p.bouzidi != p0.bouzidi || p.snapshot != p0.snapshot ||
        p.fast_stats != p0.fast_stats)
      throw ArgError("ensemble members must share precision, handler, ConstRho, Bouzidi, snapshot and statistics mode");
    if (!(p.omega > 0.0)) throw ArgError("omega must be positive");
    v[i].omega = p.omega;
    for (int f = 0; f < 6; ++f) v[i].spec[f] = ResolveSpec(&p, f);
  }
  std::vector<int> owner(nb);
  for (int i = 0; i < m; ++i)
    for (int b = begin[i]; b < begin[i + 1]; ++b) owner[b] = i;
  for (size_t e = 0; e < same.size() / 3; ++e)
    if (owner[same[3 * e]] != owner[same[3 * e + 2]]) throw ArgError("a block face joins two ensemble members");
  nmembers = m;
  member_begin.assign(begin, begin + m + 1);
  scm = v;
  sc = v[0];
  for (int b = 0; b < nb; ++b) bt[b].member = owner[b];
  d_scm = Alloc<SimConst>(m);
  Upload(d_scm, scm, stream);
  d_avg = Alloc<double>(m);
  const std::vector<double> ones(m, 1.0);
  Upload(d_avg, ones, stream);
  Upload(d_bt, bt, stream);
  Sync();
  prepared = false;
  InvalidateGraphs();
}

template <typename T>
TileArgs<T> octlb_dev::MakeTileArgs(int l) {
  TileArgs<T> a{};
  a.tiles = d_tiles + tile_begin[l];
  a.ntiles = tile_count[l];
  a.bt = d_bt;
  a.kind = d_kind;
  a.rule = d_rule;
  a.src = static_cast<const T*>(pops[par[l]]);
  a.dst = static_cast<T*>(pops[par[l] ^ 1]);
  a.shell_src = static_cast<const T*>(shells[par[l]]);
  a.shell_dst = static_cast<T*>(shells[par[l] ^ 1]);
  a.qmask = d_qmask;
  a.qoff = d_qoff;
  a.qval = d_qval;
  a.dmask = d_dmask;
  a.doff = d_doff;
  a.snap = static_cast<T*>(d_snap);
  a.stats = d_stats;
  a.block_rank = d_block_rank;
  a.stat_val = d_stat_val;
  a.stat_use = d_stat_use;
  {
    const int tpa = (NP + kTile - 1) / kTile;
    a.tiles_per_block = padded ? tpa * tpa * tpa : 1;
  }
  a.avg_rho = d_avg;
  a.scm = (decltype(a.scm))d_scm;  // read-only table (constant address space on the device)
  a.step = d_step + step_slot;
  a.want_stats = (sc.const_rho && l >= max_level) ? (fast_stats ? 1 : 2) : 0;
  a.snap_on = 1;
  a.sc = sc;
  return a;
}

// Replays the fused collide+stream kernel of level l `reps` times on the
// current state: the input buffer is only read, the output goes to the idle
// buffer (overwritten by the next step anyway), and the drag snapshot and
// ConstRho statistics are left alone.  Returns the mean launch duration (ms).
template <typename T>
double octlb_dev::TimeKernel(int l, int reps) {
  TileArgs<T> a = MakeTileArgs<T>(l);
  a.want_stats = 0;
  a.snap_on = 0;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  LaunchTile<T>(a, stream);  // warm the code object and caches
  HIP_CHECK(hipEventRecord(e0, stream));
  for (int r = 0; r < reps; ++r) LaunchTile<T>(a, stream);
  HIP_CHECK(hipEventRecord(e1, stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return static_cast<double>(ms) / reps;
}

// Coupler arguments for a coupling step run by the update launch of level l
// (its cells are in that launch's output).  timing_only: the step's writes
// into level other_level go to that level's idle buffer instead (overwritten
// by its next update), so a timing replay leaves the state as it was.
template <typename T>
CoupleArgs<T> octlb_dev::EpilogueArgs(int l, bool timing_only, int other_level) const {
  const double tau = 1.0 / sc.omega;  // level_coupler.cpp:226-228
  CoupleArgs<T> e{};
  e.plan = d_plan;
  e.pops0 = static_cast<T*>(pops[0]);
  e.pops1 = static_cast<T*>(pops[1]);
  e.prev = static_cast<T*>(d_prev);
  e.prolong_scale = static_cast<T>((tau - 0.25) / tau);
  e.restrict_scale = static_cast<T>(tau / (tau - 0.25));
  e.par = Par();
  e.par.v[l] ^= 1;
  if (timing_only) e.par.v[other_level] ^= 1;
  return e;
}

// Mean duration (ms) of `reps` back-to-back replays, state unchanged:
// what = 0: level l's fused kernel with the half-time prolongation epilogue;
// 1: with the restriction epilogue; 2: the standalone half-time prolongation
// of coarse level l; 3: the standalone restriction from fine level l.  Every
// write goes to an idle buffer (the fused kernel's output as in TimeKernel,
// the coupled level's cells through EpilogueArgs' redirection).
template <typename T>
double octlb_dev::TimeLaunches(int l, int reps, int what) {
  auto launch = [&]() {
    if (what <= 1) {
      TileArgs<T> a = MakeTileArgs<T>(l);
      a.want_stats = 0;
      a.snap_on = 0;
      if (what == 0) {
        a.ep_pro_beg = d_ep_pro_beg[l];
        a.ep_pro_pts = d_ep_pro_pts[l];
        a.ep_pro = EpilogueArgs<T>(l, true, l + 1);
      } else {
        a.ep_res_beg = d_ep_res_beg[l];
        a.ep_res_pts = d_ep_res_pts[l];
        a.ep_res = EpilogueArgs<T>(l, true, l - 1);
      }
      LaunchTile<T>(a, stream);
    } else {
      // the standalone kernels read the current state; their writes go idle
      CoupleArgs<T> c = EpilogueArgs<T>(l, true, what == 2 ? l + 1 : l - 1);
      c.par.v[l] ^= 1;  // reads of level l: the current state, as after its update
      if (what == 2) {
        LaunchProlong<T>(c, d_prolong + prolong_off[l], static_cast<int>(prolong_active[l].size()), 1, stream);
      } else {
        LaunchRestrict<T>(c, d_restrict + restrict_off[l], static_cast<int>(restrict_active[l].size()), stream);
      }
    }
    HIP_CHECK(hipGetLastError());
  };
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  launch();
  HIP_CHECK(hipEventRecord(e0, stream));
  for (int r = 0; r < reps; ++r) launch();
  HIP_CHECK(hipEventRecord(e1, stream));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return static_cast<double>(ms) / reps;
}

template <typename T>
void octlb_dev::Level(int l) {
  const int cnt = tile_count[l];
  if (cnt > 0) {
    TileArgs<T> a = MakeTileArgs<T>(l);
    a.step_next = TakeInc(true);
    a.fd_rep_level = -1;
    if (pfull_fuse_level == l) {
      const double tau = 1.0 / sc.omega;  // level_coupler.cpp:226-228
      CoupleArgs<T> c{};
      c.plan = d_plan;
      c.pops0 = static_cast<T*>(pops[0]);
      c.pops1 = static_cast<T*>(pops[1]);
      c.prev = static_cast<T*>(d_prev);
      c.prolong_scale = static_cast<T>((tau - 0.25) / tau);
      c.par = Par();
      c.par.v[l] ^= 1;  // the fine targets are in this launch's output buffer
      a.pfull = d_pfull;
      a.pc = c;
      a.pc_list = d_prolong + prolong_off[l - 1];
      a.pc_count = static_cast<int>(prolong_active[l - 1].size());
      a.pc_wgs = ((a.pc_count + 1023) / 1024 + 7) & ~7;  // keep the tiles' XCD placement
      pfull_fuse_level = -1;
      pfull_fused = true;
    }
    if (fd_pending) {
      if (fd_pend_level >= 0 && fd_pend_level != l && fd_other_cnt[l] > 0) {
        // a pass deferred past the coupling step: it runs first, in the
        // same workgroups as this launch's pass over those blocks
        a.fd_rep_level = fd_pend_level;
        fd_pending = false;
      } else {
        FlushFd<T>();
      }
    }
    if (ep_ok) {
      if ((ep_mode & 1) && l < max_level && d_ep_pro_beg[l] != nullptr) {
        a.ep_pro_beg = d_ep_pro_beg[l];
        a.ep_pro_pts = d_ep_pro_pts[l];
        a.ep_pro = EpilogueArgs<T>(l, false, l + 1);
        pro_done = true;
      }
      // restriction: only from the finest level, whose second update is the
      // last thing before it (a coarser level's nested visits come between)
      if ((ep_mode & 2) && l >= 1 && l == max_level && second_visit[l] && d_ep_res_beg[l] != nullptr) {
        a.ep_res_beg = d_ep_res_beg[l];
        a.ep_res_pts = d_ep_res_pts[l];
        a.ep_res = EpilogueArgs<T>(l, false, l - 1);
        res_done = true;
      }
    }
    if (!padded && fd_other_cnt[l] > 0) {
      // this level's FD pass over boundary blocks of other levels does not
      // depend on this level's update: it rides in the same launch
      a.fd = MakeFdArgs<T>();
      a.fd_blocks = d_fd_blocks;
      a.fd_slot = d_fd_slot;
      a.fd_idx = d_fd_idx + fd_other_off[l];
      a.fd_cnt = fd_other_cnt[l];
      a.fd_wgs = (fd_other_cnt[l] + 7) & ~7;  // keep the tiles' XCD placement
    }
    const bool timed = timing && !capturing;
    if (timed) {
      while (ev.size() < ev_used + 2) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        ev.push_back(e);
      }
      HIP_CHECK(hipEventRecord(ev[ev_used], stream));
    }
    LaunchTile<T>(a, stream);
    HIP_CHECK(hipGetLastError());
    if (timed) {
      HIP_CHECK(hipEventRecord(ev[ev_used + 1], stream));
      ev_used += 2;
      tile_launches += 1;
      const int tpa = (NP + kTile - 1) / kTile;
      const int tpb = padded ? tpa * tpa * tpa : 1;
      (void)tpb;
      tile_updates += static_cast<double>(level_blocks[l]) * static_cast<double>(ncore);
    }
    par[l] ^= 1;
    if (a.want_stats == 1) LaunchStatsFinalize<T>(d_stats, d_stat_tiles, nmembers, d_avg, stream);
    if (a.want_stats == 2) LaunchStatsSequential(d_stat_val, d_stat_use, d_stat_cells, nmembers, d_avg, stream);
  }
  if (fd_phase[3] > 0) {
    if (padded) {
      FdArgs<T> f = MakeFdArgs<T>();
      f.step_next = TakeInc(true);
      LaunchFd<T>(f, fd_phase, stream);
    } else {
      // the FD pass over this level's own boundary blocks (all of them when
      // the level has none to update) follows the update: it is launched
      // with the coupling step that comes next, or alone
      const bool split = cnt > 0;
      fd_pend_idx = d_fd_desc + (split ? fd_same_off[l] : fd_all_off);
      fd_pend_cnt = split ? fd_same_cnt[l] : nfd_blocks;
      for (int c = 0; c < kFdClasses; ++c) fd_pend_cls[c] = split ? fd_same_cls[l][c] : fd_all_cls[c];
      fd_pend_level = split ? l : -1;
      fd_pending = fd_pend_cnt > 0;
    }
  }
  if (l >= max_level) return;
  CoupleArgs<T> c{};
  c.plan = d_plan;
  c.pops0 = static_cast<T*>(pops[0]);
  c.pops1 = static_cast<T*>(pops[1]);
  c.prev = static_cast<T*>(d_prev);
  const double tau = 1.0 / sc.omega;  // level_coupler.cpp:226-228
  c.prolong_scale = static_cast<T>((tau - 0.25) / tau);
  c.restrict_scale = static_cast<T>(tau / (tau - 0.25));
  c.step = d_step + step_slot;
  const int npro = static_cast<int>(prolong_active[l].size());
  const int nres = static_cast<int>(restrict_active[l + 1].size());
  // When level l+1 is the finest and has no FD blocks of its own, its first
  // update launch is the last thing before the full-time prolongation and
  // does that prolongation itself (nothing reads the fine interface cells in
  // between; FD blocks and coupled blocks are disjoint).
  const bool fuse_full = !padded && l + 1 == max_level && npro > 0 && tile_count[l + 1] > 0 &&
                         fd_same_cnt[l + 1] == 0 && (nfd_blocks == 0 || fd_merge_ok);
  Couple<T>(c, d_prolong + prolong_off[l], npro, 1, l + 1);
  pfull_fuse_level = fuse_full ? l + 1 : -1;
  pfull_fused = false;
  Level<T>(l + 1);
  if (!pfull_fused) Couple<T>(c, d_prolong + prolong_off[l], npro, 0, l + 1);
  pfull_fused = false;
  pfull_fuse_level = -1;
  second_visit[l + 1] = true;
  Level<T>(l + 1);
  second_visit[l + 1] = false;
  Couple<T>(c, d_restrict + restrict_off[l + 1], nres, 2, -1);
}

template <typename T>
FdArgs<T> octlb_dev::MakeFdArgs() {
  FdArgs<T> f{};
  f.cells = d_fd_cells;
  f.smap = d_fd_smap;
  f.fd_index = d_fd_index;
  f.level = d_level;
  f.bt = d_bt;
  f.kind = d_kind;
  f.pops0 = static_cast<T*>(pops[0]);
  f.pops1 = static_cast<T*>(pops[1]);
  f.post = static_cast<T*>(d_fd_post);
  f.step = d_step + step_slot;
  f.scm = (decltype(f.scm))d_scm;
  f.fd_list = d_fd_list;
  f.fd_count = d_fd_count;
  f.par = Par();
  f.sc = sc;
  return f;
}

// Launches a deferred FD pass on its own.
template <typename T>
void octlb_dev::FlushFd() {
  if (!fd_pending) return;
  fd_pending = false;
  FdArgs<T> f = MakeFdArgs<T>();
  f.step_next = TakeInc(true);
  LaunchFdBlocks<T>(f, d_fd_slot, fd_pend_idx, fd_pend_cls, stream);
}

// One coupling step (mode 1 = prolongation at half time, 0 = at full time,
// 2 = restriction); a deferred FD pass rides in the same launch when its
// blocks and the coupled blocks are disjoint.
template <typename T>
void octlb_dev::Couple(CoupleArgs<T> c, const int* list, int count, int mode, int next_level) {
  c.par = Par();
  // a pending pass over one level's FD blocks can wait for the next update
  // launch when that launch is of another level (it then runs beside it)
  const bool defer = fd_pending && fd_merge_ok && fd_pend_level >= 0 && next_level >= 0 &&
                     next_level != fd_pend_level && tile_count[next_level] > 0 &&
                     fd_other_cnt[next_level] > 0;
  if ((mode == 1 && pro_done) || (mode == 2 && res_done)) {
    // the coupling step ran in the update launch before; a pending FD pass
    // still needs a launch unless the next update carries it
    pro_done = mode == 1 ? false : pro_done;
    res_done = mode == 2 ? false : res_done;
    if (!defer) FlushFd<T>();
    return;
  }
  if (defer) {
    c.step_next = TakeInc(count > 0);
    if (mode == 2) {
      LaunchRestrict<T>(c, list, count, stream);
    } else {
      LaunchProlong<T>(c, list, count, mode, stream);
    }
    return;
  }
  if (fd_pending && fd_merge_ok) {
    fd_pending = false;
    FdArgs<T> f = MakeFdArgs<T>();
    f.step_next = TakeInc(true);
    LaunchFdCouple<T>(f, d_fd_slot, fd_pend_idx, fd_pend_cnt, c, list, count, mode, stream);
    return;
  }
  FlushFd<T>();
  c.step_next = TakeInc(count > 0);
  if (mode == 2) {
    LaunchRestrict<T>(c, list, count, stream);
  } else {
    LaunchProlong<T>(c, list, count, mode, stream);
  }
}

// One coarse step.  The step counter lives in two slots: this step's kernels
// read step_slot, the first kernel launched also writes step_slot ^ 1.
// Level 0's buffer parity flips every coarse step, even when level 0 has no
// blocks, so (par[0], step_slot) is the same for every replay of a graph.
template <typename T>
void octlb_dev::CoarseStep() {
  step_slot = par[0];
  inc_next = d_step + (step_slot ^ 1);
  Level<T>(0);
  FlushFd<T>();
  if (inc_next) {
    LaunchStepInc(d_step + step_slot, inc_next, stream);
    inc_next = nullptr;
  }
  par[0] = step_slot ^ 1;
  step_slot = par[0];
}

void octlb_dev::Step() {
  if (prec == 64) {
    CoarseStep<double>();
  } else {
    CoarseStep<float>();
  }
  ++step;
}

// ----------------------------------------------------------------- C ABI
namespace {

template <typename F>
int Guard(F&& f) {
  try {
    f();
    return OCTLB_OK;
  } catch (const ArgError& e) {
    return SetError(OCTLB_EINVAL, e.what());
  } catch (const StateError& e) {
    return SetError(OCTLB_ESTATE, e.what());
  } catch (const HipError& e) {
    return SetError(OCTLB_EHIP, e.what());
  } catch (const std::runtime_error& e) {  // output files (vtk_out.cpp)
    return SetError(OCTLB_EIO, e.what());
  } catch (const std::bad_alloc&) {
    return SetError(OCTLB_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return SetError(OCTLB_EINVAL, e.what());
  }
}

void EnsureReady(octlb_dev* d) {
  if (!d->prepared) d->Prepare();
  if (!d->coupler_ready) {
    if (d->npts > 0) {
      if (d->prec == 64) {
        CoupleArgs<double> c{};
        c.plan = d->d_plan;
        c.pops0 = static_cast<double*>(d->pops[0]);
        c.pops1 = static_cast<double*>(d->pops[1]);
        c.prev = static_cast<double*>(d->d_prev);
        c.par = d->Par();
        LaunchCouplerInit<double>(c, d->npts, d->stream);
      } else {
        CoupleArgs<float> c{};
        c.plan = d->d_plan;
        c.pops0 = static_cast<float*>(d->pops[0]);
        c.pops1 = static_cast<float*>(d->pops[1]);
        c.prev = static_cast<float*>(d->d_prev);
        c.par = d->Par();
        LaunchCouplerInit<float>(c, d->npts, d->stream);
      }
      HIP_CHECK(hipGetLastError());
    }
    d->coupler_ready = true;
  }
}

template <typename T>
void ToHostBlock(octlb_dev* d, int b, double* aos) {
  const int buf = d->par[d->level[b]];
  const size_t NS = static_cast<size_t>(d->NS);
  std::vector<T> tmp(static_cast<size_t>(kQ) * d->plane);
  HIP_CHECK(hipMemcpy(tmp.data(), d->Core(buf, b), tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
  std::fill(aos, aos + NS * kQ, 0.0);
  if (d->padded) {
    for (size_t s = 0; s < NS; ++s)
      for (int i = 0; i < kQ; ++i) aos[s * kQ + i] = static_cast<double>(tmp[i * NS + s]);
    return;
  }
  for (int c = 0; c < d->ncore; ++c) {
    const int x = c / 64, y = (c / 8) % 8, z = c % 8;
    const size_t s = static_cast<size_t>(((x + 1) * kBox + y + 1) * kBox + z + 1);
    for (int i = 0; i < kQ; ++i) aos[s * kQ + i] = static_cast<double>(tmp[static_cast<size_t>(i) * d->ncore + c]);
  }
  const std::vector<int> slots = ShellSlots();
  const int sh = d->bt[b].shell;
  if (sh >= 0) {
    std::vector<T> st(static_cast<size_t>(kQ) * kApron);
    HIP_CHECK(hipMemcpy(st.data(), static_cast<char*>(d->shells[buf]) + static_cast<size_t>(sh) * kQ * kApron * sizeof(T),
                        st.size() * sizeof(T), hipMemcpyDeviceToHost));
    for (int j = 0; j < kApron; ++j)
      for (int i = 0; i < kQ; ++i) aos[static_cast<size_t>(slots[j]) * kQ + i] = static_cast<double>(st[static_cast<size_t>(i) * kApron + j]);
  }
  // Exchanged halo slots: the values the next ghost exchange writes there
  // (the neighbour's boundary layer).  The step never reads these slots
  // from the shell, so they are not kept on the device.
  std::vector<T> nbuf(static_cast<size_t>(kQ) * d->ncore);
  int loaded = -1;
  for (int j = 0; j < kApron; ++j) {
    const int s = slots[j];
    const int x = s / (kBox * kBox) - 1, y = (s / kBox) % kBox - 1, z = s % kBox - 1;
    int cc = 0;
    const int nbk = ApronNeighbour(d->bt[b], x, y, z, &cc);
    if (nbk < 0) continue;
    if (nbk != loaded) {
      HIP_CHECK(hipMemcpy(nbuf.data(), d->Core(d->par[d->level[nbk]], nbk), nbuf.size() * sizeof(T), hipMemcpyDeviceToHost));
      loaded = nbk;
    }
    for (int i = 0; i < kQ; ++i) aos[static_cast<size_t>(s) * kQ + i] = static_cast<double>(nbuf[static_cast<size_t>(i) * d->ncore + cc]);
  }
}

template <typename T>
void FromHostBlock(octlb_dev* d, int b, const double* aos) {
  const int buf = d->par[d->level[b]];
  const size_t NS = static_cast<size_t>(d->NS);
  std::vector<T> tmp(static_cast<size_t>(kQ) * d->plane);
  if (d->padded) {
    for (size_t s = 0; s < NS; ++s)
      for (int i = 0; i < kQ; ++i) tmp[i * NS + s] = static_cast<T>(aos[s * kQ + i]);
  } else {
    for (int c = 0; c < d->ncore; ++c) {
      const int x = c / 64, y = (c / 8) % 8, z = c % 8;
      const size_t s = static_cast<size_t>(((x + 1) * kBox + y + 1) * kBox + z + 1);
      for (int i = 0; i < kQ; ++i) tmp[static_cast<size_t>(i) * d->ncore + c] = static_cast<T>(aos[s * kQ + i]);
    }
  }
  HIP_CHECK(hipMemcpy(d->Core(buf, b), tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
  const int sh = d->bt[b].shell;
  if (!d->padded && sh >= 0) {
    std::vector<T> st(static_cast<size_t>(kQ) * kApron);
    const std::vector<int> slots = ShellSlots();
    for (int j = 0; j < kApron; ++j)
      for (int i = 0; i < kQ; ++i) st[static_cast<size_t>(i) * kApron + j] = static_cast<T>(aos[static_cast<size_t>(slots[j]) * kQ + i]);
    HIP_CHECK(hipMemcpy(static_cast<char*>(d->shells[buf]) + static_cast<size_t>(sh) * kQ * kApron * sizeof(T), st.data(),
                        st.size() * sizeof(T), hipMemcpyHostToDevice));
  }
}

template <typename T>
void CoreToHost(octlb_dev* d, double* out) {
  std::vector<T> tmp(static_cast<size_t>(kQ) * d->plane);
  for (int b = 0; b < d->nb; ++b) {
    const int buf = d->par[d->level[b]];
    HIP_CHECK(hipMemcpy(tmp.data(), d->Core(buf, b), tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
    for (int c = 0; c < d->ncore; ++c) {
      const int s = CoreSlot(d->sc, c);
      double* o = out + (static_cast<size_t>(b) * d->ncore + c) * kQ;
      for (int i = 0; i < kQ; ++i) o[i] = static_cast<double>(tmp[static_cast<size_t>(i) * d->plane + s]);
    }
  }
}

template <typename T>
void Drag(octlb_dev* d, const double* bbox, double* out) {
  DragArgs<T> a{};
  a.links = d->d_dlinks;
  a.bt = d->d_bt;
  a.pops0 = static_cast<const T*>(d->pops[0]);
  a.pops1 = static_cast<const T*>(d->pops[1]);
  a.snap = d->snapshot ? static_cast<const T*>(d->d_snap) : nullptr;
  a.term = d->d_dterm;
  a.use = d->d_duse;
  for (int k = 0; k < 6; ++k) a.bbox[k] = bbox[k];
  a.par = d->Par();
  a.sc = d->sc;
  if (d->ndblocks == 0) {  // no link in any block: the sum of nothing
    for (int k = 0; k < 3; ++k) out[k] = 0.0;
    return;
  }
  // the last drag workgroup writes the three sums straight into pinned host memory
  LaunchDrag<T>(a, d->d_dblock_begin, d->ndblocks, d->d_dblock_out, d->d_drag_done, d->h_out, d->stream);
  HIP_CHECK(hipGetLastError());
  d->Sync();
  for (int k = 0; k < 3; ++k) out[k] = d->h_out[k];
}

template <typename T>
void Mass(octlb_dev* d, double* out) {
  LaunchMass<T>(static_cast<const T*>(d->pops[0]), static_cast<const T*>(d->pops[1]), d->Par(), d->d_bt,
                d->d_kind, d->sc, d->nb, d->d_mass_partial, d->d_out + 3, d->stream);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(d->h_out + 3, d->d_out + 3, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  d->Sync();
  *out = d->h_out[3];
}

// rho [nb][ncore], u [nb][ncore][3] in VTK cell order, on the host
template <typename T>
void MacroFields(octlb_dev* d, double* rho, double* u) {
  const size_t cells = static_cast<size_t>(d->nb) * d->ncore;
  double* buf = DevAlloc<double>(4 * cells);
  try {
    LaunchMacro<T>(static_cast<const T*>(d->pops[0]), static_cast<const T*>(d->pops[1]), d->Par(), d->d_bt, d->sc,
                   d->nb, buf, buf + cells, d->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(rho, buf, cells * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipMemcpyAsync(u, buf + cells, 3 * cells * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    d->Sync();
  } catch (...) {
    (void)hipStreamSynchronize(d->stream);
    (void)hipFree(buf);
    throw;
  }
  HIP_CHECK(hipFree(buf));
}

void CheckDev(const octlb_dev* d) {
  if (d == nullptr) throw ArgError("null device handle");
}

}  // namespace

namespace octlb {

void ExportAdjTopo(const octlb_dev* d, AdjTopo* t) {
  if (!d || !t) throw std::invalid_argument("null argument");
  if (d->padded || d->n != kTile || d->h != 1)
    throw std::invalid_argument("adjoint: needs blocks of 8^3 cells with a one-cell halo");
  for (int lv : d->level)
    if (lv != d->max_level) throw std::invalid_argument("adjoint: one level only (no coupling adjoint)");
  if (d->nmembers != 1) throw std::invalid_argument("adjoint: ensembles are not differentiated");
  if (d->sc.handler != OCTLB_HANDLER_CONCRETE || d->sc.const_rho || !d->sc.bouzidi)
    throw std::invalid_argument("adjoint: needs the concrete handler, plain BGK and Bouzidi streaming");
  for (uint8_t k : d->kinds)
    if (IsFd(k)) throw std::invalid_argument("adjoint: FD boundary kinds are not differentiated");
  constexpr int NP = kTile + 2, NS = NP * NP * NP;
  t->nb = d->nb;
  t->prec = d->prec;
  t->omega = d->sc.omega;
  t->kinds = d->kinds;
  t->link_key.clear();
  t->link_q.clear();
  for (const auto& kv : d->links) {
    t->link_key.push_back(kv.first);
    t->link_q.push_back(kv.second);
  }
  t->origin_cw.resize(static_cast<size_t>(d->nb) * 4);
  t->halo_src.assign(static_cast<size_t>(d->nb) * NS, -1);
  for (int b = 0; b < d->nb; ++b) {
    const BlockTab& r = d->bt[b];
    for (int a = 0; a < 3; ++a) t->origin_cw[4 * b + a] = r.origin[a];
    t->origin_cw[4 * b + 3] = r.cw;
    for (int px = 0; px < NP; ++px)
      for (int py = 0; py < NP; ++py)
        for (int pz = 0; pz < NP; ++pz) {
          const int x = px - 1, y = py - 1, z = pz - 1;
          const bool core = x >= 0 && x < kTile && y >= 0 && y < kTile && z >= 0 && z < kTile;
          if (core) continue;
          int cc = 0;
          const int nbk = ApronNeighbour(r, x, y, z, &cc);
          if (nbk < 0) continue;
          const int cx = cc / (kTile * kTile), cy = (cc / kTile) % kTile, cz = cc % kTile;
          t->halo_src[static_cast<size_t>(b) * NS + (px * NP + py) * NP + pz] =
              nbk * NS + ((cx + 1) * NP + cy + 1) * NP + cz + 1;
        }
  }
}

bool AdjNeedsMultiLevel(const octlb_dev* d) {
  if (!d) return false;
  for (int lv : d->level)
    if (lv != d->max_level) return true;
  for (uint8_t k : d->kinds)
    if (IsFd(k)) return true;
  return false;
}

void ExportAdjMLTopo(octlb_dev* d, AdjTopo* t, AdjMLTopo* m) {
  if (!d || !t || !m) throw std::invalid_argument("null argument");
  if (!d->prepared) d->Prepare();
  if (d->padded || d->n != kTile || d->h != 1)
    throw std::invalid_argument("adjoint: needs blocks of 8^3 cells with a one-cell halo");
  if (d->nmembers != 1) throw std::invalid_argument("adjoint: ensembles are not differentiated");
  if (d->sc.handler != OCTLB_HANDLER_CONCRETE || d->sc.const_rho || !d->sc.bouzidi)
    throw std::invalid_argument("adjoint: needs the concrete handler, plain BGK and Bouzidi streaming");
  constexpr int NP = kTile + 2, NS = NP * NP * NP;
  t->nb = d->nb;
  t->prec = d->prec;
  t->omega = d->sc.omega;
  t->kinds = d->kinds;
  t->link_key.clear();
  t->link_q.clear();
  for (const auto& kv : d->links) {
    t->link_key.push_back(kv.first);
    t->link_q.push_back(kv.second);
  }
  t->origin_cw.resize(static_cast<size_t>(d->nb) * 4);
  t->halo_src.assign(static_cast<size_t>(d->nb) * NS, -1);
  for (int b = 0; b < d->nb; ++b) {
    const BlockTab& r = d->bt[b];
    for (int a = 0; a < 3; ++a) t->origin_cw[4 * b + a] = r.origin[a];
    t->origin_cw[4 * b + 3] = r.cw;
    for (int px = 0; px < NP; ++px)
      for (int py = 0; py < NP; ++py)
        for (int pz = 0; pz < NP; ++pz) {
          const int x = px - 1, y = py - 1, z = pz - 1;
          if (x >= 0 && x < kTile && y >= 0 && y < kTile && z >= 0 && z < kTile) continue;
          int cc = 0;
          const int nbk = ApronNeighbour(r, x, y, z, &cc);
          if (nbk < 0) continue;
          const int cx = cc / (kTile * kTile), cy = (cc / kTile) % kTile, cz = cc % kTile;
          t->halo_src[static_cast<size_t>(b) * NS + (px * NP + py) * NP + pz] =
              nbk * NS + ((cx + 1) * NP + cy + 1) * NP + cz + 1;
        }
  }
  t->level = d->level;
  m->max_level = d->max_level;
  m->level = d->level;
  m->plan = d->plan;
  m->prolong_active.assign(d->prolong_active.begin(), d->prolong_active.end());
  m->restrict_active.assign(d->restrict_active.begin(), d->restrict_active.end());
  const double tau = 1.0 / d->sc.omega;  // level_coupler.cpp:226-228
  m->prolong_scale = (tau - 0.25) / tau;
  m->restrict_scale = tau / (tau - 0.25);
  m->fd_blocks = d->fd_blocks_host;
  m->fd_slot = d->fd_slot_host;
  m->fd_list = d->fd_list_host;
  m->fd_count = d->fd_count_host;
  m->bt = d->bt;
  m->sc = d->sc;
}

}  // namespace octlb

extern "C" {

const char* octlb_last_error(void) { return g_last_error.c_str(); }

int octlb_dev_create(const octlb_topology* topo, const octlb_params* params, octlb_dev** out) {
  if (out) *out = nullptr;
  if (!topo || !params || !out) return SetError(OCTLB_EINVAL, "null argument");
  octlb_dev* d = new (std::nothrow) octlb_dev();
  if (!d) return SetError(OCTLB_ENOMEM, "host allocation failed");
  const int rc = Guard([&] { d->Build(topo, params); });
  if (rc != OCTLB_OK) {
    d->FreeAll();
    delete d;
    return rc;
  }
  *out = d;
  return OCTLB_OK;
}

int octlb_dev_create_ensemble(const octlb_topology* topo, int32_t num_members, const int32_t* member_begin,
                              const octlb_params* params, octlb_dev** out) {
  if (out) *out = nullptr;
  if (!topo || !params || !member_begin || !out) return SetError(OCTLB_EINVAL, "null argument");
  octlb_dev* d = new (std::nothrow) octlb_dev();
  if (!d) return SetError(OCTLB_ENOMEM, "host allocation failed");
  const int rc = Guard([&] {
    d->Build(topo, &params[0]);
    d->SetMembers(num_members, member_begin, params);
  });
  if (rc != OCTLB_OK) {
    d->FreeAll();
    delete d;
    return rc;
  }
  *out = d;
  return OCTLB_OK;
}

int octlb_dev_member_average_rho(octlb_dev* d, int32_t member, double* out) {
  return Guard([&] {
    CheckDev(d);
    if (!out) throw ArgError("null argument");
    if (member < 0 || member >= d->nmembers) throw ArgError("member out of range");
    d->Sync();
    HIP_CHECK(hipMemcpy(out, d->d_avg + member, sizeof(double), hipMemcpyDeviceToHost));
  });
}

// Moves the state of src onto dst, a handle built for the forest after an
// adapt (same block edge, halo, precision and relaxation rate).  Per dst
// block: a src block with the same level and bounds is copied whole (its
// halo-inclusive view); a block refined from a src block takes, in each
// fine cell, the parent cell's olb (rho, u) and non-equilibrium part with the
// coupler's prolongation rule f = feq + s_p fneq, s_p = (tau - 1/4)/tau
// (level_coupler.cpp:163-214); a block coarsened from eight src blocks takes,
// in each coarse cell, the mean rho, the momentum-conserving mean u and the
// mean fneq of its 2x2x2 fine cells times s_r = tau/(tau - 1/4).  The
// reference leaves re-adaptation to a future rebuild step (it has no
// migration code), so the two interpolation rules are this library's.
int octlb_dev_migrate(octlb_dev* dst, octlb_dev* src) {
  return Guard([&] {
    CheckDev(dst);
    CheckDev(src);
    if (dst->n != src->n || dst->h != src->h || dst->prec != src->prec || dst->sc.omega != src->sc.omega)
      throw ArgError("migrate: handles differ in block edge, halo, precision or relaxation rate");
    if (dst->nmembers != 1 || src->nmembers != 1)
      throw ArgError("migrate: ensemble handles are not migrated (members share block bounds)");
    if (dst->nb == 0 || src->nb == 0) throw ArgError("migrate: empty forest");
    src->Sync();
    dst->Sync();
    // Blocks of both forests keyed by (level, integer origin in that level's
    // block widths): a hash lookup per destination block, so the matching is
    // linear in the block count.  The populations move on the device.
    const double* sb = src->bounds.data();
    const double* db = dst->bounds.data();
    double gmin[3] = {db[0], db[1], db[2]};
    for (int k = 0; k < 3; ++k) {
      for (int b = 0; b < dst->nb; ++b) gmin[k] = std::min(gmin[k], db[6 * static_cast<size_t>(b) + k]);
      for (int b = 0; b < src->nb; ++b) gmin[k] = std::min(gmin[k], sb[6 * static_cast<size_t>(b) + k]);
    }
    const double w0 = (db[3] - db[0]) * std::ldexp(1.0, dst->level[0]);  // level-0 block width
    // key = level (5 bits) | x, y, z index of the origin in that level's
    // block widths (19 bits each)
    constexpr int64_t kIdxLimit = int64_t(1) << 19;
    auto pack = [](int64_t lv, int64_t ix, int64_t iy, int64_t iz) {
      return (static_cast<uint64_t>(lv) << 57) | (static_cast<uint64_t>(ix) << 38) |
             (static_cast<uint64_t>(iy) << 19) | static_cast<uint64_t>(iz);
    };
    auto key_of = [&](const double* bb, int lv) {
      const double w = std::ldexp(w0, -lv);
      if (std::fabs((bb[3] - bb[0]) - w) > 1e-9 * w0) throw ArgError("migrate: block width does not follow its level");
      if (lv < 0 || lv >= 31) throw ArgError("migrate: level out of range");
      std::array<int64_t, 4> key{lv, 0, 0, 0};
      for (int k = 0; k < 3; ++k) {
        const double c = (bb[k] - gmin[k]) / w;
        key[1 + k] = std::llround(c);
        if (std::fabs(c - static_cast<double>(key[1 + k])) > 1e-6) throw ArgError("migrate: block origin off its level's grid");
        if (key[1 + k] < 0 || key[1 + k] >= kIdxLimit / 2) throw ArgError("migrate: forest wider than 2^18 blocks per level");
      }
      return key;
    };
    // open-addressing table of src blocks (linear probing, power-of-two size)
    size_t cap = 16;
    while (cap < static_cast<size_t>(src->nb) * 2) cap <<= 1;
    constexpr uint64_t kEmpty = ~uint64_t(0);
    std::vector<uint64_t> keys(cap, kEmpty);
    std::vector<int> vals(cap, -1);
    const int shift = 64 - __builtin_ctzll(cap);
    auto slot_of = [&](uint64_t k) { return static_cast<size_t>((k * 0x9E3779B97F4A7C15ull) >> shift); };
    for (int s = 0; s < src->nb; ++s) {
      const auto k4 = key_of(sb + 6 * static_cast<size_t>(s), src->level[s]);
      const uint64_t k = pack(k4[0], k4[1], k4[2], k4[3]);
      size_t i = slot_of(k);
      while (keys[i] != kEmpty && keys[i] != k) i = (i + 1) & (cap - 1);
      if (keys[i] == k) throw ArgError("migrate: two source blocks share level and bounds");
      keys[i] = k;
      vals[i] = s;
    }
    auto find = [&](const std::array<int64_t, 4>& k4) {
      const uint64_t k = pack(k4[0], k4[1], k4[2], k4[3]);
      for (size_t i = slot_of(k);; i = (i + 1) & (cap - 1)) {
        if (keys[i] == k) return vals[i];
        if (keys[i] == kEmpty) return -1;
      }
    };
    std::vector<MigTask> task(static_cast<size_t>(dst->nb));
    for (int b = 0; b < dst->nb; ++b) {
      const auto key = key_of(db + 6 * static_cast<size_t>(b), dst->level[b]);
      const int lv = static_cast<int>(key[0]);
      MigTask& t = task[static_cast<size_t>(b)];
      std::fill(t.src, t.src + 8, -1);
      t.oct[0] = t.oct[1] = t.oct[2] = 0;
      t.src[0] = find(key);
      if (t.src[0] >= 0) {
        t.mode = kMigCopy;
        continue;
      }
      t.src[0] = lv > 0 ? find({lv - 1, key[1] >> 1, key[2] >> 1, key[3] >> 1}) : -1;
      if (t.src[0] >= 0) {
        t.mode = kMigRefine;
        for (int k = 0; k < 3; ++k) t.oct[k] = static_cast<int>(key[1 + k] & 1);
        continue;
      }
      t.mode = kMigCoarsen;
      for (int c = 0; c < 8; ++c) {
        t.src[c] = find({lv + 1, 2 * key[1] + (c & 1), 2 * key[2] + ((c >> 1) & 1), 2 * key[3] + (c >> 2)});
        if (t.src[c] < 0)
          throw ArgError("migrate: block " + std::to_string(b) + " has no source block, parent or eight children");
      }
    }
    const double tau = 1.0 / dst->sc.omega;
    // the new blocks grouped by rule: copies, refinements, coarsenings
    int count[3] = {0, 0, 0};
    for (const MigTask& t : task) ++count[t.mode];
    if (dst->padded && count[0] != dst->nb)
      throw ArgError("migrate: refinement and coarsening need the blocked layout (8^3 blocks, halo 1)");
    // one upload: the tasks, then the block list of each rule
    const size_t task_bytes = task.size() * sizeof(MigTask);
    std::vector<char> stage(task_bytes + task.size() * sizeof(int));
    std::memcpy(stage.data(), task.data(), task_bytes);
    int* list = reinterpret_cast<int*>(stage.data() + task_bytes);
    int at[3] = {0, count[0], count[0] + count[1]};
    for (int b = 0; b < dst->nb; ++b) list[at[task[static_cast<size_t>(b)].mode]++] = b;
    char* d_stage = DevAlloc<char>(stage.size());
    MigTask* d_task = reinterpret_cast<MigTask*>(d_stage);
    int* d_list = reinterpret_cast<int*>(d_stage + task_bytes);
    try {
      HIP_CHECK(hipMemcpyAsync(d_stage, stage.data(), stage.size(), hipMemcpyHostToDevice, dst->stream));
      auto launch = [&](auto zero) {
        using T = decltype(zero);
        MigArgs<T> a{};
        a.task = d_task;
        a.src.bt = src->d_bt;
        a.dst.bt = dst->d_bt;
        for (int k = 0; k < 2; ++k) {
          a.src.pops[k] = static_cast<T*>(src->pops[k]);
          a.src.shells[k] = static_cast<T*>(src->shells[k]);
          a.dst.pops[k] = static_cast<T*>(dst->pops[k]);
          a.dst.shells[k] = static_cast<T*>(dst->shells[k]);
        }
        a.src.par = src->Par();
        a.dst.par = dst->Par();
        a.n = dst->n;
        a.h = dst->h;
        a.NP = dst->NP;
        a.NS = dst->NS;
        a.plane = dst->plane;
        a.padded = dst->padded ? 1 : 0;
        a.sp = (tau - 0.25) / tau;
        a.sr = tau / (tau - 0.25);
        LaunchMigrate<T>(a, d_list, count, dst->stream);
        HIP_CHECK(hipGetLastError());
      };
      if (dst->prec == 64) launch(0.0); else launch(0.0f);
      dst->Sync();
    } catch (...) {
      (void)hipFree(d_stage);
      throw;
    }
    HIP_CHECK(hipFree(d_stage));
    // the time state moves with the populations: the step counter (inlet
    // ramps) and the lagged ConstRho average
    dst->step = src->step;
    HIP_CHECK(hipMemcpyAsync(dst->d_step + dst->step_slot, &dst->step, sizeof(int), hipMemcpyHostToDevice,
                             dst->stream));
    HIP_CHECK(hipMemcpyAsync(dst->d_avg, src->d_avg, sizeof(double), hipMemcpyDeviceToDevice, dst->stream));
    // the coupler's previous-step macros (level_coupler.cpp:438-478) carry
    // over when the coupling plan is unchanged; otherwise the new plan's
    // points start from the current coarse state, as at construction
    if (dst->npts > 0 && dst->plan == src->plan && src->coupler_ready) {
      HIP_CHECK(hipMemcpyAsync(dst->d_prev, src->d_prev, static_cast<size_t>(dst->npts) * kMacro * dst->elem,
                               hipMemcpyDeviceToDevice, dst->stream));
      dst->coupler_ready = true;
    } else {
      dst->coupler_ready = false;
    }
    dst->Sync();
  });
}

int octlb_dev_destroy(octlb_dev* dev) {
  if (!dev) return OCTLB_OK;
  if (dev->stream) (void)hipStreamSynchronize(dev->stream);  // teardown: nothing to report to
  dev->FreeAll();
  delete dev;
  return OCTLB_OK;
}

int octlb_dev_set_bc_kinds(octlb_dev* d, const uint8_t* kinds) {
  return Guard([&] {
    CheckDev(d);
    if (!kinds) throw ArgError("null kinds");
    for (size_t i = 0; i < d->kinds.size(); ++i)
      if (kinds[i] > OCTLB_KIND_OUTFLOW) throw ArgError("BcKind byte out of range");
    d->kinds.assign(kinds, kinds + d->kinds.size());
    HIP_CHECK(hipMemcpyAsync(d->d_kind, d->kinds.data(), d->kinds.size(), hipMemcpyHostToDevice, d->stream));
    d->Sync();
    d->prepared = false;
  });
}

int octlb_dev_set_links(octlb_dev* d, int64_t count, const int32_t* link, const double* q) {
  return Guard([&] {
    CheckDev(d);
    if (count < 0 || (count > 0 && (!link || !q))) throw ArgError("bad link arrays");
    if (!d->sc.bouzidi) throw StateError("solver built without Bouzidi links (params.bouzidi = 0)");
    for (int64_t i = 0; i < count; ++i) {
      const int b = link[3 * i], c = link[3 * i + 1], p = link[3 * i + 2];
      if (b < 0 || b >= d->nb || c < 0 || c >= d->ncore || p < 0 || p >= kQ) throw ArgError("link out of range");
      d->links[(static_cast<int64_t>(b) * d->ncore + c) * kQ + p] = q[i];
    }
    d->prepared = false;
  });
}

int octlb_dev_init_equilibrium(octlb_dev* d, double rho0, const double u0[3], uint32_t mask) {
  return Guard([&] {
    CheckDev(d);
    if (!u0) throw ArgError("null u0");
    for (int buf = 0; buf < 2; ++buf) {
      // both copies: the state is then defined whatever each level's parity
      if (d->prec == 64) {
        LaunchInitEq<double>(static_cast<double*>(d->pops[buf]), d->d_kind, d->nb, d->sc, rho0, u0, mask, d->stream);
      } else {
        LaunchInitEq<float>(static_cast<float*>(d->pops[buf]), d->d_kind, d->nb, d->sc, rho0, u0, mask, d->stream);
      }
    }
    HIP_CHECK(hipGetLastError());
    d->Sync();
    d->coupler_ready = false;
  });
}

int octlb_dev_init_field(octlb_dev* d, const double* rho, const double* u) {
  return Guard([&] {
    CheckDev(d);
    if (!rho || !u) throw ArgError("null field");
    const size_t nc = static_cast<size_t>(d->nb) * d->ncore;
    double* dr = DevAlloc<double>(nc);
    double* du = DevAlloc<double>(3 * nc);
    HIP_CHECK(hipMemcpy(dr, rho, nc * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(du, u, 3 * nc * sizeof(double), hipMemcpyHostToDevice));
    for (int buf = 0; buf < 2; ++buf) {
      if (d->prec == 64) {
        LaunchInitField<double>(static_cast<double*>(d->pops[buf]), d->nb, d->sc, dr, du, d->stream);
      } else {
        LaunchInitField<float>(static_cast<float*>(d->pops[buf]), d->nb, d->sc, dr, du, d->stream);
      }
    }
    HIP_CHECK(hipGetLastError());
    d->Sync();
    HIP_CHECK(hipFree(dr));
    HIP_CHECK(hipFree(du));
    d->coupler_ready = false;
  });
}

int octlb_dev_upload_block(octlb_dev* d, int32_t b, const double* aos) {
  return Guard([&] {
    CheckDev(d);
    if (b < 0 || b >= d->nb || !aos) throw ArgError("bad block");
    d->Sync();
    if (d->prec == 64) {
      FromHostBlock<double>(d, b, aos);
    } else {
      FromHostBlock<float>(d, b, aos);
    }
    d->coupler_ready = false;
  });
}

int octlb_dev_download_block(octlb_dev* d, int32_t b, double* aos) {
  return Guard([&] {
    CheckDev(d);
    if (b < 0 || b >= d->nb || !aos) throw ArgError("bad block");
    d->Sync();
    if (d->prec == 64) {
      ToHostBlock<double>(d, b, aos);
    } else {
      ToHostBlock<float>(d, b, aos);
    }
  });
}

int octlb_dev_download_core(octlb_dev* d, double* out) {
  return Guard([&] {
    CheckDev(d);
    if (!out) throw ArgError("null output");
    d->Sync();
    if (d->prec == 64) {
      CoreToHost<double>(d, out);
    } else {
      CoreToHost<float>(d, out);
    }
  });
}

int octlb_dev_advance(octlb_dev* d, int32_t steps) {
  return Guard([&] {
    CheckDev(d);
    if (steps < 0) throw ArgError("negative step count");
    EnsureReady(d);
    d->ev_used = 0;
    d->tile_ms = 0.0;
    d->tile_launches = 0;
    d->tile_updates = 0.0;
    for (int s = 0; s < steps; ++s) d->Step();
    HIP_CHECK(hipGetLastError());
    d->Sync();
    if (d->timing) {
      double ms = 0.0;
      for (size_t k = 0; k + 1 < d->ev_used; k += 2) {
        float e = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&e, d->ev[k], d->ev[k + 1]));
        ms += e;
      }
      d->tile_ms = ms;
    }
  });
}

namespace {

// Captures and instantiates the graph of `count` coarse steps that starts
// from level-0 parity slot & 1 (slots 0, 1: one step; 2, 3: kGraphSteps).
// Capture only records launches: the device state and the host's parity and
// step-slot bookkeeping are as before.
void CaptureGraph(octlb_dev* d, int slot, int count) {
  int saved[kMaxLevels];
  std::memcpy(saved, d->par, sizeof(saved));
  const int saved_slot = d->step_slot;
  d->par[0] = slot & 1;
  d->step_slot = d->par[0];
  hipGraph_t g = nullptr;
  d->capturing = true;
  HIP_CHECK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < count; ++k) {
    if (d->prec == 64) {
      d->CoarseStep<double>();
    } else {
      d->CoarseStep<float>();
    }
  }
  HIP_CHECK(hipStreamEndCapture(d->stream, &g));
  d->capturing = false;
  std::memcpy(d->par, saved, sizeof(saved));
  d->step_slot = saved_slot;
  HIP_CHECK(hipGraphInstantiate(&d->gexec[slot], g, nullptr, nullptr, 0));
  HIP_CHECK(hipGraphDestroy(g));
}

}  // namespace

int octlb_dev_prepare_graphs(octlb_dev* d) {
  return Guard([&] {
    CheckDev(d);
    EnsureReady(d);
    for (int slot = 0; slot < 4; ++slot)
      if (!d->gexec[slot]) CaptureGraph(d, slot, slot < 2 ? 1 : octlb_dev::kGraphSteps);
    d->Sync();
  });
}

int octlb_dev_advance_graph(octlb_dev* d, int32_t steps) {
  return Guard([&] {
    CheckDev(d);
    if (steps < 0) throw ArgError("negative step count");
    EnsureReady(d);
    // Replays captured hipGraphs: kGraphSteps coarse steps per launch while
    // enough remain (one graph launch per coarse step would leave a launch gap
    // between every step), single steps for the rest.
    while (steps > 0) {
      const int p0 = d->par[0];
      if (steps >= octlb_dev::kGraphSteps) {
        // an even number of coarse steps leaves every level's parity as it was
        if (!d->gexec[2 + p0]) CaptureGraph(d, 2 + p0, octlb_dev::kGraphSteps);
        HIP_CHECK(hipGraphLaunch(d->gexec[2 + p0], d->stream));
        d->step += octlb_dev::kGraphSteps;
        steps -= octlb_dev::kGraphSteps;
        continue;
      }
      if (!d->gexec[p0]) CaptureGraph(d, p0, 1);
      HIP_CHECK(hipGraphLaunch(d->gexec[p0], d->stream));
      d->par[0] ^= 1;  // finer levels step an even number of times per coarse step
      d->step_slot = d->par[0];
      ++d->step;
      --steps;
    }
    d->Sync();
  });
}

int octlb_dev_synchronize(octlb_dev* d) {
  return Guard([&] {
    CheckDev(d);
    d->Sync();
  });
}

int octlb_dev_set_step(octlb_dev* d, int32_t step) {
  return Guard([&] {
    CheckDev(d);
    d->step = step;
    HIP_CHECK(hipMemcpyAsync(d->d_step + d->step_slot, &d->step, sizeof(int), hipMemcpyHostToDevice, d->stream));
    d->Sync();
  });
}

int octlb_dev_set_snapshot(octlb_dev* d, int32_t on) {
  return Guard([&] {
    CheckDev(d);
    if ((on != 0) != d->snapshot) {
      d->Sync();
      d->snapshot = on != 0;
      d->prepared = false;  // drag flags and snapshot slots are laid out by Prepare
    }
  });
}

int octlb_dev_drag(octlb_dev* d, const double bbox[6], double out[3]) {
  return Guard([&] {
    CheckDev(d);
    if (!bbox || !out) throw ArgError("null argument");
    if (!d->prepared) d->Prepare();
    // the drag launch follows the steps on the solver's stream: no wait first
    if (d->prec == 64) {
      Drag<double>(d, bbox, out);
    } else {
      Drag<float>(d, bbox, out);
    }
  });
}

int octlb_dev_mass(octlb_dev* d, double* out) {
  return Guard([&] {
    CheckDev(d);
    if (!out) throw ArgError("null argument");
    d->Sync();
    if (d->prec == 64) {
      Mass<double>(d, out);
    } else {
      Mass<float>(d, out);
    }
  });
}

int octlb_dev_macro_fields(octlb_dev* d, double* rho, double* u) {
  return Guard([&] {
    CheckDev(d);
    if (!rho || !u) throw ArgError("null argument");
    d->Sync();
    if (d->prec == 64) {
      MacroFields<double>(d, rho, u);
    } else {
      MacroFields<float>(d, rho, u);
    }
  });
}

int octlb_dev_write_vtk(octlb_dev* d, int32_t iT, const char* output_dir, const char* base_name) {
  return Guard([&] {
    CheckDev(d);
    if (!output_dir || !base_name) throw ArgError("null argument");
    const size_t nc = static_cast<size_t>(d->ncore);
    std::vector<double> rho(static_cast<size_t>(d->nb) * nc), u(3 * rho.size()), p(nc);
    d->Sync();
    if (d->prec == 64) {
      MacroFields<double>(d, rho.data(), u.data());
    } else {
      MacroFields<float>(d, rho.data(), u.data());
    }
    const std::string dir(output_dir), base(base_name), step = std::to_string(iT);
    std::vector<std::string> rel(static_cast<size_t>(d->nb));
    const double cs2 = 1.0 / 3.0;  // T{1} / invCs2 (vtk_lbm_fields.h:90-101)
    for (int b = 0; b < d->nb; ++b) {
      for (size_t c = 0; c < nc; ++c) p[c] = rho[b * nc + c] * cs2;
      rel[b] = base + "_block" + std::to_string(b) + "_T" + step + ".vts";
      const char* names[2] = {"velocity", "pressure"};
      const int32_t comps[2] = {3, 1};
      const double* data[2] = {u.data() + 3 * b * nc, p.data()};
      WriteVts(dir + "/" + rel[b], d->n, d->n, d->n, d->bounds.data() + 6 * static_cast<size_t>(b), 2, names, comps,
               data);
    }
    std::vector<const char*> relp(rel.size());
    for (size_t b = 0; b < rel.size(); ++b) relp[b] = rel[b].c_str();
    char tag[16];
    std::snprintf(tag, sizeof(tag), "%05d", iT);
    WriteVtm(dir + "/" + base + "_T" + tag + ".vtm", d->nb, relp.data());
    AppendPvd(dir + "/" + base + ".pvd", iT, base + "_T" + step + ".vtm");
  });
}

int octlb_dev_average_rho(octlb_dev* d, double* out) {
  return Guard([&] {
    CheckDev(d);
    if (!out) throw ArgError("null argument");
    d->Sync();
    HIP_CHECK(hipMemcpy(out, d->d_avg, sizeof(double), hipMemcpyDeviceToHost));
  });
}

int octlb_dev_update_counts(octlb_dev* d, double* actual, double* feq) {
  return Guard([&] {
    CheckDev(d);
    double a = 0.0, f = 0.0;
    for (int b = 0; b < d->nb; ++b) {
      const int l = d->level[b];
      a += static_cast<double>(d->ncore) * std::ldexp(1.0, l);
      f += static_cast<double>(d->ncore) * std::ldexp(1.0, 3 * (d->max_level - l));
    }
    f *= std::ldexp(1.0, d->max_level);
    if (actual) *actual = a;
    if (feq) *feq = f;
  });
}

int octlb_dev_kernel_timing(octlb_dev* d, double* ms, int64_t* launches, double* updates) {
  return Guard([&] {
    CheckDev(d);
    if (ms) *ms = d->tile_ms;
    if (launches) *launches = d->tile_launches;
    if (updates) *updates = d->tile_updates;
  });
}

int octlb_dev_time_kernel(octlb_dev* d, int32_t level, int32_t reps, double* avg_ms,
                          double* cell_updates) {
  return Guard([&] {
    CheckDev(d);
    if (reps <= 0) throw ArgError("reps must be positive");
    if (level < 0 || level > d->max_level || d->tile_count[level] == 0) throw ArgError("level has no blocks");
    EnsureReady(d);
    d->Sync();
    const double ms = d->prec == 64 ? d->TimeKernel<double>(level, reps) : d->TimeKernel<float>(level, reps);
    if (avg_ms) *avg_ms = ms;
    if (cell_updates) {
      const int tpa = (d->NP + kTile - 1) / kTile;
      const int tpb = d->padded ? tpa * tpa * tpa : 1;
      (void)tpb;
      *cell_updates = static_cast<double>(d->level_blocks[level]) * static_cast<double>(d->ncore);
    }
  });
}

int octlb_dev_set_kernel_variant(octlb_dev* d, int32_t variant) {
  return Guard([&] {
    CheckDev(d);
    if (variant != 0 && variant != 1) throw ArgError("kernel variant must be 0 (pull) or 1 (push where plain)");
    if (variant != d->variant) {
      d->Sync();
      d->variant = variant;
      d->prepared = false;  // the block classes and the tile list are laid out by Prepare
    }
  });
}

int octlb_dev_kernel_variant(octlb_dev* d, int32_t* variant, int32_t* push_blocks) {
  return Guard([&] {
    CheckDev(d);
    EnsureReady(d);
    if (variant) *variant = d->variant;
    if (push_blocks) *push_blocks = d->npush_blocks;
  });
}

int octlb_dev_autotune(octlb_dev* d, int32_t reps, double* step_ms) {
  return Guard([&] {
    CheckDev(d);
    if (reps <= 0) throw ArgError("reps must be positive");
    // Per candidate: the fused kernel of every level replayed `reps` times
    // on the current state (inputs only read, outputs to the idle buffers),
    // weighted by its 2^l launches per coarse step.  The state is unchanged.
    double best = 0.0;
    int best_v = d->variant;
    for (int v = 0; v < 2; ++v) {
      d->variant = v;
      d->prepared = false;
      EnsureReady(d);
      if (v == 1 && d->npush_blocks == 0) continue;
      double ms = 0.0;
      for (int l = 0; l <= d->max_level; ++l) {
        if (d->tile_count[l] == 0) continue;
        const double t = d->prec == 64 ? d->TimeKernel<double>(l, reps) : d->TimeKernel<float>(l, reps);
        ms += t * static_cast<double>(1 << l);
      }
      if (step_ms) step_ms[v] = ms;
      if (v == 0 || ms < best) {
        best = ms;
        best_v = v;
      }
    }
    if (d->variant != best_v) {
      d->variant = best_v;
      d->prepared = false;
      EnsureReady(d);
    }
    // Coupling epilogues, per kind: on when the fused launches carrying them
    // (per coarse step) take less than the plain launches plus the separate
    // coupling kernels they replace.
    d->ep_ms[0] = d->ep_ms[1] = 0.0;
    if (d->ep_ok) {
      double pro_off = 0.0, pro_on = 0.0, res_off = 0.0, res_on = 0.0;
      for (int l = 0; l <= d->max_level; ++l) {
        if (d->tile_count[l] == 0) continue;
        auto time = [&](int what) {
          return d->prec == 64 ? d->TimeLaunches<double>(l, reps, what) : d->TimeLaunches<float>(l, reps, what);
        };
        const double t = d->prec == 64 ? d->TimeKernel<double>(l, reps) : d->TimeKernel<float>(l, reps);
        const double w = static_cast<double>(1 << l);
        if (l < d->max_level && d->d_ep_pro_beg[l] != nullptr) {
          pro_off += w * time(2);
          pro_on += w * (time(0) - t);
        }
        if (l >= 1 && l == d->max_level && d->d_ep_res_beg[l] != nullptr) {
          const double wr = static_cast<double>(1 << (l - 1));  // once per visit of the parent
          res_off += wr * time(3);
          res_on += wr * (time(1) - t);
        }
      }
      d->ep_mode = (pro_on < pro_off ? 1 : 0) | (res_on < res_off ? 2 : 0);
      d->ep_ms[0] = pro_off + res_off;
      d->ep_ms[1] = std::min(pro_on, pro_off) + std::min(res_on, res_off);
    }
  });
}

int octlb_dev_coupling_epilogues(octlb_dev* d, int32_t* mode, double ms[2]) {
  return Guard([&] {
    CheckDev(d);
    if (!d->prepared) d->Prepare();  // whether the epilogues apply is known once the tables are built
    if (mode) *mode = d->ep_ok ? d->ep_mode : 0;
    if (ms) {
      ms[0] = d->ep_ms[0];
      ms[1] = d->ep_ms[1];
    }
  });
}

int octlb_dev_set_coupling_epilogues(octlb_dev* d, int32_t mode) {
  return Guard([&] {
    CheckDev(d);
    if (mode < 0 || mode > 3) throw ArgError("epilogue mode must be 0..3");
    d->ep_mode = mode;
  });
}

int octlb_dev_enable_timing(octlb_dev* d, int32_t on) {
  return Guard([&] {
    CheckDev(d);
    d->timing = on != 0;
  });
}

}  // extern "C"
