// Discrete adjoint of the single-level time step, on the device.
//
// The reference has no adjoint.  What is differentiated is exactly the
// forward this library runs (TimeLoop::advance_one for one level with the
// concrete boundary handler, plain BGK bulk, bounce-back cells and Bouzidi
// links, time_loop.h:173-246): the objective is the momentum-exchange drag
// (drag_postprocessor.h:66-115) projected on a direction and averaged over a
// window of steps; the parameters are the initial flow velocity u0 of
// feq(rho0, u0) and every Bouzidi wall fraction q (bouzidi_pull.h:8-26).
//
// One step is y = C(x) (collide, per cell) followed by x' = M y, where M
// composes the halo exchange (faces and same-level edges, GhostSchedule) with
// the pull stream, its halo-shell wrap and the Bouzidi and bounce-back rules
// (block_lattice.cpp:503-720).  M does not depend on the state, so the host
// builds it once over the reference BlockLattice view (10^3 halo-inclusive
// slots per block) and keeps only the live part: the core cells and the
// halo slots that unexchanged shells carry into them.  State layout: per
// block, its live slots in ascending slot order as population planes,
//   state[block_offset + pop * live_slots(block) + rank(slot)]
// so that threads over a block's cells read each plane coalesced.
//   forward  KAdjCollide (first step of a sweep), then KAdjStep per step:
//            gather x' = M y through one source index per entry (Bouzidi
//            rows through a link record), store x', collide it into the
//            next y (double-buffered)
//   reverse  KAdjReverse per step: lambda_y = M^T lambda_x through one
//            reader index per entry (the few entries read more than once
//            or by a Bouzidi row go through an overflow list), then the BGK
//            vector-Jacobian product at the stored pre-collide state;
//            the drag seeds of the step ride along; KAdjLinkGrad adds
//            d out / d q on the Bouzidi rows (tiny)
// The forward keeps a checkpoint every K steps; the reverse sweep recomputes
// each segment of K states from its checkpoint, so memory is (steps/K + K)
// states.  The forward states take the handle's precision (FP32: the FP32
// path's collide arithmetic, Bouzidi rows in FP64 rounded to FP32, as the
// forward does); adjoint variables, vector-Jacobian products, objective
// values, dJ/du0 and dJ/dq are FP64, accumulated in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "d3q19.h"
#include "octlb_gpu.h"
#include "octlb_internal.h"
#include "adjoint_ml.h"
#include "adjoint_topo.h"

namespace octlb {
namespace {

#define ADJ_CHECK(expr)                                                                        \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e) +            \
                               " (adjoint.hip:" + std::to_string(__LINE__) + ")");            \
  } while (0)

// A launch's read-only inputs seen through the constant address space: their
// loads cannot alias its stores, so the compiler batches them.
#if defined(__HIP_DEVICE_COMPILE__)
#define ADJ_RO __attribute__((address_space(4)))
#else
#define ADJ_RO
#endif
template <typename P>
__device__ __forceinline__ const ADJ_RO P* Ro(const P* p) {
  return (const ADJ_RO P*)p;
}

// Consecutive workgroups (consecutive cells of a block and its neighbours
// in the cell order) go to the same XCD, whose L2 then serves the gathers.
__device__ __forceinline__ int XcdWg(int w, int n) {
  const int q8 = n >> 3;
  return (w < (q8 << 3)) ? (w & 7) * q8 + (w >> 3) : w;
}

constexpr int kN = 8, kNP = 10, kNS = kNP * kNP * kNP, kNC = kN * kN * kN;
constexpr int kThreads = 128;
constexpr uint8_t kHalo = 255;  // kind of a halo slot: carried, never collided

struct AdjLink {
  int row;      // output entry
  int bb;       // f_bb_opp source entry
  int other;    // f_interior (q >= 1/2) or f_same (q < 1/2) source entry
  int upper;    // q >= 1/2 branch
  int bb_cell, bb_pop, ot_cell, ot_pop;  // the same sources as (cell, population)
  double q;
};

struct DragTerm {
  int snap;  // post-collide entry (cell, pop i)
  int live;  // post-stream entry (cell, pop opp(i))
  double w;  // -(dir . c_i)
};

struct Cell {
  int base;    // entry of population 0
  int stride;  // live slots of the block (population plane length)
};

template <typename Ptr, typename V>
__device__ __forceinline__ void LoadCell(Ptr s, Cell c, V* f) {
#pragma unroll
  for (int i = 0; i < kQ; ++i) f[i] = s[c.base + i * c.stride];
}

// dispatch_collide for the kinds the adjoint admits (bc_dispatcher.h:32-81),
// in the state's precision like the forward path
template <typename S>
__device__ __forceinline__ void Collide(S* f, uint8_t k, double omega) {
  if (k == kBulk) {
    S rho, u[3];
    ProxyRhoU(f, rho, u);
    OlbBgk(f, rho, u, static_cast<S>(omega));
  } else if (k == kBounceBack) {
    S t[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) t[i] = f[i];
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = t[Opp(i)];
  }
}

// l <- transpose of (ProxyRhoU + OlbBgk) at the pre-collide cell f:
// l_j' = (1 - w) l_j + w (A + sum_d G_d (c_jd - u_d) / rho),
// A = sum_i l_i dfeq_i/drho, G_d = sum_i l_i dfeq_i/du_d.
__device__ __forceinline__ void BgkVjp(const double* f, double* l, double omega) {
  double rho, u[3];
  ProxyRhoU(f, rho, u);
  const double usqr = NormSqr(u);
  double A = 0.0, G0 = 0.0, G1 = 0.0, G2 = 0.0;
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const double cu = Cx(i) * u[0] + Cy(i) * u[1] + Cz(i) * u[2];
    const double w = W<double>(i);
    A += l[i] * w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usqr);
    const double g = l[i] * rho * w;
    G0 += g * (3.0 * Cx(i) + 9.0 * cu * Cx(i) - 3.0 * u[0]);
    G1 += g * (3.0 * Cy(i) + 9.0 * cu * Cy(i) - 3.0 * u[1]);
    G2 += g * (3.0 * Cz(i) + 9.0 * cu * Cz(i) - 3.0 * u[2]);
  }
  const double inv_rho = 1.0 / rho;  // one division per cell
  const double h0 = G0 * inv_rho, h1 = G1 * inv_rho, h2 = G2 * inv_rho;
  const double base_g = A - (h0 * u[0] + h1 * u[1] + h2 * u[2]);
#pragma unroll
  for (int j = 0; j < kQ; ++j) {
    const double g = base_g + (h0 * Cx(j) + h1 * Cy(j) + h2 * Cz(j));
    l[j] = (1.0 - omega) * l[j] + omega * g;
  }
}

template <typename S>
__global__ __launch_bounds__(kThreads) void KAdjInit(S* x, const Cell* cells, const uint8_t* ckind, int ncells,
                                                     double rho0, double u0x, double u0y, double u0z, uint32_t mask) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= ncells) return;
  const Cell c = cells[t];
  const uint8_t k = ckind[t];
  const bool on = k != kHalo && ((mask >> k) & 1u);
  const double u[3] = {u0x, u0y, u0z};
  const double usqr = NormSqr(u);
#pragma unroll
  for (int i = 0; i < kQ; ++i) x[c.base + i * c.stride] = on ? static_cast<S>(Feq<double>(i, rho0, u, usqr)) : S(0);
}

template <typename S>
__global__ __launch_bounds__(kThreads) void KAdjCollide(const S* x, S* y, const Cell* cells,
                                                        const uint8_t* ckind, int ncells, double omega) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= ncells) return;
  const Cell c = cells[t];
  S f[kQ];
  LoadCell(x, c, f);
  Collide(f, ckind[t], omega);
#pragma unroll
  for (int i = 0; i < kQ; ++i) y[c.base + i * c.stride] = f[i];
}

// Regular cells (every population pulled plainly from its geometric
// neighbour in this block or an exchanged neighbour block): the entry of
// population i at the cell offset by (dx, dy, dz) is computed from the
// block's 27 neighbour (base, plane length) pairs -- core cells come first in
// every block, in (x, y, z) order -- instead of loaded from the index tables.
// The host sets the flag only where this reproduces the tables entry by
// entry.
__device__ __forceinline__ int RegularEntry(const ADJ_RO int2* nbase, int ix, int iy, int iz, int i) {
  const int ox = ix < 0 ? 0 : (ix > kN - 1 ? 2 : 1);
  const int oy = iy < 0 ? 0 : (iy > kN - 1 ? 2 : 1);
  const int oz = iz < 0 ? 0 : (iz > kN - 1 ? 2 : 1);
  const int2 nb = nbase[(ox * 3 + oy) * 3 + oz];
  return nb.x + i * nb.y + (((ix & 7) * kN + (iy & 7)) * kN + (iz & 7));
}

// entry e of x' = M y, rounded to the state's precision (Bouzidi rows are
// evaluated in FP64 for every storage type, as the forward does)
template <typename S, typename PY, typename PS, typename PL>
__device__ __forceinline__ S RowValue(PY y, PS src, PL links, int e) {
  const int s = src[e];
  if (s >= 0) return y[s];
  const AdjLink L = links[-1 - s];
  const double a = y[L.bb], o = y[L.other];
  return static_cast<S>(L.upper ? BouzidiPull(a, o, o, L.q) : BouzidiPull(a, a, o, L.q));
}

// x' = M y for this cell's entries, stored when x is given, then collided
// into the next y
template <typename S>
__global__ __launch_bounds__(kThreads) void KAdjStep(const S* y, S* x, S* yn, const int* src,
                                                     const AdjLink* links, const Cell* cells, const uint8_t* ckind,
                                                     const int* cblock, const uint8_t* cflag, const int2* nbase,
                                                     int ncells, double omega) {
  const int t = XcdWg(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (t >= ncells) return;
  const Cell c = Ro(cells)[t];
  S f[kQ];
  if (Ro(cflag)[t] & 1) {
    const ADJ_RO int2* nb = Ro(nbase) + 27 * Ro(cblock)[t];
    const int r = c.base - nb[13].x, ix = r >> 6, iy = (r >> 3) & 7, iz = r & 7;
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = Ro(y)[RegularEntry(nb, ix - Cx(i), iy - Cy(i), iz - Cz(i), i)];
    if (x) {
#pragma unroll
      for (int i = 0; i < kQ; ++i) x[c.base + i * c.stride] = f[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const int e = c.base + i * c.stride;
      f[i] = RowValue<S>(Ro(y), Ro(src), Ro(links), e);
      if (x) x[e] = f[i];
    }
  }
  Collide(f, Ro(ckind)[t], omega);
#pragma unroll
  for (int i = 0; i < kQ; ++i) yn[c.base + i * c.stride] = f[i];
}

// sum of w (snap + live) over the drag terms: one partial per workgroup of
// consecutive terms (tree order fixed), summed on the host in order; the
// live value is the stream row of the post-collide state
template <typename S>
__global__ __launch_bounds__(kThreads) void KAdjDrag(const S* y, const int* src, const AdjLink* links,
                                                     const DragTerm* terms, int nterms, double* out) {
  __shared__ double red[kThreads];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  double acc = 0.0;
  if (t < nterms) {
    const DragTerm d = terms[t];
    acc = d.w * (static_cast<double>(y[d.snap]) + static_cast<double>(RowValue<S>(y, src, links, d.live)));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// live-state drag seeds: every term owns its entry (distinct pops)
__global__ __launch_bounds__(kThreads) void KAdjSeedLive(double* lx, const DragTerm* terms, int nterms, double scale) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nterms) return;
  const DragTerm d = terms[t];
  lx[d.live] += d.w * scale;
}

__device__ __forceinline__ double ReaderCoef(AdjLink L, int which) {
  if (L.upper) {
    const double a = 1.0 / (2.0 * L.q);
    return which ? 1.0 - a : a;
  }
  return which ? 1.0 - 2.0 * L.q : 2.0 * L.q;
}

// lxn = C^T (M^T lx) at the pre-collide state xp.  rd[e] >= 0: the one
// output that copies e; -1: none; -2 - k: overflow list k (several readers
// or a Bouzidi row) in ovo / ove (row, -1 for a copy or 2 link + which).
template <typename S>
__global__ __launch_bounds__(kThreads) void KAdjReverse(const double* lx, double* lxn, const S* xp,
                                                        const int* rd, const int* ovo, const int2* ove,
                                                        const AdjLink* links, const Cell* cells,
                                                        const uint8_t* ckind, const int* cblock, const uint8_t* cflag,
                                                        const int2* nbase, const int* cseed, const double* seed_w,
                                                        double snap_scale, double live_scale, int ncells,
                                                        double omega) {
  const int t = XcdWg(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (t >= ncells) return;
  const Cell c = Ro(cells)[t];
  double l[kQ];
  if (Ro(cflag)[t] & 2) {
    const ADJ_RO int2* nb = Ro(nbase) + 27 * Ro(cblock)[t];
    const int r = c.base - nb[13].x, ix = r >> 6, iy = (r >> 3) & 7, iz = r & 7;
#pragma unroll
    for (int i = 0; i < kQ; ++i) l[i] = Ro(lx)[RegularEntry(nb, ix + Cx(i), iy + Cy(i), iz + Cz(i), i)];
  } else {
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const int r = Ro(rd)[c.base + i * c.stride];
      double v = 0.0;
      if (r >= 0) {
        v = Ro(lx)[r];
      } else if (r < -1) {
        const int k = -2 - r;
        for (int j = Ro(ovo)[k]; j < Ro(ovo)[k + 1]; ++j) {
          const int2 o = Ro(ove)[j];
          v += (o.y < 0 ? 1.0 : ReaderCoef(Ro(links)[o.y >> 1], o.y & 1)) * Ro(lx)[o.x];
        }
      }
      l[i] = v;
    }
  }
  // drag cells: the step's post-collide seed enters before the collide
  // transpose, the previous step's live seed after it
  const int sk = Ro(cseed)[t];
  if (sk >= 0 && snap_scale != 0.0) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) l[i] += snap_scale * Ro(seed_w)[static_cast<size_t>(sk) * 2 * kQ + i];
  }
  const uint8_t k = Ro(ckind)[t];
  if (k == kBulk) {
    double f[kQ];
    LoadCell(Ro(xp), c, f);
    BgkVjp(f, l, omega);
  } else if (k == kBounceBack) {
    double tmp[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) tmp[i] = l[i];
#pragma unroll
    for (int i = 0; i < kQ; ++i) l[i] = tmp[Opp(i)];
  }
  if (sk >= 0 && live_scale != 0.0) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) l[i] += live_scale * Ro(seed_w)[static_cast<size_t>(sk) * 2 * kQ + kQ + i];
  }
#pragma unroll
  for (int i = 0; i < kQ; ++i) lxn[c.base + i * c.stride] = l[i];
}

template <typename S>
__device__ double PostCollide(const S* x, const Cell* cells, const uint8_t* ckind, int cell, int pop, double omega) {
  S f[kQ];
  LoadCell(x, cells[cell], f);
  Collide(f, ckind[cell], omega);
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < kQ; ++i) v = (i == pop) ? static_cast<double>(f[i]) : v;
  return v;
}

// dq += lx[row] * d BouzidiPull / d q, the pulled values recollided from the
// step's pre-collide state
template <typename S>
__global__ __launch_bounds__(kThreads) void KAdjLinkGrad(const double* lx, const S* xp, const Cell* cells,
                                                         const uint8_t* ckind, const AdjLink* links, int nlinks,
                                                         double omega, double* dq) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= nlinks) return;
  const AdjLink L = links[k];
  const double fbb = PostCollide(xp, cells, ckind, L.bb_cell, L.bb_pop, omega);
  const double fo = PostCollide(xp, cells, ckind, L.ot_cell, L.ot_pop, omega);
  const double d = L.upper ? -(fbb - fo) / (2.0 * L.q * L.q) : 2.0 * fbb - 2.0 * fo;
  dq[k] += lx[L.row] * d;
}

// dJ/du0 partial sums per workgroup: x0 = feq(rho0, u0) on the masked cells
__global__ __launch_bounds__(kThreads) void KAdjInitGrad(const double* lx, const Cell* cells, const uint8_t* ckind,
                                                         int ncells, double rho0, double u0x, double u0y, double u0z,
                                                         uint32_t mask, double* partial) {
  __shared__ double red[3][kThreads];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  double g[3] = {0.0, 0.0, 0.0};
  if (t < ncells) {
    const uint8_t k = ckind[t];
    if (k != kHalo && ((mask >> k) & 1u)) {
      const Cell c = cells[t];
      const double u[3] = {u0x, u0y, u0z};
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        const double l = lx[c.base + i * c.stride];
        const double cu = Cx(i) * u[0] + Cy(i) * u[1] + Cz(i) * u[2];
        const double s = l * rho0 * W<double>(i);
        for (int d = 0; d < 3; ++d) g[d] += s * (3.0 * Ci(i, d) + 9.0 * cu * Ci(i, d) - 3.0 * u[d]);
      }
    }
  }
  for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = g[d];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) partial[3 * blockIdx.x + d] = red[d][0];
}

inline unsigned Blocks(long long n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

// ------------------------------------------------------------------------
// Multi-level step (several refinement levels, FD inlet/outlet cells, the
// Lagrava coupler): the step is the recursion of time_loop.h:201-246 as a
// list of operations on the live state X (all levels) and the coupler's
// previous macros P.  A level visit collides the level's cells with the
// production collide (CollideCell: BGK bulk, bounce-back, the velocity-
// Dirichlet collide with the prescribed velocity at time t, no-op kinds)
// into Y and streams them back, X = M_l Y; the FD pass, prolongation and
// restriction run the production device code on the same layout
// (adjoint_ml.inc).  Core cell c of block b is entry boff[b] + i cnt[b] + c.

__device__ __forceinline__ Origin CellOrigin(const double* ocw, int b) {
  const double* o = ocw + 4 * static_cast<size_t>(b);
  return Origin{o[0], o[1], o[2], o[3]};
}

// y = C(x) over cells [c0, c1) (halo slots are carried: y = x); with a
// tape, x is also written there (the cells' entries are the level's window)
template <typename S>
__global__ __launch_bounds__(kThreads) void KMLCollide(const S* x, S* y, const Cell* cells, const uint8_t* ckind,
                                                       const int* cblock, const int* boff, const double* ocw,
                                                       const SimConst* scp, double t, int c0, int c1, S* tape) {
  const SimConst& sc = *scp;  // by pointer: a by-value copy would be staged in scratch per lane
  const int cl = c0 + static_cast<int>(blockIdx.x) * kThreads + static_cast<int>(threadIdx.x);
  if (cl >= c1) return;
  const Cell c = cells[cl];
  S f[kQ];
  LoadCell(x, c, f);
  if (tape) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) tape[c.base + i * c.stride] = f[i];
  }
  const uint8_t k = ckind[cl];
  if (k != kHalo) {
    const int b = cblock[cl];
    const int rank = c.base - boff[b];
    double ssum = 0.0;
    int scnt = 0;
    CollideCell<S>(f, k, sc, CellOrigin(ocw, b), rank >> 6, (rank >> 3) & 7, rank & 7, t, AvgArg<S>(1.0), false,
                   ssum, scnt);
  }
#pragma unroll
  for (int i = 0; i < kQ; ++i) y[c.base + i * c.stride] = f[i];
}

// x = M y over the entries of cells [c0, c1)
template <typename S>
__global__ __launch_bounds__(kThreads) void KMLGather(const S* y, S* x, const int* src, const AdjLink* links,
                                                      const Cell* cells, int c0, int c1) {
  const int cl = c0 + static_cast<int>(blockIdx.x) * kThreads + static_cast<int>(threadIdx.x);
  if (cl >= c1) return;
  const Cell c = cells[cl];
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const int e = c.base + i * c.stride;
    x[e] = RowValue<S>(y, src, links, e);
  }
}

// Velocity-Dirichlet collide VJP (CollideDirichletBoundaryCellAt,
// interpolated_velocity.h:858-884): f' = (1 - w) f + w feq(rho, u) with u
// prescribed and rho = VelocityBoundaryRhoFromPop (flat face cells) or
// CellProxy rho (edges, corners).  l is updated in place; du receives
// d/du of the prescribed velocity.
__device__ inline void VelocityCollideVjp(const double* f, double* l, double omega, const double* u, int n, int ix,
                                          int iy, int iz, double* du) {
  const int fc = (ix == 0) + (ix == n - 1) + (iy == 0) + (iy == n - 1) + (iz == 0) + (iz == n - 1);
  int dir = 0, ori = 0;
  double rho, drho_df[kQ], drho_du = 0.0;
  if (fc >= 2) {
    double r, ut[3];
    ProxyRhoU(f, r, ut);
    rho = r;
    for (int j = 0; j < kQ; ++j) drho_df[j] = 1.0;
  } else {
    FlatInfo(n, ix, iy, iz, dir, ori);
    rho = VelRhoFromPop(dir, ori, f, u);
    const double den = 1.0 + ori * u[dir];
    for (int j = 0; j < kQ; ++j) {
      const int cj = Ci(j, dir);
      drho_df[j] = (cj == 0 ? 1.0 : (cj == ori ? 2.0 : 0.0)) / den;
    }
    drho_du = -ori * rho / den;
  }
  const double usqr = NormSqr(u);
  double A = 0.0, G[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const double cu = Cx(i) * u[0] + Cy(i) * u[1] + Cz(i) * u[2];
    const double w = W<double>(i);
    A += l[i] * w * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usqr);
    const double g = l[i] * rho * w;
    G[0] += g * (3.0 * Cx(i) + 9.0 * cu * Cx(i) - 3.0 * u[0]);
    G[1] += g * (3.0 * Cy(i) + 9.0 * cu * Cy(i) - 3.0 * u[1]);
    G[2] += g * (3.0 * Cz(i) + 9.0 * cu * Cz(i) - 3.0 * u[2]);
  }
#pragma unroll
  for (int j = 0; j < kQ; ++j) l[j] = (1.0 - omega) * l[j] + omega * A * drho_df[j];
  for (int d = 0; d < 3; ++d) du[d] = omega * G[d];
  if (fc < 2) du[dir] += omega * A * drho_du;
}

// One level visit reversed over cells [c0, c1): lxn = C^T (M^T lx + snap
// seeds) at the visit's pre-collide state xp; d/d(inlet velocity) of the
// velocity cells as one partial per workgroup (fixed tree) in dui.
template <typename S>
__global__ __launch_bounds__(kThreads) void KMLVisitVjp(const double* lx, double* lxn, const S* xp, const int* rd,
                                                        const int* ovo, const int2* ove, const AdjLink* links,
                                                        const Cell* cells, const uint8_t* ckind,
                                                        const SimConst* scp, const int* cseed, const double* seed_w,
                                                        double snap_scale, int c0, int c1) {
  const int cl = c0 + static_cast<int>(blockIdx.x) * kThreads + static_cast<int>(threadIdx.x);
  if (cl >= c1) return;
  const Cell c = cells[cl];
  double l[kQ];
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const int r = rd[c.base + i * c.stride];
    double v = 0.0;
    if (r >= 0) {
      v = lx[r];
    } else if (r < -1) {
      const int k = -2 - r;
      for (int j = ovo[k]; j < ovo[k + 1]; ++j) {
        const int2 o = ove[j];
        v += (o.y < 0 ? 1.0 : ReaderCoef(links[o.y >> 1], o.y & 1)) * lx[o.x];
      }
    }
    l[i] = v;
  }
  const int sk = cseed[cl];
  if (sk >= 0 && snap_scale != 0.0) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) l[i] += snap_scale * seed_w[static_cast<size_t>(sk) * 2 * kQ + i];
  }
  const uint8_t k = ckind[cl];
  if (k == kBulk) {
    double f[kQ];
    LoadCell(xp, c, f);
    BgkVjp(f, l, scp->omega);
  } else if (k == kBounceBack) {
    double tmp[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) tmp[i] = l[i];
#pragma unroll
    for (int i = 0; i < kQ; ++i) l[i] = tmp[Opp(i)];
  }
  // velocity cells: KMLVelocityVjp takes their collide next
#pragma unroll
  for (int i = 0; i < kQ; ++i) lxn[c.base + i * c.stride] = l[i];
}

// The velocity-Dirichlet collide VJP of one level's velocity cells (list
// vc[0, nv)), in place on the adjoint KMLVisitVjp left for them; d/d(inlet
// velocity) as one partial per workgroup (fixed tree) in dui.
template <typename S>
__global__ __launch_bounds__(kThreads) void KMLVelocityVjp(double* lxn, const S* xp, const int* vc, int nv,
                                                           const Cell* cells, const int* cblock, const int* boff,
                                                           const double* ocw, const SimConst* scp, double t,
                                                           double* dui) {
  __shared__ double red[3][kThreads];
  const SimConst& sc = *scp;
  const int j = static_cast<int>(blockIdx.x) * kThreads + static_cast<int>(threadIdx.x);
  double g[3] = {0.0, 0.0, 0.0};
  if (j < nv) {
    const int cl = vc[j];
    const Cell c = cells[cl];
    double l[kQ], f[kQ];
    for (int i = 0; i < kQ; ++i) l[i] = lxn[c.base + i * c.stride];
    LoadCell(xp, c, f);
    const int b = cblock[cl];
    const int rank = c.base - boff[b];
    const int ix = rank >> 6, iy = (rank >> 3) & 7, iz = rank & 7;
    const Origin og = CellOrigin(ocw, b);
    const double org[3] = {og.x, og.y, og.z};
    double ud[3];
    PrescribedU(sc, org, og.cw, ix, iy, iz, t, ud);
    // the forward rounds the prescribed velocity to the state's precision
    const double u[3] = {static_cast<double>(static_cast<S>(ud[0])), static_cast<double>(static_cast<S>(ud[1])),
                         static_cast<double>(static_cast<S>(ud[2]))};
    double du[3];
    VelocityCollideVjp(f, l, sc.omega, u, kN, ix, iy, iz, du);
    const double fct = InletFactor(sc, org, og.cw, ix, iy, iz, t);
    for (int d = 0; d < 3; ++d) g[d] = du[d] * fct;
    for (int i = 0; i < kQ; ++i) lxn[c.base + i * c.stride] = l[i];
  }
  for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = g[d];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w)
      for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) dui[3 * blockIdx.x + d] = red[d][0];
}

// post-collide value of one population of a live cell (production collide)
template <typename S>
__device__ double MLPostCollide(const S* x, const Cell* cells, const uint8_t* ckind, const int* cblock,
                                const int* boff, const double* ocw, const SimConst& sc, double t, int cell, int pop) {
  const Cell c = cells[cell];
  S f[kQ];
  LoadCell(x, c, f);
  const uint8_t k = ckind[cell];
  if (k != kHalo) {
    const int b = cblock[cell];
    const int rank = c.base - boff[b];
    double ssum = 0.0;
    int scnt = 0;
    CollideCell<S>(f, k, sc, CellOrigin(ocw, b), rank >> 6, (rank >> 3) & 7, rank & 7, t, AvgArg<S>(1.0), false,
                   ssum, scnt);
  }
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < kQ; ++i) v = (i == pop) ? static_cast<double>(f[i]) : v;
  return v;
}

// dq += lx[row] d BouzidiPull / dq for links [k0, k1), sources recollided
// from the visit's pre-collide state
template <typename S>
__global__ __launch_bounds__(kThreads) void KMLLinkGrad(const double* lx, const S* xp, const Cell* cells,
                                                        const uint8_t* ckind, const int* cblock, const int* boff,
                                                        const double* ocw, const SimConst* scp, double t,
                                                        const AdjLink* links, int k0, int k1, double* dq) {
  const int k = k0 + static_cast<int>(blockIdx.x) * kThreads + static_cast<int>(threadIdx.x);
  if (k >= k1) return;
  const SimConst& sc = *scp;
  const AdjLink L = links[k];
  const double fbb = MLPostCollide(xp, cells, ckind, cblock, boff, ocw, sc, t, L.bb_cell, L.bb_pop);
  const double fo = MLPostCollide(xp, cells, ckind, cblock, boff, ocw, sc, t, L.ot_cell, L.ot_pop);
  const double d = L.upper ? -(fbb - fo) / (2.0 * L.q * L.q) : 2.0 * fbb - 2.0 * fo;
  dq[k] += lx[L.row] * d;
}

// sum of w (snap + live) over the drag terms after a coarse step: the
// post-collide snapshot of the drag cells' last visit (y) and the final
// post-stream state (x); one partial per workgroup
template <typename S>
__global__ __launch_bounds__(kThreads) void KMLDrag(const S* y, const S* x, const DragTerm* terms, int nterms,
                                                    double* out) {
  __shared__ double red[kThreads];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  double acc = 0.0;
  if (t < nterms) {
    const DragTerm d = terms[t];
    acc = d.w * (static_cast<double>(y[d.snap]) + static_cast<double>(x[d.live]));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// acc[0..2] += sum of part[3 j + d] over j, in a fixed order: thread t sums
// j = t, t + kAccumThreads, ... in turn, then a tree over the threads.
constexpr int kAccumThreads = 256;
__global__ __launch_bounds__(kAccumThreads) void KMLAccum(const double* part, int n, double* acc) {
  __shared__ double red[kAccumThreads][3];
  const int t = threadIdx.x;
  double s[3] = {0.0, 0.0, 0.0};
  for (int j = t; j < n; j += kAccumThreads)
    for (int d = 0; d < 3; ++d) s[d] += part[3 * j + d];
  for (int d = 0; d < 3; ++d) red[t][d] = s[d];
  __syncthreads();
  for (int w = kAccumThreads / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int d = 0; d < 3; ++d) red[t][d] += red[t + w][d];
    __syncthreads();
  }
  if (t == 0)
    for (int d = 0; d < 3; ++d) acc[d] += red[0][d];
}

}  // namespace
}  // namespace octlb

using namespace octlb;

struct octlb_adj {
  int nb = 0, ncells = 0, ntot = 0, nlinks = 0;
  double omega = 0.0;
  std::vector<Cell> cells;
  std::vector<uint8_t> ckind;
  std::vector<int> cell_block, cell_slot;
  std::vector<int> core_cell;  // [nb][512] -> cell
  std::vector<int> cblock;     // block of each cell
  std::vector<uint8_t> cflag;  // bit 0: forward rows regular, bit 1: readers regular
  std::vector<int2> nbase;     // [nb][27] (entry base, plane length) of the block at (dx, dy, dz), or -1
  std::vector<double> origin_cw;
  std::unordered_map<int64_t, int> link_index;  // (block*512 + cell)*19 + pop -> Bouzidi row
  std::vector<int> border;                      // blocks in level order
  std::vector<int> lv_cell, lv_entry, lv_link;  // per level: [begin, end) of cells, entries, links
  // multi-level: per operation, the window [win_lo, win_hi) of state entries
  // its adjoint reads, and the windows' offsets in a packed per-step tape
  std::vector<long long> win_lo, win_hi, win_off;
  std::vector<int> lv_vcell;  // per level: [begin, end) in d_vcell of its velocity-Dirichlet cells
  int* d_vcell = nullptr;
  double* d_pcm = nullptr;
  unsigned fd_level_mask = 0;      // levels holding FD blocks
  double* d_dui_fd = nullptr;      // d/d(inlet velocity) from the FD VJPs (their own accumulator)
  double* d_dui_fd_part = nullptr;
  std::vector<int> block_off, block_cnt;        // per block: first entry, plane length
  hipStream_t stream = nullptr;
  // multi-level: the FD VJP runs on `side` beside a level visit's VJP when
  // they touch different levels (fork / join events)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  Cell* d_cells = nullptr;
  int* d_cblock = nullptr;
  uint8_t* d_cflag = nullptr;
  int2* d_nbase = nullptr;
  uint8_t* d_ckind = nullptr;
  int* d_src = nullptr;
  int* d_rd = nullptr;
  int* d_ovo = nullptr;
  int2* d_ove = nullptr;
  AdjLink* d_links = nullptr;
  double* d_dq = nullptr;
  void* d_y[2] = {nullptr, nullptr};  // post-collide states (state precision)
  double* d_lx = nullptr;
  double* d_ln = nullptr;
  void* d_final = nullptr;  // post-stream state after the last run's forward
  int prec = 64;            // state precision: 64, or 32 (FP32 forward, FP64 adjoint)
  size_t selem = 8;
  std::vector<void*> owned;
  double fwd_ms = 0.0, bwd_ms = 0.0;
  bool fd_side = false;      // MLReverse: launch the FD VJP on the side stream
  bool fd_inflight = false;  // an FD VJP on the side stream is not joined yet
  bool packed_mode = false;  // multi-level: the last run kept a segment's tapes for all its steps
  long long tape_budget = -1;  // bytes allowed for a segment's packed tapes (-1: 60 % of the free memory)

  // ---- multi-level step (levels, FD cells, coupling); see KMLCollide
  enum MLOp { kOpVisit = 0, kOpFd = 1, kOpHalf = 2, kOpFull = 3, kOpRestrict = 4 };
  bool ml = false;
  int nlev = 1, nfd = 0, npts = 0;
  AdjMLTopo mtopo;
  std::vector<int2> ops;  // (op, level): level visited, coarse level coupled, fine level restricted
  int* d_boff = nullptr;
  int* d_cnt = nullptr;
  double* d_ocw = nullptr;
  uint8_t* d_kind_blk = nullptr;  // [nb][512] BcKind bytes (production layout)
  BlockTab* d_bt = nullptr;
  SimConst* d_scm = nullptr;
  int* d_fd_blocks = nullptr;
  short* d_fd_slot = nullptr;
  short* d_fd_list = nullptr;
  int4* d_fd_count = nullptr;
  int* d_fd_rbeg = nullptr;           // reverse FD stencils: CSR offsets per FD block
  FdRevEntry* d_fd_rent = nullptr;    // and entries (kFdRevMax per FD block)
  int* d_plan = nullptr;
  std::vector<int*> d_plist, d_ppts, d_pbeg, d_rlist, d_rebeg, d_recell;
  std::vector<int2*> d_rent;
  std::vector<int> np_l, npg_l, nr_l, nrc_l;
  int* d_ipts = nullptr;
  int* d_ibeg = nullptr;
  int nig = 0;
  double* d_dui = nullptr;       // d J / d(inlet velocity), accumulated
  double* d_dui_part = nullptr;  // per-workgroup partials of one kernel
  double* d_rbuf = nullptr;      // restriction: adjoint of the restricted coarse cells
  double dui[3] = {0.0, 0.0, 0.0};

  template <typename P>
  P* Alloc(size_t count) {
    void* p = nullptr;
    ADJ_CHECK(hipMalloc(&p, (count == 0 ? 1 : count) * sizeof(P)));
    owned.push_back(p);
    return static_cast<P*>(p);
  }
  void FreeAll() {
    for (void* p : owned) (void)hipFree(p);  // teardown: nothing to report to
    owned.clear();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
    ev_fork = ev_join = nullptr;
    side = nullptr;
    stream = nullptr;
  }

  void Build(const AdjTopo& t) {
    nb = t.nb;
    omega = t.omega;
    prec = t.prec;
    selem = prec == 32 ? sizeof(float) : sizeof(double);
    const long long nfull = static_cast<long long>(nb) * kQ * kNS;
    if (nfull >= (1ll << 31)) throw std::invalid_argument("adjoint: too many blocks");
    origin_cw = t.origin_cw;
    auto kind = [&](int b, int ix, int iy, int iz) {
      return t.kinds[static_cast<size_t>(b) * kNC + (ix * kN + iy) * kN + iz];
    };
    std::unordered_map<int64_t, double> qmap;
    for (size_t k = 0; k < t.link_key.size(); ++k) qmap[t.link_key[k]] = t.link_q[k];
    // ---- rows of M over the full halo view: stream_block
    // (block_lattice.cpp:503-720) after the exchange
    auto resolve = [&](int b, int slot, int pop) {
      const int hs = t.halo_src[static_cast<size_t>(b) * kNS + slot];
      const int sb = hs >= 0 ? hs / kNS : b, ss = hs >= 0 ? hs % kNS : slot;
      return (sb * kQ + pop) * kNS + ss;
    };
    auto in_np = [](int v) { return v >= 0 && v < kNP; };
    auto in_core = [](int v) { return v >= 0 && v < kN; };
    std::vector<int> srcf(static_cast<size_t>(nfull));
    std::vector<AdjLink> links;
    // blocks in level order (a stable sort: one level keeps the block order),
    // so that every level's cells, entries and links are contiguous ranges
    border.resize(nb);
    for (int b = 0; b < nb; ++b) border[b] = b;
    if (!t.level.empty())
      std::stable_sort(border.begin(), border.end(), [&](int a, int c) { return t.level[a] < t.level[c]; });
    const int nlev = t.level.empty() ? 1 : 1 + *std::max_element(t.level.begin(), t.level.end());
    auto lvl = [&](int b) { return t.level.empty() ? 0 : t.level[b]; };
    lv_link.assign(nlev + 1, 0);
    for (int bi = 0; bi < nb; ++bi) {
      const int b = border[bi];
      lv_link[lvl(b) + 1] = -1;  // filled below
      for (int hx = 0; hx < kNP; ++hx)
        for (int hy = 0; hy < kNP; ++hy)
          for (int hz = 0; hz < kNP; ++hz) {
            const int ix = hx - 1, iy = hy - 1, iz = hz - 1;
            const bool core = in_core(ix) && in_core(iy) && in_core(iz);
            const int dst = (hx * kNP + hy) * kNP + hz;
            for (int i = 0; i < kQ; ++i) {
              const int e = (b * kQ + i) * kNS + dst;
              const int sx = hx - Cx(i), sy = hy - Cy(i), sz = hz - Cz(i);
              const bool s_in = in_np(sx) && in_np(sy) && in_np(sz);
              if (!core) {
                int s;
                if (s_in) {
                  s = (sx * kNP + sy) * kNP + sz;
                } else {
                  const int off = Cx(i) * kNP * kNP + Cy(i) * kNP + Cz(i);
                  s = ((dst - off) % kNS + kNS) % kNS;
                }
                srcf[e] = resolve(b, s, i);
                continue;
              }
              const int k = kind(b, ix, iy, iz);
              if (k == kSolid) {
                srcf[e] = e;
                continue;
              }
              const int self_opp = (b * kQ + Opp(i)) * kNS + dst;
              if (!s_in) {
                srcf[e] = ReflectsOnPull(k) ? self_opp : e;
                continue;
              }
              const bool s_core = in_core(sx - 1) && in_core(sy - 1) && in_core(sz - 1);
              if (ReflectsOnPull(k) && !s_core) {
                srcf[e] = self_opp;
                continue;
              }
              const int sk = s_core ? kind(b, sx - 1, sy - 1, sz - 1) : kBulk;
              if (sk == kSolid || sk == kBouzidi) {
                const int64_t key = (static_cast<int64_t>(b) * kNC + (ix * kN + iy) * kN + iz) * kQ + i;
                const auto it = qmap.find(key);
                const double qf = it == qmap.end() ? 1.0 : it->second;
                if (qf > 1.0e-10 && qf < 1.0 - 1.0e-10) {
                  const int nx = hx + Cx(i), ny = hy + Cy(i), nz = hz + Cz(i);
                  const int nk = (in_core(nx - 1) && in_core(ny - 1) && in_core(nz - 1))
                                     ? kind(b, nx - 1, ny - 1, nz - 1)
                                     : kBulk;
                  AdjLink L{};
                  L.row = e;
                  L.bb = self_opp;
                  L.upper = qf >= 0.5 - 1.0e-14;
                  L.other = L.upper ? (nk == kBulk ? resolve(b, (nx * kNP + ny) * kNP + nz, i) : e) : e;
                  L.q = qf;
                  link_index[key] = static_cast<int>(links.size());
                  srcf[e] = -1 - static_cast<int>(links.size());
                  links.push_back(L);
                  continue;
                }
                srcf[e] = self_opp;
                continue;
              }
              srcf[e] = resolve(b, (sx * kNP + sy) * kNP + sz, i);
            }
          }
      lv_link[lvl(b) + 1] = static_cast<int>(links.size());
    }
    for (int l = 1; l <= nlev; ++l) lv_link[l] = std::max(lv_link[l], lv_link[l - 1]);
    nlinks = static_cast<int>(links.size());
    // ---- live slots: the core and whatever the core's rows reach
    std::vector<uint8_t> live(static_cast<size_t>(nb) * kNS, 0);
    std::vector<int> work;
    auto mark = [&](int e) {
      const int b = e / (kQ * kNS), slot = e % kNS;
      const size_t s = static_cast<size_t>(b) * kNS + slot;
      if (!live[s]) {
        live[s] = 1;
        work.push_back(static_cast<int>(s));
      }
    };
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c < kNC; ++c) {
        const int ix = c / (kN * kN), iy = (c / kN) % kN, iz = c % kN;
        mark((b * kQ) * kNS + ((ix + 1) * kNP + iy + 1) * kNP + iz + 1);
      }
    while (!work.empty()) {
      const int s = work.back();
      work.pop_back();
      const int b = s / kNS, slot = s % kNS;
      for (int i = 0; i < kQ; ++i) {
        const int r = srcf[(static_cast<size_t>(b) * kQ + i) * kNS + slot];
        if (r >= 0) {
          mark(r);
        } else {
          mark(links[-1 - r].bb);
          mark(links[-1 - r].other);
        }
      }
    }
    // ---- compact numbering
    std::vector<int> rank(static_cast<size_t>(nb) * kNS, -1), boff(nb), cnt(nb), first_cell(nb);
    int total = 0;
    cells.clear();
    ckind.clear();
    cell_block.clear();
    cell_slot.clear();
    core_cell.assign(static_cast<size_t>(nb) * kNC, -1);
    lv_cell.assign(nlev + 1, 0);
    lv_entry.assign(nlev + 1, 0);
    block_off.assign(nb, 0);
    block_cnt.assign(nb, 0);
    for (int bi = 0; bi < nb; ++bi) {
      const int b = border[bi];
      // core cells first in (x, y, z) order, then the live halo slots
      std::vector<int> order;
      for (int c = 0; c < kNC; ++c)
        order.push_back(((c / (kN * kN) + 1) * kNP + (c / kN) % kN + 1) * kNP + c % kN + 1);
      for (int slot = 0; slot < kNS; ++slot) {
        const int px = slot / (kNP * kNP), py = (slot / kNP) % kNP, pz = slot % kNP;
        if (live[static_cast<size_t>(b) * kNS + slot] && !(in_core(px - 1) && in_core(py - 1) && in_core(pz - 1)))
          order.push_back(slot);
      }
      int r = 0;
      for (int slot : order) rank[static_cast<size_t>(b) * kNS + slot] = r++;
      cnt[b] = r;
      boff[b] = total;
      total += kQ * r;
      first_cell[b] = static_cast<int>(cells.size());
      for (int slot : order) {
        const int rk = rank[static_cast<size_t>(b) * kNS + slot];
        const int px = slot / (kNP * kNP), py = (slot / kNP) % kNP, pz = slot % kNP;
        const bool core = in_core(px - 1) && in_core(py - 1) && in_core(pz - 1);
        if (core) core_cell[static_cast<size_t>(b) * kNC + ((px - 1) * kN + py - 1) * kN + pz - 1] =
            static_cast<int>(cells.size());
        cells.push_back(Cell{boff[b] + rk, r});
        ckind.push_back(core ? kind(b, px - 1, py - 1, pz - 1) : kHalo);
        cell_block.push_back(b);
        cell_slot.push_back(slot);
      }
      lv_cell[lvl(b) + 1] = static_cast<int>(cells.size());
      lv_entry[lvl(b) + 1] = total;
      block_off[b] = boff[b];
      block_cnt[b] = r;
    }
    for (int l = 1; l <= nlev; ++l) {
      lv_cell[l] = std::max(lv_cell[l], lv_cell[l - 1]);
      lv_entry[l] = std::max(lv_entry[l], lv_entry[l - 1]);
    }
    ntot = total;
    ncells = static_cast<int>(cells.size());
    auto cmap = [&](int e) {
      const int b = e / (kQ * kNS), i = (e / kNS) % kQ, slot = e % kNS;
      return boff[b] + i * cnt[b] + rank[static_cast<size_t>(b) * kNS + slot];
    };
    auto cell_of = [&](int e) {
      const int b = e / (kQ * kNS), slot = e % kNS;
      return first_cell[b] + rank[static_cast<size_t>(b) * kNS + slot];
    };
    std::vector<int> src(static_cast<size_t>(ntot), 0);
    for (int b = 0; b < nb; ++b)
      for (int slot = 0; slot < kNS; ++slot) {
        if (rank[static_cast<size_t>(b) * kNS + slot] < 0) continue;
        for (int i = 0; i < kQ; ++i) {
          const int e = (b * kQ + i) * kNS + slot;
          src[cmap(e)] = srcf[e] >= 0 ? cmap(srcf[e]) : srcf[e];
        }
      }
    for (AdjLink& L : links) {
      L.bb_cell = cell_of(L.bb);
      L.bb_pop = (L.bb / kNS) % kQ;
      L.ot_cell = cell_of(L.other);
      L.ot_pop = (L.other / kNS) % kQ;
      L.row = cmap(L.row);
      L.bb = cmap(L.bb);
      L.other = cmap(L.other);
    }
    // ---- readers of every entry (the transpose of M)
    std::vector<int> nread(static_cast<size_t>(ntot), 0), ncopy(static_cast<size_t>(ntot), 0);
    for (int e = 0; e < ntot; ++e) {
      if (src[e] >= 0) {
        ++nread[src[e]];
        ++ncopy[src[e]];
      } else {
        const AdjLink& L = links[-1 - src[e]];
        ++nread[L.bb];
        ++nread[L.other];
      }
    }
    std::vector<int> rd(static_cast<size_t>(ntot), -1), ovo(1, 0), ovidx(static_cast<size_t>(ntot), -1);
    for (int s = 0; s < ntot; ++s) {
      if (nread[s] == 0) continue;
      if (nread[s] == 1 && ncopy[s] == 1) continue;  // filled below
      ovidx[s] = static_cast<int>(ovo.size()) - 1;
      ovo.push_back(ovo.back() + nread[s]);
      rd[s] = -2 - ovidx[s];
    }
    std::vector<int2> ove(static_cast<size_t>(ovo.back()));
    std::vector<int> ofill(ovo.begin(), ovo.end() - 1);
    for (int e = 0; e < ntot; ++e) {
      if (src[e] >= 0) {
        const int s = src[e];
        if (ovidx[s] < 0) {
          rd[s] = e;
        } else {
          ove[ofill[ovidx[s]]++] = make_int2(e, -1);
        }
      } else {
        const int lk = -1 - src[e];
        const AdjLink& L = links[lk];
        ove[ofill[ovidx[L.bb]]++] = make_int2(e, 2 * lk);
        ove[ofill[ovidx[L.other]]++] = make_int2(e, 2 * lk + 1);
      }
    }
    // ---- regular cells: entries computable from 27 neighbour-block bases
    nbase.assign(static_cast<size_t>(nb) * 27, make_int2(-1, 0));
    for (int b = 0; b < nb; ++b) {
      for (int o = 0; o < 27; ++o) {
        const int dx = o / 9 - 1, dy = (o / 3) % 3 - 1, dz = o % 3 - 1;
        if (dx == 0 && dy == 0 && dz == 0) {
          nbase[27 * static_cast<size_t>(b) + o] = make_int2(boff[b], cnt[b]);
          continue;
        }
        // the block that supplies this block's halo slot in that direction
        const int hx = dx < 0 ? 0 : (dx > 0 ? kNP - 1 : 1), hy = dy < 0 ? 0 : (dy > 0 ? kNP - 1 : 1),
                  hz = dz < 0 ? 0 : (dz > 0 ? kNP - 1 : 1);
        const int hs = t.halo_src[static_cast<size_t>(b) * kNS + (hx * kNP + hy) * kNP + hz];
        if (hs < 0) continue;
        nbase[27 * static_cast<size_t>(b) + o] = make_int2(boff[hs / kNS], cnt[hs / kNS]);
      }
    }
    auto regular_entry = [&](int b, int ix, int iy, int iz, int i) {
      const int ox = ix < 0 ? 0 : (ix > kN - 1 ? 2 : 1), oy = iy < 0 ? 0 : (iy > kN - 1 ? 2 : 1),
                oz = iz < 0 ? 0 : (iz > kN - 1 ? 2 : 1);
      const int2 nbb = nbase[27 * static_cast<size_t>(b) + (ox * 3 + oy) * 3 + oz];
      if (nbb.x < 0) return -1;
      return nbb.x + i * nbb.y + (((ix & 7) * kN + (iy & 7)) * kN + (iz & 7));
    };
    cblock.assign(cells.size(), 0);
    cflag.assign(cells.size(), 0);
    for (size_t cl = 0; cl < cells.size(); ++cl) {
      const int b = cell_block[cl];
      cblock[cl] = b;
      const int r = cells[cl].base - boff[b];
      if (r >= kNC) continue;  // halo slot
      const int ix = r >> 6, iy = (r >> 3) & 7, iz = r & 7;
      bool fwd = true, rev = true;
      for (int i = 0; i < kQ; ++i) {
        const int e = cells[cl].base + i * cells[cl].stride;
        const int fs = regular_entry(b, ix - Cx(i), iy - Cy(i), iz - Cz(i), i);
        fwd = fwd && fs >= 0 && src[e] == fs;
        const int rs = regular_entry(b, ix + Cx(i), iy + Cy(i), iz + Cz(i), i);
        rev = rev && rs >= 0 && rd[e] == rs;
      }
      cflag[cl] = static_cast<uint8_t>((fwd ? 1 : 0) | (rev ? 2 : 0));
    }
    ADJ_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    d_cells = Alloc<Cell>(cells.size());
    d_cblock = Alloc<int>(cblock.size());
    d_cflag = Alloc<uint8_t>(cflag.size());
    d_nbase = Alloc<int2>(nbase.size());
    ADJ_CHECK(hipMemcpy(d_cblock, cblock.data(), cblock.size() * sizeof(int), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_cflag, cflag.data(), cflag.size(), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_nbase, nbase.data(), nbase.size() * sizeof(int2), hipMemcpyHostToDevice));
    d_ckind = Alloc<uint8_t>(ckind.size());
    d_src = Alloc<int>(src.size());
    d_rd = Alloc<int>(rd.size());
    d_ovo = Alloc<int>(ovo.size());
    d_ove = Alloc<int2>(ove.size());
    d_links = Alloc<AdjLink>(links.size());
    d_dq = Alloc<double>(links.size());
    const size_t st = static_cast<size_t>(ntot);
    d_y[0] = Alloc<unsigned char>(st * selem);
    d_y[1] = Alloc<unsigned char>(st * selem);
    d_lx = Alloc<double>(st);
    d_ln = Alloc<double>(st);
    d_final = Alloc<unsigned char>(st * selem);
    ADJ_CHECK(hipMemcpy(d_cells, cells.data(), cells.size() * sizeof(Cell), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_ckind, ckind.data(), ckind.size(), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_rd, rd.data(), rd.size() * sizeof(int), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_ovo, ovo.data(), ovo.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ove.empty()) ADJ_CHECK(hipMemcpy(d_ove, ove.data(), ove.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!links.empty())
      ADJ_CHECK(hipMemcpy(d_links, links.data(), links.size() * sizeof(AdjLink), hipMemcpyHostToDevice));
  }

  // drag terms of ol_drag's predicate: fluid cell, in-block solid-like
  // neighbour along c_i whose centre lies in bbox; and per seeded cell the
  // weights of its post-collide populations
  void Terms(const double* dir, const double* bbox, std::vector<DragTerm>* terms, std::vector<int>* seed_cell,
             std::vector<double>* seed_w) const {
    for (int b = 0; b < nb; ++b) {
      const double* o = &origin_cw[4 * static_cast<size_t>(b)];
      for (int ix = 0; ix < kN; ++ix)
        for (int iy = 0; iy < kN; ++iy)
          for (int iz = 0; iz < kN; ++iz) {
            const int cell = core_cell[static_cast<size_t>(b) * kNC + (ix * kN + iy) * kN + iz];
            if (ckind[cell] != kBulk) continue;
            bool seeded = false;
            for (int i = 1; i < kQ; ++i) {
              const int x = ix + Cx(i), y = iy + Cy(i), z = iz + Cz(i);
              if (x < 0 || x >= kN || y < 0 || y >= kN || z < 0 || z >= kN) continue;
              const uint8_t nk = ckind[core_cell[static_cast<size_t>(b) * kNC + (x * kN + y) * kN + z]];
              if (nk != kSolid && nk != kBouzidi) continue;
              const double px = o[0] + (x + 0.5) * o[3], py = o[1] + (y + 0.5) * o[3], pz = o[2] + (z + 0.5) * o[3];
              if (!(px >= bbox[0] && px <= bbox[3] && py >= bbox[1] && py <= bbox[4] && pz >= bbox[2] &&
                    pz <= bbox[5]))
                continue;
              const Cell c = cells[cell];
              DragTerm d;
              d.snap = c.base + i * c.stride;
              d.live = c.base + Opp(i) * c.stride;
              d.w = -(dir[0] * Cx(i) + dir[1] * Cy(i) + dir[2] * Cz(i));
              terms->push_back(d);
              if (!seeded) {  // per drag cell: post-collide weights, then live weights
                seed_cell->push_back(cell);
                seed_w->resize(seed_w->size() + 2 * kQ, 0.0);
                seeded = true;
              }
              (*seed_w)[seed_w->size() - 2 * kQ + i] += d.w;
              (*seed_w)[seed_w->size() - kQ + Opp(i)] += d.w;
            }
          }
    }
  }

  // ---- multi-level engine: tables on top of Build's (call after Build)
  template <typename P>
  P* Upload(const std::vector<P>& v) {
    P* d = Alloc<P>(v.size());
    if (!v.empty()) ADJ_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(P), hipMemcpyHostToDevice));
    return d;
  }

  void BuildML(const AdjTopo& t, const AdjMLTopo& m) {
    ml = true;
    mtopo = m;
    nlev = m.max_level + 1;
    d_boff = Upload(block_off);
    d_cnt = Upload(block_cnt);
    d_ocw = Upload(origin_cw);
    d_kind_blk = Upload(t.kinds);
    d_bt = Upload(m.bt);
    d_scm = Upload(std::vector<SimConst>{m.sc});
    nfd = static_cast<int>(m.fd_blocks.size());
    d_fd_blocks = Upload(m.fd_blocks);
    d_fd_slot = Upload(m.fd_slot);
    d_fd_list = Upload(m.fd_list);
    d_fd_count = Upload(m.fd_count);
    if (nfd > 0) {
      d_fd_rbeg = Alloc<int>(static_cast<size_t>(nfd) * kFdRevBeg);
      d_fd_rent = Alloc<FdRevEntry>(static_cast<size_t>(nfd) * kFdRevMax);
      LaunchMLFdRevBuild(d_fd_list, d_fd_count, d_fd_slot, nfd, d_fd_rbeg, d_fd_rent, stream);
      ADJ_CHECK(hipGetLastError());
    }
    npts = static_cast<int>(m.plan.size() / 12);
    d_plan = Upload(m.plan);
    // prolongation: per coarse level, the active points and their grouping by
    // coarse cell (first appearance order; points of a group in list order)
    auto cell_key = [](int b, int c) { return static_cast<int64_t>(b) * kNC + c; };
    d_plist.assign(nlev, nullptr);
    d_ppts.assign(nlev, nullptr);
    d_pbeg.assign(nlev, nullptr);
    np_l.assign(nlev, 0);
    npg_l.assign(nlev, 0);
    auto group = [&](const std::vector<int>& list, std::vector<int>* pts, std::vector<int>* beg) {
      std::unordered_map<int64_t, int> gi;
      std::vector<std::vector<int>> gs;
      for (int p : list) {
        const int* r = &m.plan[12 * static_cast<size_t>(p)];
        const int64_t key = cell_key(r[0], (r[1] * kN + r[2]) * kN + r[3]);
        auto it = gi.find(key);
        if (it == gi.end()) {
          it = gi.emplace(key, static_cast<int>(gs.size())).first;
          gs.emplace_back();
        }
        gs[it->second].push_back(p);
      }
      pts->clear();
      beg->assign(1, 0);
      for (const auto& g : gs) {
        pts->insert(pts->end(), g.begin(), g.end());
        beg->push_back(static_cast<int>(pts->size()));
      }
      return static_cast<int>(gs.size());
    };
    for (int l = 0; l + 1 < nlev; ++l) {
      const std::vector<int>& list = m.prolong_active[l];
      np_l[l] = static_cast<int>(list.size());
      d_plist[l] = Upload(list);
      std::vector<int> pts, beg;
      npg_l[l] = group(list, &pts, &beg);
      d_ppts[l] = Upload(pts);
      d_pbeg[l] = Upload(beg);
    }
    {
      int most = 1;
      for (int l = 0; l < nlev; ++l) most = std::max(most, np_l[l]);
      d_pcm = Alloc<double>(static_cast<size_t>(most) * kMacro);  // per-point macro adjoints of a prolongation VJP
    }
    {
      std::vector<int> all(npts), pts, beg;
      for (int p = 0; p < npts; ++p) all[p] = p;
      nig = group(all, &pts, &beg);
      d_ipts = Upload(pts);
      d_ibeg = Upload(beg);
    }
    // restriction: per fine level, the active points and, per fine cell they
    // read, the (point, neighbour slot) entries in point order
    d_rlist.assign(nlev, nullptr);
    d_rent.assign(nlev, nullptr);
    d_rebeg.assign(nlev, nullptr);
    d_recell.assign(nlev, nullptr);
    nr_l.assign(nlev, 0);
    nrc_l.assign(nlev, 0);
    size_t rmax = 1;
    for (int l = 1; l < nlev; ++l) {
      const std::vector<int>& list = m.restrict_active[l];
      nr_l[l] = static_cast<int>(list.size());
      rmax = std::max(rmax, list.size());
      d_rlist[l] = Upload(list);
      std::unordered_map<int64_t, int> gi;
      std::vector<std::vector<int2>> gs;
      std::vector<int> cells_bc;
      for (int j = 0; j < static_cast<int>(list.size()); ++j) {
        const int* r = &m.plan[12 * static_cast<size_t>(list[j])];
        for (int jp = 0; jp < kQ; ++jp) {
          const int x = r[5] + Cx(jp), y = r[6] + Cy(jp), z = r[7] + Cz(jp);
          if (x < 0 || x >= kN || y < 0 || y >= kN || z < 0 || z >= kN) continue;
          const int c = (x * kN + y) * kN + z;
          const int64_t key = cell_key(r[4], c);
          auto it = gi.find(key);
          if (it == gi.end()) {
            it = gi.emplace(key, static_cast<int>(gs.size())).first;
            gs.emplace_back();
            cells_bc.push_back(r[4]);
            cells_bc.push_back(c);
          }
          gs[it->second].push_back(make_int2(j, jp));
        }
      }
      std::vector<int2> ent;
      std::vector<int> beg(1, 0);
      for (const auto& g : gs) {
        ent.insert(ent.end(), g.begin(), g.end());
        beg.push_back(static_cast<int>(ent.size()));
      }
      nrc_l[l] = static_cast<int>(gs.size());
      d_rent[l] = Upload(ent);
      d_rebeg[l] = Upload(beg);
      d_recell[l] = Upload(cells_bc);
    }
    d_rbuf = Alloc<double>(rmax * kQ);
    d_dui = Alloc<double>(3);
    d_dui_fd = Alloc<double>(3);
    d_dui_fd_part = Alloc<double>(3 * static_cast<size_t>(std::max(nfd, 1)));
    fd_level_mask = 0;
    for (int b : m.fd_blocks) fd_level_mask |= 1u << m.level[b];
    ADJ_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    ADJ_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    ADJ_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    // per level, its velocity-Dirichlet cells (their collide VJP runs apart)
    {
      std::vector<int> vc;
      lv_vcell.assign(nlev + 1, 0);
      for (int l = 0; l < nlev; ++l) {
        lv_vcell[l] = static_cast<int>(vc.size());
        for (int cl = lv_cell[l]; cl < lv_cell[l + 1]; ++cl)
          if (ckind[cl] == kVelocity) vc.push_back(cl);
      }
      lv_vcell[nlev] = static_cast<int>(vc.size());
      d_vcell = Upload(vc);
    }
    d_dui_part = Alloc<double>(3 * static_cast<size_t>(std::max<long long>(Blocks(ncells), std::max(nfd, 1))));
    // the coarse step as operations (TimeLoop::advance, time_loop.h:201-240)
    ops.clear();
    std::function<void(int)> adv = [&](int l) {
      ops.push_back(make_int2(kOpVisit, l));
      if (nfd > 0) ops.push_back(make_int2(kOpFd, l));
      if (l + 1 >= nlev) return;
      ops.push_back(make_int2(kOpHalf, l));
      adv(l + 1);
      ops.push_back(make_int2(kOpFull, l));
      adv(l + 1);
      ops.push_back(make_int2(kOpRestrict, l + 1));
    };
    adv(0);
    // what each operation's adjoint reads of its input state: a contiguous
    // window of entries (its level; for the FD pass the span of the FD
    // blocks), plus the prev macros for the prolongations
    long long fd_lo = 0, fd_hi = 0;
    for (size_t k = 0; k < m.fd_blocks.size(); ++k) {
      const int b = m.fd_blocks[k];
      const long long lo = block_off[b], hi = lo + static_cast<long long>(kQ) * block_cnt[b];
      fd_lo = k == 0 ? lo : std::min(fd_lo, lo);
      fd_hi = k == 0 ? hi : std::max(fd_hi, hi);
    }
    win_lo.assign(ops.size(), 0);
    win_hi.assign(ops.size(), 0);
    win_off.assign(ops.size() + 1, 0);
    for (size_t k = 0; k < ops.size(); ++k) {
      const int2 op = ops[k];
      win_lo[k] = op.x == kOpFd ? fd_lo : lv_entry[op.y];
      win_hi[k] = op.x == kOpFd ? fd_hi : lv_entry[op.y + 1];
      win_off[k + 1] = win_off[k] + (win_hi[k] - win_lo[k]);
    }
  }

  template <typename S>
  FdArgs<S> MLFdArgs() const {
    FdArgs<S> a{};
    a.kind = d_kind_blk;
    a.bt = d_bt;
    a.scm = (decltype(a.scm))d_scm;
    a.fd_list = d_fd_list;
    a.fd_count = d_fd_count;
    a.sc = mtopo.sc;
    return a;
  }

  // operation k forward on (x, y, prev) at time t; when tape is given (in
  // state coordinates: entry e of the window at tape[e]), the window of the
  // state that the operation's adjoint reads is copied there first, and for a
  // prolongation the prev macros into ptape
  template <typename S>
  void MLApply(size_t k, S* x, S* y, S* prev, double t, S* tape, S* ptape) {
    const size_t pm = static_cast<size_t>(npts) * kMacro;
    const int2 op = ops[k];
    const int l = op.y;
    // a visit's window is the entries of its level's cells: its collide tapes them
    if (tape && op.x != kOpVisit && win_hi[k] > win_lo[k])
      ADJ_CHECK(hipMemcpyAsync(tape + win_lo[k], x + win_lo[k], static_cast<size_t>(win_hi[k] - win_lo[k]) * sizeof(S),
                               hipMemcpyDeviceToDevice, stream));
    switch (op.x) {
      case kOpVisit: {
        const int c0 = lv_cell[l], c1 = lv_cell[l + 1];
        if (c1 <= c0) break;
        hipLaunchKernelGGL(KMLCollide<S>, dim3(Blocks(c1 - c0)), dim3(kThreads), 0, stream, x, y, d_cells, d_ckind,
                           d_cblock, d_boff, d_ocw, d_scm, t, c0, c1, tape);
        hipLaunchKernelGGL(KMLGather<S>, dim3(Blocks(c1 - c0)), dim3(kThreads), 0, stream, y, x, d_src, d_links,
                           d_cells, c0, c1);
        break;
      }
      case kOpFd:
        LaunchMLFd<S>(MLFdArgs<S>(), d_fd_blocks, d_fd_slot, nfd, x, d_boff, d_cnt, t, stream);
        break;
      case kOpHalf:
      case kOpFull:
        if (tape && pm) ADJ_CHECK(hipMemcpyAsync(ptape, prev, pm * sizeof(S), hipMemcpyDeviceToDevice, stream));
        LaunchMLProlong<S>(d_plan, d_plist[l], np_l[l], op.x == kOpHalf ? 1 : 0, x, d_boff, d_cnt, prev,
                           mtopo.prolong_scale, stream);
        break;
      case kOpRestrict:
        LaunchMLRestrict<S>(d_plan, d_rlist[l], nr_l[l], x, d_boff, d_cnt, mtopo.restrict_scale, stream);
        break;
    }
    ADJ_CHECK(hipGetLastError());
  }

  // the adjoint of operation op (its tape as MLApply left it)
  template <typename S>
  void MLReverse(int2 op, double t, const S* tape, const S* ptape, double* lp, double snap_scale,
                 const int* cseed, const double* seed_w) {
    const int l = op.y;
    switch (op.x) {
      case kOpVisit: {
        const int c0 = lv_cell[l], c1 = lv_cell[l + 1];
        if (c1 <= c0) break;
        const int k0 = lv_link[l], k1 = lv_link[l + 1];
        if (k1 > k0)
          hipLaunchKernelGGL(KMLLinkGrad<S>, dim3(Blocks(k1 - k0)), dim3(kThreads), 0, stream, d_lx, tape, d_cells,
                             d_ckind, d_cblock, d_boff, d_ocw, d_scm, t, d_links, k0, k1, d_dq);
        hipLaunchKernelGGL(KMLVisitVjp<S>, dim3(Blocks(c1 - c0)), dim3(kThreads), 0, stream, d_lx, d_ln, tape, d_rd,
                           d_ovo, d_ove, d_links, d_cells, d_ckind, d_scm, cseed, seed_w, snap_scale, c0, c1);
        const int nv = lv_vcell[l + 1] - lv_vcell[l];
        if (nv > 0) {
          const unsigned nwg = Blocks(nv);
          hipLaunchKernelGGL(KMLVelocityVjp<S>, dim3(nwg), dim3(kThreads), 0, stream, d_ln, tape,
                             d_vcell + lv_vcell[l], nv, d_cells, d_cblock, d_boff, d_ocw, d_scm, t, d_dui_part);
          hipLaunchKernelGGL(KMLAccum, dim3(1), dim3(kAccumThreads), 0, stream, d_dui_part, static_cast<int>(nwg),
                             d_dui);
        }
        // the level's new adjoint is in d_ln: copy it back, or, when the other
        // levels hold fewer entries, copy those across and swap the buffers
        const size_t mine = static_cast<size_t>(lv_entry[l + 1] - lv_entry[l]);
        const size_t rest = static_cast<size_t>(ntot) - mine;
        if (rest < mine) {
          if (fd_inflight) {  // the FD VJP beside this visit writes entries of the other levels
            ADJ_CHECK(hipStreamWaitEvent(stream, ev_join, 0));
            fd_inflight = false;
          }
          if (lv_entry[l] > 0)
            ADJ_CHECK(hipMemcpyAsync(d_ln, d_lx, static_cast<size_t>(lv_entry[l]) * sizeof(double),
                                     hipMemcpyDeviceToDevice, stream));
          if (lv_entry[l + 1] < ntot)
            ADJ_CHECK(hipMemcpyAsync(d_ln + lv_entry[l + 1], d_lx + lv_entry[l + 1],
                                     static_cast<size_t>(ntot - lv_entry[l + 1]) * sizeof(double),
                                     hipMemcpyDeviceToDevice, stream));
          std::swap(d_lx, d_ln);
        } else {
          ADJ_CHECK(hipMemcpyAsync(d_lx + lv_entry[l], d_ln + lv_entry[l], mine * sizeof(double),
                                   hipMemcpyDeviceToDevice, stream));
        }
        break;
      }
      case kOpFd: {
        hipStream_t fs = fd_side ? side : stream;
        LaunchMLFdVjp<S>(MLFdArgs<S>(), d_fd_blocks, d_fd_slot, nfd, tape, d_lx, d_boff, d_cnt, t, d_fd_rbeg,
                         d_fd_rent, d_dui_fd_part, fs);
        hipLaunchKernelGGL(KMLAccum, dim3(1), dim3(kAccumThreads), 0, fs, d_dui_fd_part, nfd, d_dui_fd);
        break;
      }
      case kOpHalf:
      case kOpFull:
        LaunchMLProlongVjp<S>(d_plan, d_ppts[l], d_pbeg[l], npg_l[l], np_l[l], op.x == kOpHalf ? 1 : 0, tape, d_lx,
                              d_boff, d_cnt, ptape, lp, mtopo.prolong_scale, d_pcm, stream);
        break;
      case kOpRestrict:
        LaunchMLRestrictTake(d_plan, d_rlist[l], nr_l[l], d_lx, d_boff, d_cnt, d_rbuf, stream);
        LaunchMLRestrictVjp<S>(d_plan, d_rlist[l], d_rent[l], d_rebeg[l], d_recell[l], nrc_l[l], tape, d_lx, d_boff,
                               d_cnt, d_rbuf, mtopo.restrict_scale, stream);
        break;
    }
    ADJ_CHECK(hipGetLastError());
  }

  // Multi-level run: steps are coarse steps (t = 0 for the first); J is the
  // mean of dir . F after each coarse step of the window; dJ/du0 of the
  // initial feq(rho0, u0) on the masked cells, dJ/d(inlet velocity) in dui,
  // dJ/dq in d_dq.  Memory: checkpoints every K steps, the K states of one
  // segment, and per step one tape per operation.
  template <typename S>
  void RunMLT(int steps, int w0, int w1, int K, const double* dir, const double* bbox, double rho0, const double* u0,
              uint32_t mask, double* J, double* du0) {
    std::vector<DragTerm> terms;
    std::vector<int> seed_cell;
    std::vector<double> seed_w;
    Terms(dir, bbox, &terms, &seed_cell, &seed_w);
    const int nterms = static_cast<int>(terms.size()), nseed = static_cast<int>(seed_cell.size());
    const int nck = (steps + K - 1) / K;
    const size_t st = static_cast<size_t>(ntot), pm = static_cast<size_t>(npts) * kMacro;
    const size_t pmn = std::max<size_t>(pm, 1);
    std::vector<void*> run_owned;
    auto alloc = [&](size_t bytes) {
      void* p = nullptr;
      ADJ_CHECK(hipMalloc(&p, bytes == 0 ? 1 : bytes));
      run_owned.push_back(p);
      return p;
    };
    try {
      DragTerm* d_terms = static_cast<DragTerm*>(alloc(terms.size() * sizeof(DragTerm)));
      std::vector<int> cseed(static_cast<size_t>(ncells), -1);
      for (int k = 0; k < nseed; ++k) cseed[seed_cell[k]] = k;
      int* d_cseed = static_cast<int*>(alloc(cseed.size() * sizeof(int)));
      double* d_seed_w = static_cast<double*>(alloc(std::max<size_t>(seed_w.size(), 1) * sizeof(double)));
      const int ndg = std::max(1, static_cast<int>(Blocks(nterms)));
      double* d_jv = static_cast<double*>(alloc(static_cast<size_t>(steps + 1) * ndg * sizeof(double)));
      S* x = static_cast<S*>(alloc(st * sizeof(S)));
      S* y = static_cast<S*>(alloc(st * sizeof(S)));
      S* pv = static_cast<S*>(alloc(pmn * sizeof(S)));
      S* ck = static_cast<S*>(alloc(static_cast<size_t>(nck) * st * sizeof(S)));
      S* ckp = static_cast<S*>(alloc(static_cast<size_t>(nck) * pmn * sizeof(S)));
      const size_t nops = ops.size();
      // Reverse sweep, per checkpoint segment.  Packed mode (when K steps of
      // tapes fit in a share of the free device memory): one forward pass
      // over the segment writes every step's tapes, then the steps are
      // reversed.  Otherwise the segment's states are recomputed first and
      // each step is re-run from its state to tape it just before its reverse.
      const size_t step_tape = static_cast<size_t>(win_off[nops]);
      const size_t step_ptape = nops * pmn;
      size_t free_b = 0, total_b = 0;
      ADJ_CHECK(hipMemGetInfo(&free_b, &total_b));
      const double packed_bytes =
          static_cast<double>(K) * static_cast<double>(step_tape + step_ptape) * static_cast<double>(sizeof(S));
      const double budget = tape_budget < 0 ? 0.6 * static_cast<double>(free_b) : static_cast<double>(tape_budget);
      const bool packed = packed_bytes <= budget;
      S* seg = packed ? nullptr : static_cast<S*>(alloc(static_cast<size_t>(K) * st * sizeof(S)));
      S* segp = packed ? nullptr : static_cast<S*>(alloc(static_cast<size_t>(K) * pmn * sizeof(S)));
      S* tape = static_cast<S*>(alloc((packed ? static_cast<size_t>(K) * step_tape : nops * st) * sizeof(S)));
      S* ptape = static_cast<S*>(alloc((packed ? static_cast<size_t>(K) : 1) * step_ptape * sizeof(S)));
      // tape of operation k at step j of the segment, in state coordinates
      auto tape_of = [&](int j, size_t k) -> S* {
        if (!packed) return tape + k * st;
        const uintptr_t at = reinterpret_cast<uintptr_t>(tape + static_cast<size_t>(j) * step_tape + win_off[k]);
        return reinterpret_cast<S*>(at - static_cast<uintptr_t>(win_lo[k]) * sizeof(S));
      };
      auto ptape_of = [&](int j, size_t k) -> S* {
        return ptape + (packed ? static_cast<size_t>(j) * step_ptape : 0) + k * pmn;
      };
      packed_mode = packed;
      double* lp = static_cast<double*>(alloc(pmn * sizeof(double)));
      const unsigned ncb = Blocks(ncells);
      double* d_part = static_cast<double*>(alloc(static_cast<size_t>(ncb) * 3 * sizeof(double)));
      ADJ_CHECK(hipMemcpyAsync(d_cseed, cseed.data(), cseed.size() * sizeof(int), hipMemcpyHostToDevice, stream));
      if (nterms) {
        ADJ_CHECK(hipMemcpyAsync(d_terms, terms.data(), terms.size() * sizeof(DragTerm), hipMemcpyHostToDevice, stream));
        ADJ_CHECK(hipMemcpyAsync(d_seed_w, seed_w.data(), seed_w.size() * sizeof(double), hipMemcpyHostToDevice,
                                 stream));
      }
      ADJ_CHECK(hipMemsetAsync(d_jv, 0, static_cast<size_t>(steps + 1) * ndg * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(d_dq, 0, static_cast<size_t>(nlinks > 0 ? nlinks : 1) * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(d_dui, 0, 3 * sizeof(double), stream));
      if (d_dui_fd) ADJ_CHECK(hipMemsetAsync(d_dui_fd, 0, 3 * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(x, 0, st * sizeof(S), stream));
      ADJ_CHECK(hipMemsetAsync(y, 0, st * sizeof(S), stream));
      hipEvent_t ev[3];
      for (auto& e : ev) ADJ_CHECK(hipEventCreate(&e));
      ADJ_CHECK(hipEventRecord(ev[0], stream));
      auto step_fwd = [&](S* xs, S* ps, double t) {
        for (size_t k = 0; k < ops.size(); ++k) MLApply<S>(k, xs, y, ps, t, nullptr, nullptr);
      };
      // ---- forward sweep: initial state, coupler init, checkpoints, J
      hipLaunchKernelGGL(KAdjInit<S>, dim3(ncb), dim3(kThreads), 0, stream, x, d_cells, d_ckind, ncells, rho0, u0[0],
                         u0[1], u0[2], mask);
      LaunchMLCouplerInit<S>(d_plan, npts, x, d_boff, d_cnt, pv, stream);
      for (int s = 1; s <= steps; ++s) {
        if ((s - 1) % K == 0) {
          const int c = (s - 1) / K;
          ADJ_CHECK(hipMemcpyAsync(ck + static_cast<size_t>(c) * st, x, st * sizeof(S), hipMemcpyDeviceToDevice, stream));
          if (pm)
            ADJ_CHECK(hipMemcpyAsync(ckp + static_cast<size_t>(c) * pmn, pv, pm * sizeof(S), hipMemcpyDeviceToDevice,
                                     stream));
        }
        step_fwd(x, pv, static_cast<double>(s - 1));
        if (s > w0 && s <= w1 && nterms)
          hipLaunchKernelGGL(KMLDrag<S>, dim3(ndg), dim3(kThreads), 0, stream, y, x, d_terms, nterms,
                             d_jv + static_cast<size_t>(s) * ndg);
      }
      ADJ_CHECK(hipMemcpyAsync(d_final, x, st * sizeof(S), hipMemcpyDeviceToDevice, stream));
      ADJ_CHECK(hipEventRecord(ev[1], stream));
      // ---- reverse, one checkpoint segment at a time, one step at a time
      const double scale = 1.0 / static_cast<double>(w1 - w0);
      ADJ_CHECK(hipMemsetAsync(d_lx, 0, st * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(d_ln, 0, st * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(lp, 0, pmn * sizeof(double), stream));
      auto reverse_step = [&](int s, int j) {
        const double t = static_cast<double>(s - 1);
        const bool win = s > w0 && s <= w1 && nterms > 0;
        if (win) hipLaunchKernelGGL(KAdjSeedLive, dim3(Blocks(nterms)), dim3(kThreads), 0, stream, d_lx, d_terms,
                                    nterms, scale);
        std::vector<char> seeded(nlev, 0);
        bool joined = true;
        for (size_t k = nops; k-- > 0;) {
          double ss = 0.0;
          if (ops[k].x == kOpVisit && win && !seeded[ops[k].y]) {
            ss = scale;  // the level's last visit took the drag snapshot
            seeded[ops[k].y] = 1;
          }
          // an FD VJP followed (in reverse order) by a visit of a level without
          // FD blocks: the two touch disjoint adjoint entries and run side by side
          const bool fork = ops[k].x == kOpFd && k > 0 && ops[k - 1].x == kOpVisit &&
                            !((fd_level_mask >> ops[k - 1].y) & 1u);
          if (fork) {
            ADJ_CHECK(hipEventRecord(ev_fork, stream));
            ADJ_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
            fd_side = true;
          }
          MLReverse<S>(ops[k], t, tape_of(j, k), ptape_of(j, k), lp, ss, d_cseed, d_seed_w);
          fd_side = false;
          if (fork) {
            ADJ_CHECK(hipEventRecord(ev_join, side));
            joined = false;
            fd_inflight = true;
          } else if (!joined) {
            ADJ_CHECK(hipStreamWaitEvent(stream, ev_join, 0));
            joined = true;
            fd_inflight = false;
          }
        }
      };
      for (int c = nck - 1; c >= 0; --c) {
        const int s0 = c * K, s1 = std::min(steps, s0 + K);
        if (packed) {
          // one pass over the segment from its checkpoint, taping every step
          ADJ_CHECK(hipMemcpyAsync(x, ck + static_cast<size_t>(c) * st, st * sizeof(S), hipMemcpyDeviceToDevice, stream));
          if (pm)
            ADJ_CHECK(hipMemcpyAsync(pv, ckp + static_cast<size_t>(c) * pmn, pm * sizeof(S), hipMemcpyDeviceToDevice,
                                     stream));
          for (int s = s0 + 1; s <= s1; ++s)
            for (size_t k = 0; k < nops; ++k)
              MLApply<S>(k, x, y, pv, static_cast<double>(s - 1), tape_of(s - 1 - s0, k), ptape_of(s - 1 - s0, k));
          for (int s = s1; s > s0; --s) reverse_step(s, s - 1 - s0);
          continue;
        }
        // segment states x_{s0} .. x_{s1 - 1} (and their prev macros)
        ADJ_CHECK(hipMemcpyAsync(seg, ck + static_cast<size_t>(c) * st, st * sizeof(S), hipMemcpyDeviceToDevice, stream));
        if (pm)
          ADJ_CHECK(hipMemcpyAsync(segp, ckp + static_cast<size_t>(c) * pmn, pm * sizeof(S), hipMemcpyDeviceToDevice,
                                   stream));
        for (int j = 1; j < s1 - s0; ++j) {
          ADJ_CHECK(hipMemcpyAsync(seg + static_cast<size_t>(j) * st, seg + static_cast<size_t>(j - 1) * st,
                                   st * sizeof(S), hipMemcpyDeviceToDevice, stream));
          if (pm)
            ADJ_CHECK(hipMemcpyAsync(segp + static_cast<size_t>(j) * pmn, segp + static_cast<size_t>(j - 1) * pmn,
                                     pm * sizeof(S), hipMemcpyDeviceToDevice, stream));
          step_fwd(seg + static_cast<size_t>(j) * st, segp + static_cast<size_t>(j) * pmn,
                   static_cast<double>(s0 + j - 1));
        }
        for (int s = s1; s > s0; --s) {
          // re-run step s from x_{s-1}, taping what each operation's adjoint reads
          ADJ_CHECK(hipMemcpyAsync(x, seg + static_cast<size_t>(s - 1 - s0) * st, st * sizeof(S),
                                   hipMemcpyDeviceToDevice, stream));
          if (pm)
            ADJ_CHECK(hipMemcpyAsync(pv, segp + static_cast<size_t>(s - 1 - s0) * pmn, pm * sizeof(S),
                                     hipMemcpyDeviceToDevice, stream));
          for (size_t k = 0; k < nops; ++k)
            MLApply<S>(k, x, y, pv, static_cast<double>(s - 1), tape_of(0, k), ptape_of(0, k));
          reverse_step(s, 0);
        }
      }
      // the initial prev macros came from the initial state
      LaunchMLCouplerInitVjp<S>(d_plan, d_ipts, d_ibeg, nig, ck, d_lx, d_boff, d_cnt, lp, stream);
      hipLaunchKernelGGL(KAdjInitGrad, dim3(ncb), dim3(kThreads), 0, stream, d_lx, d_cells, d_ckind, ncells, rho0,
                         u0[0], u0[1], u0[2], mask, d_part);
      ADJ_CHECK(hipGetLastError());
      ADJ_CHECK(hipEventRecord(ev[2], stream));
      std::vector<double> jv(static_cast<size_t>(steps + 1) * ndg), part(static_cast<size_t>(ncb) * 3);
      ADJ_CHECK(hipMemcpyAsync(jv.data(), d_jv, jv.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      ADJ_CHECK(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      double dui_fd[3] = {0.0, 0.0, 0.0};
      ADJ_CHECK(hipMemcpyAsync(dui, d_dui, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
      if (d_dui_fd) ADJ_CHECK(hipMemcpyAsync(dui_fd, d_dui_fd, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
      ADJ_CHECK(hipStreamSynchronize(stream));
      for (int d = 0; d < 3; ++d) dui[d] += dui_fd[d];
      float a = 0.f, b = 0.f;
      ADJ_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
      ADJ_CHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
      fwd_ms = a;
      bwd_ms = b;
      for (auto& e : ev) ADJ_CHECK(hipEventDestroy(e));
      double js = 0.0;
      for (int s = w0 + 1; s <= w1; ++s) {
        double step_sum = 0.0;
        for (int g = 0; g < ndg; ++g) step_sum += jv[static_cast<size_t>(s) * ndg + g];
        js += step_sum;
      }
      *J = js * scale;
      du0[0] = du0[1] = du0[2] = 0.0;
      for (unsigned g = 0; g < ncb; ++g)
        for (int d = 0; d < 3; ++d) du0[d] += part[3 * g + d];
    } catch (...) {
      (void)hipStreamSynchronize(stream);
      for (void* p : run_owned) (void)hipFree(p);
      throw;
    }
    ADJ_CHECK(hipStreamSynchronize(stream));
    for (void* p : run_owned) ADJ_CHECK(hipFree(p));
  }

  void Run(int steps, int w0, int w1, int K, const double* dir, const double* bbox, double rho0, const double* u0,
           uint32_t mask, double* J, double* du0) {
    if (ml) {
      if (prec == 32) {
        RunMLT<float>(steps, w0, w1, K, dir, bbox, rho0, u0, mask, J, du0);
      } else {
        RunMLT<double>(steps, w0, w1, K, dir, bbox, rho0, u0, mask, J, du0);
      }
      return;
    }
    dui[0] = dui[1] = dui[2] = 0.0;
    if (prec == 32) {
      RunT<float>(steps, w0, w1, K, dir, bbox, rho0, u0, mask, J, du0);
    } else {
      RunT<double>(steps, w0, w1, K, dir, bbox, rho0, u0, mask, J, du0);
    }
  }

  template <typename S>
  void RunT(int steps, int w0, int w1, int K, const double* dir, const double* bbox, double rho0, const double* u0,
            uint32_t mask, double* J, double* du0) {
    S* ys[2] = {static_cast<S*>(d_y[0]), static_cast<S*>(d_y[1])};
    S* xfinal = static_cast<S*>(d_final);
    std::vector<DragTerm> terms;
    std::vector<int> seed_cell;
    std::vector<double> seed_w;
    Terms(dir, bbox, &terms, &seed_cell, &seed_w);
    const int nterms = static_cast<int>(terms.size()), nseed = static_cast<int>(seed_cell.size());
    const int nck = (steps + K - 1) / K;
    const size_t st = static_cast<size_t>(ntot);
    std::vector<void*> run_owned;
    auto alloc = [&](size_t bytes) {
      void* p = nullptr;
      ADJ_CHECK(hipMalloc(&p, bytes == 0 ? 1 : bytes));
      run_owned.push_back(p);
      return p;
    };
    try {
      DragTerm* d_terms = static_cast<DragTerm*>(alloc(terms.size() * sizeof(DragTerm)));
      std::vector<int> cseed(static_cast<size_t>(ncells), -1);
      for (int k = 0; k < nseed; ++k) cseed[seed_cell[k]] = k;
      int* d_cseed = static_cast<int*>(alloc(cseed.size() * sizeof(int)));
      double* d_seed_w = static_cast<double*>(alloc(seed_w.size() * sizeof(double)));
      const int ndg = std::max(1, static_cast<int>(Blocks(nterms)));  // drag partials per step
      double* d_jv = static_cast<double*>(alloc(static_cast<size_t>(steps + 1) * ndg * sizeof(double)));
      S* ck = static_cast<S*>(alloc(static_cast<size_t>(nck) * st * sizeof(S)));
      S* seg = static_cast<S*>(alloc(static_cast<size_t>(K) * st * sizeof(S)));
      const unsigned ncb = Blocks(ncells);
      double* d_part = static_cast<double*>(alloc(static_cast<size_t>(ncb) * 3 * sizeof(double)));
      if (nterms) {
        ADJ_CHECK(hipMemcpyAsync(d_terms, terms.data(), terms.size() * sizeof(DragTerm), hipMemcpyHostToDevice, stream));
        ADJ_CHECK(hipMemcpyAsync(d_cseed, cseed.data(), cseed.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        ADJ_CHECK(hipMemcpyAsync(d_seed_w, seed_w.data(), seed_w.size() * sizeof(double), hipMemcpyHostToDevice,
                                 stream));
      }
      ADJ_CHECK(hipMemsetAsync(d_jv, 0, static_cast<size_t>(steps + 1) * ndg * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(d_dq, 0, static_cast<size_t>(nlinks > 0 ? nlinks : 1) * sizeof(double), stream));
      hipEvent_t ev[3];
      for (auto& e : ev) ADJ_CHECK(hipEventCreate(&e));
      ADJ_CHECK(hipEventRecord(ev[0], stream));
      // ---- forward: y_{s+1} = C(M y_s); x_s is stored only where the
      // reverse sweep starts from it: the checkpoints x_{cK} and every state
      // of the last segment (which is then not recomputed)
      const int s_last = (nck - 1) * K;
      auto xdest = [&](int s) -> S* {
        if (s == steps) return xfinal;
        if (s >= s_last) return seg + static_cast<size_t>(s - s_last) * st;
        return s % K == 0 ? ck + static_cast<size_t>(s / K) * st : nullptr;
      };
      hipLaunchKernelGGL(KAdjInit<S>, dim3(ncb), dim3(kThreads), 0, stream, xdest(0), d_cells, d_ckind, ncells, rho0,
                         u0[0], u0[1], u0[2], mask);
      hipLaunchKernelGGL(KAdjCollide<S>, dim3(ncb), dim3(kThreads), 0, stream, xdest(0), ys[0], d_cells, d_ckind,
                         ncells, omega);
      int cur = 0;
      for (int s = 1; s <= steps; ++s) {
        hipLaunchKernelGGL(KAdjStep<S>, dim3(ncb), dim3(kThreads), 0, stream, ys[cur], xdest(s), ys[cur ^ 1], d_src,
                           d_links, d_cells, d_ckind, d_cblock, d_cflag, d_nbase, ncells, omega);
        if (s > w0 && s <= w1 && nterms)
          hipLaunchKernelGGL(KAdjDrag<S>, dim3(ndg), dim3(kThreads), 0, stream, ys[cur], d_src, d_links, d_terms, nterms,
                             d_jv + static_cast<size_t>(s) * ndg);
        cur ^= 1;
      }
      ADJ_CHECK(hipEventRecord(ev[1], stream));
      // ---- reverse, one checkpoint segment at a time
      const double scale = 1.0 / static_cast<double>(w1 - w0);
      ADJ_CHECK(hipMemsetAsync(d_lx, 0, st * sizeof(double), stream));
      if (steps > w0 && steps <= w1 && nterms)  // live seed of the last step; the others ride in KAdjReverse
        hipLaunchKernelGGL(KAdjSeedLive, dim3(Blocks(nterms)), dim3(kThreads), 0, stream, d_lx, d_terms, nterms, scale);
      for (int c = nck - 1; c >= 0; --c) {
        const int s0 = c * K, s1 = std::min(steps, s0 + K);
        // segment states: x_{s0} in its checkpoint (or seg for the last
        // segment), x_{s0+j} in seg[j], recomputed unless the forward kept them
        S* x0 = c == nck - 1 ? seg : ck + static_cast<size_t>(c) * st;
        auto xs = [&](int j) { return j == 0 ? x0 : seg + static_cast<size_t>(j) * st; };
        if (c != nck - 1 && s1 - s0 > 1) {
          hipLaunchKernelGGL(KAdjCollide<S>, dim3(ncb), dim3(kThreads), 0, stream, x0, ys[0], d_cells, d_ckind, ncells,
                             omega);
          int yc = 0;
          for (int j = 1; j < s1 - s0; ++j) {
            hipLaunchKernelGGL(KAdjStep<S>, dim3(ncb), dim3(kThreads), 0, stream, ys[yc], xs(j), ys[yc ^ 1], d_src,
                               d_links, d_cells, d_ckind, d_cblock, d_cflag, d_nbase, ncells, omega);
            yc ^= 1;
          }
        }
        for (int s = s1; s > s0; --s) {
          const S* xp = xs(s - 1 - s0);
          auto win = [&](int st) { return st > w0 && st <= w1 && nterms > 0; };
          if (nlinks)
            hipLaunchKernelGGL(KAdjLinkGrad<S>, dim3(Blocks(nlinks)), dim3(kThreads), 0, stream, d_lx, xp, d_cells,
                               d_ckind, d_links, nlinks, omega, d_dq);
          hipLaunchKernelGGL(KAdjReverse<S>, dim3(ncb), dim3(kThreads), 0, stream, d_lx, d_ln, xp, d_rd, d_ovo, d_ove,
                             d_links, d_cells, d_ckind, d_cblock, d_cflag, d_nbase, d_cseed, d_seed_w,
                             win(s) ? scale : 0.0, win(s - 1) ? scale : 0.0, ncells, omega);
          std::swap(d_lx, d_ln);
        }
      }
      hipLaunchKernelGGL(KAdjInitGrad, dim3(ncb), dim3(kThreads), 0, stream, d_lx, d_cells, d_ckind, ncells, rho0,
                         u0[0], u0[1], u0[2], mask, d_part);
      ADJ_CHECK(hipGetLastError());
      ADJ_CHECK(hipEventRecord(ev[2], stream));
      std::vector<double> jv(static_cast<size_t>(steps + 1) * ndg), part(static_cast<size_t>(ncb) * 3);
      ADJ_CHECK(hipMemcpyAsync(jv.data(), d_jv, jv.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      ADJ_CHECK(hipMemcpyAsync(part.data(), d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      ADJ_CHECK(hipStreamSynchronize(stream));
      float a = 0.f, b = 0.f;
      ADJ_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
      ADJ_CHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
      fwd_ms = a;
      bwd_ms = b;
      for (auto& e : ev) ADJ_CHECK(hipEventDestroy(e));
      double js = 0.0;
      for (int s = w0 + 1; s <= w1; ++s) {
        double step_sum = 0.0;
        for (int g = 0; g < ndg; ++g) step_sum += jv[static_cast<size_t>(s) * ndg + g];
        js += step_sum;
      }
      *J = js * scale;
      du0[0] = du0[1] = du0[2] = 0.0;
      for (unsigned g = 0; g < ncb; ++g)
        for (int d = 0; d < 3; ++d) du0[d] += part[3 * g + d];
    } catch (...) {
      (void)hipStreamSynchronize(stream);
      for (void* p : run_owned) (void)hipFree(p);
      throw;
    }
    ADJ_CHECK(hipStreamSynchronize(stream));
    for (void* p : run_owned) ADJ_CHECK(hipFree(p));
  }
};

namespace {
template <typename F>
int AdjGuard(F&& f) {
  try {
    f();
  } catch (const std::invalid_argument& e) {
    return SetError(OCTLB_EINVAL, e.what());
  } catch (const std::bad_alloc&) {
    return SetError(OCTLB_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return SetError(OCTLB_EHIP, e.what());
  }
  return OCTLB_OK;
}
}  // namespace

extern "C" {

int octlb_adj_create(octlb_dev* dev, octlb_adj** out) {
  if (out) *out = nullptr;
  if (!dev || !out) return SetError(OCTLB_EINVAL, "null argument");
  octlb_adj* a = new (std::nothrow) octlb_adj();
  if (!a) return SetError(OCTLB_ENOMEM, "host allocation failed");
  const int rc = AdjGuard([&] {
    AdjTopo t;
    if (AdjNeedsMultiLevel(dev)) {
      AdjMLTopo m;
      ExportAdjMLTopo(dev, &t, &m);
      a->Build(t);
      a->BuildML(t, m);
    } else {
      ExportAdjTopo(dev, &t);
      a->Build(t);
    }
  });
  if (rc != OCTLB_OK) {
    a->FreeAll();
    delete a;
    return rc;
  }
  *out = a;
  return OCTLB_OK;
}

int octlb_adj_destroy(octlb_adj* adj) {
  if (!adj) return OCTLB_OK;
  if (adj->stream) (void)hipStreamSynchronize(adj->stream);  // teardown: nothing to report to
  adj->FreeAll();
  delete adj;
  return OCTLB_OK;
}

int octlb_adj_run(octlb_adj* adj, int32_t steps, int32_t win_begin, int32_t win_end, int32_t checkpoint,
                  const double dir[3], const double bbox[6], double rho0, const double u0[3], uint32_t kind_mask,
                  int64_t num_links, const int32_t* link, double* J, double dJ_du0[3], double* dJ_dq) {
  return AdjGuard([&] {
    if (!adj || !dir || !bbox || !u0 || !J || !dJ_du0) throw std::invalid_argument("null argument");
    if (num_links < 0 || (num_links > 0 && (!link || !dJ_dq))) throw std::invalid_argument("bad link list");
    if (steps < 1 || win_begin < 0 || win_end <= win_begin || win_end > steps)
      throw std::invalid_argument("adjoint: need 0 <= win_begin < win_end <= steps");
    if (checkpoint < 1) throw std::invalid_argument("adjoint: checkpoint interval must be positive");
    const int K = std::min<int>(checkpoint, steps);
    adj->Run(steps, win_begin, win_end, K, dir, bbox, rho0, u0, kind_mask, J, dJ_du0);
    std::vector<double> dq(static_cast<size_t>(adj->nlinks));
    if (adj->nlinks)
      ADJ_CHECK(hipMemcpy(dq.data(), adj->d_dq, dq.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < num_links; ++k) {
      const int b = link[3 * k], c = link[3 * k + 1], i = link[3 * k + 2];
      double g = 0.0;
      if (b >= 0 && b < adj->nb && c >= 0 && c < kNC && i >= 0 && i < kQ) {
        const auto it = adj->link_index.find((static_cast<int64_t>(b) * kNC + c) * kQ + i);
        if (it != adj->link_index.end()) g = dq[static_cast<size_t>(it->second)];
      }
      dJ_dq[k] = g;
    }
  });
}

int octlb_adj_final_core(octlb_adj* adj, double* out) {
  return AdjGuard([&] {
    if (!adj || !out) throw std::invalid_argument("null argument");
    std::vector<double> s(static_cast<size_t>(adj->ntot));
    if (adj->prec == 32) {
      std::vector<float> f(s.size());
      ADJ_CHECK(hipMemcpy(f.data(), adj->d_final, f.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < f.size(); ++k) s[k] = static_cast<double>(f[k]);
    } else {
      ADJ_CHECK(hipMemcpy(s.data(), adj->d_final, s.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int b = 0; b < adj->nb; ++b)
      for (int c = 0; c < kNC; ++c) {
        const Cell cl = adj->cells[adj->core_cell[static_cast<size_t>(b) * kNC + c]];
        for (int i = 0; i < kQ; ++i)
          out[(static_cast<size_t>(b) * kNC + c) * kQ + i] = s[static_cast<size_t>(cl.base) + static_cast<size_t>(i) * cl.stride];
      }
  });
}

int octlb_adj_inlet_grad(const octlb_adj* adj, double out[3]) {
  return AdjGuard([&] {
    if (!adj || !out) throw std::invalid_argument("null argument");
    for (int d = 0; d < 3; ++d) out[d] = adj->dui[d];
  });
}

int octlb_adj_timing(const octlb_adj* adj, double* forward_ms, double* reverse_ms) {
  return AdjGuard([&] {
    if (!adj) throw std::invalid_argument("null argument");
    if (forward_ms) *forward_ms = adj->fwd_ms;
    if (reverse_ms) *reverse_ms = adj->bwd_ms;
  });
}

int octlb_adj_set_tape_budget(octlb_adj* adj, int64_t bytes, int32_t* packed_last_run) {
  return AdjGuard([&] {
    if (!adj) throw std::invalid_argument("null argument");
    if (bytes < -1) throw std::invalid_argument("tape budget must be -1 (automatic) or >= 0 bytes");
    adj->tape_budget = bytes;
    if (packed_last_run) *packed_last_run = adj->packed_mode ? 1 : 0;
  });
}

}  // extern "C"
