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
// Layout: one FP64 state holds every block's 10^3 halo-inclusive slots, the
// reference BlockLattice's view of a block, as
//   state[(block * 19 + pop) * 1000 + slot]
// so that threads over slots read each population plane coalesced.  One step
// is y = C(x) (collide, per cell) followed by x' = M y, where M composes the
// halo exchange (faces and same-level edges, GhostSchedule) with the pull
// stream, its halo-shell wrap and the Bouzidi and bounce-back rules
// (block_lattice.cpp:503-720).  M does not depend on the state, so the host
// builds it once as rows -- a source entry per output entry, or a Bouzidi
// link -- together with its transpose in CSR form:
//   forward   KAdjCollide, KAdjStream (rows)
//   adjoint   KAdjStreamT (CSR of M^T), KAdjLinkGrad (d out / d q on the
//             Bouzidi rows), KAdjCollideT (the BGK vector-Jacobian product
//             at the stored pre-collide state)
// with drag seeds on the post-collide snapshot and the live state.  The
// forward keeps a checkpoint every K steps; the reverse sweep recomputes each
// segment of K states from its checkpoint, so memory is (steps/K + K) states.
// Objective values, dJ/du0 and dJ/dq are accumulated in FP64 in fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "d3q19.h"
#include "octlb_gpu.h"
#include "octlb_internal.h"

namespace octlb {
namespace {

#define ADJ_CHECK(expr)                                                                        \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e) +            \
                               " (adjoint.hip:" + std::to_string(__LINE__) + ")");            \
  } while (0)

constexpr int kN = 8, kNP = 10, kNS = kNP * kNP * kNP, kNC = kN * kN * kN;
constexpr int kThreads = 256;
constexpr uint8_t kHalo = 255;  // slot kind of a halo slot

struct AdjLink {
  int row;     // output entry
  int bb;      // f_bb_opp source entry
  int other;   // f_interior (q >= 1/2) or f_same (q < 1/2) source entry
  int upper;   // q >= 1/2 branch
  double q;
};

struct DragTerm {
  int snap;  // post-collide entry (cell, pop i)
  int live;  // post-stream entry (cell, pop opp(i))
  double w;  // -(dir . c_i)
};

__device__ __forceinline__ void LoadCell(const double* s, size_t base, double* f) {
#pragma unroll
  for (int i = 0; i < kQ; ++i) f[i] = s[base + static_cast<size_t>(i) * kNS];
}

// dispatch_collide for the kinds the adjoint admits (bc_dispatcher.h:32-81)
__device__ __forceinline__ void Collide(double* f, uint8_t k, double omega) {
  if (k == kBulk) {
    double rho, u[3];
    ProxyRhoU(f, rho, u);
    OlbBgk(f, rho, u, omega);
  } else if (k == kBounceBack) {
    double t[kQ];
#pragma unroll
    for (int i = 0; i < kQ; ++i) t[i] = f[i];
#pragma unroll
    for (int i = 0; i < kQ; ++i) f[i] = t[Opp(i)];
  }
}

__global__ __launch_bounds__(kThreads) void KAdjInit(double* x, const uint8_t* kslot, int nslots, double rho0,
                                                     double u0x, double u0y, double u0z, uint32_t mask) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nslots) return;
  const size_t base = static_cast<size_t>(t / kNS) * kQ * kNS + t % kNS;
  const uint8_t k = kslot[t];
  const bool on = k != kHalo && ((mask >> k) & 1u);
  const double u[3] = {u0x, u0y, u0z};
  const double usqr = NormSqr(u);
#pragma unroll
  for (int i = 0; i < kQ; ++i) x[base + static_cast<size_t>(i) * kNS] = on ? Feq(i, rho0, u, usqr) : 0.0;
}

__global__ __launch_bounds__(kThreads) void KAdjCollide(const double* x, double* y, const uint8_t* kslot,
                                                        int nslots, double omega) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nslots) return;
  const size_t base = static_cast<size_t>(t / kNS) * kQ * kNS + t % kNS;
  double f[kQ];
  LoadCell(x, base, f);
  Collide(f, kslot[t], omega);
#pragma unroll
  for (int i = 0; i < kQ; ++i) y[base + static_cast<size_t>(i) * kNS] = f[i];
}

__global__ __launch_bounds__(kThreads) void KAdjStream(const double* y, double* x, const int* src,
                                                       const AdjLink* links, int ntot) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= ntot) return;
  const int s = src[e];
  if (s >= 0) {
    x[e] = y[s];
  } else {
    const AdjLink L = links[-1 - s];
    x[e] = L.upper ? BouzidiPull(y[L.bb], y[L.other], y[L.other], L.q)
                   : BouzidiPull(y[L.bb], y[L.bb], y[L.other], L.q);
  }
}

// sum of w (snap + live) over the drag terms, fixed order (one workgroup)
__global__ __launch_bounds__(kThreads) void KAdjDrag(const double* x, const double* y, const DragTerm* terms,
                                                     int nterms, double* out) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int t = threadIdx.x; t < nterms; t += kThreads) {
    const DragTerm d = terms[t];
    acc += d.w * (y[d.snap] + x[d.live]);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// drag seeds: every term owns its snapshot and live entries (distinct pops)
__global__ __launch_bounds__(kThreads) void KAdjSeed(double* l, const DragTerm* terms, int nterms, double scale,
                                                     int live) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nterms) return;
  const DragTerm d = terms[t];
  l[live ? d.live : d.snap] += d.w * scale;
}

// ly = M^T lx: each source entry gathers the outputs that read it
__global__ __launch_bounds__(kThreads) void KAdjStreamT(const double* lx, double* ly, const int* off,
                                                        const int2* ent, const AdjLink* links, int ntot) {
  const int s = blockIdx.x * kThreads + threadIdx.x;
  if (s >= ntot) return;
  double acc = 0.0;
  for (int k = off[s]; k < off[s + 1]; ++k) {
    const int2 r = ent[k];
    const double v = lx[r.x];
    if (r.y < 0) {
      acc += v;
    } else {
      const AdjLink L = links[r.y >> 1];
      double c;
      if (L.upper) {
        const double a = 1.0 / (2.0 * L.q);
        c = (r.y & 1) ? 1.0 - a : a;
      } else {
        c = (r.y & 1) ? 1.0 - 2.0 * L.q : 2.0 * L.q;
      }
      acc += c * v;
    }
  }
  ly[s] = acc;
}

__device__ double PostCollide(const double* x, const uint8_t* kslot, int entry, double omega) {
  const int b = entry / (kQ * kNS), rem = entry % (kQ * kNS);
  const int pop = rem / kNS, slot = rem % kNS;
  double f[kQ];
  LoadCell(x, static_cast<size_t>(b) * kQ * kNS + slot, f);
  Collide(f, kslot[b * kNS + slot], omega);
  double v = 0.0;
#pragma unroll
  for (int i = 0; i < kQ; ++i) v = (i == pop) ? f[i] : v;
  return v;
}

// dq += lx[row] * d BouzidiPull / d q, with the pulled values recollided from
// the pre-collide state of the step
__global__ __launch_bounds__(kThreads) void KAdjLinkGrad(const double* lx, const double* xp, const uint8_t* kslot,
                                                         const AdjLink* links, int nlinks, double omega,
                                                         double* dq) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= nlinks) return;
  const AdjLink L = links[k];
  const double fbb = PostCollide(xp, kslot, L.bb, omega);
  const double fo = PostCollide(xp, kslot, L.other, omega);
  const double d = L.upper ? -(fbb - fo) / (2.0 * L.q * L.q) : 2.0 * fbb - 2.0 * fo;
  dq[k] += lx[L.row] * d;
}

// lx = C^T ly at the pre-collide state xp.  BGK with (rho, u) from the cell:
// lx_j = (1 - w) ly_j + w (A + sum_d G_d (c_jd - u_d) / rho),
// A = sum_i ly_i dfeq_i/drho, G_d = sum_i ly_i dfeq_i/du_d.
__global__ __launch_bounds__(kThreads) void KAdjCollideT(const double* xp, const double* ly, double* lx,
                                                         const uint8_t* kslot, int nslots, double omega) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= nslots) return;
  const size_t base = static_cast<size_t>(t / kNS) * kQ * kNS + t % kNS;
  const uint8_t k = kslot[t];
  double l[kQ];
  LoadCell(ly, base, l);
  if (k == kBulk) {
    double f[kQ];
    LoadCell(xp, base, f);
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
    const double base_g = A - (G0 * u[0] + G1 * u[1] + G2 * u[2]) / rho;
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
      const double g = base_g + (G0 * Cx(j) + G1 * Cy(j) + G2 * Cz(j)) / rho;
      lx[base + static_cast<size_t>(j) * kNS] = (1.0 - omega) * l[j] + omega * g;
    }
  } else if (k == kBounceBack) {
#pragma unroll
    for (int j = 0; j < kQ; ++j) lx[base + static_cast<size_t>(j) * kNS] = l[Opp(j)];
  } else {
#pragma unroll
    for (int j = 0; j < kQ; ++j) lx[base + static_cast<size_t>(j) * kNS] = l[j];
  }
}

// dJ/du0 partial sums per workgroup: x0 = feq(rho0, u0) on the masked cells
__global__ __launch_bounds__(kThreads) void KAdjInitGrad(const double* lx, const uint8_t* kslot, int nslots,
                                                         double rho0, double u0x, double u0y, double u0z,
                                                         uint32_t mask, double* partial) {
  __shared__ double red[3][kThreads];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  double g[3] = {0.0, 0.0, 0.0};
  if (t < nslots) {
    const uint8_t k = kslot[t];
    if (k != kHalo && ((mask >> k) & 1u)) {
      const size_t base = static_cast<size_t>(t / kNS) * kQ * kNS + t % kNS;
      const double u[3] = {u0x, u0y, u0z};
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        const double l = lx[base + static_cast<size_t>(i) * kNS];
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

}  // namespace
}  // namespace octlb

using namespace octlb;

struct octlb_adj {
  int nb = 0, nslots = 0, ntot = 0, nlinks = 0;
  double omega = 0.0;
  std::vector<uint8_t> kslot;            // [nb][1000]
  std::vector<double> origin_cw;
  std::unordered_map<int64_t, int> link_index;  // (block*512 + cell)*19 + pop -> Bouzidi row
  hipStream_t stream = nullptr;
  uint8_t* d_kslot = nullptr;
  int* d_src = nullptr;
  int* d_off = nullptr;
  int2* d_ent = nullptr;
  AdjLink* d_links = nullptr;
  double* d_dq = nullptr;
  double* d_x = nullptr;
  double* d_y = nullptr;
  double* d_lx = nullptr;
  double* d_ly = nullptr;
  double* d_final = nullptr;  // post-stream state after the last run's forward
  std::vector<void*> owned;
  double fwd_ms = 0.0, bwd_ms = 0.0;

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
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  void Build(const AdjTopo& t) {
    nb = t.nb;
    omega = t.omega;
    nslots = nb * kNS;
    if (static_cast<long long>(nb) * kQ * kNS >= (1ll << 31)) throw std::invalid_argument("adjoint: too many blocks");
    ntot = nb * kQ * kNS;
    origin_cw = t.origin_cw;
    kslot.assign(static_cast<size_t>(nslots), kHalo);
    auto kind = [&](int b, int ix, int iy, int iz) { return t.kinds[static_cast<size_t>(b) * kNC + (ix * kN + iy) * kN + iz]; };
    for (int b = 0; b < nb; ++b)
      for (int ix = 0; ix < kN; ++ix)
        for (int iy = 0; iy < kN; ++iy)
          for (int iz = 0; iz < kN; ++iz)
            kslot[static_cast<size_t>(b) * kNS + ((ix + 1) * kNP + iy + 1) * kNP + iz + 1] = kind(b, ix, iy, iz);
    std::unordered_map<int64_t, double> q;
    for (size_t k = 0; k < t.link_key.size(); ++k) q[t.link_key[k]] = t.link_q[k];
    // rows of M: stream_block (block_lattice.cpp:503-720) over the exchanged halo
    auto resolve = [&](int b, int slot, int pop) {
      const int hs = t.halo_src[static_cast<size_t>(b) * kNS + slot];
      const int sb = hs >= 0 ? hs / kNS : b, ss = hs >= 0 ? hs % kNS : slot;
      return (sb * kQ + pop) * kNS + ss;
    };
    auto in_np = [](int v) { return v >= 0 && v < kNP; };
    auto in_core = [](int v) { return v >= 0 && v < kN; };
    std::vector<int> src(static_cast<size_t>(ntot));
    std::vector<AdjLink> links;
    for (int b = 0; b < nb; ++b)
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
                src[e] = resolve(b, s, i);
                continue;
              }
              const int k = kind(b, ix, iy, iz);
              if (k == kSolid) {
                src[e] = e;
                continue;
              }
              const int self_opp = (b * kQ + Opp(i)) * kNS + dst;
              if (!s_in) {
                src[e] = ReflectsOnPull(k) ? self_opp : e;
                continue;
              }
              const bool s_core = in_core(sx - 1) && in_core(sy - 1) && in_core(sz - 1);
              if (ReflectsOnPull(k) && !s_core) {
                src[e] = self_opp;
                continue;
              }
              const int sk = s_core ? kind(b, sx - 1, sy - 1, sz - 1) : kBulk;
              if (sk == kSolid || sk == kBouzidi) {
                const int64_t key = (static_cast<int64_t>(b) * kNC + (ix * kN + iy) * kN + iz) * kQ + i;
                const auto it = q.find(key);
                const double qf = it == q.end() ? 1.0 : it->second;
                if (qf > 1.0e-10 && qf < 1.0 - 1.0e-10) {
                  const int nx = hx + Cx(i), ny = hy + Cy(i), nz = hz + Cz(i);
                  const int nk = (in_core(nx - 1) && in_core(ny - 1) && in_core(nz - 1)) ? kind(b, nx - 1, ny - 1, nz - 1)
                                                                                       : kBulk;
                  AdjLink L;
                  L.row = e;
                  L.bb = self_opp;
                  L.upper = qf >= 0.5 - 1.0e-14;
                  L.other = L.upper ? (nk == kBulk ? resolve(b, (nx * kNP + ny) * kNP + nz, i) : e) : e;
                  L.q = qf;
                  link_index[key] = static_cast<int>(links.size());
                  src[e] = -1 - static_cast<int>(links.size());
                  links.push_back(L);
                  continue;
                }
                src[e] = self_opp;
                continue;
              }
              src[e] = resolve(b, (sx * kNP + sy) * kNP + sz, i);
            }
          }
    nlinks = static_cast<int>(links.size());
    // transpose in CSR form
    std::vector<int> off(static_cast<size_t>(ntot) + 1, 0);
    for (int e = 0; e < ntot; ++e) {
      if (src[e] >= 0) {
        ++off[src[e] + 1];
      } else {
        const AdjLink& L = links[-1 - src[e]];
        ++off[L.bb + 1];
        ++off[L.other + 1];
      }
    }
    for (int s = 0; s < ntot; ++s) off[s + 1] += off[s];
    std::vector<int2> ent(static_cast<size_t>(off[ntot]));
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int e = 0; e < ntot; ++e) {
      if (src[e] >= 0) {
        ent[fill[src[e]]++] = make_int2(e, -1);
      } else {
        const int lk = -1 - src[e];
        const AdjLink& L = links[lk];
        ent[fill[L.bb]++] = make_int2(e, 2 * lk);
        ent[fill[L.other]++] = make_int2(e, 2 * lk + 1);
      }
    }
    ADJ_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    d_kslot = Alloc<uint8_t>(kslot.size());
    d_src = Alloc<int>(src.size());
    d_off = Alloc<int>(off.size());
    d_ent = Alloc<int2>(ent.size());
    d_links = Alloc<AdjLink>(links.size());
    d_dq = Alloc<double>(links.size());
    const size_t st = static_cast<size_t>(ntot);
    d_x = Alloc<double>(st);
    d_y = Alloc<double>(st);
    d_lx = Alloc<double>(st);
    d_ly = Alloc<double>(st);
    d_final = Alloc<double>(st);
    ADJ_CHECK(hipMemcpy(d_kslot, kslot.data(), kslot.size(), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
    ADJ_CHECK(hipMemcpy(d_off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ent.empty()) ADJ_CHECK(hipMemcpy(d_ent, ent.data(), ent.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!links.empty())
      ADJ_CHECK(hipMemcpy(d_links, links.data(), links.size() * sizeof(AdjLink), hipMemcpyHostToDevice));
  }

  // drag terms of ol_drag's predicate: fluid cell, in-block solid-like
  // neighbour along c_i whose centre lies in bbox
  std::vector<DragTerm> Terms(const double* dir, const double* bbox) const {
    std::vector<DragTerm> out;
    for (int b = 0; b < nb; ++b) {
      const double* o = &origin_cw[4 * static_cast<size_t>(b)];
      for (int ix = 0; ix < kN; ++ix)
        for (int iy = 0; iy < kN; ++iy)
          for (int iz = 0; iz < kN; ++iz) {
            const int slot = ((ix + 1) * kNP + iy + 1) * kNP + iz + 1;
            if (kslot[static_cast<size_t>(b) * kNS + slot] != kBulk) continue;
            for (int i = 1; i < kQ; ++i) {
              const int x = ix + Cx(i), y = iy + Cy(i), z = iz + Cz(i);
              if (x < 0 || x >= kN || y < 0 || y >= kN || z < 0 || z >= kN) continue;
              const uint8_t nk = kslot[static_cast<size_t>(b) * kNS + ((x + 1) * kNP + y + 1) * kNP + z + 1];
              if (nk != kSolid && nk != kBouzidi) continue;
              const double px = o[0] + (x + 0.5) * o[3], py = o[1] + (y + 0.5) * o[3], pz = o[2] + (z + 0.5) * o[3];
              if (!(px >= bbox[0] && px <= bbox[3] && py >= bbox[1] && py <= bbox[4] && pz >= bbox[2] &&
                    pz <= bbox[5]))
                continue;
              DragTerm d;
              d.snap = (b * kQ + i) * kNS + slot;
              d.live = (b * kQ + Opp(i)) * kNS + slot;
              d.w = -(dir[0] * Cx(i) + dir[1] * Cy(i) + dir[2] * Cz(i));
              out.push_back(d);
            }
          }
    }
    return out;
  }

  void Step(const double* x, double* y, double* xo) {
    hipLaunchKernelGGL(KAdjCollide, dim3(Blocks(nslots)), dim3(kThreads), 0, stream, x, y, d_kslot, nslots, omega);
    hipLaunchKernelGGL(KAdjStream, dim3(Blocks(ntot)), dim3(kThreads), 0, stream, y, xo, d_src, d_links, ntot);
  }

  void Run(int steps, int w0, int w1, int K, const double* dir, const double* bbox, double rho0, const double* u0,
           uint32_t mask, double* J, double* du0) {
    const std::vector<DragTerm> terms = Terms(dir, bbox);
    const int nterms = static_cast<int>(terms.size());
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
      double* d_jv = static_cast<double*>(alloc(static_cast<size_t>(steps + 1) * sizeof(double)));
      double* ck = static_cast<double*>(alloc(static_cast<size_t>(nck) * st * sizeof(double)));
      double* seg = static_cast<double*>(alloc(static_cast<size_t>(K) * st * sizeof(double)));
      const unsigned nig = Blocks(nslots);
      double* d_part = static_cast<double*>(alloc(static_cast<size_t>(nig) * 3 * sizeof(double)));
      if (nterms) ADJ_CHECK(hipMemcpyAsync(d_terms, terms.data(), terms.size() * sizeof(DragTerm), hipMemcpyHostToDevice, stream));
      ADJ_CHECK(hipMemsetAsync(d_jv, 0, static_cast<size_t>(steps + 1) * sizeof(double), stream));
      ADJ_CHECK(hipMemsetAsync(d_dq, 0, static_cast<size_t>(nlinks > 0 ? nlinks : 1) * sizeof(double), stream));
      hipEvent_t ev[3];
      for (auto& e : ev) ADJ_CHECK(hipEventCreate(&e));
      ADJ_CHECK(hipEventRecord(ev[0], stream));
      // forward with checkpoints
      hipLaunchKernelGGL(KAdjInit, dim3(nig), dim3(kThreads), 0, stream, d_x, d_kslot, nslots, rho0, u0[0], u0[1],
                         u0[2], mask);
      for (int s = 1; s <= steps; ++s) {
        if ((s - 1) % K == 0)
          ADJ_CHECK(hipMemcpyAsync(ck + static_cast<size_t>((s - 1) / K) * st, d_x, st * sizeof(double),
                                   hipMemcpyDeviceToDevice, stream));
        Step(d_x, d_y, d_lx);  // the adjoint buffer is free until the reverse sweep
        std::swap(d_x, d_lx);
        if (s > w0 && s <= w1 && nterms)
          hipLaunchKernelGGL(KAdjDrag, dim3(1), dim3(kThreads), 0, stream, d_x, d_y, d_terms, nterms, d_jv + s);
      }
      ADJ_CHECK(hipMemcpyAsync(d_final, d_x, st * sizeof(double), hipMemcpyDeviceToDevice, stream));
      ADJ_CHECK(hipEventRecord(ev[1], stream));
      // reverse sweep, one checkpoint segment at a time
      const double scale = 1.0 / static_cast<double>(w1 - w0);
      ADJ_CHECK(hipMemsetAsync(d_lx, 0, st * sizeof(double), stream));
      for (int c = nck - 1; c >= 0; --c) {
        const int s0 = c * K, s1 = std::min(steps, s0 + K);
        ADJ_CHECK(hipMemcpyAsync(seg, ck + static_cast<size_t>(c) * st, st * sizeof(double), hipMemcpyDeviceToDevice,
                                 stream));
        for (int j = 1; j < s1 - s0; ++j) Step(seg + (j - 1) * st, d_y, seg + j * st);
        for (int s = s1; s > s0; --s) {
          const double* xp = seg + static_cast<size_t>(s - 1 - s0) * st;
          const bool win = s > w0 && s <= w1 && nterms;
          if (win) hipLaunchKernelGGL(KAdjSeed, dim3(Blocks(nterms)), dim3(kThreads), 0, stream, d_lx, d_terms, nterms, scale, 1);
          hipLaunchKernelGGL(KAdjStreamT, dim3(Blocks(ntot)), dim3(kThreads), 0, stream, d_lx, d_ly, d_off, d_ent,
                             d_links, ntot);
          if (nlinks)
            hipLaunchKernelGGL(KAdjLinkGrad, dim3(Blocks(nlinks)), dim3(kThreads), 0, stream, d_lx, xp, d_kslot,
                               d_links, nlinks, omega, d_dq);
          if (win) hipLaunchKernelGGL(KAdjSeed, dim3(Blocks(nterms)), dim3(kThreads), 0, stream, d_ly, d_terms, nterms, scale, 0);
          hipLaunchKernelGGL(KAdjCollideT, dim3(nig), dim3(kThreads), 0, stream, xp, d_ly, d_lx, d_kslot, nslots, omega);
        }
      }
      hipLaunchKernelGGL(KAdjInitGrad, dim3(nig), dim3(kThreads), 0, stream, d_lx, d_kslot, nslots, rho0, u0[0],
                         u0[1], u0[2], mask, d_part);
      ADJ_CHECK(hipGetLastError());
      ADJ_CHECK(hipEventRecord(ev[2], stream));
      std::vector<double> jv(static_cast<size_t>(steps + 1)), part(static_cast<size_t>(nig) * 3);
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
      for (int s = w0 + 1; s <= w1; ++s) js += jv[s];
      *J = js * scale;
      du0[0] = du0[1] = du0[2] = 0.0;
      for (unsigned g = 0; g < nig; ++g)
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
    ExportAdjTopo(dev, &t);
    a->Build(t);
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
    ADJ_CHECK(hipMemcpy(s.data(), adj->d_final, s.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int b = 0; b < adj->nb; ++b)
      for (int c = 0; c < kNC; ++c) {
        const int ix = c / (kN * kN), iy = (c / kN) % kN, iz = c % kN;
        const int slot = ((ix + 1) * kNP + iy + 1) * kNP + iz + 1;
        for (int i = 0; i < kQ; ++i)
          out[(static_cast<size_t>(b) * kNC + c) * kQ + i] = s[(static_cast<size_t>(b) * kQ + i) * kNS + slot];
      }
  });
}

int octlb_adj_timing(const octlb_adj* adj, double* forward_ms, double* reverse_ms) {
  return AdjGuard([&] {
    if (!adj) throw std::invalid_argument("null argument");
    if (forward_ms) *forward_ms = adj->fwd_ms;
    if (reverse_ms) *reverse_ms = adj->bwd_ms;
  });
}

}  // extern "C"
